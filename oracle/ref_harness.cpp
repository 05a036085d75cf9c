// ref_harness.cpp -- compiles the reference's own SU(2) hot-path source files (included from
// the reference tree at build time, never copied) into oracle/_ref/libovlref.so so that tests can
// generate golden vectors from the original code.  TEST INFRASTRUCTURE ONLY.
//
// Only files that need nothing beyond the standard library are included.  The prelude below
// supplies what the reference's Definitions.cpp (:16-18, :33-35, :46-48, :64-67), 3DGrid.cpp and
// Simulation.cpp (:194-200) define for them (type aliases, the complex unit, three macros, the
// group flags and the fermion parameters).  No MPI, GSL or FFTW header is stubbed: files that need
// them (OverlapOperator.cpp's random vector, Dynamics.cpp, ChernSimonsOperators.cpp with its
// missing plaquette macros) are not built.
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>

typedef int INT;
typedef double DOUBLE;
typedef std::complex<DOUBLE> COMPLEX;
static const int OUTPUT_PRECISION = 15;
#define ComplexI COMPLEX(0.0, 1.0)
#define SQR(x) ((x) * (x))
#define SQR_ABS(x) (SQR(real(x)) + SQR(imag(x)))
#define MOD(x, y) ((x) >= (0) ? ((x) % (y)) : (((x) + (y)) % (y)))
#define DELTA(x, y) (((x) == (y)) ? (1) : (0))
#define U1_FLAG 111
#define SU2_FLAG 222
#define SU3_FLAG 333
#define SU_Nc_FLAG SU2_FLAG

#include "LATTICE/3DGrid.cpp"
#include "DiracAlgebra.cpp"
#include "SUNc/Definitions.cpp"

DOUBLE mFermion = 0.0;
DOUBLE rWilson = 1.0;
DOUBLE MDWHeight = 1.4;

#ifndef VECDEF_H_GUARD
#define VECDEF_H_GUARD
typedef struct { COMPLEX f[4]; } wf;
typedef struct { wf c[2]; } wcf;
#endif
#ifndef GAUGEF_H_GUARD
#define GAUGEF_H_GUARD
typedef struct { COMPLEX c[2]; } u_c;
typedef struct { u_c c[2]; } u_cc;
#endif

#include "FERMIONS/WILSON/OPERATOR/WilsonOperator.cpp"
#include "FERMIONS/OVERLAP/OPERATOR/RitzSolver.cpp"
#include "FERMIONS/OVERLAP/OPERATOR/ZolotarevCoefficients.cpp"
#include "FERMIONS/OVERLAP/OPERATOR/OverlapU.cpp"

extern "C" {

void ref_set_params(const int *n, double r_wilson, double m_fermion, double m_dw)
{
    Lattice::N[0] = n[0];
    Lattice::N[1] = n[1];
    Lattice::N[2] = n[2];
    Lattice::Volume = n[0] * n[1] * n[2];
    rWilson = r_wilson;
    mFermion = m_fermion;
    MDWHeight = m_dw;
}

// WilsonEvoOperator::Compute (WilsonOperator.cpp:30)
void ref_wilson_evo(const double *ucc, const double *phi, double *psi)
{
    WilsonEvoOperator::Compute((u_cc *)ucc, (wcf *)phi, (wcf *)psi);
}

// OverlapOperations::ZolotarevCoefficients (ZolotarevCoefficients.cpp:71)
double ref_zolotarev(double amin, double amax, double errtol, int no, double *bz, double *dz)
{
    return OverlapOperations::ZolotarevCoefficients(amin, amax, errtol, no, bz, dz);
}

// OverlapOperations::Zolotrr (OverlapU.cpp:5) on the flat layout
void ref_zolotrr(const double *ucc, const double *bz, const double *dz, int no, double cgerr, int nitcgm,
                 const double *ini_flat, double *chi_flat)
{
    int mvfc = Lattice::N[0] * Lattice::N[1] * Lattice::N[2] * Nc * DiracAlgebra::SpinorComponents;
    OverlapOperations::Zolotrr((u_cc *)ucc, (DOUBLE *)bz, (DOUBLE *)dz, (COMPLEX *)chi_flat, (COMPLEX *)ini_flat, 0.0, 0.0,
                               no, cgerr, mvfc, nitcgm);
}

// OverlapOperator::Compute (OverlapOperator.cpp:25-64) minus the unused random vector (:29-47)
void ref_overlap_evo(const double *ucc, const double *phi, double *psi, double lmin, double lmax)
{
    const INT no = 25;
    DOUBLE bz[25], dz[25];
    INT mvfc = Lattice::N[0] * Lattice::N[1] * Lattice::N[2] * Nc * DiracAlgebra::SpinorComponents;
    OverlapOperations::ZolotarevCoefficients(lmin, lmax, std::pow(10.0, -12.0), no, bz, dz);
    OverlapOperations::OverlapU((u_cc *)ucc, (wcf *)phi, (wcf *)psi, bz, dz, no, mvfc, lmin, lmax, 1000,
                                std::pow(10.0, -16.0));
}

// OverlapOperations::RitzSolver (RitzSolver.cpp:92), neigen = 1
double ref_ritz(const double *ucc, const double *vec_flat, double delitr, int nitlz, double sign, int nrestart)
{
    int mvfc = Lattice::N[0] * Lattice::N[1] * Lattice::N[2] * Nc * DiracAlgebra::SpinorComponents;
    return OverlapOperations::RitzSolver((u_cc *)ucc, (COMPLEX *)vec_flat, 1, delitr, nitlz, sign, nrestart, mvfc);
}

// OverlapOperations::WilsonUU (RitzSolver.cpp:49) on the flat layout
void ref_wilson_uu(const double *ucc, const double *a_flat, double *b_flat)
{
    OverlapOperations::WilsonUU((u_cc *)ucc, (COMPLEX *)a_flat, (COMPLEX *)b_flat);
}

// SUNcAlgebra::Operations::MatrixIExp (SU2/AlgebraOperations.cpp:37)
void ref_su2_iexp(double c, const double *alpha, double *out)
{
    SUNcAlgebra::Operations::MatrixIExp(c, (DOUBLE *)alpha, out);
}

// SUNcGroup::Operations::UU (SU2/GroupOperations.cpp:103)
void ref_su2_uu(const double *a, const double *b, double *out)
{
    SUNcGroup::Operations::UU((DOUBLE *)a, (DOUBLE *)b, out);
}

// SUNcGroup::Operations::GetMatrix (SU2/GroupOperations.cpp:213), complex out[4] column-major
void ref_su2_getmatrix(const double *u, double *out)
{
    SUNcGroup::Operations::GetMatrix(u, (COMPLEX *)out);
}

// SUNcGroup::Operations::UnitarityNorm (SU2/GroupOperations.cpp:172)
double ref_su2_unitarity_norm(const double *u)
{
    return SUNcGroup::Operations::UnitarityNorm((DOUBLE *)u);
}
}
