/*
 * ovl_oracle.h -- FP64 CPU restatement of the OVERLAP fermion hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * product (overlap_amd/) and the timed CPU baseline of bench.py.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product never links or calls it.
 *
 * Layouts follow the reference exactly:
 *   site        s = x + Nx*(y + Ny*z)                      (FermionField.cpp:31-33)
 *   spinor AoS  psi[(s*Nc + colour)*4 + spin]  (wcf.c[i].f[alpha], FermionField.cpp:8-9)
 *   flat vector m = s + V*(spin + 4*colour)               (OverlapU.cpp:163, RitzSolver.cpp:60)
 *   links SU(2) 4 reals at 4*(mu + 3*s), U = u3 + i u_a sigma_a (SU2/GroupOperations.cpp:4-8)
 *   links SU(3) 9 complex at 9*(mu + 3*s), element (i,j) at i+3j (SU3/GroupOperations.cpp:28-30)
 *   u_cc        ucc[(mu*V + s)*Nc*Nc + i*Nc + j] = U_mu(s)(i,j)  (OverlapEvolutionOperator.cpp:23-51)
 *   E field     e[a + (Nc^2-1)*(mu + 3*s)]                (ElectricFields.cpp:36-41)
 *
 * Functions with no golden vector or closed-form check are marked
 * "parity unpinned" at their definition.
 */
#ifndef OVL_ORACLE_H
#define OVL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nc;            /* colours: 2 or 3 */
    int n[3];          /* Nx, Ny, Nz */
    double r_wilson;   /* rWilson  (Simulation.cpp:194) */
    double m_fermion;  /* mFermion (Simulation.cpp:194) */
    double m_dw;       /* MDWHeight (Simulation.cpp:200); the overlap rho */
} orc_lat;

/* complex numbers are passed as interleaved (re, im) doubles */

/* ---- gauge data ---- */
void orc_u1_phases(const int n[3], int nb, double q, double *phase /* 2*3V */);
void orc_links_to_ucc(const orc_lat *L, const double *links, const double *phase, double *ucc);
void orc_su2_iexp(double c, const double *alpha, double *out4);
void orc_su3_iexp(double c, const double *alpha, double *out18);
void orc_update_links(const orc_lat *L, double *links, const double *efield, double dt);
void orc_efield_update(const orc_lat *L, const double *links, double *efield, double dt);
/* every link projected back onto SU(Nc): SU(2) V/|V|, SU(3) polar factor with det 1 */
void orc_reunitarize(const orc_lat *L, double *links);

/* ---- operators (one mode, AoS) ---- */
void orc_wilson_evo(const orc_lat *L, const double *ucc, const double *phi, double *psi);
void orc_hw(const orc_lat *L, const double *ucc, const double *phi, double *psi);
double orc_hw_norm_bound(const orc_lat *L);

/* Chebyshev sign function */
void orc_cheb_coeffs(int n, double cut, double *c);
void orc_sign_cheb(const orc_lat *L, const double *ucc, double lam, double cut, int n,
                   const double *c, const double *phi, double *out);
void orc_overlap_evo_cheb(const orc_lat *L, const double *ucc, double lam, double cut, int n,
                          const double *c, const double *phi, double *psi);

/* Zolotarev / multi-shift CG sign function (reference algorithm) */
double orc_zolotarev(double amin, double amax, double errtol, int no, double *bz, double *dz);
int orc_zolotrr(const orc_lat *L, const double *ucc, const double *bz, const double *dz, int no,
                double cgerr, int nitcgm, const double *ini_aos, double *chi_aos);
int orc_overlap_evo_zolo(const orc_lat *L, const double *ucc, const double *phi, double *psi,
                         double lmin, double lmax);
double orc_ritz(const orc_lat *L, const double *ucc, const double *vec_flat, double delitr,
                int nitlz, double sign, int nrestart);

/* ---- leapfrog (all modes; modes stored one after another, AoS each) ---- */
enum { ORC_WILSON = 0, ORC_OVERLAP_CHEB = 1, ORC_OVERLAP_ZOLO = 2 };
typedef struct {
    int kind;          /* ORC_WILSON / ORC_OVERLAP_CHEB / ORC_OVERLAP_ZOLO */
    double dt;         /* full step */
    int update_links;  /* 1: U <- exp(-i dt/2 E) U after each half step (OverlapUpdate.cpp:111-127); 2: E too */
    int nb; double q;  /* U(1) flux quanta and charge folded into the stencil links */
    int cheb_n; double cheb_cut;
    double zolo_lmin, zolo_lmax;
    int threads;       /* OpenMP threads over modes (0 = default) */
} orc_step_opts;
void orc_leapfrog(const orc_lat *L, double *links, double *efield, int nmodes,
                  double *psi, double *psimid, const orc_step_opts *o, int nsteps);

/* ---- reductions ---- */
void orc_plaquette_energy(const orc_lat *L, const double *links, const double *efield,
                          double *e_el, double *e_mag);
void orc_gauss_violation(const orc_lat *L, const double *links, const double *efield,
                         double *gmax, double *grms);
void orc_unitarity_violation(const orc_lat *L, const double *links, double *umax, double *urms);
/* GaussLawRestoration::UpdateStep (GaussRestore.cpp:115-120); returns max |G^a| before the step */
double orc_gauss_restore_step(const orc_lat *L, const double *links, double *efield, double gamma);
/* GaussLawRestoration::Restore (GaussRestore.cpp:129-160); returns iterations, *vmax final max |G^a| */
int orc_gauss_restore(const orc_lat *L, const double *links, double *efield, double tol, int max_iter,
                      double *vmax);
double orc_edotb(const orc_lat *L, const double *links, const double *efield);
void orc_currents(const orc_lat *L, const double *psi, const double *hpsi, int nmodes,
                  int first_mode, double *jv, double *ja);

#ifdef __cplusplus
}
#endif
#endif
