// ovl_reference_api.hpp -- the reference's C++ call surface of the fermion hot path, on top of
// the C ABI (include/ovl.h).  Header-only; link with libovl.so.
//
// The functions keep the reference's names, argument order and argument meaning and take the
// caller's own GaugeLinks / ElectricFields / FermionField objects (any class with the members
// the reference's classes have: N[3], Volume, U or E arrays, NumberOfSamples, GetMode(s)), so
// the reference's objects bind directly:
//
//   WilsonEvoOperator::Compute(u_cc*, wcf*, wcf*)                 WilsonOperator.cpp:30
//   OverlapOperator::Compute(u_cc*, wcf*, wcf*, LambdaMin, LambdaMax)  OverlapOperator.cpp:25
//   OverlapOperations::ZolotarevCoefficients(...)                ZolotarevCoefficients.cpp:71
//   Dynamics::UpdateGaugeLinks(xLow..zHigh, U, E, dTauStep)      Dynamics.cpp:97
//   Dynamics::Fermions::Update(U, E, Psi, PsiMid)                OverlapUpdate.cpp:175 / WILSON/Update.cpp:48
//   ChernSimonsNumber::NCsDot(U, E)                              ChernSimonsDerivative.cpp:31
//   Fermions::Observables::ComputeCurrents(U, Psi, jV, jA)       OverlapCurrents.cpp:292
//   GaussLawRestoration::Restore(U, E)                            GaussRestore.cpp:129
//   IO::SaveConfiguration / IO::LoadConfiguration(U, E, files)    SaveConfiguration.cpp:3, LoadConfiguration.cpp:3
//   InitialConditions::SetOverlapQEDEigenfunctions(Psi, PsiMid)  SetOverlapQEDEigenfunctions.cpp:58
//
// Error behaviour: the reference exits or prints on bad input / non-convergence; here every
// failure throws ovl_ref::Error carrying the C-ABI status code and message (non-convergence of
// the multi-shift CG is reported through LastStatus() and stderr, as the reference prints it).
//
// State the reference keeps in globals (Lattice::N, mFermion, rWilson, MDWHeight, dTau,
// tSphaleron, the Ritz bounds) is held in ovl_ref::Settings.  By default every call uploads its
// host objects and downloads the results (safe drop-in); Settings::resident = true keeps the
// device copies between Update() calls and only downloads on ovl_ref::Sync().
#pragma once

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ovl.h"

namespace ovl_ref {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

struct Settings {
    int nc = 2;
    int n[3] = {24, 24, 24};                               // Lattice::N (3DGrid.cpp:14)
    double r_wilson = 1.0, m_fermion = 0.0, m_dw = 1.4;  // Simulation.cpp:194-200
    double dTau = 0.02;                                    // Dynamics.cpp:17
    double tSphaleron = 6.0;                               // Simulation.cpp:168
    int fermion = OVL_OVERLAP;                             // FERMION_FLAG (Simulation.cpp:3)
    ovl_sign_opts sign{};                                  // default: reference Zolotarev + Ritz
    int mode_block = 0;
    bool resident = false;
    int bflux = 1;                                         // BFlux (Simulation.cpp:196), QED flux quanta
    double qcharge = 1.0;                                  // Qs (Simulation.cpp:197)
    double gauss_tol = 1e-12;                              // 10^(1 - MAX_DIGITS_PRECISION), GaussRestore.cpp:145
    Settings()
    {
        ovl_sign_defaults(&sign);
        sign.method = OVL_SIGN_ZOLOTAREV;
        sign.lmin = sign.lmax = 0.0;  // Ritz bounds each half step (OverlapUpdate.cpp:181,195)
    }
};

inline Settings &settings()
{
    static Settings s;
    return s;
}

struct State {
    ovl_ctx *ctx = nullptr;
    int nc = 0, n[3] = {0, 0, 0}, modes = 0;
    int steps = 0;       // tSteps (Dynamics.cpp:14)
    int last_status = OVL_OK;
    bool bound = false;  // device holds the resident copies
};

inline State &state()
{
    static State s;
    return s;
}

inline void check(int rc)
{
    State &st = state();
    st.last_status = rc;
    if (rc != OVL_OK && rc != OVL_E_CONVERGE) throw Error(rc, st.ctx ? ovl_error(st.ctx) : "ovl error");
    if (rc == OVL_E_CONVERGE) std::fprintf(stderr, "Too many CG iteration steps (%s)\n", ovl_error(st.ctx));
}

inline int LastStatus() { return state().last_status; }

// (re)creates the device context for a lattice of n[0..2] sites and `modes` mode functions
inline ovl_ctx *context(const int *n, int modes)
{
    State &st = state();
    Settings &cfg = settings();
    if (st.ctx && st.nc == cfg.nc && st.n[0] == n[0] && st.n[1] == n[1] && st.n[2] == n[2] && st.modes >= modes)
        return st.ctx;
    if (st.ctx) ovl_destroy(st.ctx);
    ovl_params p{};
    p.nc = cfg.nc;
    p.n[0] = n[0];
    p.n[1] = n[1];
    p.n[2] = n[2];
    p.n_modes = modes;
    p.r_wilson = cfg.r_wilson;
    p.m_fermion = cfg.m_fermion;
    p.m_dw = cfg.m_dw;
    p.mode_block = cfg.mode_block;
    st.ctx = nullptr;
    int rc = ovl_create(&p, &st.ctx);
    if (rc != OVL_OK) throw Error(rc, "ovl_create failed (no HIP device or invalid lattice)");
    st.nc = cfg.nc;
    st.n[0] = n[0];
    st.n[1] = n[1];
    st.n[2] = n[2];
    st.modes = modes;
    st.bound = false;
    return st.ctx;
}

inline void Shutdown()
{
    State &st = state();
    if (st.ctx) ovl_destroy(st.ctx);
    st = State();
}

// ---------------------------------------------------------------- operators (one mode, host arrays)
namespace WilsonEvoOperator {
// psi = -i h_W phi (WilsonOperator.cpp:30-114); u: u_cc[3V], phi/psi: wcf[V]
template <class UCC, class WCF> void Compute(const int *N, UCC *u, WCF *phi, WCF *psi)
{
    ovl_ctx *c = context(N, 1);
    check(ovl_wilson_evo_compute(c, reinterpret_cast<const double *>(u), reinterpret_cast<const double *>(phi),
                                 reinterpret_cast<double *>(psi)));
}
// the reference's signature: the lattice comes from Settings::n (the reference reads Lattice::N)
template <class UCC, class WCF> void Compute(UCC *u, WCF *phi, WCF *psi) { Compute(settings().n, u, phi, psi); }
}  // namespace WilsonEvoOperator

namespace OverlapOperator {
// psi = -i h_ov phi (OverlapOperator.cpp:25-64) with the 25-pole Zolotarev sign of the reference
template <class UCC, class WCF> void Compute(const int *N, UCC *u, WCF *phi, WCF *psi, double LambdaMin, double LambdaMax)
{
    ovl_ctx *c = context(N, 1);
    check(ovl_overlap_operator_compute(c, reinterpret_cast<const double *>(u), reinterpret_cast<const double *>(phi),
                                       reinterpret_cast<double *>(psi), LambdaMin, LambdaMax));
}
template <class UCC, class WCF> void Compute(UCC *u, WCF *phi, WCF *psi, double LambdaMin, double LambdaMax)
{
    Compute(settings().n, u, phi, psi, LambdaMin, LambdaMax);
}
}  // namespace OverlapOperator

namespace OverlapOperations {
inline double ZolotarevCoefficients(double amin, double amax, double errtol, int no, double *bz, double *dz)
{
    return ovl_zolotarev_coefficients(amin, amax, errtol, no, bz, dz);
}
}  // namespace OverlapOperations

// ---------------------------------------------------------------- device mirrors of host objects
template <class GL> inline size_t link_doubles(GL *U) { return (size_t)U->Volume * 3 * (settings().nc == 2 ? 4 : 18); }
template <class EF> inline size_t efield_doubles(EF *E) { return (size_t)E->Volume * 3 * (settings().nc * settings().nc - 1); }

template <class GL, class EF, class FF> void upload(GL *U, EF *E, FF *Psi, FF *PsiMid)
{
    ovl_ctx *c = context(U->N, Psi->NumberOfSamples);
    check(ovl_set_links(c, reinterpret_cast<const double *>(U->U)));
    if (E) check(ovl_set_efield(c, reinterpret_cast<const double *>(E->E)));
    for (int s = 0; s < Psi->NumberOfSamples; ++s) {
        check(ovl_set_modes(c, OVL_PSI, reinterpret_cast<const double *>(Psi->GetMode(s)), s, 1));
        if (PsiMid) check(ovl_set_modes(c, OVL_PSIMID, reinterpret_cast<const double *>(PsiMid->GetMode(s)), s, 1));
    }
    state().bound = true;
}

template <class GL, class EF, class FF> void download(GL *U, EF *E, FF *Psi, FF *PsiMid)
{
    ovl_ctx *c = state().ctx;
    check(ovl_get_links(c, reinterpret_cast<double *>(U->U)));
    if (E) check(ovl_get_efield(c, reinterpret_cast<double *>(E->E)));
    for (int s = 0; s < Psi->NumberOfSamples; ++s) {
        check(ovl_get_modes(c, OVL_PSI, reinterpret_cast<double *>(Psi->GetMode(s)), s, 1));
        if (PsiMid) check(ovl_get_modes(c, OVL_PSIMID, reinterpret_cast<double *>(PsiMid->GetMode(s)), s, 1));
    }
}

// writes the device copies back into the caller's objects (resident mode)
template <class GL, class EF, class FF> void Sync(GL *U, EF *E, FF *Psi, FF *PsiMid)
{
    if (state().bound) download(U, E, Psi, PsiMid);
}

// ---------------------------------------------------------------- dynamics
namespace Dynamics {

// Dynamics.cpp:97-138 over the whole lattice (the reference always passes the full range)
template <class GL, class EF>
void UpdateGaugeLinks(int xLow, int xHigh, int yLow, int yHigh, int zLow, int zHigh, GL *U, EF *E, double dTauStep)
{
    if (xLow != 0 || yLow != 0 || zLow != 0 || xHigh != U->N[0] - 1 || yHigh != U->N[1] - 1 || zHigh != U->N[2] - 1)
        throw Error(OVL_E_ARG, "UpdateGaugeLinks: only the full lattice range is supported");
    ovl_ctx *c = context(U->N, 1);
    check(ovl_set_links(c, reinterpret_cast<const double *>(U->U)));
    check(ovl_set_efield(c, reinterpret_cast<const double *>(E->E)));
    check(ovl_update_links(c, dTauStep));
    check(ovl_get_links(c, reinterpret_cast<double *>(U->U)));
}

namespace Fermions {
// OverlapUpdate.cpp:175-211 / WILSON/Update.cpp:48-79: one staggered leapfrog step of all modes,
// links moving with E fixed while Time() < tSphaleron (Dynamics.cpp:20-22)
template <class GL, class EF, class FF> void Update(GL *U, EF *E, FF *Psi, FF *PsiMid)
{
    Settings &cfg = settings();
    State &st = state();
    if (!(cfg.resident && st.bound)) upload(U, E, Psi, PsiMid);
    ovl_step_opts o{};
    o.kind = cfg.fermion;
    o.dt = cfg.dTau;
    o.gauge = (st.steps * cfg.dTau < cfg.tSphaleron) ? OVL_GAUGE_LINKS : OVL_GAUGE_FROZEN;
    o.sign = cfg.sign;
    o.use_graph = 0;
    check(ovl_fermion_update(st.ctx, &o, 1));
    st.steps++;
    if (!cfg.resident) download(U, E, Psi, PsiMid);
}
}  // namespace Fermions
}  // namespace Dynamics

// ---------------------------------------------------------------- observables
namespace ChernSimonsNumber {
// ChernSimonsDerivative.cpp:31-49
template <class GL, class EF> double NCsDot(GL *U, EF *E)
{
    ovl_ctx *c = context(U->N, 1);
    if (!(settings().resident && state().bound)) {
        check(ovl_set_links(c, reinterpret_cast<const double *>(U->U)));
        check(ovl_set_efield(c, reinterpret_cast<const double *>(E->E)));
    }
    double r = 0;
    check(ovl_ncs_dot(c, &r));
    return r;
}
}  // namespace ChernSimonsNumber

namespace Fermions {
namespace Observables {
// OverlapCurrents.cpp:5-296: per-site vector / axial currents summed over modes, written to
// jV, jA [4*site + mu] (the reference writes them to text files on rank 0)
template <class GL, class FF> double ComputeCurrents(GL *U, FF *Psi, double *jV, double *jA, int firstMode = 0)
{
    ovl_ctx *c = context(U->N, Psi->NumberOfSamples);
    if (!(settings().resident && state().bound)) {
        check(ovl_set_links(c, reinterpret_cast<const double *>(U->U)));
        for (int s = 0; s < Psi->NumberOfSamples; ++s)
            check(ovl_set_modes(c, OVL_PSI, reinterpret_cast<const double *>(Psi->GetMode(s)), s, 1));
    }
    double q = 0;
    check(ovl_currents(c, OVL_PSI, &settings().sign, firstMode, jV, jA, &q));
    return q;
}
}  // namespace Observables
}  // namespace Fermions

namespace GaussLawRestoration {
// GaussRestore.cpp:129-160: relaxation of E until max |G^a| <= Settings::gauss_tol (10000 iterations
// at most); returns the iteration count
template <class GL, class EF> int Restore(GL *U, EF *E)
{
    ovl_ctx *c = context(U->N, 1);
    check(ovl_set_links(c, reinterpret_cast<const double *>(U->U)));
    check(ovl_set_efield(c, reinterpret_cast<const double *>(E->E)));
    int it = 0;
    double vmax = 0;
    check(ovl_gauss_restore(c, 0, settings().gauss_tol, 10000, &it, &vmax));
    check(ovl_get_efield(c, reinterpret_cast<double *>(E->E)));
    return it;
}
}  // namespace GaussLawRestoration

namespace IO {
// SaveConfiguration.cpp:3-63 (E file optional: SaveGaugeLinks) and LoadConfiguration.cpp:3-144;
// the reference composes the file names from its output directory and seed, here they are given
template <class GL, class EF> void SaveConfiguration(GL *U, EF *E, const std::string &ufile, const std::string &efile)
{
    ovl_ctx *c = context(U->N, 1);
    check(ovl_set_links(c, reinterpret_cast<const double *>(U->U)));
    if (E) check(ovl_set_efield(c, reinterpret_cast<const double *>(E->E)));
    check(ovl_save_configuration(c, 0, ufile.c_str(), E ? efile.c_str() : nullptr));
}
template <class GL, class EF> void LoadConfiguration(GL *U, EF *E, const std::string &ufile, const std::string &efile)
{
    ovl_ctx *c = context(U->N, 1);
    check(ovl_load_configuration(c, 0, ufile.c_str(), E ? efile.c_str() : nullptr));
    check(ovl_get_links(c, reinterpret_cast<double *>(U->U)));
    if (E) check(ovl_get_efield(c, reinterpret_cast<double *>(E->E)));
    state().bound = false;
}
}  // namespace IO

namespace InitialConditions {
// SetOverlapQEDEigenfunctions.cpp:58-86: eigenbasis of the overlap Hamiltonian in the U(1)
// background (Settings::bflux, Settings::qcharge), Psi = sqrt(V) v, PsiMid = Psi exp(-i omega dTau/2)
template <class FF> void SetOverlapQEDEigenfunctions(FF *Psi, FF *PsiMid, int firstMode = 0)
{
    Settings &cfg = settings();
    ovl_ctx *c = context(Psi->N, Psi->NumberOfSamples);
    check(ovl_qed_eigenbasis(c, &cfg.sign, cfg.bflux, cfg.qcharge, cfg.dTau, firstMode, nullptr, nullptr));
    for (int s = 0; s < Psi->NumberOfSamples; ++s) {
        check(ovl_get_modes(c, OVL_PSI, reinterpret_cast<double *>(Psi->GetMode(s)), s, 1));
        check(ovl_get_modes(c, OVL_PSIMID, reinterpret_cast<double *>(PsiMid->GetMode(s)), s, 1));
    }
    state().bound = false;
}
}  // namespace InitialConditions

}  // namespace ovl_ref
