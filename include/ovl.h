/*
 * ovl.h -- C ABI of the MI355X (gfx950) implementation of the OVERLAP fermion hot path.
 *
 * The reference exposes this path as free C++ functions on GaugeLinks / ElectricFields /
 * FermionField objects (no plugin or FFI layer).  Each entry point below names the reference
 * interface it replaces.  All functions return an int status (OVL_OK = 0, negative = error);
 * ovl_error() gives the message of the last failure on a context (ovl_error(NULL): of the last
 * failed context-free call on the calling thread).
 *
 * Ownership: device memory is owned by the context; host pointers are borrowed for the
 * duration of the call only.  One context per host thread; every call is ordered on the
 * context's HIP stream and returns after the stream has drained, except where noted.
 *
 * Host layouts are the reference's:
 *   site      s = x + Nx*(y + Ny*z)                               (FermionField.cpp:31-33)
 *   links     SU(2): 4 doubles at 4*(mu+3s), U = u3 + i u_a sigma_a (SU2/GroupOperations.cpp:4-8)
 *             SU(3): 9 complex at 9*(mu+3s), (i,j) at i+3j          (SU3/GroupOperations.cpp:28-30)
 *   E field   e[a + (Nc^2-1)*(mu+3s)]                              (ElectricFields.cpp:36-41)
 *   mode      AoS complex psi[(s*Nc + colour)*4 + spin]           (FermionField.cpp:8-9,537-545)
 *   u_cc      complex ucc[((mu*V + s)*Nc + i)*Nc + j] = U_mu(s)(i,j)  (OverlapEvolutionOperator.cpp:3-4,23-51)
 * Complex numbers are interleaved (re, im) doubles.
 */
#ifndef OVL_H
#define OVL_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ovl_ctx ovl_ctx;

enum {
    OVL_OK = 0,
    OVL_E_ARG = -1,      /* invalid argument (sizes, indices, NULL, non-finite input) */
    OVL_E_HIP = -2,      /* HIP runtime error / no device */
    OVL_E_STATE = -3,    /* call order (e.g. no links uploaded) */
    OVL_E_CONVERGE = -4, /* iterative solver hit its iteration cap */
    OVL_E_NOMEM = -5,    /* device allocation failed */
    OVL_E_SPECTRUM = -6, /* invalid spectral interval (lmin <= 0 or lmin >= lmax) */
    OVL_E_NONFINITE = -7 /* a leapfrog call left Inf/NaN in Psi or PsiMid (a step far beyond the
                            leapfrog's stability bound dt |omega| < 2) */
};

/* fields held on the device: the reference's Fermions::Psi and Fermions::PsiMid
 * (FermionField.cpp:628), plus one scratch field of the same shape */
enum { OVL_PSI = 0, OVL_PSIMID = 1, OVL_SCRATCH = 2 };

enum { OVL_SIGN_CHEBYSHEV = 0, OVL_SIGN_ZOLOTAREV = 1 };
enum { OVL_PREC_FP64 = 0, OVL_PREC_FP32 = 1, OVL_PREC_MIXED = 2 };
enum { OVL_WILSON = 0, OVL_OVERLAP = 1 };
enum { OVL_GAUGE_FROZEN = 0, OVL_GAUGE_LINKS = 1, OVL_GAUGE_FULL = 2 };

typedef struct {
    int nc;              /* 2 or 3 colours */
    int n[3];            /* Nx, Ny, Nz (each >= 3) */
    int n_modes;         /* mode functions held per field */
    double r_wilson;     /* rWilson   (Simulation.cpp:194) */
    double m_fermion;    /* mFermion  (Simulation.cpp:194) */
    double m_dw;         /* MDWHeight (Simulation.cpp:200), the overlap rho */
    int mode_block;      /* modes per SoA block: 16, 32 or 64; 0 = by the mode count; -1 = timed: a
                            Clenshaw step of every storage precision is timed for each block size
                            on this lattice (unit links, zero fields) and the fastest kept; the
                            choice leads ovl_autotune's report */
    int reunitarize;     /* 1: project links back onto SU(Nc) after every link update */
    int traversal;       /* order in which the stencil launches walk (mode, site) lanes over the 8 XCDs:
                            0 = automatic (ovl_autotune picks among the others),
                            1 = lexicographic (hardware round-robin of workgroups over XCDs),
                            2 = one spatial box of sites per XCD (every mode block in turn),
                            3 = 1/8 of the mode blocks per XCD over the whole lattice (y slabs,
                                z-major); needs a multiple of 8 mode blocks */
} ovl_params;

/* sign function eps(h_W) options */
typedef struct {
    int method;          /* OVL_SIGN_CHEBYSHEV (default) or OVL_SIGN_ZOLOTAREV (reference algorithm) */
    int precision;       /* OVL_PREC_FP64 / OVL_PREC_FP32 / OVL_PREC_MIXED (Chebyshev only) */
    int degree;          /* Chebyshev degree n of P(x) ~ x^-1/2 */
    double cut;          /* spectrum cut: approximation on |h_W| in [cut, 1] * lambda_max */
    double tol;          /* MIXED: target relative error; selects where each storage tier starts */
    int head;            /* MIXED: first FP32 step, i.e. the length of the FP64 + 48-bit head
                            (>= 0 overrides tol) */
    double lmin, lmax;   /* ZOLOTAREV: interval of h_W^2; lmin = lmax = 0 -> Ritz estimate */
    int mode_chunk;      /* mode blocks per recurrence pass (0 = all) */
    int fp16_tail;       /* MIXED: 1 = store the far tail of the recurrence in FP16 (scaled by powers
                            of two, FP32 arithmetic), 0 = FP32 is the lowest tier */
    int head16;          /* MIXED with fp16_tail: first FP16 step (> 0 overrides tol; 0 = from tol) */
    int fp48;            /* MIXED: 1 = store the FP64 steps whose error bound allows it in 48 bits (the
                            top 48 bits of the double, rounded; FP64 arithmetic), 0 = plain FP64 */
    int head48;          /* MIXED with fp48: first 48-bit step (> 0 overrides tol; 0 = from tol) */
    int guard;           /* MIXED: a-posteriori accuracy guard (0 = off).  A check takes one mode
                            block of the input (Psi for the leapfrog; a different block each time;
                            in an ensemble the one flat block of every system) through the mixed
                            sign and through the FP64 sign; while a mode's relative difference (per
                            system in an ensemble) exceeds tol, the ladder is tightened a level (head overrides
                            dropped, then tol divided by 16 per level, at last plain FP64) and
                            checked again.  A check runs before the first step of an update call (or
                            before an apply) when the links or these ladder options changed since
                            the last check, or when at least `guard` leapfrog steps (applies) went
                            by.  The tightening is kept by the context until new links are set.
                            Default 100. */
    int relax;           /* MIXED with guard: levels (0..8) by which a check may loosen the ladder on
                            this lattice, each planned for 4x the tol of the last: a level whose
                            check stays within tol/2 with 4x its error to spare is tried one looser
                            on every mode block and kept while that check stays within tol/2; later
                            checks tighten again as above.  Default 0 (never loosen). */
} ovl_sign_opts;

typedef struct {
    int kind;            /* OVL_WILSON or OVL_OVERLAP */
    double dt;           /* full leapfrog step */
    /* sign.precision also selects the spinor arithmetic of OVL_WILSON: OVL_PREC_FP64 steps the FP64
     * fields, OVL_PREC_FP32 and OVL_PREC_MIXED step FP32 copies of Psi and PsiMid (converted when
     * the call starts, written back to the FP64 fields when it ends); links, E and every
     * observable stay FP64 */
    int gauge;           /* OVL_GAUGE_FROZEN / _LINKS (U <- exp(-i dt/2 E) U, E fixed) / _FULL */
    ovl_sign_opts sign;
    int use_graph;       /* capture one full step into a hipGraph and replay it; the context keeps
                            the graph, and a later call replays it without capturing while these
                            options, the FP16 scales of the sign, the context's buffers and the
                            stencil tuning are unchanged */
} ovl_step_opts;

/* ---- context ---- */
int ovl_create(const ovl_params *p, ovl_ctx **out);
int ovl_destroy(ovl_ctx *ctx);
const char *ovl_error(const ovl_ctx *ctx);
int ovl_device_name(char *buf, int len);
void ovl_sign_defaults(ovl_sign_opts *o);
/* the precision ladder the Chebyshev sign uses for these options (host arithmetic, no device):
 * steps k < *head48 in FP64, *head48 <= k < *head in 48-bit storage, *head <= k < *head16 in
 * FP32, k >= *head16 in FP16 storage */
int ovl_sign_plan(const ovl_sign_opts *o, int *head48, int *head, int *head16);

/* ---- gauge field (GaugeLinks / ElectricFields objects, LATTICE/GaugeLinks.cpp:11-89,
 *      LATTICE/ElectricFields.cpp:9-88) ---- */
int ovl_set_links(ovl_ctx *ctx, const double *links);
int ovl_get_links(ovl_ctx *ctx, double *links);
int ovl_set_efield(ovl_ctx *ctx, const double *e);
int ovl_get_efield(ovl_ctx *ctx, double *e);
/* U(1) background of QED::SetConstantMagneticField (QED/QED.cpp:131-153), charge q; folded into
 * the stencil links (the reference applies it only to its initial eigenbasis) */
int ovl_set_u1_flux(ovl_ctx *ctx, int nb, double q);

/* ---- mode functions (FermionField, FermionField.cpp:12-623) ---- */
int ovl_set_modes(ovl_ctx *ctx, int field, const double *aos, int first, int count);
int ovl_get_modes(ovl_ctx *ctx, int field, double *aos, int first, int count);
int ovl_copy_field(ovl_ctx *ctx, int src, int dst);

/* ---- operators on device fields, all modes at once ---- */
/* out = -i h_W in : WilsonEvoOperator::Compute (WilsonOperator.cpp:30-114), batched */
int ovl_apply_wilson(ovl_ctx *ctx, int in, int out);
/* out = -i h_ov in : OverlapOperator::Compute (OverlapOperator.cpp:25-64), batched */
int ovl_apply_overlap(ovl_ctx *ctx, int in, int out, const ovl_sign_opts *o);
/* out = eps(h_W) in */
int ovl_apply_sign(ovl_ctx *ctx, int in, int out, const ovl_sign_opts *o);
/* rel = max over modes ||eps(eps in) - in|| / ||in||  (OverlapSignTest.cpp:155-165) */
int ovl_sign_residual(ovl_ctx *ctx, int in, const ovl_sign_opts *o, double *rel);
/* Ritz extreme eigenvalue of +-h_W^2 (RitzSolver.cpp:92-297, ComputeRitzEigenvalues
 * OverlapUpdate.cpp:137-173); start: flat-layout host vector (m = s + V(spin + 4 colour)) */
int ovl_ritz(ovl_ctx *ctx, const double *start_flat, double sign, double delitr, int nitlz,
             int nrestart, double *eigenvalue);

/* ---- single-mode drop-ins with the reference's host arguments ---- */
/* WilsonEvoOperator::Compute(u_cc*, wcf* phi, wcf* psi) (WilsonOperator.cpp:30) */
int ovl_wilson_evo_compute(ovl_ctx *ctx, const double *ucc, const double *phi, double *psi);
/* OverlapOperator::Compute(u_cc*, wcf* phi, wcf* psi, LambdaMin, LambdaMax) (OverlapOperator.cpp:25):
 * 25-pole Zolotarev, multi-shift CG 1e-16 / 1000 iterations, as the reference */
int ovl_overlap_operator_compute(ovl_ctx *ctx, const double *ucc, const double *phi, double *psi,
                                 double lmin, double lmax);
/* OverlapOperations::ZolotarevCoefficients (ZolotarevCoefficients.cpp:71-175): host arithmetic */
double ovl_zolotarev_coefficients(double amin, double amax, double errtol, int no, double *bz, double *dz);

/* ---- dynamics ---- */
/* Dynamics::UpdateGaugeLinks (Dynamics.cpp:97-138): U_mu(x) <- exp(-i dt E_mu(x)) U_mu(x) */
int ovl_update_links(ovl_ctx *ctx, double dt);
/* Dynamics::UpdateElectricFields (Dynamics.cpp:145-197): E += dt * (-dH/dU) */
int ovl_update_efield(ovl_ctx *ctx, double dt);
/* Dynamics::Fermions::Update(U, E, Psi, PsiMid) (OverlapUpdate.cpp:175-211, WILSON/Update.cpp:48-79),
 * nsteps times */
int ovl_fermion_update(ovl_ctx *ctx, const ovl_step_opts *o, int nsteps);

/* ---- observables (FP64, deterministic reductions) ---- */
int ovl_energy(ovl_ctx *ctx, double *e_el, double *e_mag);              /* Bulk (BulkObservables.cpp) */
int ovl_gauss_violation(ovl_ctx *ctx, double *vmax, double *vrms);      /* GaussLawViolation.cpp:13-61 */
int ovl_unitarity_violation(ovl_ctx *ctx, double *vmax, double *vrms);  /* UnitarityViolation.cpp:13-51 */
/* ChernSimonsNumber::NCsDot (ChernSimonsDerivative.cpp:31-49) = improved sum E.B / (8 pi^2) */
int ovl_ncs_dot(ovl_ctx *ctx, double *ncsdot);
/* Fermions::Observables::ComputeCurrents (OverlapCurrents.cpp:5-296): jv, ja [4*site + mu]
 * (host arrays or NULL), axial charge = sum_x ja[4x]; first_mode = global index of mode 0 */
int ovl_currents(ovl_ctx *ctx, int field, const ovl_sign_opts *o, int first_mode, double *jv,
                 double *ja, double *axial_charge);
/* Dynamics::Fermions::Update x nsteps with Fermions::Observables::ComputeCurrents after every
 * `every`-th step (the reference's driver does both, Simulation.cpp:466-480).  The currents reuse
 * the overlap apply of the leapfrog's second half step, -i h_ov[U(t + dt/2)] Psi(t + dt): the same
 * operator the reference's ComputeCurrents applies after Update, since it reads the stencil links
 * of the last EvolutionOperator::Setup (OverlapUpdate.cpp:116, OverlapCurrents.cpp:134).  So no
 * extra overlap apply is made.  Chebyshev overlap, one lattice.  Outputs per observation (host
 * arrays of nsteps / every entries, or NULL): the axial charge sum_x ja[4x] and dN_CS/dt
 * (ovl_ncs_dot); jv, ja (4V, or NULL) hold the per-site currents of the last observation. */
typedef struct {
    int every;
    int first_mode;        /* global index of mode 0 (commutator weights) */
    int count;             /* out: observations written */
    double *axial_charge;
    double *ncs_dot;
    double *jv, *ja;
} ovl_observe;
int ovl_fermion_update_observed(ovl_ctx *ctx, const ovl_step_opts *o, int nsteps, ovl_observe *obs);
/* per-mode squared norms of a field (host array of n_modes doubles) */
int ovl_mode_norms(ovl_ctx *ctx, int field, double *norm2);

/* ---- ensembles: many independent lattices in one context ----
 * No counterpart in the reference, which evolves one lattice per program run
 * (Simulation.cpp:466-470 calls Dynamics::Fermions::Update on the one global GaugeLinks /
 * FermionField set); an ensemble context replaces running that program once per configuration.
 * All systems share Nc, the fermion parameters and the mode count; each has its own lattice
 * size, links, E field, U(1) flux and modes.  The systems' data are concatenated in system order:
 * ovl_set_links / ovl_get_links / ovl_set_efield / ovl_get_efield take the concatenation of the
 * per-system arrays, ovl_mode_norms returns n_systems * n_modes values (system-major).
 * ovl_copy_field, ovl_apply_wilson, ovl_apply_overlap / ovl_apply_sign / ovl_sign_residual
 * (Chebyshev sign), ovl_fermion_update, ovl_update_links / ovl_update_efield and ovl_autotune act
 * on every system at once: each stencil application is one kernel launch over all systems.
 * ovl_set_modes / ovl_get_modes, the single-lattice observables, ovl_currents, ovl_ritz, the
 * Zolotarev sign and the single-mode drop-ins return OVL_E_STATE on an ensemble; use the
 * per-system calls below (which also accept system 0 of an ordinary context). */
typedef struct {
    int nc;              /* 2 or 3 colours */
    int n_systems;       /* number of lattices (1 .. 65536) */
    const int *dims;     /* 3 * n_systems lattice sizes Nx, Ny, Nz (each >= 3), system-major */
    int n_modes;         /* mode functions per system and field */
    double r_wilson, m_fermion, m_dw;
    int mode_block;      /* 16, 32 or 64; 0 = by the mode count; -1 = timed (as ovl_params) */
    int reunitarize;
} ovl_ensemble_params;
int ovl_create_ensemble(const ovl_ensemble_params *p, ovl_ctx **out);
int ovl_set_system_links(ovl_ctx *ctx, int sys, const double *links);
int ovl_get_system_links(ovl_ctx *ctx, int sys, double *links);
int ovl_set_system_efield(ovl_ctx *ctx, int sys, const double *e);
int ovl_get_system_efield(ovl_ctx *ctx, int sys, double *e);
int ovl_set_system_u1_flux(ovl_ctx *ctx, int sys, int nb, double q);
int ovl_set_system_modes(ovl_ctx *ctx, int sys, int field, const double *aos, int first, int count);
int ovl_get_system_modes(ovl_ctx *ctx, int sys, int field, double *aos, int first, int count);
/* obs[OVL_OBS_COUNT] of one system: electric and magnetic energy, Gauss-law violation max and
 * rms, unitarity violation max and rms, dN_CS/dt -- the quantities of ovl_energy,
 * ovl_gauss_violation, ovl_unitarity_violation and ovl_ncs_dot */
enum { OVL_OBS_E_EL = 0, OVL_OBS_E_MAG, OVL_OBS_GAUSS_MAX, OVL_OBS_GAUSS_RMS, OVL_OBS_UNIT_MAX,
       OVL_OBS_UNIT_RMS, OVL_OBS_NCS_DOT, OVL_OBS_COUNT };
int ovl_system_observables(ovl_ctx *ctx, int sys, double *obs);
/* Allocation-free variant of ovl_system_observables.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under hipGraph capture, serving loops).  It
 * takes no context: `links` and `efield` are DEVICE arrays in the host layouts above (nc colours,
 * lattice n[0] x n[1] x n[2]), `obs` a DEVICE array of OVL_OBS_COUNT doubles that receives the
 * values ovl_system_observables returns for the same fields, bit for bit.  The kernels are queued on
 * `stream` (a hipStream_t of the current device; NULL = its default stream) and the call returns
 * without waiting.  It
 * calls no allocator: the reductions' partial results live in a static 32 KiB array per device,
 * shared by every call on that device, so calls that overlap on one device (other streams, other
 * threads) race and must be ordered by the caller.  At most OVL_NOALLOC_MAX_SITES sites; larger
 * lattices, nc other than 2/3 or an extent below 3 return OVL_E_ARG, and ovl_error(NULL) then gives
 * the message (per thread). */
#define OVL_NOALLOC_MAX_SITES (1L << 26)
int ovl_observables_noalloc(int nc, const int *n, const double *links, const double *efield, double *obs,
                            void *stream);
/* stencil work balance of an ensemble: max over XCDs of the workgroups assigned / mean (1 = even) */
int ovl_ensemble_balance(ovl_ctx *ctx, double *imbalance);

/* GaussLawRestoration::Restore (GAUSSLAWFIXING/GaussRestore.cpp:129-160) on system `sys` (0 for an
 * ordinary context): relaxation sweeps E_mu(x) -= gamma (G(x) - U_mu(x) G(x+mu) U_mu(x)^+) with
 * gamma = 7/24, 7/48, 7/72 per iteration until max |G^a| <= tol (the reference's 10^(1 - digits)) or
 * max_iter iterations (the reference: 10000); iters and vmax (violation of the final field) may be NULL */
int ovl_gauss_restore(ovl_ctx *ctx, int sys, double tol, int max_iter, int *iters, double *vmax);

/* OverlapQEDEigenfunctions::Setup + InitialConditions::SetOverlapQEDEigenfunctions
 * (INITIALCONDITIONS/FERMIONS/BASIS/OverlapQEDEigenfunctions.cpp:141-591,
 * SetOverlapQEDEigenfunctions.cpp:5-56): the overlap Hamiltonian of the U(1) background alone (flux
 * nb, charge q; colour 0 carries the phase and the other colours have no hopping, as the reference's
 * ConvertU1GaugeLinks sets them).  The reference diagonalises the dense D x D matrix (D = 4 Nc V);
 * here its exact block structure is used: colours do not mix, colours >= 1 reduce to one 4x4 spin
 * block, and colour 0 commutes with translations in z, so it splits into n2 z-momentum blocks of
 * size 4 n0 n1, assembled from overlap applies to the 4 n0 n1 unit sources of the z = 0 plane and
 * diagonalised together on the device (rocSOLVER zheevd, strided batched).  Mode s of Psi and PsiMid
 * (global index g = first_mode + s: particle g % 2, level g / 2) becomes sqrt(V) times eigenvector
 * D/2 - 1 - level (particle) or D/2 + level (antiparticle) of the merged descending spectrum,
 * PsiMid = Psi exp(-i omega dt/2); within a degenerate level the basis is the block one (plane
 * waves in z), any orthonormal basis of the level being an eigenbasis.  Limits: 4 n0 n1 <= 16384 and
 * n2 (4 n0 n1)^2 complex <= 32 GB.  Overwrites the scratch field; the SU(Nc) links are untouched.
 * eigenvalues (D doubles, descending) and hermiticity (max |B - B^+| over the assembled blocks) may
 * be NULL. */
int ovl_qed_eigenbasis(ovl_ctx *ctx, const ovl_sign_opts *o, int nb, double q, double dt, int first_mode,
                       double *eigenvalues, double *hermiticity);

/* ---- configuration files ----
 * IO::SaveConfiguration / IO::SaveGaugeLinks (IO/SaveConfiguration.cpp:3-63) and
 * IO::LoadConfiguration (IO/LoadConfiguration.cpp:3-144) for system `sys` (0 for an ordinary
 * context): text, one line per link "x y z mu <link reals>" and "x y z mu <E components>", sites
 * in storage order, 17 significant digits (FP64 round trip).  efile may be NULL (links only).
 * Loading checks that every link appears exactly once (the reference exits on a wrong count). */
int ovl_save_configuration(ovl_ctx *ctx, int sys, const char *ufile, const char *efile);
int ovl_load_configuration(ovl_ctx *ctx, int sys, const char *ufile, const char *efile);

/* ---- timing of the last ovl_apply_* / ovl_fermion_update call ---- */
typedef struct {
    double total_ms;          /* event-timed wall time of the call on the stream */
    double hop_ms;            /* summed event time of the stencil kernels */
    long long hop_launches;   /* number of stencil kernel launches */
    double hop_bytes;         /* algorithmic bytes of those launches */
    int fp64_steps, fp32_steps, fp16_steps, fp48_steps;
    /* per stencil-kernel variant: [0] FP64 in/out, [1] FP32 in/out, [2] FP32 in / FP64 out,
     * [3] FP16 storage in/out, [4] 48-bit storage in/out */
    double var_ms[5];
    long long var_launches[5];
    double var_bytes[5];
    int graph_captures;       /* hipGraph captures the call made (a graph-mode ovl_fermion_update
                                 whose step is unchanged since the previous one replays that graph: 0) */
    int guard_checks;         /* accuracy-guard checks the call made (each: one mode block, or all of
                                 them for a loosening try, mixed and FP64 sign; their kernels are
                                 not counted above) */
    int guard_level;          /* tightening of the mixed ladder after the call (0 = as requested;
                                 negative: loosened by that many levels) */
    double guard_err;         /* largest relative mixed-vs-FP64 difference of the last check */
} ovl_timing;
int ovl_get_timing(ovl_ctx *ctx, ovl_timing *t);
int ovl_set_profiling(ovl_ctx *ctx, int on);

/* ---- tuning (no counterpart in the reference) ----
 * times the traversal, store and occupancy variants of the stencil kernel on this context's
 * lattice and modes for every storage precision, keeps the fastest for later calls and writes a
 * one-line report */
int ovl_autotune(ovl_ctx *ctx, char *report, int len);
/* the stencil kernel's tuning per storage precision (0 = FP64, 1 = FP32, 2 = FP16, 3 = 48-bit), as
 * ovl_autotune leaves it: lane traversal (1 lexicographic, 2 XCD boxes, 3 XCD mode split; see
 * ovl_params.traversal; for an ensemble 1 = the static work table, 4 = the persistent work-queue
 * kernel, whose `slabs` is the number of tiles a workgroup claims at a time), y slabs of the
 * traversal, streaming output stores (0/1), occupancy
 * variant (minimum waves per SIMD, 1-4; 5 = one workgroup per CU through an LDS reservation, where
 * the device allows 96 KiB per workgroup).  Setting one replays a saved tuning without the sweep. */
typedef struct {
    int traversal;
    int slabs;
    int nt_store;
    int waves_per_eu;
} ovl_tuning;
int ovl_get_tuning(ovl_ctx *ctx, int precision, ovl_tuning *t);
int ovl_set_tuning(ovl_ctx *ctx, int precision, const ovl_tuning *t);

#ifdef __cplusplus
}
#endif
#endif
