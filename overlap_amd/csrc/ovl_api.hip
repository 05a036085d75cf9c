// Host side of the C ABI (include/ovl.h): device state, the sign-function drivers and the
// leapfrog.  Scalar recurrences (Clenshaw coefficients, multi-shift CG and Ritz scalars) run on
// the host in FP64; every vector operation runs in the gfx950 kernels of ovl_kernels.hip.
#include "../../include/ovl.h"
#include "ovl_internal.hpp"

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

using ovl::Geom;
using ovl::HopArgs;
using ovl::P32;
using ovl::P64;
using ovl::P16;
using ovl::P48;
using ovl::NUM_PREC;
using ovl::SysDesc;
typedef std::complex<double> cplx;

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

// stencil variants timed separately: [0] FP64, [1] FP32, [2] FP32 in / FP64 out, [3] FP16, [4] 48-bit
constexpr int NVAR = 5;

// The host-side inputs of a captured leapfrog step: the arguments of its kernels are built from
// the options, the sign plan (ladder and FP16 scales), the context's device pointers and the
// stencil tuning.  A kept graph is replayed only while all of them are unchanged; the contents of
// device memory (links, fields, phases, site orders) are read by every replay afresh.
struct StepKey {
    ovl_step_opts o{};
    int head48 = 0, head = 0, head16 = 0;
    double sx = 0.0;
    std::vector<double> sk, sw;
    std::vector<const void *> ptrs;
    std::vector<int> ints;
};

struct ovl_ctx {
    ovl_params prm{};
    std::string mb_report;  // the timed mode-block choice (mode_block -1), if one was made
    Geom g{};
    int ncomp = 0;
    size_t field64_bytes = 0, field32_bytes = 0;  // full field (all blocks)
    size_t blk64_elems = 0, blk32_elems = 0;      // per mode block (double2 / float4 units)
    size_t blk16_elems = 0;                       // per mode block (16-byte FP16 units)
    size_t blk48_elems = 0;                       // per mode block (16-byte units of the 48-bit layout)
    hipStream_t st = nullptr;
    std::string err;

    double *links = nullptr, *efield = nullptr;  // master FP64, reference layout
    double *ul64 = nullptr;                      // stencil links FP64 (double2)
    float *ul32 = nullptr;                       // stencil links FP32 (float2)
    double *phase = nullptr;                     // U(1) phases double2[3V] or null
    int nb = 0;
    double q = 1.0;
    bool links_set = false, efield_set = false, stencil_dirty = true;
    bool stencil_custom = false;  // the stencil links hold a special operator (QED eigenbasis)

    void *field[3] = {nullptr, nullptr, nullptr};  // Psi, PsiMid, scratch (FP64)
    // FP32 copies of Psi and PsiMid that the FP32 Wilson leapfrog steps in (allocated on first use;
    // filled from the FP64 fields when a call starts and written back when it ends)
    void *f32[2] = {nullptr, nullptr};

    // a-posteriori accuracy guard of the mixed sign: tightening level of the ladder (kept until new
    // links are set), the next mode block to sample, two FP64 one-block results, steps (or applies)
    // since the last check, whether the links or the ladder options changed since it (guard_opts:
    // the options it checked), the checks of the current call and the largest difference of the last
    int guard_level = 0, guard_next = 0, guard_since = 0, guard_checks = 0;
    bool guard_fresh = true;
    ovl_sign_opts guard_opts{};
    double guard_err = 0.0;
    void *guard_buf[2] = {nullptr, nullptr};
    size_t guard_bytes = 0;  // size of each guard buffer

    // Clenshaw work buffers (sized for `work_blocks` mode blocks)
    int work_blocks = 0;
    void *w64 = nullptr, *a64 = nullptr, *b64 = nullptr;
    void *x32 = nullptr, *w32 = nullptr, *a32 = nullptr, *b32 = nullptr;
    void *x16 = nullptr, *w16 = nullptr, *a16 = nullptr, *b16 = nullptr;
    void *x48 = nullptr, *w48 = nullptr, *a48 = nullptr, *b48 = nullptr;

    // reductions / small scratch
    double *partial = nullptr;
    size_t partial_len = 0;
    bool boxes_ok = false;     // the lattice splits into 8 equal boxes (TRAV_BOXES possible)
    // per input precision (P64, P32, P16, P48) of the stencil, set by ovl_autotune:
    int *order[NUM_PREC] = {nullptr, nullptr, nullptr, nullptr};  // device site order (null: lexicographic)
    bool order_set[NUM_PREC] = {false, false, false, false};
    int trav[NUM_PREC] = {ovl::TRAV_LEX, ovl::TRAV_LEX, ovl::TRAV_LEX, ovl::TRAV_LEX};  // lane traversal
    int slabs[NUM_PREC] = {1, 1, 1, 1};         // y slabs (TRAV_MODES) or sub-slabs per box (TRAV_BOXES)
    int nt_out[NUM_PREC] = {0, 0, 0, 0};        // streaming output stores
    int wpe[NUM_PREC] = {1, 1, 1, 1};           // occupancy variant
    bool tuned = false;                         // ovl_set_tuning or ovl_autotune has run
    double step_xmax = 0.0;  // bound of the mode norms during ovl_fermion_update (FP16 scales)
    // the last sign polynomial (a function of degree and cut only): its Chebyshev coefficients and
    // the Clenshaw bounds of the ladder (t0 = -1) and of the FP16 scales (t0 = beta), empty until
    // first needed; every sign apply plans, and the coefficients alone cost milliseconds of host time
    int memo_n = -1;
    double memo_cut = 0.0;
    std::vector<double> memo_coef, memo_B1, memo_Bb;
    double *dout = nullptr;  // 16 doubles
    double *gam = nullptr;   // current coupling matrices
    double *jv = nullptr, *ja = nullptr;
    double *staging = nullptr;
    size_t staging_bytes = 0;

    // multi-shift CG
    int ms_no = 0;
    void *ms_ps = nullptr, *ms_r = nullptr, *ms_q = nullptr, *ms_w = nullptr, *ms_chi = nullptr;
    double *ms_coef = nullptr;
    int *ms_active = nullptr;

    // Ritz vectors (one block)
    void *rz[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};

    // timing
    bool profiling = false;
    std::vector<hipEvent_t> evpool;
    std::vector<int> evvar;  // stencil variant of each recorded event pair
    size_t evused = 0;
    long long var_launches[NVAR] = {};
    double var_bytes[NVAR] = {};
    long long hop_launches = 0;
    double hop_bytes = 0;
    int fp64_steps = 0, fp32_steps = 0, fp16_steps = 0, fp48_steps = 0;
    ovl_timing last{};
    hipEvent_t t0 = nullptr, t1 = nullptr;
    bool capturing = false;
    int captures = 0;  // hipGraph captures made by the current call

    // the last captured leapfrog step, kept for later calls with the same key, and its per-step
    // counter increments (stencil launches and bytes, per variant, tier steps)
    hipGraphExec_t step_exec = nullptr;
    StepKey step_key;
    long long step_dl = 0, step_dvl[NVAR] = {};
    double step_db = 0, step_dvb[NVAR] = {};
    int step_dp[4] = {};

    // lattices held: one for an ordinary context, or the systems of an ensemble, stored one after
    // another in every field (site-blocks q0..), link and E-field array (link index l0..)
    bool ensemble = false;
    std::vector<SysDesc> sys;
    std::vector<int> sys_nb;            // U(1) flux quanta per system (0 = no phases)
    std::vector<char> sys_links_set;
    long link_sites = 0;                // sites summed over the systems
    SysDesc *d_sys = nullptr;           // device copy of sys (ensemble stencil)
    int2 *d_work = nullptr;             // ensemble stencil work table
    int ens_grid = 0;
    int2 *d_qwork = nullptr;            // the same tiles queue by queue (XCD k's systems together)
    int *d_qoff = nullptr;              // queue k: d_qwork[d_qoff[k] .. d_qoff[k + 1])
    int *d_qctr = nullptr;              // claim counters [8] and done counters [8] of the queues
    double ens_imbalance = 1.0;
};

namespace {

#define HIPCHK(ctx, call)                                                                                  \
    do {                                                                                                   \
        hipError_t e_ = (call);                                                                            \
        if (e_ != hipSuccess) {                                                                            \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                                \
            return OVL_E_HIP;                                                                              \
        }                                                                                                  \
    } while (0)

#define RETCHK(expr)              \
    do {                          \
        int r_ = (expr);          \
        if (r_ != OVL_OK) return r_; \
    } while (0)

int fail(ovl_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

int dalloc(ovl_ctx *c, void **p, size_t bytes)
{
    if (*p) return OVL_OK;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(c, OVL_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    }
    return OVL_OK;
}

void dfree(void *&p)
{
    if (p) hipFree(p);
    p = nullptr;
}

double lambda_bound(const ovl_ctx *c)
{
    // ||h_W|| <= |A|/2 + 3: every direction's hopping block squares to 1 for unitary links
    return std::fabs(3.0 * c->prm.r_wilson + c->prm.m_fermion - c->prm.m_dw) + 3.0;
}

double amass(const ovl_ctx *c) { return 6.0 * c->prm.r_wilson + 2.0 * c->prm.m_fermion - 2.0 * c->prm.m_dw; }

// per-system views of the concatenated arrays
int link_reals(const ovl_ctx *c) { return c->g.nc == 2 ? 4 : 18; }
int n_adj(const ovl_ctx *c) { return c->g.nc * c->g.nc - 1; }

Geom sys_geom(const ovl_ctx *c, int s)
{
    Geom g = c->g;
    const SysDesc &d = c->sys[s];
    g.n0 = d.n0;
    g.n1 = d.n1;
    g.n2 = d.n2;
    g.V = d.V;
    g.nblk = d.nblk;
    return g;
}

double *sys_links(const ovl_ctx *c, int s) { return c->links + (size_t)c->sys[s].l0 * link_reals(c); }
double *sys_efield(const ovl_ctx *c, int s) { return c->efield + (size_t)c->sys[s].l0 * n_adj(c); }
double *sys_ul64(const ovl_ctx *c, int s) { return c->ul64 + (size_t)c->sys[s].l0 * c->g.nc * c->g.nc * 2; }
float *sys_ul32(const ovl_ctx *c, int s) { return c->ul32 + (size_t)c->sys[s].l0 * c->g.nc * c->g.nc * 2; }
const double *sys_phase(const ovl_ctx *c, int s)
{
    return (c->phase && c->sys_nb[s] != 0) ? c->phase + (size_t)c->sys[s].l0 * 2 : nullptr;
}
// FP64 field of the context -> first site-block of system s
void *sys_field(const ovl_ctx *c, const void *f, int s)
{
    return (char *)f + (size_t)c->sys[s].q0 * c->ncomp * c->g.mb * 16;
}

int nsys(const ovl_ctx *c) { return (int)c->sys.size(); }

int check_sys(ovl_ctx *c, int s)
{
    if (s < 0 || s >= nsys(c)) return fail(c, OVL_E_ARG, "system index out of range");
    return OVL_OK;
}

int single_lattice(ovl_ctx *c, const char *what)
{
    if (c->ensemble)
        return fail(c, OVL_E_STATE, std::string(what) + " acts on one lattice; on an ensemble use the per-system calls");
    return OVL_OK;
}

int check_field(ovl_ctx *c, int f)
{
    if (f < 0 || f > 2) return fail(c, OVL_E_ARG, "field index must be OVL_PSI, OVL_PSIMID or OVL_SCRATCH");
    return OVL_OK;
}

int ensure_stencil(ovl_ctx *c)
{
    if (c->stencil_custom) return OVL_OK;  // stencil links set directly for a special operator
    if (!c->links_set) return fail(c, OVL_E_STATE, "no gauge links uploaded (ovl_set_links)");
    if (c->stencil_dirty) {
        for (int s = 0; s < nsys(c); ++s)
            HIPCHK(c, ovl::launch_link_setup(sys_geom(c, s), sys_links(c, s), sys_phase(c, s), sys_ul64(c, s),
                                             sys_ul32(c, s), c->st));
        c->stencil_dirty = false;
    }
    return OVL_OK;
}

void *blk_ptr(void *base, int prec, size_t blk, const ovl_ctx *c)
{
    if (!base) return nullptr;
    if (prec == P64) return (char *)base + blk * c->blk64_elems * 16;
    if (prec == P16) return (char *)base + blk * c->blk16_elems * 16;
    if (prec == P48) return (char *)base + blk * c->blk48_elems * 16;
    return (char *)base + blk * c->blk32_elems * 16;
}

// the next event of the pool (nullptr when no event can be created: that launch goes untimed)
hipEvent_t next_event(ovl_ctx *c)
{
    if (c->evused == c->evpool.size()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->evpool.push_back(e);
    }
    return c->evpool[c->evused++];
}

// algorithmic HBM bytes of one stencil launch: input stencil read once, pointwise operands and
// output once, links once per launch
double hop_alg_bytes(const ovl_ctx *c, const HopArgs &a, int pin, int pout)
{
    const double sites_modes = (double)a.nblk * c->g.V * c->g.mb;
    auto real_bytes = [](int p) { return p == P64 ? 8.0 : (p == P32 ? 4.0 : (p == P48 ? 6.0 : 2.0)); };
    const double Sin = c->ncomp * 2.0 * real_bytes(pin);
    const double Sout = c->ncomp * 2.0 * real_bytes(pout);
    double b = sites_modes * (Sin + Sout);
    if (a.x) b += sites_modes * Sout;
    if (a.b1) b += sites_modes * Sout;
    if (a.b2) b += sites_modes * Sout;
    if (a.tgt) b += sites_modes * Sout;
    // FP16 uses the FP32 links, the 48-bit layout the FP64 links
    b += 3.0 * c->link_sites * c->g.nc * c->g.nc * 2.0 * ((pin == P64 || pin == P48) ? 8 : 4);
    return b;
}

int hop(ovl_ctx *c, const Geom &g, int pin, int pout, const HopArgs &a)
{
    const int var = (pin == P64) ? 0 : (pin == P16 ? 3 : (pin == P48 ? 4 : (pout == P32 ? 1 : 2)));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->profiling && !c->capturing) {
        e0 = next_event(c);
        e1 = e0 ? next_event(c) : nullptr;
        if (e1) {
            c->evvar.resize(c->evused / 2);
            c->evvar[c->evused / 2 - 1] = var;
            HIPCHK(c, hipEventRecord(e0, c->st));
        } else if (e0) {
            c->evused--;  // an unpaired event stays unused
        }
    }
    HIPCHK(c, ovl::launch_hop(g, pin, pout, a, c->wpe[pin], c->st));
    if (e1) HIPCHK(c, hipEventRecord(e1, c->st));
    const double b = hop_alg_bytes(c, a, pin, pout);
    c->hop_launches++;
    c->hop_bytes += b;
    c->var_launches[var]++;
    c->var_bytes[var] += b;
    return OVL_OK;
}

HopArgs hop_args(const ovl_ctx *c, int prec)
{
    HopArgs a{};
    a.links = (prec == P64 || prec == P48) ? (const void *)c->ul64 : (const void *)c->ul32;
    a.amass = amass(c);
    a.nblk = c->g.nblk;
    a.trav = c->trav[prec];
    a.nt_out = c->nt_out[prec];
    a.order = c->order[prec];
    if (c->ensemble) {
        a.trav = c->trav[prec] == ovl::TRAV_QUEUE ? ovl::TRAV_QUEUE : ovl::TRAV_LEX;
        a.order = nullptr;
        a.ens_sys = c->d_sys;
        a.ens_work = c->d_work;
        a.ens_grid = c->ens_grid;
        if (a.trav == ovl::TRAV_QUEUE) {
            a.ens_qwork = c->d_qwork;
            a.ens_qoff = c->d_qoff;
            a.ens_qctr = c->d_qctr;
            a.ens_per = c->slabs[prec];
        }
    }
    return a;
}

// Site traversal for the XCD-aware stencil launches: the lattice is cut into 8 boxes (one per XCD,
// px*py*pz = 8 dividing the extents); inside a box the shorter of the y/z box extents is cut into
// `sub` slabs, each walked with x fastest, then that extent, then the other one, so the three
// planes a site's neighbours span (b_x * slab sites each) fit the XCD's L2.  Empty when no split fits.
std::vector<int> xcd_site_order(int n0, int n1, int n2, int sub = 1)
{
    const int n[3] = {n0, n1, n2};
    int best[3] = {0, 0, 0};
    long best_cost = -1;
    for (int px = 1; px <= 8; px *= 2)
        for (int py = 1; px * py <= 8; py *= 2) {
            const int pz = 8 / (px * py);
            if (n0 % px || n1 % py || n2 % pz) continue;
            // prefer compact boxes: minimise the box surface
            const long bx = n0 / px, by = n1 / py, bz = n2 / pz;
            const long cost = bx * by + by * bz + bz * bx;
            if (best_cost < 0 || cost < best_cost) {
                best_cost = cost;
                best[0] = px; best[1] = py; best[2] = pz;
            }
        }
    std::vector<int> order;
    if (best_cost < 0) return order;
    const int b[3] = {n[0] / best[0], n[1] / best[1], n[2] / best[2]};
    const bool y_mid = b[1] <= b[2];
    const int slow = y_mid ? b[2] : b[1], mid = y_mid ? b[1] : b[2];
    sub = std::max(1, std::min(sub, mid));
    order.reserve((size_t)n0 * n1 * n2);
    for (int kz = 0; kz < best[2]; ++kz)
        for (int ky = 0; ky < best[1]; ++ky)
            for (int kx = 0; kx < best[0]; ++kx)
                for (int t = 0; t < sub; ++t)
                    for (int i = 0; i < slow; ++i)
                        for (int j = t * mid / sub; j < (t + 1) * mid / sub; ++j)
                            for (int x = 0; x < b[0]; ++x) {
                                const int y = y_mid ? j : i, z = y_mid ? i : j;
                                const int gx = kx * b[0] + x, gy = ky * b[1] + y, gz = kz * b[2] + z;
                                order.push_back(gx + n0 * (gy + n1 * gz));
                            }
    return order;
}

// Site order of the mode-split traversal: the lattice is cut into `t` slabs along y and each slab
// is walked z-major (x fastest), so that the planes z-1, z, z+1 of one slab -- the reuse distance
// of the z neighbours -- are what the XCD's L2 has to hold.
std::vector<int> slab_site_order(int n0, int n1, int n2, int t)
{
    std::vector<int> ord;
    ord.reserve((size_t)n0 * n1 * n2);
    for (int k = 0; k < t; ++k) {
        const int y0 = (int)((long)k * n1 / t), y1 = (int)((long)(k + 1) * n1 / t);
        for (int z = 0; z < n2; ++z)
            for (int y = y0; y < y1; ++y)
                for (int x = 0; x < n0; ++x) ord.push_back(x + n0 * (y + n1 * z));
    }
    return ord;
}

// slab count that keeps three slab planes of one mode block of precision `prec` within ~3 MB
// of the 4 MB L2 of an XCD
int default_slabs(const ovl_ctx *c, int prec)
{
    const double site_bytes = (double)c->g.mb * c->ncomp * (prec == P64 ? 16 : (prec == P32 ? 8 : (prec == P48 ? 12 : 4)));
    int t = 1;
    while (t < c->g.n1 && 3.0 * c->g.n0 * ((c->g.n1 + t - 1) / t) * site_bytes > 3.0 * (1 << 20)) ++t;
    return t;
}

// select the traversal of precision `prec`; uploads the slab order when it changes
int set_traversal(ovl_ctx *c, int prec, int trav, int slabs)
{
    if (c->ensemble) {
        // ensembles: the static table (TRAV_LEX) or the work queues with `slabs` tiles per claim
        c->trav[prec] = trav == ovl::TRAV_QUEUE ? ovl::TRAV_QUEUE : ovl::TRAV_LEX;
        c->slabs[prec] = c->trav[prec] == ovl::TRAV_QUEUE ? std::max(1, std::min(slabs, 64)) : 1;
        return OVL_OK;
    }
    if (trav == ovl::TRAV_QUEUE) trav = ovl::TRAV_LEX;
    if (trav == ovl::TRAV_BOXES && !c->boxes_ok) trav = ovl::TRAV_LEX;
    slabs = std::max(1, std::min(slabs, c->g.n1));
    if (trav == ovl::TRAV_LEX) slabs = 1;
    if (c->order_set[prec] && trav == c->trav[prec] && slabs == c->slabs[prec]) return OVL_OK;
    if (c->order[prec]) {
        HIPCHK(c, hipStreamSynchronize(c->st));
        hipFree(c->order[prec]);
        c->order[prec] = nullptr;
    }
    c->trav[prec] = trav;
    c->slabs[prec] = slabs;
    c->order_set[prec] = true;
    std::vector<int> ord;
    if (trav == ovl::TRAV_BOXES) ord = xcd_site_order(c->g.n0, c->g.n1, c->g.n2, slabs);
    else if (trav == ovl::TRAV_MODES && slabs > 1) ord = slab_site_order(c->g.n0, c->g.n1, c->g.n2, slabs);
    if (ord.empty()) return OVL_OK;
    RETCHK(dalloc(c, (void **)&c->order[prec], sizeof(int) * ord.size()));
    HIPCHK(c, hipMemcpy(c->order[prec], ord.data(), sizeof(int) * ord.size(), hipMemcpyHostToDevice));
    return OVL_OK;
}

void begin_timing(ovl_ctx *c)
{
    c->evused = 0;
    c->hop_launches = 0;
    c->hop_bytes = 0;
    for (int v = 0; v < NVAR; ++v) { c->var_launches[v] = 0; c->var_bytes[v] = 0; }
    c->fp64_steps = c->fp32_steps = c->fp16_steps = c->fp48_steps = 0;
    c->captures = 0;
    if (!c->capturing) hipEventRecord(c->t0, c->st);
}

int end_timing(ovl_ctx *c)
{
    HIPCHK(c, hipEventRecord(c->t1, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    float ms = 0;
    hipEventElapsedTime(&ms, c->t0, c->t1);
    c->last.total_ms = ms;
    double hop_ms = 0, var_ms[NVAR] = {};
    for (size_t i = 0; i + 1 < c->evused; i += 2) {
        float t = 0;
        hipEventElapsedTime(&t, c->evpool[i], c->evpool[i + 1]);
        hop_ms += t;
        var_ms[c->evvar[i / 2]] += t;
    }
    c->last.hop_ms = hop_ms;
    for (int v = 0; v < NVAR; ++v) {
        c->last.var_ms[v] = var_ms[v];
        c->last.var_launches[v] = c->var_launches[v];
        c->last.var_bytes[v] = c->var_bytes[v];
    }
    c->last.hop_launches = c->hop_launches;
    c->last.hop_bytes = c->hop_bytes;
    c->last.fp64_steps = c->fp64_steps;
    c->last.fp32_steps = c->fp32_steps;
    c->last.fp16_steps = c->fp16_steps;
    c->last.fp48_steps = c->fp48_steps;
    c->last.graph_captures = c->captures;
    c->last.guard_checks = c->guard_checks;
    c->last.guard_level = c->guard_level;
    c->last.guard_err = c->guard_err;
    return OVL_OK;
}

// ---------------------------------------------------------------------------------------------
// Chebyshev approximation of x^{-1/2} on [cut^2, 1] (FP64, host)
// ---------------------------------------------------------------------------------------------
std::vector<double> cheb_coeffs(int n, double cut)
{
    // c_k = (2/N) sum_j f(x(t_j)) cos(k theta_j) at N Chebyshev nodes, c_0 halved
    const int N = 8 * (n + 1) + 64;
    const double d2 = cut * cut;
    std::vector<long double> acc(n + 1, 0.0L);
    for (int j = 0; j < N; ++j) {
        const double th = M_PI * (j + 0.5) / N;
        const double t = std::cos(th);
        const double x = 0.5 * ((1.0 - d2) * t + (1.0 + d2));
        const long double f = 1.0L / std::sqrt((long double)x);
        for (int k = 0; k <= n; ++k) acc[k] += f * (long double)std::cos(k * th);
    }
    std::vector<double> c(n + 1);
    for (int k = 0; k <= n; ++k) c[k] = (double)(2.0L * acc[k] / N);
    c[0] *= 0.5;
    return c;
}

// B_k = max over t in [t0, 1] of |b_k(t)|: the scalar Clenshaw intermediates of the polynomial
// (they decay geometrically with k).  t0 = -1 gives the bound on the approximation interval;
// t0 = T(0) < -1 covers the whole spectrum of T = alpha h^2 + beta (h^2 >= 0).
std::vector<double> clenshaw_sup(const std::vector<double> &c, double t0)
{
    const int n = (int)c.size() - 1;
    const int NT = 2049;
    std::vector<double> t(NT), b1(NT, 0.0), b2(NT, 0.0), B(n + 2, 0.0);
    for (int i = 0; i < NT; ++i) t[i] = 0.5 * (1.0 + t0) + 0.5 * (1.0 - t0) * std::cos(M_PI * i / (NT - 1));
    for (int k = n; k >= 0; --k) {
        double mx = 0;
        for (int i = 0; i < NT; ++i) {
            const double bk = c[k] + 2.0 * t[i] * b1[i] - b2[i];
            b2[i] = b1[i];
            b1[i] = bk;
            mx = std::max(mx, std::fabs(bk));
        }
        B[k] = mx;
    }
    return B;
}

// Precision ladder of the mixed recurrence: the storage of b_k (k >= first) in a format with unit
// roundoff u perturbs the result by about kappa * u * B_first.  kappa is set per format above the
// largest error measured for that format alone (storage roundings of the whole tail from `first`
// on, host model and GPU, random SU(2)/SU(3) fields and a 6^3 SU(3) lattice with 8 flux quanta, the
// worst case): FP32 0.1-0.7 u32 B and FP16 up to 0.58 u16 B -> kappa 1; 48-bit storage 0.013-0.052
// u48 B (its tier sits at low k) -> kappa 0.1.  The mixed sign's accuracy guard checks the result
// against FP64 on the lattice at hand and tightens the ladder where this model falls short (it does
// when the spectrum of h_W reaches below the cut: B_k is then larger than on [-1, 1]).  The coherent
// worst case (errors aligned with the extreme eigenvector and summed linearly over steps) would ask
// for 1e4x more and is not what rounding produces.  Returns the smallest `first` >= 1 such that
// every step k >= first satisfies the bound (n + 1: none).
int tier_start(const std::vector<double> &B, int n, double u, double tol, double kappa)
{
    int first = n + 1;
    for (int k = n; k >= 1; --k) {
        if (kappa * u * B[k] > tol) break;
        first = k;
    }
    return first;
}

int ensure_work(ovl_ctx *c, int blocks, bool need32, bool need16 = false, bool need48 = false)
{
    if (blocks > c->work_blocks) {
        dfree(c->w64); dfree(c->a64); dfree(c->b64);
        dfree(c->x32); dfree(c->w32); dfree(c->a32); dfree(c->b32);
        dfree(c->x16); dfree(c->w16); dfree(c->a16); dfree(c->b16);
        dfree(c->x48); dfree(c->w48); dfree(c->a48); dfree(c->b48);
        c->work_blocks = blocks;
    }
    const size_t b64 = (size_t)c->work_blocks * c->blk64_elems * 16;
    const size_t b32 = (size_t)c->work_blocks * c->blk32_elems * 16;
    const size_t b16 = (size_t)c->work_blocks * c->blk16_elems * 16;
    const size_t b48 = (size_t)c->work_blocks * c->blk48_elems * 16;
    if (need48) {
        RETCHK(dalloc(c, &c->x48, b48));
        RETCHK(dalloc(c, &c->w48, b48));
        RETCHK(dalloc(c, &c->a48, b48));
        RETCHK(dalloc(c, &c->b48, b48));
    }
    RETCHK(dalloc(c, &c->w64, b64));
    RETCHK(dalloc(c, &c->a64, b64));
    RETCHK(dalloc(c, &c->b64, b64));
    if (need32 || need16) {
        RETCHK(dalloc(c, &c->x32, b32));
        RETCHK(dalloc(c, &c->w32, b32));
        RETCHK(dalloc(c, &c->a32, b32));
        RETCHK(dalloc(c, &c->b32, b32));
    }
    if (need16) {
        RETCHK(dalloc(c, &c->x16, b16));
        RETCHK(dalloc(c, &c->w16, b16));
        RETCHK(dalloc(c, &c->a16, b16));
        RETCHK(dalloc(c, &c->b16, b16));
    }
    return OVL_OK;
}

struct SignPlan {
    std::vector<double> coef;
    // k < head48: FP64, head48 <= k < head: 48-bit, head <= k < head16: FP32, k >= head16: FP16
    int n = 0, head48 = 0, head = 0, head16 = 0;
    double lam = 0, alpha = 0, beta = 0;
    // FP16 tier: b_k is stored as sk[k] * b_k, w = h b_{k+1} as sw[k] * w, x as sx * x (powers of two)
    std::vector<double> sk, sw;
    double sx = 1.0;
};

// error-model factors of the storage tiers (tier_start)
constexpr double KAPPA32 = 1.0, KAPPA16 = 1.0, KAPPA48 = 0.1;

// the precision ladder: k < head48 FP64, head48 <= k < head 48-bit storage, head <= k < head16
// FP32, k >= head16 FP16 storage
// B1: clenshaw_sup(coef, -1) when the caller holds it (null: computed here for the mixed ladder)
void plan_tiers(const ovl_sign_opts *o, const std::vector<double> &coef, int *head48, int *head, int *head16,
                const std::vector<double> *B1 = nullptr)
{
    const int n = (int)coef.size() - 1;
    const double u32 = 5.96e-8, u16 = 4.89e-4, u48 = 7.28e-12;  // 2^-24 (+), 2^-11 (+), 2^-37
    const double tol = o->tol > 0 ? o->tol : 1e-10;
    *head16 = n + 1;
    if (o->precision == OVL_PREC_FP64) {
        *head = n + 1;
        *head48 = n + 1;
    } else if (o->precision == OVL_PREC_FP32) {
        *head = 0;
        *head48 = 0;
    } else {
        const std::vector<double> B = B1 ? *B1 : clenshaw_sup(coef, -1.0);
        const bool fp16 = o->fp16_tail != 0, fp48 = o->fp48 != 0;
        // error budget: 1/4 to each of the FP16 and 48-bit tiers that are on, the rest to FP32
        const double s16 = fp16 ? 0.25 : 0.0, s48 = fp48 ? 0.25 : 0.0, s32 = 1.0 - s16 - s48;
        if (o->head >= 0) *head = std::min(o->head, n + 1);
        else *head = tier_start(B, n, u32, s32 * tol, KAPPA32);
        if (fp16) {
            if (o->head16 > 0) *head16 = std::min(o->head16, n + 1);
            else *head16 = tier_start(B, n, u16, s16 * tol, KAPPA16);
            // at least one FP32 step between the FP16 tail and the FP64 head
            *head16 = std::min(std::max(*head16, *head + 1), n + 1);
        }
        *head48 = *head;
        if (fp48 && *head > 1) {
            const int h = (o->head48 > 0) ? o->head48 : tier_start(B, n, u48, s48 * tol, KAPPA48);
            // the last step (k = 0) stays FP64: its output feeds the final FP64 stencil
            *head48 = std::min(std::max(h, 1), *head);
        }
    }
}

// power of two that maps values bounded by `bound` to at most 2^8 (FP16 keeps 2^15 of headroom
// above that and its subnormal floor 2^-24 far below the unit roundoff of the stored values)
double fp16_scale(double bound)
{
    if (!(bound > 0.0) || !std::isfinite(bound)) return 1.0;
    return std::ldexp(1.0, 8 - (int)std::ceil(std::log2(bound)));
}

// xmax: an upper bound of the 2-norm of every mode of the input (needed for the FP16 scales)
int make_plan(ovl_ctx *c, const ovl_sign_opts *o, SignPlan &p, double xmax = 0.0)
{
    if (o->degree < 1 || o->degree > 100000) return fail(c, OVL_E_ARG, "Chebyshev degree must be in [1, 100000]");
    if (!(o->cut > 0.0 && o->cut < 1.0)) return fail(c, OVL_E_SPECTRUM, "cut must be in (0, 1)");
    p.n = o->degree;
    if (c->memo_n != o->degree || c->memo_cut != o->cut) {
        c->memo_coef = cheb_coeffs(o->degree, o->cut);
        c->memo_B1.clear();
        c->memo_Bb.clear();
        c->memo_n = o->degree;
        c->memo_cut = o->cut;
    }
    p.coef = c->memo_coef;
    p.lam = lambda_bound(c);
    const double d2 = o->cut * o->cut;
    p.alpha = 2.0 / (p.lam * p.lam * (1.0 - d2));
    p.beta = -(1.0 + d2) / (1.0 - d2);
    if (o->precision == OVL_PREC_MIXED && c->memo_B1.empty()) c->memo_B1 = clenshaw_sup(p.coef, -1.0);
    plan_tiers(o, p.coef, &p.head48, &p.head, &p.head16, o->precision == OVL_PREC_MIXED ? &c->memo_B1 : nullptr);
    if (p.head16 <= p.n) {
        if (!(xmax > 0.0) || !std::isfinite(xmax)) return fail(c, OVL_E_ARG, "FP16 tier needs a bound of the input norms");
        if (c->memo_Bb.empty()) c->memo_Bb = clenshaw_sup(p.coef, p.beta);  // beta depends on cut only
        const std::vector<double> &Ball = c->memo_Bb;
        p.sk.assign(p.n + 2, 1.0);
        p.sw.assign(p.n + 2, 1.0);
        for (int k = 0; k <= p.n; ++k) {
            p.sk[k] = fp16_scale(Ball[k] * xmax);
            p.sw[k] = fp16_scale(p.lam * Ball[k + 1] * xmax);
        }
        p.sx = fp16_scale(xmax);
    }
    return OVL_OK;
}

// Clenshaw recurrence for one chunk of mode blocks:  p = P(T) x,  T = alpha h_W^2 + beta.
// Steps k >= head16 store their vectors in FP16 (scaled), head <= k < head16 in FP32,
// head48 <= k < head in the 48-bit layout, k < head48 in FP64.  Returns the buffer and precision
// of p.
int clenshaw_chunk(ovl_ctx *c, const Geom &gc, const SignPlan &pl, const void *x64, void **pout, int *pprec)
{
    const int n = pl.n;
    const bool any16 = pl.head16 <= n;
    const bool any32 = pl.head <= n;
    const bool any48 = pl.head48 < pl.head && pl.head48 <= n;
    RETCHK(ensure_work(c, gc.nblk, any32, any16, any48));
    if (any32) HIPCHK(c, ovl::launch_convert(gc, gc.nblk, x64, P64, c->x32, P32, c->st));
    if (any16) HIPCHK(c, ovl::launch_convert(gc, gc.nblk, x64, P64, c->x16, P16, c->st, pl.sx));
    if (any48) HIPCHK(c, ovl::launch_convert(gc, gc.nblk, x64, P64, c->x48, P48, c->st));
    // buffer rotation: b1 = b_{k+1}, b2 = b_{k+2} (null = zero), stored times sc1 / sc2
    void *b1 = nullptr, *b2 = nullptr;
    double sc1 = 1.0, sc2 = 1.0;
    int prec_b = -1;  // precision of b1/b2
    for (int k = n; k >= 0; --k) {
        const int pk = (k < pl.head48) ? P64 : (k < pl.head ? P48 : (k < pl.head16 ? P32 : P16));
        if (pk == P64) c->fp64_steps++;
        else if (pk == P48) c->fp48_steps++;
        else if (pk == P32) c->fp32_steps++;
        else c->fp16_steps++;
        auto pick = [pk](void *p64, void *p48, void *p32, void *p16) {
            return pk == P64 ? p64 : (pk == P48 ? p48 : (pk == P32 ? p32 : p16));
        };
        void *xk = pick((void *)x64, c->x48, c->x32, c->x16);
        void *A = pick(c->a64, c->a48, c->a32, c->a16);
        void *B = pick(c->b64, c->b48, c->b32, c->b16);
        void *W = pick(c->w64, c->w48, c->w32, c->w16);
        const double sk = (pk == P16) ? pl.sk[k] : 1.0, sw = (pk == P16) ? pl.sw[k] : 1.0;
        const double sx = (pk == P16) ? pl.sx : 1.0;
        if (b1 && prec_b != pk) {
            // FP16 -> FP32 -> (48-bit ->) FP64 transitions: promote b_{k+1}, b_{k+2} (and drop the
            // FP16 scale)
            HIPCHK(c, ovl::launch_convert(gc, gc.nblk, b1, prec_b, A, pk, c->st, 1.0 / sc1));
            if (b2) HIPCHK(c, ovl::launch_convert(gc, gc.nblk, b2, prec_b, B, pk, c->st, 1.0 / sc2));
            b1 = A;
            b2 = b2 ? B : nullptr;
            sc1 = sc2 = 1.0;
            prec_b = pk;
        }
        if (!b1) {  // b_n = c_n x
            if (pk == P16 || pk == P48)
                HIPCHK(c, ovl::launch_convert(gc, gc.nblk, x64, P64, A, pk, c->st, pl.coef[k] * sk));
            else
                HIPCHK(c, ovl::launch_lincomb(gc, gc.nblk, pk, A, xk, pl.coef[k], 0.0, nullptr, 0, 0, c->st));
            b1 = A;
            sc1 = sk;
            prec_b = pk;
            continue;
        }
        // w = h_W b_{k+1}
        HopArgs a = hop_args(c, pk);
        a.nblk = gc.nblk;
        a.in = b1;
        a.out = W;
        a.sh_r = sw / sc1;
        RETCHK(hop(c, gc, pk, pk, a));
        // b_k = c_k x + 2 alpha h_W w + 2 beta b_{k+1} - b_{k+2}    (k >= 1)
        // p   = c_0 x +   alpha h_W w +   beta b_1     - b_2        (k == 0)
        const double f = (k == 0) ? 1.0 : 2.0;
        void *out = b2 ? b2 : ((b1 == A) ? B : A);
        HopArgs q = hop_args(c, pk);
        q.nblk = gc.nblk;
        q.in = W;
        q.out = out;
        q.sh_r = f * pl.alpha * sk / sw;
        q.x = xk;
        q.sx_r = pl.coef[k] * sk / sx;
        q.b1 = b1;
        q.s1_r = f * pl.beta * sk / sc1;
        q.b2 = b2;
        q.s2_r = -sk / sc2;
        RETCHK(hop(c, gc, pk, pk, q));
        b2 = b1;
        sc2 = sc1;
        b1 = out;
        sc1 = sk;
    }
    *pout = b1;
    *pprec = prec_b;
    return OVL_OK;
}

// result = sh * h_W(p) + sx * G(x) + st * tgt for every chunk; kind selects the combination
enum FinalKind { FIN_SIGN = 0, FIN_OVERLAP = 1, FIN_OVERLAP_AXPY = 2 };

// xmax: bound of the 2-norm of every input mode (used by the FP16 tier only)
// hpsi (FIN_OVERLAP_AXPY only, optional): also store -i h_ov in (the term added to out times dt)
// blocks [blk0, blk0 + nblk) of the fields (nblk < 0: all from blk0 on); out_rel: out64 is a buffer
// whose first block receives block blk0
int sign_driver(ovl_ctx *c, const ovl_sign_opts *o, const void *in64, void *out64, int kind, double dt,
                double xmax, void *hpsi = nullptr, int blk0 = 0, int nblk = -1, bool out_rel = false)
{
    RETCHK(ensure_stencil(c));
    SignPlan pl;
    RETCHK(make_plan(c, o, pl, xmax));
    const int bend = (nblk < 0) ? c->g.nblk : std::min(c->g.nblk, blk0 + nblk);
    const int chunk = (o->mode_chunk > 0) ? std::min(o->mode_chunk, bend - blk0) : bend - blk0;
    const double M = c->prm.m_dw, m = c->prm.m_fermion;
    for (int b0 = blk0; b0 < bend; b0 += chunk) {
        Geom gc = c->g;
        gc.nblk = std::min(chunk, bend - b0);
        const void *x = blk_ptr((void *)in64, P64, b0, c);
        void *y = blk_ptr(out64, P64, out_rel ? b0 - blk0 : b0, c);
        void *p = nullptr;
        int pp = P64;
        RETCHK(clenshaw_chunk(c, gc, pl, x, &p, &pp));
        HopArgs a = hop_args(c, pp);
        a.nblk = gc.nblk;
        a.in = p;
        a.out = y;
        if (kind == FIN_SIGN) {
            a.sh_r = 1.0 / pl.lam;
        } else {
            // -i h_ov = -i M [ (1 + m/2) gamma0 + (1 - m/2) eps ]   (OverlapU.cpp:188-201)
            const double s = (kind == FIN_OVERLAP_AXPY) ? dt : 1.0;
            a.sh_i = -s * M * (1.0 - 0.5 * m) / pl.lam;
            a.x = x;
            a.x_gamma0 = 1;
            a.sx_i = -s * M * (1.0 + 0.5 * m);
            if (kind == FIN_OVERLAP_AXPY) {
                a.tgt = y;
                a.st_r = 1.0;
                if (hpsi) {
                    a.out2 = blk_ptr(hpsi, P64, b0, c);
                    a.out2_scale = 1.0 / dt;
                }
            }
        }
        RETCHK(hop(c, gc, pp, P64, a));
    }
    return OVL_OK;
}

// ---------------------------------------------------------------------------------------------
// Zolotarev rational approximation and multi-shift CG (reference algorithm, FP64)
// ---------------------------------------------------------------------------------------------

// CompElliptic (ZolotarevCoefficients.cpp:5-49): complete elliptic integral by Carlson's R_F
double elliptic_k(double kappa, double tol)
{
    double X = 0.0, Y = 1.0 - kappa, Z = 1.0, mean = (X + Y + Z) / 3.0;
    double dx = 2.0 - (mean + X) / mean, dy = 2.0 - (mean + Y) / mean, dz = 2.0 - (mean + Z) / mean;
    for (double dev = tol; dev >= tol;) {
        mean = (X + Y + Z) / 3.0;
        dx = 2.0 - (mean + X) / mean;
        dy = 2.0 - (mean + Y) / mean;
        dz = 2.0 - (mean + Z) / mean;
        dev = std::max(std::fabs(dx), std::max(std::fabs(dy), std::fabs(dz)));
        const double sx = std::sqrt(X), sy = std::sqrt(Y), sz = std::sqrt(Z);
        const double lam = sx * (sy + sz) + sy * sz;
        X = 0.25 * (X + lam);
        Y = 0.25 * (Y + lam);
        Z = 0.25 * (Z + lam);
    }
    const double e2 = dx * dy - dz * dz, e3 = dx * dy * dz;
    return (1.0 + (e2 / 24.0 - 0.1 - e3 / 44.0) * e2 + e3 / 14.0) / std::sqrt(mean);
}

}  // namespace

extern "C" double ovl_zolotarev_coefficients(double amin, double amax, double errtol, int no, double *bz, double *dz)
{
    // ZolotarevCoefficients.cpp:71-175.  The single-precision square roots of the reference
    // (:97, :106, :148) and its fmin start value 1.0 (:136) are kept so that the coefficients
    // are the reference's bit for bit.
    if (no < 1 || !(amin > 0) || !(amax > amin) || !bz || !dz) return NAN;
    const double kappa = std::sqrt(1.0 - amin / amax);
    const double K = elliptic_k(kappa, errtol);
    const double step = 0.5 * K / no;
    std::vector<double> c(2 * no), al(4000), ga(4000);
    for (int l = 0; l < 2 * no - 1; ++l) {
        const double u = (l + 1) * step;
        // descending Landen / AGM for the Jacobi amplitude, then tan^2
        al[0] = 1.0;
        double b = (double)sqrtf((float)(1.0 - kappa));
        ga[0] = std::sqrt(kappa);
        int i = 1;
        double p2 = 1.0;
        while (std::fabs(ga[i - 1] / al[i - 1]) >= errtol) {
            const double bprev = b;
            al[i] = 0.5 * (al[i - 1] + bprev);
            b = (double)sqrtf((float)(al[i - 1] * bprev));
            ga[i] = 0.5 * (al[i - 1] - bprev);
            ++i;
            p2 *= 2.0;
        }
        --i;
        double phi = p2 * al[i] * u;
        for (; i >= 1; --i) phi = 0.5 * (phi + std::asin(ga[i] * std::sin(phi) / al[i]));
        c[l] = std::tan(phi) * std::tan(phi);
    }
    for (int l = 0; l < no; ++l) {
        dz[l] = c[2 * l] * amin;
        // coeffb (ZolotarevCoefficients.cpp:51-69)
        double prod = 1.0;
        int j = 0;
        for (int i = 0; i < no - 1; ++i) {
            prod *= (c[2 * i + 1] - c[2 * l]);
            if (i == l) ++j;
            prod /= (c[2 * j] - c[2 * l]);
            ++j;
        }
        bz[l] = prod;
    }
    const int np = 5000;
    const double dxl = (std::log(amax) - std::log(amin)) / np;
    double y = std::log(amin) - dxl, fmin = std::pow(1.0, 10.0), fmax = 0.0;
    for (int i = 0; i < np; ++i) {
        y += dxl;
        const double x = std::exp(y);
        double z = 0.0;
        for (int l = 0; l < no; ++l) z += bz[l] / (x + dz[l]);
        z *= (double)sqrtf((float)x);
        fmax = std::max(fmax, z);
        fmin = std::min(fmin, z);
    }
    const double anorm = 2.0 / (fmin + fmax);
    for (int k = 0; k < no; ++k) bz[k] *= anorm;
    return (fmax - fmin) / (fmax + fmin);
}

namespace {

int mode_dot(ovl_ctx *c, const Geom &g, const void *a, const void *b, std::vector<cplx> &out)
{
    HIPCHK(c, ovl::launch_mode_dot(g, a, b, c->partial, c->dout + 16, c->st));
    const int nm = g.nblk * g.mb;
    out.resize(nm);
    HIPCHK(c, hipMemcpyAsync(out.data(), c->dout + 16, sizeof(cplx) * nm, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return OVL_OK;
}

// per-mode <a,b> of every system (system-major, n_modes per system)
int mode_dot_all(ovl_ctx *c, const void *a, const void *b, std::vector<cplx> &out)
{
    out.clear();
    std::vector<cplx> d;
    for (int s = 0; s < nsys(c); ++s) {
        RETCHK(mode_dot(c, sys_geom(c, s), sys_field(c, a, s), sys_field(c, b, s), d));
        out.insert(out.end(), d.begin(), d.begin() + c->prm.n_modes);
    }
    return OVL_OK;
}

// max over all modes of all systems of the 2-norm of an FP64 field (0 when the FP16 tier is off)
int input_bound(ovl_ctx *c, const ovl_sign_opts *o, const void *f, double *xmax)
{
    *xmax = 0.0;
    if (o->method != OVL_SIGN_CHEBYSHEV || o->precision != OVL_PREC_MIXED || !o->fp16_tail) return OVL_OK;
    std::vector<cplx> d;
    RETCHK(mode_dot_all(c, f, f, d));
    double m = 0;
    for (const cplx &v : d) m = std::max(m, v.real());
    *xmax = std::sqrt(m);
    if (!(*xmax > 0.0)) *xmax = 1.0;  // all-zero input: any scale works
    return OVL_OK;
}

// ---- a-posteriori accuracy guard of the mixed sign ----
bool guarded(const ovl_sign_opts *o)
{
    return o->method == OVL_SIGN_CHEBYSHEV && o->precision == OVL_PREC_MIXED && o->guard > 0;
}

// a check is due when the links or the options that set the ladder changed since the last check, or
// at least `guard` steps (applies) went by
bool guard_due(const ovl_ctx *c, const ovl_sign_opts *o)
{
    if (!guarded(o)) return false;
    const ovl_sign_opts &g = c->guard_opts;
    const bool same = g.degree == o->degree && g.cut == o->cut && g.tol == o->tol && g.head == o->head &&
                      g.fp16_tail == o->fp16_tail && g.head16 == o->head16 && g.fp48 == o->fp48 &&
                      g.head48 == o->head48 && g.precision == o->precision && g.method == o->method &&
                      g.relax == o->relax;
    return c->guard_fresh || !same || c->guard_since >= o->guard;
}

// the options a guarded mixed sign runs with at the context's level: level 0 as requested; level >= 1
// without head overrides (the ladder follows tol); level l >= 2 planned for tol / 16^(l - 1) (each
// tier starts about 26 steps earlier at the geometric decay of the Clenshaw intermediates); the
// last level is plain FP64.  Negative levels (only with o->relax) loosen: level -l is planned for
// tol 4^l (each tier about 13 steps later).
constexpr int GUARD_LEVELS = 6;
constexpr double RELAX_KEEP = 0.5;  // a loosened ladder is kept while the check stays within tol / 2
ovl_sign_opts guarded_opts(const ovl_ctx *c, const ovl_sign_opts *o)
{
    ovl_sign_opts e = *o;
    if (!guarded(o) || c->guard_level == 0) return e;
    e.head = -1;
    e.head16 = 0;
    e.head48 = 0;
    const double tol = o->tol > 0 ? o->tol : 1e-10;
    if (c->guard_level < 0) e.tol = tol * std::pow(4.0, -(double)c->guard_level);
    else e.tol = tol * std::pow(16.0, -(double)(c->guard_level - 1));
    if (c->guard_level >= GUARD_LEVELS) e.precision = OVL_PREC_FP64;
    return e;
}

// One mode block of in64 (the next in turn) through the mixed sign and through the FP64 sign: while
// some mode's relative difference exceeds tol, the ladder is tightened a level and checked again.
// *eff receives the options to run with.  The check's kernels stay out of the call's stencil
// counters and events; its time is part of the call.  A non-finite FP64 result is an error (the
// field itself has overflowed); a non-finite mixed one (FP16 range) is a failed check.
int guard_check(ovl_ctx *c, const ovl_sign_opts *o, const void *in64, double xmax, ovl_sign_opts *eff)
{
    *eff = guarded_opts(c, o);
    if (!guarded(o)) return OVL_OK;
    const int blk = c->guard_next++ % c->g.nblk;
    const bool multi = nsys(c) > 1;  // an ensemble: its one flat block holds every system
    // the check's two buffers, grown on demand (one mode block, or all of them for a loosening try)
    auto ensure_buffers = [&](size_t bytes) -> int {
        if (c->guard_bytes >= bytes) return OVL_OK;
        HIPCHK(c, hipStreamSynchronize(c->st));
        for (int f = 0; f < 2; ++f) {
            if (c->guard_buf[f]) HIPCHK(c, hipFree(c->guard_buf[f]));
            c->guard_buf[f] = nullptr;
        }
        c->guard_bytes = 0;
        for (int f = 0; f < 2; ++f) RETCHK(dalloc(c, &c->guard_buf[f], bytes));
        c->guard_bytes = bytes;
        return OVL_OK;
    };
    RETCHK(ensure_buffers(c->blk64_elems * 16));
    const long long hl = c->hop_launches;
    const double hb = c->hop_bytes;
    long long vl[NVAR];
    double vb[NVAR];
    for (int v = 0; v < NVAR; ++v) { vl[v] = c->var_launches[v]; vb[v] = c->var_bytes[v]; }
    const int st[4] = {c->fp64_steps, c->fp32_steps, c->fp16_steps, c->fp48_steps};
    const bool prof = c->profiling;
    c->profiling = false;
    const double tol = o->tol > 0 ? o->tol : 1e-10;
    int r = OVL_OK;
    // loosening (o->relax > 0): a level that passes with 4x its error still within tol / 2 is tried
    // one level looser, on every mode block (a looser ladder must hold for all modes, not only for
    // the block in turn); it is kept if that check stays within tol / 2, else the context returns to
    // the level that passed.  Errors are per mode, and in an ensemble per mode of each system.
    bool trying = false;
    int passed_level = c->guard_level;
    double passed_err = 0.0;
    for (;;) {
        *eff = guarded_opts(c, o);
        if (eff->precision != OVL_PREC_MIXED) break;  // plain FP64: nothing left to compare
        ovl_sign_opts o64 = *eff;
        o64.precision = OVL_PREC_FP64;
        const bool all = trying && c->g.nblk > 1;
        const int b0 = all ? 0 : blk, nb = all ? c->g.nblk : 1;
        Geom gk = c->g;
        gk.nblk = nb;
        r = ensure_buffers(c->blk64_elems * 16 * (size_t)nb);
        if (r == OVL_OK)
            r = sign_driver(c, eff, in64, c->guard_buf[0], FIN_SIGN, 0.0, xmax, nullptr, b0, nb, true);
        if (r == OVL_OK) r = sign_driver(c, &o64, in64, c->guard_buf[1], FIN_SIGN, 0.0, xmax, nullptr, b0, nb, true);
        if (r == OVL_OK) {
            const hipError_t e = ovl::launch_lincomb(gk, nb, P64, c->guard_buf[0], c->guard_buf[0], 1.0, 0.0,
                                                     c->guard_buf[1], -1.0, 0.0, c->st);
            if (e != hipSuccess) r = fail(c, OVL_E_HIP, std::string("guard: ") + hipGetErrorString(e));
        }
        std::vector<cplx> dd, yy;
        if (multi) {
            if (r == OVL_OK) r = mode_dot_all(c, c->guard_buf[0], c->guard_buf[0], dd);
            if (r == OVL_OK) r = mode_dot_all(c, c->guard_buf[1], c->guard_buf[1], yy);
        } else {
            if (r == OVL_OK) r = mode_dot(c, gk, c->guard_buf[0], c->guard_buf[0], dd);
            if (r == OVL_OK) r = mode_dot(c, gk, c->guard_buf[1], c->guard_buf[1], yy);
        }
        if (r != OVL_OK) break;
        double worst = 0.0;
        bool ok = true;
        for (size_t i = 0; i < dd.size() && r == OVL_OK; ++i) {
            if (!std::isfinite(yy[i].real())) r = fail(c, OVL_E_ARG, "non-finite mode function (the leapfrog diverged)");
            else if (!std::isfinite(dd[i].real())) ok = false;
            else if (yy[i].real() > 0) worst = std::max(worst, std::sqrt(dd[i].real() / yy[i].real()));
        }
        if (r != OVL_OK) break;
        c->guard_checks++;
        c->guard_err = ok ? worst : INFINITY;
        if (trying && !(ok && worst <= RELAX_KEEP * tol)) {  // the looser ladder misses: back
            c->guard_level = passed_level;
            c->guard_err = passed_err;
            *eff = guarded_opts(c, o);
            break;
        }
        if (ok && worst <= (trying ? RELAX_KEEP : 1.0) * tol) {
            if (o->relax > 0 && c->guard_level <= 0 && c->guard_level > -o->relax && 4.0 * worst <= RELAX_KEEP * tol) {
                trying = true;
                passed_level = c->guard_level;
                passed_err = worst;
                c->guard_level--;
                continue;
            }
            break;
        }
        c->guard_level++;
        // level 1 only drops head overrides: without any it is level 0 again
        if (c->guard_level == 1 && o->head < 0 && o->head16 <= 0 && o->head48 <= 0) c->guard_level = 2;
    }
    c->profiling = prof;
    if (r == OVL_OK) {
        c->guard_fresh = false;
        c->guard_since = 0;
        c->guard_opts = *o;
    }
    c->hop_launches = hl;
    c->hop_bytes = hb;
    for (int v = 0; v < NVAR; ++v) { c->var_launches[v] = vl[v]; c->var_bytes[v] = vb[v]; }
    c->fp64_steps = st[0];
    c->fp32_steps = st[1];
    c->fp16_steps = st[2];
    c->fp48_steps = st[3];
    return r;
}

// (h_W^2 + d0) v  ->  q   via two stencil launches, FP64
int apply_h2(ovl_ctx *c, const Geom &g, const void *v, void *w, void *q, double d0)
{
    HopArgs a = hop_args(c, P64);
    a.nblk = g.nblk;
    a.in = v;
    a.out = w;
    a.sh_r = 1.0;
    RETCHK(hop(c, g, P64, P64, a));
    HopArgs b = hop_args(c, P64);
    b.nblk = g.nblk;
    b.in = w;
    b.out = q;
    b.sh_r = 1.0;
    if (d0 != 0.0) {
        b.x = v;
        b.sx_r = d0;
    }
    return hop(c, g, P64, P64, b);
}

// Zolotrr (OverlapU.cpp:5-149) for all modes of `in`: chi = sum_k bz_k (h^2 + dz_k)^{-1} in.
// Per-mode scalars live on the host; the chi accumulation is folded into the shift update.
int mscg(ovl_ctx *c, const Geom &g, const void *in, void *chi, const double *bz, const double *dz, int no,
         double cgerr, int nitcgm, int *max_iter, int *nonconv)
{
    const size_t fb = (size_t)g.nblk * c->blk64_elems * 16;      // stride of one shift vector for g
    const size_t fmax = (size_t)c->g.nblk * c->blk64_elems * 16;  // allocation: the whole context
    if (c->ms_no < no) {
        dfree(c->ms_ps);
        c->ms_no = no;
    }
    RETCHK(dalloc(c, &c->ms_ps, fmax * no));
    RETCHK(dalloc(c, &c->ms_r, fmax));
    RETCHK(dalloc(c, &c->ms_q, fmax));
    RETCHK(dalloc(c, &c->ms_w, fmax));
    const int nm = g.nblk * g.mb;
    RETCHK(dalloc(c, (void **)&c->ms_coef, sizeof(double) * (size_t)c->g.nblk * c->g.mb * 3 * 32));
    RETCHK(dalloc(c, (void **)&c->ms_active, sizeof(int) * (size_t)c->g.nblk * c->g.mb));
    HIPCHK(c, ovl::launch_lincomb(g, g.nblk, P64, c->ms_r, in, 1, 0, nullptr, 0, 0, c->st));
    for (int k = 0; k < no; ++k)
        HIPCHK(c, ovl::launch_lincomb(g, g.nblk, P64, (char *)c->ms_ps + k * fb, in, 1, 0, nullptr, 0, 0, c->st));
    HIPCHK(c, ovl::launch_lincomb(g, g.nblk, P64, chi, in, 0, 0, nullptr, 0, 0, c->st));
    std::vector<cplx> d;
    RETCHK(mode_dot(c, g, c->ms_r, c->ms_r, d));
    std::vector<double> rn(nm), bn(nm), betam(nm, 1.0), a0(nm, 0.0);
    std::vector<double> z((size_t)nm * no, 1.0), gc((size_t)nm * no, 1.0), be((size_t)nm * no, 0.0), al((size_t)nm * no, 0.0);
    std::vector<int> act(nm), it(nm, 0);
    std::vector<double> coefA(nm), coef((size_t)nm * 3 * no, 0.0);
    for (int i = 0; i < nm; ++i) {
        rn[i] = d[i].real();
        bn[i] = std::sqrt(rn[i]);
        act[i] = rn[i] > cgerr * bn[i];
    }
    *max_iter = 0;
    *nonconv = 0;
    while (std::any_of(act.begin(), act.end(), [](int v) { return v != 0; })) {
        HIPCHK(c, hipMemcpyAsync(c->ms_active, act.data(), sizeof(int) * nm, hipMemcpyHostToDevice, c->st));
        RETCHK(apply_h2(c, g, c->ms_ps, c->ms_w, c->ms_q, dz[0]));
        std::vector<cplx> pq;
        RETCHK(mode_dot(c, g, c->ms_ps, c->ms_q, pq));
        for (int i = 0; i < nm; ++i) coefA[i] = act[i] ? rn[i] / pq[i].real() : 0.0;
        HIPCHK(c, hipMemcpyAsync(c->ms_coef, coefA.data(), sizeof(double) * nm, hipMemcpyHostToDevice, c->st));
        HIPCHK(c, ovl::launch_axpy_modes(g, c->ms_r, c->ms_q, c->ms_coef, 1, -1.0, c->ms_active, c->st));
        std::vector<cplx> rr;
        RETCHK(mode_dot(c, g, c->ms_r, c->ms_r, rr));
        std::vector<int> nact(nm, 0);
        for (int i = 0; i < nm; ++i) {
            if (!act[i]) continue;
            double *zi = &z[(size_t)i * no], *bi = &be[(size_t)i * no], *ai = &al[(size_t)i * no], *gi = &gc[(size_t)i * no];
            bi[0] = coefA[i];
            for (int k = 1; k < no; ++k) {
                const double dr = bi[0] * a0[i] * (1.0 - zi[k]) + betam[i] * (1.0 + (dz[k] - dz[0]) * bi[0]);
                zi[k] = betam[i] / dr;
                bi[k] = bi[0] * zi[k];
            }
            a0[i] = rr[i].real() / rn[i];
            rn[i] = rr[i].real();
            for (int k = 0; k < no; ++k) {
                ai[k] = a0[i] * zi[k] * bi[k] / bi[0];
                gi[k] = gi[k] * zi[k];
                coef[((size_t)i * no + k) * 3 + 0] = bz[k] * bi[k];
                coef[((size_t)i * no + k) * 3 + 1] = ai[k];
                coef[((size_t)i * no + k) * 3 + 2] = gi[k];
            }
            betam[i] = bi[0];
            it[i]++;
            nact[i] = (rn[i] > cgerr * bn[i]) && it[i] < nitcgm;
            if (it[i] >= nitcgm && rn[i] > cgerr * bn[i]) (*nonconv)++;
            *max_iter = std::max(*max_iter, it[i]);
        }
        HIPCHK(c, hipMemcpyAsync(c->ms_coef, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice, c->st));
        HIPCHK(c, ovl::launch_mscg_update(g, no, (double *)c->ms_ps, (double *)chi, (const double *)c->ms_r, c->ms_coef,
                                          c->ms_active, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        act = nact;
    }
    return OVL_OK;
}

int ritz_bounds(ovl_ctx *c, double *lmin, double *lmax);

// -i h_ov in -> out (or out += dt (-i h_ov in)) with the Zolotarev sign, all modes of g
int zolo_overlap(ovl_ctx *c, const Geom &g, const void *in, void *out, double lmin, double lmax, int axpy, double dt)
{
    if (lmin == 0.0 && lmax == 0.0) RETCHK(ritz_bounds(c, &lmin, &lmax));
    if (!(lmin > 0.0) || !(lmax > lmin) || !std::isfinite(lmax))
        return fail(c, OVL_E_SPECTRUM, "Zolotarev interval needs 0 < lmin < lmax");
    RETCHK(ensure_stencil(c));
    const int no = 25;
    double bz[25], dz[25];
    ovl_zolotarev_coefficients(lmin, lmax, 1e-12, no, bz, dz);
    const size_t fb = (size_t)g.nblk * c->blk64_elems * 16;
    RETCHK(dalloc(c, &c->ms_chi, (size_t)c->g.nblk * c->blk64_elems * 16));
    int mx = 0, nc = 0;
    RETCHK(mscg(c, g, in, c->ms_chi, bz, dz, no, 1e-16, 1000, &mx, &nc));
    (void)fb;
    // psi = -i h_W chi, combined as OverlapU.cpp:188-201
    const double M = c->prm.m_dw, m = c->prm.m_fermion, s = axpy ? dt : 1.0;
    HopArgs a = hop_args(c, P64);
    a.nblk = g.nblk;
    a.in = c->ms_chi;
    a.out = out;
    a.sh_i = -s * M * (1.0 - 0.5 * m);
    a.x = in;
    a.x_gamma0 = 1;
    a.sx_i = -s * M * (1.0 + 0.5 * m);
    if (axpy) {
        a.tgt = out;
        a.st_r = 1.0;
    }
    RETCHK(hop(c, g, P64, P64, a));
    if (nc) return fail(c, OVL_E_CONVERGE, "multi-shift CG reached 1000 iterations for " + std::to_string(nc) + " mode(s)");
    return OVL_OK;
}

}  // namespace

// =============================================================================================
// C ABI
// =============================================================================================
extern "C" {

void ovl_sign_defaults(ovl_sign_opts *o)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->method = OVL_SIGN_CHEBYSHEV;
    o->precision = OVL_PREC_FP64;
    o->degree = 256;
    o->cut = 0.05;
    o->tol = 1e-10;
    o->head = -1;
    o->fp16_tail = 1;
    o->head16 = 0;
    o->fp48 = 1;
    o->head48 = 0;
    o->guard = 100;
}

int ovl_sign_plan(const ovl_sign_opts *o, int *head48, int *head, int *head16)
{
    if (!o || !head48 || !head || !head16) return OVL_E_ARG;
    if (o->method != OVL_SIGN_CHEBYSHEV || o->degree < 1 || o->degree > 100000 || !(o->cut > 0.0 && o->cut < 1.0))
        return OVL_E_ARG;
    plan_tiers(o, cheb_coeffs(o->degree, o->cut), head48, head, head16);
    return OVL_OK;
}

int ovl_device_name(char *buf, int len)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return OVL_E_HIP;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, 0) != hipSuccess) return OVL_E_HIP;
    if (buf && len > 0) std::snprintf(buf, len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return OVL_OK;
}

// Allocations and device tables common to ordinary and ensemble contexts; c->g, c->sys and
// c->link_sites are filled in.  On failure the caller destroys c.
static int ctx_init(ovl_ctx *c)
{
    const int mb = c->g.mb, nc = c->g.nc;
    c->ncomp = 4 * nc;
    c->blk64_elems = (size_t)c->g.V * c->ncomp * mb;
    c->blk32_elems = (size_t)c->g.V * (c->ncomp / 2) * mb;
    c->blk16_elems = (size_t)c->g.V * (c->ncomp / 4) * mb;
    c->blk48_elems = 3 * c->blk16_elems;
    c->field64_bytes = c->blk64_elems * 16 * c->g.nblk;
    c->field32_bytes = c->blk32_elems * 16 * c->g.nblk;
    c->sys_nb.assign(c->sys.size(), 0);
    c->sys_links_set.assign(c->sys.size(), 0);
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) return OVL_E_HIP;
    if (hipEventCreate(&c->t0) != hipSuccess || hipEventCreate(&c->t1) != hipSuccess) return OVL_E_HIP;
    const size_t L = (size_t)c->link_sites;
    const size_t esz = sizeof(double) * n_adj(c) * 3 * L;
    RETCHK(dalloc(c, (void **)&c->links, sizeof(double) * link_reals(c) * 3 * L));
    RETCHK(dalloc(c, (void **)&c->efield, esz));
    RETCHK(dalloc(c, (void **)&c->ul64, 16 * 3 * L * nc * nc));
    RETCHK(dalloc(c, (void **)&c->ul32, 8 * 3 * L * nc * nc));
    for (int f = 0; f < 3; ++f) RETCHK(dalloc(c, &c->field[f], c->field64_bytes));
    for (int f = 0; f < 3; ++f) HIPCHK(c, hipMemsetAsync(c->field[f], 0, c->field64_bytes, c->st));
    HIPCHK(c, hipMemsetAsync(c->efield, 0, esz, c->st));
    size_t plen = std::max(std::max(ovl::mode_dot_partial_len(c->g), ovl::reduce_partial_len(c->g)),
                           std::max(ovl::currents_partial_len(c->g), ovl::eig_partial_len()));
    int maxblk = c->g.nblk;
    for (int s = 0; s < nsys(c); ++s) {
        const Geom gs = sys_geom(c, s);
        plen = std::max(plen, std::max(std::max(ovl::mode_dot_partial_len(gs), ovl::reduce_partial_len(gs)),
                                       ovl::currents_partial_len(gs)));
        maxblk = std::max(maxblk, gs.nblk);
    }
    c->partial_len = plen;
    RETCHK(dalloc(c, (void **)&c->partial, sizeof(double) * c->partial_len));
    RETCHK(dalloc(c, (void **)&c->dout, sizeof(double) * (16 + 2 * (size_t)maxblk * mb)));
    if (!c->ensemble && (c->prm.traversal == 0 || c->prm.traversal == 2))
        c->boxes_ok = !xcd_site_order(c->g.n0, c->g.n1, c->g.n2).empty();
    if (!c->ensemble) {
        // traversal 0 (automatic): mode split where the mode blocks divide over the XCDs, else
        // boxes; ovl_autotune may change it
        for (int p = 0; p < NUM_PREC; ++p) {
            int t = ovl::TRAV_LEX;
            if (c->prm.traversal == 0) t = (c->g.nblk % ovl::NUM_XCD == 0) ? ovl::TRAV_MODES : ovl::TRAV_BOXES;
            if (c->prm.traversal == 2) t = ovl::TRAV_BOXES;
            if (c->prm.traversal == 3) t = ovl::TRAV_MODES;
            RETCHK(set_traversal(c, p, t, t == ovl::TRAV_MODES ? default_slabs(c, p) : 1));
        }
    }
    if (c->ensemble) {
        // Work table of the ensemble stencil: the hardware hands workgroup b to XCD b % 8, so
        // placing a system's workgroups at b = k + 8 j keeps the whole system (and the L2 reuse
        // of its neighbour loads) on XCD k.  Systems go largest first to the least loaded XCD;
        // shorter XCD lists are padded with idle entries.
        const int ns = nsys(c);
        std::vector<long> wg(ns);
        std::vector<int> idx(ns);
        for (int s = 0; s < ns; ++s) {
            const SysDesc &d = c->sys[s];
            wg[s] = ((long)d.V * d.nblk * mb + ovl::STENCIL_TPB - 1) / ovl::STENCIL_TPB;
            idx[s] = s;
        }
        std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return wg[x] > wg[y]; });
        std::vector<std::vector<int2>> bins(ovl::NUM_XCD);
        std::vector<long> load(ovl::NUM_XCD, 0);
        for (int s : idx) {
            const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            for (long j = 0; j < wg[s]; ++j) bins[k].push_back(make_int2(s, (int)j));
            load[k] += wg[s];
        }
        const long lmax = *std::max_element(load.begin(), load.end());
        long ltot = 0;
        for (long l : load) ltot += l;
        if ((long)ovl::NUM_XCD * lmax > (1L << 30)) return fail(c, OVL_E_ARG, "ensemble too large for one launch");
        std::vector<int2> tab((size_t)ovl::NUM_XCD * lmax, make_int2(-1, 0));
        for (int k = 0; k < ovl::NUM_XCD; ++k)
            for (size_t j = 0; j < bins[k].size(); ++j) tab[(size_t)k + ovl::NUM_XCD * j] = bins[k][j];
        c->ens_grid = (int)tab.size();
        c->ens_imbalance = (double)lmax * ovl::NUM_XCD / (double)ltot;
        RETCHK(dalloc(c, (void **)&c->d_work, sizeof(int2) * tab.size()));
        RETCHK(dalloc(c, (void **)&c->d_sys, sizeof(SysDesc) * c->sys.size()));
        for (SysDesc &d : c->sys) {
            d.dv_V = ovl::make_fastdiv((unsigned)d.V);
            d.dv_n01 = ovl::make_fastdiv((unsigned)(d.n0 * d.n1));
            d.dv_n0 = ovl::make_fastdiv((unsigned)d.n0);
        }
        HIPCHK(c, hipMemcpy(c->d_work, tab.data(), sizeof(int2) * tab.size(), hipMemcpyHostToDevice));
        // the work queues: the same tiles, XCD k's in one run
        std::vector<int2> qw;
        std::vector<int> qoff(ovl::NUM_XCD + 1, 0);
        for (int k = 0; k < ovl::NUM_XCD; ++k) {
            qoff[k] = (int)qw.size();
            qw.insert(qw.end(), bins[k].begin(), bins[k].end());
        }
        qoff[ovl::NUM_XCD] = (int)qw.size();
        RETCHK(dalloc(c, (void **)&c->d_qwork, sizeof(int2) * std::max<size_t>(1, qw.size())));
        RETCHK(dalloc(c, (void **)&c->d_qoff, sizeof(int) * qoff.size()));
        RETCHK(dalloc(c, (void **)&c->d_qctr, sizeof(int) * 2 * ovl::NUM_XCD));
        HIPCHK(c, hipMemcpy(c->d_qwork, qw.data(), sizeof(int2) * qw.size(), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->d_qoff, qoff.data(), sizeof(int) * qoff.size(), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemset(c->d_qctr, 0, sizeof(int) * 2 * ovl::NUM_XCD));
        HIPCHK(c, hipMemcpy(c->d_sys, c->sys.data(), sizeof(SysDesc) * c->sys.size(), hipMemcpyHostToDevice));
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    return OVL_OK;
}

static int pick_mode_block(int mode_block, int n_modes)
{
    if (mode_block > 0) return mode_block;
    return n_modes >= 64 ? 64 : (n_modes > 16 ? 32 : 16);
}

// mode_block = -1: each block size (16, 32, 64 modes; not one whose padding would double the lanes)
// gets a context of its own (`make`) with unit links and zero fields; one Clenshaw step of every
// storage precision is timed at the default traversal, two waves per SIMD and streaming stores (the
// common choice of the autotuner), and the size with the smallest sum is returned.  The fields are
// memory-bound and their contents do not change the time.
static int time_clenshaw_step(ovl_ctx *c, int prec, void *x, void *b1, void *b2, void *w, double *ms_per_step);

static int timed_mode_block(int n_modes, const std::function<int(int, ovl_ctx **)> &make, int *mb_out,
                            std::string *report)
{
    double best = 1e30;
    int best_mb = pick_mode_block(0, n_modes);
    std::string rep = "mode block:";
    for (int mb : {16, 32, 64}) {
        const int nblk = (n_modes + mb - 1) / mb;
        if (mb > 16 && (long)nblk * mb >= 2L * n_modes) continue;
        ovl_ctx *c = nullptr;
        int rc = make(mb, &c);
        if (rc != OVL_OK) {
            if (c) ovl_destroy(c);
            continue;  // e.g. too large for one launch: not a candidate
        }
        double sum = 0;
        rc = ensure_stencil(c);
        if (rc == OVL_OK) rc = ensure_work(c, c->g.nblk, true, true, true);
        for (int prec = 0; prec < NUM_PREC && rc == OVL_OK; ++prec) {
            c->wpe[prec] = 2;
            c->nt_out[prec] = 1;
            auto pick = [prec](void *p64, void *p32, void *p16, void *p48) {
                return prec == P64 ? p64 : (prec == P32 ? p32 : (prec == P16 ? p16 : p48));
            };
            double ms = 0;
            rc = time_clenshaw_step(c, prec, pick(c->field[OVL_PSI], c->x32, c->x16, c->x48),
                                    pick(c->a64, c->a32, c->a16, c->a48), pick(c->b64, c->b32, c->b16, c->b48),
                                    pick(c->w64, c->w32, c->w16, c->w48), &ms);
            sum += ms;
        }
        ovl_destroy(c);
        if (rc != OVL_OK) continue;
        char buf[64];
        std::snprintf(buf, sizeof buf, " %d %.1f us", mb, 1e3 * sum);
        rep += buf;
        if (sum < best) {
            best = sum;
            best_mb = mb;
        }
    }
    rep += ", chosen " + std::to_string(best_mb);
    *mb_out = best_mb;
    if (report) *report = rep;
    return OVL_OK;
}

// unit links in the reference format for n links
static std::vector<double> unit_links(int nc, long n)
{
    std::vector<double> u((size_t)n * (nc == 2 ? 4 : 18), 0.0);
    for (long i = 0; i < n; ++i) {
        if (nc == 2) u[4 * i + 3] = 1.0;
        else u[18 * i] = u[18 * i + 8] = u[18 * i + 16] = 1.0;
    }
    return u;
}

int ovl_create(const ovl_params *p, ovl_ctx **out)
{
    if (!p || !out) return OVL_E_ARG;
    *out = nullptr;
    if (p->nc != 2 && p->nc != 3) return OVL_E_ARG;
    if (p->n[0] < 3 || p->n[1] < 3 || p->n[2] < 3) return OVL_E_ARG;
    if (p->n_modes < 1) return OVL_E_ARG;
    if (p->mode_block != -1 && p->mode_block != 0 && p->mode_block != 16 && p->mode_block != 32 &&
        p->mode_block != 64)
        return OVL_E_ARG;
    if (p->traversal < 0 || p->traversal > 3) return OVL_E_ARG;
    const long V = (long)p->n[0] * p->n[1] * p->n[2];
    if (V > (1L << 26)) return OVL_E_ARG;
    std::string mb_report;
    int mb = pick_mode_block(p->mode_block, p->n_modes);
    if (p->mode_block == -1) {
        const std::vector<double> ul = unit_links(p->nc, 3 * V);
        auto make = [&](int m, ovl_ctx **cc) {
            ovl_params q = *p;
            q.mode_block = m;
            RETCHK(ovl_create(&q, cc));
            return ovl_set_links(*cc, ul.data());
        };
        RETCHK(timed_mode_block(p->n_modes, make, &mb, &mb_report));
    }
    const int nblk = (p->n_modes + mb - 1) / mb;
    if ((long)nblk * V * mb > (1L << 31) - 1) return OVL_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return OVL_E_HIP;
    ovl_ctx *c = new ovl_ctx();
    c->prm = *p;
    c->prm.mode_block = mb;
    c->mb_report = mb_report;
    c->g.nc = p->nc;
    c->g.n0 = p->n[0];
    c->g.n1 = p->n[1];
    c->g.n2 = p->n[2];
    c->g.V = (int)V;
    c->g.mb = mb;
    c->g.nblk = nblk;
    c->sys.push_back(SysDesc{p->n[0], p->n[1], p->n[2], (int)V, nblk, 0, 0, 0});
    c->link_sites = V;
    const int rc = ctx_init(c);
    if (rc != OVL_OK) {
        ovl_destroy(c);
        return rc;
    }
    *out = c;
    return OVL_OK;
}

int ovl_create_ensemble(const ovl_ensemble_params *p, ovl_ctx **out)
{
    if (!p || !out || !p->dims) return OVL_E_ARG;
    *out = nullptr;
    if (p->nc != 2 && p->nc != 3) return OVL_E_ARG;
    if (p->n_systems < 1 || p->n_systems > 65536 || p->n_modes < 1) return OVL_E_ARG;
    if (p->mode_block != -1 && p->mode_block != 0 && p->mode_block != 16 && p->mode_block != 32 &&
        p->mode_block != 64)
        return OVL_E_ARG;
    std::string mb_report;
    int mb = pick_mode_block(p->mode_block, p->n_modes);
    if (p->mode_block == -1) {
        long sites = 0;
        for (int s = 0; s < p->n_systems; ++s) {
            const int *n = p->dims + 3 * s;
            sites += (long)n[0] * n[1] * n[2];
        }
        if (sites > (1L << 29)) return OVL_E_ARG;
        const std::vector<double> ul = unit_links(p->nc, 3 * sites);
        auto make = [&](int m, ovl_ctx **cc) {
            ovl_ensemble_params q = *p;
            q.mode_block = m;
            RETCHK(ovl_create_ensemble(&q, cc));
            return ovl_set_links(*cc, ul.data());
        };
        RETCHK(timed_mode_block(p->n_modes, make, &mb, &mb_report));
    }
    const int nblk = (p->n_modes + mb - 1) / mb;
    std::vector<SysDesc> sys;
    long q0 = 0, sites = 0;
    for (int s = 0; s < p->n_systems; ++s) {
        const int *n = p->dims + 3 * s;
        if (n[0] < 3 || n[1] < 3 || n[2] < 3) return OVL_E_ARG;
        const long V = (long)n[0] * n[1] * n[2];
        if (V > (1L << 26) || (long)nblk * V * mb > (1L << 31) - 1) return OVL_E_ARG;
        sys.push_back(SysDesc{n[0], n[1], n[2], (int)V, nblk, (int)q0, (int)(3 * sites), 0});
        q0 += (long)nblk * V;
        sites += V;
        if (q0 * mb > (1L << 31) - 1 || 3 * sites > (1L << 31) - 1) return OVL_E_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return OVL_E_HIP;
    ovl_ctx *c = new ovl_ctx();
    c->ensemble = true;
    c->mb_report = mb_report;
    c->prm.nc = p->nc;
    c->prm.n[0] = (int)q0;  // the concatenation seen as one flat run of site-blocks
    c->prm.n[1] = 1;
    c->prm.n[2] = 1;
    c->prm.n_modes = p->n_modes;
    c->prm.r_wilson = p->r_wilson;
    c->prm.m_fermion = p->m_fermion;
    c->prm.m_dw = p->m_dw;
    c->prm.mode_block = mb;
    c->prm.reunitarize = p->reunitarize;
    c->prm.traversal = 1;
    // pointwise work (conversions, linear combinations) runs over all systems as one flat field
    c->g.nc = p->nc;
    c->g.n0 = (int)q0;
    c->g.n1 = 1;
    c->g.n2 = 1;
    c->g.V = (int)q0;
    c->g.mb = mb;
    c->g.nblk = 1;
    c->sys = sys;
    c->link_sites = sites;
    const int rc = ctx_init(c);
    if (rc != OVL_OK) {
        ovl_destroy(c);
        return rc;
    }
    *out = c;
    return OVL_OK;
}

int ovl_destroy(ovl_ctx *c)
{
    if (!c) return OVL_OK;
    if (c->st) hipStreamSynchronize(c->st);
    void *ps[] = {c->links, c->efield, c->ul64, c->ul32, c->phase, c->field[0], c->field[1], c->field[2], c->w64, c->a64,
                  c->b64, c->x32, c->w32, c->a32, c->b32, c->partial, c->dout, c->gam, c->jv, c->ja, c->staging,
                  c->ms_ps, c->ms_r, c->ms_q, c->ms_w, c->ms_chi, c->ms_coef, c->ms_active, c->rz[0], c->rz[1],
                  c->rz[2], c->rz[3], c->rz[4], c->rz[5], c->d_sys, c->d_work, c->d_qwork, c->d_qoff, c->d_qctr,
                  c->order[0], c->order[1],
                  c->order[2], c->order[3], c->x16, c->w16, c->a16, c->b16, c->x48, c->w48, c->a48, c->b48,
                  c->f32[0], c->f32[1], c->guard_buf[0], c->guard_buf[1]};
    if (c->step_exec) hipGraphExecDestroy(c->step_exec);
    for (void *p : ps)
        if (p) hipFree(p);
    for (auto e : c->evpool) hipEventDestroy(e);
    if (c->t0) hipEventDestroy(c->t0);
    if (c->t1) hipEventDestroy(c->t1);
    if (c->st) hipStreamDestroy(c->st);
    delete c;
    return OVL_OK;
}

static thread_local std::string g_free_err;  // last failure of a context-free call on this thread
const char *ovl_error(const ovl_ctx *c)
{
    if (c) return c->err.c_str();
    return g_free_err.empty() ? "null context" : g_free_err.c_str();
}

static int stage(ovl_ctx *c, size_t bytes)
{
    if (c->staging_bytes >= bytes) return OVL_OK;
    if (c->staging) hipFree(c->staging);
    c->staging = nullptr;
    c->staging_bytes = 0;
    RETCHK(dalloc(c, (void **)&c->staging, bytes));
    c->staging_bytes = bytes;
    return OVL_OK;
}

static bool all_finite(const double *p, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i])) return false;
    return true;
}

// *ok = no Inf/NaN among the n doubles at device address x, nor at y when given (one streaming
// reduction over both, then a synchronisation)
static int device_finite(ovl_ctx *c, const void *x, const void *y, long n, bool *ok)
{
    std::vector<double> h((size_t)ovl::eig_partial_blocks());
    HIPCHK(c, ovl::launch_nonfinite(x, y, n, c->partial, c->st));
    HIPCHK(c, hipMemcpyAsync(h.data(), c->partial, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    *ok = true;
    for (double v : h)
        if (v != 0.0) *ok = false;
    return OVL_OK;
}

static void refresh_links_set(ovl_ctx *c)
{
    c->links_set = true;
    for (char f : c->sys_links_set) c->links_set = c->links_set && f;
}

// host <-> device copies of `n` doubles at device offset `off` (elements)
static int put_doubles(ovl_ctx *c, double *dst, const double *src, size_t n, const char *what)
{
    if (!all_finite(src, n)) return fail(c, OVL_E_ARG, std::string("non-finite ") + what);
    HIPCHK(c, hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return OVL_OK;
}

static int get_doubles(ovl_ctx *c, double *dst, const double *src, size_t n)
{
    HIPCHK(c, hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return OVL_OK;
}

int ovl_set_links(ovl_ctx *c, const double *links)
{
    if (c) {  // a new gauge configuration: the guard starts again from the requested ladder
        c->guard_level = 0;
        c->guard_fresh = true;
    }
    if (!c || !links) return OVL_E_ARG;
    RETCHK(put_doubles(c, c->links, links, (size_t)link_reals(c) * 3 * c->link_sites, "gauge link"));
    std::fill(c->sys_links_set.begin(), c->sys_links_set.end(), 1);
    refresh_links_set(c);
    c->stencil_dirty = true;
    return OVL_OK;
}

int ovl_get_links(ovl_ctx *c, double *links)
{
    if (!c || !links) return OVL_E_ARG;
    return get_doubles(c, links, c->links, (size_t)link_reals(c) * 3 * c->link_sites);
}

int ovl_set_efield(ovl_ctx *c, const double *e)
{
    if (!c || !e) return OVL_E_ARG;
    RETCHK(put_doubles(c, c->efield, e, (size_t)n_adj(c) * 3 * c->link_sites, "electric field"));
    c->efield_set = true;
    return OVL_OK;
}

int ovl_get_efield(ovl_ctx *c, double *e)
{
    if (!c || !e) return OVL_E_ARG;
    return get_doubles(c, e, c->efield, (size_t)n_adj(c) * 3 * c->link_sites);
}

int ovl_set_system_links(ovl_ctx *c, int sys, const double *links)
{
    if (!c || !links) return OVL_E_ARG;
    c->guard_level = 0;
    c->guard_fresh = true;
    RETCHK(check_sys(c, sys));
    RETCHK(put_doubles(c, sys_links(c, sys), links, (size_t)link_reals(c) * 3 * c->sys[sys].V, "gauge link"));
    c->sys_links_set[sys] = 1;
    refresh_links_set(c);
    c->stencil_dirty = true;
    return OVL_OK;
}

int ovl_get_system_links(ovl_ctx *c, int sys, double *links)
{
    if (!c || !links) return OVL_E_ARG;
    RETCHK(check_sys(c, sys));
    return get_doubles(c, links, sys_links(c, sys), (size_t)link_reals(c) * 3 * c->sys[sys].V);
}

int ovl_set_system_efield(ovl_ctx *c, int sys, const double *e)
{
    if (!c || !e) return OVL_E_ARG;
    RETCHK(check_sys(c, sys));
    RETCHK(put_doubles(c, sys_efield(c, sys), e, (size_t)n_adj(c) * 3 * c->sys[sys].V, "electric field"));
    c->efield_set = true;
    return OVL_OK;
}

int ovl_get_system_efield(ovl_ctx *c, int sys, double *e)
{
    if (!c || !e) return OVL_E_ARG;
    RETCHK(check_sys(c, sys));
    return get_doubles(c, e, sys_efield(c, sys), (size_t)n_adj(c) * 3 * c->sys[sys].V);
}

// U(1) phases of system s (QED.cpp:131-153), multiplied into its stencil links
static int set_flux(ovl_ctx *c, int sys, int nb, double q)
{
    c->stencil_dirty = true;
    c->sys_nb[sys] = nb;
    if (nb == 0) {
        bool any = false;
        for (int v : c->sys_nb) any = any || v != 0;
        if (!any && c->phase) {
            HIPCHK(c, hipStreamSynchronize(c->st));
            hipFree(c->phase);
            c->phase = nullptr;
        }
        return OVL_OK;
    }
    const SysDesc &d = c->sys[sys];
    const int n0 = d.n0, n1 = d.n1, V = d.V;
    std::vector<double> ph(6 * (size_t)V);
    for (int s = 0; s < V; ++s) {
        const int z = s / (n0 * n1), y = (s - z * n0 * n1) / n0, x = s - n0 * (y + n1 * z);
        // QED.cpp:136-143: U_y = exp(-2 pi i Nb x/(Nx Ny)), U_x(Nx-1,y) = exp(2 pi i Nb y Nx/(Nx Ny))
        const double ay = -2.0 * M_PI * (double)(nb * x) / (double)(n0 * n1);
        const double ax = (x == n0 - 1) ? 2.0 * M_PI * (double)(nb * y * n0) / (double)(n0 * n1) : 0.0;
        const cplx px = std::exp(cplx(0, q * ax)), py = std::exp(cplx(0, q * ay));
        ph[2 * (0 * (size_t)V + s)] = px.real();
        ph[2 * (0 * (size_t)V + s) + 1] = px.imag();
        ph[2 * (1 * (size_t)V + s)] = py.real();
        ph[2 * (1 * (size_t)V + s) + 1] = py.imag();
        ph[2 * (2 * (size_t)V + s)] = 1.0;
        ph[2 * (2 * (size_t)V + s) + 1] = 0.0;
    }
    if (!c->phase) RETCHK(dalloc(c, (void **)&c->phase, sizeof(double) * 6 * (size_t)c->link_sites));
    HIPCHK(c, hipMemcpyAsync(c->phase + (size_t)d.l0 * 2, ph.data(), sizeof(double) * ph.size(), hipMemcpyHostToDevice,
                             c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return OVL_OK;
}

int ovl_set_u1_flux(ovl_ctx *c, int nb, double q)
{
    if (!c || !std::isfinite(q)) return OVL_E_ARG;
    c->nb = nb;
    c->q = q;
    for (int s = 0; s < nsys(c); ++s) RETCHK(set_flux(c, s, nb, q));
    return OVL_OK;
}

int ovl_set_system_u1_flux(ovl_ctx *c, int sys, int nb, double q)
{
    if (!c || !std::isfinite(q)) return OVL_E_ARG;
    RETCHK(check_sys(c, sys));
    return set_flux(c, sys, nb, q);
}

static int modes_io(ovl_ctx *c, int sys, int field, double *aos, int first, int count, bool upload)
{
    RETCHK(check_field(c, field));
    if (first < 0 || count < 1 || first + count > c->prm.n_modes) return fail(c, OVL_E_ARG, "mode range out of bounds");
    const Geom g = sys_geom(c, sys);
    void *f = sys_field(c, c->field[field], sys);
    const size_t n = (size_t)count * g.V * c->ncomp * 2;
    RETCHK(stage(c, n * sizeof(double)));
    if (upload) {
        if (!all_finite(aos, n)) return fail(c, OVL_E_ARG, "non-finite spinor component");
        HIPCHK(c, hipMemcpyAsync(c->staging, aos, n * sizeof(double), hipMemcpyHostToDevice, c->st));
        HIPCHK(c, ovl::launch_aos_to_soa(g, c->staging, first, count, f, c->st));
    } else {
        HIPCHK(c, ovl::launch_soa_to_aos(g, f, first, count, c->staging, c->st));
        HIPCHK(c, hipMemcpyAsync(aos, c->staging, n * sizeof(double), hipMemcpyDeviceToHost, c->st));
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    return OVL_OK;
}

int ovl_set_modes(ovl_ctx *c, int field, const double *aos, int first, int count)
{
    if (!c || !aos) return OVL_E_ARG;
    RETCHK(single_lattice(c, "ovl_set_modes"));
    return modes_io(c, 0, field, const_cast<double *>(aos), first, count, true);
}

int ovl_get_modes(ovl_ctx *c, int field, double *aos, int first, int count)
{
    if (!c || !aos) return OVL_E_ARG;
    RETCHK(single_lattice(c, "ovl_get_modes"));
    return modes_io(c, 0, field, aos, first, count, false);
}

int ovl_set_system_modes(ovl_ctx *c, int sys, int field, const double *aos, int first, int count)
{
    if (!c || !aos) return OVL_E_ARG;
    RETCHK(check_sys(c, sys));
    return modes_io(c, sys, field, const_cast<double *>(aos), first, count, true);
}

int ovl_get_system_modes(ovl_ctx *c, int sys, int field, double *aos, int first, int count)
{
    if (!c || !aos) return OVL_E_ARG;
    RETCHK(check_sys(c, sys));
    return modes_io(c, sys, field, aos, first, count, false);
}

int ovl_copy_field(ovl_ctx *c, int src, int dst)
{
    if (!c) return OVL_E_ARG;
    RETCHK(check_field(c, src));
    RETCHK(check_field(c, dst));
    if (src == dst) return OVL_OK;
    HIPCHK(c, ovl::launch_lincomb(c->g, c->g.nblk, P64, c->field[dst], c->field[src], 1, 0, nullptr, 0, 0, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return OVL_OK;
}

int ovl_apply_wilson(ovl_ctx *c, int in, int out)
{
    if (!c) return OVL_E_ARG;
    RETCHK(check_field(c, in));
    RETCHK(check_field(c, out));
    if (in == out) return fail(c, OVL_E_ARG, "in and out must be different fields");
    RETCHK(ensure_stencil(c));
    begin_timing(c);
    HopArgs a = hop_args(c, P64);
    a.in = c->field[in];
    a.out = c->field[out];
    a.sh_i = -1.0;  // -i h_W
    RETCHK(hop(c, c->g, P64, P64, a));
    return end_timing(c);
}

static int check_sign_opts(ovl_ctx *c, const ovl_sign_opts *o)
{
    if (!o) return fail(c, OVL_E_ARG, "null sign options");
    if (o->method != OVL_SIGN_CHEBYSHEV && o->method != OVL_SIGN_ZOLOTAREV) return fail(c, OVL_E_ARG, "unknown sign method");
    if (o->precision < OVL_PREC_FP64 || o->precision > OVL_PREC_MIXED) return fail(c, OVL_E_ARG, "unknown precision");
    if (o->method == OVL_SIGN_ZOLOTAREV && o->precision != OVL_PREC_FP64)
        return fail(c, OVL_E_ARG, "the Zolotarev sign runs in FP64 only");
    if (o->method == OVL_SIGN_ZOLOTAREV) RETCHK(single_lattice(c, "the Zolotarev sign"));
    if (o->relax < 0 || o->relax > 8) return fail(c, OVL_E_ARG, "relax must be 0..8 levels");
    return OVL_OK;
}

int ovl_apply_overlap(ovl_ctx *c, int in, int out, const ovl_sign_opts *o)
{
    if (!c) return OVL_E_ARG;
    RETCHK(check_field(c, in));
    RETCHK(check_field(c, out));
    RETCHK(check_sign_opts(c, o));
    if (in == out) return fail(c, OVL_E_ARG, "in and out must be different fields");
    begin_timing(c);
    if (o->method == OVL_SIGN_ZOLOTAREV) {
        int r = zolo_overlap(c, c->g, c->field[in], c->field[out], o->lmin, o->lmax, 0, 0.0);
        end_timing(c);
        return r;
    }
    double xmax;
    RETCHK(input_bound(c, o, c->field[in], &xmax));
    c->guard_checks = 0;
    begin_timing(c);
    ovl_sign_opts e = guarded_opts(c, o);
    if (guard_due(c, o)) RETCHK(guard_check(c, o, c->field[in], xmax, &e));
    RETCHK(sign_driver(c, &e, c->field[in], c->field[out], FIN_OVERLAP, 0.0, xmax));
    c->guard_since++;
    return end_timing(c);
}

int ovl_apply_sign(ovl_ctx *c, int in, int out, const ovl_sign_opts *o)
{
    if (!c) return OVL_E_ARG;
    RETCHK(check_field(c, in));
    RETCHK(check_field(c, out));
    RETCHK(check_sign_opts(c, o));
    if (o->method != OVL_SIGN_CHEBYSHEV) return fail(c, OVL_E_ARG, "ovl_apply_sign supports the Chebyshev sign");
    if (in == out) return fail(c, OVL_E_ARG, "in and out must be different fields");
    double xmax;
    RETCHK(input_bound(c, o, c->field[in], &xmax));
    c->guard_checks = 0;
    begin_timing(c);
    ovl_sign_opts e = guarded_opts(c, o);
    if (guard_due(c, o)) RETCHK(guard_check(c, o, c->field[in], xmax, &e));
    RETCHK(sign_driver(c, &e, c->field[in], c->field[out], FIN_SIGN, 0.0, xmax));
    c->guard_since++;
    return end_timing(c);
}

int ovl_mode_norms(ovl_ctx *c, int field, double *norm2)
{
    if (!c || !norm2) return OVL_E_ARG;
    RETCHK(check_field(c, field));
    std::vector<cplx> d;
    RETCHK(mode_dot_all(c, c->field[field], c->field[field], d));
    for (size_t i = 0; i < d.size(); ++i) norm2[i] = d[i].real();
    return OVL_OK;
}

int ovl_sign_residual(ovl_ctx *c, int in, const ovl_sign_opts *o, double *rel)
{
    // ||eps(eps x) - x|| / ||x||  (OverlapSignTest.cpp:155-165), max over modes
    if (!c || !rel) return OVL_E_ARG;
    RETCHK(check_field(c, in));
    RETCHK(check_sign_opts(c, o));
    void *t1 = nullptr, *t2 = nullptr;
    RETCHK(dalloc(c, &t1, c->field64_bytes));
    if (dalloc(c, &t2, c->field64_bytes) != OVL_OK) {
        hipFree(t1);
        return OVL_E_NOMEM;
    }
    double xmax = 0;
    int r = input_bound(c, o, c->field[in], &xmax);
    // |eps(x)| <= ~|x| (the polynomial sign is bounded by about 1 on the whole spectrum)
    if (r == OVL_OK) r = sign_driver(c, o, c->field[in], t1, FIN_SIGN, 0.0, xmax);
    if (r == OVL_OK) r = sign_driver(c, o, t1, t2, FIN_SIGN, 0.0, 2.0 * xmax);
    std::vector<cplx> dx, dd;
    if (r == OVL_OK) {
        const hipError_t e = ovl::launch_lincomb(c->g, c->g.nblk, P64, t2, t2, 1, 0, c->field[in], -1, 0, c->st);
        if (e != hipSuccess) r = fail(c, OVL_E_HIP, std::string("sign residual: ") + hipGetErrorString(e));
    }
    if (r == OVL_OK) {
        r = mode_dot_all(c, t2, t2, dd);
        if (r == OVL_OK) r = mode_dot_all(c, c->field[in], c->field[in], dx);
    }
    hipStreamSynchronize(c->st);
    hipFree(t1);
    hipFree(t2);
    if (r != OVL_OK) return r;
    double worst = 0;
    for (size_t i = 0; i < dx.size(); ++i)
        if (dx[i].real() > 0) worst = std::max(worst, std::sqrt(dd[i].real() / dx[i].real()));
    *rel = worst;
    return OVL_OK;
}

// ---- single-mode drop-ins on the first mode block of the scratch field ----
static int dropin_setup(ovl_ctx *c, const double *ucc, const double *phi)
{
    if (!ucc || !phi) return fail(c, OVL_E_ARG, "null argument");
    RETCHK(single_lattice(c, "the single-mode drop-ins"));
    const size_t nu = (size_t)3 * c->g.V * c->g.nc * c->g.nc * 2, np = (size_t)c->g.V * c->ncomp * 2;
    // the inputs are checked for Inf/NaN on the device after the upload (a host scan of the 8.6 MB
    // of a 24^3 SU(3) call costs about as long as the copy itself)
    bool ok = true;
    RETCHK(stage(c, std::max(nu, np) * sizeof(double)));
    HIPCHK(c, hipMemcpyAsync(c->staging, ucc, nu * sizeof(double), hipMemcpyHostToDevice, c->st));
    RETCHK(device_finite(c, c->staging, nullptr, (long)nu, &ok));
    if (!ok) return fail(c, OVL_E_ARG, "non-finite input (u_cc)");
    HIPCHK(c, ovl::launch_ucc_to_links(c->g, c->staging, c->ul64, c->ul32, c->st));
    c->stencil_dirty = true;  // the next batched call re-derives the stencil from the master links
    Geom g1 = c->g;
    g1.nblk = 1;
    HIPCHK(c, hipMemsetAsync(c->field[OVL_SCRATCH], 0, c->blk64_elems * 16, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, hipMemcpyAsync(c->staging, phi, np * sizeof(double), hipMemcpyHostToDevice, c->st));
    RETCHK(device_finite(c, c->staging, nullptr, (long)np, &ok));
    if (!ok) return fail(c, OVL_E_ARG, "non-finite input (phi)");
    HIPCHK(c, ovl::launch_aos_to_soa(g1, c->staging, 0, 1, c->field[OVL_SCRATCH], c->st));
    return OVL_OK;
}

static int dropin_fetch(ovl_ctx *c, const void *field, double *psi)
{
    Geom g1 = c->g;
    g1.nblk = 1;
    const size_t np = (size_t)c->g.V * c->ncomp * 2;
    HIPCHK(c, ovl::launch_soa_to_aos(g1, field, 0, 1, c->staging, c->st));
    HIPCHK(c, hipMemcpyAsync(psi, c->staging, np * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return OVL_OK;
}

int ovl_wilson_evo_compute(ovl_ctx *c, const double *ucc, const double *phi, double *psi)
{
    if (!c || !psi) return OVL_E_ARG;
    RETCHK(dropin_setup(c, ucc, phi));
    RETCHK(ensure_work(c, 1, false));
    Geom g1 = c->g;
    g1.nblk = 1;
    HopArgs a = hop_args(c, P64);
    a.nblk = 1;
    a.in = c->field[OVL_SCRATCH];
    a.out = c->w64;
    a.sh_i = -1.0;
    RETCHK(hop(c, g1, P64, P64, a));
    return dropin_fetch(c, c->w64, psi);
}

int ovl_overlap_operator_compute(ovl_ctx *c, const double *ucc, const double *phi, double *psi, double lmin,
                                 double lmax)
{
    if (!c || !psi) return OVL_E_ARG;
    if (!(lmin > 0.0) || !(lmax > lmin)) return fail(c, OVL_E_SPECTRUM, "Zolotarev interval needs 0 < lmin < lmax");
    RETCHK(dropin_setup(c, ucc, phi));
    RETCHK(ensure_work(c, 1, false));
    Geom g1 = c->g;
    g1.nblk = 1;
    // zolo_overlap's ensure_stencil() must not overwrite the uploaded u_cc links
    const bool saved_set = c->links_set;
    c->links_set = true;
    c->stencil_dirty = false;
    int r = zolo_overlap(c, g1, c->field[OVL_SCRATCH], c->w64, lmin, lmax, 0, 0.0);
    c->links_set = saved_set;
    c->stencil_dirty = true;
    if (r != OVL_OK && r != OVL_E_CONVERGE) return r;
    RETCHK(dropin_fetch(c, c->w64, psi));
    return r;
}

// ---- Ritz (RitzSolver.cpp:92-297), neigen = 1, on mode 0 of one block ----
int ovl_ritz(ovl_ctx *c, const double *start_flat, double sign, double delitr, int nitlz, int nrestart,
             double *eigenvalue)
{
    if (!c || !start_flat || !eigenvalue) return OVL_E_ARG;
    RETCHK(single_lattice(c, "ovl_ritz"));
    RETCHK(ensure_stencil(c));
    Geom g = c->g;
    g.nblk = 1;
    const size_t fb = c->blk64_elems * 16;
    for (int i = 0; i < 6; ++i) RETCHK(dalloc(c, &c->rz[i], fb));
    void *xr = c->rz[0], *yr = c->rz[1], *zr = c->rz[2], *gr = c->rz[3], *pr = c->rz[4], *w = c->rz[5];
    // flat (m = s + V(spin + 4 colour)) -> AoS -> SoA mode 0
    const int V = c->g.V, nc = c->g.nc;
    std::vector<double> aos((size_t)V * c->ncomp * 2);
    for (int s = 0; s < V; ++s)
        for (int k = 0; k < 4; ++k)
            for (int j = 0; j < nc; ++j) {
                const size_t m = s + (size_t)V * (k + 4 * j);
                aos[2 * ((s * nc + j) * 4 + k)] = start_flat[2 * m];
                aos[2 * ((s * nc + j) * 4 + k) + 1] = start_flat[2 * m + 1];
            }
    RETCHK(stage(c, aos.size() * sizeof(double)));
    for (void *v : {xr, yr, zr, gr, pr, w}) HIPCHK(c, hipMemsetAsync(v, 0, fb, c->st));
    HIPCHK(c, hipMemcpyAsync(c->staging, aos.data(), aos.size() * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, ovl::launch_aos_to_soa(g, c->staging, 0, 1, xr, c->st));
    auto dot = [&](const void *a, const void *b, cplx &res) -> int {
        std::vector<cplx> d;
        RETCHK(mode_dot(c, g, a, b, d));
        res = d[0];
        return OVL_OK;
    };
    auto lin = [&](void *out, const void *a, cplx ca, const void *b, cplx cb) -> int {
        HIPCHK(c, ovl::launch_lincomb(g, 1, P64, out, a, ca.real(), ca.imag(), b, cb.real(), cb.imag(), c->st));
        return OVL_OK;
    };
    auto uu = [&](const void *v, void *out) -> int {  // out = sign * h^2 v
        RETCHK(apply_h2(c, g, v, w, out, 0.0));
        return lin(out, out, sign, nullptr, 0.0);
    };
    cplx t;
    RETCHK(dot(xr, xr, t));
    double xnorm = 1.0 / std::sqrt(t.real());
    RETCHK(lin(xr, xr, xnorm, nullptr, 0.0));
    int ittot = 0;
    bool done = false;
    cplx xynorm = 0, oxynorm = 0, xpnorm = 0, amb = 0;
    double gnorm = 0, gnormim = 0, pnorm = 0;
    while (!done) {
        RETCHK(uu(xr, yr));
        RETCHK(dot(xr, yr, xynorm));
        RETCHK(lin(gr, yr, 1.0, xr, -xynorm));
        RETCHK(lin(pr, gr, 1.0, nullptr, 0.0));
        RETCHK(dot(gr, gr, t));
        gnormim = t.real();
        pnorm = std::sqrt(gnormim);
        RETCHK(lin(pr, pr, 1.0 / pnorm, nullptr, 0.0));
        for (int itt = 1; itt <= nrestart; ++itt) {
            ittot++;
            RETCHK(uu(pr, zr));
            amb = xynorm;
            RETCHK(dot(pr, zr, t));
            amb -= t.real();
            RETCHK(dot(xr, zr, t));
            const double ccc = t.real();
            const double cox = 2.0 * ccc / amb.real();
            const double sig1 = -std::copysign(1.0, amb.real());
            const double sig2 = std::copysign(1.0, cox) * sig1;
            const double cox2 = cox * cox;
            double ct, stt;
            if (cox2 <= 1.0e-3) {
                // series of RitzSolver.cpp:199-200
                const double sqp = 1.0 - 0.125 * cox2 + 0.0859375 * cox2 * cox2 - 0.0673828125 * cox2 * cox2 * cox2 +
                                   0.0562439 * cox2 * cox2 * cox2 * cox2 + 0.0487 * cox2 * cox2 * cox2 * cox2 * cox2;
                const double sqm = 0.5 * sig1 * sig2 * cox *
                                   (1.0 - 0.375 * cox2 + 0.2421875 * cox2 * cox2 - 0.182617875 * cox2 * cox2 * cox2 +
                                    0.1482849 * cox2 * cox2 * cox2 * cox2 - 0.12575 * cox2 * cox2 * cox2 * cox2 * cox2 +
                                    0.1097 * cox2 * cox2 * cox2 * cox2 * cox2 * cox2);
                if (sig1 < 0.0) { ct = sqm; stt = sig2 * sqp; }
                else { ct = sqp; stt = sig2 * sqm; }
            } else {
                const double sox = sig1 / std::sqrt(1.0 + cox2);
                ct = std::sqrt((1.0 + sox) * 0.5);
                stt = std::sqrt((1.0 - sox) * 0.5) * sig2;
            }
            RETCHK(lin(xr, xr, ct, pr, stt));
            RETCHK(lin(yr, yr, ct, zr, stt));
            RETCHK(dot(xr, xr, t));
            xnorm = t.real();
            oxynorm = xynorm;
            RETCHK(dot(xr, yr, xynorm));
            xnorm = 1.0 / std::sqrt(xnorm);
            xynorm *= xnorm;
            RETCHK(lin(xr, xr, xnorm, nullptr, 0.0));
            RETCHK(lin(gr, yr, 1.0, xr, -xynorm));
            RETCHK(dot(gr, gr, t));
            gnorm = t.real();
            RETCHK(dot(xr, pr, xpnorm));
            if (xynorm.real() > oxynorm.real()) { done = true; break; }
            if (gnorm < delitr * sign * xynorm.real()) { done = true; break; }
            const double beta = gnorm / gnormim * pnorm;
            // pr = gr + beta (pr - xr xpnorm)
            RETCHK(lin(pr, pr, beta, xr, -beta * xpnorm));
            RETCHK(lin(pr, pr, 1.0, gr, 1.0));
            RETCHK(dot(pr, pr, t));
            pnorm = std::sqrt(t.real());
            RETCHK(lin(pr, pr, 1.0 / pnorm, nullptr, 0.0));
            gnormim = gnorm;
            if (ittot >= nitlz) { done = true; break; }
        }
    }
    *eigenvalue = xynorm.real();
    return OVL_OK;
}

}  // extern "C"

namespace {
// MaxEValue = -1.05 Ritz(-1), MinEValue = 0.95 Ritz(+1) with a normalised uniform start vector,
// 100 restarts, 2000 iterations, 1e-16 (OverlapUpdate.cpp:130-173); the start vector comes from a
// fixed-seed 64-bit Mersenne twister instead of the reference's GSL stream
int ritz_bounds(ovl_ctx *c, double *lmin, double *lmax)
{
    const size_t n = (size_t)c->g.V * c->ncomp;
    std::vector<double> v(2 * n);
    unsigned long long st = 0x9E3779B97F4A7C15ull;
    double nrm = 0;
    for (size_t i = 0; i < 2 * n; ++i) {
        st ^= st << 13; st ^= st >> 7; st ^= st << 17;  // xorshift64
        v[i] = (double)(st >> 11) * (1.0 / 9007199254740992.0);
        nrm += v[i] * v[i];
    }
    nrm = std::sqrt(nrm);
    for (auto &x : v) x /= nrm;
    double a = 0, b = 0;
    RETCHK(ovl_ritz(c, v.data(), -1.0, 1e-16, 2000, 100, &a));
    RETCHK(ovl_ritz(c, v.data(), 1.0, 1e-16, 2000, 100, &b));
    *lmax = -1.05 * a;
    *lmin = 0.95 * b;
    return OVL_OK;
}
}  // namespace

extern "C" {

// ---- dynamics ----
int ovl_update_links(ovl_ctx *c, double dt)
{
    if (!c || !std::isfinite(dt)) return OVL_E_ARG;
    if (!c->links_set) return fail(c, OVL_E_STATE, "no gauge links uploaded");
    for (int s = 0; s < nsys(c); ++s)
        HIPCHK(c, ovl::launch_link_update(sys_geom(c, s), sys_links(c, s), sys_efield(c, s), dt, c->prm.reunitarize,
                                          c->st));
    c->stencil_dirty = true;
    HIPCHK(c, hipStreamSynchronize(c->st));
    return OVL_OK;
}

int ovl_update_efield(ovl_ctx *c, double dt)
{
    if (!c || !std::isfinite(dt)) return OVL_E_ARG;
    if (!c->links_set) return fail(c, OVL_E_STATE, "no gauge links uploaded");
    for (int s = 0; s < nsys(c); ++s)
        HIPCHK(c, ovl::launch_efield_update(sys_geom(c, s), sys_links(c, s), sys_efield(c, s), dt, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return OVL_OK;
}

static int gauge_half(ovl_ctx *c, int mode, double h)
{
    if (mode == OVL_GAUGE_FROZEN) return OVL_OK;
    for (int s = 0; s < nsys(c); ++s) {
        const Geom g = sys_geom(c, s);
        if (mode == OVL_GAUGE_FULL)
            HIPCHK(c, ovl::launch_efield_update(g, sys_links(c, s), sys_efield(c, s), h, c->st));
        HIPCHK(c, ovl::launch_link_update(g, sys_links(c, s), sys_efield(c, s), h, c->prm.reunitarize, c->st));
        HIPCHK(c, ovl::launch_link_setup(g, sys_links(c, s), sys_phase(c, s), sys_ul64(c, s), sys_ul32(c, s), c->st));
    }
    return OVL_OK;
}

// the Wilson leapfrog steps FP32 spinor fields unless the options ask for FP64
static bool wilson_fp32(const ovl_step_opts *o) { return o->kind == OVL_WILSON && o->sign.precision != OVL_PREC_FP64; }

// target += dt * (-i h src)   (OverlapUpdate.cpp:85-124, WILSON/Update.cpp:6-45)
static int fermion_half(ovl_ctx *c, const ovl_step_opts *o, int src, int tgt, void *hpsi = nullptr)
{
    if (o->kind == OVL_WILSON) {
        // one stencil launch with the axpy in its epilogue: FP64 fields, or the FP32 copies (3S + L/M
        // bytes per site-mode at S = 96 B for SU(3) instead of 192)
        const int pr = wilson_fp32(o) ? P32 : P64;
        HopArgs a = hop_args(c, pr);
        a.in = pr == P32 ? c->f32[src] : c->field[src];
        a.out = pr == P32 ? c->f32[tgt] : c->field[tgt];
        a.tgt = a.out;
        a.st_r = 1.0;
        a.sh_i = -o->dt;
        return hop(c, c->g, pr, pr, a);
    }
    if (o->sign.method == OVL_SIGN_ZOLOTAREV)
        return zolo_overlap(c, c->g, c->field[src], c->field[tgt], o->sign.lmin, o->sign.lmax, 1, o->dt);
    return sign_driver(c, &o->sign, c->field[src], c->field[tgt], FIN_OVERLAP_AXPY, o->dt, c->step_xmax, hpsi);
}

// every field of the step options; a new option must be added here (the sizes below catch it)
static_assert(sizeof(ovl_sign_opts) == 88 && sizeof(ovl_step_opts) == 120, "compare every option in same_step_opts");
static bool same_step_opts(const ovl_step_opts &a, const ovl_step_opts &b)
{
    const ovl_sign_opts &x = a.sign, &y = b.sign;
    return a.kind == b.kind && a.dt == b.dt && a.gauge == b.gauge && a.use_graph == b.use_graph &&
           x.method == y.method && x.precision == y.precision && x.degree == y.degree && x.cut == y.cut &&
           x.tol == y.tol && x.head == y.head && x.lmin == y.lmin && x.lmax == y.lmax &&
           x.mode_chunk == y.mode_chunk && x.fp16_tail == y.fp16_tail && x.head16 == y.head16 &&
           x.fp48 == y.fp48 && x.head48 == y.head48 && x.guard == y.guard && x.relax == y.relax;
}

static StepKey make_step_key(const ovl_ctx *c, const ovl_step_opts *o, const SignPlan *pl)
{
    StepKey k;
    k.o = *o;
    if (pl) {
        k.head48 = pl->head48;
        k.head = pl->head;
        k.head16 = pl->head16;
        k.sx = pl->sx;
        k.sk = pl->sk;
        k.sw = pl->sw;
    }
    k.ptrs = {c->links, c->efield, c->ul64, c->ul32, c->phase, c->field[0], c->field[1], c->field[2],
              c->w64, c->a64, c->b64, c->x32, c->w32, c->a32, c->b32, c->x16, c->w16, c->a16, c->b16,
              c->x48, c->w48, c->a48, c->b48, c->partial, c->d_sys, c->d_work, c->d_qwork, c->d_qoff, c->d_qctr,
              c->order[0], c->order[1],
              c->order[2], c->order[3], c->f32[0], c->f32[1]};
    k.ints = {c->work_blocks, c->ens_grid};
    for (int nb : c->sys_nb) k.ints.push_back(nb != 0);  // systems whose links take the phases
    for (int p = 0; p < NUM_PREC; ++p)
        for (int v : {c->trav[p], c->slabs[p], c->nt_out[p], c->wpe[p]}) k.ints.push_back(v);
    return k;
}

static bool same_step_key(const StepKey &a, const StepKey &b)
{
    return same_step_opts(a.o, b.o) && a.head48 == b.head48 && a.head == b.head && a.head16 == b.head16 &&
           a.sx == b.sx && a.sk == b.sk && a.sw == b.sw && a.ptrs == b.ptrs && a.ints == b.ints;
}

static void drop_step_graph(ovl_ctx *c)
{
    if (c->step_exec) {
        hipStreamSynchronize(c->st);
        hipGraphExecDestroy(c->step_exec);
    }
    c->step_exec = nullptr;
    c->step_key = StepKey{};
}

// observe: the second half step also leaves -i h_ov[U(t + dt/2)] Psi(t + dt) in the scratch field
static int one_step(ovl_ctx *c, const ovl_step_opts *o, bool observe = false)
{
    RETCHK(fermion_half(c, o, OVL_PSIMID, OVL_PSI));
    RETCHK(gauge_half(c, o->gauge, 0.5 * o->dt));
    RETCHK(fermion_half(c, o, OVL_PSI, OVL_PSIMID, observe ? c->field[OVL_SCRATCH] : nullptr));
    RETCHK(gauge_half(c, o->gauge, 0.5 * o->dt));
    return OVL_OK;
}

static int update_impl(ovl_ctx *c, const ovl_step_opts *o, int nsteps, bool observe_last);

int ovl_fermion_update(ovl_ctx *c, const ovl_step_opts *o, int nsteps)
{
    if (!c || !o || nsteps < 0) return OVL_E_ARG;
    c->guard_checks = 0;
    return update_impl(c, o, nsteps, false);
}

// Psi and PsiMid after a leapfrog call: a call whose fields turned non-finite (a step far beyond the
// leapfrog's stability bound) fails instead of handing Inf/NaN back silently.  One reduction per
// field and call.
static int check_fields_finite(ovl_ctx *c)
{
    const long n = (long)(c->field64_bytes / sizeof(double));
    bool ok = true;
    RETCHK(device_finite(c, c->field[OVL_PSI], c->field[OVL_PSIMID], n, &ok));
    if (!ok)
        return fail(c, OVL_E_NONFINITE, "Psi or PsiMid holds Inf/NaN after the leapfrog (dt beyond its stability bound?)");
    return OVL_OK;
}

// ||h_ov|| <= M (|1 + m/2| + |1 - m/2|) for |eps| <= 1 (OverlapU.cpp:188-201); the staggered leapfrog
// keeps the mode norms bounded only while dt ||h|| < 2
static double overlap_norm_bound(const ovl_ctx *c)
{
    const double m = c->prm.m_fermion, M = std::fabs(c->prm.m_dw);
    return M * (std::fabs(1.0 + 0.5 * m) + std::fabs(1.0 - 0.5 * m));
}

// nsteps leapfrog steps; with observe_last the last step runs with stream launches and leaves
// -i h_ov Psi in the scratch field
static int update_impl(ovl_ctx *c, const ovl_step_opts *o, int nsteps, bool observe_last)
{
    if (!c || !o || nsteps < 0 || !std::isfinite(o->dt)) return OVL_E_ARG;
    if (o->kind != OVL_WILSON && o->kind != OVL_OVERLAP) return fail(c, OVL_E_ARG, "unknown fermion kind");
    if (o->gauge < OVL_GAUGE_FROZEN || o->gauge > OVL_GAUGE_FULL) return fail(c, OVL_E_ARG, "unknown gauge mode");
    if (o->kind == OVL_OVERLAP) RETCHK(check_sign_opts(c, &o->sign));
    c->step_xmax = 0.0;
    SignPlan pl;
    const ovl_step_opts *const req = o;  // the options as requested (o becomes the guarded ones)
    ovl_step_opts oe = *o;
    const bool cheb = o->kind == OVL_OVERLAP && o->sign.method == OVL_SIGN_CHEBYSHEV;
    if (cheb) {
        // bound of the mode norms for the FP16 scales: the staggered leapfrog conserves the norms
        // for dt |omega| < 2, so twice the larger of the two fields at the start is a safe bound.
        // Rounded up to a power of two: the norms of the two fields oscillate by O(dt omega) from
        // step to step, and the FP16 scales (and with them a kept step graph) should not follow.
        double b0, b1;
        RETCHK(input_bound(c, &o->sign, c->field[OVL_PSI], &b0));
        RETCHK(input_bound(c, &o->sign, c->field[OVL_PSIMID], &b1));
        const double bnd = 2.0 * std::max(b0, b1);
        int ex = 0;
        const double mant = std::frexp(bnd, &ex);  // bnd = mant 2^ex, mant in [1/2, 1)
        c->step_xmax = (bnd > 0.0 && std::isfinite(bnd)) ? std::ldexp(1.0, mant == 0.5 ? ex - 1 : ex) : bnd;
        // when a check is due (new links or ladder options, or `guard` steps since the last) the
        // guard takes Psi's next mode block through both signs before the first step; the steps
        // run with the ladder as tightened
        oe.sign = guarded_opts(c, &o->sign);
        if (guard_due(c, &o->sign)) RETCHK(guard_check(c, &o->sign, c->field[OVL_PSI], c->step_xmax, &oe.sign));
        // FP16 storage is scaled for mode norms within twice their start: beyond the leapfrog's
        // stability bound (dt ||h_ov|| >= 2, norms growing) the ladder stops at FP32
        if (std::fabs(o->dt) * overlap_norm_bound(c) >= 2.0) oe.sign.fp16_tail = 0;
        o = &oe;
        RETCHK(make_plan(c, &o->sign, pl, c->step_xmax));
    }
    RETCHK(ensure_stencil(c));
    const bool w32 = wilson_fp32(o) && nsteps > 0;
    if (w32)
        for (int f = 0; f < 2; ++f)
            if (!c->f32[f]) RETCHK(dalloc(c, &c->f32[f], c->field32_bytes));
    begin_timing(c);
    if (observe_last && nsteps > 0) {
        RETCHK(update_impl(c, req, nsteps - 1, false));
        begin_timing(c);
        RETCHK(one_step(c, o, true));
        c->guard_since += 1;
        RETCHK(end_timing(c));
        return check_fields_finite(c);
    }
    // FP32 Wilson: the call steps the FP32 copies of Psi and PsiMid and writes them back at the end
    if (w32)
        for (int f = 0; f < 2; ++f)
            HIPCHK(c, ovl::launch_convert(c->g, c->g.nblk, c->field[f], P64, c->f32[f], P32, c->st));
    auto finish = [&]() -> int {
        c->guard_since += nsteps;
        if (w32)
            for (int f = 0; f < 2; ++f)
                HIPCHK(c, ovl::launch_convert(c->g, c->g.nblk, c->f32[f], P32, c->field[f], P64, c->st));
        RETCHK(end_timing(c));
        return nsteps > 0 ? check_fields_finite(c) : OVL_OK;
    };
    const bool graph = o->use_graph && !(o->kind == OVL_OVERLAP && o->sign.method == OVL_SIGN_ZOLOTAREV) && nsteps > 0;
    if (!graph) {
        for (int s = 0; s < nsteps; ++s) RETCHK(one_step(c, o));
        return finish();
    }
    // one full step (both half steps and both gauge half steps) as a graph, replayed nsteps times;
    // every buffer the captured step touches is allocated before the capture.  The instantiated
    // graph stays in the context and later calls replay it while the step's key is unchanged.
    const bool mixed = o->kind == OVL_OVERLAP && o->sign.precision == OVL_PREC_MIXED;
    RETCHK(ensure_work(c, (o->sign.mode_chunk > 0) ? std::min(o->sign.mode_chunk, c->g.nblk) : c->g.nblk,
                       o->kind == OVL_OVERLAP && o->sign.precision != OVL_PREC_FP64, mixed && o->sign.fp16_tail,
                       mixed && o->sign.fp48));
    const StepKey key = make_step_key(c, o, cheb ? &pl : nullptr);
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, hipEventRecord(c->t0, c->st));
    if (!(c->step_exec && same_step_key(key, c->step_key))) {
        drop_step_graph(c);
        hipGraph_t graph_h = nullptr;
        hipGraphExec_t exec = nullptr;
        c->capturing = true;
        hipError_t e = hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
            c->capturing = false;
            return fail(c, OVL_E_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
        }
        const long long l0 = c->hop_launches;
        const double b0 = c->hop_bytes;
        long long vl0[NVAR];
        double vb0[NVAR];
        for (int v = 0; v < NVAR; ++v) { vl0[v] = c->var_launches[v]; vb0[v] = c->var_bytes[v]; }
        const int s0[4] = {c->fp64_steps, c->fp32_steps, c->fp16_steps, c->fp48_steps};
        int r = one_step(c, o);
        hipError_t ee = hipStreamEndCapture(c->st, &graph_h);
        c->capturing = false;
        // the capture only recorded: the counters advance with the replays below
        c->step_dl = c->hop_launches - l0;
        c->step_db = c->hop_bytes - b0;
        for (int v = 0; v < NVAR; ++v) {
            c->step_dvl[v] = c->var_launches[v] - vl0[v];
            c->step_dvb[v] = c->var_bytes[v] - vb0[v];
            c->var_launches[v] = vl0[v];
            c->var_bytes[v] = vb0[v];
        }
        c->hop_launches = l0;
        c->hop_bytes = b0;
        c->step_dp[0] = c->fp64_steps - s0[0];
        c->step_dp[1] = c->fp32_steps - s0[1];
        c->step_dp[2] = c->fp16_steps - s0[2];
        c->step_dp[3] = c->fp48_steps - s0[3];
        c->fp64_steps = s0[0];
        c->fp32_steps = s0[1];
        c->fp16_steps = s0[2];
        c->fp48_steps = s0[3];
        if (r != OVL_OK) {
            if (graph_h) hipGraphDestroy(graph_h);
            return r;
        }
        if (ee != hipSuccess) return fail(c, OVL_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ee));
        e = hipGraphInstantiate(&exec, graph_h, nullptr, nullptr, 0);
        hipGraphDestroy(graph_h);
        if (e != hipSuccess) return fail(c, OVL_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        c->step_exec = exec;
        c->step_key = key;
        c->captures++;
    }
    hipError_t e = hipSuccess;
    for (int s = 0; s < nsteps && e == hipSuccess; ++s) e = hipGraphLaunch(c->step_exec, c->st);
    if (e != hipSuccess) {
        drop_step_graph(c);
        return fail(c, OVL_E_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
    }
    c->hop_launches += c->step_dl * nsteps;
    c->hop_bytes += c->step_db * nsteps;
    for (int v = 0; v < NVAR; ++v) {
        c->var_launches[v] += c->step_dvl[v] * nsteps;
        c->var_bytes[v] += c->step_dvb[v] * nsteps;
    }
    c->fp64_steps += c->step_dp[0] * nsteps;
    c->fp32_steps += c->step_dp[1] * nsteps;
    c->fp16_steps += c->step_dp[2] * nsteps;
    c->fp48_steps += c->step_dp[3] * nsteps;
    return finish();
}

// ---- observables ----
static int fetch_out(ovl_ctx *c, double *h, int n)
{
    HIPCHK(c, hipMemcpyAsync(h, c->dout, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return OVL_OK;
}

static int obs_energy(ovl_ctx *c, int s, double *e_el, double *e_mag)
{
    if (!c->sys_links_set[s]) return fail(c, OVL_E_STATE, "no gauge links uploaded");
    HIPCHK(c, ovl::launch_energy(sys_geom(c, s), sys_links(c, s), sys_efield(c, s), c->partial, c->dout, c->st));
    double h[2];
    RETCHK(fetch_out(c, h, 2));
    *e_el = h[0];
    *e_mag = h[1];
    return OVL_OK;
}

static int obs_gauss(ovl_ctx *c, int s, double *vmax, double *vrms)
{
    if (!c->sys_links_set[s]) return fail(c, OVL_E_STATE, "no gauge links uploaded");
    HIPCHK(c, ovl::launch_gauss(sys_geom(c, s), sys_links(c, s), sys_efield(c, s), c->partial, c->dout, nullptr,
                                c->st));
    double h[2];
    RETCHK(fetch_out(c, h, 2));
    *vmax = h[0];
    *vrms = std::sqrt(h[1] / (n_adj(c) * (double)c->sys[s].V));
    return OVL_OK;
}

static int obs_unitarity(ovl_ctx *c, int s, double *vmax, double *vrms)
{
    if (!c->sys_links_set[s]) return fail(c, OVL_E_STATE, "no gauge links uploaded");
    HIPCHK(c, ovl::launch_unitarity(sys_geom(c, s), sys_links(c, s), c->partial, c->dout, c->st));
    double h[2];
    RETCHK(fetch_out(c, h, 2));
    *vmax = h[0];
    *vrms = std::sqrt(h[1] / (3.0 * c->sys[s].V));
    return OVL_OK;
}

static int obs_ncs_dot(ovl_ctx *c, int s, double *ncsdot)
{
    if (!c->sys_links_set[s]) return fail(c, OVL_E_STATE, "no gauge links uploaded");
    HIPCHK(c, ovl::launch_edotb(sys_geom(c, s), sys_links(c, s), sys_efield(c, s), c->partial, c->dout, c->st));
    double h[1];
    RETCHK(fetch_out(c, h, 1));
    *ncsdot = 0.125 * h[0] / (M_PI * M_PI);  // ChernSimonsDerivative.cpp:47
    return OVL_OK;
}

int ovl_energy(ovl_ctx *c, double *e_el, double *e_mag)
{
    if (!c || !e_el || !e_mag) return OVL_E_ARG;
    RETCHK(single_lattice(c, "ovl_energy"));
    return obs_energy(c, 0, e_el, e_mag);
}

int ovl_gauss_violation(ovl_ctx *c, double *vmax, double *vrms)
{
    if (!c || !vmax || !vrms) return OVL_E_ARG;
    RETCHK(single_lattice(c, "ovl_gauss_violation"));
    return obs_gauss(c, 0, vmax, vrms);
}

int ovl_unitarity_violation(ovl_ctx *c, double *vmax, double *vrms)
{
    if (!c || !vmax || !vrms) return OVL_E_ARG;
    RETCHK(single_lattice(c, "ovl_unitarity_violation"));
    return obs_unitarity(c, 0, vmax, vrms);
}

int ovl_ncs_dot(ovl_ctx *c, double *ncsdot)
{
    if (!c || !ncsdot) return OVL_E_ARG;
    RETCHK(single_lattice(c, "ovl_ncs_dot"));
    return obs_ncs_dot(c, 0, ncsdot);
}

int ovl_system_observables(ovl_ctx *c, int sys, double *obs)
{
    if (!c || !obs) return OVL_E_ARG;
    RETCHK(check_sys(c, sys));
    RETCHK(obs_energy(c, sys, &obs[OVL_OBS_E_EL], &obs[OVL_OBS_E_MAG]));
    RETCHK(obs_gauss(c, sys, &obs[OVL_OBS_GAUSS_MAX], &obs[OVL_OBS_GAUSS_RMS]));
    RETCHK(obs_unitarity(c, sys, &obs[OVL_OBS_UNIT_MAX], &obs[OVL_OBS_UNIT_RMS]));
    return obs_ncs_dot(c, sys, &obs[OVL_OBS_NCS_DOT]);
}

// ---- allocation-free observables ----
// The reductions' per-workgroup partial results (ovl::reduce_partial_len doubles) and the four raw
// results live in this file-scope device array instead of a context's buffers.  Each device holds
// its own copy; host code uses only the address hipGetSymbolAddress returns for the current device.
// Every reduction writes all the partial results its finishing pass reads (a fixed grid), so the
// array needs no clearing between calls; calls that overlap on one device share it and race.
constexpr size_t NOALLOC_SCRATCH_DOUBLES = 4096;  // 32 KiB per device
constexpr size_t NOALLOC_RAW = NOALLOC_SCRATCH_DOUBLES - 8;  // raw results: scratch[NOALLOC_RAW .. +7)
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[NOALLOC_MAX_DEVICES];  // device address of `scratch`, per device

static int free_fail(int code, const std::string &msg)
{
    g_free_err = msg;
    return code;
}

// the device address of `scratch` on the current device, looked up once per device
static int noalloc_scratch(double **out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return free_fail(OVL_E_HIP, "no current HIP device");
    if (dev >= NOALLOC_MAX_DEVICES) return free_fail(OVL_E_HIP, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return free_fail(OVL_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return OVL_OK;
}

// raw reduction results -> the quantities of ovl_system_observables, in the same arithmetic
__global__ void k_obs_noalloc_finish(const double *raw, double *obs, double n_adj_v, double three_v)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    obs[OVL_OBS_E_EL] = raw[0];
    obs[OVL_OBS_E_MAG] = raw[1];
    obs[OVL_OBS_GAUSS_MAX] = raw[2];
    obs[OVL_OBS_GAUSS_RMS] = sqrt(raw[3] / n_adj_v);
    obs[OVL_OBS_UNIT_MAX] = raw[4];
    obs[OVL_OBS_UNIT_RMS] = sqrt(raw[5] / three_v);
    obs[OVL_OBS_NCS_DOT] = 0.125 * raw[6] / (M_PI * M_PI);  // ChernSimonsDerivative.cpp:47
}

int ovl_observables_noalloc(int nc, const int *n, const double *links, const double *efield, double *obs,
                            void *stream)
{
    if ((nc != 2 && nc != 3) || !n || !links || !efield || !obs)
        return free_fail(OVL_E_ARG, "ovl_observables_noalloc: nc must be 2 or 3 and every pointer set");
    if (n[0] < 3 || n[1] < 3 || n[2] < 3)
        return free_fail(OVL_E_ARG, "ovl_observables_noalloc: every lattice extent must be >= 3");
    const long V = (long)n[0] * n[1] * n[2];
    if (V > OVL_NOALLOC_MAX_SITES)
        return free_fail(OVL_E_ARG, "ovl_observables_noalloc: " + std::to_string(V) + " sites exceed the maximum of " +
                                        std::to_string((long)OVL_NOALLOC_MAX_SITES));
    Geom g{};
    g.nc = nc;
    g.n0 = n[0];
    g.n1 = n[1];
    g.n2 = n[2];
    g.V = (int)V;
    g.mb = 16;
    g.nblk = 1;
    if (ovl::reduce_partial_len(g) > NOALLOC_RAW)
        return free_fail(OVL_E_ARG, "ovl_observables_noalloc: reduction partials exceed the static scratch");
    double *s = nullptr;
    RETCHK(noalloc_scratch(&s));
    double *raw = s + NOALLOC_RAW;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t e = ovl::launch_energy(g, links, efield, s, raw + 0, st);
    if (e == hipSuccess) e = ovl::launch_gauss(g, links, efield, s, raw + 2, nullptr, st);
    if (e == hipSuccess) e = ovl::launch_unitarity(g, links, s, raw + 4, st);
    if (e == hipSuccess) e = ovl::launch_edotb(g, links, efield, s, raw + 6, st);
    if (e == hipSuccess) {
        k_obs_noalloc_finish<<<1, 64, 0, st>>>(raw, obs, (double)(nc * nc - 1) * (double)V, 3.0 * (double)V);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return free_fail(OVL_E_HIP, std::string("ovl_observables_noalloc: ") + hipGetErrorString(e));
    return OVL_OK;
}

// GaussLawRestoration::Restore (GaussRestore.cpp:129-160): relaxation sweeps of E with step
// sizes 7/24, 7/48, 7/72 until max_a,x |G^a(x)| <= tol (checked on the violation measured at the
// start of the last sweep, as the reference does) or max_iter iterations
int ovl_gauss_restore(ovl_ctx *c, int sys, double tol, int max_iter, int *iters, double *vmax)
{
    if (!c || !(tol >= 0.0) || max_iter < 0) return OVL_E_ARG;
    RETCHK(check_sys(c, sys));
    if (!c->sys_links_set[sys]) return fail(c, OVL_E_STATE, "no gauge links uploaded");
    const Geom g = sys_geom(c, sys);
    double *gf = nullptr;
    RETCHK(dalloc(c, (void **)&gf, sizeof(double) * n_adj(c) * (size_t)g.V));
    const double *L = sys_links(c, sys);
    double *E = sys_efield(c, sys);
    auto deviation = [&](double *mx) -> int {
        HIPCHK(c, ovl::launch_gauss(g, L, E, c->partial, c->dout, gf, c->st));
        if (mx) {
            double h[2];
            RETCHK(fetch_out(c, h, 2));
            *mx = h[0];
        }
        return OVL_OK;
    };
    double mx = 0;
    int it = 0, r = deviation(&mx);
    const double gam[3] = {7.0 / 24.0, 7.0 / 48.0, 7.0 / 72.0};
    while (r == OVL_OK && it < max_iter && mx > tol) {
        for (int k = 0; k < 3 && r == OVL_OK; ++k) {
            r = deviation(k == 2 ? &mx : nullptr);
            if (r == OVL_OK && ovl::launch_gauss_restore(g, L, gf, E, gam[k], c->st) != hipSuccess)
                r = fail(c, OVL_E_HIP, "gauss restore launch failed");
        }
        ++it;
    }
    if (r == OVL_OK) r = deviation(&mx);  // violation of the final field
    hipStreamSynchronize(c->st);
    hipFree(gf);
    if (iters) *iters = it;
    if (vmax) *vmax = mx;
    return r;
}

int ovl_ensemble_balance(ovl_ctx *c, double *imbalance)
{
    if (!c || !imbalance) return OVL_E_ARG;
    *imbalance = c->ens_imbalance;
    return OVL_OK;
}

static void build_gammas(std::vector<double> &out)
{
    // OverlapCurrents.cpp:14-97 coupling matrices from DiracAlgebra.cpp:5-9
    const cplx I(0, 1);
    cplx g0[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, -1, 0}, {0, 0, 0, -1}};
    cplx gx[4][4] = {{0, 0, 0, 1}, {0, 0, 1, 0}, {0, -1, 0, 0}, {-1, 0, 0, 0}};
    cplx gy[4][4] = {{0, 0, 0, -I}, {0, 0, I, 0}, {0, I, 0, 0}, {-I, 0, 0, 0}};
    cplx gz[4][4] = {{0, 0, 1, 0}, {0, 0, 0, -1}, {-1, 0, 0, 0}, {0, 1, 0, 0}};
    cplx g5[4][4] = {{0, 0, 1, 0}, {0, 0, 0, 1}, {1, 0, 0, 0}, {0, 1, 0, 0}};
    cplx(*gm[4])[4] = {g0, gx, gy, gz};
    out.assign(4 * 4 * 16 * 2, 0.0);
    for (int mu = 0; mu < 4; ++mu)
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) {
                cplx v = 0, ax = 0, v2 = 0, a2 = 0;
                for (int g = 0; g < 4; ++g) {
                    v += g0[a][g] * gm[mu][g][b];
                    for (int d = 0; d < 4; ++d) {
                        ax += g0[a][g] * gm[mu][g][d] * g5[d][b];
                        v2 += cplx(0, -0.5) * g0[a][g] * gm[mu][g][d] * g0[d][b];
                        for (int e = 0; e < 4; ++e) a2 += cplx(0, -0.5) * g0[a][g] * gm[mu][g][d] * g5[d][e] * g0[e][b];
                    }
                }
                const cplx vals[4] = {v, ax, v2, a2};
                for (int k = 0; k < 4; ++k) {
                    const size_t idx = ((k * 4 + mu) * 4 + a) * 4 + b;
                    out[2 * idx] = vals[k].real();
                    out[2 * idx + 1] = vals[k].imag();
                }
            }
}

static int currents_from(ovl_ctx *c, const void *psi, int first_mode, double *jv, double *ja, double *axial_charge);

int ovl_currents(ovl_ctx *c, int field, const ovl_sign_opts *o, int first_mode, double *jv, double *ja,
                 double *axial_charge)
{
    if (!c) return OVL_E_ARG;
    RETCHK(check_field(c, field));
    RETCHK(check_sign_opts(c, o));
    if (field == OVL_SCRATCH) return fail(c, OVL_E_ARG, "currents use the scratch field for -i h_ov psi");
    RETCHK(single_lattice(c, "ovl_currents"));
    if (o->method == OVL_SIGN_ZOLOTAREV)
        RETCHK(zolo_overlap(c, c->g, c->field[field], c->field[OVL_SCRATCH], o->lmin, o->lmax, 0, 0.0));
    else
    {
        double xmax;
        RETCHK(input_bound(c, o, c->field[field], &xmax));
        RETCHK(sign_driver(c, o, c->field[field], c->field[OVL_SCRATCH], FIN_OVERLAP, 0.0, xmax));
    }
    return currents_from(c, c->field[field], first_mode, jv, ja, axial_charge);
}

// per-site currents from a field and -i h_ov of it in the scratch field (OverlapCurrents.cpp:137-228)
static int currents_from(ovl_ctx *c, const void *psi, int first_mode, double *jv, double *ja, double *axial_charge)
{
    if (!c->gam) {
        std::vector<double> g;
        build_gammas(g);
        RETCHK(dalloc(c, (void **)&c->gam, sizeof(double) * g.size()));
        HIPCHK(c, hipMemcpyAsync(c->gam, g.data(), sizeof(double) * g.size(), hipMemcpyHostToDevice, c->st));
    }
    RETCHK(dalloc(c, (void **)&c->jv, sizeof(double) * 4 * c->g.V));
    RETCHK(dalloc(c, (void **)&c->ja, sizeof(double) * 4 * c->g.V));
    HIPCHK(c, ovl::launch_currents(c->g, psi, c->field[OVL_SCRATCH], c->gam, first_mode, c->partial, c->jv, c->ja,
                                   c->st));
    std::vector<double> hv(4 * (size_t)c->g.V), ha(4 * (size_t)c->g.V);
    HIPCHK(c, hipMemcpyAsync(hv.data(), c->jv, sizeof(double) * hv.size(), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(ha.data(), c->ja, sizeof(double) * ha.size(), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (jv) std::memcpy(jv, hv.data(), sizeof(double) * hv.size());
    if (ja) std::memcpy(ja, ha.data(), sizeof(double) * ha.size());
    if (axial_charge) {
        double s = 0;
        for (int i = 0; i < c->g.V; ++i) s += ha[4 * i];
        *axial_charge = s;
    }
    return OVL_OK;
}

// Dynamics::Fermions::Update with Fermions::Observables::ComputeCurrents every `every` steps
// (Simulation.cpp:466-480), the currents taken from the overlap apply the leapfrog already does:
// the second half step applies -i h_ov with the links of the middle of the step to Psi(t + dt),
// which is what the reference's ComputeCurrents uses after Update (it reads EvolutionOperator::Ucc,
// last set up before that half step, OverlapUpdate.cpp:116 and OverlapCurrents.cpp:134).
int ovl_fermion_update_observed(ovl_ctx *c, const ovl_step_opts *o, int nsteps, ovl_observe *obs)
{
    if (!c || !o || !obs || nsteps < 0) return OVL_E_ARG;
    RETCHK(single_lattice(c, "ovl_fermion_update_observed"));
    obs->count = 0;
    if (obs->every <= 0) return ovl_fermion_update(c, o, nsteps);
    if (o->kind != OVL_OVERLAP || o->sign.method != OVL_SIGN_CHEBYSHEV)
        return fail(c, OVL_E_ARG, "observed updates take the currents from the Chebyshev overlap apply");
    // the step keeps -i h_ov Psi as (dt term) / dt: dt = 0 leaves nothing to recover it from
    if (!(o->dt != 0.0)) return fail(c, OVL_E_ARG, "observed updates need dt != 0 (the currents reuse the step's dt-scaled overlap term)");
    c->guard_checks = 0;
    int done = 0;
    while (done < nsteps) {
        const int chunk = std::min(obs->every, nsteps - done);
        const bool observe = chunk == obs->every;
        RETCHK(update_impl(c, o, chunk, observe));
        done += chunk;
        if (!observe) break;
        double q5 = 0, ncs = 0;
        RETCHK(currents_from(c, c->field[OVL_PSI], obs->first_mode, obs->jv, obs->ja, &q5));
        RETCHK(obs_ncs_dot(c, 0, &ncs));
        if (obs->axial_charge) obs->axial_charge[obs->count] = q5;
        if (obs->ncs_dot) obs->ncs_dot[obs->count] = ncs;
        obs->count++;
    }
    return OVL_OK;
}

// ---- configuration files (IO/SaveConfiguration.cpp:3-63, IO/LoadConfiguration.cpp:3-144) ----
// One line per link, sites in storage order (z, y, x outermost to innermost, then mu):
//   links   "x y z mu v_0 v_1 ... " with the link's stored reals (SU(2): u0..u3; SU(3): re im of
//           the 9 entries, column-major), each followed by a space (MatrixToString)
//   E field "x y z mu E_0 ... E_{Nc^2-2}"
// written with 17 significant digits so that FP64 values read back bit for bit.
int ovl_save_configuration(ovl_ctx *c, int sys, const char *ufile, const char *efile)
{
    if (!c || !ufile) return OVL_E_ARG;
    RETCHK(check_sys(c, sys));
    const SysDesc &d = c->sys[sys];
    const int lr = link_reals(c), ng = n_adj(c);
    std::vector<double> u((size_t)lr * 3 * d.V), e;
    RETCHK(get_doubles(c, u.data(), sys_links(c, sys), u.size()));
    if (efile) {
        e.resize((size_t)ng * 3 * d.V);
        RETCHK(get_doubles(c, e.data(), sys_efield(c, sys), e.size()));
    }
    FILE *fu = std::fopen(ufile, "w");
    if (!fu) return fail(c, OVL_E_ARG, std::string("cannot open ") + ufile);
    FILE *fe = efile ? std::fopen(efile, "w") : nullptr;
    if (efile && !fe) {
        std::fclose(fu);
        return fail(c, OVL_E_ARG, std::string("cannot open ") + efile);
    }
    for (int z = 0; z < d.n2; ++z)
        for (int y = 0; y < d.n1; ++y)
            for (int x = 0; x < d.n0; ++x)
                for (int mu = 0; mu < 3; ++mu) {
                    const size_t l = (size_t)mu + 3 * ((size_t)x + (size_t)d.n0 * (y + (size_t)d.n1 * z));
                    std::fprintf(fu, "%d %d %d %d ", x, y, z, mu);
                    for (int a = 0; a < lr; ++a) std::fprintf(fu, "%.17g ", u[l * lr + a]);
                    std::fputc('\n', fu);
                    if (fe) {
                        std::fprintf(fe, "%d %d %d %d", x, y, z, mu);
                        for (int a = 0; a < ng; ++a) std::fprintf(fe, " %.17g", e[l * ng + a]);
                        std::fputc('\n', fe);
                    }
                }
    const bool ok = std::fclose(fu) == 0 && (!fe || std::fclose(fe) == 0);
    if (!ok) return fail(c, OVL_E_ARG, "write error");
    return OVL_OK;
}

// reads "x y z mu v_0 .. v_{nv-1}" lines into out[nv * (mu + 3 site)]; every link exactly once
static int read_link_file(ovl_ctx *c, const SysDesc &d, const char *path, int nv, std::vector<double> &out)
{
    std::ifstream in(path);
    if (!in) return fail(c, OVL_E_ARG, std::string("cannot open ") + path);
    out.assign((size_t)nv * 3 * d.V, 0.0);
    std::vector<char> seen((size_t)3 * d.V, 0);
    std::string line;
    long count = 0, lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
        std::istringstream ls(line);
        int x, y, z, mu;
        if (!(ls >> x >> y >> z >> mu) || x < 0 || x >= d.n0 || y < 0 || y >= d.n1 || z < 0 || z >= d.n2 || mu < 0 ||
            mu > 2)
            return fail(c, OVL_E_ARG, std::string(path) + ":" + std::to_string(lineno) + ": bad site or direction");
        const size_t l = (size_t)mu + 3 * ((size_t)x + (size_t)d.n0 * (y + (size_t)d.n1 * z));
        for (int a = 0; a < nv; ++a)
            if (!(ls >> out[l * nv + a]) || !std::isfinite(out[l * nv + a]))
                return fail(c, OVL_E_ARG, std::string(path) + ":" + std::to_string(lineno) + ": expected " +
                                              std::to_string(nv) + " finite values");
        if (seen[l]) return fail(c, OVL_E_ARG, std::string(path) + ":" + std::to_string(lineno) + ": link given twice");
        seen[l] = 1;
        ++count;
    }
    if (count != 3L * d.V)
        return fail(c, OVL_E_ARG, std::string(path) + ": " + std::to_string(count) + " links read, " +
                                      std::to_string(3L * d.V) + " expected");
    return OVL_OK;
}

int ovl_load_configuration(ovl_ctx *c, int sys, const char *ufile, const char *efile)
{
    if (!c || !ufile) return OVL_E_ARG;
    c->guard_level = 0;
    c->guard_fresh = true;
    RETCHK(check_sys(c, sys));
    const SysDesc d = c->sys[sys];
    std::vector<double> u, e;
    RETCHK(read_link_file(c, d, ufile, link_reals(c), u));
    if (efile) RETCHK(read_link_file(c, d, efile, n_adj(c), e));
    RETCHK(ovl_set_system_links(c, sys, u.data()));
    if (efile) RETCHK(ovl_set_system_efield(c, sys, e.data()));
    return OVL_OK;
}

// ---- initial eigenbasis in a constant magnetic field ----
// OverlapQEDEigenfunctions::Setup (INITIALCONDITIONS/FERMIONS/BASIS/OverlapQEDEigenfunctions.cpp:141-591)
// and InitialConditions::SetOverlapQEDEigenfunctions (SetOverlapQEDEigenfunctions.cpp:5-56).  The
// reference applies the overlap operator to the columns of a random orthonormal basis Q and forms
// H = i (A Q) Q^+; the unit basis gives the same matrix without the O(D^3) product.
int ovl_qed_eigenbasis(ovl_ctx *c, const ovl_sign_opts *o, int nb, double q, double dt, int first_mode,
                       double *eigenvalues, double *hermiticity)
{
    if (!c || !o || !std::isfinite(q) || !std::isfinite(dt) || first_mode < 0) return OVL_E_ARG;
    RETCHK(single_lattice(c, "ovl_qed_eigenbasis"));
    RETCHK(check_sign_opts(c, o));
    const int nc = c->g.nc, V = c->g.V, n0 = c->g.n0, n1 = c->g.n1, n2 = c->g.n2;
    const long D = 4L * nc * V;
    // colour-0 block of one z momentum: (x, y, spin), Dz = 4 n0 n1; all n2 of them on the device at once
    const long Dz = 4L * n0 * n1;
    if (Dz > 16384) return fail(c, OVL_E_ARG, "4 n0 n1 above 16384: the xy blocks of the eigenproblem are too large");
    if ((double)n2 * Dz * Dz * 16.0 > 32e9) return fail(c, OVL_E_ARG, "the n2 blocks of 4 n0 n1 squared exceed 32 GB");
    const int nm = c->prm.n_modes;
    // every requested mode must name an eigenvector: global mode g = particle + 2 level with
    // level < 2 Nc V = D/2 (GetLocalModeQuantumNumbers, FermionField.cpp:63-74, which reports an
    // indexing error when g / 2 carries past the last level)
    if ((long)first_mode + nm > D)
        return fail(c, OVL_E_ARG, "mode index beyond the D/2 levels of the eigenbasis (first_mode + n_modes > 4 Nc V)");
    // U(1)-only stencil links (ConvertU1GaugeLinks, OverlapQEDEigenfunctions.cpp:20-48): colour (0,0)
    // carries the phase of QED::SetConstantMagneticField (QED.cpp:136-143), every other entry is 0
    std::vector<double> ucc((size_t)3 * V * nc * nc * 2, 0.0);
    for (int st = 0; st < V; ++st) {
        const int z = st / (n0 * n1), y = (st - z * n0 * n1) / n0, x = st - n0 * (y + n1 * z);
        (void)z;
        const double ay = -2.0 * M_PI * (double)(nb * x) / (double)(n0 * n1);
        const double ax = (x == n0 - 1) ? 2.0 * M_PI * (double)(nb * y * n0) / (double)(n0 * n1) : 0.0;
        const double ph[3] = {q * ax, q * ay, 0.0};
        for (int mu = 0; mu < 3; ++mu) {
            const size_t e = ((size_t)(mu * V + st) * nc + 0) * nc + 0;
            ucc[2 * e] = std::cos(ph[mu]);
            ucc[2 * e + 1] = std::sin(ph[mu]);
        }
    }
    RETCHK(stage(c, ucc.size() * sizeof(double)));
    HIPCHK(c, hipMemcpyAsync(c->staging, ucc.data(), ucc.size() * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, ovl::launch_ucc_to_links(c->g, c->staging, c->ul64, c->ul32, c->st));
    c->stencil_custom = true;
    c->stencil_dirty = true;  // afterwards the stencil is re-derived from the master links
    struct Restore {
        ovl_ctx *c;
        ~Restore() { c->stencil_custom = false; }
    } restore{c};
    // a context never tuned (automatic traversal, no ovl_set_tuning / ovl_autotune) runs the
    // eigenbasis's FP64 stencil as the autotuner picks it for large lattices: XCD boxes, streaming
    // stores, two waves per SIMD (bit-identical to every other variant); its own tuning comes back
    // afterwards
    struct Retune {
        ovl_ctx *c;
        bool on;
        ovl_tuning saved;
        ~Retune()
        {
            if (!on) return;
            ovl_set_tuning(c, P64, &saved);
            c->tuned = false;
        }
    } retune{c, false, {}};
    if (c->prm.traversal == 0 && c->boxes_ok && !c->tuned) {
        RETCHK(ovl_get_tuning(c, P64, &retune.saved));
        ovl_tuning t = retune.saved;
        t.traversal = 2;
        t.slabs = 1;
        t.nt_store = 1;
        t.waves_per_eu = 2;
        RETCHK(ovl_set_tuning(c, P64, &t));
        retune.on = true;
    }

    // Structure of H = i (-i h_ov) on this background (exact, not an approximation):
    //  * colours do not mix (the link matrices are diag(phase, 0, ..)); colours >= 1 have no hopping,
    //    so there H is one 4x4 spin matrix, the same at every site and colour;
    //  * the phases do not depend on z, so H commutes with translations in z: on colour 0 it is block
    //    diagonal in the z momentum k = 2 pi j / n2, with blocks
    //      B_j[(x', y', s'), (x, y, s)] = sum_z' H[(x', y', z', s'), (x, y, 0, s)] exp(-i k z').
    // So H is applied to the Dz unit sources of the z = 0 plane (colour 0) and to the 4 spin sources of
    // site 0, colour 1, the n2 Hermitian blocks of size Dz are diagonalised in one batched rocSOLVER
    // call, and the spectrum is merged: the same eigenvalues and eigenspaces as the dense matrix of
    // the reference, at Nc n2 times fewer overlap applies and an (Nc n2)^3 smaller eigensolve.
    const size_t per_mode = (size_t)V * 4 * nc;  // complex entries of one mode
    // AoS index of (site, colour, spin) in a host mode
    auto aos = [&](long site, int colour, int spin) { return (size_t)((site * nc + colour) * 4 + spin); };
    std::vector<cplx> tw((size_t)n2 * n2);  // exp(-2 pi i j z / n2), symmetric in (j, z)
    for (int j = 0; j < n2; ++j)
        for (int z = 0; z < n2; ++z) tw[(size_t)j * n2 + z] = std::polar(1.0, -2.0 * M_PI * (double)((long)j * z % n2) / n2);
    cplx spin_block[16];  // colour >= 1, column-major
    double leak = 0;      // largest entry outside the structure above
    const long nsrc = Dz + (nc > 1 ? 4 : 0);
    const size_t plane = (size_t)n0 * n1;
    // Zolotarev: one spectral interval for every batch (the Ritz estimate of the U(1) operator once)
    double zlmin = o->lmin, zlmax = o->lmax;
    if (o->method == OVL_SIGN_ZOLOTAREV && zlmin == 0.0 && zlmax == 0.0) RETCHK(ritz_bounds(c, &zlmin, &zlmax));

    // device buffers: the n2 blocks (column-major, block j at j Dz^2), the z-momentum planes of one
    // batch (the FP64 work field w64: per mode block n2 planes of plane * 4 Nc * MB complex, the size
    // of one field), the DFT matrix, the eigensolver's outputs
    const int nprob = nc > 1 ? 2 : 1;
    std::vector<double> Wz((size_t)n2 * Dz), W4(4, 0.0);
    void *dA = nullptr, *dTW = nullptr, *dW = nullptr, *dE = nullptr, *dInfo = nullptr, *dS = nullptr, *dM = nullptr;
    rocblas_handle handle = nullptr;
    auto release = [&]() {
        hipStreamSynchronize(c->st);
        if (handle) rocblas_destroy_handle(handle);
        handle = nullptr;
        dfree(dA); dfree(dTW); dfree(dW); dfree(dE); dfree(dInfo); dfree(dS); dfree(dM);
    };
    struct Release {
        decltype(release) &f;
        ~Release() { f(); }
    } release_on_exit{release};
    RETCHK(ensure_work(c, c->g.nblk, false));
    RETCHK(dalloc(c, &dA, sizeof(cplx) * (size_t)n2 * Dz * Dz));
    RETCHK(dalloc(c, &dTW, sizeof(cplx) * tw.size()));
    RETCHK(dalloc(c, &dW, sizeof(double) * Wz.size()));
    RETCHK(dalloc(c, &dE, sizeof(double) * Wz.size()));
    RETCHK(dalloc(c, &dInfo, sizeof(int) * n2));
    RETCHK(dalloc(c, &dS, sizeof(cplx) * 16 + sizeof(double) * 8));  // 4x4 spin problem: matrix, W, E
    HIPCHK(c, hipMemcpyAsync(dTW, tw.data(), sizeof(cplx) * tw.size(), hipMemcpyHostToDevice, c->st));
    if (rocblas_create_handle(&handle) != rocblas_status_success) {
        handle = nullptr;
        return fail(c, OVL_E_HIP, "rocblas_create_handle failed");
    }
    rocblas_set_stream(handle, c->st);
    // host maxima of the per-workgroup partial results written to c->partial
    auto partial_max = [&](int stride, int which, double *out) -> int {
        const int nb = ovl::eig_partial_blocks();
        std::vector<double> h((size_t)nb * stride);
        HIPCHK(c, hipMemcpyAsync(h.data(), c->partial, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        double m = 0;
        for (int b = 0; b < nb; ++b) m = std::max(m, h[(size_t)b * stride + which]);
        *out = m;
        return OVL_OK;
    };
    const Geom &g = c->g;
    const long ncol = (long)plane * c->ncomp * g.mb;  // entries of one z plane of one mode block
    std::vector<cplx> spin_resp;
    for (long k0 = 0; k0 < nsrc; k0 += nm) {
        const int cnt = (int)std::min<long>(nm, nsrc - k0);
        const int nsb = (int)std::max<long>(0, std::min<long>(cnt, Dz - k0));  // colour-0 sources of the batch
        HIPCHK(c, hipMemsetAsync(c->field[OVL_SCRATCH], 0, c->field64_bytes, c->st));
        HIPCHK(c, ovl::launch_eig_sources(g, c->field[OVL_SCRATCH], k0, cnt, Dz, c->st));
        if (o->method == OVL_SIGN_ZOLOTAREV)
            RETCHK(zolo_overlap(c, g, c->field[OVL_SCRATCH], c->field[OVL_PSIMID], zlmin, zlmax, 0, 0.0));
        else
            RETCHK(sign_driver(c, o, c->field[OVL_SCRATCH], c->field[OVL_PSIMID], FIN_OVERLAP, 0.0, 1.0));
        if (nsb > 0) {
            // z momenta of every mode block: out[blk][j] = i sum_z r[blk][z] exp(-2 pi i j z / n2), one
            // strided batched GEMM over the mode blocks (column-major: out (ncol x n2) = R (ncol x n2,
            // leading dimension ncol: the z planes of a mode block follow each other) times TW)
            const rocblas_double_complex one_i = {0.0, 1.0}, zero = {0.0, 0.0};
            const rocblas_status rs = rocblas_zgemm_strided_batched(
                handle, rocblas_operation_none, rocblas_operation_none, (rocblas_int)ncol, n2, n2, &one_i,
                (const rocblas_double_complex *)c->field[OVL_PSIMID], (rocblas_int)ncol, (rocblas_stride)(ncol * n2),
                (const rocblas_double_complex *)dTW, n2, 0, &zero, (rocblas_double_complex *)c->w64,
                (rocblas_int)ncol, (rocblas_stride)(ncol * n2), g.nblk);
            if (rs != rocblas_status_success) return fail(c, OVL_E_HIP, "zgemm: status " + std::to_string((int)rs));
            HIPCHK(c, ovl::launch_eig_gather(g, c->w64, dA, k0, nsb, c->st));
            // colours >= 1 of the colour-0 responses must vanish
            HIPCHK(c, ovl::launch_eig_leak(g, c->field[OVL_PSIMID], nsb, c->partial, c->st));
            double lk = 0;
            RETCHK(partial_max(1, 0, &lk));
            leak = std::max(leak, lk);
        }
        if (nsb < cnt) {
            // the 4 spin sources of colour 1 at site 0: responses on the host
            spin_resp.assign((size_t)(cnt - nsb) * per_mode, cplx(0, 0));
            RETCHK(modes_io(c, 0, OVL_PSIMID, reinterpret_cast<double *>(spin_resp.data()), nsb, cnt - nsb, false));
            for (int j = nsb; j < cnt; ++j) {
                const int sc = (int)(k0 + j - Dz);
                const cplx *r = spin_resp.data() + (size_t)(j - nsb) * per_mode;
                for (long site = 0; site < V; ++site)
                    for (int col = 0; col < nc; ++col)
                        for (int sp = 0; sp < 4; ++sp) {
                            const cplx v = r[aos(site, col, sp)];
                            if (site == 0 && col == 1) spin_block[sc * 4 + sp] = cplx(0, 1) * v;
                            else leak = std::max(leak, std::abs(v) <= 1.7976931348623157e308 ? std::abs(v) : INFINITY);
                        }
            }
        }
    }
    // Hermiticity of every block (the reference prints entries with |H - H^+|^2 > 1e-12), then the
    // Hermitian part
    double asym = 0, nonfinite = 0;
    HIPCHK(c, ovl::launch_eig_hermitize(dA, Dz, n2, c->partial, c->st));
    RETCHK(partial_max(2, 0, &asym));
    RETCHK(partial_max(2, 1, &nonfinite));
    if (nc > 1) {
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j <= i; ++j) {
                cplx &a = spin_block[j * 4 + i], &b = spin_block[i * 4 + j];
                if (!(std::isfinite(a.real()) && std::isfinite(a.imag()) && std::isfinite(b.real()) && std::isfinite(b.imag())))
                    nonfinite = 1;
                asym = std::max(asym, std::abs(a - std::conj(b)));
                const cplx h = 0.5 * (a + std::conj(b));
                a = h;
                b = std::conj(h);
            }
    }
    if (hermiticity) *hermiticity = asym;
    if (nonfinite != 0) return fail(c, OVL_E_SPECTRUM, "the assembled overlap operator has a non-finite entry");
    if (!(leak <= 1e-12))
        return fail(c, OVL_E_SPECTRUM, "the U(1) overlap operator mixes colours (entry " + std::to_string(leak) + ")");

    // diagonalisation on the device (the reference: LAPACK zheev of the dense matrix,
    // MISC/LapackDiagonalization.cpp:68-120): all z-momentum blocks in one batched divide and
    // conquer, in place; the 4x4 spin block of the other colours as one more small problem
    std::string why;
    for (int prob = 0; prob < nprob && why.empty(); ++prob) {
        const long n = prob == 0 ? Dz : 4;
        const int count = prob == 0 ? n2 : 1;
        double *w = prob == 0 ? Wz.data() : W4.data();
        void *mat = prob == 0 ? dA : dS;
        double *wdev = prob == 0 ? (double *)dW : (double *)((cplx *)dS + 16);
        double *edev = prob == 0 ? (double *)dE : wdev + 4;
        std::vector<int> info(count, -1);
        hipError_t he = hipSuccess;
        if (prob == 1) he = hipMemcpyAsync(mat, spin_block, sizeof(spin_block), hipMemcpyHostToDevice, c->st);
        rocblas_status rs = rocblas_status_success;
        if (he == hipSuccess)
            rs = rocsolver_zheevd_strided_batched(handle, rocblas_evect_original, rocblas_fill_lower, (rocblas_int)n,
                                                  (rocblas_double_complex *)mat, (rocblas_int)n, (rocblas_stride)(n * n),
                                                  wdev, (rocblas_stride)n, edev, (rocblas_stride)n,
                                                  (rocblas_int *)dInfo, count);
        if (he == hipSuccess && rs == rocblas_status_success)
            he = hipMemcpyAsync(w, wdev, sizeof(double) * (size_t)count * n, hipMemcpyDeviceToHost, c->st);
        if (he == hipSuccess && rs == rocblas_status_success && prob == 1)
            he = hipMemcpyAsync(spin_block, mat, sizeof(spin_block), hipMemcpyDeviceToHost, c->st);
        if (he == hipSuccess && rs == rocblas_status_success)
            he = hipMemcpyAsync(info.data(), dInfo, sizeof(int) * count, hipMemcpyDeviceToHost, c->st);
        if (he == hipSuccess) he = hipStreamSynchronize(c->st);
        if (he != hipSuccess) { why = std::string("zheevd: ") + hipGetErrorString(he); break; }
        if (rs != rocblas_status_success) { why = "zheevd: status " + std::to_string((int)rs); break; }
        for (int b = 0; b < count; ++b)
            if (info[b] != 0) why = "zheevd: info " + std::to_string(info[b]) + " in block " + std::to_string(b);
        for (long i = 0; i < (long)count * n; ++i)
            if (!std::isfinite(w[i])) why = "zheevd: non-finite eigenvalue";
    }
    if (!why.empty()) return fail(c, OVL_E_CONVERGE, "eigensolver failed: " + why);
    HIPCHK(c, ovl::launch_nonfinite(dA, nullptr, (long)n2 * Dz * Dz * 2, c->partial, c->st));
    double bad = 0;
    RETCHK(partial_max(1, 0, &bad));
    if (bad != 0) return fail(c, OVL_E_CONVERGE, "eigensolver failed: non-finite eigenvector entries");
    for (const cplx &v : spin_block)
        if (!(std::isfinite(v.real()) && std::isfinite(v.imag())))
            return fail(c, OVL_E_CONVERGE, "eigensolver failed: non-finite eigenvector entries");

    // the merged spectrum in descending order, as the reference stores it (LapackDiagonalization.cpp:
    // 103-112); ties (degenerate levels) are ordered by block and by index, the reference's order
    // within a degenerate eigenspace is LAPACK's and any orthonormal basis of it is an eigenbasis.
    // Entry: {value, block (0..n2-1: z momentum; -1: colours >= 1), index (block eigenvector or
    // 4 * copy + spin eigenvector)}
    struct Ent { double w; int blk; long idx; };
    std::vector<Ent> ent;
    ent.reserve((size_t)D);
    for (int jz = 0; jz < n2; ++jz)
        for (long l = 0; l < Dz; ++l) ent.push_back({Wz[(size_t)jz * Dz + l], jz, l});
    for (long copy = 0; copy < (long)V * (nc - 1); ++copy)
        for (int l = 0; l < 4; ++l) ent.push_back({W4[l], -1, 4 * copy + l});
    std::stable_sort(ent.begin(), ent.end(), [](const Ent &a, const Ent &b) { return a.w > b.w; });
    if ((long)ent.size() != D) return fail(c, OVL_E_STATE, "eigenbasis size mismatch");
    if (eigenvalues)
        for (long b = 0; b < D; ++b) eigenvalues[b] = ent[b].w;

    // Psi(s) = sqrt(V) v_b, PsiMid(s) = Psi(s) exp(-i omega dt / 2), b from (particle, level) of the
    // global mode index (FermionField.cpp:63-74, SetOverlapQEDEigenfunctions.cpp:14-44); the fields
    // are assembled on the device from the eigenvector blocks, one descriptor per mode
    std::vector<ovl::EigMode> em((size_t)nm);
    const double norm = std::sqrt((double)V);
    for (int sm = 0; sm < nm; ++sm) {
        const long gm = (long)first_mode + sm, particle = gm % 2, level = (gm / 2) % (2L * V * nc);
        const long b = particle == 0 ? D / 2 - 1 - level : D / 2 + level;
        const Ent &e = ent[b];
        const cplx rot = std::exp(cplx(0, -0.5 * e.w * dt));
        ovl::EigMode &d = em[(size_t)sm];
        d = ovl::EigMode{};
        d.rot_r = rot.real();
        d.rot_i = rot.imag();
        if (e.blk >= 0) {
            // exp(i k z) chi(x, y, s) sqrt(V / n2) on colour 0, chi column idx of block blk
            d.blk = e.blk;
            d.idx = (int)e.idx;
        } else {
            // spin eigenvector at one (site, colour >= 1)
            const long copy = e.idx / 4;
            const int l = (int)(e.idx % 4);
            d.blk = -1;
            d.idx = (int)(copy % V);
            d.col = 1 + (int)(copy / V);
            for (int sp = 0; sp < 4; ++sp) {
                const cplx v = norm * spin_block[l * 4 + sp];
                d.v[2 * sp] = v.real();
                d.v[2 * sp + 1] = v.imag();
            }
        }
    }
    RETCHK(dalloc(c, &dM, sizeof(ovl::EigMode) * em.size()));
    HIPCHK(c, hipMemcpyAsync(dM, em.data(), sizeof(ovl::EigMode) * em.size(), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, ovl::launch_eig_modes(g, (const ovl::EigMode *)dM, nm, dA, dTW, norm / std::sqrt((double)n2),
                                    c->field[OVL_PSI], c->field[OVL_PSIMID], c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return OVL_OK;
}

int ovl_get_timing(ovl_ctx *c, ovl_timing *t)
{
    if (!c || !t) return OVL_E_ARG;
    *t = c->last;
    return OVL_OK;
}

int ovl_set_profiling(ovl_ctx *c, int on)
{
    if (!c) return OVL_E_ARG;
    c->profiling = on != 0;
    return OVL_OK;
}

// public traversal numbers (ovl_params.traversal) <-> TRAV_*
static int trav_public(int t)
{
    return t == ovl::TRAV_BOXES ? 2 : (t == ovl::TRAV_MODES ? 3 : (t == ovl::TRAV_QUEUE ? 4 : 1));
}

int ovl_get_tuning(ovl_ctx *c, int precision, ovl_tuning *t)
{
    if (!c || !t || precision < 0 || precision >= NUM_PREC) return OVL_E_ARG;
    t->traversal = trav_public(c->trav[precision]);
    t->slabs = c->slabs[precision];
    t->nt_store = c->nt_out[precision];
    t->waves_per_eu = c->wpe[precision];
    return OVL_OK;
}

int ovl_set_tuning(ovl_ctx *c, int precision, const ovl_tuning *t)
{
    if (!c || !t || precision < 0 || precision >= NUM_PREC) return OVL_E_ARG;
    if (t->traversal < 1 || t->traversal > 4 || t->slabs < 1 || t->waves_per_eu < 1 || t->waves_per_eu > 5 ||
        (t->nt_store != 0 && t->nt_store != 1))
        return fail(c, OVL_E_ARG, "tuning out of range");
    if ((t->traversal == 4) != c->ensemble && t->traversal != 1)
        return fail(c, OVL_E_ARG, "traversals 2 and 3 are for single lattices, 4 (work queues) for ensembles");
    const int trav = t->traversal == 2 ? ovl::TRAV_BOXES
                   : (t->traversal == 3 ? ovl::TRAV_MODES : (t->traversal == 4 ? ovl::TRAV_QUEUE : ovl::TRAV_LEX));
    RETCHK(set_traversal(c, precision, trav, t->slabs));
    c->nt_out[precision] = t->nt_store;
    c->wpe[precision] = t->waves_per_eu;
    c->tuned = true;
    return OVL_OK;
}

// Time of one Clenshaw step of storage precision prec (the pure launch w = h b1, then the epilogue
// b2 = 0.1 h w + 0.01 x + 0.5 b1 - 0.5 b2) with the context's current tuning: three steps after a
// warm-up pass, in ms per step.
static int time_clenshaw_step(ovl_ctx *c, int prec, void *x, void *b1, void *b2, void *w, double *ms_per_step)
{
    float ms = 0;
    for (int rep_i = 0; rep_i < 2; ++rep_i) {  // first pass warms up
        HIPCHK(c, hipEventRecord(c->t0, c->st));
        for (int k = 0; k < 3; ++k) {
            HopArgs a = hop_args(c, prec);
            a.in = b1;
            a.out = w;
            a.sh_r = 1.0;
            RETCHK(hop(c, c->g, prec, prec, a));
            HopArgs q = hop_args(c, prec);
            q.in = w;
            q.out = b2;
            q.sh_r = 0.1;
            q.x = x;
            q.sx_r = 0.01;
            q.b1 = b1;
            q.s1_r = 0.5;
            q.b2 = b2;
            q.s2_r = -0.5;
            RETCHK(hop(c, c->g, prec, prec, q));
        }
        HIPCHK(c, hipEventRecord(c->t1, c->st));
        HIPCHK(c, hipEventSynchronize(c->t1));
        HIPCHK(c, hipEventElapsedTime(&ms, c->t0, c->t1));
    }
    *ms_per_step = ms / 3.0;
    return OVL_OK;
}

int ovl_autotune(ovl_ctx *c, char *report, int len)
{
    // Times the Clenshaw step (pure stencil + epilogue stencil) of every lane traversal (when the
    // context was created with traversal 0; y-slab counts of the mode split around the L2 estimate)
    // and every occupancy variant of the stencil kernel, for every storage precision, on this
    // context's lattice and modes, and keeps the fastest per precision.  Field contents are left
    // unchanged except the work buffers.
    if (!c) return OVL_E_ARG;
    RETCHK(ensure_stencil(c));
    RETCHK(ensure_work(c, c->g.nblk, true, true, true));
    HIPCHK(c, ovl::launch_convert(c->g, c->g.nblk, c->field[OVL_PSI], P64, c->x48, P48, c->st));
    HIPCHK(c, ovl::launch_convert(c->g, c->g.nblk, c->field[OVL_PSIMID], P64, c->a48, P48, c->st));
    HIPCHK(c, ovl::launch_convert(c->g, c->g.nblk, c->field[OVL_PSI], P64, c->b48, P48, c->st));
    HIPCHK(c, ovl::launch_convert(c->g, c->g.nblk, c->field[OVL_PSI], P64, c->x32, P32, c->st));
    HIPCHK(c, ovl::launch_convert(c->g, c->g.nblk, c->field[OVL_PSI], P64, c->x16, P16, c->st, 64.0));
    HIPCHK(c, ovl::launch_convert(c->g, c->g.nblk, c->field[OVL_PSIMID], P64, c->a16, P16, c->st, 64.0));
    HIPCHK(c, ovl::launch_convert(c->g, c->g.nblk, c->field[OVL_PSI], P64, c->b16, P16, c->st, 64.0));
    HIPCHK(c, ovl::launch_convert(c->g, c->g.nblk, c->field[OVL_PSIMID], P64, c->a32, P32, c->st));
    HIPCHK(c, ovl::launch_convert(c->g, c->g.nblk, c->field[OVL_PSI], P64, c->b32, P32, c->st));
    HIPCHK(c, ovl::launch_lincomb(c->g, c->g.nblk, P64, c->a64, c->field[OVL_PSIMID], 1, 0, nullptr, 0, 0, c->st));
    HIPCHK(c, ovl::launch_lincomb(c->g, c->g.nblk, P64, c->b64, c->field[OVL_PSI], 1, 0, nullptr, 0, 0, c->st));
    std::string rep;
    const bool prof = c->profiling;
    c->profiling = false;
    static const char *pname[NUM_PREC] = {"fp64", "fp32", "fp16", "fp48"};
    // occupancy variant 5 reserves 96 KiB of LDS per workgroup (one workgroup per CU)
    int dev = 0, max_lds = 0;
    HIPCHK(c, hipGetDevice(&dev));
    HIPCHK(c, hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const bool one_wg = max_lds >= 96 * 1024;
    auto trav_name = [](const ovl_ctx *cc, int prec) {
        if (cc->trav[prec] == ovl::TRAV_QUEUE) return std::string("queue/") + std::to_string(cc->slabs[prec]);
        if (cc->trav[prec] == ovl::TRAV_MODES) return std::string("modes/") + std::to_string(cc->slabs[prec]);
        if (cc->trav[prec] == ovl::TRAV_BOXES) return std::string("boxes/") + std::to_string(cc->slabs[prec]);
        return std::string("lex");
    };
    for (int prec = 0; prec < NUM_PREC; ++prec) {
        // candidate traversals (trav, slabs)
        std::vector<std::pair<int, int>> cand;
        if (c->ensemble) {
            // the static table, and the work queues claiming 1, 4 or 16 tiles at a time
            cand.push_back({ovl::TRAV_LEX, 1});
            for (int t : {1, 4, 16}) cand.push_back({ovl::TRAV_QUEUE, t});
        } else if (c->prm.traversal != 0) {
            cand.push_back({c->trav[prec], c->slabs[prec]});
        } else {
            if (c->boxes_ok)
                for (int t : {1, 2, 3})
                    cand.push_back({ovl::TRAV_BOXES, t});
            for (int t : {1, 2, 3, 4, 6})
                if (t <= c->g.n1) cand.push_back({ovl::TRAV_MODES, t});
        }
        auto pick = [prec](void *p64, void *p32, void *p16, void *p48) {
            return prec == P64 ? p64 : (prec == P32 ? p32 : (prec == P16 ? p16 : p48));
        };
        void *x = pick(c->field[OVL_PSI], c->x32, c->x16, c->x48);
        void *b1 = pick(c->a64, c->a32, c->a16, c->a48);
        void *b2 = pick(c->b64, c->b32, c->b16, c->b48);
        void *w = pick(c->w64, c->w32, c->w16, c->w48);
        double best = 1e30;
        int best_v = 1, best_nt = 0;
        std::pair<int, int> best_t = cand[0];
        for (const auto &tr : cand) {
            RETCHK(set_traversal(c, prec, tr.first, tr.second));
            for (int nt = 0; nt <= 1; ++nt)
                for (int v = 1; v <= (one_wg ? 5 : 4); ++v) {
                    c->wpe[prec] = v;
                    c->nt_out[prec] = nt;
                    double ms = 0;
                    RETCHK(time_clenshaw_step(c, prec, x, b1, b2, w, &ms));
                    char buf[96];
                    std::snprintf(buf, sizeof buf, "%s %s%s wpe=%d %.1f us; ", pname[prec],
                                  trav_name(c, prec).c_str(), nt ? " nt" : "", v, 1e3 * ms);
                    rep += buf;
                    if (ms < best) {
                        best = ms;
                        best_v = v;
                        best_nt = nt;
                        best_t = {c->trav[prec], c->slabs[prec]};
                    }
                }
        }
        RETCHK(set_traversal(c, prec, best_t.first, best_t.second));
        c->wpe[prec] = best_v;
        c->nt_out[prec] = best_nt;
    }
    c->profiling = prof;
    c->tuned = true;
    if (!c->mb_report.empty()) rep = c->mb_report + "; " + rep;
    rep += "chosen";
    for (int prec = 0; prec < NUM_PREC; ++prec) {
        char buf[96];
        std::snprintf(buf, sizeof buf, "%s %s %s%s wpe=%d", prec == P64 ? "" : ",", pname[prec],
                      trav_name(c, prec).c_str(), c->nt_out[prec] ? " nt" : "", c->wpe[prec]);
        rep += buf;
    }
    if (report && len > 0) std::snprintf(report, len, "%s", rep.c_str());
    return OVL_OK;
}

}  // extern "C"
