// Internal declarations shared by the host driver and the gfx950 kernels.
//
// Device layout of a mode-function field ("mode-block SoA"):
//   modes are grouped in blocks of MB (16, 32 or 64); inside a block the mode index is the
//   fastest-running index so a wavefront reads MB consecutive modes of one site per load.
//   FP64: double2 per complex, element (blk, site, comp, m) at ((blk*V + site)*NCOMP + comp)*MB + m
//   FP32: float4 per two complex,      (blk, site, comp/2, m) at ((blk*V + site)*NCOMP/2 + comp/2)*MB + m
//   FP16: 8 halves per four complex,   (blk, site, comp/4, m) at ((blk*V + site)*NCOMP/4 + comp/4)*MB + m
//   comp = spin*Nc + colour, NCOMP = 4*Nc.
// Stencil links: [mu][site][Nc][Nc] complex, row-major, FP64 (double2) and FP32 (float2) copies,
// each already multiplied by its U(1) phase.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace ovl {

// Division by a launch constant d without a division sequence: for 0 <= n < 2^31,
// n / d = umulhi(n, mul) >> shift with l = ceil(log2 d), mul = ceil(2^(31+l) / d) < 2^32, shift = l - 1
// (Granlund and Montgomery, "Division by invariant integers using multiplication", 1994: m d -
// 2^(31+l) < d <= 2^l); d = 1 is the identity.  On a wave-uniform n the product is one scalar
// instruction.
struct FastDiv {
    unsigned mul;
    int shift;
    unsigned d;
};
inline FastDiv make_fastdiv(unsigned d)
{
    FastDiv f{0u, 0, d};
    if (d <= 1) return f;
    int l = 0;
    while ((1ull << l) < d) ++l;
    f.mul = (unsigned)(((1ull << (31 + l)) + d - 1) / d);
    f.shift = l - 1;
    return f;
}

// One lattice of an ensemble context: its fields are site-blocks [q0, q0 + nblk*V) of the
// concatenated fields (block-major inside the system), its links start at link index l0 (= 3 x
// the number of sites of the systems before it).
struct SysDesc {
    int n0, n1, n2, V;
    int nblk;
    int q0;
    int l0;
    int pad;
    FastDiv dv_V, dv_n01, dv_n0;  // V, n0 n1, n0 (filled before the table goes to the device)
};

struct Geom {
    int nc, n0, n1, n2, V;
    int mb;      // modes per block
    int nblk;    // mode blocks held per field
};

// pointwise epilogue of the stencil kernel:
//   out = sh * h_W(in) + sx * G(x) + s1 * b1 + s2 * b2 + st * tgt      (G = 1 or gamma0)
// the five operands are block-relative pointers (already offset to the first block launched);
// a null pointer drops its term.
struct HopArgs {
    const void *in;
    void *out;
    const void *links;
    const void *x;
    const void *b1;
    const void *b2;
    const void *tgt;
    double sh_r, sh_i, sx_r, sx_i, s1_r, s1_i, s2_r, s2_i, st_r, st_i;
    double amass;     // A = 6r + 2m - 2M of h_W = (1/2) gamma0 [A - hopping]
    int x_gamma0;     // 1: G = gamma0
    int nt_out;       // 1: non-temporal (streaming) stores of `out`
    void *out2;        // optional second output: out2_scale * (sh*h_W(in) + sx*G(x)), before b1/b2/tgt
    double out2_scale;
    int nblk;         // blocks covered by this launch
    int trav;         // lane traversal, TRAV_* (falls back to TRAV_LEX where the launch does not split evenly)
    const int *order; // TRAV_BOXES: XCD k's box is order[k V/8 .. (k+1) V/8); TRAV_MODES: site order
                      // inside a mode block (null = lexicographic)
    // ensemble launch (null = one lattice): system table and per-workgroup work table
    // (int2 {system, workgroup within the system}, system -1 = idle), ens_grid workgroups
    const SysDesc *ens_sys;
    const int2 *ens_work;
    int ens_grid;
    // ensemble work queues (trav == TRAV_QUEUE): the tiles {system, workgroup within the system} of
    // XCD k's systems are ens_qwork[ens_qoff[k] .. ens_qoff[k + 1]); a persistent grid of workgroups
    // serves them, workgroup b the queue of XCD b % 8, claiming ens_per tiles at a time from the
    // counter ens_qctr[k] (ens_qctr[8 + k] counts the workgroups done; the last one of a queue
    // resets both for the next launch)
    const int2 *ens_qwork;
    const int *ens_qoff;
    int *ens_qctr;
    int ens_per;
    // lane coordinates of a single-lattice launch (set by the launcher): V, nblk V / 8, V / 8,
    // n0 n1, n0; lanes are counted in 32 bits (nblk V MB < 2^31)
    FastDiv dv_V, dv_per, dv_v8, dv_n01, dv_n0;
};

// lane traversals of the stencil launch (see lane_coords)
enum { TRAV_LEX = 0, TRAV_BOXES = 1, TRAV_MODES = 2, TRAV_QUEUE = 3 };

// number of XCDs the hardware round-robins workgroups over (MI355X: 8)
constexpr int NUM_XCD = 8;
// threads per workgroup of the stencil kernels
constexpr int STENCIL_TPB = 256;

// field storage precisions; FP16 fields hold values scaled by a power of two chosen by the host
// (the stencil computes in FP32 and the scale is folded into the epilogue coefficients).  P48
// stores a double without its low 16 mantissa bits (sign, exponent, 36 mantissa bits, rounded to
// nearest: unit roundoff 2^-37) and computes in FP64.
enum Prec { P64 = 0, P32 = 1, P16 = 2, P48 = 3 };
constexpr int NUM_PREC = 4;

// launchers (ovl_kernels.hip); all enqueue on `st` and return hipError_t
// wpe: occupancy variant of the stencil kernel (1 = compiler default, 2/3/4 = at least that many
// waves per SIMD)
hipError_t launch_hop(const Geom &g, int pin, int pout, const HopArgs &a, int wpe, hipStream_t st);
// dst = scale * src between precisions: P64->P32, P32->P64, P64->P16, P16->P32, P64->P48,
// P32->P48, P48->P64
hipError_t launch_convert(const Geom &g, int nblk, const void *src, int psrc, void *dst, int pdst, hipStream_t st,
                          double scale = 1.0);
// out = ca * a + cb * b (same precision p), b may be null
hipError_t launch_lincomb(const Geom &g, int nblk, int p, void *out, const void *a, double ca_r, double ca_i,
                          const void *b, double cb_r, double cb_i, hipStream_t st);
hipError_t launch_aos_to_soa(const Geom &g, const double *aos, int first, int count, void *field, hipStream_t st);
hipError_t launch_soa_to_aos(const Geom &g, const void *field, int first, int count, double *aos, hipStream_t st);
hipError_t launch_link_setup(const Geom &g, const double *links, const double *phase, double *ul64, float *ul32,
                             hipStream_t st);
hipError_t launch_ucc_to_links(const Geom &g, const double *ucc, double *ul64, float *ul32, hipStream_t st);
hipError_t launch_link_update(const Geom &g, double *links, const double *efield, double dt, int reunit,
                              hipStream_t st);
hipError_t launch_efield_update(const Geom &g, const double *links, double *efield, double dt, hipStream_t st);
// per-mode inner products <a,b> = sum conj(a) b over one field (FP64 fields), nmodes results
// written as (re, im) pairs to `out` (device); `partial` scratch of partial_len doubles
hipError_t launch_mode_dot(const Geom &g, const void *a, const void *b, double *partial, double *out,
                           hipStream_t st);
size_t mode_dot_partial_len(const Geom &g);
// global observables, written to device `out`
hipError_t launch_energy(const Geom &g, const double *links, const double *efield, double *partial, double *out,
                         hipStream_t st);
// gfield (optional): G^a(x) stored as [site][a]
hipError_t launch_gauss(const Geom &g, const double *links, const double *efield, double *partial, double *out,
                        double *gfield, hipStream_t st);
// E_mu(x) -= gamma (G(x) - U_mu(x) G(x+mu) U_mu(x)^+) from a stored G
hipError_t launch_gauss_restore(const Geom &g, const double *links, const double *gfield, double *efield, double gamma,
                                hipStream_t st);
hipError_t launch_unitarity(const Geom &g, const double *links, double *partial, double *out, hipStream_t st);
hipError_t launch_edotb(const Geom &g, const double *links, const double *efield, double *partial, double *out,
                        hipStream_t st);
size_t reduce_partial_len(const Geom &g);
// per-site currents summed over modes: jv, ja [4*site + mu] (device); psi, hpsi FP64 fields
hipError_t launch_currents(const Geom &g, const void *psi, const void *hpsi, const double *gam, int first_mode,
                           double *partial, double *jv, double *ja, hipStream_t st);
size_t currents_partial_len(const Geom &g);
// multi-shift CG helpers (FP64 fields)
hipError_t launch_mscg_update(const Geom &g, int nshift, double *ps, double *chi, const double *r,
                              const double *coef, const int *active, hipStream_t st);
hipError_t launch_axpy_modes(const Geom &g, void *y, const void *x, const double *coef, int coef_stride,
                             double sign, const int *active, hipStream_t st);

// QED eigenbasis (FP64 fields): unit sources of one batch into a zeroed field (mode mm < cnt:
// source k0 + mm, colour 0 of the z = 0 plane below dz = 4 n0 n1, colour 1 of site 0 above);
// gather of the nsb colour-0 sources of a batch from the z-momentum planes `out` (per mode block,
// n2 planes of n0 n1 4 Nc MB complex) into the column-major dz x dz blocks A; per-workgroup
// partial results (eig_partial_blocks() of them, two per workgroup for the hermitize) of the
// largest colour >= 1 entry of the first nvalid modes, of the Hermitian defect (and a non-finite
// flag) while replacing each block by its Hermitian part, and of a non-finite flag over n doubles
hipError_t launch_eig_sources(const Geom &g, void *field, long k0, int cnt, long dz, hipStream_t st);
hipError_t launch_eig_gather(const Geom &g, const void *out, void *A, long k0, int nsb, hipStream_t st);
hipError_t launch_eig_leak(const Geom &g, const void *field, int nvalid, double *partial, hipStream_t st);
hipError_t launch_eig_hermitize(void *A, long dz, int n2, double *partial, hipStream_t st);
// partial[0..eig_partial_blocks()) = 1 in blocks that saw an Inf/NaN among the n doubles of x, or of
// y when y is not null (the two arrays are read in one pass); x and y 16-byte aligned
hipError_t launch_nonfinite(const void *x, const void *y, long n, double *partial, hipStream_t st);
// one requested mode of the eigenbasis: blk >= 0: f exp(i k_blk z) chi(x, y, spin) on colour 0, chi
// the eigenvector in column idx of z-momentum block blk; blk = -1: the four spin components v at
// site idx, colour col >= 1.  The second field holds the mode times rot.
struct EigMode {
    int blk, idx, col, pad;
    double rot_r, rot_i;
    double v[8];
};
// Psi and PsiMid of nm modes from the eigenvector blocks A (n2 column-major 4 n0 n1 squared blocks)
// and tw[j n2 + z] = exp(-2 pi i j z / n2); writes every entry of both FP64 fields (modes >= nm: 0)
hipError_t launch_eig_modes(const Geom &g, const EigMode *modes, int nm, const void *A, const void *tw, double f,
                            void *psi, void *mid, hipStream_t st);
size_t eig_partial_len();
int eig_partial_blocks();

}  // namespace ovl
