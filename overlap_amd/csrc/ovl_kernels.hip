// gfx950 kernels of the OVERLAP fermion hot path.
//
// All spinor kernels are HBM-bandwidth bound: one lane owns one (site, mode) pair, a wavefront
// covers 64 consecutive (mode, site) lanes of one mode block, so every 16-byte load of a spinor
// component is one contiguous MB*16-byte segment.  Links are shared by all modes of a site and
// read once per wavefront (MB = 64) or once per MB-lane group.
#include "ovl_internal.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <type_traits>
#include <unordered_set>

namespace ovl {

namespace {

constexpr int TPB = STENCIL_TPB;  // threads per block for the streaming kernels

// 48-bit storage of a double: its top 48 bits (sign, exponent, 36 mantissa bits) after rounding
// the dropped 16 bits to nearest.  Only a tag for the templates: the bits live in 16-byte units.
struct f48 { unsigned short h[3]; };
static_assert(sizeof(f48) == 6, "f48 tag size selects the 48-bit layout");

// storage type -> arithmetic type (C) and stencil link type (L).  FP16 storage computes in FP32
// with the FP32 links, 48-bit storage in FP64 with the FP64 links.
template <typename T> struct Vec;
template <> struct Vec<double> { using C = double; using L = double2; };
template <> struct Vec<float> { using C = float; using L = float2; };
template <> struct Vec<_Float16> { using C = float; using L = float2; };
template <> struct Vec<f48> { using C = double; using L = double2; };

// 16 bytes of FP16 storage: 4 complex numbers
typedef _Float16 h8v __attribute__((ext_vector_type(8)));
typedef unsigned u4v __attribute__((ext_vector_type(4)));

// 48-bit units: 4 complex (v0..v7 = re0, im0, .., re3, im3) take three 16-byte units holding
// six 64-bit words W0..W5 (W0, W1 | W2, W3 | W4, W5).  W_i (i < 6) is v_i's sign, exponent and top
// 36 mantissa bits in bits 16..63 and one 16-bit piece of v6 (W0..W2) or v7 (W3..W5) in bits 0..15:
// the piece of W_i, read as the low mantissa bits of v_i, is accounted for when v_i's top bits are
// rounded, so W_i read as a double IS v_i to within 2^-37 (no decoding).  v6 and v7 are their top
// 48 bits (rounded to nearest) reassembled from the three pieces.
__device__ __forceinline__ unsigned long long r48(double v)
{
    // top 48 bits rounded to nearest (+2^15 then truncation, on the sign-magnitude pattern)
    return ((unsigned long long)__double_as_longlong(v) + 0x8000ull) & ~0xffffull;
}
__device__ __forceinline__ unsigned long long w48(double v, unsigned long long piece)
{
    // top 48 bits t of v such that t | piece is the nearest such pattern to v's magnitude
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned long long sgn = b & 0x8000000000000000ull, mag = b ^ sgn;
    const unsigned long long t = (mag < 0x10000ull) ? 0ull : ((mag - piece + 0x8000ull) & ~0xffffull);
    return sgn | t | piece;
}
__device__ __forceinline__ double w_as_double(unsigned lo, unsigned hi)
{
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// v6 / v7 from the low 16 bits of the three words that carry its pieces
__device__ __forceinline__ double pieces_as_double(unsigned w0, unsigned w1, unsigned w2)
{
    // high word: low half of w1 below low half of w2, one byte permute
    return w_as_double(w0 << 16, __builtin_amdgcn_perm(w2, w1, 0x05040100u));
}
// TRUNC: the six words keep their top 48 bits as they are (v_i's low 16 bits are replaced by the
// piece: within 2^-36 instead of 2^-37) and v6 / v7 are truncated to 48 bits; two instructions per
// word instead of about twelve.  Used for the intermediate w = h b of a Clenshaw step.
template <bool TRUNC = false, typename C>
__device__ __forceinline__ void pack48(const C *re, const C *im, int u, u4v &A, u4v &B, u4v &M)
{
    double v[8];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        v[2 * j] = (double)re[4 * u + j];
        v[2 * j + 1] = (double)im[4 * u + j];
    }
    unsigned long long w[6];
    if constexpr (TRUNC) {
        const unsigned long long r6 = (unsigned long long)__double_as_longlong(v[6]);
        const unsigned long long r7 = (unsigned long long)__double_as_longlong(v[7]);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            w[i] = ((unsigned long long)__double_as_longlong(v[i]) & ~0xffffull) | ((r6 >> (16 * (i + 1))) & 0xffffull);
            w[i + 3] =
                ((unsigned long long)__double_as_longlong(v[i + 3]) & ~0xffffull) | ((r7 >> (16 * (i + 1))) & 0xffffull);
        }
    } else {
        const unsigned long long r6 = r48(v[6]), r7 = r48(v[7]);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            w[i] = w48(v[i], (r6 >> (16 * (i + 1))) & 0xffffull);
            w[i + 3] = w48(v[i + 3], (r7 >> (16 * (i + 1))) & 0xffffull);
        }
    }
    A = {(unsigned)w[0], (unsigned)(w[0] >> 32), (unsigned)w[1], (unsigned)(w[1] >> 32)};
    B = {(unsigned)w[2], (unsigned)(w[2] >> 32), (unsigned)w[3], (unsigned)(w[3] >> 32)};
    M = {(unsigned)w[4], (unsigned)(w[4] >> 32), (unsigned)w[5], (unsigned)(w[5] >> 32)};
}
template <typename C>
__device__ __forceinline__ void unpack48(const u4v &A, const u4v &B, const u4v &M, int u, C *re, C *im)
{
    re[4 * u + 0] = w_as_double(A.x, A.y);
    im[4 * u + 0] = w_as_double(A.z, A.w);
    re[4 * u + 1] = w_as_double(B.x, B.y);
    im[4 * u + 1] = w_as_double(B.z, B.w);
    re[4 * u + 2] = w_as_double(M.x, M.y);
    im[4 * u + 2] = w_as_double(M.z, M.w);
    re[4 * u + 3] = pieces_as_double(A.x, A.z, B.x);
    im[4 * u + 3] = pieces_as_double(B.z, M.x, M.z);
}

// ---------------------------------------------------------------------------------------------
// spinor load / store for one lane
// ---------------------------------------------------------------------------------------------
// FP16 layout: 16-byte units of 4 complex, (blk, site, comp/4, m) at ((blk*V + site)*NCOMP/4 + comp/4)*MB + m
// 48-bit layout: three 16-byte units j = A, B, M per 4 complex at
//   (((blk*V + site)*NCOMP/4 + comp/4)*3 + j)*MB + m
template <int NC, int MB, typename T, typename C>
__device__ __forceinline__ void load_sp(const void *__restrict__ base, int V, int blk, int s, int m, C *re, C *im)
{
    constexpr int NCOMP = 4 * NC;
    if constexpr (sizeof(T) == 6) {
        const u4v *p = reinterpret_cast<const u4v *>(base) + ((size_t)(blk * V + s) * (NCOMP / 4)) * 3 * MB + m;
#pragma unroll
        for (int u = 0; u < NCOMP / 4; ++u) {
            const u4v A = p[(size_t)(3 * u) * MB], B = p[(size_t)(3 * u + 1) * MB], M = p[(size_t)(3 * u + 2) * MB];
            unpack48(A, B, M, u, re, im);
        }
    } else if constexpr (sizeof(T) == 2) {
        const h8v *p = reinterpret_cast<const h8v *>(base) + ((size_t)(blk * V + s) * (NCOMP / 4)) * MB + m;
#pragma unroll
        for (int u = 0; u < NCOMP / 4; ++u) {
            const h8v t = p[(size_t)u * MB];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                re[4 * u + j] = (C)t[2 * j];
                im[4 * u + j] = (C)t[2 * j + 1];
            }
        }
    } else if constexpr (sizeof(T) == 8) {
        const double2 *p = reinterpret_cast<const double2 *>(base) + ((size_t)(blk * V + s) * NCOMP) * MB + m;
#pragma unroll
        for (int c = 0; c < NCOMP; ++c) {
            double2 t = p[(size_t)c * MB];
            re[c] = t.x;
            im[c] = t.y;
        }
    } else {
        const float4 *p = reinterpret_cast<const float4 *>(base) + ((size_t)(blk * V + s) * (NCOMP / 2)) * MB + m;
#pragma unroll
        for (int c = 0; c < NCOMP / 2; ++c) {
            float4 t = p[(size_t)c * MB];
            re[2 * c] = t.x;
            im[2 * c] = t.y;
            re[2 * c + 1] = t.z;
            im[2 * c + 1] = t.w;
        }
    }
}

// streaming (non-temporal) variants for operands read or written exactly once per launch, so that
// they do not evict the stencil input whose neighbours are re-read from L2
typedef double d2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));

template <int NC, int MB, typename T, typename C>
__device__ __forceinline__ void load_sp_nt(const void *__restrict__ base, int V, int blk, int s, int m, C *re, C *im)
{
    constexpr int NCOMP = 4 * NC;
    if constexpr (sizeof(T) == 6) {
        const u4v *p = reinterpret_cast<const u4v *>(base) + ((size_t)(blk * V + s) * (NCOMP / 4)) * 3 * MB + m;
#pragma unroll
        for (int u = 0; u < NCOMP / 4; ++u) {
            const u4v A = __builtin_nontemporal_load(p + (size_t)(3 * u) * MB);
            const u4v B = __builtin_nontemporal_load(p + (size_t)(3 * u + 1) * MB);
            const u4v M = __builtin_nontemporal_load(p + (size_t)(3 * u + 2) * MB);
            unpack48(A, B, M, u, re, im);
        }
    } else if constexpr (sizeof(T) == 2) {
        const h8v *p = reinterpret_cast<const h8v *>(base) + ((size_t)(blk * V + s) * (NCOMP / 4)) * MB + m;
#pragma unroll
        for (int u = 0; u < NCOMP / 4; ++u) {
            const h8v t = __builtin_nontemporal_load(p + (size_t)u * MB);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                re[4 * u + j] = (C)t[2 * j];
                im[4 * u + j] = (C)t[2 * j + 1];
            }
        }
    } else if constexpr (sizeof(T) == 8) {
        const d2v *p = reinterpret_cast<const d2v *>(base) + ((size_t)(blk * V + s) * NCOMP) * MB + m;
#pragma unroll
        for (int c = 0; c < NCOMP; ++c) {
            const d2v t = __builtin_nontemporal_load(p + (size_t)c * MB);
            re[c] = t.x;
            im[c] = t.y;
        }
    } else {
        const f4v *p = reinterpret_cast<const f4v *>(base) + ((size_t)(blk * V + s) * (NCOMP / 2)) * MB + m;
#pragma unroll
        for (int c = 0; c < NCOMP / 2; ++c) {
            const f4v t = __builtin_nontemporal_load(p + (size_t)c * MB);
            re[2 * c] = t.x;
            im[2 * c] = t.y;
            re[2 * c + 1] = t.z;
            im[2 * c + 1] = t.w;
        }
    }
}

// the output is read by the very next launch, so it is stored normally (it may stay in the MALL)
template <int NC, int MB, typename T, typename C>
__device__ __forceinline__ h8v pack_h8(const C *re, const C *im, int u)
{
    h8v v;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        v[2 * j] = (_Float16)re[4 * u + j];
        v[2 * j + 1] = (_Float16)im[4 * u + j];
    }
    return v;
}

template <int NC, int MB, typename T, bool TRUNC = false, typename C>
__device__ __forceinline__ void store_sp(void *__restrict__ base, int V, int blk, int s, int m, const C *re,
                                         const C *im)
{
    constexpr int NCOMP = 4 * NC;
    if constexpr (sizeof(T) == 6) {
        u4v *p = reinterpret_cast<u4v *>(base) + ((size_t)(blk * V + s) * (NCOMP / 4)) * 3 * MB + m;
#pragma unroll
        for (int u = 0; u < NCOMP / 4; ++u) {
            u4v A, B, M;
            pack48<TRUNC>(re, im, u, A, B, M);
            p[(size_t)(3 * u) * MB] = A;
            p[(size_t)(3 * u + 1) * MB] = B;
            p[(size_t)(3 * u + 2) * MB] = M;
        }
    } else if constexpr (sizeof(T) == 2) {
        h8v *p = reinterpret_cast<h8v *>(base) + ((size_t)(blk * V + s) * (NCOMP / 4)) * MB + m;
#pragma unroll
        for (int u = 0; u < NCOMP / 4; ++u) p[(size_t)u * MB] = pack_h8<NC, MB, T>(re, im, u);
    } else if constexpr (sizeof(T) == 8) {
        double2 *p = reinterpret_cast<double2 *>(base) + ((size_t)(blk * V + s) * NCOMP) * MB + m;
#pragma unroll
        for (int c = 0; c < NCOMP; ++c) p[(size_t)c * MB] = make_double2(re[c], im[c]);
    } else {
        float4 *p = reinterpret_cast<float4 *>(base) + ((size_t)(blk * V + s) * (NCOMP / 2)) * MB + m;
#pragma unroll
        for (int c = 0; c < NCOMP / 2; ++c)
            p[(size_t)c * MB] = make_float4(re[2 * c], im[2 * c], re[2 * c + 1], im[2 * c + 1]);
    }
}

// streaming store of an output that the launch never reads back (keeps L2 for the stencil input)
template <int NC, int MB, typename T, bool TRUNC = false, typename C>
__device__ __forceinline__ void store_sp_nt(void *__restrict__ base, int V, int blk, int s, int m, const C *re,
                                            const C *im)
{
    constexpr int NCOMP = 4 * NC;
    if constexpr (sizeof(T) == 6) {
        u4v *p = reinterpret_cast<u4v *>(base) + ((size_t)(blk * V + s) * (NCOMP / 4)) * 3 * MB + m;
#pragma unroll
        for (int u = 0; u < NCOMP / 4; ++u) {
            u4v A, B, M;
            pack48<TRUNC>(re, im, u, A, B, M);
            __builtin_nontemporal_store(A, p + (size_t)(3 * u) * MB);
            __builtin_nontemporal_store(B, p + (size_t)(3 * u + 1) * MB);
            __builtin_nontemporal_store(M, p + (size_t)(3 * u + 2) * MB);
        }
    } else if constexpr (sizeof(T) == 2) {
        h8v *p = reinterpret_cast<h8v *>(base) + ((size_t)(blk * V + s) * (NCOMP / 4)) * MB + m;
#pragma unroll
        for (int u = 0; u < NCOMP / 4; ++u) __builtin_nontemporal_store(pack_h8<NC, MB, T>(re, im, u), p + (size_t)u * MB);
    } else if constexpr (sizeof(T) == 8) {
        d2v *p = reinterpret_cast<d2v *>(base) + ((size_t)(blk * V + s) * NCOMP) * MB + m;
#pragma unroll
        for (int c = 0; c < NCOMP; ++c) {
            d2v v = {re[c], im[c]};
            __builtin_nontemporal_store(v, p + (size_t)c * MB);
        }
    } else {
        f4v *p = reinterpret_cast<f4v *>(base) + ((size_t)(blk * V + s) * (NCOMP / 2)) * MB + m;
#pragma unroll
        for (int c = 0; c < NCOMP / 2; ++c) {
            f4v v = {re[2 * c], im[2 * c], re[2 * c + 1], im[2 * c + 1]};
            __builtin_nontemporal_store(v, p + (size_t)c * MB);
        }
    }
}

// Lane -> (mode m, site s, block blk) for the three traversals (TRAV_*).  The hardware hands
// workgroup b to XCD b % 8; the XCD traversals undo that round-robin so that XCD k runs the
// contiguous k-th eighth of the logical lanes:
//   TRAV_BOXES: the eighth is XCD k's spatial box of sites (order[k V/8 .. (k+1) V/8)), every
//     mode block in turn, so x/y/z neighbours are re-read from that XCD's L2 except at the box
//     faces;
//   TRAV_MODES: the eighth is a contiguous range of (mode block, site-order position) in
//     block-major order -- nblk/8 whole blocks when there are at least 8, else 1/(8/nblk) of one
//     block's site order (a z range, or a range of y slabs) -- with sites in the order of `order`
//     (y slabs, z-major) or lexicographic, so few sites are fetched by two XCDs and the
//     z-neighbour planes of a slab stay in L2.
__device__ __forceinline__ unsigned fdiv(unsigned n, const FastDiv &f)
{
    return f.d <= 1 ? n : (__umulhi(n, f.mul) >> f.shift);
}

template <int MB>
__device__ __forceinline__ bool lane_coords(const HopArgs &a, int V, int &m, int &s, int &blk)
{
    const unsigned total = (unsigned)a.nblk * (unsigned)V * MB;  // < 2^31, checked by the launcher
    unsigned lb = blockIdx.x;
    if (a.trav != TRAV_LEX) lb = (blockIdx.x % NUM_XCD) * (gridDim.x / NUM_XCD) + blockIdx.x / NUM_XCD;
    const unsigned gl = lb * TPB + threadIdx.x;
    if (gl >= total) return false;
    m = (int)(gl % MB);
    unsigned r = gl / MB;
    if constexpr (MB == 64) r = __builtin_amdgcn_readfirstlane(r);  // one site per wavefront
    if (a.trav == TRAV_BOXES) {
        const unsigned xc = fdiv(r, a.dv_per);  // XCD of the traversal slot (nblk V / 8 slots each)
        const unsigned rr = r - xc * a.dv_per.d;
        const unsigned b = fdiv(rr, a.dv_v8);
        blk = (int)b;
        s = a.order[xc * a.dv_v8.d + (rr - b * a.dv_v8.d)];
    } else {
        const unsigned b = fdiv(r, a.dv_V);
        const unsigned i = r - b * (unsigned)V;
        blk = (int)b;
        s = (a.trav == TRAV_MODES && a.order) ? a.order[i] : (int)i;
    }
    return true;
}

// ---------------------------------------------------------------------------------------------
// one direction of the Wilson hopping term (spin-projected):
//   forward  (sg=+1): h = u + i sigma_k l,  v = U h,      X_u -= v, X_l += i sigma_k v
//   backward (sg=-1): g = u - i sigma_k l,  v = U^+ g,    X_u -= v, X_l -= i sigma_k v
// which is -(1 +- i gamma_k) U phi  in the Dirac basis of DiracAlgebra.cpp:5-9.
// ---------------------------------------------------------------------------------------------
// a * b + c rounded once: the complex multiply-accumulates of the colour mat-vec as two FMAs per
// real part (the compiler does not contract a sum of two products into a running sum by itself)
__device__ __forceinline__ double fmac(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmac(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// u + S l (S = +-1) rounded once in T.  Operands read from FP16 storage are widened inside one
// mixed-precision FMA, u * one + S l with one = 1 held in a register the compiler cannot see
// through (a literal 1 turns the FMA back into an add of two separately widened values): the same
// value, without the two conversion instructions.
template <int S, typename T, typename TP>
__device__ __forceinline__ T padd(TP u, TP l, float one)
{
    if constexpr (std::is_same<TP, _Float16>::value)
        return S > 0 ? __builtin_fmaf((float)u, one, (float)l) : __builtin_fmaf((float)u, one, -(float)l);
    else
        return u + S * l;
}
__device__ __forceinline__ float opaque_one()
{
    float one;
    asm("s_mov_b32 %0, 1.0" : "=s"(one));
    return one;
}

// TP: element type of the input spinor as loaded (the storage type for FP16, else T)
template <int NC, int K, int SG, typename TP, typename T, typename LT>
__device__ __forceinline__ void hop_dir(const TP *pr, const TP *pi, const LT *__restrict__ U, T *xr, T *xi)
{
    // spinor component index: spin*NC + colour; u = spins 0,1, l = spins 2,3
    T hr[2][NC], hi[2][NC];
    const float one = std::is_same<TP, _Float16>::value ? opaque_one() : 1.0f;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const TP u0r = pr[0 * NC + c], u0i = pi[0 * NC + c];
        const TP u1r = pr[1 * NC + c], u1i = pi[1 * NC + c];
        const TP l0r = pr[2 * NC + c], l0i = pi[2 * NC + c];
        const TP l1r = pr[3 * NC + c], l1i = pi[3 * NC + c];
        if constexpr (K == 0) {  // i sigma_x l = (i l1, i l0)
            hr[0][c] = padd<-SG, T>(u0r, l1i, one); hi[0][c] = padd<SG, T>(u0i, l1r, one);
            hr[1][c] = padd<-SG, T>(u1r, l0i, one); hi[1][c] = padd<SG, T>(u1i, l0r, one);
        } else if constexpr (K == 1) {  // i sigma_y l = (l1, -l0)
            hr[0][c] = padd<SG, T>(u0r, l1r, one); hi[0][c] = padd<SG, T>(u0i, l1i, one);
            hr[1][c] = padd<-SG, T>(u1r, l0r, one); hi[1][c] = padd<-SG, T>(u1i, l0i, one);
        } else {  // i sigma_z l = (i l0, -i l1)
            hr[0][c] = padd<-SG, T>(u0r, l0i, one); hi[0][c] = padd<SG, T>(u0i, l0r, one);
            hr[1][c] = padd<SG, T>(u1r, l1i, one); hi[1][c] = padd<-SG, T>(u1i, l1r, one);
        }
    }
    T vr[2][NC], vi[2][NC];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            T ar = 0, ai = 0;
#pragma unroll
            for (int j = 0; j < NC; ++j) {
                if constexpr (SG > 0) {
                    const LT u = U[i * NC + j];
                    const T ur = (T)u.x, ui = (T)u.y;
                    ar = fmac(ur, hr[s][j], ar);
                    ar = fmac(-ui, hi[s][j], ar);
                    ai = fmac(ur, hi[s][j], ai);
                    ai = fmac(ui, hr[s][j], ai);
                } else {
                    const LT u = U[j * NC + i];  // (U^+)_{ij} = conj(U_{ji})
                    const T ur = (T)u.x, ui = (T)u.y;
                    ar = fmac(ur, hr[s][j], ar);
                    ar = fmac(ui, hi[s][j], ar);
                    ai = fmac(ur, hi[s][j], ai);
                    ai = fmac(-ui, hr[s][j], ai);
                }
            }
            vr[s][i] = ar;
            vi[s][i] = ai;
        }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        xr[0 * NC + c] -= vr[0][c]; xi[0 * NC + c] -= vi[0][c];
        xr[1 * NC + c] -= vr[1][c]; xi[1 * NC + c] -= vi[1][c];
        // X_l += SG * i sigma_k v
        if constexpr (K == 0) {  // (i v1, i v0)
            xr[2 * NC + c] -= SG * vi[1][c]; xi[2 * NC + c] += SG * vr[1][c];
            xr[3 * NC + c] -= SG * vi[0][c]; xi[3 * NC + c] += SG * vr[0][c];
        } else if constexpr (K == 1) {  // (v1, -v0)
            xr[2 * NC + c] += SG * vr[1][c]; xi[2 * NC + c] += SG * vi[1][c];
            xr[3 * NC + c] -= SG * vr[0][c]; xi[3 * NC + c] -= SG * vi[0][c];
        } else {  // (i v0, -i v1)
            xr[2 * NC + c] -= SG * vi[0][c]; xi[2 * NC + c] += SG * vr[0][c];
            xr[3 * NC + c] += SG * vi[1][c]; xi[3 * NC + c] -= SG * vr[1][c];
        }
    }
}

// One (site, mode) lane of the stencil: out = sh*h_W(in) + sx*G(x) + s1*b1 + s2*b2 + st*tgt,
// h_W = (1/2) gamma0 [A in - sum_k (1+i gamma_k) U in(x+k) + (1-i gamma_k) U^+ in(x-k)]
// (WilsonOperator.cpp:30-114 computes -i h_W of one mode; here every mode of every block).
// The site s lives on an n0 x n1 x n2 lattice; its spinors are site-block q0 + s of the fields
// (q0 = first site-block of this lane's mode block), its links start at site index l0.
// EPI = false: a pure launch (out = sh h_W(in)); the pointwise operands are then known to be absent
// at compile time, which frees the registers their arguments would hold
template <int NC, typename TI, typename TO, int MB, bool EPI = true>
__device__ __forceinline__ void hop_site(const HopArgs &a, int n0, int n1, int n2, int s, int x, int y, int z, int q0,
                                         int l0, int m)
{
    constexpr int NCOMP = 4 * NC;
    using LT = typename Vec<TI>::L;
    using CI = typename Vec<TI>::C;  // arithmetic type of the stencil
    using CO = typename Vec<TO>::C;  // arithmetic type of the epilogue
    const int V = n0 * n1 * n2;

    const LT *__restrict__ L = reinterpret_cast<const LT *>(a.links) + (size_t)l0 * NC * NC;
    constexpr int NN = NC * NC;

    // FP16 input stays in FP16 registers until the arithmetic widens it (hop_dir's projection)
    using PT = typename std::conditional<std::is_same<TI, _Float16>::value, _Float16, CI>::type;
    CI xr[NCOMP], xi[NCOMP];
    PT pr[NCOMP], pi[NCOMP];
    load_sp<NC, MB, TI>(a.in, 0, 0, q0 + s, m, pr, pi);
    const CI am = (CI)a.amass;
#pragma unroll
    for (int c = 0; c < NCOMP; ++c) { xr[c] = am * (CI)pr[c]; xi[c] = am * (CI)pi[c]; }

    {  // x direction
        const int sp = s + ((x + 1 == n0) ? 1 - n0 : 1);
        const int sm = s + ((x == 0) ? n0 - 1 : -1);
        load_sp<NC, MB, TI>(a.in, 0, 0, q0 + sp, m, pr, pi);
        hop_dir<NC, 0, 1>(pr, pi, L + ((size_t)0 * V + s) * NN, xr, xi);
        load_sp<NC, MB, TI>(a.in, 0, 0, q0 + sm, m, pr, pi);
        hop_dir<NC, 0, -1>(pr, pi, L + ((size_t)0 * V + sm) * NN, xr, xi);
    }
    {  // y direction
        const int sp = s + ((y + 1 == n1) ? (1 - n1) * n0 : n0);
        const int sm = s + ((y == 0) ? (n1 - 1) * n0 : -n0);
        load_sp<NC, MB, TI>(a.in, 0, 0, q0 + sp, m, pr, pi);
        hop_dir<NC, 1, 1>(pr, pi, L + ((size_t)1 * V + s) * NN, xr, xi);
        load_sp<NC, MB, TI>(a.in, 0, 0, q0 + sm, m, pr, pi);
        hop_dir<NC, 1, -1>(pr, pi, L + ((size_t)1 * V + sm) * NN, xr, xi);
    }
    {  // z direction
        const int sp = s + ((z + 1 == n2) ? (1 - n2) * n0 * n1 : n0 * n1);
        const int sm = s + ((z == 0) ? (n2 - 1) * n0 * n1 : -n0 * n1);
        load_sp<NC, MB, TI>(a.in, 0, 0, q0 + sp, m, pr, pi);
        hop_dir<NC, 2, 1>(pr, pi, L + ((size_t)2 * V + s) * NN, xr, xi);
        load_sp<NC, MB, TI>(a.in, 0, 0, q0 + sm, m, pr, pi);
        hop_dir<NC, 2, -1>(pr, pi, L + ((size_t)2 * V + sm) * NN, xr, xi);
    }

    // h_W = (1/2) gamma0 X, then the epilogue in the output precision
    CO orr[NCOMP], oi[NCOMP];
    const CO shr = (CO)a.sh_r, shi = (CO)a.sh_i;
#pragma unroll
    for (int c = 0; c < NCOMP; ++c) {
        const CO g = (c < 2 * NC) ? (CO)0.5 : (CO)-0.5;
        const CO hr = g * (CO)xr[c], hi = g * (CO)xi[c];
        orr[c] = shr * hr - shi * hi;
        oi[c] = shr * hi + shi * hr;
    }
    const int q = q0 + s;
    if constexpr (EPI) {
    CO tr[NCOMP], ti[NCOMP];
    if (a.x) {
        load_sp_nt<NC, MB, TO>(a.x, 0, 0, q, m, tr, ti);
        const CO cr = (CO)a.sx_r, ci = (CO)a.sx_i;
#pragma unroll
        for (int c = 0; c < NCOMP; ++c) {
            const CO g = (a.x_gamma0 && c >= 2 * NC) ? (CO)-1 : (CO)1;
            const CO vr = g * tr[c], vi = g * ti[c];
            orr[c] += cr * vr - ci * vi;
            oi[c] += cr * vi + ci * vr;
        }
    }
    if (a.out2) {  // sh*h_W(in) + sx*G(x) alone, times out2_scale (the overlap term of a leapfrog axpy)
        CO ur[NCOMP], ui[NCOMP];
        const CO f = (CO)a.out2_scale;
#pragma unroll
        for (int c = 0; c < NCOMP; ++c) { ur[c] = f * orr[c]; ui[c] = f * oi[c]; }
        store_sp<NC, MB, TO>(a.out2, 0, 0, q, m, ur, ui);
    }
    if (a.b1) {
        load_sp_nt<NC, MB, TO>(a.b1, 0, 0, q, m, tr, ti);
        const CO cr = (CO)a.s1_r, ci = (CO)a.s1_i;
#pragma unroll
        for (int c = 0; c < NCOMP; ++c) { orr[c] += cr * tr[c] - ci * ti[c]; oi[c] += cr * ti[c] + ci * tr[c]; }
    }
    if (a.b2) {
        load_sp_nt<NC, MB, TO>(a.b2, 0, 0, q, m, tr, ti);
        const CO cr = (CO)a.s2_r, ci = (CO)a.s2_i;
#pragma unroll
        for (int c = 0; c < NCOMP; ++c) { orr[c] += cr * tr[c] - ci * ti[c]; oi[c] += cr * ti[c] + ci * tr[c]; }
    }
    if (a.tgt) {
        load_sp_nt<NC, MB, TO>(a.tgt, 0, 0, q, m, tr, ti);
        const CO cr = (CO)a.st_r, ci = (CO)a.st_i;
#pragma unroll
        for (int c = 0; c < NCOMP; ++c) { orr[c] += cr * tr[c] - ci * ti[c]; oi[c] += cr * ti[c] + ci * tr[c]; }
    }
    }
    // the pure launch's output w is an intermediate of the step: 48-bit storage truncates it
    if (a.nt_out) store_sp_nt<NC, MB, TO, !EPI>(a.out, 0, 0, q, m, orr, oi);
    else store_sp<NC, MB, TO, !EPI>(a.out, 0, 0, q, m, orr, oi);
}

// WPE: minimum waves per SIMD the register allocation must allow (1 = compiler's choice); the
// stencil hoists its neighbour loads, so this trades memory-level parallelism per wave against
// waves in flight.  The host picks the variant per problem shape (autotune).
template <int NC, typename TI, typename TO, int MB, int WPE, bool EPI>
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(WPE))) void k_hop(HopArgs a, int n0, int n1,
                                                                                       int n2)
{
    const int V = n0 * n1 * n2;
    int m, s, blk;
    if (!lane_coords<MB>(a, V, m, s, blk)) return;
    if constexpr (MB == 64) {
        // one site per wavefront: make the site wave-uniform so that the links (shared by all
        // modes of the site) come through scalar loads instead of 64 identical vector lanes
        s = __builtin_amdgcn_readfirstlane(s);
        blk = __builtin_amdgcn_readfirstlane(blk);
    }
    // site coordinates (s = x + n0 (y + n1 z))
    const int z = (int)fdiv((unsigned)s, a.dv_n01);
    const int rem = s - z * (int)a.dv_n01.d;
    const int y = (int)fdiv((unsigned)rem, a.dv_n0);
    const int x = rem - y * n0;
    hop_site<NC, TI, TO, MB, EPI>(a, n0, n1, n2, s, x, y, z, blk * V, 0, m);
}

// Ensemble launch: many independent lattices of different sizes in one grid.  One tile = TPB lanes
// of one system: workgroup w.y of system w.x (lanes (site, mode) in block-major order, as k_hop's
// lexicographic traversal).
template <int NC, typename TI, typename TO, int MB, bool EPI>
__device__ __forceinline__ void hop_ens_tile(const HopArgs &a, int2 w)
{
    const SysDesc d = a.ens_sys[w.x];
    const unsigned gl = (unsigned)w.y * TPB + threadIdx.x;  // lane within the system
    if (gl >= (unsigned)(d.V * d.nblk * MB)) return;
    const int m = (int)(gl % MB);
    unsigned r = gl / MB;
    if constexpr (MB == 64) r = __builtin_amdgcn_readfirstlane(r);  // one site per wavefront
    const int blk = (int)fdiv(r, d.dv_V);
    const int s = (int)r - blk * d.V;
    const int z = (int)fdiv((unsigned)s, d.dv_n01);
    const int rem = s - z * d.n0 * d.n1;
    const int y = (int)fdiv((unsigned)rem, d.dv_n0);
    hop_site<NC, TI, TO, MB, EPI>(a, d.n0, d.n1, d.n2, s, rem - y * d.n0, y, z, d.q0 + blk * d.V, d.l0, m);
}

// Static table (ens_qctr null): workgroup b runs tile ens_work[b]; the host lays the table out so
// that the tiles of one system all run on one XCD (whole systems per XCD L2, balanced by size).
// Work queues (ens_qctr set): a persistent grid; workgroup b claims ens_per tiles at a time from the
// queue of XCD b % 8 until it is empty, so the tiles of an XCD's systems go to whichever of its
// workgroups is free.  Every workgroup leaves when its queue is exhausted (no workgroup waits on
// another); the last one of each queue to leave resets the queue's counters for the next launch.
template <int NC, typename TI, typename TO, int MB, int WPE, bool EPI>
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(WPE))) void k_hop_ens(HopArgs a)
{
    if (!a.ens_qctr) {
        const int2 w = a.ens_work[blockIdx.x];
        if (w.x >= 0) hop_ens_tile<NC, TI, TO, MB, EPI>(a, w);
        return;
    }
    const int q = (int)(blockIdx.x % NUM_XCD);
    const int beg = a.ens_qoff[q], len = a.ens_qoff[q + 1] - beg;
    __shared__ int claim;
    for (;;) {
        if (threadIdx.x == 0) claim = atomicAdd(&a.ens_qctr[q], 1);
        __syncthreads();
        const int t0 = claim * a.ens_per;
        __syncthreads();  // every lane has read `claim` before the next claim overwrites it
        if (t0 >= len) break;
        const int t1 = min(len, t0 + a.ens_per);
        for (int t = t0; t < t1; ++t) hop_ens_tile<NC, TI, TO, MB, EPI>(a, a.ens_qwork[beg + t]);
    }
    if (threadIdx.x == 0) {
        const int wgs = (int)(gridDim.x / NUM_XCD);
        if (atomicAdd(&a.ens_qctr[NUM_XCD + q], 1) == wgs - 1) {
            atomicExch(&a.ens_qctr[q], 0);
            atomicExch(&a.ens_qctr[NUM_XCD + q], 0);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// pointwise kernels
// ---------------------------------------------------------------------------------------------
// precision conversion between the FP64 (double2) and FP32 (float4) layouts
// precision conversions between the field layouts, dst = scale * src.  A FP32 unit (float4) holds
// complex pairs 2r, 2r+1 and a FP16 unit (8 halves) complex 4r..4r+3 of the FP64 layout, where
// r = unit index / MB; the mode index is unit index % MB in all three.
__global__ __launch_bounds__(TPB) void k_d2f(const double2 *__restrict__ src, float4 *__restrict__ dst, long npairs,
                                             int mb, double scale)
{
    long i = (long)blockIdx.x * TPB + threadIdx.x;
    if (i >= npairs) return;
    const long m = i % mb, q = i / mb;
    const double2 a = src[(2 * q) * mb + m], b = src[(2 * q + 1) * mb + m];
    dst[i] = make_float4((float)(scale * a.x), (float)(scale * a.y), (float)(scale * b.x), (float)(scale * b.y));
}
__global__ __launch_bounds__(TPB) void k_f2d(const float4 *__restrict__ src, double2 *__restrict__ dst, long npairs,
                                             int mb, double scale)
{
    long i = (long)blockIdx.x * TPB + threadIdx.x;
    if (i >= npairs) return;
    const long m = i % mb, q = i / mb;
    const float4 t = src[i];
    dst[(2 * q) * mb + m] = make_double2(scale * t.x, scale * t.y);
    dst[(2 * q + 1) * mb + m] = make_double2(scale * t.z, scale * t.w);
}
__global__ __launch_bounds__(TPB) void k_d2h(const double2 *__restrict__ src, h8v *__restrict__ dst, long nunits,
                                             int mb, double scale)
{
    long i = (long)blockIdx.x * TPB + threadIdx.x;
    if (i >= nunits) return;
    const long m = i % mb, q = i / mb;
    h8v v;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const double2 a = src[(4 * q + j) * mb + m];
        v[2 * j] = (_Float16)(float)(scale * a.x);
        v[2 * j + 1] = (_Float16)(float)(scale * a.y);
    }
    dst[i] = v;
}
__global__ __launch_bounds__(TPB) void k_h2f(const h8v *__restrict__ src, float4 *__restrict__ dst, long nunits,
                                             int mb, float scale)
{
    long i = (long)blockIdx.x * TPB + threadIdx.x;
    if (i >= nunits) return;
    const long m = i % mb, q = i / mb;
    const h8v t = src[i];
    dst[(2 * q) * mb + m] = make_float4(scale * (float)t[0], scale * (float)t[1], scale * (float)t[2], scale * (float)t[3]);
    dst[(2 * q + 1) * mb + m] =
        make_float4(scale * (float)t[4], scale * (float)t[5], scale * (float)t[6], scale * (float)t[7]);
}

// one group of 4 complex numbers (group q of the FP64 numbering: complex 4q..4q+3 of a mode
// block's (site, comp) sequence, mode m) in any of the four layouts, as doubles
template <typename T>
__device__ __forceinline__ void load_group(const void *__restrict__ base, long q, long m, int mb, double *re, double *im)
{
    if constexpr (sizeof(T) == 6) {
        const u4v *p = reinterpret_cast<const u4v *>(base);
        unpack48(p[(3 * q) * mb + m], p[(3 * q + 1) * mb + m], p[(3 * q + 2) * mb + m], 0, re, im);
    } else if constexpr (sizeof(T) == 8) {
        const double2 *p = reinterpret_cast<const double2 *>(base);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const double2 t = p[(4 * q + j) * mb + m];
            re[j] = t.x;
            im[j] = t.y;
        }
    } else if constexpr (sizeof(T) == 4) {
        const float4 *p = reinterpret_cast<const float4 *>(base);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const float4 t = p[(2 * q + j) * mb + m];
            re[2 * j] = t.x;
            im[2 * j] = t.y;
            re[2 * j + 1] = t.z;
            im[2 * j + 1] = t.w;
        }
    } else {
        const h8v t = reinterpret_cast<const h8v *>(base)[q * mb + m];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            re[j] = (double)t[2 * j];
            im[j] = (double)t[2 * j + 1];
        }
    }
}
template <typename T>
__device__ __forceinline__ void store_group(void *__restrict__ base, long q, long m, int mb, const double *re,
                                            const double *im)
{
    if constexpr (sizeof(T) == 6) {
        u4v A, B, M;
        pack48(re, im, 0, A, B, M);
        u4v *p = reinterpret_cast<u4v *>(base);
        p[(3 * q) * mb + m] = A;
        p[(3 * q + 1) * mb + m] = B;
        p[(3 * q + 2) * mb + m] = M;
    } else if constexpr (sizeof(T) == 8) {
        double2 *p = reinterpret_cast<double2 *>(base);
#pragma unroll
        for (int j = 0; j < 4; ++j) p[(4 * q + j) * mb + m] = make_double2(re[j], im[j]);
    } else if constexpr (sizeof(T) == 4) {
        float4 *p = reinterpret_cast<float4 *>(base);
#pragma unroll
        for (int j = 0; j < 2; ++j)
            p[(2 * q + j) * mb + m] =
                make_float4((float)re[2 * j], (float)im[2 * j], (float)re[2 * j + 1], (float)im[2 * j + 1]);
    } else {
        h8v v;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            v[2 * j] = (_Float16)re[j];
            v[2 * j + 1] = (_Float16)im[j];
        }
        reinterpret_cast<h8v *>(base)[q * mb + m] = v;
    }
}
// dst = scale * src, one (group, mode) per lane (the conversions to and from the 48-bit layout)
template <typename TS, typename TD>
__global__ __launch_bounds__(TPB) void k_convert(const void *__restrict__ src, void *__restrict__ dst, long nlanes,
                                                 int mb, double scale)
{
    const long i = (long)blockIdx.x * TPB + threadIdx.x;
    if (i >= nlanes) return;
    const long m = i % mb, q = i / mb;
    double re[4], im[4];
    load_group<TS>(src, q, m, mb, re, im);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        re[j] *= scale;
        im[j] *= scale;
    }
    store_group<TD>(dst, q, m, mb, re, im);
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_lincomb(T *__restrict__ out, const T *__restrict__ a, double car, double cai,
                                                 const T *__restrict__ b, double cbr, double cbi, long ncplx)
{
    long i = (long)blockIdx.x * TPB + threadIdx.x;
    if (i >= ncplx) return;
    const T ar = a[2 * i], ai = a[2 * i + 1];
    T r = (T)car * ar - (T)cai * ai, im = (T)car * ai + (T)cai * ar;
    if (b) {
        const T br = b[2 * i], bi = b[2 * i + 1];
        r += (T)cbr * br - (T)cbi * bi;
        im += (T)cbr * bi + (T)cbi * br;
    }
    out[2 * i] = r;
    out[2 * i + 1] = im;
}

// host AoS (count modes starting at `first`) <-> device SoA FP64
__global__ __launch_bounds__(TPB) void k_aos_to_soa(const double2 *__restrict__ aos, double2 *__restrict__ f, int first,
                                                    int count, int V, int ncomp, int mb)
{
    long i = (long)blockIdx.x * TPB + threadIdx.x;
    long tot = (long)count * V * ncomp;
    if (i >= tot) return;
    // host AoS component = colour*4 + spin; device component = spin*Nc + colour
    const int ca = (int)(i % ncomp);
    const int c = (ca % 4) * (ncomp / 4) + ca / 4;
    const long q = i / ncomp;
    const int s = (int)(q % V);
    const int k = (int)(q / V);
    const int mode = first + k, blk = mode / mb, m = mode % mb;
    f[((size_t)(blk * V + s) * ncomp + c) * mb + m] = aos[i];
}

__global__ __launch_bounds__(TPB) void k_soa_to_aos(const double2 *__restrict__ f, double2 *__restrict__ aos, int first,
                                                    int count, int V, int ncomp, int mb)
{
    long i = (long)blockIdx.x * TPB + threadIdx.x;
    long tot = (long)count * V * ncomp;
    if (i >= tot) return;
    // host AoS component = colour*4 + spin; device component = spin*Nc + colour
    const int ca = (int)(i % ncomp);
    const int c = (ca % 4) * (ncomp / 4) + ca / 4;
    const long q = i / ncomp;
    const int s = (int)(q % V);
    const int k = (int)(q / V);
    const int mode = first + k, blk = mode / mb, m = mode % mb;
    aos[i] = f[((size_t)(blk * V + s) * ncomp + c) * mb + m];
}

// ---------------------------------------------------------------------------------------------
// gauge-link algebra (FP64)
// ---------------------------------------------------------------------------------------------
struct cd { double r, i; };
__device__ __forceinline__ cd cmul(cd a, cd b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ cd cadd(cd a, cd b) { return {a.r + b.r, a.i + b.i}; }

template <int NC> struct Mat { cd e[NC * NC]; };

template <int NC> __device__ __forceinline__ Mat<NC> mmul(const Mat<NC> &a, const Mat<NC> &b)
{
    Mat<NC> c;
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
        for (int j = 0; j < NC; ++j) {
            cd s = {0, 0};
#pragma unroll
            for (int k = 0; k < NC; ++k) s = cadd(s, cmul(a.e[i * NC + k], b.e[k * NC + j]));
            c.e[i * NC + j] = s;
        }
    return c;
}

template <int NC> __device__ __forceinline__ Mat<NC> mdag(const Mat<NC> &a)
{
    Mat<NC> c;
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
        for (int j = 0; j < NC; ++j) c.e[i * NC + j] = {a.e[j * NC + i].r, -a.e[j * NC + i].i};
    return c;
}

template <int NC> __device__ __forceinline__ Mat<NC> mid()
{
    Mat<NC> c;
#pragma unroll
    for (int i = 0; i < NC * NC; ++i) c.e[i] = {0, 0};
#pragma unroll
    for (int i = 0; i < NC; ++i) c.e[i * NC + i] = {1, 0};
    return c;
}

// master link (reference format) -> row-major matrix
template <int NC> __device__ __forceinline__ Mat<NC> get_link(const double *__restrict__ links, int s, int mu)
{
    Mat<NC> m;
    if constexpr (NC == 2) {
        const double *u = links + 4 * (mu + 3 * s);  // GetMatrix, SU2/GroupOperations.cpp:213-225
        m.e[0] = {u[3], u[2]};
        m.e[1] = {u[1], u[0]};
        m.e[2] = {-u[1], u[0]};
        m.e[3] = {u[3], -u[2]};
    } else {
        const double *u = links + 18 * (mu + 3 * s);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) m.e[i * 3 + j] = {u[2 * (i + 3 * j)], u[2 * (i + 3 * j) + 1]};
    }
    return m;
}

// generators t^a as (row, col, value) entries; value = (re, im) times 1/2 (SU(3): lambda_8/(2 sqrt 3))
template <int NC> __device__ __forceinline__ Mat<NC> alg_to_mat(const double *alpha)
{
    Mat<NC> m;
    if constexpr (NC == 2) {  // alpha_a sigma_a / 2
        m.e[0] = {0.5 * alpha[2], 0};
        m.e[3] = {-0.5 * alpha[2], 0};
        m.e[1] = {0.5 * alpha[0], -0.5 * alpha[1]};
        m.e[2] = {0.5 * alpha[0], 0.5 * alpha[1]};
    } else {  // alpha_a lambda_a / 2
        const double r3 = 0.5 / sqrt(3.0);
        m.e[0] = {0.5 * alpha[2] + r3 * alpha[7], 0};
        m.e[4] = {-0.5 * alpha[2] + r3 * alpha[7], 0};
        m.e[8] = {-2.0 * r3 * alpha[7], 0};
        m.e[1] = {0.5 * alpha[0], -0.5 * alpha[1]};
        m.e[3] = {0.5 * alpha[0], 0.5 * alpha[1]};
        m.e[2] = {0.5 * alpha[3], -0.5 * alpha[4]};
        m.e[6] = {0.5 * alpha[3], 0.5 * alpha[4]};
        m.e[5] = {0.5 * alpha[5], -0.5 * alpha[6]};
        m.e[7] = {0.5 * alpha[5], 0.5 * alpha[6]};
    }
    return m;
}

// ReTr(i t^a V) for all generators
template <int NC> __device__ __forceinline__ void retr_it(const Mat<NC> &v, double *out)
{
    // tr(t^a V) = sum_ij t_ij V_ji; ReTr(i X) = -Im tr(X)
    if constexpr (NC == 2) {
        // t0: (0,1)=(1,0)=1/2 -> tr = (V10 + V01)/2
        out[0] = -0.5 * (v.e[2].i + v.e[1].i);
        // t1: (0,1)=-i/2, (1,0)=i/2 -> tr = (-i V10 + i V01)/2 ; Im = (V01.r - V10.r)/2
        out[1] = -0.5 * (v.e[1].r - v.e[2].r);
        out[2] = -0.5 * (v.e[0].i - v.e[3].i);
    } else {
        out[0] = -0.5 * (v.e[3].i + v.e[1].i);
        out[1] = -0.5 * (v.e[1].r - v.e[3].r);
        out[2] = -0.5 * (v.e[0].i - v.e[4].i);
        out[3] = -0.5 * (v.e[6].i + v.e[2].i);
        out[4] = -0.5 * (v.e[2].r - v.e[6].r);
        out[5] = -0.5 * (v.e[7].i + v.e[5].i);
        out[6] = -0.5 * (v.e[5].r - v.e[7].r);
        const double r3 = 0.5 / sqrt(3.0);
        out[7] = -r3 * (v.e[0].i + v.e[4].i - 2.0 * v.e[8].i);
    }
}

// exp(i c alpha^a t^a): scaling and squaring of the Taylor series (FP64)
template <int NC> __device__ Mat<NC> iexp(double c, const double *alpha)
{
    Mat<NC> a = alg_to_mat<NC>(alpha);
    double nrm = 0;
#pragma unroll
    for (int i = 0; i < NC * NC; ++i) {
        a.e[i] = {-c * a.e[i].i, c * a.e[i].r};  // i c A
        nrm += a.e[i].r * a.e[i].r + a.e[i].i * a.e[i].i;
    }
    nrm = sqrt(nrm);
    int sq = 0;
    while (nrm > 0.25 && sq < 64) { nrm *= 0.5; ++sq; }
    const double sc = ldexp(1.0, -sq);
#pragma unroll
    for (int i = 0; i < NC * NC; ++i) { a.e[i].r *= sc; a.e[i].i *= sc; }
    // Horner: I + a(1 + a/2(1 + a/3(...)))
    Mat<NC> r = mid<NC>();
    for (int k = 20; k >= 1; --k) {
        Mat<NC> t = mmul<NC>(a, r);
        const double f = 1.0 / k;
        r = mid<NC>();
#pragma unroll
        for (int i = 0; i < NC * NC; ++i) { r.e[i].r += f * t.e[i].r; r.e[i].i += f * t.e[i].i; }
    }
    for (int k = 0; k < sq; ++k) r = mmul<NC>(r, r);
    return r;
}

__device__ __forceinline__ void store_link3(double *links, int s, int mu, const Mat<3> &m)
{
    double *u = links + 18 * (mu + 3 * s);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            u[2 * (i + 3 * j)] = m.e[i * 3 + j].r;
            u[2 * (i + 3 * j) + 1] = m.e[i * 3 + j].i;
        }
}

// project onto SU(3): the unitary polar factor P = V (V^+ V)^{-1/2} (the U(3) matrix closest to V,
// of largest Re Tr(P^+ V)) by the Newton iteration X <- (X + X^{-+})/2, quadratically convergent
// from any nonsingular V (Higham 1986; six steps take a drift of 1e-3 far below rounding), then
// P det(P)^{-1/3} with the principal cube root so that det = 1.  The reference's SU(3)
// MaxTraceProjection (SU3/GroupOperations.cpp:510-533, Cabibbo-Marinari sweeps over SU(3)) agrees
// to second order in the distance of V from the group.
__device__ __forceinline__ cd cdiv(cd a, cd b)
{
    const double d = 1.0 / (b.r * b.r + b.i * b.i);
    return {(a.r * b.r + a.i * b.i) * d, (a.i * b.r - a.r * b.i) * d};
}
__device__ __forceinline__ cd csub(cd a, cd b) { return {a.r - b.r, a.i - b.i}; }

__device__ Mat<3> reunit3(Mat<3> x)
{
    for (int it = 0; it < 6; ++it) {
        const cd *a = x.e;
        const cd c00 = csub(cmul(a[4], a[8]), cmul(a[5], a[7]));
        const cd c01 = csub(cmul(a[5], a[6]), cmul(a[3], a[8]));
        const cd c02 = csub(cmul(a[3], a[7]), cmul(a[4], a[6]));
        const cd det = cadd(cadd(cmul(a[0], c00), cmul(a[1], c01)), cmul(a[2], c02));
        Mat<3> inv;  // row-major inverse: adjugate / det
        inv.e[0] = cdiv(c00, det);
        inv.e[3] = cdiv(c01, det);
        inv.e[6] = cdiv(c02, det);
        inv.e[1] = cdiv(csub(cmul(a[2], a[7]), cmul(a[1], a[8])), det);
        inv.e[4] = cdiv(csub(cmul(a[0], a[8]), cmul(a[2], a[6])), det);
        inv.e[7] = cdiv(csub(cmul(a[1], a[6]), cmul(a[0], a[7])), det);
        inv.e[2] = cdiv(csub(cmul(a[1], a[5]), cmul(a[2], a[4])), det);
        inv.e[5] = cdiv(csub(cmul(a[2], a[3]), cmul(a[0], a[5])), det);
        inv.e[8] = cdiv(csub(cmul(a[0], a[4]), cmul(a[1], a[3])), det);
        Mat<3> nx;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {  // (X^{-+})_{ij} = conj((X^{-1})_{ji})
                const cd v = x.e[i * 3 + j], w = inv.e[j * 3 + i];
                nx.e[i * 3 + j] = {0.5 * (v.r + w.r), 0.5 * (v.i - w.i)};
            }
        x = nx;
    }
    const cd *a = x.e;
    const cd det = cadd(cadd(cmul(a[0], csub(cmul(a[4], a[8]), cmul(a[5], a[7]))),
                             cmul(a[1], csub(cmul(a[5], a[6]), cmul(a[3], a[8])))),
                        cmul(a[2], csub(cmul(a[3], a[7]), cmul(a[4], a[6]))));
    const double ph = -atan2(det.i, det.r) / 3.0;
    const cd f = {cos(ph), sin(ph)};
    for (int k = 0; k < 9; ++k) x.e[k] = cmul(x.e[k], f);
    return x;
}

// stencil links from master links and U(1) phases (ConvertGaugeLinks,
// OverlapEvolutionOperator.cpp:23-51, with the QED phase folded in)
template <int NC>
__global__ __launch_bounds__(TPB) void k_link_setup(const double *__restrict__ links, const double2 *__restrict__ phase,
                                                    double2 *__restrict__ u64, float2 *__restrict__ u32, int V)
{
    const int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= 3 * V) return;
    const int mu = i / V, s = i - mu * V;
    Mat<NC> m = get_link<NC>(links, s, mu);
    cd ph = {1, 0};
    if (phase) ph = {phase[i].x, phase[i].y};
#pragma unroll
    for (int k = 0; k < NC * NC; ++k) {
        cd v = cmul(ph, m.e[k]);
        u64[(size_t)i * NC * NC + k] = make_double2(v.r, v.i);
        u32[(size_t)i * NC * NC + k] = make_float2((float)v.r, (float)v.i);
    }
}

// stencil links from a host-layout u_cc array (already in [mu*V+s][i][j] order)
__global__ __launch_bounds__(TPB) void k_ucc_setup(const double2 *__restrict__ ucc, double2 *__restrict__ u64,
                                                   float2 *__restrict__ u32, long n)
{
    const long i = (long)blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    const double2 v = ucc[i];
    u64[i] = v;
    u32[i] = make_float2((float)v.x, (float)v.y);
}

// Dynamics.cpp:97-138: U <- exp(i cU E) U with cU = -dt
template <int NC>
__global__ __launch_bounds__(TPB) void k_link_update(double *__restrict__ links, const double *__restrict__ e,
                                                     double dt, int V, int reunit)
{
    const int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= 3 * V) return;
    const int s = i / 3, mu = i - 3 * s;  // link (s, mu) at index mu + 3 s
    constexpr int NG = NC * NC - 1;
    double al[NG];
#pragma unroll
    for (int a = 0; a < NG; ++a) al[a] = e[NG * (mu + 3 * s) + a];
    const double c = -dt;
    if constexpr (NC == 2) {
        // MatrixIExp (SU2/AlgebraOperations.cpp:37-55) and UU (SU2/GroupOperations.cpp:103-110)
        const double d = sqrt((c * al[0]) * (c * al[0]) + (c * al[1]) * (c * al[1]) + (c * al[2]) * (c * al[2]));
        double up[4];
        if (d != 0) {
            const double mm = c * sin(0.5 * d) / d;
            up[0] = mm * al[0]; up[1] = mm * al[1]; up[2] = mm * al[2]; up[3] = cos(0.5 * d);
        } else {
            up[0] = up[1] = up[2] = 0; up[3] = 1;
        }
        double *u = links + 4 * (mu + 3 * s);
        const double o0 = u[0], o1 = u[1], o2 = u[2], o3 = u[3];
        double n0 = up[3] * o0 + up[0] * o3 - (up[1] * o2 - up[2] * o1);
        double n1 = up[3] * o1 + up[1] * o3 + (up[0] * o2 - up[2] * o0);
        double n2 = up[3] * o2 + up[2] * o3 - (up[0] * o1 - up[1] * o0);
        double n3 = up[3] * o3 - up[0] * o0 - up[1] * o1 - up[2] * o2;
        if (reunit) {  // MaxTraceProjection (SU2/GroupOperations.cpp:261-270)
            const double nn = sqrt(n0 * n0 + n1 * n1 + n2 * n2 + n3 * n3);
            n0 /= nn; n1 /= nn; n2 /= nn; n3 /= nn;
        }
        u[0] = n0; u[1] = n1; u[2] = n2; u[3] = n3;
    } else {
        Mat<3> ex = iexp<3>(c, al);
        Mat<3> u = get_link<3>(links, s, mu);
        Mat<3> r = mmul<3>(ex, u);
        if (reunit) r = reunit3(r);
        store_link3(links, s, mu, r);
    }
}

__device__ __forceinline__ int nbr(int s, int dmu, int step, int n0, int n1, int n2)
{
    int z = s / (n0 * n1), rem = s - z * n0 * n1, y = rem / n0, x = rem - y * n0;
    if (dmu == 0) { x = (x + step) % n0; if (x < 0) x += n0; }
    else if (dmu == 1) { y = (y + step) % n1; if (y < 0) y += n1; }
    else { z = (z + step) % n2; if (z < 0) z += n2; }
    return x + n0 * (y + n1 * z);
}

// E_mu(x) += dt * (-2) sum_{nu != mu} [ReTr(i t^a U_{mu nu}(x)) + ReTr(i t^a U_{mu,-nu}(x))]
// (Hamilton's equation of H = 1/2 E^2 + 2 sum ReTr(1 - U_plaq) for U' = -i E U; the
// reference's plaquette macros used by Dynamics.cpp:145-197 are absent)
template <int NC>
__global__ __launch_bounds__(TPB) void k_efield_update(const double *__restrict__ links, double *__restrict__ e,
                                                       double dt, int n0, int n1, int n2)
{
    const int V = n0 * n1 * n2;
    const int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= 3 * V) return;
    const int s = i / 3, mu = i - 3 * s;
    constexpr int NG = NC * NC - 1;
    double f[NG];
#pragma unroll
    for (int a = 0; a < NG; ++a) f[a] = 0;
    const Mat<NC> umu = get_link<NC>(links, s, mu);
    const int sp = nbr(s, mu, 1, n0, n1, n2);
    for (int nu = 0; nu < 3; ++nu) {
        if (nu == mu) continue;
        // forward: U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
        const int snu = nbr(s, nu, 1, n0, n1, n2);
        Mat<NC> p = mmul<NC>(umu, get_link<NC>(links, sp, nu));
        p = mmul<NC>(p, mdag<NC>(get_link<NC>(links, snu, mu)));
        p = mmul<NC>(p, mdag<NC>(get_link<NC>(links, s, nu)));
        double t[NG];
        retr_it<NC>(p, t);
#pragma unroll
        for (int a = 0; a < NG; ++a) f[a] += t[a];
        // backward: U_mu(x) U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu)
        const int smn = nbr(sp, nu, -1, n0, n1, n2);
        const int sm = nbr(s, nu, -1, n0, n1, n2);
        Mat<NC> q = mmul<NC>(umu, mdag<NC>(get_link<NC>(links, smn, nu)));
        q = mmul<NC>(q, mdag<NC>(get_link<NC>(links, sm, mu)));
        q = mmul<NC>(q, get_link<NC>(links, sm, nu));
        retr_it<NC>(q, t);
#pragma unroll
        for (int a = 0; a < NG; ++a) f[a] += t[a];
    }
#pragma unroll
    for (int a = 0; a < NG; ++a) e[NG * (mu + 3 * s) + a] += -2.0 * dt * f[a];
}

// ---------------------------------------------------------------------------------------------
// deterministic reductions: fixed grid, per-block partials, one fixed-order final pass
// ---------------------------------------------------------------------------------------------
constexpr int RED_BLOCKS = 512;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// block reduction of NV values (sum or max per value); result valid in thread 0
template <int NV> __device__ void block_reduce(double *v, const int *is_max)
{
    __shared__ double sh[TPB / 64][NV];
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = is_max[k] ? wave_max(v[k]) : wave_sum(v[k]);
    if (l == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) sh[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double r = sh[0][k];
            for (int j = 1; j < TPB / 64; ++j) r = is_max[k] ? fmax(r, sh[j][k]) : r + sh[j][k];
            v[k] = r;
        }
    }
}

template <int NV>
__global__ __launch_bounds__(TPB) void k_finish(const double *__restrict__ partial, int nblocks, double *out, int m0,
                                                int m1, int m2, int m3)
{
    const int is_max[4] = {m0, m1, m2, m3};
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double r = is_max[k] ? 0.0 : 0.0;
        for (int b = threadIdx.x; b < nblocks; b += TPB) {
            const double p = partial[(size_t)b * NV + k];
            r = is_max[k] ? fmax(r, p) : r + p;
        }
        v[k] = r;
    }
    block_reduce<NV>(v, is_max);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) out[k] = v[k];
}

template <int NC> __device__ double plaq_retr_1mu(const double *links, int s, int mu, int nu, int n0, int n1, int n2)
{
    const int sp = nbr(s, mu, 1, n0, n1, n2), snu = nbr(s, nu, 1, n0, n1, n2);
    Mat<NC> p = mmul<NC>(get_link<NC>(links, s, mu), get_link<NC>(links, sp, nu));
    p = mmul<NC>(p, mdag<NC>(get_link<NC>(links, snu, mu)));
    p = mmul<NC>(p, mdag<NC>(get_link<NC>(links, s, nu)));
    double tr = 0;
#pragma unroll
    for (int i = 0; i < NC; ++i) tr += 1.0 - p.e[i * NC + i].r;
    return tr;
}

// energy: (1/2) sum E^2 and (1/2) sum_planes 4 ReTr(1 - U_plaq)   (BulkObservables.cpp:29-31,84-86)
template <int NC>
__global__ __launch_bounds__(TPB) void k_energy(const double *__restrict__ links, const double *__restrict__ e,
                                                double *partial, int n0, int n1, int n2)
{
    const int V = n0 * n1 * n2;
    constexpr int NG = NC * NC - 1;
    double v[2] = {0, 0};
    for (int s = blockIdx.x * TPB + threadIdx.x; s < V; s += RED_BLOCKS * TPB) {
        for (int k = 0; k < 3 * NG; ++k) { const double q = e[(size_t)3 * NG * s + k]; v[0] += q * q; }
        v[1] += 4.0 * (plaq_retr_1mu<NC>(links, s, 1, 2, n0, n1, n2) + plaq_retr_1mu<NC>(links, s, 2, 0, n0, n1, n2) +
                       plaq_retr_1mu<NC>(links, s, 0, 1, n0, n1, n2));
    }
    v[0] *= 0.5;
    v[1] *= 0.5;
    const int ism[2] = {0, 0};
    block_reduce<2>(v, ism);
    if (threadIdx.x == 0) { partial[2 * blockIdx.x] = v[0]; partial[2 * blockIdx.x + 1] = v[1]; }
}

// components X^a = 2 tr(t^a X) of a Hermitian traceless matrix, via ReTr(i t^a (-i X))
template <int NC> __device__ __forceinline__ void alg_components(const Mat<NC> &X, double *out)
{
    Mat<NC> mx;
#pragma unroll
    for (int k = 0; k < NC * NC; ++k) mx.e[k] = {X.e[k].i, -X.e[k].r};
    retr_it<NC>(mx, out);
#pragma unroll
    for (int a = 0; a < NC * NC - 1; ++a) out[a] *= 2.0;
}

// Gauss law G(x) = sum_mu E_mu(x) - U_mu(x-mu)^+ E_mu(x-mu) U_mu(x-mu), components G^a
template <int NC>
__device__ void gauss_site(const double *__restrict__ links, const double *__restrict__ e, int s, int n0, int n1,
                           int n2, double *comp)
{
    constexpr int NG = NC * NC - 1;
    Mat<NC> G;
#pragma unroll
    for (int k = 0; k < NC * NC; ++k) G.e[k] = {0, 0};
    for (int mu = 0; mu < 3; ++mu) {
        const int sm = nbr(s, mu, -1, n0, n1, n2);
        Mat<NC> e0 = alg_to_mat<NC>(e + NG * (mu + 3 * s));
        Mat<NC> e1 = alg_to_mat<NC>(e + NG * (mu + 3 * sm));
        Mat<NC> u = get_link<NC>(links, sm, mu);
        Mat<NC> t = mmul<NC>(mmul<NC>(mdag<NC>(u), e1), u);
#pragma unroll
        for (int k = 0; k < NC * NC; ++k) G.e[k] = {G.e[k].r + e0.e[k].r - t.e[k].r, G.e[k].i + e0.e[k].i - t.e[k].i};
    }
    alg_components<NC>(G, comp);
}

// max |G^a| and sum G^2 (GaussLawViolation.cpp:13-61); with g != null also stores G^a(x)
template <int NC>
__global__ __launch_bounds__(TPB) void k_gauss(const double *__restrict__ links, const double *__restrict__ e,
                                               double *partial, double *__restrict__ g, int n0, int n1, int n2)
{
    const int V = n0 * n1 * n2;
    constexpr int NG = NC * NC - 1;
    double v[2] = {0, 0};
    for (int s = blockIdx.x * TPB + threadIdx.x; s < V; s += RED_BLOCKS * TPB) {
        double comp[NG];
        gauss_site<NC>(links, e, s, n0, n1, n2, comp);
#pragma unroll
        for (int a = 0; a < NG; ++a) {
            if (g) g[(size_t)NG * s + a] = comp[a];
            const double q = fabs(comp[a]);
            v[0] = fmax(v[0], q);
            v[1] += q * q;
        }
    }
    const int ism[2] = {1, 0};
    block_reduce<2>(v, ism);
    if (threadIdx.x == 0) { partial[2 * blockIdx.x] = v[0]; partial[2 * blockIdx.x + 1] = v[1]; }
}

// one relaxation step of GaussLawRestoration::UpdateElectricFields (GaussRestore.cpp:74-112):
//   E_mu(x) -= gamma (G(x) - U_mu(x) G(x+mu) U_mu(x)^+)
// i.e. gradient descent on (1/2) sum_x G(x)^2 in E
template <int NC>
__global__ __launch_bounds__(TPB) void k_gauss_restore(const double *__restrict__ links, const double *__restrict__ g,
                                                       double *__restrict__ e, double gamma, int n0, int n1, int n2)
{
    const int V = n0 * n1 * n2;
    constexpr int NG = NC * NC - 1;
    const int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= 3 * V) return;
    const int s = i / 3, mu = i - 3 * s;
    const int sp = nbr(s, mu, 1, n0, n1, n2);
    const Mat<NC> u = get_link<NC>(links, s, mu);
    const Mat<NC> t = mmul<NC>(mmul<NC>(u, alg_to_mat<NC>(g + (size_t)NG * sp)), mdag<NC>(u));
    double up[NG];
    alg_components<NC>(t, up);
#pragma unroll
    for (int a = 0; a < NG; ++a) e[(size_t)NG * (mu + 3 * s) + a] -= gamma * (g[(size_t)NG * s + a] - up[a]);
}

template <int NC>
__global__ __launch_bounds__(TPB) void k_unitarity(const double *__restrict__ links, double *partial, int V)
{
    double v[2] = {0, 0};
    for (int i = blockIdx.x * TPB + threadIdx.x; i < 3 * V; i += RED_BLOCKS * TPB) {
        const int s = i / 3, mu = i - 3 * s;
        double nrm;
        if constexpr (NC == 2) {  // UnitarityNorm (SU2/GroupOperations.cpp:172-184)
            const double *u = links + 4 * (mu + 3 * s);
            const double c3 = u[3] * u[3] + u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
            nrm = fabs(c3 - 1.0);
        } else {
            Mat<3> u = get_link<3>(links, s, mu);
            Mat<3> p = mmul<3>(u, mdag<3>(u));
            nrm = 0;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    const double re = p.e[a * 3 + b].r - (a == b ? 1.0 : 0.0), im = p.e[a * 3 + b].i;
                    nrm += re * re + im * im;
                }
            nrm = sqrt(nrm);
        }
        v[0] = fmax(v[0], nrm);
        v[1] += nrm * nrm;
    }
    const int ism[2] = {1, 0};
    block_reduce<2>(v, ism);
    if (threadIdx.x == 0) { partial[2 * blockIdx.x] = v[0]; partial[2 * blockIdx.x + 1] = v[1]; }
}

// improved sum_x E.B (ChernSimonsOperators.cpp:228-439)
template <int NC> __device__ __forceinline__ Mat<NC> L_(const double *links, int s, int mu) { return get_link<NC>(links, s, mu); }

template <int NC> __device__ __forceinline__ Mat<NC> transport(const Mat<NC> &u, const Mat<NC> &q)
{
    return mmul<NC>(mmul<NC>(u, q), mdag<NC>(u));
}

template <int NC>
__device__ __forceinline__ void acc(Mat<NC> &a, const Mat<NC> &b, double w)
{
#pragma unroll
    for (int k = 0; k < NC * NC; ++k) { a.e[k].r += w * b.e[k].r; a.e[k].i += w * b.e[k].i; }
}

template <int NC>
__device__ double edotb_site(const double *links, const double *e, int s, int n0, int n1, int n2)
{
    constexpr int NG = NC * NC - 1;
    double total = 0;
    for (int mu = 0; mu < 3; ++mu) {
        const int a = (mu + 1) % 3, b = (mu + 2) % 3;
        // site at (i along a, j along b)
        auto at = [&](int i, int j) {
            int t = s;
            if (i) t = nbr(t, a, i, n0, n1, n2);
            if (j) t = nbr(t, b, j, n0, n1, n2);
            return t;
        };
        auto atm = [&](int k) { return nbr(s, mu, k, n0, n1, n2); };
        // E_mu transported average
        Mat<NC> ea;
#pragma unroll
        for (int k = 0; k < NC * NC; ++k) ea.e[k] = {0, 0};
        acc<NC>(ea, alg_to_mat<NC>(e + NG * (mu + 3 * s)), 7.0 / 12.0);
        acc<NC>(ea, transport<NC>(L_<NC>(links, s, mu), alg_to_mat<NC>(e + NG * (mu + 3 * atm(1)))), -1.0 / 12.0);
        const Mat<NC> um = mdag<NC>(L_<NC>(links, atm(-1), mu));
        acc<NC>(ea, transport<NC>(um, alg_to_mat<NC>(e + NG * (mu + 3 * atm(-1)))), 7.0 / 12.0);
        const Mat<NC> um2 = mmul<NC>(um, mdag<NC>(L_<NC>(links, atm(-2), mu)));
        acc<NC>(ea, transport<NC>(um2, alg_to_mat<NC>(e + NG * (mu + 3 * atm(-2)))), -1.0 / 12.0);
        // E^a = 2 tr(t^a E) = 2 ReTr(i t^a (-i E))
        Mat<NC> mie;
#pragma unroll
        for (int k = 0; k < NC * NC; ++k) mie.e[k] = {ea.e[k].i, -ea.e[k].r};
        double Ea[NG];
        retr_it<NC>(mie, Ea);

        // B from the (a, b) plane
        const Mat<NC> uaM = mdag<NC>(L_<NC>(links, at(-1, 0), a));
        const Mat<NC> ubM = mdag<NC>(L_<NC>(links, at(0, -1), b));
        const Mat<NC> ua2M = mmul<NC>(uaM, mdag<NC>(L_<NC>(links, at(-2, 0), a)));
        const Mat<NC> ub2M = mmul<NC>(ubM, mdag<NC>(L_<NC>(links, at(0, -2), b)));
        Mat<NC> ub;
#pragma unroll
        for (int k = 0; k < NC * NC; ++k) ub.e[k] = {0, 0};
        auto box = [&](int i, int j, int dag12) {  // UUDD from (i,j) or DDUU
            (void)dag12;
            return mmul<NC>(mmul<NC>(L_<NC>(links, at(i, j), a), L_<NC>(links, at(i + 1, j), b)),
                            mmul<NC>(mdag<NC>(L_<NC>(links, at(i, j + 1), a)), mdag<NC>(L_<NC>(links, at(i, j), b))));
        };
        acc<NC>(ub, box(0, 0, 0), 5.0 / 12.0);
        acc<NC>(ub, transport<NC>(uaM, box(-1, 0, 0)), 5.0 / 12.0);
        {  // lower left: U_a(-1,0)^+ U_b(-1,-1)^+ U_a(-1,-1) U_b(0,-1)
            Mat<NC> p = mmul<NC>(mdag<NC>(L_<NC>(links, at(-1, 0), a)), mdag<NC>(L_<NC>(links, at(-1, -1), b)));
            p = mmul<NC>(p, mmul<NC>(L_<NC>(links, at(-1, -1), a), L_<NC>(links, at(0, -1), b)));
            acc<NC>(ub, p, 5.0 / 12.0);
        }
        acc<NC>(ub, transport<NC>(ubM, box(0, -1, 0)), 5.0 / 12.0);
        // rectangles: six links (offset i, offset j, direction b?, dagger?)
        struct R6 { signed char o[6][2]; unsigned char d[6]; unsigned char g[6]; unsigned char t; };
        const R6 R[8] = {
            {{{0, 0}, {1, 0}, {2, 0}, {1, 1}, {0, 1}, {0, 0}}, {0, 0, 1, 0, 0, 1}, {0, 0, 0, 1, 1, 1}, 0},
            {{{0, 0}, {1, 0}, {1, 1}, {0, 2}, {0, 1}, {0, 0}}, {0, 1, 1, 0, 1, 1}, {0, 0, 0, 1, 1, 1}, 0},
            {{{-1, 0}, {0, 0}, {0, 1}, {-1, 2}, {-1, 1}, {-1, 0}}, {0, 1, 1, 0, 1, 1}, {0, 0, 0, 1, 1, 1}, 1},
            {{{-2, 0}, {-1, 0}, {0, 0}, {-1, 1}, {-2, 1}, {-2, 0}}, {0, 0, 1, 0, 0, 1}, {0, 0, 0, 1, 1, 1}, 2},
            {{{-1, 0}, {-1, -1}, {-1, -2}, {-1, -2}, {0, -2}, {0, -1}}, {0, 1, 1, 0, 1, 1}, {1, 1, 1, 0, 0, 0}, 0},
            {{{-1, 0}, {-2, 0}, {-2, -1}, {-2, -1}, {-1, -1}, {0, -1}}, {0, 0, 1, 0, 0, 1}, {1, 1, 1, 0, 0, 0}, 0},
            {{{0, -1}, {1, -1}, {2, -1}, {1, 0}, {0, 0}, {0, -1}}, {0, 0, 1, 0, 0, 1}, {0, 0, 0, 1, 1, 1}, 3},
            {{{0, -2}, {1, -2}, {1, -1}, {0, 0}, {0, -1}, {0, -2}}, {0, 1, 1, 0, 1, 1}, {0, 0, 0, 1, 1, 1}, 4},
        };
        for (int r = 0; r < 8; ++r) {
            Mat<NC> p = mid<NC>();
            for (int k = 0; k < 6; ++k) {
                Mat<NC> l = L_<NC>(links, at(R[r].o[k][0], R[r].o[k][1]), R[r].d[k] ? b : a);
                p = mmul<NC>(p, R[r].g[k] ? mdag<NC>(l) : l);
            }
            if (R[r].t == 1) p = transport<NC>(uaM, p);
            if (R[r].t == 2) p = transport<NC>(ua2M, p);
            if (R[r].t == 3) p = transport<NC>(ubM, p);
            if (R[r].t == 4) p = transport<NC>(ub2M, p);
            acc<NC>(ub, p, -1.0 / 24.0);
        }
        double Bb[NG];
        retr_it<NC>(ub, Bb);
#pragma unroll
        for (int c = 0; c < NG; ++c) total += 2.0 * Ea[c] * 2.0 * Bb[c];
    }
    return total;
}

template <int NC>
__global__ __launch_bounds__(TPB) void k_edotb(const double *__restrict__ links, const double *__restrict__ e,
                                               double *partial, int n0, int n1, int n2)
{
    const int V = n0 * n1 * n2;
    double v[1] = {0};
    for (int s = blockIdx.x * TPB + threadIdx.x; s < V; s += RED_BLOCKS * TPB)
        v[0] += edotb_site<NC>(links, e, s, n0, n1, n2);
    const int ism[1] = {0};
    block_reduce<1>(v, ism);
    if (threadIdx.x == 0) partial[blockIdx.x] = v[0];
}

// per-mode inner products: partial[chunk][mode] then a fixed-order sum over chunks
constexpr int DOT_CHUNKS = 128;

template <int NC, int MB>
__global__ __launch_bounds__(TPB) void k_mode_dot(const double2 *__restrict__ a, const double2 *__restrict__ b,
                                                  double2 *__restrict__ partial, int V, int nblk)
{
    // grid: (DOT_CHUNKS, nblk); each thread: one mode (m) and a strided subset of (site, comp)
    constexpr int NCOMP = 4 * NC;
    constexpr int LPM = TPB / MB;  // lanes per mode inside the block
    const int m = threadIdx.x % MB, lane = threadIdx.x / MB;
    const int blk = blockIdx.y;
    const long nel = (long)V * NCOMP;
    double sr = 0, si = 0;
    for (long q = (long)blockIdx.x * LPM + lane; q < nel; q += (long)DOT_CHUNKS * LPM) {
        const size_t idx = ((size_t)blk * nel + q) * MB + m;
        const double2 x = a[idx], y = b[idx];
        sr += x.x * y.x + x.y * y.y;
        si += x.x * y.y - x.y * y.x;
    }
    __shared__ double2 sh[TPB];
    sh[threadIdx.x] = make_double2(sr, si);
    __syncthreads();
    if (lane == 0) {
        double rr = 0, ii = 0;
        for (int l = 0; l < LPM; ++l) { rr += sh[l * MB + m].x; ii += sh[l * MB + m].y; }
        partial[((size_t)blockIdx.x) * nblk * MB + (size_t)blk * MB + m] = make_double2(rr, ii);
    }
}

__global__ __launch_bounds__(TPB) void k_mode_dot_finish(const double2 *__restrict__ partial, double2 *__restrict__ out,
                                                         int nmodes_pad)
{
    const int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= nmodes_pad) return;
    double r = 0, im = 0;
    for (int c = 0; c < DOT_CHUNKS; ++c) { r += partial[(size_t)c * nmodes_pad + i].x; im += partial[(size_t)c * nmodes_pad + i].y; }
    out[i] = make_double2(r, im);
}

// currents (OverlapCurrents.cpp:137-228), per (site, mode) bilinears summed over the modes of a block.
// gam: [kind][mu][a][b] complex, kind 0 = g0 g^mu, 1 = g0 g^mu g5, 2 = -(i/2) g0 g^mu g0,
// 3 = -(i/2) g0 g^mu g5 g0 (built on the host from the Dirac matrices as the reference does)
template <int NC, int MB>
__global__ __launch_bounds__(TPB) void k_currents(const double2 *__restrict__ psi, const double2 *__restrict__ hpsi,
                                                  const double2 *__restrict__ gam, double *__restrict__ partial, int V,
                                                  int nblk, int first_mode)
{
    constexpr int SPB = TPB / MB;
    const int m = threadIdx.x % MB, sl = threadIdx.x / MB;
    const long site_g = (long)blockIdx.x * SPB + sl;  // over nblk * V
    const bool ok = site_g < (long)nblk * V;
    const int blk = ok ? (int)(site_g / V) : 0, s = ok ? (int)(site_g % V) : 0;
    double j[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (ok) {
        const double w = (((blk * MB + m) + first_mode) % 2 == 0) ? 1.0 : -1.0;
        for (int c = 0; c < NC; ++c) {
            double2 p[4], h[4];
            for (int sp = 0; sp < 4; ++sp) {
                const size_t idx = ((size_t)(blk * V + s) * 4 * NC + sp * NC + c) * MB + m;
                p[sp] = psi[idx];
                h[sp] = hpsi[idx];
            }
            for (int mu = 0; mu < 4; ++mu)
                for (int a = 0; a < 4; ++a)
                    for (int b = 0; b < 4; ++b) {
                        // conj(p_a) p_b and conj(p_a) h_b
                        const double ppr = p[a].x * p[b].x + p[a].y * p[b].y, ppi = p[a].x * p[b].y - p[a].y * p[b].x;
                        const double phr = p[a].x * h[b].x + p[a].y * h[b].y, phi = p[a].x * h[b].y - p[a].y * h[b].x;
                        const double2 gv = gam[((0 * 4 + mu) * 4 + a) * 4 + b], ga = gam[((1 * 4 + mu) * 4 + a) * 4 + b];
                        const double2 gv2 = gam[((2 * 4 + mu) * 4 + a) * 4 + b], ga2 = gam[((3 * 4 + mu) * 4 + a) * 4 + b];
                        j[mu] += w * (gv.x * ppr - gv.y * ppi + gv2.x * phr - gv2.y * phi);
                        j[4 + mu] += w * (ga.x * ppr - ga.y * ppi + ga2.x * phr - ga2.y * phi);
                    }
        }
    }
    // sum over the MB modes of this site (lanes of one site are contiguous), fixed order
    __shared__ double sh[TPB][8];
    for (int k = 0; k < 8; ++k) sh[threadIdx.x][k] = j[k];
    __syncthreads();
    if (m == 0 && ok) {
        double t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int q = 0; q < MB; ++q)
            for (int k = 0; k < 8; ++k) t[k] += sh[sl * MB + q][k];
        for (int k = 0; k < 8; ++k) partial[((size_t)blk * V + s) * 8 + k] = t[k];
    }
}

__global__ __launch_bounds__(TPB) void k_currents_finish(const double *__restrict__ partial, double *jv, double *ja,
                                                         int V, int nblk, double pre)
{
    const int s = blockIdx.x * TPB + threadIdx.x;
    if (s >= V) return;
    double t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = 0; b < nblk; ++b)
        for (int k = 0; k < 8; ++k) t[k] += partial[((size_t)b * V + s) * 8 + k];
    for (int k = 0; k < 4; ++k) { jv[4 * s + k] = pre * t[k]; ja[4 * s + k] = pre * t[4 + k]; }
}

// multi-shift CG vector update (OverlapU.cpp:164-177), with chi += bz_k beta_k p_k accumulated
// directly:  for each shift k: chi += cw_k p_k ; p_k = a_k p_k + g_k r   (per-mode coefficients)
__global__ __launch_bounds__(TPB) void k_mscg_update2(int nshift, double2 *__restrict__ ps, double2 *__restrict__ chi,
                                                      const double2 *__restrict__ r, const double *__restrict__ coef,
                                                      const int *__restrict__ active, long per_blk, int nblk, int mb)
{
    // element i = (blk, q, m) with per_blk = V*NCOMP elements per mode block
    const long i = (long)blockIdx.x * TPB + threadIdx.x;
    const long nel = per_blk * nblk * mb;
    if (i >= nel) return;
    const int m = (int)(i % mb);
    const int blk = (int)(i / ((long)per_blk * mb));
    const int mode = blk * mb + m;
    if (!active[mode]) return;
    const double *c = coef + (size_t)mode * 3 * nshift;
    const double2 rv = r[i];
    double2 x = chi[i];
    for (int k = 0; k < nshift; ++k) {
        double2 p = ps[(size_t)k * nel + i];
        x.x += c[3 * k] * p.x;
        x.y += c[3 * k] * p.y;
        p.x = c[3 * k + 1] * p.x + c[3 * k + 2] * rv.x;
        p.y = c[3 * k + 1] * p.y + c[3 * k + 2] * rv.y;
        ps[(size_t)k * nel + i] = p;
    }
    chi[i] = x;
}

// y += sign * coef[mode*stride] * x for active modes (FP64 fields)
__global__ __launch_bounds__(TPB) void k_axpy_modes(double2 *__restrict__ y, const double2 *__restrict__ x,
                                                    const double *__restrict__ coef, int stride, double sign,
                                                    const int *__restrict__ active, long per_blk, int nblk, int mb)
{
    const long i = (long)blockIdx.x * TPB + threadIdx.x;
    const long nel = per_blk * nblk * mb;
    if (i >= nel) return;
    const int m = (int)(i % mb);
    const int blk = (int)(i / ((long)per_blk * mb));
    const int mode = blk * mb + m;
    if (active && !active[mode]) return;
    const double c = sign * coef[(size_t)mode * stride];
    const double2 xv = x[i];
    double2 yv = y[i];
    yv.x += c * xv.x;
    yv.y += c * xv.y;
    y[i] = yv;
}

inline unsigned nblocks(long n) { return (unsigned)((n + TPB - 1) / TPB); }

// occupancy variant 5: the wpe = 1 kernel with an LDS reservation it does not use, larger than
// half of a CU's 160 KiB, so that one workgroup (one wave per SIMD) runs per CU: fewer sites in
// flight per XCD, so the neighbour planes of the stencil input fit in L2 beside them
constexpr unsigned LDS_ONE_WG = 96 * 1024;
hipError_t reserve_lds(const void *kernel)
{
    static std::mutex mu;
    static std::unordered_set<const void *> done;
    std::lock_guard<std::mutex> lock(mu);
    if (!done.insert(kernel).second) return hipSuccess;
    return hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, LDS_ONE_WG);
}

template <int NC, typename TI, typename TO, int MB, bool EPI>
void hop_wpe_epi(const HopArgs &b, unsigned nb, int wpe, const Geom &g, hipStream_t st)
{
    switch (wpe) {
    case 2: k_hop<NC, TI, TO, MB, 2, EPI><<<nb, TPB, 0, st>>>(b, g.n0, g.n1, g.n2); break;
    case 3: k_hop<NC, TI, TO, MB, 3, EPI><<<nb, TPB, 0, st>>>(b, g.n0, g.n1, g.n2); break;
    case 4: k_hop<NC, TI, TO, MB, 4, EPI><<<nb, TPB, 0, st>>>(b, g.n0, g.n1, g.n2); break;
    case 5:
        (void)reserve_lds(reinterpret_cast<const void *>(k_hop<NC, TI, TO, MB, 1, EPI>));
        k_hop<NC, TI, TO, MB, 1, EPI><<<nb, TPB, LDS_ONE_WG, st>>>(b, g.n0, g.n1, g.n2);
        break;
    default: k_hop<NC, TI, TO, MB, 1, EPI><<<nb, TPB, 0, st>>>(b, g.n0, g.n1, g.n2); break;
    }
}

template <int NC, typename TI, typename TO, int MB>
void hop_wpe(const HopArgs &b, unsigned nb, int wpe, const Geom &g, hipStream_t st)
{
    // the pure launch of the Clenshaw step (no pointwise operand) has a variant of its own
    if (b.x || b.b1 || b.b2 || b.tgt || b.out2) hop_wpe_epi<NC, TI, TO, MB, true>(b, nb, wpe, g, st);
    else hop_wpe_epi<NC, TI, TO, MB, false>(b, nb, wpe, g, st);
}

template <int NC, typename TI, typename TO, int MB, bool EPI>
void hop_ens_wpe_epi(const HopArgs &b, int wpe, hipStream_t st)
{
    unsigned nb = (unsigned)b.ens_grid;
    if (b.ens_qctr) {
        // persistent grid: the resident workgroups of the variant on every CU, a multiple of the XCDs
        int dev = 0, cus = 0, per_cu = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
        const void *f = wpe == 2 ? reinterpret_cast<const void *>(k_hop_ens<NC, TI, TO, MB, 2, EPI>)
                      : wpe == 3 ? reinterpret_cast<const void *>(k_hop_ens<NC, TI, TO, MB, 3, EPI>)
                      : wpe == 4 ? reinterpret_cast<const void *>(k_hop_ens<NC, TI, TO, MB, 4, EPI>)
                                 : reinterpret_cast<const void *>(k_hop_ens<NC, TI, TO, MB, 1, EPI>);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, TPB, wpe == 5 ? LDS_ONE_WG : 0) != hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        nb = (unsigned)std::max(NUM_XCD, (cus * per_cu) / NUM_XCD * NUM_XCD);
    }
    switch (wpe) {
    case 2: k_hop_ens<NC, TI, TO, MB, 2, EPI><<<nb, TPB, 0, st>>>(b); break;
    case 3: k_hop_ens<NC, TI, TO, MB, 3, EPI><<<nb, TPB, 0, st>>>(b); break;
    case 4: k_hop_ens<NC, TI, TO, MB, 4, EPI><<<nb, TPB, 0, st>>>(b); break;
    case 5:
        (void)reserve_lds(reinterpret_cast<const void *>(k_hop_ens<NC, TI, TO, MB, 1, EPI>));
        k_hop_ens<NC, TI, TO, MB, 1, EPI><<<nb, TPB, LDS_ONE_WG, st>>>(b);
        break;
    default: k_hop_ens<NC, TI, TO, MB, 1, EPI><<<nb, TPB, 0, st>>>(b); break;
    }
}
template <int NC, typename TI, typename TO, int MB>
void hop_ens_wpe(const HopArgs &b, int wpe, hipStream_t st)
{
    // the pure launch of the Clenshaw step (no pointwise operand) has a variant of its own
    if (b.x || b.b1 || b.b2 || b.tgt || b.out2) hop_ens_wpe_epi<NC, TI, TO, MB, true>(b, wpe, st);
    else hop_ens_wpe_epi<NC, TI, TO, MB, false>(b, wpe, st);
}

template <int NC, typename TI, typename TO>
hipError_t hop_mb(const Geom &g, const HopArgs &a, int wpe, hipStream_t st)
{
    if (a.ens_sys) {
        if (a.ens_grid <= 0) return hipErrorInvalidValue;
        switch (g.mb) {
        case 16: hop_ens_wpe<NC, TI, TO, 16>(a, wpe, st); break;
        case 32: hop_ens_wpe<NC, TI, TO, 32>(a, wpe, st); break;
        case 64: hop_ens_wpe<NC, TI, TO, 64>(a, wpe, st); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    const long total = (long)a.nblk * g.V * g.mb;
    if (total >= (1l << 31)) return hipErrorInvalidValue;  // lanes are counted in 32 bits
    const unsigned nb = nblocks(total);
    HopArgs b = a;
    b.dv_V = make_fastdiv((unsigned)g.V);
    b.dv_per = make_fastdiv((unsigned)((long)a.nblk * g.V / NUM_XCD));
    b.dv_v8 = make_fastdiv((unsigned)(g.V / NUM_XCD));
    b.dv_n01 = make_fastdiv((unsigned)(g.n0 * g.n1));
    b.dv_n0 = make_fastdiv((unsigned)g.n0);
    // the XCD traversals need whole workgroups spread evenly over the XCDs
    const bool even = total % ((long)TPB * NUM_XCD) == 0;
    if (b.trav == TRAV_BOXES && (!even || g.V % NUM_XCD != 0 || !b.order)) b.trav = TRAV_LEX;
    if (b.trav == TRAV_MODES && !even) b.trav = TRAV_LEX;
    switch (g.mb) {
    case 16: hop_wpe<NC, TI, TO, 16>(b, nb, wpe, g, st); break;
    case 32: hop_wpe<NC, TI, TO, 32>(b, nb, wpe, g, st); break;
    case 64: hop_wpe<NC, TI, TO, 64>(b, nb, wpe, g, st); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// QED eigenbasis: unit sources, z-momentum blocks and their checks
// ---------------------------------------------------------------------------------------------
// One batch of unit sources in a zeroed FP64 field: mode mm < cnt gets a 1 at source src = k0 + mm,
// i.e. colour 0, spin src % 4 at site src / 4 of the z = 0 plane when src < dz, else colour 1,
// spin src - dz at site 0.
template <int NC, int MB>
__global__ __launch_bounds__(TPB) void k_eig_sources(double2 *field, int V, long k0, int cnt, long dz)
{
    constexpr int NCOMP = 4 * NC;
    const int mm = blockIdx.x * TPB + threadIdx.x;
    if (mm >= cnt) return;
    const long src = k0 + mm;
    const int blk = mm / MB, m = mm - blk * MB;
    const long site = src < dz ? src / 4 : 0;
    const int comp = src < dz ? (int)(src % 4) * NC : (int)(src - dz) * NC + 1;
    field[(((size_t)blk * V + site) * NCOMP + comp) * MB + m] = make_double2(1.0, 0.0);
}

// A[j dz^2 + (k0 + mm) dz + row] = out[blk][j ncol + (xy NCOMP + sp NC) MB + m] for row = 4 xy + sp
// and the first nsb modes mm (blk = mm / MB, m = mm % MB) of a batch; `out` holds per mode block
// the n2 z-momentum planes of ncol = plane NCOMP MB entries each
template <int NC, int MB>
__global__ __launch_bounds__(TPB) void k_eig_gather(const double2 *__restrict__ out, double2 *__restrict__ A,
                                                    long plane, int n2, long k0, int nsb)
{
    constexpr int NCOMP = 4 * NC;
    const long dz = 4 * plane, ncol = plane * NCOMP * MB;
    const long total = (long)n2 * nsb * dz;
    for (long i = (long)blockIdx.x * TPB + threadIdx.x; i < total; i += (long)gridDim.x * TPB) {
        const long row = i % dz, r2 = i / dz;
        const int mm = (int)(r2 % nsb), j = (int)(r2 / nsb);
        const int blk = mm / MB, m = mm - blk * MB;
        const long xy = row / 4;
        const int sp = (int)(row % 4);
        A[(size_t)j * dz * dz + (size_t)(k0 + mm) * dz + row] =
            out[((size_t)blk * n2 + j) * ncol + (size_t)(xy * NCOMP + sp * NC) * MB + m];
    }
}

__device__ __forceinline__ double abs_or_inf(double2 x)
{
    const double a = sqrt(x.x * x.x + x.y * x.y);
    return (a <= 1.7976931348623157e308) ? a : __builtin_inf();  // NaN and inf count as inf
}

// largest |entry| of colours >= 1 among the first nvalid modes of an FP64 field (non-finite entries
// count as inf); one partial maximum per workgroup
template <int NC, int MB>
__global__ __launch_bounds__(TPB) void k_eig_leak(const double2 *__restrict__ f, int V, int nblk, int nvalid,
                                                  double *partial)
{
    constexpr int NCOMP = 4 * NC;
    double v[1] = {0.0};
    const long total = (long)nblk * V * NCOMP * MB;
    for (long i = (long)blockIdx.x * TPB + threadIdx.x; i < total; i += (long)gridDim.x * TPB) {
        const int m = (int)(i % MB);
        const long r = i / MB;
        const int comp = (int)(r % NCOMP), blk = (int)(r / NCOMP / V);
        if (comp % NC == 0 || blk * MB + m >= nvalid) continue;
        v[0] = fmax(v[0], abs_or_inf(f[i]));
    }
    const int ism[1] = {1};
    block_reduce<1>(v, ism);
    if (threadIdx.x == 0) partial[blockIdx.x] = v[0];
}

// Hermitian part of each of the n2 column-major dz x dz blocks, in place: the pair (row, col),
// (col, row) with row >= col is handled by one thread.  partial[2 b] = largest |A_rc - conj(A_cr)|,
// partial[2 b + 1] = 1 when an entry is not finite.
__global__ __launch_bounds__(TPB) void k_eig_hermitize(double2 *A, long dz, int n2, double *partial)
{
    double v[2] = {0.0, 0.0};
    const long total = (long)n2 * dz * dz;
    for (long i = (long)blockIdx.x * TPB + threadIdx.x; i < total; i += (long)gridDim.x * TPB) {
        const long row = i % dz, r2 = i / dz;
        const long col = r2 % dz, j = r2 / dz;
        if (row < col) continue;
        double2 *b = A + (size_t)j * dz * dz;
        const double2 x = b[(size_t)col * dz + row], y = b[(size_t)row * dz + col];
        const double2 d = make_double2(x.x - y.x, x.y + y.y);  // x - conj(y)
        const double a = abs_or_inf(d);
        if (a > 1.7976931348623157e308 || abs_or_inf(x) > 1.7976931348623157e308 ||
            abs_or_inf(y) > 1.7976931348623157e308)
            v[1] = 1.0;
        else
            v[0] = fmax(v[0], a);
        const double2 h = make_double2(0.5 * (x.x + y.x), 0.5 * (x.y - y.y));
        b[(size_t)col * dz + row] = h;
        b[(size_t)row * dz + col] = make_double2(h.x, -h.y);
    }
    const int ism[2] = {1, 1};
    block_reduce<2>(v, ism);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = v[0];
        partial[2 * blockIdx.x + 1] = v[1];
    }
}

// Psi and PsiMid of the requested modes, every entry of both fields (modes >= nm are zero): mode
// e.blk >= 0 is f conj(tw[e.blk n2 + z]) A[e.blk dz^2 + e.idx dz + 4 xy + spin] on colour 0 (site =
// z plane + xy), mode e.blk = -1 is e.v at site e.idx, colour e.col; PsiMid = Psi rot
template <int NC, int MB>
__global__ __launch_bounds__(TPB) void k_eig_modes(const EigMode *__restrict__ modes, int nm,
                                                   const double2 *__restrict__ A, const double2 *__restrict__ tw,
                                                   double f, int V, long plane, int n2, int nblk,
                                                   double2 *__restrict__ psi, double2 *__restrict__ mid)
{
    constexpr int NCOMP = 4 * NC;
    const long dz = 4 * plane;
    const long total = (long)nblk * V * NCOMP * MB;
    for (long i = (long)blockIdx.x * TPB + threadIdx.x; i < total; i += (long)gridDim.x * TPB) {
        const int m = (int)(i % MB);
        const long r = i / MB;
        const int comp = (int)(r % NCOMP);
        const long r2 = r / NCOMP;
        const int site = (int)(r2 % V), blk = (int)(r2 / V);
        const int mode = blk * MB + m, sp = comp / NC, colour = comp - sp * NC;
        double vr = 0.0, vi = 0.0, rr = 0.0, ri = 0.0;
        if (mode < nm) {
            const EigMode &e = modes[mode];
            rr = e.rot_r;
            ri = e.rot_i;
            if (e.blk >= 0) {
                if (colour == 0) {
                    const long z = site / plane, xy = site - z * plane;
                    const double2 t = tw[(size_t)e.blk * n2 + z];
                    const double ph_r = f * t.x, ph_i = -(f * t.y);
                    const double2 a = A[(size_t)e.blk * dz * dz + (size_t)e.idx * dz + xy * 4 + sp];
                    vr = ph_r * a.x - ph_i * a.y;
                    vi = ph_r * a.y + ph_i * a.x;
                }
            } else if (site == e.idx && colour == e.col) {
                vr = e.v[2 * sp];
                vi = e.v[2 * sp + 1];
            }
        }
        psi[i] = make_double2(vr, vi);
        mid[i] = make_double2(vr * rr - vi * ri, vr * ri + vi * rr);
    }
}

// Inf or NaN: the 11 exponent bits all set (read from the high word of each double)
__device__ __forceinline__ unsigned nonfinite_bits(d2v v)
{
    const unsigned hx = (unsigned)((unsigned long long)__double_as_longlong(v.x) >> 32);
    const unsigned hy = (unsigned)((unsigned long long)__double_as_longlong(v.y) >> 32);
    return (unsigned)((hx & 0x7ff00000u) == 0x7ff00000u) | (unsigned)((hy & 0x7ff00000u) == 0x7ff00000u);
}

// 1 in partial[blockIdx.x] when one of the n doubles of a, or of b (when given), is not finite.  A
// streaming pass over one index range (b's 16-byte units follow a's): four 16-byte loads in flight
// per lane, read once (non-temporal).
__global__ __launch_bounds__(TPB) void k_nonfinite(const double *__restrict__ a, const double *__restrict__ b, long n,
                                                   double *partial)
{
    constexpr int U = 4;
    const long n2 = n / 2;                // whole double2 units per array
    const long total = b ? 2 * n2 : n2;   // units of a, then of b
    const d2v *__restrict__ a2 = reinterpret_cast<const d2v *>(a);
    const d2v *__restrict__ b2 = reinterpret_cast<const d2v *>(b);
    const long stride = (long)gridDim.x * TPB;
    long i = (long)blockIdx.x * TPB + threadIdx.x;
    unsigned bad = 0;
    for (; i + (U - 1) * stride < total; i += U * stride) {
        d2v v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long j = i + u * stride;
            v[u] = __builtin_nontemporal_load(j < n2 ? a2 + j : b2 + (j - n2));
        }
#pragma unroll
        for (int u = 0; u < U; ++u) bad |= nonfinite_bits(v[u]);
    }
    for (; i < total; i += stride) bad |= nonfinite_bits(i < n2 ? a2[i] : b2[i - n2]);
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {  // an odd last double of each array
        bad |= nonfinite_bits(d2v{a[n - 1], 0.0});
        if (b) bad |= nonfinite_bits(d2v{b[n - 1], 0.0});
    }
    double v[1] = {bad ? 1.0 : 0.0};
    const int ism[1] = {1};
    block_reduce<1>(v, ism);
    if (threadIdx.x == 0) partial[blockIdx.x] = v[0];
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------------
hipError_t launch_hop(const Geom &g, int pin, int pout, const HopArgs &a, int wpe, hipStream_t st)
{
    if (g.nc == 2) {
        if (pin == P64 && pout == P64) return hop_mb<2, double, double>(g, a, wpe, st);
        if (pin == P32 && pout == P32) return hop_mb<2, float, float>(g, a, wpe, st);
        if (pin == P32 && pout == P64) return hop_mb<2, float, double>(g, a, wpe, st);
        if (pin == P16 && pout == P16) return hop_mb<2, _Float16, _Float16>(g, a, wpe, st);
        if (pin == P48 && pout == P48) return hop_mb<2, f48, f48>(g, a, wpe, st);
    } else if (g.nc == 3) {
        if (pin == P64 && pout == P64) return hop_mb<3, double, double>(g, a, wpe, st);
        if (pin == P32 && pout == P32) return hop_mb<3, float, float>(g, a, wpe, st);
        if (pin == P32 && pout == P64) return hop_mb<3, float, double>(g, a, wpe, st);
        if (pin == P16 && pout == P16) return hop_mb<3, _Float16, _Float16>(g, a, wpe, st);
        if (pin == P48 && pout == P48) return hop_mb<3, f48, f48>(g, a, wpe, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_convert(const Geom &g, int nblk, const void *src, int psrc, void *dst, int pdst, hipStream_t st,
                          double scale)
{
    const long npairs = (long)nblk * g.V * (4 * g.nc / 2) * g.mb;
    const long nunits = npairs / 2;
    if (psrc == P64 && pdst == P32)
        k_d2f<<<nblocks(npairs), TPB, 0, st>>>((const double2 *)src, (float4 *)dst, npairs, g.mb, scale);
    else if (psrc == P32 && pdst == P64)
        k_f2d<<<nblocks(npairs), TPB, 0, st>>>((const float4 *)src, (double2 *)dst, npairs, g.mb, scale);
    else if (psrc == P64 && pdst == P16)
        k_d2h<<<nblocks(nunits), TPB, 0, st>>>((const double2 *)src, (h8v *)dst, nunits, g.mb, scale);
    else if (psrc == P16 && pdst == P32)
        k_h2f<<<nblocks(nunits), TPB, 0, st>>>((const h8v *)src, (float4 *)dst, nunits, g.mb, (float)scale);
    else if (psrc == P64 && pdst == P48)
        k_convert<double, f48><<<nblocks(nunits), TPB, 0, st>>>(src, dst, nunits, g.mb, scale);
    else if (psrc == P32 && pdst == P48)
        k_convert<float, f48><<<nblocks(nunits), TPB, 0, st>>>(src, dst, nunits, g.mb, scale);
    else if (psrc == P48 && pdst == P64)
        k_convert<f48, double><<<nblocks(nunits), TPB, 0, st>>>(src, dst, nunits, g.mb, scale);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_lincomb(const Geom &g, int nblk, int p, void *out, const void *a, double car, double cai,
                          const void *b, double cbr, double cbi, hipStream_t st)
{
    const long ncplx = (long)nblk * g.V * 4 * g.nc * g.mb;
    if (p == P64)
        k_lincomb<double><<<nblocks(ncplx), TPB, 0, st>>>((double *)out, (const double *)a, car, cai, (const double *)b,
                                                          cbr, cbi, ncplx);
    else
        k_lincomb<float><<<nblocks(ncplx), TPB, 0, st>>>((float *)out, (const float *)a, car, cai, (const float *)b, cbr,
                                                         cbi, ncplx);
    return hipGetLastError();
}

hipError_t launch_aos_to_soa(const Geom &g, const double *aos, int first, int count, void *field, hipStream_t st)
{
    const long tot = (long)count * g.V * 4 * g.nc;
    k_aos_to_soa<<<nblocks(tot), TPB, 0, st>>>((const double2 *)aos, (double2 *)field, first, count, g.V, 4 * g.nc, g.mb);
    return hipGetLastError();
}

hipError_t launch_soa_to_aos(const Geom &g, const void *field, int first, int count, double *aos, hipStream_t st)
{
    const long tot = (long)count * g.V * 4 * g.nc;
    k_soa_to_aos<<<nblocks(tot), TPB, 0, st>>>((const double2 *)field, (double2 *)aos, first, count, g.V, 4 * g.nc, g.mb);
    return hipGetLastError();
}

hipError_t launch_link_setup(const Geom &g, const double *links, const double *phase, double *ul64, float *ul32,
                             hipStream_t st)
{
    const unsigned nb = nblocks(3L * g.V);
    if (g.nc == 2)
        k_link_setup<2><<<nb, TPB, 0, st>>>(links, (const double2 *)phase, (double2 *)ul64, (float2 *)ul32, g.V);
    else
        k_link_setup<3><<<nb, TPB, 0, st>>>(links, (const double2 *)phase, (double2 *)ul64, (float2 *)ul32, g.V);
    return hipGetLastError();
}

hipError_t launch_ucc_to_links(const Geom &g, const double *ucc, double *ul64, float *ul32, hipStream_t st)
{
    const long n = 3L * g.V * g.nc * g.nc;
    k_ucc_setup<<<nblocks(n), TPB, 0, st>>>((const double2 *)ucc, (double2 *)ul64, (float2 *)ul32, n);
    return hipGetLastError();
}

hipError_t launch_link_update(const Geom &g, double *links, const double *efield, double dt, int reunit,
                              hipStream_t st)
{
    const unsigned nb = nblocks(3L * g.V);
    if (g.nc == 2) k_link_update<2><<<nb, TPB, 0, st>>>(links, efield, dt, g.V, reunit);
    else k_link_update<3><<<nb, TPB, 0, st>>>(links, efield, dt, g.V, reunit);
    return hipGetLastError();
}

hipError_t launch_efield_update(const Geom &g, const double *links, double *efield, double dt, hipStream_t st)
{
    const unsigned nb = nblocks(3L * g.V);
    if (g.nc == 2) k_efield_update<2><<<nb, TPB, 0, st>>>(links, efield, dt, g.n0, g.n1, g.n2);
    else k_efield_update<3><<<nb, TPB, 0, st>>>(links, efield, dt, g.n0, g.n1, g.n2);
    return hipGetLastError();
}

size_t mode_dot_partial_len(const Geom &g) { return (size_t)DOT_CHUNKS * g.nblk * g.mb * 2; }

hipError_t launch_mode_dot(const Geom &g, const void *a, const void *b, double *partial, double *out, hipStream_t st)
{
    dim3 grid(DOT_CHUNKS, g.nblk);
    const auto A = (const double2 *)a;
    const auto B = (const double2 *)b;
    const auto P = (double2 *)partial;
#define DOT(NC, MB) k_mode_dot<NC, MB><<<grid, TPB, 0, st>>>(A, B, P, g.V, g.nblk)
    if (g.nc == 2) {
        if (g.mb == 16) DOT(2, 16); else if (g.mb == 32) DOT(2, 32); else DOT(2, 64);
    } else {
        if (g.mb == 16) DOT(3, 16); else if (g.mb == 32) DOT(3, 32); else DOT(3, 64);
    }
#undef DOT
    const int np = g.nblk * g.mb;
    k_mode_dot_finish<<<nblocks(np), TPB, 0, st>>>(P, (double2 *)out, np);
    return hipGetLastError();
}

size_t reduce_partial_len(const Geom &) { return (size_t)RED_BLOCKS * 4; }

hipError_t launch_energy(const Geom &g, const double *links, const double *efield, double *partial, double *out,
                         hipStream_t st)
{
    if (g.nc == 2) k_energy<2><<<RED_BLOCKS, TPB, 0, st>>>(links, efield, partial, g.n0, g.n1, g.n2);
    else k_energy<3><<<RED_BLOCKS, TPB, 0, st>>>(links, efield, partial, g.n0, g.n1, g.n2);
    k_finish<2><<<1, TPB, 0, st>>>(partial, RED_BLOCKS, out, 0, 0, 0, 0);
    return hipGetLastError();
}

hipError_t launch_gauss(const Geom &g, const double *links, const double *efield, double *partial, double *out,
                        double *gfield, hipStream_t st)
{
    if (g.nc == 2) k_gauss<2><<<RED_BLOCKS, TPB, 0, st>>>(links, efield, partial, gfield, g.n0, g.n1, g.n2);
    else k_gauss<3><<<RED_BLOCKS, TPB, 0, st>>>(links, efield, partial, gfield, g.n0, g.n1, g.n2);
    k_finish<2><<<1, TPB, 0, st>>>(partial, RED_BLOCKS, out, 1, 0, 0, 0);
    return hipGetLastError();
}

hipError_t launch_gauss_restore(const Geom &g, const double *links, const double *gfield, double *efield, double gamma,
                                hipStream_t st)
{
    const unsigned nb = nblocks(3L * g.V);
    if (g.nc == 2) k_gauss_restore<2><<<nb, TPB, 0, st>>>(links, gfield, efield, gamma, g.n0, g.n1, g.n2);
    else k_gauss_restore<3><<<nb, TPB, 0, st>>>(links, gfield, efield, gamma, g.n0, g.n1, g.n2);
    return hipGetLastError();
}

hipError_t launch_unitarity(const Geom &g, const double *links, double *partial, double *out, hipStream_t st)
{
    if (g.nc == 2) k_unitarity<2><<<RED_BLOCKS, TPB, 0, st>>>(links, partial, g.V);
    else k_unitarity<3><<<RED_BLOCKS, TPB, 0, st>>>(links, partial, g.V);
    k_finish<2><<<1, TPB, 0, st>>>(partial, RED_BLOCKS, out, 1, 0, 0, 0);
    return hipGetLastError();
}

hipError_t launch_edotb(const Geom &g, const double *links, const double *efield, double *partial, double *out,
                        hipStream_t st)
{
    if (g.nc == 2) k_edotb<2><<<RED_BLOCKS, TPB, 0, st>>>(links, efield, partial, g.n0, g.n1, g.n2);
    else k_edotb<3><<<RED_BLOCKS, TPB, 0, st>>>(links, efield, partial, g.n0, g.n1, g.n2);
    k_finish<1><<<1, TPB, 0, st>>>(partial, RED_BLOCKS, out, 0, 0, 0, 0);
    return hipGetLastError();
}

size_t currents_partial_len(const Geom &g) { return (size_t)g.nblk * g.V * 8; }

hipError_t launch_currents(const Geom &g, const void *psi, const void *hpsi, const double *gam, int first_mode,
                           double *partial, double *jv, double *ja, hipStream_t st)
{
    const long nsites = (long)g.nblk * g.V;
    const auto P = (const double2 *)psi;
    const auto H = (const double2 *)hpsi;
#define CUR(NC, MB) k_currents<NC, MB><<<(unsigned)((nsites + TPB / MB - 1) / (TPB / MB)), TPB, 0, st>>>(P, H, (const double2 *)gam, partial, g.V, g.nblk, first_mode)
    if (g.nc == 2) {
        if (g.mb == 16) CUR(2, 16); else if (g.mb == 32) CUR(2, 32); else CUR(2, 64);
    } else {
        if (g.mb == 16) CUR(3, 16); else if (g.mb == 32) CUR(3, 32); else CUR(3, 64);
    }
#undef CUR
    k_currents_finish<<<nblocks(g.V), TPB, 0, st>>>(partial, jv, ja, g.V, g.nblk, 1.0 / (2.0 * g.V));
    return hipGetLastError();
}

hipError_t launch_mscg_update(const Geom &g, int nshift, double *ps, double *chi, const double *r, const double *coef,
                              const int *active, hipStream_t st)
{
    const long per_blk = (long)g.V * 4 * g.nc;
    const long nel = per_blk * g.nblk * g.mb;
    k_mscg_update2<<<nblocks(nel), TPB, 0, st>>>(nshift, (double2 *)ps, (double2 *)chi, (const double2 *)r, coef, active,
                                                 per_blk, g.nblk, g.mb);
    return hipGetLastError();
}

hipError_t launch_axpy_modes(const Geom &g, void *y, const void *x, const double *coef, int coef_stride, double sign,
                             const int *active, hipStream_t st)
{
    const long per_blk = (long)g.V * 4 * g.nc;
    const long nel = per_blk * g.nblk * g.mb;
    k_axpy_modes<<<nblocks(nel), TPB, 0, st>>>((double2 *)y, (const double2 *)x, coef, coef_stride, sign, active, per_blk,
                                               g.nblk, g.mb);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// QED eigenbasis launchers
// ---------------------------------------------------------------------------------------------
#define OVL_EIG_DISPATCH(KERNEL, GRID, ...)                                       \
    do {                                                                          \
        if (g.nc == 2 && g.mb == 16) KERNEL<2, 16><<<GRID, TPB, 0, st>>>(__VA_ARGS__); \
        else if (g.nc == 2 && g.mb == 32) KERNEL<2, 32><<<GRID, TPB, 0, st>>>(__VA_ARGS__); \
        else if (g.nc == 2 && g.mb == 64) KERNEL<2, 64><<<GRID, TPB, 0, st>>>(__VA_ARGS__); \
        else if (g.nc == 3 && g.mb == 16) KERNEL<3, 16><<<GRID, TPB, 0, st>>>(__VA_ARGS__); \
        else if (g.nc == 3 && g.mb == 32) KERNEL<3, 32><<<GRID, TPB, 0, st>>>(__VA_ARGS__); \
        else if (g.nc == 3 && g.mb == 64) KERNEL<3, 64><<<GRID, TPB, 0, st>>>(__VA_ARGS__); \
        else return hipErrorInvalidValue;                                         \
    } while (0)

hipError_t launch_eig_sources(const Geom &g, void *field, long k0, int cnt, long dz, hipStream_t st)
{
    if (cnt < 1 || cnt > g.nblk * g.mb) return hipErrorInvalidValue;
    const unsigned grid = (unsigned)((cnt + TPB - 1) / TPB);
    OVL_EIG_DISPATCH(k_eig_sources, grid, (double2 *)field, g.V, k0, cnt, dz);
    return hipGetLastError();
}

hipError_t launch_eig_gather(const Geom &g, const void *out, void *A, long k0, int nsb, hipStream_t st)
{
    const long plane = (long)g.n0 * g.n1;
    if (nsb < 1 || nsb > g.nblk * g.mb) return hipErrorInvalidValue;
    const long total = (long)g.n2 * nsb * 4 * plane;
    const unsigned grid = (unsigned)std::min<long>((total + TPB - 1) / TPB, 65536);
    OVL_EIG_DISPATCH(k_eig_gather, grid, (const double2 *)out, (double2 *)A, plane, g.n2, k0, nsb);
    return hipGetLastError();
}

size_t eig_partial_len() { return (size_t)2 * RED_BLOCKS; }
int eig_partial_blocks() { return RED_BLOCKS; }

hipError_t launch_eig_leak(const Geom &g, const void *field, int nvalid, double *partial, hipStream_t st)
{
    OVL_EIG_DISPATCH(k_eig_leak, RED_BLOCKS, (const double2 *)field, g.V, g.nblk, nvalid, partial);
    return hipGetLastError();
}

hipError_t launch_eig_hermitize(void *A, long dz, int n2, double *partial, hipStream_t st)
{
    k_eig_hermitize<<<RED_BLOCKS, TPB, 0, st>>>((double2 *)A, dz, n2, partial);
    return hipGetLastError();
}

hipError_t launch_nonfinite(const void *x, const void *y, long n, double *partial, hipStream_t st)
{
    if (!x || n < 0) return hipErrorInvalidValue;
    k_nonfinite<<<RED_BLOCKS, TPB, 0, st>>>((const double *)x, (const double *)y, n, partial);
    return hipGetLastError();
}

hipError_t launch_eig_modes(const Geom &g, const EigMode *modes, int nm, const void *A, const void *tw, double f,
                            void *psi, void *mid, hipStream_t st)
{
    if (nm < 1 || nm > g.nblk * g.mb) return hipErrorInvalidValue;
    const long total = (long)g.nblk * g.V * 4 * g.nc * g.mb;
    const unsigned grid = (unsigned)std::min<long>((total + TPB - 1) / TPB, 65536);
    OVL_EIG_DISPATCH(k_eig_modes, grid, modes, nm, (const double2 *)A, (const double2 *)tw, f, g.V,
                     (long)g.n0 * g.n1, g.n2, g.nblk, (double2 *)psi, (double2 *)mid);
    return hipGetLastError();
}
#undef OVL_EIG_DISPATCH

}  // namespace ovl
