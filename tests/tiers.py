"""Storage-rounding model of the tiered Chebyshev sign, for tests only.

The mixed-precision sign evaluates eps(h_W) x = h_W P(T) x / lam with the Clenshaw recurrence
b_k = c_k x + 2 (alpha h_W^2 + beta) b_{k+1} - b_{k+2} (the oracle's orc_sign_cheb; the
reference evaluates the same sign with Zolotarev poles, OverlapOperator.cpp), storing each step's
vectors in a format that depends on k: FP64 (and the 48-bit layout, whose rounding, below
2^-37, is not modelled), FP32 or FP16.  FP16 vectors are stored times powers of two chosen from
bounds of the recurrence.

`sign_tiers` repeats that sequence of roundings in numpy with the oracle's FP64 h_W, so the GPU
result of a plan with an FP16 tier can be compared with the same roundings made on the host: the
two differ by the FP32 arithmetic of the GPU stencil (relative 2^-24 per operation), where the
plain FP64 form differs by the FP16 roundings themselves (2^-11)."""
import numpy as np

import orc


def q16(v):
    """Round to FP16 (via FP32, as the device stores FP32 results), nearest-even."""
    return np.asarray(v, np.float64).astype(np.float32).astype(np.float16).astype(np.float64)


def q32(v):
    return np.asarray(v, np.float64).astype(np.float32).astype(np.float64)


def qc(q, z):
    return q(z.real) + 1j * q(z.imag)


def clenshaw_sup(coef, t0, nt=2049):
    """B_k = max over nt Chebyshev-spaced t in [t0, 1] of |b_k(t)| of the scalar recurrence."""
    n = len(coef) - 1
    t = 0.5 * (1.0 + t0) + 0.5 * (1.0 - t0) * np.cos(np.pi * np.arange(nt) / (nt - 1))
    b1 = np.zeros(nt)
    b2 = np.zeros(nt)
    B = np.zeros(n + 2)
    for k in range(n, -1, -1):
        bk = coef[k] + 2.0 * t * b1 - b2
        b2, b1 = b1, bk
        B[k] = np.abs(bk).max()
    return B


def pow2_scale(bound):
    """Power of two mapping values bounded by `bound` to at most 2^8."""
    if not (bound > 0.0) or not np.isfinite(bound):
        return 1.0
    return float(np.ldexp(1.0, 8 - int(np.ceil(np.log2(bound)))))


def sign_tiers(L, ucc, lam, cut, coef, X, plan):
    """eps(h_W) x for every row of X with the storage roundings of `plan` = (head48, head,
    head16): k < head FP64, head <= k < head16 FP32, k >= head16 FP16 (power-of-two scales from
    the bound of the input norms over all rows)."""
    n = len(coef) - 1
    _, head, head16 = plan
    d2 = cut * cut
    alpha = 2.0 / (lam * lam * (1.0 - d2))
    beta = -(1.0 + d2) / (1.0 - d2)
    Ball = clenshaw_sup(coef, beta)
    xmax = max(np.linalg.norm(x) for x in X)
    sk = [pow2_scale(Ball[k] * xmax) for k in range(n + 1)]
    sw = [pow2_scale(lam * Ball[k + 1] * xmax) for k in range(n + 1)]
    sx = pow2_scale(xmax)

    def tier(k):
        return 64 if k < head else (32 if k < head16 else 16)

    Q = {64: lambda z: z, 32: lambda z: qc(q32, z), 16: lambda z: qc(q16, z)}
    out = []
    for x in X:
        b1 = b2 = None
        s1 = s2 = 1.0
        pb = None
        for k in range(n, -1, -1):
            p = tier(k)
            skk, sww, sxx = (sk[k], sw[k], sx) if p == 16 else (1.0, 1.0, 1.0)
            xk = Q[p](sxx * x)
            if b1 is not None and pb != p:  # promotion drops the FP16 scale (exact)
                b1 = Q[p](b1 / s1)
                b2 = Q[p](b2 / s2) if b2 is not None else None
                s1 = s2 = 1.0
                pb = p
            if b1 is None:  # b_n = c_n x
                b1 = Q[p](coef[k] * skk * x) if p == 16 else Q[p](coef[k] * xk)
                s1, pb = skk, p
                continue
            w = Q[p](sww / s1 * orc.hw(L, ucc, b1))
            f = 1.0 if k == 0 else 2.0
            bk = f * alpha * skk / sww * orc.hw(L, ucc, w) + coef[k] * skk / sxx * xk + f * beta * skk / s1 * b1
            if b2 is not None:
                bk = bk - skk / s2 * b2
            b2, s2 = b1, s1
            b1, s1 = Q[p](bk), skk
        out.append(orc.hw(L, ucc, b1) / lam)
    return np.array(out)
