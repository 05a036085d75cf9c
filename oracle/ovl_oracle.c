/*
 * ovl_oracle.c -- FP64 CPU restatement of the OVERLAP fermion hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Each function
 * cites the reference file:line it restates.  Layouts: see ovl_oracle.h.
 *
 * Pinning (what each function is checked against, tests/):
 *   orc_wilson_evo        golden vectors of the compiled reference (oracle/_ref), free-field
 *                         closed form, Hermiticity, gauge covariance
 *   orc_zolotarev         golden vectors of the compiled reference
 *   orc_zolotrr / orc_overlap_evo_zolo / orc_ritz   golden vectors of the compiled reference
 *   orc_sign_cheb         dense eigendecomposition sign(H) on small lattices, eps^2 = 1
 *   orc_su2_iexp          compiled reference MatrixIExp; orc_su3_iexp: SU(2) embedding,
 *                         unitarity, det = 1, exp(A)exp(-A) = 1, small-angle series
 *   orc_update_links      exact constant-E solution, unitarity
 *   orc_reunitarize       SVD polar factor with the determinant phase removed (numpy)
 *   orc_efield_update     Gauss-law conservation and energy conservation  (the reference's
 *                         plaquette macros are absent: formula restated from the Hamiltonian)
 *   orc_edotb             gauge invariance, E=0, linearity in E
 *   orc_gauss_restore     finite-difference gradient of (1/2) sum G^2 (the step direction) and
 *                         convergence to G = 0 (tests); the reference macro for G is absent
 *   orc_currents          vector charge = sum_s w_s |psi_s|^2 / (2V) identity; hermiticity
 */
#include "ovl_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

#define NCMAX 3
#define NG_MAX 8

static inline int lat_vol(const orc_lat *L) { return L->n[0] * L->n[1] * L->n[2]; }

static inline int site_of(const int *n, int x, int y, int z)
{
    x %= n[0]; if (x < 0) x += n[0];
    y %= n[1]; if (y < 0) y += n[1];
    z %= n[2]; if (z < 0) z += n[2];
    return x + n[0] * (y + n[1] * z);
}

static inline void coords_of(const int *n, int s, int *x, int *y, int *z)
{
    *z = s / (n[0] * n[1]);
    *y = (s - *z * n[0] * n[1]) / n[0];
    *x = s - n[0] * (*y + n[1] * *z);
}

/* WilsonOperator.cpp:5-15 (iup) and :18-28 (idn); mu = 0,1,2 here (reference passes mu+1) */
static int iup(const int *n, int n1, int mu)
{
    int z = n1 / (n[1] * n[0]);
    int y = (n1 - z * n[1] * n[0]) / n[0];
    int x = n1 - z * n[1] * n[0] - y * n[0];
    if (mu == 0) return n1 + (x + 1) % n[0] - x;
    if (mu == 1) return n1 + n[0] * ((y + 1) % n[1] - y);
    return n1 + n[0] * n[1] * ((z + 1) % n[2] - z);
}

static int idn(const int *n, int n1, int mu)
{
    int z = n1 / (n[1] * n[0]);
    int y = (n1 - z * n[1] * n[0]) / n[0];
    int x = n1 - z * n[1] * n[0] - y * n[0];
    if (mu == 0) return n1 + (x - 1 + n[0]) % n[0] - x;
    if (mu == 1) return n1 + n[0] * ((y - 1 + n[1]) % n[1] - y);
    return n1 + n[0] * n[1] * ((z - 1 + n[2]) % n[2] - z);
}

/* ------------------------------------------------------------------------- */
/* small complex matrices (row-major, nc x nc)                                */
/* ------------------------------------------------------------------------- */
static void m_mul(int nc, const cplx *a, const cplx *b, cplx *c)
{
    cplx t[9];
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < nc; j++) {
            cplx s = 0;
            for (int k = 0; k < nc; k++) s += a[i * nc + k] * b[k * nc + j];
            t[i * nc + j] = s;
        }
    memcpy(c, t, sizeof(cplx) * nc * nc);
}

static void m_dag(int nc, const cplx *a, cplx *c)
{
    cplx t[9];
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < nc; j++) t[i * nc + j] = conj(a[j * nc + i]);
    memcpy(c, t, sizeof(cplx) * nc * nc);
}

static void m_id(int nc, cplx *a)
{
    for (int i = 0; i < nc * nc; i++) a[i] = 0;
    for (int i = 0; i < nc; i++) a[i * nc + i] = 1;
}

/* generators t^a: SU(2) sigma_a/2 (SU2/AlgebraOperations.cpp:6), SU(3) lambda_a/2
 * (SU3/AlgebraOperations.cpp:6-25) */
static void gen_matrix(int nc, int a, cplx *t)
{
    for (int i = 0; i < nc * nc; i++) t[i] = 0;
    if (nc == 2) {
        if (a == 0) { t[1] = 0.5; t[2] = 0.5; }
        if (a == 1) { t[1] = -0.5 * I; t[2] = 0.5 * I; }
        if (a == 2) { t[0] = 0.5; t[3] = -0.5; }
        return;
    }
    switch (a) {
    case 0: t[1] = 0.5; t[3] = 0.5; break;
    case 1: t[1] = -0.5 * I; t[3] = 0.5 * I; break;
    case 2: t[0] = 0.5; t[4] = -0.5; break;
    case 3: t[2] = 0.5; t[6] = 0.5; break;
    case 4: t[2] = -0.5 * I; t[6] = 0.5 * I; break;
    case 5: t[5] = 0.5; t[7] = 0.5; break;
    case 6: t[5] = -0.5 * I; t[7] = 0.5 * I; break;
    case 7: {
        double s = 0.5 / sqrt(3.0);
        t[0] = s; t[4] = s; t[8] = -2.0 * s;
    } break;
    }
}

static inline int ngen(int nc) { return nc * nc - 1; }

/* alpha^a t^a */
static void alg_matrix(int nc, const double *alpha, cplx *m)
{
    cplx t[9];
    for (int i = 0; i < nc * nc; i++) m[i] = 0;
    for (int a = 0; a < ngen(nc); a++) {
        gen_matrix(nc, a, t);
        for (int i = 0; i < nc * nc; i++) m[i] += alpha[a] * t[i];
    }
}

/* ReTr(i t^a V) -- SU2/GroupOperations.cpp:78-84 for Nc=2 */
static void retr_itu(int nc, const cplx *v, double *out)
{
    cplx t[9];
    for (int a = 0; a < ngen(nc); a++) {
        gen_matrix(nc, a, t);
        cplx s = 0;
        for (int i = 0; i < nc; i++)
            for (int k = 0; k < nc; k++) s += t[i * nc + k] * v[k * nc + i];
        out[a] = creal(I * s);
    }
}

/* link at (site, mu) as a row-major matrix; SU(2) via GetMatrix (SU2/GroupOperations.cpp:213-225) */
static void link_matrix(int nc, const double *links, int s, int mu, cplx *m)
{
    if (nc == 2) {
        const double *u = links + 4 * (mu + 3 * s);
        m[0] = u[3] + I * u[2];
        m[1] = u[1] + I * u[0];
        m[2] = -u[1] + I * u[0];
        m[3] = u[3] - I * u[2];
    } else {
        const double *u = links + 18 * (mu + 3 * s);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                int k = i + 3 * j;
                m[i * 3 + j] = u[2 * k] + I * u[2 * k + 1];
            }
    }
}

static void link_store(int nc, double *links, int s, int mu, const cplx *m)
{
    if (nc == 2) {
        /* GetGaugeLink (SU2/GroupOperations.cpp:227-239) */
        double *u = links + 4 * (mu + 3 * s);
        u[0] = creal(0.5 * I * (-m[2] - m[1]));
        u[1] = creal(0.5 * (-m[2] + m[1]));
        u[2] = creal(0.5 * I * (m[0] - m[3]));
        u[3] = creal(0.5 * (m[0] + m[3]));
    } else {
        double *u = links + 18 * (mu + 3 * s);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                int k = i + 3 * j;
                u[2 * k] = creal(m[i * 3 + j]);
                u[2 * k + 1] = cimag(m[i * 3 + j]);
            }
    }
}

/* ------------------------------------------------------------------------- */
/* gauge data                                                                 */
/* ------------------------------------------------------------------------- */

/* QED.cpp:131-153: U_y(x,y) = exp(-2 pi i Nb x/(Nx Ny)), U_x(Nx-1,y) = exp(2 pi i Nb y Nx/(Nx Ny)),
 * all others 1; the charge q multiplies the phase angle (q = 1 is the reference). */
void orc_u1_phases(const int n[3], int nb, double q, double *phase)
{
    int V = n[0] * n[1] * n[2];
    cplx *p = (cplx *)phase;
    for (int s = 0; s < V; s++) {
        int x, y, z;
        coords_of(n, s, &x, &y, &z);
        double ay = -2.0 * M_PI * (double)(nb * x) / (double)(n[0] * n[1]);
        double ax = (x == n[0] - 1) ? 2.0 * M_PI * (double)(nb * y * n[0]) / (double)(n[0] * n[1]) : 0.0;
        p[0 * V + s] = cexp(I * q * ax);
        p[1 * V + s] = cexp(I * q * ay);
        p[2 * V + s] = 1.0;
    }
}

/* OverlapEvolutionOperator.cpp:23-51 ConvertGaugeLinks: Ucc[s + mu V].c[i].c[j] = U(i,j),
 * times the U(1) phase of that link when phase != NULL. */
void orc_links_to_ucc(const orc_lat *L, const double *links, const double *phase, double *ucc_)
{
    int V = lat_vol(L), nc = L->nc;
    cplx *ucc = (cplx *)ucc_;
    const cplx *ph = (const cplx *)phase;
    for (int s = 0; s < V; s++)
        for (int mu = 0; mu < 3; mu++) {
            cplx m[9];
            link_matrix(nc, links, s, mu, m);
            cplx f = ph ? ph[mu * V + s] : 1.0;
            for (int k = 0; k < nc * nc; k++) ucc[(size_t)(mu * V + s) * nc * nc + k] = f * m[k];
        }
}

/* SU2/AlgebraOperations.cpp:37-55 MatrixIExp: exp(i c alpha^a sigma_a/2) in 4-real format */
void orc_su2_iexp(double c, const double *alpha, double *e)
{
    double d = sqrt((c * alpha[0]) * (c * alpha[0]) + (c * alpha[1]) * (c * alpha[1]) +
                    (c * alpha[2]) * (c * alpha[2]));
    double arg = 0.5 * d;
    if (d != 0) {
        double m = c * sin(arg) / d;
        e[0] = m * alpha[0];
        e[1] = m * alpha[1];
        e[2] = m * alpha[2];
        e[3] = cos(arg);
    } else {
        e[0] = e[1] = e[2] = 0.0;
        e[3] = 1.0;
    }
}

/* exp(i c alpha^a lambda_a/2) for SU(3) by scaling and squaring of the Taylor series
 * (independent of the reference's eigendecomposition, SU3/AlgebraOperations.cpp:141-182).
 * Output in the reference SU(3) format: element (i,j) at i+3j. */
void orc_su3_iexp(double c, const double *alpha, double *out)
{
    cplx a[9], x[9], term[9], sum[9];
    alg_matrix(3, alpha, a);
    double nrm = 0;
    for (int i = 0; i < 9; i++) { a[i] *= I * c; nrm += creal(a[i] * conj(a[i])); }
    nrm = sqrt(nrm);
    int sq = 0;
    while (nrm > 0.25 && sq < 64) { nrm *= 0.5; sq++; }
    double scale = ldexp(1.0, -sq);
    for (int i = 0; i < 9; i++) x[i] = a[i] * scale;
    m_id(3, sum);
    m_id(3, term);
    for (int k = 1; k <= 24; k++) {
        m_mul(3, term, x, term);
        for (int i = 0; i < 9; i++) { term[i] /= (double)k; sum[i] += term[i]; }
    }
    for (int s = 0; s < sq; s++) m_mul(3, sum, sum, sum);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            out[2 * (i + 3 * j)] = creal(sum[i * 3 + j]);
            out[2 * (i + 3 * j) + 1] = cimag(sum[i * 3 + j]);
        }
}

/* Dynamics.cpp:97-138 UpdateGaugeLinks: U_mu(x) <- exp(i cU E_mu(x)) U_mu(x),
 * cU = -dt (Minkowski metric, a = 1; Dynamics.cpp:102-104). */
void orc_update_links(const orc_lat *L, double *links, const double *efield, double dt)
{
    int V = lat_vol(L), nc = L->nc, ng = ngen(nc);
    double cu = -dt;
    for (int s = 0; s < V; s++)
        for (int mu = 0; mu < 3; mu++) {
            const double *e = efield + ng * (mu + 3 * s);
            if (nc == 2) {
                /* SU2/GroupOperations.cpp:103-110 UU */
                double up[4], old[4];
                double *u = links + 4 * (mu + 3 * s);
                orc_su2_iexp(cu, e, up);
                memcpy(old, u, sizeof old);
                u[0] = up[3] * old[0] + up[0] * old[3] - (up[1] * old[2] - up[2] * old[1]);
                u[1] = up[3] * old[1] + up[1] * old[3] + (up[0] * old[2] - up[2] * old[0]);
                u[2] = up[3] * old[2] + up[2] * old[3] - (up[0] * old[1] - up[1] * old[0]);
                u[3] = up[3] * old[3] - up[0] * old[0] - up[1] * old[1] - up[2] * old[2];
            } else {
                double ex[18];
                cplx m[9], u[9];
                orc_su3_iexp(cu, e, ex);
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++)
                        m[i * 3 + j] = ex[2 * (i + 3 * j)] + I * ex[2 * (i + 3 * j) + 1];
                link_matrix(3, links, s, mu, u);
                m_mul(3, m, u, u);
                link_store(3, links, s, mu, u);
            }
        }
}

/* Projection of every link back onto SU(Nc).  SU(2): V/|V| in the 4-real format, the
 * reference's MaxTraceProjection (SU2/GroupOperations.cpp:261-270).  SU(3): the unitary polar
 * factor P = V (V^+ V)^{-1/2} (the U(3) matrix of largest Re Tr(P^+ V)) by the Newton iteration
 * X <- (X + X^{-+})/2 (Higham, "Computing the polar decomposition", 1986), then P det(P)^{-1/3}
 * with the principal cube root.  The reference's SU(3) MaxTraceProjection
 * (SU3/GroupOperations.cpp:510-533, seven Cabibbo-Marinari sweeps) maximises Re Tr over SU(3)
 * instead; both agree to second order in the distance of V from the group. */
static void inv3(const cplx *a, cplx *r)
{
    cplx c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
    cplx det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    r[0] = c00 / det; r[3] = c01 / det; r[6] = c02 / det;
    r[1] = (a[2] * a[7] - a[1] * a[8]) / det;
    r[4] = (a[0] * a[8] - a[2] * a[6]) / det;
    r[7] = (a[1] * a[6] - a[0] * a[7]) / det;
    r[2] = (a[1] * a[5] - a[2] * a[4]) / det;
    r[5] = (a[2] * a[3] - a[0] * a[5]) / det;
    r[8] = (a[0] * a[4] - a[1] * a[3]) / det;
}

void orc_reunitarize(const orc_lat *L, double *links)
{
    int V = lat_vol(L), nc = L->nc;
    for (int s = 0; s < V; s++)
        for (int mu = 0; mu < 3; mu++) {
            if (nc == 2) {
                double *u = links + 4 * (mu + 3 * s);
                double nrm = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2] + u[3] * u[3]);
                for (int k = 0; k < 4; k++) u[k] /= nrm;
                continue;
            }
            cplx x[9], xi[9];
            link_matrix(3, links, s, mu, x);
            for (int it = 0; it < 6; it++) {
                inv3(x, xi);
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) x[i * 3 + j] = 0.5 * (x[i * 3 + j] + conj(xi[j * 3 + i]));
            }
            cplx det = x[0] * (x[4] * x[8] - x[5] * x[7]) - x[1] * (x[3] * x[8] - x[5] * x[6]) +
                       x[2] * (x[3] * x[7] - x[4] * x[6]);
            cplx f = cexp(-I * carg(det) / 3.0);
            for (int i = 0; i < 9; i++) x[i] *= f;
            link_store(3, links, s, mu, x);
        }
}

/* ------------------------------------------------------------------------- */
/* Wilson operator                                                            */
/* ------------------------------------------------------------------------- */

/* WilsonOperator.cpp:30-114, generalised from 2 to Nc colours: psi = -i h_W phi with
 * h_W = (1/2) gamma0 [ A phi - sum_k (1 + i gamma_k) U phi(x+k) + (1 - i gamma_k) U^+ phi(x-k) ],
 * A = 6 r + 2 m - 2 M.  AoS layout [site][colour][spin]. */
void orc_wilson_evo(const orc_lat *L, const double *ucc_, const double *phi_, double *psi_)
{
    const cplx *u = (const cplx *)ucc_;
    const cplx *phi = (const cplx *)phi_;
    cplx *psi = (cplx *)psi_;
    const int nc = L->nc, V = lat_vol(L), nn = nc * nc;
    const double amass = 6.0 * L->r_wilson + 2.0 * L->m_fermion - 2.0 * L->m_dw;

    for (int ix = 0; ix < V; ix++)
        for (int c = 0; c < nc; c++)
            for (int jf = 0; jf < 4; jf++)
                psi[(ix * nc + c) * 4 + jf] = amass * phi[(ix * nc + c) * 4 + jf];

    for (int ix = 0; ix < V; ix++) {
        cplx d[NCMAX][4][3];
        for (int jf = 0; jf < 4; jf++)
            for (int mu = 0; mu < 3; mu++) {
                cplx a[NCMAX], b[NCMAX];
                int jx = iup(L->n, ix, mu);
                int bc = ix + mu * V;
                for (int i = 0; i < nc; i++) {
                    cplx s = 0;
                    for (int j = 0; j < nc; j++) s += u[(size_t)bc * nn + i * nc + j] * phi[(jx * nc + j) * 4 + jf];
                    a[i] = s;
                }
                jx = idn(L->n, ix, mu);
                bc = jx + mu * V;
                for (int i = 0; i < nc; i++) {
                    cplx s = 0;
                    for (int j = 0; j < nc; j++) s += conj(u[(size_t)bc * nn + j * nc + i]) * phi[(jx * nc + j) * 4 + jf];
                    b[i] = s;
                }
                for (int i = 0; i < nc; i++) {
                    d[i][jf][mu] = a[i] - b[i];
                    psi[(ix * nc + i) * 4 + jf] = psi[(ix * nc + i) * 4 + jf] - a[i] - b[i];
                }
            }
        for (int i = 0; i < nc; i++) {
            cplx *p = psi + (ix * nc + i) * 4;
            /* WilsonOperator.cpp:90-93 */
            p[0] = p[0] - I * d[i][3][0] - 1.0 * d[i][3][1] - I * d[i][2][2];
            p[1] = p[1] - I * d[i][2][0] + 1.0 * d[i][2][1] + I * d[i][3][2];
            p[2] = p[2] + I * d[i][1][0] + 1.0 * d[i][1][1] + I * d[i][0][2];
            p[3] = p[3] + I * d[i][0][0] - 1.0 * d[i][0][1] - I * d[i][1][2];
            /* WilsonOperator.cpp:101-104 */
            p[0] = -I * p[0] * 0.5;
            p[1] = -I * p[1] * 0.5;
            p[2] = I * p[2] * 0.5;
            p[3] = I * p[3] * 0.5;
        }
    }
}

/* Hermitian h_W = i * (-i h_W) */
void orc_hw(const orc_lat *L, const double *ucc, const double *phi, double *psi_)
{
    cplx *psi = (cplx *)psi_;
    orc_wilson_evo(L, ucc, phi, psi_);
    int n = lat_vol(L) * L->nc * 4;
    for (int i = 0; i < n; i++) psi[i] *= I;
}

/* ||h_W|| <= |A|/2 + 3: each direction's hopping block squares to 1 for unitary links. */
double orc_hw_norm_bound(const orc_lat *L)
{
    double a = 3.0 * L->r_wilson + L->m_fermion - L->m_dw;
    return fabs(a) + 3.0;
}

/* ------------------------------------------------------------------------- */
/* Chebyshev sign function                                                    */
/* ------------------------------------------------------------------------- */

/* Chebyshev coefficients of f(x) = x^{-1/2} on x in [cut^2, 1], variable t = (2x - 1 - cut^2)/(1 - cut^2).
 * c[0] is already halved so that P(t) = sum_k c_k T_k(t). */
void orc_cheb_coeffs(int n, double cut, double *c)
{
    int nodes = 8 * (n + 1) + 64;
    double d2 = cut * cut;
    for (int k = 0; k <= n; k++) {
        long double s = 0;
        for (int j = 0; j < nodes; j++) {
            double th = M_PI * (j + 0.5) / nodes;
            double t = cos(th);
            double x = 0.5 * ((1.0 - d2) * t + (1.0 + d2));
            s += (long double)(1.0 / sqrt(x)) * (long double)cos(k * th);
        }
        c[k] = (double)(2.0L * s / nodes);
    }
    c[0] *= 0.5;
}

/* out = eps(h_W) phi = h_W * (1/lam) * P(T) phi with T = alpha h_W^2 + beta (Clenshaw). */
void orc_sign_cheb(const orc_lat *L, const double *ucc, double lam, double cut, int n,
                   const double *coef, const double *phi_, double *out_)
{
    const cplx *x = (const cplx *)phi_;
    int N = lat_vol(L) * L->nc * 4;
    double d2 = cut * cut;
    double alpha = 2.0 / (lam * lam * (1.0 - d2));
    double beta = -(1.0 + d2) / (1.0 - d2);
    cplx *b1 = calloc(N, sizeof(cplx)), *b2 = calloc(N, sizeof(cplx)), *bk = calloc(N, sizeof(cplx));
    cplx *w = malloc(N * sizeof(cplx)), *w2 = malloc(N * sizeof(cplx));
    for (int k = n; k >= 1; k--) {
        orc_hw(L, ucc, (double *)b1, (double *)w);
        orc_hw(L, ucc, (double *)w, (double *)w2);
        for (int i = 0; i < N; i++) bk[i] = coef[k] * x[i] + 2.0 * (alpha * w2[i] + beta * b1[i]) - b2[i];
        cplx *t = b2; b2 = b1; b1 = bk; bk = t;
    }
    orc_hw(L, ucc, (double *)b1, (double *)w);
    orc_hw(L, ucc, (double *)w, (double *)w2);
    for (int i = 0; i < N; i++) w[i] = (coef[0] * x[i] + alpha * w2[i] + beta * b1[i] - b2[i]) / lam;
    orc_hw(L, ucc, (double *)w, out_);
    free(b1); free(b2); free(bk); free(w); free(w2);
}

/* OverlapU.cpp:188-201 combination with the Chebyshev sign:
 * psi = M [ -i gamma0 (1 + m/2) phi + (1 - m/2) (-i eps phi) ] */
static void overlap_combine(const orc_lat *L, const cplx *phi, const cplx *meps, cplx *psi)
{
    int V = lat_vol(L), nc = L->nc;
    double M = L->m_dw, m = L->m_fermion;
    for (int ix = 0; ix < V; ix++)
        for (int c = 0; c < nc; c++)
            for (int a = 0; a < 4; a++) {
                int k = (ix * nc + c) * 4 + a;
                cplx g = (a < 2) ? -I : I;
                psi[k] = M * (g * (1.0 + 0.5 * m) * phi[k] + (1.0 - 0.5 * m) * meps[k]);
            }
}

void orc_overlap_evo_cheb(const orc_lat *L, const double *ucc, double lam, double cut, int n,
                          const double *coef, const double *phi, double *psi)
{
    int N = lat_vol(L) * L->nc * 4;
    cplx *e = malloc(N * sizeof(cplx));
    orc_sign_cheb(L, ucc, lam, cut, n, coef, phi, (double *)e);
    for (int i = 0; i < N; i++) e[i] *= -I;
    overlap_combine(L, (const cplx *)phi, e, (cplx *)psi);
    free(e);
}

/* ------------------------------------------------------------------------- */
/* Zolotarev rational approximation (reference algorithm)                     */
/* ------------------------------------------------------------------------- */

/* ZolotarevCoefficients.cpp:5-49 CompElliptic (Carlson R_F, restated literally) */
static double comp_elliptic(double kappa, double err)
{
    double mu, xn, yn, zn, xndev, yndev, zndev, epsilon, e2, e3, s, k, xr, yr, zr, lambda;
    xn = 0.0; yn = 1.0 - kappa; zn = 1.0; epsilon = err;
    mu = (xn + yn + zn) / 3.0;
    xndev = 2.0 - (mu + xn) / mu;
    yndev = 2.0 - (mu + yn) / mu;
    zndev = 2.0 - (mu + zn) / mu;
    while (epsilon >= err) {
        mu = (xn + yn + zn) / 3.0;
        xndev = 2.0 - (mu + xn) / mu;
        yndev = 2.0 - (mu + yn) / mu;
        zndev = 2.0 - (mu + zn) / mu;
        epsilon = fabs(xndev);
        if (fabs(yndev) > epsilon) epsilon = fabs(yndev);
        if (fabs(zndev) > epsilon) epsilon = fabs(zndev);
        xr = sqrt(xn); yr = sqrt(yn); zr = sqrt(zn);
        lambda = xr * (yr + zr) + yr * zr;
        xn = 0.25 * (xn + lambda);
        yn = 0.25 * (yn + lambda);
        zn = 0.25 * (zn + lambda);
    }
    e2 = xndev * yndev - zndev * zndev;
    e3 = xndev * yndev * zndev;
    s = 1.0 + (e2 / 24.0 - 0.1 - e3 / 44.0) * e2 + e3 / 14.0;
    k = s / sqrt(mu);
    return k;
}

/* ZolotarevCoefficients.cpp:51-69 coeffb */
static double coeffb(const double *cval, int n, int norder)
{
    double prodnr = 1.0;
    int j = 0;
    for (int i = 0; i < norder - 1; i++) {
        prodnr = prodnr * (cval[2 * i + 1] - cval[2 * n]);
        if (i == n) j = j + 1;
        prodnr = prodnr / (cval[2 * j] - cval[2 * n]);
        j++;
    }
    return prodnr;
}

/* ZolotarevCoefficients.cpp:71-175, including the single-precision square roots at :97,106,148
 * and the fmin = pow(1,10) = 1 start value at :136.  Returns the max relative error eps. */
double orc_zolotarev(double amin, double amax, double errtol, int no, double *bz, double *dz)
{
    const int maxzolot = 4000;
    double *alpha = malloc(sizeof(double) * maxzolot), *gamma = malloc(sizeof(double) * maxzolot);
    double *c = malloc(sizeof(double) * 2 * no);
    double kappa = sqrt(1.0 - amin / amax);
    double K = comp_elliptic(kappa, errtol);
    double zn = 0.5 * K / no, xn, betak, tempbeta, pow2, phi, yn, fmin, fmax, anorm, eps, dx;
    for (int l = 0; l < 2 * no - 1; l++) {
        xn = (l + 1) * zn;
        alpha[0] = 1.0;
        betak = sqrtf((float)(1.0 - kappa));
        gamma[0] = sqrt(kappa);
        int i = 1;
        pow2 = 1.0;
        while (fabs(gamma[i - 1] / alpha[i - 1]) >= errtol) {
            tempbeta = betak;
            alpha[i] = (alpha[i - 1] + tempbeta) * 0.5;
            betak = sqrtf((float)(alpha[i - 1] * tempbeta));
            gamma[i] = (alpha[i - 1] - tempbeta) * 0.5;
            i++;
            pow2 = pow2 * 2.0;
        }
        i = i - 1;
        phi = pow2 * alpha[i] * xn;
        while (i >= 1) {
            phi = 0.5 * (phi + asin(gamma[i] * sin(phi) / alpha[i]));
            i = i - 1;
        }
        c[l] = tan(phi) * tan(phi);
    }
    for (int l = 0; l < no; l++) {
        dz[l] = c[2 * l] * amin;
        bz[l] = coeffb(c, l, no);
    }
    int npoints = 5000;
    dx = (log(amax) - log(amin)) / npoints;
    yn = log(amin) - dx;
    fmin = pow(1.0, 10.0);
    fmax = 0.0;
    for (int i = 0; i < npoints; i++) {
        yn = yn + dx;
        xn = exp(yn);
        zn = 0.0;
        for (int l = 0; l < no; l++) zn = zn + bz[l] / (xn + dz[l]);
        zn = zn * sqrtf((float)xn);
        if (zn > fmax) fmax = zn;
        if (zn < fmin) fmin = zn;
    }
    anorm = 2.0 / (fmin + fmax);
    eps = (fmax - fmin) / (fmax + fmin);
    for (int k = 0; k < no; k++) bz[k] = anorm * bz[k];
    free(c); free(alpha); free(gamma);
    return eps;
}

/* RitzSolver.cpp:36-45 norm: Re sum conj(v1) v2 */
static double rnorm2(const cplx *v1, const cplx *v2, int n)
{
    cplx t = 0;
    for (int i = 0; i < n; i++) t += v2[i] * conj(v1[i]);
    return creal(t);
}

static cplx cnorm2(const cplx *v1, const cplx *v2, int n)
{
    cplx t = 0;
    for (int i = 0; i < n; i++) t += v2[i] * conj(v1[i]);
    return t;
}

/* RitzSolver.cpp:49-89 WilsonUU: b = -(-i h)^2 a, on the flat layout m = s + V(spin + 4 colour) */
static void wilson_uu(const orc_lat *L, const double *ucc, const cplx *a, cplx *b, cplx *A, cplx *chi, cplx *B)
{
    int V = lat_vol(L), nc = L->nc;
    for (int i = 0; i < V; i++)
        for (int k = 0; k < 4; k++)
            for (int j = 0; j < nc; j++) A[(i * nc + j) * 4 + k] = a[i + V * (k + j * 4)];
    orc_wilson_evo(L, ucc, (const double *)A, (double *)chi);
    orc_wilson_evo(L, ucc, (const double *)chi, (double *)B);
    for (int i = 0; i < V; i++)
        for (int k = 0; k < 4; k++)
            for (int j = 0; j < nc; j++) b[i + V * (k + j * 4)] = -B[(i * nc + j) * 4 + k];
}

/* OverlapU.cpp:5-149 Zolotrr: multi-shift CG for (H^2 + dz_k) x_k = ini, chi = sum_k bz_k x_k.
 * ini/chi in AoS here; transposed to the flat layout internally as OverlapU.cpp:160-185 does.
 * Returns the number of CG iterations. */
int orc_zolotrr(const orc_lat *L, const double *ucc, const double *bz, const double *dz, int no,
                double cgerr, int nitcgm, const double *ini_aos, double *chi_aos)
{
    int V = lat_vol(L), nc = L->nc, mvfc = V * nc * 4;
    cplx *x = calloc((size_t)mvfc * no, sizeof(cplx)), *ps = malloc(sizeof(cplx) * (size_t)mvfc * no);
    cplx *r = malloc(sizeof(cplx) * mvfc), *q = malloc(sizeof(cplx) * mvfc), *spbuf = malloc(sizeof(cplx) * mvfc);
    cplx *ini = malloc(sizeof(cplx) * mvfc), *chi = malloc(sizeof(cplx) * mvfc);
    cplx *tA = malloc(sizeof(cplx) * mvfc), *tC = malloc(sizeof(cplx) * mvfc), *tB = malloc(sizeof(cplx) * mvfc);
    double *z = malloc(sizeof(double) * no), *alphas = malloc(sizeof(double) * no), *betas = malloc(sizeof(double) * no),
           *gcoeff = malloc(sizeof(double) * no);
    const cplx *in = (const cplx *)ini_aos;
    for (int i = 0; i < V; i++)
        for (int k = 0; k < 4; k++)
            for (int j = 0; j < nc; j++) ini[i + V * (k + j * 4)] = in[(i * nc + j) * 4 + k];

    for (int j = 0; j < mvfc; j++) chi[j] = 0.0;
    for (int j = 0; j < mvfc; j++) {
        r[j] = ini[j];
        for (int k = 0; k < no; k++) ps[j + (size_t)k * mvfc] = ini[j];
    }
    double rn = rnorm2(r, r, mvfc);
    int it = 0;
    double bnorm = sqrt(rn), betam = 1.0;
    alphas[0] = 0.0;
    for (int k = 0; k < no; k++) { gcoeff[k] = 1.0; z[k] = 1.0; }
    while (rn > cgerr * bnorm) {
        for (int j = 0; j < mvfc; j++) spbuf[j] = ps[j];
        wilson_uu(L, ucc, spbuf, q, tA, tC, tB);
        for (int j = 0; j < mvfc; j++) q[j] = q[j] + ps[j] * dz[0];
        betas[0] = rn / rnorm2(spbuf, q, mvfc);
        for (int j = 0; j < mvfc; j++) r[j] = r[j] - betas[0] * q[j];
        for (int k = 1; k < no; k++) {
            double dr = betas[0] * alphas[0] * (1.0 - z[k]) + betam * (1.0 + (dz[k] - dz[0]) * betas[0]);
            z[k] = betam / dr;
            betas[k] = betas[0] * z[k];
        }
        alphas[0] = rnorm2(r, r, mvfc) / rn;
        rn = rnorm2(r, r, mvfc);
        for (int k = 0; k < no; k++) {
            for (int j = 0; j < mvfc; j++) x[j + (size_t)k * mvfc] = x[j + (size_t)k * mvfc] + betas[k] * ps[j + (size_t)k * mvfc];
            alphas[k] = alphas[0] * z[k] * betas[k] / betas[0];
            gcoeff[k] = gcoeff[k] * z[k];
            for (int j = 0; j < mvfc; j++) ps[j + (size_t)k * mvfc] = ps[j + (size_t)k * mvfc] * alphas[k] + r[j] * gcoeff[k];
        }
        betam = betas[0];
        it++;
        if (it >= nitcgm) break;
    }
    for (int k = 0; k < no; k++)
        for (int j = 0; j < mvfc; j++) chi[j] = chi[j] + x[j + (size_t)k * mvfc] * bz[k];
    cplx *out = (cplx *)chi_aos;
    for (int i = 0; i < V; i++)
        for (int k = 0; k < 4; k++)
            for (int j = 0; j < nc; j++) out[(i * nc + j) * 4 + k] = chi[i + V * (k + j * 4)];
    free(x); free(ps); free(r); free(q); free(spbuf); free(ini); free(chi); free(tA); free(tB); free(tC);
    free(z); free(alphas); free(betas); free(gcoeff);
    return it;
}

/* OverlapOperator.cpp:25-64 (without the unused random vector of :29-47) and
 * OverlapU.cpp:152-208: 25 Zolotarev poles, tolerance 1e-12, CG 1e-16 / 1000 iterations. */
int orc_overlap_evo_zolo(const orc_lat *L, const double *ucc, const double *phi, double *psi,
                         double lmin, double lmax)
{
    const int no = 25;
    double bz[25], dz[25];
    int N = lat_vol(L) * L->nc * 4;
    orc_zolotarev(lmin, lmax, 1e-12, no, bz, dz);
    cplx *chi = malloc(N * sizeof(cplx)), *e = malloc(N * sizeof(cplx));
    int it = orc_zolotrr(L, ucc, bz, dz, no, 1e-16, 1000, phi, (double *)chi);
    orc_wilson_evo(L, ucc, (const double *)chi, (double *)e);
    overlap_combine(L, (const cplx *)phi, e, (cplx *)psi);
    free(chi); free(e);
    return it;
}

/* RitzSolver.cpp:92-297 RitzSolver, neigen = 1 (orthog is then a no-op, RitzSolver.cpp:6-23).
 * vec_flat is the normalised start vector on the flat layout. */
double orc_ritz(const orc_lat *L, const double *ucc, const double *vec_flat, double delitr,
                int nitlz, double sign, int nrestart)
{
    int V = lat_vol(L), nc = L->nc, mvfc = V * nc * 4;
    const cplx *vec = (const cplx *)vec_flat;
    cplx *xr = malloc(sizeof(cplx) * mvfc), *yr = malloc(sizeof(cplx) * mvfc), *zr = malloc(sizeof(cplx) * mvfc);
    cplx *gr = malloc(sizeof(cplx) * mvfc), *pr = malloc(sizeof(cplx) * mvfc);
    cplx *tA = malloc(sizeof(cplx) * mvfc), *tC = malloc(sizeof(cplx) * mvfc), *tB = malloc(sizeof(cplx) * mvfc);
    double xnorm, gnorm, pnorm, gnormim, beta, ct, st, ccc, cox, cox2, sox, sqp, sqm, sig1, sig2;
    cplx xpnorm, xynorm, amb, oxynorm;
    int ittot, cond;

    for (int i = 0; i < mvfc; i++) xr[i] = vec[i];
    xnorm = rnorm2(xr, xr, mvfc);
    xnorm = 1.0 / sqrt(xnorm);
    for (int i = 0; i < mvfc; i++) xr[i] = xr[i] * xnorm;
    ittot = 0;
    cond = 0;
    while (cond >= 0) {
        wilson_uu(L, ucc, xr, yr, tA, tC, tB);
        for (int i = 0; i < mvfc; i++) yr[i] = sign * yr[i];
        xynorm = cnorm2(xr, yr, mvfc);
        for (int i = 0; i < mvfc; i++) { gr[i] = yr[i] - xynorm * xr[i]; pr[i] = gr[i]; }
        gnormim = rnorm2(gr, gr, mvfc);
        pnorm = sqrt(gnormim);
        for (int i = 0; i < mvfc; i++) pr[i] = pr[i] / pnorm;
        for (int itt = 1; itt <= nrestart; itt++) {
            ittot = ittot + 1;
            wilson_uu(L, ucc, pr, zr, tA, tC, tB);
            for (int i = 0; i < mvfc; i++) zr[i] = sign * zr[i];
            amb = xynorm;
            amb = amb - rnorm2(pr, zr, mvfc);
            ccc = creal(cnorm2(xr, zr, mvfc));
            cox = 2.0 * ccc / creal(amb);
            sig1 = -copysign(1.0, creal(amb));
            sig2 = copysign(1.0, cox) * sig1;
            cox2 = cox * cox;
            if (cox2 <= 1.0e-3) {
                sqp = 1.0 - 0.125 * cox2 + 0.0859375 * cox2 * cox2 - 0.0673828125 * cox2 * cox2 * cox2 +
                      0.0562439 * cox2 * cox2 * cox2 * cox2 + 0.0487 * cox2 * cox2 * cox2 * cox2 * cox2;
                sqm = 0.5 * sig1 * sig2 * cox *
                      (1.0 - 0.375 * cox2 + 0.2421875 * cox2 * cox2 - 0.182617875 * cox2 * cox2 * cox2 +
                       0.1482849 * cox2 * cox2 * cox2 * cox2 - 0.12575 * cox2 * cox2 * cox2 * cox2 * cox2 +
                       0.1097 * cox2 * cox2 * cox2 * cox2 * cox2 * cox2);
                if (sig1 < 0.0) { ct = sqm; st = sig2 * sqp; }
                else { ct = sqp; st = sig2 * sqm; }
            } else {
                sox = sig1 / sqrt(1.0 + cox2);
                ct = sqrt((1.0 + sox) * 0.5);
                st = sqrt((1.0 - sox) * 0.5) * sig2;
            }
            for (int i = 0; i < mvfc; i++) {
                xr[i] = xr[i] * ct + pr[i] * st;
                yr[i] = yr[i] * ct + zr[i] * st;
            }
            xnorm = rnorm2(xr, xr, mvfc);
            oxynorm = xynorm;
            xynorm = cnorm2(xr, yr, mvfc);
            xnorm = 1.0 / sqrt(xnorm);
            xynorm = xynorm * xnorm;
            for (int i = 0; i < mvfc; i++) { xr[i] = xr[i] * xnorm; gr[i] = yr[i] - xynorm * xr[i]; }
            gnorm = rnorm2(gr, gr, mvfc);
            xpnorm = cnorm2(xr, pr, mvfc);
            if (creal(xynorm) > creal(oxynorm)) { cond = -1; break; }
            if (gnorm < delitr * sign * creal(xynorm)) { cond = -1; break; }
            beta = gnorm / gnormim * pnorm;
            for (int i = 0; i < mvfc; i++) pr[i] = gr[i] + beta * (pr[i] - xr[i] * xpnorm);
            pnorm = sqrt(rnorm2(pr, pr, mvfc));
            for (int i = 0; i < mvfc; i++) pr[i] = pr[i] / pnorm;
            gnormim = gnorm;
            if (ittot >= nitlz) { cond = -1; break; }
        }
    }
    double omeg1 = creal(xynorm);
    free(xr); free(yr); free(zr); free(gr); free(pr); free(tA); free(tB); free(tC);
    return omeg1;
}

/* ------------------------------------------------------------------------- */
/* leapfrog                                                                   */
/* ------------------------------------------------------------------------- */

static void apply_op(const orc_lat *L, const double *ucc, const orc_step_opts *o, double lam,
                     const double *coef, const double *in, double *out)
{
    if (o->kind == ORC_WILSON) orc_wilson_evo(L, ucc, in, out);
    else if (o->kind == ORC_OVERLAP_CHEB) orc_overlap_evo_cheb(L, ucc, lam, o->cheb_cut, o->cheb_n, coef, in, out);
    else orc_overlap_evo_zolo(L, ucc, in, out, o->zolo_lmin, o->zolo_lmax);
}

/* OverlapUpdate.cpp:85-124 PerformOverlapUpdate / WILSON/Update.cpp:6-45: target += dt * op(source) */
static void half_step(const orc_lat *L, const double *ucc, const orc_step_opts *o, double lam,
                      const double *coef, int nmodes, const double *src, double *tgt)
{
    int N = lat_vol(L) * L->nc * 4;
    int i;
#ifdef _OPENMP
    int nt = o->threads > 0 ? o->threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
#endif
    for (i = 0; i < nmodes; i++) {
        double *d = malloc(sizeof(double) * 2 * N);
        apply_op(L, ucc, o, lam, coef, src + (size_t)2 * N * i, d);
        double *t = tgt + (size_t)2 * N * i;
        for (int k = 0; k < 2 * N; k++) t[k] += o->dt * d[k];
        free(d);
    }
}

/* OverlapUpdate.cpp:175-211 / WILSON/Update.cpp:48-79 Dynamics::Fermions::Update, nsteps times.
 * psi, psimid: nmodes consecutive AoS mode functions.  update_links: 0 frozen gauge field,
 * 1 links move with E fixed (the reference), 2 E moves too (orc_efield_update before each link half step). */
void orc_leapfrog(const orc_lat *L, double *links, double *efield, int nmodes,
                  double *psi, double *psimid, const orc_step_opts *o, int nsteps)
{
    int V = lat_vol(L), nc = L->nc;
    double *ucc = malloc(sizeof(double) * 2 * (size_t)3 * V * nc * nc);
    double *ph = NULL;
    if (o->nb != 0) {
        ph = malloc(sizeof(double) * 2 * 3 * V);
        orc_u1_phases(L->n, o->nb, o->q, ph);
    }
    double *coef = NULL, lam = orc_hw_norm_bound(L);
    if (o->kind == ORC_OVERLAP_CHEB) {
        coef = malloc(sizeof(double) * (o->cheb_n + 1));
        orc_cheb_coeffs(o->cheb_n, o->cheb_cut, coef);
    }
    for (int st = 0; st < nsteps; st++) {
        orc_links_to_ucc(L, links, ph, ucc);
        half_step(L, ucc, o, lam, coef, nmodes, psimid, psi);
        if (o->update_links == 2) orc_efield_update(L, links, efield, 0.5 * o->dt);
        if (o->update_links) orc_update_links(L, links, efield, 0.5 * o->dt);
        orc_links_to_ucc(L, links, ph, ucc);
        half_step(L, ucc, o, lam, coef, nmodes, psi, psimid);
        if (o->update_links == 2) orc_efield_update(L, links, efield, 0.5 * o->dt);
        if (o->update_links) orc_update_links(L, links, efield, 0.5 * o->dt);
    }
    free(ucc); free(ph); free(coef);
}

/* ------------------------------------------------------------------------- */
/* reductions                                                                 */
/* ------------------------------------------------------------------------- */

static void lnk(const orc_lat *L, const double *links, int x, int y, int z, int mu, cplx *m)
{
    link_matrix(L->nc, links, site_of(L->n, x, y, z), mu, m);
}

/* plaquette U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+ (UUDD, AdvancedOperations.cpp:190-202) */
static void plaq(const orc_lat *L, const double *links, int x, int y, int z, int mu, int nu, cplx *p)
{
    int d[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    int nc = L->nc;
    cplx a[9], b[9], c[9], e[9];
    lnk(L, links, x, y, z, mu, a);
    lnk(L, links, x + d[mu][0], y + d[mu][1], z + d[mu][2], nu, b);
    lnk(L, links, x + d[nu][0], y + d[nu][1], z + d[nu][2], mu, c);
    lnk(L, links, x, y, z, nu, e);
    m_mul(nc, a, b, p);
    m_dag(nc, c, c);
    m_mul(nc, p, c, p);
    m_dag(nc, e, e);
    m_mul(nc, p, e, p);
}

/* E-field half of the gauge leapfrog.  Dynamics.cpp:145-197 updates E from plaquette traces
 * through macros that are absent from the reference, so the force is restated from the
 * Hamiltonian H = 1/2 sum E^2 + 2 sum ReTr(1 - U_plaq) with U' = -i E U (Dynamics.cpp:97-138):
 *   E^a_mu(x) += -2 dt sum_{nu != mu} [ReTr(i t^a U_{mu,nu}(x)) + ReTr(i t^a U_{mu,-nu}(x))]
 * Pinned by Gauss-law and energy conservation (tests), not by a reference vector. */
void orc_efield_update(const orc_lat *L, const double *links, double *efield, double dt)
{
    int V = lat_vol(L), nc = L->nc, ng = ngen(nc);
    int d[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    double *f = calloc((size_t)3 * V * ng, sizeof(double));
    for (int s = 0; s < V; s++) {
        int x, y, z;
        coords_of(L->n, s, &x, &y, &z);
        for (int mu = 0; mu < 3; mu++)
            for (int nu = 0; nu < 3; nu++) {
                if (nu == mu) continue;
                cplx p[9], a[9], b[9], c[9], e[9];
                double t[8];
                plaq(L, links, x, y, z, mu, nu, p);
                retr_itu(nc, p, t);
                for (int k = 0; k < ng; k++) f[ng * (mu + 3 * s) + k] += t[k];
                /* U_mu(x) U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu) */
                lnk(L, links, x, y, z, mu, a);
                lnk(L, links, x + d[mu][0] - d[nu][0], y + d[mu][1] - d[nu][1], z + d[mu][2] - d[nu][2], nu, b);
                lnk(L, links, x - d[nu][0], y - d[nu][1], z - d[nu][2], mu, c);
                lnk(L, links, x - d[nu][0], y - d[nu][1], z - d[nu][2], nu, e);
                m_dag(nc, b, b);
                m_dag(nc, c, c);
                m_mul(nc, a, b, p);
                m_mul(nc, p, c, p);
                m_mul(nc, p, e, p);
                retr_itu(nc, p, t);
                for (int k = 0; k < ng; k++) f[ng * (mu + 3 * s) + k] += t[k];
            }
    }
    for (int i = 0; i < 3 * V * ng; i++) efield[i] += -2.0 * dt * f[i];
    free(f);
}

/* Energy of the Kogut-Susskind Hamiltonian in the normalisation of BulkObservables.cpp:29-31,84-86:
 * E_el = 1/2 sum E^2,  E_mag = 1/2 sum_planes 4 ReTr(1 - U_plaq) (for Nc = 2: ReTrIDMinusU = tr(1-U)).
 * The plaquette macro of the reference is absent: plaquettes are the standard UUDD products. */
void orc_plaquette_energy(const orc_lat *L, const double *links, const double *efield,
                          double *e_el, double *e_mag)
{
    int V = lat_vol(L), nc = L->nc, ng = ngen(nc);
    double sel = 0, smag = 0;
    for (int i = 0; i < V * 3 * ng; i++) sel += efield[i] * efield[i];
    for (int s = 0; s < V; s++) {
        int x, y, z;
        coords_of(L->n, s, &x, &y, &z);
        int planes[3][2] = {{1, 2}, {2, 0}, {0, 1}};
        for (int p = 0; p < 3; p++) {
            cplx u[9];
            plaq(L, links, x, y, z, planes[p][0], planes[p][1], u);
            double tr = 0;
            for (int i = 0; i < nc; i++) tr += creal(1.0 - u[i * nc + i]);
            smag += 4.0 * tr;
        }
    }
    *e_el = 0.5 * sel;
    *e_mag = 0.5 * smag;
}

/* Gauss-law violation G^a(x) = sum_mu [E_mu(x) - U_mu(x-mu)^+ E_mu(x-mu) U_mu(x-mu)]^a,
 * max |G^a| and sqrt(sum G^2/(Ng V)) as GaussLawViolation.cpp:13-61 reduces them.  The reference's
 * COMPUTE_GAUSS_VIOLATION macro is absent; the transport follows ChernSimonsOperators.cpp:263-269. */
void orc_gauss_violation(const orc_lat *L, const double *links, const double *efield,
                         double *gmax, double *grms)
{
    int V = lat_vol(L), nc = L->nc, ng = ngen(nc);
    int d[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    double mx = 0, sq = 0;
    for (int s = 0; s < V; s++) {
        int x, y, z;
        coords_of(L->n, s, &x, &y, &z);
        cplx G[9] = {0};
        for (int mu = 0; mu < 3; mu++) {
            cplx e0[9], e1[9], u[9], ud[9], t[9];
            alg_matrix(nc, efield + ng * (mu + 3 * s), e0);
            int sm = site_of(L->n, x - d[mu][0], y - d[mu][1], z - d[mu][2]);
            alg_matrix(nc, efield + ng * (mu + 3 * sm), e1);
            link_matrix(nc, links, sm, mu, u);
            m_dag(nc, u, ud);
            m_mul(nc, ud, e1, t);
            m_mul(nc, t, u, t);
            for (int i = 0; i < nc * nc; i++) G[i] += e0[i] - t[i];
        }
        /* components: G^a = 2 tr(t^a G) */
        for (int a = 0; a < ng; a++) {
            cplx ta[9];
            gen_matrix(nc, a, ta);
            cplx tr = 0;
            for (int i = 0; i < nc; i++)
                for (int k = 0; k < nc; k++) tr += ta[i * nc + k] * G[k * nc + i];
            double g = fabs(2.0 * creal(tr));
            if (g > mx) mx = g;
            sq += g * g;
        }
    }
    *gmax = mx;
    *grms = sqrt(sq / (ng * V));
}

/* G^a(x) of every site, signed ([site][a]), as orc_gauss_violation forms it */
static void gauss_field(const orc_lat *L, const double *links, const double *efield, double *g)
{
    int V = lat_vol(L), nc = L->nc, ng = ngen(nc);
    int d[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int s = 0; s < V; s++) {
        int x, y, z;
        coords_of(L->n, s, &x, &y, &z);
        cplx G[9] = {0};
        for (int mu = 0; mu < 3; mu++) {
            cplx e0[9], e1[9], u[9], ud[9], t[9];
            alg_matrix(nc, efield + ng * (mu + 3 * s), e0);
            int sm = site_of(L->n, x - d[mu][0], y - d[mu][1], z - d[mu][2]);
            alg_matrix(nc, efield + ng * (mu + 3 * sm), e1);
            link_matrix(nc, links, sm, mu, u);
            m_dag(nc, u, ud);
            m_mul(nc, ud, e1, t);
            m_mul(nc, t, u, t);
            for (int i = 0; i < nc * nc; i++) G[i] += e0[i] - t[i];
        }
        for (int a = 0; a < ng; a++) {
            cplx ta[9], tr = 0;
            gen_matrix(nc, a, ta);
            for (int i = 0; i < nc; i++)
                for (int k = 0; k < nc; k++) tr += ta[i * nc + k] * G[k * nc + i];
            g[ng * s + a] = 2.0 * creal(tr);
        }
    }
}

/* GaussLawRestoration::UpdateStep (GaussRestore.cpp:115-120): UpdateDeviation stores G^a(x)
 * (:13-61), UpdateElectricFields sets E_mu(x) -= gamma (G(x) - V G(x+mu) V^+) with V = U_mu(x)
 * (:74-112, AdjointMultiplication SU2/AlgebraOperations.cpp:76-83).  Returns max |G^a| of the
 * field before the update. */
double orc_gauss_restore_step(const orc_lat *L, const double *links, double *efield, double gamma)
{
    int V = lat_vol(L), nc = L->nc, ng = ngen(nc);
    int d[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    double *g = malloc(sizeof(double) * (size_t)ng * V);
    double mx = 0;
    gauss_field(L, links, efield, g);
    for (int i = 0; i < ng * V; i++) mx = fmax(mx, fabs(g[i]));
    for (int s = 0; s < V; s++) {
        int x, y, z;
        coords_of(L->n, s, &x, &y, &z);
        for (int mu = 0; mu < 3; mu++) {
            int sp = site_of(L->n, x + d[mu][0], y + d[mu][1], z + d[mu][2]);
            cplx u[9], ud[9], gm[9], t[9];
            link_matrix(nc, links, s, mu, u);
            m_dag(nc, u, ud);
            alg_matrix(nc, g + ng * sp, gm);
            m_mul(nc, u, gm, t);
            m_mul(nc, t, ud, t);
            for (int a = 0; a < ng; a++) {
                cplx ta[9], tr = 0;
                gen_matrix(nc, a, ta);
                for (int i = 0; i < nc; i++)
                    for (int j = 0; j < nc; j++) tr += ta[i * nc + j] * t[j * nc + i];
                efield[ng * (mu + 3 * s) + a] -= gamma * (g[ng * s + a] - 2.0 * creal(tr));
            }
        }
    }
    free(g);
    return mx;
}

/* GaussLawRestoration::Restore (GaussRestore.cpp:129-160): the three steps gamma = 7/24, 7/48,
 * 7/72 are repeated while the maximum measured by the last step exceeds tol and fewer than
 * max_iter iterations ran.  The reference's maximum is max_a,x G^a(x) (signed, through a macro
 * absent from the reference); this one is max |G^a(x)|.  Returns the iteration count; *vmax =
 * max |G^a| of the final field. */
int orc_gauss_restore(const orc_lat *L, const double *links, double *efield, double tol, int max_iter,
                      double *vmax)
{
    int V = lat_vol(L), ng = ngen(L->nc);
    double *g = malloc(sizeof(double) * (size_t)ng * V);
    const double gam[3] = {7.0 / 24.0, 7.0 / 48.0, 7.0 / 72.0};
    double mx = 0;
    gauss_field(L, links, efield, g);
    for (int i = 0; i < ng * V; i++) mx = fmax(mx, fabs(g[i]));
    int it = 0;
    while (it < max_iter && mx > tol) {
        for (int k = 0; k < 3; k++) {
            double m = orc_gauss_restore_step(L, links, efield, gam[k]);
            if (k == 2) mx = m;
        }
        it++;
    }
    gauss_field(L, links, efield, g);
    mx = 0;
    for (int i = 0; i < ng * V; i++) mx = fmax(mx, fabs(g[i]));
    *vmax = mx;
    free(g);
    return it;
}

/* UnitarityViolation.cpp:13-51 with UnitarityNorm (SU2/GroupOperations.cpp:172-184) generalised:
 * || U U^+ - 1 ||_F in the group's own format (for SU(2): 4-real quaternion components). */
void orc_unitarity_violation(const orc_lat *L, const double *links, double *umax, double *urms)
{
    int V = lat_vol(L), nc = L->nc;
    double mx = 0, sq = 0;
    for (int s = 0; s < V; s++)
        for (int mu = 0; mu < 3; mu++) {
            double nrm = 0;
            if (nc == 2) {
                /* UD(V,V): quaternion components of V V^+ minus the unit (0,0,0,1) */
                const double *v = links + 4 * (mu + 3 * s);
                double c0 = -v[3] * v[0] + v[0] * v[3] + (v[1] * v[2] - v[2] * v[1]);
                double c1 = -v[3] * v[1] + v[1] * v[3] - (v[0] * v[2] - v[2] * v[0]);
                double c2 = -v[3] * v[2] + v[2] * v[3] + (v[0] * v[1] - v[1] * v[0]);
                double c3 = v[3] * v[3] + v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                nrm = sqrt(c0 * c0 + c1 * c1 + c2 * c2 + (c3 - 1.0) * (c3 - 1.0));
            } else {
                cplx u[9], ud[9], p[9];
                link_matrix(3, links, s, mu, u);
                m_dag(3, u, ud);
                m_mul(3, u, ud, p);
                for (int i = 0; i < 3; i++) p[i * 3 + i] -= 1.0;
                for (int i = 0; i < 9; i++) nrm += creal(p[i] * conj(p[i]));
                nrm = sqrt(nrm);
            }
            if (nrm > mx) mx = nrm;
            sq += nrm * nrm;
        }
    *umax = mx;
    *urms = sqrt(sq / (3.0 * V));
}

/* ChernSimonsOperators.cpp:228-439 ImprovedOperators::ComputeEDotB, restated in matrix form:
 * E: 4-point transported average (-1/12, 7/12, 7/12, -1/12); B = 2 ReTr(i t^a Ubar) with
 * Ubar = 5/12 (4 clover plaquettes) - 1/24 (8 rectangles), all transported to x. */
static void getl(const orc_lat *L, const double *links, int x, int y, int z, int mu, cplx *m)
{
    lnk(L, links, x, y, z, mu, m);
}

static void getlp(const orc_lat *L, const double *links, int x, int y, int z, const int *da, const int *db,
                  int i, int j, int mu, cplx *m)
{
    lnk(L, links, x + i * da[0] + j * db[0], y + i * da[1] + j * db[1], z + i * da[2] + j * db[2], mu, m);
}

static void prod4(int nc, const cplx *a, const cplx *b, const cplx *c, const cplx *d, int da, int db, int dc, int dd, cplx *out)
{
    cplx t[9], u[9];
    if (da) m_dag(nc, a, t); else memcpy(t, a, sizeof(cplx) * nc * nc);
    if (db) { m_dag(nc, b, u); m_mul(nc, t, u, t); } else m_mul(nc, t, b, t);
    if (dc) { m_dag(nc, c, u); m_mul(nc, t, u, t); } else m_mul(nc, t, c, t);
    if (dd) { m_dag(nc, d, u); m_mul(nc, t, u, t); } else m_mul(nc, t, d, t);
    memcpy(out, t, sizeof(cplx) * nc * nc);
}

static void prod6(int nc, const cplx *m[6], const int dg[6], cplx *out)
{
    cplx t[9], u[9];
    m_id(nc, t);
    for (int k = 0; k < 6; k++) {
        if (dg[k]) { m_dag(nc, m[k], u); m_mul(nc, t, u, t); }
        else m_mul(nc, t, m[k], t);
    }
    memcpy(out, t, sizeof(cplx) * nc * nc);
}

static void transport(int nc, const cplx *u, cplx *q) /* q <- u q u^+ (ParallelTransport, :225-233) */
{
    cplx ud[9];
    m_dag(nc, u, ud);
    m_mul(nc, u, q, q);
    m_mul(nc, q, ud, q);
}

double orc_edotb(const orc_lat *L, const double *links, const double *efield)
{
    int V = lat_vol(L), nc = L->nc, ng = ngen(nc), nn = nc * nc;
    double total = 0;
    for (int s = 0; s < V; s++) {
        int x, y, z;
        coords_of(L->n, s, &x, &y, &z);
        double B[3][8], Ev[3][8];
        for (int mu = 0; mu < 3; mu++) {
            /* axes: mu and its two partners in the cyclic order used for B_mu (yz, zx, xy) */
            int a = (mu + 1) % 3, b = (mu + 2) % 3;
            int dm[3] = {0, 0, 0}, da[3] = {0, 0, 0}, db[3] = {0, 0, 0};
            dm[mu] = 1; da[a] = 1; db[b] = 1;
#define P(ox, oy, oz) x + (ox), y + (oy), z + (oz)
            /* ---- E_mu transported average ---- */
            cplx acc[9] = {0}, em[9], u1[9], u2[9], tmp[9];
            alg_matrix(nc, efield + ng * (mu + 3 * s), em);
            for (int i = 0; i < nn; i++) acc[i] += (7.0 / 12.0) * em[i];
            getl(L, links, x, y, z, mu, u1);                       /* U(x) E(x+mu) U(x)^+ */
            alg_matrix(nc, efield + ng * (mu + 3 * site_of(L->n, P(dm[0], dm[1], dm[2]))), em);
            transport(nc, u1, em);
            for (int i = 0; i < nn; i++) acc[i] += -(1.0 / 12.0) * em[i];
            getl(L, links, P(-dm[0], -dm[1], -dm[2]), mu, u1);    /* U(x-mu)^+ E(x-mu) U(x-mu) */
            m_dag(nc, u1, u1);
            alg_matrix(nc, efield + ng * (mu + 3 * site_of(L->n, P(-dm[0], -dm[1], -dm[2]))), em);
            transport(nc, u1, em);
            for (int i = 0; i < nn; i++) acc[i] += (7.0 / 12.0) * em[i];
            getl(L, links, P(-2 * dm[0], -2 * dm[1], -2 * dm[2]), mu, u2); /* DD(U(x-mu), U(x-2mu)) */
            getl(L, links, P(-dm[0], -dm[1], -dm[2]), mu, tmp);
            m_dag(nc, tmp, tmp);
            m_dag(nc, u2, u2);
            m_mul(nc, tmp, u2, u2);
            alg_matrix(nc, efield + ng * (mu + 3 * site_of(L->n, P(-2 * dm[0], -2 * dm[1], -2 * dm[2]))), em);
            transport(nc, u2, em);
            for (int i = 0; i < nn; i++) acc[i] += -(1.0 / 12.0) * em[i];
            for (int c = 0; c < ng; c++) {
                cplx ta[9];
                gen_matrix(nc, c, ta);
                cplx tr = 0;
                for (int i = 0; i < nc; i++)
                    for (int k = 0; k < nc; k++) tr += ta[i * nc + k] * acc[k * nc + i];
                Ev[mu][c] = 2.0 * creal(tr);
            }
            /* ---- B_mu from the (a,b) plane ---- */
            cplx ubar[9] = {0}, box[9], l[6][9];
            cplx uaM[9], ubM[9], ua2M[9], ub2M[9];
            getl(L, links, P(-da[0], -da[1], -da[2]), a, uaM); m_dag(nc, uaM, uaM);
            getl(L, links, P(-db[0], -db[1], -db[2]), b, ubM); m_dag(nc, ubM, ubM);
            getl(L, links, P(-2 * da[0], -2 * da[1], -2 * da[2]), a, tmp); m_dag(nc, tmp, tmp);
            m_mul(nc, uaM, tmp, ua2M);
            getl(L, links, P(-2 * db[0], -2 * db[1], -2 * db[2]), b, tmp); m_dag(nc, tmp, tmp);
            m_mul(nc, ubM, tmp, ub2M);
            /* upper right box */
            getlp(L, links, x, y, z, da, db, 0, 0, a, l[0]); getlp(L, links, x, y, z, da, db, 1, 0, b, l[1]);
            getlp(L, links, x, y, z, da, db, 0, 1, a, l[2]); getlp(L, links, x, y, z, da, db, 0, 0, b, l[3]);
            prod4(nc, l[0], l[1], l[2], l[3], 0, 0, 1, 1, box);
            for (int i = 0; i < nn; i++) ubar[i] += (5.0 / 12.0) * box[i];
            /* upper left box */
            getlp(L, links, x, y, z, da, db, -1, 0, a, l[0]); getlp(L, links, x, y, z, da, db, 0, 0, b, l[1]);
            getlp(L, links, x, y, z, da, db, -1, 1, a, l[2]); getlp(L, links, x, y, z, da, db, -1, 0, b, l[3]);
            prod4(nc, l[0], l[1], l[2], l[3], 0, 0, 1, 1, box);
            transport(nc, uaM, box);
            for (int i = 0; i < nn; i++) ubar[i] += (5.0 / 12.0) * box[i];
            /* lower left box (DDUU) */
            getlp(L, links, x, y, z, da, db, -1, 0, a, l[0]); getlp(L, links, x, y, z, da, db, -1, -1, b, l[1]);
            getlp(L, links, x, y, z, da, db, -1, -1, a, l[2]); getlp(L, links, x, y, z, da, db, 0, -1, b, l[3]);
            prod4(nc, l[0], l[1], l[2], l[3], 1, 1, 0, 0, box);
            for (int i = 0; i < nn; i++) ubar[i] += (5.0 / 12.0) * box[i];
            /* lower right box */
            getlp(L, links, x, y, z, da, db, 0, -1, a, l[0]); getlp(L, links, x, y, z, da, db, 1, -1, b, l[1]);
            getlp(L, links, x, y, z, da, db, 0, 0, a, l[2]); getlp(L, links, x, y, z, da, db, 0, -1, b, l[3]);
            prod4(nc, l[0], l[1], l[2], l[3], 0, 0, 1, 1, box);
            transport(nc, ubM, box);
            for (int i = 0; i < nn; i++) ubar[i] += (5.0 / 12.0) * box[i];
            /* eight rectangles */
            struct { int o[6][2]; int dir[6]; int dg[6]; int tr; } R[8] = {
                /* X2Y: U_a(0,0) U_a(1,0) U_b(2,0) U_a(1,1)^+ U_a(0,1)^+ U_b(0,0)^+ */
                {{{0, 0}, {1, 0}, {2, 0}, {1, 1}, {0, 1}, {0, 0}}, {0, 0, 1, 0, 0, 1}, {0, 0, 0, 1, 1, 1}, 0},
                /* XY2: U_a(0,0) U_b(1,0) U_b(1,1) U_a(0,2)^+ U_b(0,1)^+ U_b(0,0)^+ */
                {{{0, 0}, {1, 0}, {1, 1}, {0, 2}, {0, 1}, {0, 0}}, {0, 1, 1, 0, 1, 1}, {0, 0, 0, 1, 1, 1}, 0},
                /* MXY2: U_a(-1,0) U_b(0,0) U_b(0,1) U_a(-1,2)^+ U_b(-1,1)^+ U_b(-1,0)^+ ; transport uaM */
                {{{-1, 0}, {0, 0}, {0, 1}, {-1, 2}, {-1, 1}, {-1, 0}}, {0, 1, 1, 0, 1, 1}, {0, 0, 0, 1, 1, 1}, 1},
                /* MX2Y: U_a(-2,0) U_a(-1,0) U_b(0,0) U_a(-1,1)^+ U_a(-2,1)^+ U_b(-2,0)^+ ; transport ua2M */
                {{{-2, 0}, {-1, 0}, {0, 0}, {-1, 1}, {-2, 1}, {-2, 0}}, {0, 0, 1, 0, 0, 1}, {0, 0, 0, 1, 1, 1}, 2},
                /* MXMY2: U_a(-1,0)^+ U_b(-1,-1)^+ U_b(-1,-2)^+ U_a(-1,-2) U_b(0,-2) U_b(0,-1) */
                {{{-1, 0}, {-1, -1}, {-1, -2}, {-1, -2}, {0, -2}, {0, -1}}, {0, 1, 1, 0, 1, 1}, {1, 1, 1, 0, 0, 0}, 0},
                /* MX2MY: U_a(-1,0)^+ U_a(-2,0)^+ U_b(-2,-1)^+ U_a(-2,-1) U_a(-1,-1) U_b(0,-1) */
                {{{-1, 0}, {-2, 0}, {-2, -1}, {-2, -1}, {-1, -1}, {0, -1}}, {0, 0, 1, 0, 0, 1}, {1, 1, 1, 0, 0, 0}, 0},
                /* X2MY: U_a(0,-1) U_a(1,-1) U_b(2,-1) U_a(1,0)^+ U_a(0,0)^+ U_b(0,-1)^+ ; transport ubM */
                {{{0, -1}, {1, -1}, {2, -1}, {1, 0}, {0, 0}, {0, -1}}, {0, 0, 1, 0, 0, 1}, {0, 0, 0, 1, 1, 1}, 3},
                /* XMY2: U_a(0,-2) U_b(1,-2) U_b(1,-1) U_a(0,0)^+ U_b(0,-1)^+ U_b(0,-2)^+ ; transport ub2M */
                {{{0, -2}, {1, -2}, {1, -1}, {0, 0}, {0, -1}, {0, -2}}, {0, 1, 1, 0, 1, 1}, {0, 0, 0, 1, 1, 1}, 4},
            };
            for (int r = 0; r < 8; r++) {
                const cplx *mp[6];
                cplx rect[9];
                for (int k = 0; k < 6; k++) {
                    int dir = R[r].dir[k] ? b : a;
                    getlp(L, links, x, y, z, da, db, R[r].o[k][0], R[r].o[k][1], dir, l[k]);
                    mp[k] = l[k];
                }
                prod6(nc, mp, R[r].dg, rect);
                if (R[r].tr == 1) transport(nc, uaM, rect);
                if (R[r].tr == 2) transport(nc, ua2M, rect);
                if (R[r].tr == 3) transport(nc, ubM, rect);
                if (R[r].tr == 4) transport(nc, ub2M, rect);
                for (int i = 0; i < nn; i++) ubar[i] += -(1.0 / 24.0) * rect[i];
            }
#undef P
            double rt[8];
            retr_itu(nc, ubar, rt);
            for (int c = 0; c < ng; c++) B[mu][c] = 2.0 * rt[c];
        }
        for (int mu = 0; mu < 3; mu++)
            for (int c = 0; c < ng; c++) total += Ev[mu][c] * B[mu][c];
    }
    return total;
}

/* OverlapCurrents.cpp:137-228: per-site vector and axial currents
 * j^mu = 1/(2V) sum_s w_s Re[ psi^+ G^mu psi + psi^+ G2^mu (h psi) ], w_s = +1 (even global s) / -1,
 * G_V = g0 g^mu, G_A = g0 g^mu g5, G2 = -(i/2) G g0;  hpsi = -i h_ov psi of the same mode.
 * jv, ja: [4*site + mu] accumulated (not cleared). */
void orc_currents(const orc_lat *L, const double *psi_, const double *hpsi_, int nmodes,
                  int first_mode, double *jv, double *ja)
{
    static const cplx g0[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, -1, 0}, {0, 0, 0, -1}};
    static const cplx gx[4][4] = {{0, 0, 0, 1}, {0, 0, 1, 0}, {0, -1, 0, 0}, {-1, 0, 0, 0}};
    const cplx gy[4][4] = {{0, 0, 0, -I}, {0, 0, I, 0}, {0, I, 0, 0}, {-I, 0, 0, 0}};
    static const cplx gz[4][4] = {{0, 0, 1, 0}, {0, 0, 0, -1}, {-1, 0, 0, 0}, {0, 1, 0, 0}};
    static const cplx g5[4][4] = {{0, 0, 1, 0}, {0, 0, 0, 1}, {1, 0, 0, 0}, {0, 1, 0, 0}};
    const cplx(*gm[4])[4] = {g0, gx, gy, gz};
    cplx GV[4][4][4], GA[4][4][4], GV2[4][4][4], GA2[4][4][4];
    for (int mu = 0; mu < 4; mu++)
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) {
                cplx v = 0, ax = 0, v2 = 0, a2 = 0;
                for (int g = 0; g < 4; g++) {
                    v += g0[a][g] * gm[mu][g][b];
                    for (int d = 0; d < 4; d++) {
                        ax += g0[a][g] * gm[mu][g][d] * g5[d][b];
                        v2 += -0.5 * I * g0[a][g] * gm[mu][g][d] * g0[d][b];
                        for (int e = 0; e < 4; e++) a2 += -0.5 * I * g0[a][g] * gm[mu][g][d] * g5[d][e] * g0[e][b];
                    }
                }
                GV[mu][a][b] = v; GA[mu][a][b] = ax; GV2[mu][a][b] = v2; GA2[mu][a][b] = a2;
            }
    int V = lat_vol(L), nc = L->nc, N = V * nc * 4;
    double pre = 1.0 / (2.0 * V);
    for (int s = 0; s < nmodes; s++) {
        const cplx *p = (const cplx *)psi_ + (size_t)N * s;
        const cplx *h = (const cplx *)hpsi_ + (size_t)N * s;
        double w = ((s + first_mode) % 2 == 0) ? 1.0 : -1.0;
        for (int ix = 0; ix < V; ix++) {
            cplx jvv[4] = {0}, jaa[4] = {0};
            for (int i = 0; i < nc; i++)
                for (int a = 0; a < 4; a++)
                    for (int b = 0; b < 4; b++) {
                        cplx pa = conj(p[(ix * nc + i) * 4 + a]);
                        cplx pb = p[(ix * nc + i) * 4 + b];
                        cplx hb = h[(ix * nc + i) * 4 + b];
                        for (int mu = 0; mu < 4; mu++) {
                            jvv[mu] += GV[mu][a][b] * pa * pb * w + GV2[mu][a][b] * pa * hb * w;
                            jaa[mu] += GA[mu][a][b] * pa * pb * w + GA2[mu][a][b] * pa * hb * w;
                        }
                    }
            for (int mu = 0; mu < 4; mu++) {
                jv[mu + 4 * ix] += pre * creal(jvv[mu]);
                ja[mu + 4 * ix] += pre * creal(jaa[mu]);
            }
        }
    }
}
