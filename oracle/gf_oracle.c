/*
 * gf_oracle.c -- CPU FP64 restatement of Gaugefields.jl's SU(3) gauge-link hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library; the HIP product never uses it.
 *
 * Every function names the reference file:line it restates (paths relative to the
 * Gaugefields.jl v1.0.5 source tree).  The restatement follows the reference's formulas,
 * not its rounding order.  It is pinned by tests/test_oracle.py against: the reference's
 * instanton plaquette golden (test/init.jl:351-371, test/sun_embedded_instanton.jl:145),
 * cold-start identities, scipy.linalg.expm / expm_frechet, central finite differences of
 * the potential (force, stout force), a numpy whole-lattice brute force, and gauge
 * invariance / covariance.
 *
 * Layouts are the reference's Julia layouts (see gf_oracle.h).
 */
#include "gf_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;
typedef struct {
    cplx e[3][3];
} m3;

/* ------------------------------------------------------------------------- */
/* 3x3 complex algebra                                                        */
/* ------------------------------------------------------------------------- */
static inline m3 m3_zero(void) {
    m3 r;
    memset(&r, 0, sizeof r);
    return r;
}
static inline m3 m3_id(void) {
    m3 r = m3_zero();
    r.e[0][0] = r.e[1][1] = r.e[2][2] = 1.0;
    return r;
}
static inline m3 mul(m3 a, m3 b) { /* a b */
    m3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            r.e[i][j] = a.e[i][0] * b.e[0][j] + a.e[i][1] * b.e[1][j] + a.e[i][2] * b.e[2][j];
    return r;
}
static inline m3 mul_nd(m3 a, m3 b) { /* a b^dag */
    m3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            r.e[i][j] = a.e[i][0] * conj(b.e[j][0]) + a.e[i][1] * conj(b.e[j][1]) +
                        a.e[i][2] * conj(b.e[j][2]);
    return r;
}
static inline m3 mul_dn(m3 a, m3 b) { /* a^dag b */
    m3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            r.e[i][j] = conj(a.e[0][i]) * b.e[0][j] + conj(a.e[1][i]) * b.e[1][j] +
                        conj(a.e[2][i]) * b.e[2][j];
    return r;
}
static inline m3 dag(m3 a) {
    m3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.e[i][j] = conj(a.e[j][i]);
    return r;
}
static inline m3 add(m3 a, m3 b) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) a.e[i][j] += b.e[i][j];
    return a;
}
static inline m3 scale(cplx s, m3 a) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) a.e[i][j] *= s;
    return a;
}
static inline cplx trace(m3 a) { return a.e[0][0] + a.e[1][1] + a.e[2][2]; }

/* column-major load/store of one matrix (Julia (3,3) slice) */
static inline m3 ld(const double *p) {
    m3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.e[i][j] = CMPLX(p[2 * (i + 3 * j)], p[2 * (i + 3 * j) + 1]);
    return r;
}
static inline void st(double *p, m3 a) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            p[2 * (i + 3 * j)] = creal(a.e[i][j]);
            p[2 * (i + 3 * j) + 1] = cimag(a.e[i][j]);
        }
}

/* ------------------------------------------------------------------------- */
/* lattice geometry                                                           */
/* ------------------------------------------------------------------------- */
typedef struct {
    int n[4];
    long V;
    long *up; /* up[mu*V + s] = s + mu-hat */
    long *dn; /* dn[mu*V + s] = s - mu-hat */
} lat;

/* periodic shift, shifted_U! (gaugefields_4D_nowing.jl:380-412) */
static lat lat_make(const int *dims) {
    lat L;
    L.V = 1;
    for (int m = 0; m < 4; m++) {
        L.n[m] = dims[m];
        L.V *= dims[m];
    }
    L.up = (long *)malloc(sizeof(long) * 4 * L.V);
    L.dn = (long *)malloc(sizeof(long) * 4 * L.V);
    for (long s = 0; s < L.V; s++) {
        int x[4];
        long r = s;
        for (int m = 0; m < 4; m++) {
            x[m] = (int)(r % L.n[m]);
            r /= L.n[m];
        }
        for (int mu = 0; mu < 4; mu++) {
            for (int d = -1; d <= 1; d += 2) {
                int y[4] = {x[0], x[1], x[2], x[3]};
                y[mu] = (y[mu] + d + L.n[mu]) % L.n[mu];
                long t = y[0] + (long)L.n[0] * (y[1] + (long)L.n[1] * (y[2] + (long)L.n[2] * y[3]));
                if (d > 0)
                    L.up[mu * L.V + s] = t;
                else
                    L.dn[mu * L.V + s] = t;
            }
        }
    }
    return L;
}
static void lat_free(lat *L) {
    free(L->up);
    free(L->dn);
}
#define UP(L, mu, s) ((L)->up[(long)(mu) * (L)->V + (s)])
#define DN(L, mu, s) ((L)->dn[(long)(mu) * (L)->V + (s)])
#define LINK(U, L, mu, s) ld((U) + 18 * ((long)(mu) * (L)->V + (s)))
#define LINKP(U, L, mu, s) ((U) + 18 * ((long)(mu) * (L)->V + (s)))

static int g_threads = 0;
void gfo_set_threads(int n) { g_threads = n; }
int gfo_get_threads(void) {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}
#ifdef _OPENMP
#define OMP_NT num_threads(gfo_get_threads())
#endif

/* deterministic reduction: per-thread static blocks, summed in thread order */
typedef double (*site_fn)(const lat *L, long s, const void *ctx);
static double lattice_sum(const lat *L, site_fn f, const void *ctx) {
    int nt = gfo_get_threads();
    double *part = (double *)calloc((size_t)nt, sizeof(double));
#ifdef _OPENMP
#pragma omp parallel OMP_NT
#endif
    {
        int tid = 0, n = 1;
#ifdef _OPENMP
        tid = omp_get_thread_num();
        n = omp_get_num_threads();
#endif
        long lo = L->V * tid / n, hi = L->V * (tid + 1) / n;
        double acc = 0.0;
        for (long s = lo; s < hi; s++) acc += f(L, s, ctx);
        part[tid] = acc;
    }
    double tot = 0.0;
    for (int t = 0; t < nt; t++) tot += part[t];
    free(part);
    return tot;
}

/* ------------------------------------------------------------------------- */
/* TA projection and su(3) basis                                              */
/* ------------------------------------------------------------------------- */
static const double SR3I = 0.57735026918962576451; /* 1/sqrt(3) */
static const double GFO_PI = 3.14159265358979323846;

/* Traceless_antihermitian_add! for NC=3 (TA_gaugefields_4D_serial.jl:181-269):
 * c_a += factor * coefficient of TA(v) in the basis i lambda_a / 2. */
static inline void ta_add(m3 v, double factor, double *c) {
    double i11 = cimag(v.e[0][0]), i22 = cimag(v.e[1][1]), i33 = cimag(v.e[2][2]);
    double tri = (i11 + i22 + i33) / 3.0;
    cplx y11 = I * (i11 - tri), y22 = I * (i22 - tri), y33 = I * (i33 - tri);
    cplx x12 = v.e[0][1] - conj(v.e[1][0]);
    cplx x13 = v.e[0][2] - conj(v.e[2][0]);
    cplx x23 = v.e[1][2] - conj(v.e[2][1]);
    cplx y12 = 0.5 * x12, y13 = 0.5 * x13, y23 = 0.5 * x23;
    cplx y21 = -0.5 * conj(x12), y31 = -0.5 * conj(x13), y32 = -0.5 * conj(x23);
    c[0] += (cimag(y12) + cimag(y21)) * factor;
    c[1] += (creal(y12) - creal(y21)) * factor;
    c[2] += (cimag(y11) - cimag(y22)) * factor;
    c[3] += (cimag(y13) + cimag(y31)) * factor;
    c[4] += (creal(y13) - creal(y31)) * factor;
    c[5] += (cimag(y23) + cimag(y32)) * factor;
    c[6] += (creal(y23) - creal(y32)) * factor;
    c[7] += SR3I * (cimag(y11) + cimag(y22) - 2.0 * cimag(y33)) * factor;
}

/* Traceless_antihermitian! as a matrix (gaugefields_4D_nowing.jl:1253-1343):
 * (M - M^dag)/2 - tr(M - M^dag)/6 */
static inline m3 ta_mat(m3 v) {
    double tri = (cimag(v.e[0][0]) + cimag(v.e[1][1]) + cimag(v.e[2][2])) / 3.0;
    m3 r;
    for (int i = 0; i < 3; i++) {
        r.e[i][i] = I * (cimag(v.e[i][i]) - tri);
        for (int j = i + 1; j < 3; j++) {
            cplx x = 0.5 * (v.e[i][j] - conj(v.e[j][i]));
            r.e[i][j] = x;
            r.e[j][i] = -conj(x);
        }
    }
    return r;
}

/* Hermitian H = (t/2) sum_a c_a lambda_a, element layout of the legacy exptU!
 * (TA_gaugefields_4D_serial.jl:790-840: v1..v18 built from c1..c8) */
static inline m3 herm_from_coeffs(const double *c, double t) {
    double c1 = 0.5 * t * c[0], c2 = 0.5 * t * c[1], c3 = 0.5 * t * c[2], c4 = 0.5 * t * c[3];
    double c5 = 0.5 * t * c[4], c6 = 0.5 * t * c[5], c7 = 0.5 * t * c[6], c8 = 0.5 * t * c[7];
    m3 h;
    h.e[0][0] = c3 + SR3I * c8;
    h.e[0][1] = CMPLX(c1, -c2);
    h.e[0][2] = CMPLX(c4, -c5);
    h.e[1][0] = CMPLX(c1, c2);
    h.e[1][1] = -c3 + SR3I * c8;
    h.e[1][2] = CMPLX(c6, -c7);
    h.e[2][0] = CMPLX(c4, c5);
    h.e[2][1] = CMPLX(c6, c7);
    h.e[2][2] = -2.0 * SR3I * c8;
    return h;
}

/* ------------------------------------------------------------------------- */
/* Cayley-Hamilton exponential (Morningstar & Peardon, PRD 69 054501 (2004))  */
/* ------------------------------------------------------------------------- */
/* calc_coefficients_Q (AbstractGaugefields.jl:3284-3343) for Hermitian traceless Q:
 * exp(iQ) = f0 + f1 Q + f2 Q^2 and, when b != NULL, the derivative coefficients
 * b[0][j] = d f_j / d c1, b[1][j] = d f_j / d c0 (the reference's b1j, b2j).
 * The reference evaluates the formulas directly for any sign of c0; here the published
 * algorithm's c0 -> -c0 symmetry is used (f_j(-c0) = (-1)^j f_j*(c0),
 * b_ij(-c0) = (-1)^(i+j+1) b_ij*(c0)) so that 9u^2 - w^2 >= 2 c1 never vanishes, and
 * sin(w)/w, cos(w)/w^2 - sin(w)/w^3 use their Taylor series for |w| < 0.05. */
static void ch_coeffs(m3 Q, cplx f[3], cplx b[2][3]) {
    m3 Q2 = mul(Q, Q);
    double c1 = 0.5 * creal(trace(Q2));
    cplx det = Q.e[0][0] * (Q.e[1][1] * Q.e[2][2] - Q.e[1][2] * Q.e[2][1]) -
               Q.e[0][1] * (Q.e[1][0] * Q.e[2][2] - Q.e[1][2] * Q.e[2][0]) +
               Q.e[0][2] * (Q.e[1][0] * Q.e[2][1] - Q.e[1][1] * Q.e[2][0]);
    double c0 = creal(det);
    int neg = c0 < 0.0;
    double c0a = fabs(c0);
    double c0max = 2.0 * pow(c1 / 3.0, 1.5);
    double ratio = c0max > 0.0 ? c0a / c0max : 1.0;
    if (ratio > 1.0) ratio = 1.0;
    double th = acos(ratio);
    double u = sqrt(c1 / 3.0) * cos(th / 3.0);
    double w = sqrt(c1) * sin(th / 3.0);
    double w2 = w * w, u2 = u * u;
    double xi0, xi1;
    if (fabs(w) < 0.05) {
        xi0 = 1.0 - w2 / 6.0 * (1.0 - w2 / 20.0 * (1.0 - w2 / 42.0));
        xi1 = -1.0 / 3.0 + w2 / 30.0 * (1.0 - w2 / 28.0 * (1.0 - w2 / 54.0));
    } else {
        xi0 = sin(w) / w;
        xi1 = cos(w) / w2 - sin(w) / (w2 * w);
    }
    double cw = cos(w);
    cplx emiu = cexp(-I * u), e2iu = cexp(2.0 * I * u);
    cplx h0 = (u2 - w2) * e2iu + emiu * (8.0 * u2 * cw + 2.0 * I * u * (3.0 * u2 + w2) * xi0);
    cplx h1 = 2.0 * u * e2iu - emiu * (2.0 * u * cw - I * (3.0 * u2 - w2) * xi0);
    cplx h2 = e2iu - emiu * (cw + 3.0 * I * u * xi0);
    double den = 9.0 * u2 - w2;
    f[0] = h0 / den;
    f[1] = h1 / den;
    f[2] = h2 / den;
    if (b) {
        cplx r1[3], r2[3];
        r1[0] = 2.0 * (u + I * (u2 - w2)) * e2iu +
                2.0 * emiu *
                    (4.0 * u * (2.0 - I * u) * cw + I * (9.0 * u2 + w2 - I * u * (3.0 * u2 + w2)) * xi0);
        r1[1] = 2.0 * (1.0 + 2.0 * I * u) * e2iu +
                emiu * (-2.0 * (1.0 - I * u) * cw + I * (6.0 * u + I * (w2 - 3.0 * u2)) * xi0);
        r1[2] = 2.0 * I * e2iu + I * emiu * (cw - 3.0 * (1.0 - I * u) * xi0);
        r2[0] = -2.0 * e2iu + 2.0 * I * u * emiu * (cw + (1.0 + 4.0 * I * u) * xi0 + 3.0 * u2 * xi1);
        r2[1] = -I * emiu * (cw + (1.0 + 2.0 * I * u) * xi0 - 3.0 * u2 * xi1);
        r2[2] = emiu * (xi0 - 3.0 * I * u * xi1);
        double d2 = 2.0 * den * den;
        for (int j = 0; j < 3; j++) {
            b[0][j] = (2.0 * u * r1[j] + (3.0 * u2 - w2) * r2[j] - 2.0 * (15.0 * u2 + w2) * f[j]) / d2;
            b[1][j] = (r1[j] - 3.0 * u * r2[j] - 24.0 * u * f[j]) / d2;
        }
    }
    if (neg) {
        for (int j = 0; j < 3; j++) {
            double sg = (j % 2) ? -1.0 : 1.0;
            f[j] = sg * conj(f[j]);
            if (b) {
                b[0][j] = sg * conj(b[0][j]);
                b[1][j] = -sg * conj(b[1][j]);
            }
        }
    }
}

/* c1 below which exp(iQ) uses 1 + iQ - Q^2/2 (the formulas divide by 9u^2-w^2 ~ 2 c1) */
#define CH_EXP_TINY 1e-20
/* c1 below which C star dexp(Q)/dQ uses its power series (the b_ij formulas cancel) */
#define CH_DER_SERIES 1e-3
#define CH_DER_TERMS 9

/* exp(iH) for Hermitian traceless H */
static m3 expi_herm(m3 H) {
    m3 H2 = mul(H, H);
    double c1 = 0.5 * creal(trace(H2));
    m3 r;
    if (c1 < CH_EXP_TINY) {
        r = add(m3_id(), add(scale(I, H), scale(-0.5, H2)));
        return r;
    }
    cplx f[3];
    ch_coeffs(H, f, NULL);
    r = add(scale(f[0], m3_id()), add(scale(f[1], H), scale(f[2], H2)));
    return r;
}

/* exptU! for TA fields (TA_gaugefields_4D_serial.jl:760-1077): exp(t sum_a c_a i lambda_a/2).
 * The reference diagonalizes H; the same matrix exponential is formed here with the
 * Cayley-Hamilton coefficients of calc_coefficients_Q. */
static m3 exp_ta(const double *c, double t) { return expi_herm(herm_from_coeffs(c, t)); }

/* CdexpQdQ! (stout_fast.jl:888-946) with construct_B1B2!/construct_trCB1B2/
 * construct_CdeQdQ_3! (stout_fast.jl:1031-1081): L such that d tr(C e^Q) = tr(L dQ) for
 * anti-Hermitian traceless Q.  For c1 < CH_DER_SERIES the same derivative is summed from
 * the series  L = sum_{n>=1} 1/n! sum_{k<n} Q^(n-1-k) C Q^k  (the reference instead
 * leaves the output unwritten for tr Q^2 <= 1e-18). */
static m3 cdexpqdq(m3 C, m3 Qa) {
    m3 Qh = scale(-I, Qa); /* Q / i */
    m3 Qh2 = mul(Qh, Qh);
    double c1 = 0.5 * creal(trace(Qh2));
    if (c1 < CH_DER_SERIES) {
        m3 L = m3_zero(), Ln = C, CXn = C;
        double fact = 1.0;
        for (int n = 1; n <= CH_DER_TERMS; n++) {
            fact *= n;
            L = add(L, scale(1.0 / fact, Ln));
            CXn = mul(CXn, Qa);
            Ln = add(mul(Qa, Ln), CXn);
        }
        return L;
    }
    cplx f[3], b[2][3];
    ch_coeffs(Qh, f, b);
    m3 B1 = add(scale(b[0][0], m3_id()), add(scale(b[0][1], Qh), scale(b[0][2], Qh2)));
    m3 B2 = add(scale(b[1][0], m3_id()), add(scale(b[1][1], Qh), scale(b[1][2], Qh2)));
    cplx t1 = trace(mul(C, B1)), t2 = trace(mul(C, B2));
    m3 L = add(scale(t1, Qh), scale(t2, Qh2));
    L = add(L, scale(f[1], C));
    L = add(L, scale(f[2], add(mul(Qh, C), mul(C, Qh))));
    return scale(-I, L); /* divide by i */
}

/* public single-matrix entry points (used by the tests to pin the primitives) */
void gfo_ta_coeffs(const double *m, double *c8) {
    for (int a = 0; a < 8; a++) c8[a] = 0.0;
    ta_add(ld(m), 1.0, c8);
}
void gfo_ta_matrix(const double *m, double *out) { st(out, ta_mat(ld(m))); }
void gfo_exp_ta(const double *c8, double t, double *out) { st(out, exp_ta(c8, t)); }
void gfo_exp_antiherm(const double *q, double *out) { st(out, expi_herm(scale(-I, ld(q)))); }
void gfo_cdexpqdq(const double *c, const double *q, double *out) { st(out, cdexpqdq(ld(c), ld(q))); }

/* ------------------------------------------------------------------------- */
/* staples                                                                    */
/* ------------------------------------------------------------------------- */
/* construct_staple! (AbstractGaugefields.jl:2873-2923):
 * V_mu(x) = sum_{nu != mu} U_nu(x) U_mu(x+nu) U_nu(x+mu)^dag */
static m3 staple_upper(const lat *L, const double *U, int mu, long s) {
    m3 V = m3_zero();
    for (int nu = 0; nu < 4; nu++) {
        if (nu == mu) continue;
        m3 U1U2 = mul(LINK(U, L, nu, s), LINK(U, L, mu, UP(L, nu, s)));
        V = add(V, mul_nd(U1U2, LINK(U, L, nu, UP(L, mu, s))));
    }
    return V;
}

/* loops_staple_prime / make_plaq_staple_prime (autostaples/wilsonloops.jl:468-485),
 * evaluated by construct_double_staple! (AbstractGaugefields.jl:2856-2871); also the
 * stout C_mu / rho of calc_C! (stout_fast.jl:603-624) for the plaquette loop set:
 * V_mu(x) = sum_{nu != mu} [U_nu(x) U_mu(x+nu) U_nu(x+mu)^dag
 *                          + U_nu(x-nu)^dag U_mu(x-nu) U_nu(x-nu+mu)] */
static m3 staple_full(const lat *L, const double *U, int mu, long s) {
    m3 V = m3_zero();
    for (int nu = 0; nu < 4; nu++) {
        if (nu == mu) continue;
        m3 a = mul(LINK(U, L, nu, s), LINK(U, L, mu, UP(L, nu, s)));
        V = add(V, mul_nd(a, LINK(U, L, nu, UP(L, mu, s))));
        long sm = DN(L, nu, s);
        m3 b = mul_dn(LINK(U, L, nu, sm), LINK(U, L, mu, sm));
        V = add(V, mul(b, LINK(U, L, nu, UP(L, mu, sm))));
    }
    return V;
}

/* ------------------------------------------------------------------------- */
/* plaquette, action                                                          */
/* ------------------------------------------------------------------------- */
static double plaq_site(const lat *L, long s, const void *ctx) {
    const double *U = (const double *)ctx;
    double acc = 0.0;
    for (int mu = 0; mu < 4; mu++) {
        m3 V = staple_upper(L, U, mu, s);
        acc += creal(trace(mul_nd(LINK(U, L, mu, s), V)));
    }
    return acc;
}

/* calculate_Plaquette (AbstractGaugefields.jl:2684-2699): real(0.5 sum_mu tr(U_mu V_mu^dag))
 * = sum_{x, mu<nu} Re tr P_munu(x).  measure_plaquette (API.jl:395-404) divides by 6*V*3. */
double gfo_plaquette_sum(const int *dims, const double *U) {
    lat L = lat_make(dims);
    double v = 0.5 * lattice_sum(&L, plaq_site, U);
    lat_free(&L);
    return v;
}

/* md_potential(::GaugeAction) (molecular_dynamics.jl:247-249) for push!(action, coeff,
 * plaquette ∪ plaquette^dag): evaluate_GaugeAction (GaugeActions.jl:132-142) sums coeff*tr
 * over 12 oriented loops per site = 2 coeff sum Re tr P; potential = -that / NC.  The same
 * normalization is restated in test/enzyme_md.jl:17-40. */
double gfo_potential(const int *dims, const double *U, double coeff) {
    return -2.0 * coeff * gfo_plaquette_sum(dims, U) / 3.0;
}

/* ------------------------------------------------------------------------- */
/* derivative, force, kick, exp update                                        */
/* ------------------------------------------------------------------------- */
/* calc_dSdUμ! (GaugeActions.jl:95-123) with the six plaquette staples per direction that
 * Wilsonloop.make_staple yields for plaquette ∪ plaquette^dag (GaugeActions.jl:48-62):
 * coeff * sum_{nu != mu} [U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
 *                        + U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu)]
 * so that tr(U_mu(x) dSdU_mu(x)) closes the plaquettes through U_mu(x). */
static m3 dsdu_site(const lat *L, const double *U, int mu, long s) {
    m3 D = m3_zero();
    for (int nu = 0; nu < 4; nu++) {
        if (nu == mu) continue;
        m3 a = mul_nd(LINK(U, L, nu, UP(L, mu, s)), LINK(U, L, mu, UP(L, nu, s)));
        D = add(D, mul_nd(a, LINK(U, L, nu, s)));
        long sm = DN(L, nu, s);
        m3 b = mul(LINK(U, L, mu, sm), LINK(U, L, nu, UP(L, mu, sm)));
        D = add(D, mul_dn(b, LINK(U, L, nu, sm)));
    }
    return D;
}

void gfo_dsdu(const int *dims, const double *U, double coeff, double *out) {
    lat L = lat_make(dims);
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
    for (long s = 0; s < L.V; s++)
        for (int mu = 0; mu < 4; mu++) st(LINKP(out, &L, mu, s), scale(coeff, dsdu_site(&L, U, mu, s)));
    lat_free(&L);
}

/* md_force!(::GaugeAction) (molecular_dynamics.jl:251-267):
 * F_mu = clear; Traceless_antihermitian_add!(F_mu, -1/NC, U_mu * dSdU_mu) */
static void force_into(const lat *L, const double *U, double coeff, double *F) {
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
    for (long s = 0; s < L->V; s++)
        for (int mu = 0; mu < 4; mu++) {
            double *c = F + 8 * ((long)mu * L->V + s);
            for (int a = 0; a < 8; a++) c[a] = 0.0;
            m3 D = scale(coeff, dsdu_site(L, U, mu, s));
            ta_add(mul(LINK(U, L, mu, s), D), -1.0 / 3.0, c);
        }
}
void gfo_force(const int *dims, const double *U, double coeff, double *F) {
    lat L = lat_make(dims);
    force_into(&L, U, coeff, F);
    lat_free(&L);
}

/* update_momenta! kick (molecular_dynamics.jl:539-551): add_U!(P, eps, F)
 * (TA_gaugefields_4D_serial.jl:150-171) */
static void kick(long n8, double *P, const double *F, double eps) {
    for (long i = 0; i < n8; i++) P[i] += eps * F[i];
}
void gfo_kick(const int *dims, double *P, const double *F, double eps) {
    long V = (long)dims[0] * dims[1] * dims[2] * dims[3];
    kick(32 * V, P, F, eps);
}

/* update_gaugefields! (molecular_dynamics.jl:513-531): U_mu <- exp(eps P_mu) U_mu */
static void exp_update(const lat *L, double *U, const double *P, double eps) {
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
    for (long s = 0; s < L->V; s++)
        for (int mu = 0; mu < 4; mu++) {
            m3 E = exp_ta(P + 8 * ((long)mu * L->V + s), eps);
            st(LINKP(U, L, mu, s), mul(E, LINK(U, L, mu, s)));
        }
}
void gfo_exp_update(const int *dims, double *U, const double *P, double eps) {
    lat L = lat_make(dims);
    exp_update(&L, U, P, eps);
    lat_free(&L);
}

/* ------------------------------------------------------------------------- */
/* the same MD update in the reference's call structure (CPU baseline)         */
/* ------------------------------------------------------------------------- */
/* update_momenta! then update_gaugefields! as the reference sequences them
 * (molecular_dynamics.jl:513-551): md_force! (:251-267) per direction mu runs calc_dSdUμ!
 * (GaugeActions.jl:95-123), whose six staples evaluate_gaugelinks! forms as whole-lattice
 * products of (shifted) link fields into temporaries that are then added (add_U!), then
 * mul!(work, U_mu, dSdU), clear_U!(F_mu) and Traceless_antihermitian_add!(F_mu, -1/NC, work);
 * add_U!(P_mu, eps, F_mu); then per mu exptU! into a field, mul!(W, E, U_mu) and
 * substitute_U!(U_mu, W).  Every step is one OpenMP pass over the lattice through full 3x3
 * temporaries, 104 passes per update: the memory traffic of the reference's structure, not
 * of the fused restatement above.  Same formulas as gfo_force / gfo_kick / gfo_exp_update. */
typedef struct {
    int m1, d1, m2, d2; /* site x + d1 m1-hat + d2 m2-hat (d = 0: no shift) */
} shift2;
static inline long shifted(const lat *L, long s, shift2 h) {
    if (h.d1 > 0) s = UP(L, h.m1, s);
    if (h.d1 < 0) s = DN(L, h.m1, s);
    if (h.d2 > 0) s = UP(L, h.m2, s);
    if (h.d2 < 0) s = DN(L, h.m2, s);
    return s;
}
/* out(x) = opA(A(x + ha)) opB(B(x + hb)), op = identity or ^dag (mode bit 1: A^dag, 2: B^dag) */
static void pass_mul(const lat *L, double *out, const double *A, shift2 ha, const double *B, shift2 hb, int mode) {
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
    for (long s = 0; s < L->V; s++) {
        m3 a = ld(A + 18 * shifted(L, s, ha)), b = ld(B + 18 * shifted(L, s, hb));
        m3 r = mode == 0 ? mul(a, b) : mode == 1 ? mul_dn(a, b) : mode == 2 ? mul_nd(a, b) : mul(dag(a), dag(b));
        st(out + 18 * s, r);
    }
}
static void pass_add(const lat *L, double *D, double coeff, const double *T) {
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
    for (long i = 0; i < 18 * L->V; i++) D[i] += coeff * T[i];
}
static void pass_clear(long n, double *D) {
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
    for (long i = 0; i < n; i++) D[i] = 0.0;
}
static void pass_ta_add(const lat *L, double *F, double factor, const double *W) {
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
    for (long s = 0; s < L->V; s++) ta_add(ld(W + 18 * s), factor, F + 8 * s);
}
static void pass_kick(long n, double *P, const double *F, double eps) {
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
    for (long i = 0; i < n; i++) P[i] += eps * F[i];
}
static void pass_exp(const lat *L, double *E, const double *P, double eps) {
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
    for (long s = 0; s < L->V; s++) st(E + 18 * s, exp_ta(P + 8 * s, eps));
}
static void pass_copy(long n, double *dst, const double *src) {
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
    for (long i = 0; i < n; i++) dst[i] = src[i];
}

void gfo_md_update_unfused(const int *dims, double *U, double *P, double coeff, double eps_p, double eps_u) {
    lat L = lat_make(dims);
    const long V = L.V, n18 = 18 * V;
    double *T1 = (double *)malloc(sizeof(double) * n18), *T2 = (double *)malloc(sizeof(double) * n18);
    double *D = (double *)malloc(sizeof(double) * n18), *W = (double *)malloc(sizeof(double) * n18);
    double *F = (double *)malloc(sizeof(double) * 32 * V);
    const shift2 none = {0, 0, 0, 0};
    for (int mu = 0; mu < 4; mu++) {
        double *Umu = U + 18 * (long)mu * V;
        pass_clear(n18, D);
        for (int nu = 0; nu < 4; nu++) {
            if (nu == mu) continue;
            double *Unu = U + 18 * (long)nu * V;
            /* U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag */
            pass_mul(&L, T1, Unu, (shift2){mu, 1, 0, 0}, Umu, (shift2){nu, 1, 0, 0}, 2);
            pass_mul(&L, T2, T1, none, Unu, none, 2);
            pass_add(&L, D, coeff, T2);
            /* U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu) = (U_mu(x-nu) U_nu(x-nu+mu))^dag U_nu(x-nu) */
            pass_mul(&L, T1, Umu, (shift2){nu, -1, 0, 0}, Unu, (shift2){nu, -1, mu, 1}, 0);
            pass_mul(&L, T2, T1, none, Unu, (shift2){nu, -1, 0, 0}, 1);
            pass_add(&L, D, coeff, T2);
        }
        pass_mul(&L, W, Umu, none, D, none, 0);
        pass_clear(8 * V, F + 8 * (long)mu * V);
        pass_ta_add(&L, F + 8 * (long)mu * V, -1.0 / 3.0, W);
    }
    for (int mu = 0; mu < 4; mu++) pass_kick(8 * V, P + 8 * (long)mu * V, F + 8 * (long)mu * V, eps_p);
    for (int mu = 0; mu < 4; mu++) {
        double *Umu = U + 18 * (long)mu * V;
        pass_exp(&L, T1, P + 8 * (long)mu * V, eps_u);
        pass_mul(&L, W, T1, none, Umu, none, 0);
        pass_copy(n18, Umu, W);
    }
    free(T1);
    free(T2);
    free(D);
    free(W);
    free(F);
    lat_free(&L);
}

/* p*p (TA_gaugefields_4D_serial.jl:106-129): sum of squares of all coefficients;
 * md_hamiltonian uses real(p*p)/2 (molecular_dynamics.jl:494-505). */
double gfo_momenta_dot(const int *dims, const double *P) {
    long V = (long)dims[0] * dims[1] * dims[2] * dims[3];
    double s = 0.0;
    for (long i = 0; i < 32 * V; i++) s += P[i] * P[i];
    return s;
}

/* ------------------------------------------------------------------------- */
/* Wilson flow                                                                */
/* ------------------------------------------------------------------------- */
/* add_force!(F, U; plaqonly=true) (AbstractGaugefields.jl:2717-2762) into cleared F:
 * F_mu = TA(U_mu V_mu^dag), V = loops_staple_prime sum. */
static void flow_force(const lat *L, const double *U, double *F) {
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
    for (long s = 0; s < L->V; s++)
        for (int mu = 0; mu < 4; mu++) {
            double *c = F + 8 * ((long)mu * L->V + s);
            for (int a = 0; a < 8; a++) c[a] = 0.0;
            m3 V = staple_full(L, U, mu, s);
            ta_add(mul_nd(LINK(U, L, mu, s), V), 1.0, c);
        }
}
void gfo_flow_force(const int *dims, const double *U, double *F) {
    lat L = lat_make(dims);
    flow_force(&L, U, F);
    lat_free(&L);
}

/* exp_aF_U! (AbstractGaugefields.jl:2810-2841): W_mu = exp(a F_mu) U_mu */
static void exp_aF_U(const lat *L, double *W, double a, const double *F, const double *U) {
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
    for (long s = 0; s < L->V; s++)
        for (int mu = 0; mu < 4; mu++) {
            m3 E = exp_ta(F + 8 * ((long)mu * L->V + s), a);
            st(LINKP(W, L, mu, s), mul(E, LINK(U, L, mu, s)));
        }
}

/* Plaquette energy density (the reference has no built-in E(t); Luscher, JHEP 08 (2010)
 * 071): E = (2/V) sum_{x, mu<nu} Re tr(1 - P_munu(x)) = 36 (1 - plaquette/(18 V)). */
double gfo_energy_density(const int *dims, const double *U) {
    double V = (double)dims[0] * dims[1] * dims[2] * dims[3];
    return 2.0 * (18.0 - gfo_plaquette_sum(dims, U) / V);
}

/* flow!(U, ::Gradientflow) (gradientflow.jl:171-238): Luscher RK3,
 * W1 = exp(-eps/4 F0) U;  W2 = exp(-8eps/9 F1 + 17eps/36 F0) W1;
 * U  = exp(-3eps/4 F2 + 8eps/9 F1 - 17eps/36 F0) W2.
 * E_out[k] = energy density after step (k+1)*every (when every > 0). */
void gfo_flow(const int *dims, double *U, double eps, int nsteps, int every, double *E_out) {
    lat L = lat_make(dims);
    long n18 = 18 * 4 * L.V, n8 = 8 * 4 * L.V;
    double *W1 = (double *)malloc(sizeof(double) * n18);
    double *W2 = (double *)malloc(sizeof(double) * n18);
    double *F0 = (double *)malloc(sizeof(double) * n8);
    double *F1 = (double *)malloc(sizeof(double) * n8);
    double *F2 = (double *)malloc(sizeof(double) * n8);
    double *Ft = (double *)malloc(sizeof(double) * n8);
    int k = 0;
    for (int step = 1; step <= nsteps; step++) {
        flow_force(&L, U, F0);
        exp_aF_U(&L, W1, -eps / 4.0, F0, U);
        flow_force(&L, W1, F1);
        for (long i = 0; i < n8; i++) Ft[i] = 0.0;
        kick(n8, Ft, F1, -(8.0 * eps / 9.0));
        kick(n8, Ft, F0, 17.0 * eps / 36.0);
        exp_aF_U(&L, W2, 1.0, Ft, W1);
        flow_force(&L, W2, F2);
        for (long i = 0; i < n8; i++) Ft[i] = 0.0;
        kick(n8, Ft, F2, -(3.0 * eps / 4.0));
        kick(n8, Ft, F1, 8.0 * eps / 9.0);
        kick(n8, Ft, F0, -(17.0 * eps / 36.0));
        exp_aF_U(&L, U, 1.0, Ft, W2);
        if (every > 0 && step % every == 0 && E_out) E_out[k++] = gfo_energy_density(dims, U);
    }
    free(W1);
    free(W2);
    free(F0);
    free(F1);
    free(F2);
    free(Ft);
    lat_free(&L);
}

/* ------------------------------------------------------------------------- */
/* stout smearing                                                             */
/* ------------------------------------------------------------------------- */
/* forward! (stout_fast.jl:250-274) with calc_C! (stout_fast.jl:603-624) for the plaquette
 * loop set: C_mu = rho V_mu (staple_full), Q_mu = TA(C_mu U_mu^dag),
 * Uout_mu = exp(Q_mu) U_mu.  C (optional) receives C_mu. */
static void stout_forward(const lat *L, const double *Uin, double rho, double *Uout, double *C) {
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
    for (long s = 0; s < L->V; s++)
        for (int mu = 0; mu < 4; mu++) {
            m3 U = LINK(Uin, L, mu, s);
            m3 Cm = scale(rho, staple_full(L, Uin, mu, s));
            m3 Q = ta_mat(mul_nd(Cm, U));
            m3 E = expi_herm(scale(-I, Q));
            st(LINKP(Uout, L, mu, s), mul(E, U));
            if (C) st(LINKP(C, L, mu, s), Cm);
        }
}
void gfo_stout_forward(const int *dims, const double *Uin, double rho, double *Uout, double *C) {
    lat L = lat_make(dims);
    stout_forward(&L, Uin, rho, Uout, C);
    lat_free(&L);
}

/* backward_dSdUαUβρ_add! (stout_fast.jl:317-407) for one plaquette stout level, with the
 * convention of the reference's matrix derivatives (dS = 2 Re tr(D dU)):
 *   Din_mu  = Dout_mu e^{Q_mu}                                  calc_dSdu1!      :629-631
 *   L       = CdexpQdQ(U_mu Dout_mu, Q_mu)                       calc_dSdQ!       :636-675
 *   Lam     = TA(L)                                              calc_dSdΩ!       :682-684
 *   G_mu    = U_mu^dag Lam                                       calc_dSdC!       :686-688
 *   Din_mu += (Lam C_mu)^dag                                     calc_dSdUdag!    :690-694
 *   Din_nu += rho sum_j B_j G_mu(x - pos_j) A_j  (and the C^dag terms)
 *                                                   calc_dSdUν_fromdSCμ_add! :712-785
 * The last line's terms are those of Wilsonloop.derive_U applied to the six plaquette
 * staples of C_mu (stout_dataset.jl:10-105); written out for output direction a at site y
 * and staple plane b != a they are
 *   (i)   U_b(y-b+a)^dag G_a(y-b) U_b(y-b)         (ii)  U_b(y+a) G_a(y+b) U_b(y)^dag
 *   (iii) U_b(y+a) U_a(y+b)^dag G_b(y)             (iv)  U_b(y-b+a)^dag U_a(y-b)^dag G_b(y-b)^dag
 *   (v)   G_b(y+a)^dag U_a(y+b)^dag U_b(y)^dag     (vi)  G_b(y+a-b) U_a(y-b)^dag U_b(y-b)      */
static void stout_backward(const lat *L, const double *Uin, double rho, const double *Dout, double *Din) {
    long n18 = 18 * 4 * L->V;
    double *G = (double *)malloc(sizeof(double) * n18);
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
    for (long s = 0; s < L->V; s++)
        for (int mu = 0; mu < 4; mu++) {
            m3 U = LINK(Uin, L, mu, s);
            m3 Cm = scale(rho, staple_full(L, Uin, mu, s));
            m3 Q = ta_mat(mul_nd(Cm, U));
            m3 E = expi_herm(scale(-I, Q));
            m3 Do = LINK(Dout, L, mu, s);
            m3 D = mul(Do, E);
            m3 Lm = ta_mat(cdexpqdq(mul(U, Do), Q));
            st(LINKP(G, L, mu, s), mul_dn(U, Lm));
            D = add(D, dag(mul(Lm, Cm)));
            st(LINKP(Din, L, mu, s), D);
        }
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
    for (long y = 0; y < L->V; y++)
        for (int a = 0; a < 4; a++) {
            m3 acc = m3_zero();
            for (int b = 0; b < 4; b++) {
                if (b == a) continue;
                long ymb = DN(L, b, y), ypa = UP(L, a, y), ypb = UP(L, b, y);
                long ymbpa = UP(L, a, ymb);
                /* (i) */
                acc = add(acc, mul(mul_dn(LINK(Uin, L, b, ymbpa), LINK(G, L, a, ymb)), LINK(Uin, L, b, ymb)));
                /* (ii) */
                acc = add(acc, mul_nd(mul(LINK(Uin, L, b, ypa), LINK(G, L, a, ypb)), LINK(Uin, L, b, y)));
                /* (iii) */
                acc = add(acc, mul(mul_nd(LINK(Uin, L, b, ypa), LINK(Uin, L, a, ypb)), LINK(G, L, b, y)));
                /* (iv) */
                acc = add(acc, mul_nd(mul_dn(LINK(Uin, L, b, ymbpa), dag(LINK(Uin, L, a, ymb))), LINK(G, L, b, ymb)));
                /* (v) */
                acc = add(acc, mul_nd(mul_dn(LINK(G, L, b, ypa), dag(LINK(Uin, L, a, ypb))), LINK(Uin, L, b, y)));
                /* (vi) */
                long ypamb = DN(L, b, ypa);
                acc = add(acc, mul(mul_nd(LINK(G, L, b, ypamb), LINK(Uin, L, a, ymb)), LINK(Uin, L, b, ymb)));
            }
            double *p = LINKP(Din, L, a, y);
            st(p, add(ld(p), scale(rho, acc)));
        }
    free(G);
}
void gfo_stout_backward(const int *dims, const double *Uin, double rho, const double *Dout, double *Din) {
    lat L = lat_make(dims);
    stout_backward(&L, Uin, rho, Dout, Din);
    lat_free(&L);
}

/* calc_smearedU / apply_neuralnet! (Abstractsmearing.jl:237-272, 468-485): nlev stout
 * layers applied in sequence. */
void gfo_stout_smear(const int *dims, const double *U, int nlev, double rho, double *Uout) {
    lat L = lat_make(dims);
    long n18 = 18 * 4 * L.V;
    double *a = (double *)malloc(sizeof(double) * n18);
    memcpy(a, U, sizeof(double) * n18);
    for (int l = 0; l < nlev; l++) {
        stout_forward(&L, a, rho, Uout, NULL);
        memcpy(a, Uout, sizeof(double) * n18);
    }
    if (nlev == 0) memcpy(Uout, U, sizeof(double) * n18);
    free(a);
    lat_free(&L);
}

double gfo_stout_potential(const int *dims, const double *U, int nlev, double rho, double coeff) {
    long V = (long)dims[0] * dims[1] * dims[2] * dims[3];
    double *S = (double *)malloc(sizeof(double) * 72 * V);
    gfo_stout_smear(dims, U, nlev, rho, S);
    double v = gfo_potential(dims, S, coeff);
    free(S);
    return v;
}

/* Stout-smeared force (test/HMCstout_test_nowing.jl:99-118): smear (storing every level,
 * Abstractsmearing.jl:280-314), calc_dSdUμ! on the smeared links, back_prop through the
 * levels (Abstractsmearing.jl:352-411), F_mu = -(1/NC) TA(U_mu dSdU_bare_mu). */
static void stout_force(const lat *L, const double *U, int nlev, double rho, double coeff, double *F) {
    long n18 = 18 * 4 * L->V;
    double *lev = (double *)malloc(sizeof(double) * n18 * (nlev + 1));
    double *D = (double *)malloc(sizeof(double) * n18);
    double *D2 = (double *)malloc(sizeof(double) * n18);
    memcpy(lev, U, sizeof(double) * n18);
    for (int l = 0; l < nlev; l++) stout_forward(L, lev + n18 * l, rho, lev + n18 * (l + 1), NULL);
    const double *Us = lev + n18 * nlev;
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
    for (long s = 0; s < L->V; s++)
        for (int mu = 0; mu < 4; mu++) st(LINKP(D, L, mu, s), scale(coeff, dsdu_site(L, Us, mu, s)));
    for (int l = nlev - 1; l >= 0; l--) {
        stout_backward(L, lev + n18 * l, rho, D, D2);
        double *t = D;
        D = D2;
        D2 = t;
    }
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
    for (long s = 0; s < L->V; s++)
        for (int mu = 0; mu < 4; mu++) {
            double *c = F + 8 * ((long)mu * L->V + s);
            for (int a = 0; a < 8; a++) c[a] = 0.0;
            ta_add(mul(LINK(U, L, mu, s), LINK(D, L, mu, s)), -1.0 / 3.0, c);
        }
    free(lev);
    free(D);
    free(D2);
}
void gfo_stout_force(const int *dims, const double *U, int nlev, double rho, double coeff, double *F) {
    lat L = lat_make(dims);
    stout_force(&L, U, nlev, rho, coeff, F);
    lat_free(&L);
}

/* ------------------------------------------------------------------------- */
/* molecular dynamics                                                         */
/* ------------------------------------------------------------------------- */
typedef struct {
    const lat *L;
    const int *dims;
    double coeff, rho;
    int nlev;
    double *F;
} md_ctx;

/* update_momenta! (molecular_dynamics.jl:539-551) */
static void md_P(md_ctx *m, double *U, double *P, double eps) {
    if (m->nlev > 0)
        stout_force(m->L, U, m->nlev, m->rho, m->coeff, m->F);
    else
        force_into(m->L, U, m->coeff, m->F);
    kick(32 * m->L->V, P, m->F, eps);
}
/* update_gaugefields! (molecular_dynamics.jl:513-531) */
static void md_Q(md_ctx *m, double *U, double *P, double eps) { exp_update(m->L, U, P, eps); }

static double md_H(md_ctx *m, const double *U, const double *P) {
    double V = m->nlev > 0 ? gfo_stout_potential(m->dims, U, m->nlev, m->rho, m->coeff)
                           : gfo_potential(m->dims, U, m->coeff);
    return V + 0.5 * gfo_momenta_dot(m->dims, P);
}

/* md_trajectory! (molecular_dynamics.jl:712-730) with md_step! QPQ / PQP
 * (molecular_dynamics.jl:604-616) and the Omelyan 2MN user integrator of
 * docs/src/howtoimplement.md:94-112.  H = {H0, H1, H1 - H0}. */
void gfo_md_trajectory(const int *dims, double *U, double *P, double coeff, int nlev, double rho,
                       int integrator, double lambda, int steps, double tau, double *H) {
    lat L = lat_make(dims);
    md_ctx m = {&L, dims, coeff, rho, nlev, (double *)malloc(sizeof(double) * 32 * L.V)};
    double h0 = md_H(&m, U, P);
    double eps = tau / steps;
    for (int k = 0; k < steps; k++) {
        if (integrator == 0) {
            md_Q(&m, U, P, eps / 2);
            md_P(&m, U, P, eps);
            md_Q(&m, U, P, eps / 2);
        } else if (integrator == 1) {
            md_P(&m, U, P, eps / 2);
            md_Q(&m, U, P, eps);
            md_P(&m, U, P, eps / 2);
        } else if (integrator == 2) {
            md_P(&m, U, P, lambda * eps);
            md_Q(&m, U, P, 0.5 * eps);
            md_P(&m, U, P, (1.0 - 2.0 * lambda) * eps);
            md_Q(&m, U, P, 0.5 * eps);
            md_P(&m, U, P, lambda * eps);
        } else {
            /* 2MN with the closing kick of a step and the opening kick of the next one applied
             * as one kick of 2 lambda eps: the links do not change between the two, so the
             * sequence is the same map; one force evaluation per step fewer */
            md_P(&m, U, P, (k == 0 ? 1.0 : 2.0) * lambda * eps);
            md_Q(&m, U, P, 0.5 * eps);
            md_P(&m, U, P, (1.0 - 2.0 * lambda) * eps);
            md_Q(&m, U, P, 0.5 * eps);
            if (k == steps - 1) md_P(&m, U, P, lambda * eps);
        }
    }
    double h1 = md_H(&m, U, P);
    H[0] = h0;
    H[1] = h1;
    H[2] = h1 - h0;
    free(m.F);
    lat_free(&L);
}

/* ------------------------------------------------------------------------- */
/* Field strengths: topological charge and clover energy density              */
/* ------------------------------------------------------------------------- */
/* Product of links along a Wilson line from site s (evaluate_gaugelinks!,
 * AbstractGaugefields.jl:1797-1860, on Wilsonline paths): a step (mu, +n) multiplies
 * U_mu(y) and moves y -> y + mu, n times; a step (mu, -n) moves y -> y - mu and multiplies
 * U_mu(y)^dag, n times. */
typedef struct {
    int mu, len;
} step;
static m3 path_product(const lat *L, const double *U, long s, const step *p, int n) {
    m3 r = m3_id();
    long y = s;
    for (int k = 0; k < n; k++) {
        int mu = p[k].mu, len = p[k].len;
        for (int q = 0; q < abs(len); q++) {
            if (len > 0) {
                r = mul(r, LINK(U, L, mu, y));
                y = UP(L, mu, y);
            } else {
                y = DN(L, mu, y);
                r = mul_nd(r, LINK(U, L, mu, y));
            }
        }
    }
    return r;
}

/* Loop sets per (mu, nu): plaquette (AbstractGaugefields.jl:1208), the four clover leaves
 * (make_cloverloops, autostaples/wilsonloops.jl:166-177), the eight 2x1 rectangles
 * (_rectangle_loops, AbstractGaugefields.jl:1314-1330). */
static int loops_of(int method, int mu, int nu, step out[8][4]) {
    const int m = mu, n = nu;
    if (method == GFO_TOPO_PLAQUETTE) {
        step a[4] = {{m, 1}, {n, 1}, {m, -1}, {n, -1}};
        memcpy(out[0], a, sizeof a);
        return 1;
    }
    if (method == GFO_TOPO_CLOVER) {
        step l[4][4] = {{{m, 1}, {n, 1}, {m, -1}, {n, -1}},
                        {{n, 1}, {m, -1}, {n, -1}, {m, 1}},
                        {{n, -1}, {m, 1}, {n, 1}, {m, -1}},
                        {{m, -1}, {n, -1}, {m, 1}, {n, 1}}};
        memcpy(out, l, sizeof l);
        return 4;
    }
    step r[8][4] = {{{m, 2}, {n, 1}, {m, -2}, {n, -1}}, {{n, 1}, {m, -2}, {n, -1}, {m, 2}},
                    {{n, -1}, {m, 2}, {n, 1}, {m, -2}}, {{m, -2}, {n, -1}, {m, 2}, {n, 1}},
                    {{m, 1}, {n, 2}, {m, -1}, {n, -2}}, {{n, 2}, {m, -1}, {n, -2}, {m, 1}},
                    {{n, -2}, {m, 1}, {n, 2}, {m, -1}}, {{m, -1}, {n, -2}, {m, 1}, {n, 2}}};
    memcpy(out, r, sizeof r);
    return 8;
}

/* F_munu(x) = TA(sum of the method's loops at x)
 * (_plaquette/_clover/_rectangle_field_strengths, AbstractGaugefields.jl:1195-1215,
 * 1255-1275, 1332-1351) */
static m3 field_strength(const lat *L, const double *U, int method, int mu, int nu, long s) {
    step loops[8][4];
    int nl = loops_of(method, mu, nu, loops);
    m3 acc = m3_zero();
    for (int k = 0; k < nl; k++) acc = add(acc, path_product(L, U, s, loops[k], 4));
    return ta_mat(acc);
}

/* Levi-Civita symbol (_epsilon_tensor_4d, AbstractGaugefields.jl:1172-1183) */
static int eps4(int a, int b, int c, int d) {
    int p[4] = {a, b, c, d};
    for (int i = 0; i < 4; i++)
        for (int j = i + 1; j < 4; j++)
            if (p[i] == p[j]) return 0;
    int inv = 0;
    for (int i = 0; i < 4; i++)
        for (int j = i + 1; j < 4; j++) inv += p[i] > p[j];
    return (inv % 2) ? -1 : 1;
}

/* q(x) = -norm * Re sum_{mu nu rho sigma} eps tr(F_munu F_rhosigma) / (32 pi^2 n_loops^2),
 * norm 1 (plaquette, clover) or 2 (rectangle) (_plaquette/_clover/_rectangle_topological_
 * charge_density, AbstractGaugefields.jl:1217-1245, 1277-1306, 1353-1385) */
static void topo_density_basic(const lat *L, const double *U, int method, double *q) {
    const int nl = method == GFO_TOPO_PLAQUETTE ? 1 : (method == GFO_TOPO_CLOVER ? 4 : 8);
    const double norm = method == GFO_TOPO_RECTANGLE ? 2.0 : 1.0;
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
    for (long s = 0; s < L->V; s++) {
        m3 F[4][4];
        for (int mu = 0; mu < 4; mu++)
            for (int nu = 0; nu < 4; nu++)
                if (mu != nu) F[mu][nu] = field_strength(L, U, method, mu, nu, s);
        cplx acc = 0.0;
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++)
                for (int c = 0; c < 4; c++)
                    for (int d = 0; d < 4; d++) {
                        int e = eps4(a, b, c, d);
                        if (e) acc += e * trace(mul(F[a][b], F[c][d]));
                    }
        q[s] = -norm * creal(acc) / (32.0 * GFO_PI * GFO_PI * nl * nl);
    }
}

/* topological_charge_density(U; method) (AbstractGaugefields.jl:1430-1446): plaquette,
 * clover, improved = 5/3 clover - 1/12 rectangle (_improved_topological_charge_density,
 * :1393-1397); also the rectangle density itself.  Returns the sum Q. */
double gfo_topological_charge_density(const int *dims, const double *U, int method, double *q) {
    lat L = lat_make(dims);
    double *buf = q ? q : (double *)malloc(sizeof(double) * L.V);
    if (method == GFO_TOPO_IMPROVED) {
        double *r = (double *)malloc(sizeof(double) * L.V);
        topo_density_basic(&L, U, GFO_TOPO_CLOVER, buf);
        topo_density_basic(&L, U, GFO_TOPO_RECTANGLE, r);
        for (long s = 0; s < L.V; s++) buf[s] = (5.0 / 3.0) * buf[s] + (-1.0 / 12.0) * r[s];
        free(r);
    } else {
        topo_density_basic(&L, U, method, buf);
    }
    double Q = 0.0;
    for (long s = 0; s < L.V; s++) Q += buf[s];
    if (!q) free(buf);
    lat_free(&L);
    return Q;
}

/* clover energy density (calculate_energy_density, samples/measurements/energydensity.jl:
 * 32-76): sum_{mu != nu} -Re tr(F_munu F_munu)/2 / (V * 4^2), F = TA(clover) */
static double clover_energy_site(const lat *L, long s, const void *ctx) {
    const double *U = (const double *)ctx;
    double w = 0.0;
    for (int mu = 0; mu < 4; mu++)
        for (int nu = 0; nu < 4; nu++)
            if (mu != nu) {
                m3 F = field_strength(L, U, GFO_TOPO_CLOVER, mu, nu, s);
                w += -creal(trace(mul(F, F))) / 2.0;
            }
    return w;
}
double gfo_energy_density_clover(const int *dims, const double *U) {
    lat L = lat_make(dims);
    double w = lattice_sum(&L, clover_energy_site, U);
    double E = w / ((double)L.V * 16.0);
    lat_free(&L);
    return E;
}

/* ------------------------------------------------------------------------- */
/* Counter-based random fields (Philox4x32-10)                                */
/* ------------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11 "Parallel random numbers: as easy as
 * 1, 2, 3"): ten rounds of two 32x32->64 multiplies (0xD2511F53, 0xCD9E8D57) with the Weyl
 * key schedule (0x9E3779B9, 0xBB67AE85). */
void gfo_philox4x32_10(const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1, n3 = (uint32_t)p0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* Stream of one link: key = seed, counter = (site, tag << 24 | mu << 16 | block, sweep).
 * Gaussian momenta (tag 1): block b = 0..3 gives the Box-Muller pair (p_2b, p_2b+1) from two
 * 53-bit uniforms; hot-start links (tag 2): blocks 0..4 give 18 uniforms in (0, 1). */
static void link_block(uint64_t seed, uint64_t sweep, long s, int mu, int tag, int block, uint32_t w[4]) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const uint32_t ctr[4] = {(uint32_t)s, ((uint32_t)tag << 24) | ((uint32_t)mu << 16) | (uint32_t)block,
                             (uint32_t)sweep, (uint32_t)(sweep >> 32)};
    gfo_philox4x32_10(ctr, key, w);
}
static inline double unit53(uint32_t hi, uint32_t lo) {
    return ((double)(hi >> 5) * 67108864.0 + (double)(lo >> 6) + 0.5) * (1.0 / 9007199254740992.0);
}
static inline double unit32(uint32_t w) { return ((double)w + 0.5) * (1.0 / 4294967296.0); }

/* gaussian_momenta! (API.jl:341-392; gauss_distribution!, TA_gaugefields_4D_MPILattice.jl:
 * 157-193): p_a ~ N(0, sigma^2) for every link, from the stream above */
void gfo_gaussian_momenta(const int *dims, double sigma, uint64_t seed, uint64_t sweep, double *P) {
    long V = (long)dims[0] * dims[1] * dims[2] * dims[3];
    const double two_pi = 2.0 * GFO_PI;
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
    for (long s = 0; s < V; s++)
        for (int mu = 0; mu < 4; mu++) {
            double *p = P + 8 * ((long)mu * V + s);
            for (int b = 0; b < 4; b++) {
                uint32_t w[4];
                link_block(seed, sweep, s, mu, 1, b, w);
                const double r = sqrt(-2.0 * log(unit53(w[0], w[1]))), ph = two_pi * unit53(w[2], w[3]);
                p[2 * b] = sigma * r * cos(ph);
                p[2 * b + 1] = sigma * r * sin(ph);
            }
        }
}

/* Hot start (randomGaugefields / hot condition, gaugefields_4D_wing.jl:295-322 with
 * normalize3!, AbstractGaugefields.jl:3378-3457): entries (u - 1/2) + i (u' - 1/2), element
 * (i, j) from uniforms 2(3i + j), 2(3i + j) + 1; rows 0 and 1 Gram-Schmidt normalized, row 2 =
 * conj(row0 x row1). */
void gfo_hot_links(const int *dims, uint64_t seed, double *U) {
    long V = (long)dims[0] * dims[1] * dims[2] * dims[3];
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
    for (long s = 0; s < V; s++)
        for (int mu = 0; mu < 4; mu++) {
            double u[20];
            for (int b = 0; b < 5; b++) {
                uint32_t w[4];
                link_block(seed, 0, s, mu, 2, b, w);
                for (int k = 0; k < 4; k++) u[4 * b + k] = unit32(w[k]);
            }
            cplx r0[3], r1[3];
            for (int j = 0; j < 3; j++) {
                r0[j] = CMPLX(u[2 * j] - 0.5, u[2 * j + 1] - 0.5);
                r1[j] = CMPLX(u[2 * (3 + j)] - 0.5, u[2 * (3 + j) + 1] - 0.5);
            }
            double n0 = 0.0;
            for (int j = 0; j < 3; j++) n0 += creal(r0[j] * conj(r0[j]));
            n0 = 1.0 / sqrt(n0);
            for (int j = 0; j < 3; j++) r0[j] *= n0;
            cplx d = 0.0;
            for (int j = 0; j < 3; j++) d += r1[j] * conj(r0[j]);
            for (int j = 0; j < 3; j++) r1[j] -= d * r0[j];
            double n1 = 0.0;
            for (int j = 0; j < 3; j++) n1 += creal(r1[j] * conj(r1[j]));
            n1 = 1.0 / sqrt(n1);
            for (int j = 0; j < 3; j++) r1[j] *= n1;
            m3 m;
            for (int j = 0; j < 3; j++) {
                m.e[0][j] = r0[j];
                m.e[1][j] = r1[j];
            }
            for (int k = 0; k < 3; k++) {
                const int a = (k + 1) % 3, b = (k + 2) % 3;
                m.e[2][k] = conj(r0[a] * r1[b] - r0[b] * r1[a]);
            }
            st(U + 18 * ((long)mu * V + s), m);
        }
}

/* ------------------------------------------------------------------------- */
/* Heatbath and overrelaxation (Cabibbo-Marinari over SU(2) subgroups)       */
/* ------------------------------------------------------------------------- */
/* project_onto_SU2! (heatbath/portable/kernels.jl:12-23): S -> [[a, -conj b], [b, conj a]],
 * a = (S11 + conj S22)/2, b = (S21 - conj S12)/2 (a multiple of an SU(2) matrix) */
typedef struct {
    cplx e[2][2];
} m2;
static inline m2 su2_project(m2 S) {
    const cplx a = 0.5 * S.e[0][0] + 0.5 * conj(S.e[1][1]);
    const cplx b = 0.5 * S.e[1][0] - 0.5 * conj(S.e[0][1]);
    m2 r;
    r.e[0][0] = a;
    r.e[1][0] = b;
    r.e[0][1] = -conj(b);
    r.e[1][1] = conj(a);
    return r;
}
static inline m2 mul2(m2 a, m2 b) {
    m2 r;
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) r.e[i][j] = a.e[i][0] * b.e[0][j] + a.e[i][1] * b.e[1][j];
    return r;
}
static const int SUBGROUP[3][2] = {{0, 1}, {1, 2}, {0, 2}}; /* (1,2), (2,3), (1,3) */

/* u <- embed(K at rows/cols n, m) u */
static m3 su2_left_mul(m2 K, int n, int m, m3 u) {
    m3 r = u;
    for (int j = 0; j < 3; j++) {
        r.e[n][j] = K.e[0][0] * u.e[n][j] + K.e[0][1] * u.e[m][j];
        r.e[m][j] = K.e[1][0] * u.e[n][j] + K.e[1][1] * u.e[m][j];
    }
    return r;
}

static m3 reunit_rows(m3 u) { /* Gram-Schmidt rows 0, 1; row 2 = conj(row0 x row1) */
    double n0 = 0.0;
    for (int j = 0; j < 3; j++) n0 += creal(u.e[0][j] * conj(u.e[0][j]));
    n0 = 1.0 / sqrt(n0);
    for (int j = 0; j < 3; j++) u.e[0][j] *= n0;
    cplx d = 0.0;
    for (int j = 0; j < 3; j++) d += u.e[1][j] * conj(u.e[0][j]);
    for (int j = 0; j < 3; j++) u.e[1][j] -= d * u.e[0][j];
    double n1 = 0.0;
    for (int j = 0; j < 3; j++) n1 += creal(u.e[1][j] * conj(u.e[1][j]));
    n1 = 1.0 / sqrt(n1);
    for (int j = 0; j < 3; j++) u.e[1][j] *= n1;
    for (int k = 0; k < 3; k++) {
        const int a = (k + 1) % 3, b = (k + 2) % 3;
        u.e[2][k] = conj(u.e[0][a] * u.e[1][b] - u.e[0][b] * u.e[1][a]);
    }
    return u;
}

/* Kennedy-Pendleton SU(2) heatbath (_su2_update_kp_core!, heatbath/portable/kernels.jl:63-150)
 * for the projected staple S (2x2, a multiple of SU(2)) and coupling beta / NC:
 * k = 2 (beta/3) rho; tries draw (R, R', R'', R''') from block `try`; after acceptance
 * (phi, cos theta) come from block 0xFFFF.  Returns 1 and K on success. */
static int kp_update(uint64_t seed, uint64_t sweep, long s, int mu, int tag, double beta, int itmax, m2 S,
                     m2 *K) {
    const double rho0 = creal(S.e[0][0] + S.e[1][1]) / 2, rho1 = -cimag(S.e[0][1] + S.e[1][0]) / 2;
    const double rho2 = creal(S.e[1][0] - S.e[0][1]) / 2, rho3 = cimag(S.e[1][1] - S.e[0][0]) / 2;
    const double rho = sqrt(rho0 * rho0 + rho1 * rho1 + rho2 * rho2 + rho3 * rho3);
    const cplx det = S.e[0][0] * S.e[1][1] - S.e[0][1] * S.e[1][0];
    m2 V0;
    V0.e[0][0] = rho * S.e[1][1] / det;
    V0.e[0][1] = -rho * S.e[0][1] / det;
    V0.e[1][0] = -rho * S.e[1][0] / det;
    V0.e[1][1] = rho * S.e[0][0] / det;
    const double k = 2.0 * (beta / 3.0) * rho;
    double delta = 0.0;
    int ok = 0;
    for (int t = 0; t < itmax && !ok; t++) {
        uint32_t w[4];
        link_block(seed, sweep, s, mu, tag, t, w);
        const double X = -log(unit32(w[0])) / k, Xp = -log(unit32(w[1])) / k;
        const double c = cos(2.0 * GFO_PI * unit32(w[2]));
        delta = Xp + X * c * c;
        const double r3 = unit32(w[3]);
        ok = r3 * r3 <= 1.0 - 0.5 * delta;
    }
    if (!ok) return 0;
    uint32_t w[4];
    link_block(seed, sweep, s, mu, tag, 0xFFFF, w);
    const double a1 = 1.0 - delta, rr = sqrt(1.0 - a1 * a1);
    const double phi = 2.0 * GFO_PI * unit32(w[0]);
    const double ct = 2.0 * (unit32(w[1]) - 0.5), st = sqrt(1.0 - ct * ct);
    const double a2 = rr * cos(phi) * st, a3 = rr * sin(phi) * st, a4 = rr * ct;
    m2 tmp;
    tmp.e[0][0] = CMPLX(a1, a4);
    tmp.e[0][1] = CMPLX(a3, a2);
    tmp.e[1][0] = CMPLX(-a3, a2);
    tmp.e[1][1] = CMPLX(a1, -a4);
    m2 Un = mul2(tmp, V0);
    const cplx al = 0.5 * Un.e[0][0] + 0.5 * conj(Un.e[1][1]);
    const cplx be = 0.5 * Un.e[1][0] - 0.5 * conj(Un.e[0][1]);
    const double d = creal(al * conj(al)) + creal(be * conj(be));
    K->e[0][0] = al / d;
    K->e[1][0] = be / d;
    K->e[0][1] = -conj(be) / d;
    K->e[1][1] = conj(al) / d;
    return 1;
}

/* SU3update_matrix! (heatbath/portable/kernels.jl:199-300): subgroups (1,2), (2,3), (1,3),
 * stream tags 3, 4, 5; returns 0 and leaves u unchanged if a subgroup is not accepted */
static int su3_heatbath_link(uint64_t seed, uint64_t sweep, long s, int mu, double beta, int itmax, m3 *u,
                             m3 V) {
    m3 w = *u;
    for (int g = 0; g < 3; g++) {
        const int n = SUBGROUP[g][0], m = SUBGROUP[g][1];
        const m3 UV = mul(w, V);
        m2 S;
        S.e[0][0] = UV.e[n][n];
        S.e[0][1] = UV.e[n][m];
        S.e[1][0] = UV.e[m][n];
        S.e[1][1] = UV.e[m][m];
        m2 K;
        if (!kp_update(seed, sweep, s, mu, 3 + g, beta, itmax, su2_project(S), &K)) return 0;
        w = su2_left_mul(K, n, m, w);
    }
    *u = reunit_rows(w);
    return 1;
}

/* overrelaxation subgroup step (_overrelaxation_subgroup!, heatbathmodule.jl:1248-1280):
 * w = P_SU2(sub(u V)), h = (w w)^dag with normalized columns, u <- embed(h) u */
static m3 su3_or_subgroup(m3 u, m3 V, int n, int m) {
    const m3 UV = mul(u, V);
    m2 S;
    S.e[0][0] = UV.e[n][n];
    S.e[0][1] = UV.e[n][m];
    S.e[1][0] = UV.e[m][n];
    S.e[1][1] = UV.e[m][m];
    const m2 w = su2_project(S);
    m2 h;
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) h.e[i][j] = conj(w.e[0][i]) * conj(w.e[j][0]) + conj(w.e[1][i]) * conj(w.e[j][1]);
    for (int c = 0; c < 2; c++) { /* Gram-Schmidt on the columns */
        if (c == 1) {
            const cplx p = conj(h.e[0][0]) * h.e[0][1] + conj(h.e[1][0]) * h.e[1][1];
            h.e[0][1] -= p * h.e[0][0];
            h.e[1][1] -= p * h.e[1][0];
        }
        const double nn = 1.0 / sqrt(creal(h.e[0][c] * conj(h.e[0][c])) + creal(h.e[1][c] * conj(h.e[1][c])));
        h.e[0][c] *= nn;
        h.e[1][c] *= nn;
    }
    return su2_left_mul(h, n, m, u);
}

/* SUN_overrelaxation_rng! (heatbathmodule.jl:1295-1321): NC = 3 steps; step i draws the
 * subgroup from block i of stream tag 6: n = 1 + floor(2 w0 / 2^32), m = n + 1 +
 * floor((3 - n) w1 / 2^32) (the reference's nonuniform choice: (1,2) 1/4, (1,3) 1/4,
 * (2,3) 1/2); the link is re-projected afterwards. */
static m3 su3_overrelax_link(uint64_t seed, uint64_t sweep, long s, int mu, m3 u, m3 V) {
    for (int i = 0; i < 3; i++) {
        uint32_t w[4];
        link_block(seed, sweep, s, mu, 6, i, w);
        const int n = (int)(((uint64_t)w[0] * 2u) >> 32);         /* 0 or 1 */
        const int m = n + 1 + (int)(((uint64_t)w[1] * (uint32_t)(2 - n)) >> 32);
        u = su3_or_subgroup(u, V, n, m);
    }
    return reunit_rows(u);
}

static int site_parity(const lat *L, long s) {
    long r = s;
    int p = 0;
    for (int m = 0; m < 4; m++) {
        p += (int)(r % L->n[m]);
        r /= L->n[m];
    }
    return p & 1;
}

/* heatbath!(U, temps, beta; seed, sweep) (heatbathmodule.jl:2119-2163): for mu = 1..4, even
 * then odd sites: V = staple sum (loops_staple, in-tree make_plaq_staple autostaples/
 * wilsonloops.jl:395-412), SU(3) heatbath of every link of that colour.  mode 0 heatbath,
 * 1 overrelaxation.  Returns the number of links left unchanged by a failed heatbath. */
long gfo_heatbath_sweep(const int *dims, double *U, double beta, uint64_t seed, uint64_t sweep, int itmax,
                        int mode) {
    if (itmax > 0xFFFE) itmax = 0xFFFE; /* 16-bit try index; block 0xFFFF follows an acceptance */
    lat L = lat_make(dims);
    long failures = 0;
    m3 *newu = (m3 *)malloc(sizeof(m3) * L.V);
    char *did = (char *)malloc(L.V);
    for (int mu = 0; mu < 4; mu++)
        for (int par = 0; par < 2; par++) {
            long fail_here = 0;
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static) reduction(+ : fail_here)
#endif
            for (long s = 0; s < L.V; s++) {
                did[s] = 0;
                if (site_parity(&L, s) != par) continue;
                const m3 V = dag(staple_full(&L, U, mu, s));
                m3 u = LINK(U, &L, mu, s);
                if (mode == 0) {
                    if (su3_heatbath_link(seed, sweep, s, mu, beta, itmax, &u, V)) {
                        newu[s] = u;
                        did[s] = 1;
                    } else {
                        fail_here++;
                    }
                } else {
                    newu[s] = su3_overrelax_link(seed, sweep, s, mu, u, V);
                    did[s] = 1;
                }
            }
            failures += fail_here;
            for (long s = 0; s < L.V; s++)
                if (did[s]) st(LINKP(U, &L, mu, s), newu[s]);
        }
    free(newu);
    free(did);
    lat_free(&L);
    return failures;
}

/* Kennedy-Pendleton sampler alone: a0 of n accepted SU(2) elements at coupling k (for the
 * distribution test p(a0) ~ sqrt(1 - a0^2) exp(k a0)); stream tag 7, site = sample index */
void gfo_kp_samples(double k, long n, uint64_t seed, double *a0) {
    for (long i = 0; i < n; i++) {
        m2 S; /* rho = k * 3 / (2 beta) with beta = 3/2: S = rho * identity */
        S.e[0][0] = k;
        S.e[1][1] = k;
        S.e[0][1] = S.e[1][0] = 0.0;
        m2 K;
        if (!kp_update(seed, 0, i, 0, 7, 1.5, 1000000, S, &K)) {
            a0[i] = NAN;
            continue;
        }
        a0[i] = creal(K.e[0][0]);
    }
}
