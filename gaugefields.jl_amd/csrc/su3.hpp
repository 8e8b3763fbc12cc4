// su3.hpp -- device-side SU(3) algebra for the gfx950 kernels.
//
// 3x3 complex matrices live in registers (row-major, 9 complex = 18 reals).  The storage
// type T is double (FP64 form) or float (FP32-bulk form); the Cayley-Hamilton coefficients
// f_j, b_ij and all global sums are always evaluated in double.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

namespace gf {

template <typename T>
struct cx {
    T r, i;
};

template <typename T>
struct M3 {
    cx<T> m[9];  // m[3*i + j] = element (i, j)
};

// ---------------------------------------------------------------------------
// Scalars (step sizes, rho, coupling factors, 1/3, 1/sqrt(3)) are applied in double and the
// product rounded once to T.  A float-rounded scalar would scale every element of a field by
// the same relative error (float(-1/3) is 3e-8 too large): a systematic force-scale bias of
// that size makes an FP32-bulk trajectory drift from energy conservation by ~1e-8 of the
// action change, visible in dH at 32^4 and above.  For T = double these are a*x and fma.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T scale1(double a, T x) {
    return (T)(a * (double)x);
}
template <typename T>
__device__ __forceinline__ T fma1(double a, T x, T y) {
    return (T)fma(a, (double)x, (double)y);
}

// true when pred holds on every active lane of the wave (a wave-uniform branch condition)
__device__ __forceinline__ bool wave_all(bool pred) { return __builtin_amdgcn_ballot_w64(!pred) == 0ull; }

// ---------------------------------------------------------------------------
// complex helpers
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void cmac(cx<T>& c, const cx<T>& a, const cx<T>& b) {  // c += a b
    c.r = fma(a.r, b.r, c.r);
    c.r = fma(-a.i, b.i, c.r);
    c.i = fma(a.r, b.i, c.i);
    c.i = fma(a.i, b.r, c.i);
}
template <typename T>
__device__ __forceinline__ void cmac_nc(cx<T>& c, const cx<T>& a, const cx<T>& b) {  // c += a conj(b)
    c.r = fma(a.r, b.r, c.r);
    c.r = fma(a.i, b.i, c.r);
    c.i = fma(a.i, b.r, c.i);
    c.i = fma(-a.r, b.i, c.i);
}
template <typename T>
__device__ __forceinline__ void cmac_cn(cx<T>& c, const cx<T>& a, const cx<T>& b) {  // c += conj(a) b
    c.r = fma(a.r, b.r, c.r);
    c.r = fma(a.i, b.i, c.r);
    c.i = fma(a.r, b.i, c.i);
    c.i = fma(-a.i, b.r, c.i);
}

// ---------------------------------------------------------------------------
// matrix products (all results by value; everything stays in registers)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ M3<T> zero3() {
    M3<T> z;
#pragma unroll
    for (int k = 0; k < 9; k++) z.m[k].r = z.m[k].i = T(0);
    return z;
}
template <typename T>
__device__ __forceinline__ M3<T> ident3() {
    M3<T> z = zero3<T>();
    z.m[0].r = z.m[4].r = z.m[8].r = T(1);
    return z;
}
template <typename T>
__device__ __forceinline__ M3<T> mul(const M3<T>& a, const M3<T>& b) {  // a b
    M3<T> c = zero3<T>();
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int k = 0; k < 3; k++) cmac(c.m[3 * i + j], a.m[3 * i + k], b.m[3 * k + j]);
    return c;
}
template <typename T>
__device__ __forceinline__ M3<T> mul_nd(const M3<T>& a, const M3<T>& b) {  // a b^dag
    M3<T> c = zero3<T>();
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int k = 0; k < 3; k++) cmac_nc(c.m[3 * i + j], a.m[3 * i + k], b.m[3 * j + k]);
    return c;
}
template <typename T>
__device__ __forceinline__ M3<T> mul_dn(const M3<T>& a, const M3<T>& b) {  // a^dag b
    M3<T> c = zero3<T>();
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int k = 0; k < 3; k++) cmac_cn(c.m[3 * i + j], a.m[3 * k + i], b.m[3 * k + j]);
    return c;
}
template <typename T>
__device__ __forceinline__ void acc_mul(M3<T>& c, const M3<T>& a, const M3<T>& b) {  // c += a b
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int k = 0; k < 3; k++) cmac(c.m[3 * i + j], a.m[3 * i + k], b.m[3 * k + j]);
}
template <typename T>
__device__ __forceinline__ void acc_mul_nd(M3<T>& c, const M3<T>& a, const M3<T>& b) {  // c += a b^dag
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int k = 0; k < 3; k++) cmac_nc(c.m[3 * i + j], a.m[3 * i + k], b.m[3 * j + k]);
}
template <typename T>
__device__ __forceinline__ void acc_mul_dn(M3<T>& c, const M3<T>& a, const M3<T>& b) {  // c += a^dag b
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int k = 0; k < 3; k++) cmac_cn(c.m[3 * i + j], a.m[3 * k + i], b.m[3 * k + j]);
}
template <typename T>
__device__ __forceinline__ M3<T> dag(const M3<T>& a) {
    M3<T> c;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            c.m[3 * i + j].r = a.m[3 * j + i].r;
            c.m[3 * i + j].i = -a.m[3 * j + i].i;
        }
    return c;
}
template <typename T>
__device__ __forceinline__ void acc(M3<T>& c, const M3<T>& a) {
#pragma unroll
    for (int k = 0; k < 9; k++) {
        c.m[k].r += a.m[k].r;
        c.m[k].i += a.m[k].i;
    }
}
template <typename T>
__device__ __forceinline__ void acc_dag(M3<T>& c, const M3<T>& a) {  // c += a^dag
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            c.m[3 * i + j].r += a.m[3 * j + i].r;
            c.m[3 * i + j].i -= a.m[3 * j + i].i;
        }
}
template <typename T>
__device__ __forceinline__ M3<T> scal(T s, const M3<T>& a) {
    M3<T> c;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        c.m[k].r = s * a.m[k].r;
        c.m[k].i = s * a.m[k].i;
    }
    return c;
}
// Re tr(a b^dag) = sum_ij Re(a_ij conj(b_ij))
template <typename T>
__device__ __forceinline__ T retr_nd(const M3<T>& a, const M3<T>& b) {
    T s = T(0);
#pragma unroll
    for (int k = 0; k < 9; k++) {
        s = fma(a.m[k].r, b.m[k].r, s);
        s = fma(a.m[k].i, b.m[k].i, s);
    }
    return s;
}

template <typename T>
__device__ __forceinline__ M3<double> to_f64(const M3<T>& a) {
    M3<double> c;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        c.m[k].r = (double)a.m[k].r;
        c.m[k].i = (double)a.m[k].i;
    }
    return c;
}

// Projection back to SU(3) as normalize3! + m3complv3! do (AbstractGaugefields.jl:3378-3457):
// Gram-Schmidt on rows 1 and 2, row 3 = conj(row1 x row2).  The FP32-bulk kernels apply it to
// every link they update so that float rounding cannot accumulate into a unitarity drift.  The
// projection itself runs in double and only its result is rounded: a float Gram-Schmidt leaves
// the row norms about 1e-7 too large on average (1/sqrt rounds asymmetrically next to 1), a bias
// that shows up directly in the plaquette.
__device__ __forceinline__ void reunitarize_f64(M3<double>& u) {
    cx<double> r1[3], r2[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        r1[c] = u.m[c];
        r2[c] = u.m[3 + c];
    }
    double w2 = 0.0, w1r = 0.0, w1i = 0.0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        w2 += r1[c].r * r1[c].r + r1[c].i * r1[c].i;
        w1r += r2[c].r * r1[c].r + r2[c].i * r1[c].i;  // row2 . conj(row1)
        w1i += r2[c].i * r1[c].r - r2[c].r * r1[c].i;
    }
    // Rows that are unit to within |d| <= 2^-14 (every updated link: an exactly unitary factor
    // times a stored link) take 1/(1+d) and 1/sqrt(1+d) from their series to d^4 (truncation
    // below 2e-21); a wave with any other row redoes them with the division and square roots.
    // 1/(1+d) = 1 - d + d^2 - d^3 + d^4,  1/sqrt(1+d) = 1 - d/2 + 3d^2/8 - 5d^3/16 + 35d^4/128
    auto rcp1 = [](double d) { return fma(fma(fma(d - 1.0, d, 1.0), d, -1.0), d, 1.0); };
    auto rsq1 = [](double d) { return fma(fma(fma(fma(d, 35.0 / 128.0, -5.0 / 16.0), d, 3.0 / 8.0), d, -0.5), d, 1.0); };
    const bool exact = !wave_all(fabs(w2 - 1.0) <= 0x1p-14);
    double inv2 = rcp1(w2 - 1.0);
    if (exact) inv2 = 1.0 / w2;
    w1r = -w1r * inv2;
    w1i = -w1i * inv2;
    double w3 = 0.0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double xr = r2[c].r + w1r * r1[c].r - w1i * r1[c].i;
        const double xi = r2[c].i + w1r * r1[c].i + w1i * r1[c].r;
        r2[c].r = xr;
        r2[c].i = xi;
        w3 += xr * xr + xi * xi;
    }
    double n1 = rsq1(w2 - 1.0), n2 = rsq1(w3 - 1.0);
    if (exact || !wave_all(fabs(w3 - 1.0) <= 0x1p-14)) {
        n1 = 1.0 / sqrt(w2);
        n2 = 1.0 / sqrt(w3);
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        r1[c].r *= n1;
        r1[c].i *= n1;
        r2[c].r *= n2;
        r2[c].i *= n2;
        u.m[c] = r1[c];
        u.m[3 + c] = r2[c];
    }
    // row3_k = conj(row1_{k+1} row2_{k+2} - row1_{k+2} row2_{k+1})
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int a = (k + 1) % 3, b = (k + 2) % 3;
        const double re = r1[a].r * r2[b].r - r1[a].i * r2[b].i - (r1[b].r * r2[a].r - r1[b].i * r2[a].i);
        const double im = r1[a].r * r2[b].i + r1[a].i * r2[b].r - (r1[b].r * r2[a].i + r1[b].i * r2[a].r);
        u.m[6 + k].r = re;
        u.m[6 + k].i = -im;
    }
}

template <typename T>
__device__ __forceinline__ void reunitarize(M3<T>& u) {
    M3<double> d = to_f64(u);
    reunitarize_f64(d);
#pragma unroll
    for (int k = 0; k < 9; k++) {
        u.m[k].r = (T)d.m[k].r;
        u.m[k].i = (T)d.m[k].i;
    }
}

// ---------------------------------------------------------------------------
// su(3) projections (basis i lambda_a / 2, order of TA_gaugefields_4D_serial.jl:243-260)
// ---------------------------------------------------------------------------
#define GF_SR3I 0.57735026918962576451

// coefficients of TA(v) = (v - v^dag)/2 - tr(v - v^dag)/6
template <typename T>
__device__ __forceinline__ void ta_coeffs(const M3<T>& v, T c[8]) {
    c[0] = v.m[1].i + v.m[3].i;
    c[1] = v.m[1].r - v.m[3].r;
    c[2] = v.m[0].i - v.m[4].i;
    c[3] = v.m[2].i + v.m[6].i;
    c[4] = v.m[2].r - v.m[6].r;
    c[5] = v.m[5].i + v.m[7].i;
    c[6] = v.m[5].r - v.m[7].r;
    c[7] = scale1(GF_SR3I, v.m[0].i + v.m[4].i - T(2) * v.m[8].i);
}

// TA(v) as a matrix
template <typename T>
__device__ __forceinline__ M3<T> ta_mat(const M3<T>& v) {
    M3<T> r;
    const T tri = (T)((double)(v.m[0].i + v.m[4].i + v.m[8].i) / 3.0);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        r.m[4 * i].r = T(0);
        r.m[4 * i].i = v.m[4 * i].i - tri;
    }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = i + 1; j < 3; j++) {
            T xr = T(0.5) * (v.m[3 * i + j].r - v.m[3 * j + i].r);
            T xi = T(0.5) * (v.m[3 * i + j].i + v.m[3 * j + i].i);
            r.m[3 * i + j].r = xr;
            r.m[3 * i + j].i = xi;
            r.m[3 * j + i].r = -xr;
            r.m[3 * j + i].i = xi;
        }
    return r;
}

// Hermitian H = (t/2) sum_a c_a lambda_a
template <typename T>
__device__ __forceinline__ M3<T> herm_from_coeffs(const T c[8], T t) {
    const T h = T(0.5) * t;
    const T s3 = T(GF_SR3I);
    T c1 = h * c[0], c2 = h * c[1], c3 = h * c[2], c4 = h * c[3];
    T c5 = h * c[4], c6 = h * c[5], c7 = h * c[6], c8 = h * c[7];
    M3<T> H;
    H.m[0] = {c3 + s3 * c8, T(0)};
    H.m[1] = {c1, -c2};
    H.m[2] = {c4, -c5};
    H.m[3] = {c1, c2};
    H.m[4] = {-c3 + s3 * c8, T(0)};
    H.m[5] = {c6, -c7};
    H.m[6] = {c4, c5};
    H.m[7] = {c6, c7};
    H.m[8] = {T(-2) * s3 * c8, T(0)};
    return H;
}

// ---------------------------------------------------------------------------
// Hermitian matrices by their independent entries: real diagonal d[k], upper triangle
// u[0] = (0,1), u[1] = (0,2), u[2] = (1,2).  Products with them skip the zero imaginary
// parts of the diagonal (10 instead of 12 real operations per output element).
// ---------------------------------------------------------------------------
template <typename T>
struct Herm {
    T d[3];
    cx<T> u[3];
};
// sum_a c_a lambda_a / 2 (the Hermitian part of a traceless anti-Hermitian i sum_a c_a lambda_a / 2)
template <typename T>
__device__ __forceinline__ Herm<T> herm_of_coeffs(const T c[8]) {
    const T s8 = scale1(GF_SR3I, c[7]);
    Herm<T> h;
    h.d[0] = T(0.5) * (c[2] + s8);
    h.d[1] = T(0.5) * (-c[2] + s8);
    h.d[2] = -s8;
    h.u[0] = {T(0.5) * c[0], T(-0.5) * c[1]};
    h.u[1] = {T(0.5) * c[3], T(-0.5) * c[4]};
    h.u[2] = {T(0.5) * c[5], T(-0.5) * c[6]};
    return h;
}
__host__ __device__ constexpr int herm_up(int k, int j) { return k + j - 1; }  // (0,1)->0 (0,2)->1 (1,2)->2
// element (k, j), k != j
template <typename T>
__device__ __forceinline__ cx<T> herm_off(const Herm<T>& h, int k, int j) {
    if (k < j) return h.u[herm_up(k, j)];
    cx<T> c = h.u[herm_up(j, k)];
    c.i = -c.i;
    return c;
}
template <typename T>
__device__ __forceinline__ M3<T> mul_mh(const M3<T>& m, const Herm<T>& h) {  // m h
    M3<T> c;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cx<T> t;
            t.r = m.m[3 * i + j].r * h.d[j];
            t.i = m.m[3 * i + j].i * h.d[j];
#pragma unroll
            for (int k = 0; k < 3; k++)
                if (k != j) cmac(t, m.m[3 * i + k], herm_off(h, k, j));
            c.m[3 * i + j] = t;
        }
    return c;
}
template <typename T>
__device__ __forceinline__ M3<T> mul_hm(const Herm<T>& h, const M3<T>& m) {  // h m
    M3<T> c;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cx<T> t;
            t.r = h.d[i] * m.m[3 * i + j].r;
            t.i = h.d[i] * m.m[3 * i + j].i;
#pragma unroll
            for (int k = 0; k < 3; k++)
                if (k != i) cmac(t, herm_off(h, i, k), m.m[3 * k + j]);
            c.m[3 * i + j] = t;
        }
    return c;
}
template <typename T>
__device__ __forceinline__ M3<T> mul_dh(const M3<T>& m, const Herm<T>& h) {  // m^dag h
    M3<T> c;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cx<T> t;
            t.r = m.m[3 * j + i].r * h.d[j];
            t.i = -m.m[3 * j + i].i * h.d[j];
#pragma unroll
            for (int k = 0; k < 3; k++)
                if (k != j) cmac_cn(t, m.m[3 * k + i], herm_off(h, k, j));
            c.m[3 * i + j] = t;
        }
    return c;
}
// i m
template <typename T>
__device__ __forceinline__ M3<T> times_i(const M3<T>& m) {
    M3<T> c;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        c.m[k].r = -m.m[k].i;
        c.m[k].i = m.m[k].r;
    }
    return c;
}

// ---------------------------------------------------------------------------
// Cayley-Hamilton coefficients in double (Morningstar & Peardon, PRD 69 054501)
// ---------------------------------------------------------------------------
struct dc {
    double r, i;
};
__device__ __forceinline__ dc dcm(dc a, dc b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ dc dcs(double s, dc a) { return {s * a.r, s * a.i}; }
__device__ __forceinline__ dc dca(dc a, dc b) { return {a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ dc dcsub(dc a, dc b) { return {a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ dc dcconj(dc a) { return {a.r, -a.i}; }

// H^2 for a Hermitian H (upper triangle read; the result is Hermitian with a real diagonal)
template <typename T>
__device__ __forceinline__ M3<T> herm_sq(const M3<T>& H) {
    auto h = [&](int i, int k) -> cx<T> {
        if (i <= k) return H.m[3 * i + k];
        cx<T> c = H.m[3 * k + i];
        c.i = -c.i;
        return c;
    };
    M3<T> S;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        T d = T(0);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const cx<T> a = h(i, k);
            d = fma(a.r, a.r, d);
            if (k != i) d = fma(a.i, a.i, d);
        }
        S.m[4 * i].r = d;
        S.m[4 * i].i = T(0);
    }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = i + 1; j < 3; j++) {
            cx<T> t;
            t.r = t.i = T(0);
#pragma unroll
            for (int k = 0; k < 3; k++) cmac(t, h(i, k), h(k, j));
            S.m[3 * i + j] = t;
            S.m[3 * j + i].r = t.r;
            S.m[3 * j + i].i = -t.i;
        }
    return S;
}

// c1 = tr(H^2)/2 and c0 = det H of a Hermitian traceless H, evaluated in double
template <typename T>
__device__ __forceinline__ void herm_invariants(const M3<T>& H, double& c0, double& c1) {
    double h00 = H.m[0].r, h11 = H.m[4].r, h22 = H.m[8].r;
    double ar = H.m[1].r, ai = H.m[1].i;  // H01
    double br = H.m[2].r, bi = H.m[2].i;  // H02
    double cr = H.m[5].r, ci = H.m[5].i;  // H12
    double a2 = ar * ar + ai * ai, b2 = br * br + bi * bi, c2 = cr * cr + ci * ci;
    c1 = 0.5 * (h00 * h00 + h11 * h11 + h22 * h22) + a2 + b2 + c2;
    // det = h00 h11 h22 - h00|c|^2 - h11|b|^2 - h22|a|^2 + 2 Re(a c conj(b))
    double re_acb = (ar * cr - ai * ci) * br + (ar * ci + ai * cr) * bi;
    c0 = h00 * h11 * h22 - h00 * c2 - h11 * b2 - h22 * a2 + 2.0 * re_acb;
}

// sin x and cos x.  For |x| <= 1 their Taylor series to x^19 and x^20 (the first omitted terms
// are below 2e-20 and 9e-22), evaluated by Horner in x^2; the library's sincos otherwise.  The
// Cayley-Hamilton arguments of MD steps, flow stages and stout levels are well inside |x| <= 1
// (theta/3 always is: at most pi/6), where this is a few FMAs instead of a range reduction.
__device__ __forceinline__ void sincos_small(double x, double* s, double* c) {
    if (fabs(x) <= 1.0) {
        const double x2 = x * x;
        double ps = -1.0 / 121645100408832000.0;  // -1/19!
        ps = fma(ps, x2, 1.0 / 355687428096000.0);
        ps = fma(ps, x2, -1.0 / 1307674368000.0);
        ps = fma(ps, x2, 1.0 / 6227020800.0);
        ps = fma(ps, x2, -1.0 / 39916800.0);
        ps = fma(ps, x2, 1.0 / 362880.0);
        ps = fma(ps, x2, -1.0 / 5040.0);
        ps = fma(ps, x2, 1.0 / 120.0);
        ps = fma(ps, x2, -1.0 / 6.0);
        *s = fma(ps * x2, x, x);
        double pc = 1.0 / 2432902008176640000.0;  // 1/20!
        pc = fma(pc, x2, -1.0 / 6402373705728000.0);
        pc = fma(pc, x2, 1.0 / 20922789888000.0);
        pc = fma(pc, x2, -1.0 / 87178291200.0);
        pc = fma(pc, x2, 1.0 / 479001600.0);
        pc = fma(pc, x2, -1.0 / 3628800.0);
        pc = fma(pc, x2, 1.0 / 40320.0);
        pc = fma(pc, x2, -1.0 / 720.0);
        pc = fma(pc, x2, 1.0 / 24.0);
        pc = fma(pc, x2, -0.5);
        *c = fma(pc, x2, 1.0);
    } else {
        sincos(x, s, c);
    }
}

// f_j (and, if WANT_B, b_1j = df_j/dc1, b_2j = df_j/dc0) for exp(iH) = f0 + f1 H + f2 H^2.
// The c0 -> -c0 symmetry keeps 9u^2 - w^2 >= 2 c1; sin(w)/w and cos(w)/w^2 - sin(w)/w^3
// switch to their Taylor series for |w| < 0.05.  Requires c1 > 0.
template <bool WANT_B>
__device__ __forceinline__ void ch_coeffs(double c0, double c1, dc f[3], dc b1[3], dc b2[3]) {
    const bool neg = c0 < 0.0;
    const double c0a = fabs(c0);
    const double c13 = c1 * (1.0 / 3.0);
    const double c0max = 2.0 * c13 * sqrt(c13);
    double ratio = c0a / c0max;
    ratio = ratio > 1.0 ? 1.0 : ratio;
    const double th = acos(ratio) * (1.0 / 3.0);
    double sth, cth;
    sincos_small(th, &sth, &cth);
    const double u = sqrt(c13) * cth;
    const double w = sqrt(c1) * sth;
    const double u2 = u * u, w2 = w * w;
    double xi0, xi1 = 0.0;
    double sw, cw;
    sincos_small(w, &sw, &cw);
    if (fabs(w) < 0.05) {
        xi0 = 1.0 - w2 * (1.0 / 6.0) * (1.0 - w2 * (1.0 / 20.0) * (1.0 - w2 * (1.0 / 42.0)));
        if (WANT_B) xi1 = -1.0 / 3.0 + w2 * (1.0 / 30.0) * (1.0 - w2 * (1.0 / 28.0) * (1.0 - w2 * (1.0 / 54.0)));
    } else {
        xi0 = sw / w;
        if (WANT_B) xi1 = cw / w2 - sw / (w2 * w);
    }
    double su, cu;
    sincos_small(u, &su, &cu);
    const double s2u = 2.0 * su * cu, c2u = 1.0 - 2.0 * su * su;  // double-angle forms
    const dc emiu = {cu, -su}, e2iu = {c2u, s2u};
    // h0 = (u^2 - w^2) e2iu + emiu (8u^2 cos w + 2iu(3u^2 + w^2) xi0)
    dc h0 = dca(dcs(u2 - w2, e2iu), dcm(emiu, dc{8.0 * u2 * cw, 2.0 * u * (3.0 * u2 + w2) * xi0}));
    // h1 = 2u e2iu - emiu (2u cos w - i(3u^2 - w^2) xi0)
    dc h1 = dcsub(dcs(2.0 * u, e2iu), dcm(emiu, dc{2.0 * u * cw, -(3.0 * u2 - w2) * xi0}));
    // h2 = e2iu - emiu (cos w + 3iu xi0)
    dc h2 = dcsub(e2iu, dcm(emiu, dc{cw, 3.0 * u * xi0}));
    const double iden = 1.0 / (9.0 * u2 - w2);
    f[0] = dcs(iden, h0);
    f[1] = dcs(iden, h1);
    f[2] = dcs(iden, h2);
    if (WANT_B) {
        dc r1[3], r2[3];
        // r10 = 2(u + i(u^2 - w^2)) e2iu + 2 emiu (4u(2 - iu) cos w + i(9u^2 + w^2 - iu(3u^2 + w^2)) xi0)
        {
            dc t1 = dcm(dc{2.0 * u, 2.0 * (u2 - w2)}, e2iu);
            dc a = {8.0 * u * cw, -4.0 * u2 * cw};                       // 4u(2 - iu) cos w
            dc bb = {u * (3.0 * u2 + w2) * xi0, (9.0 * u2 + w2) * xi0};  // i(9u^2+w^2 - iu(3u^2+w^2)) xi0
            r1[0] = dca(t1, dcs(2.0, dcm(emiu, dca(a, bb))));
        }
        // r11 = 2(1 + 2iu) e2iu + emiu (-2(1 - iu) cos w + i(6u + i(w^2 - 3u^2)) xi0)
        {
            dc t1 = dcm(dc{2.0, 4.0 * u}, e2iu);
            dc a = {-2.0 * cw, 2.0 * u * cw};
            dc bb = {-(w2 - 3.0 * u2) * xi0, 6.0 * u * xi0};
            r1[1] = dca(t1, dcm(emiu, dca(a, bb)));
        }
        // r12 = 2i e2iu + i emiu (cos w - 3(1 - iu) xi0)
        {
            dc t1 = dcm(dc{0.0, 2.0}, e2iu);
            dc in = {cw - 3.0 * xi0, 3.0 * u * xi0};
            r1[2] = dca(t1, dcm(dc{0.0, 1.0}, dcm(emiu, in)));
        }
        // r20 = -2 e2iu + 2iu emiu (cos w + (1 + 4iu) xi0 + 3u^2 xi1)
        {
            dc in = {cw + xi0 + 3.0 * u2 * xi1, 4.0 * u * xi0};
            r2[0] = dca(dcs(-2.0, e2iu), dcm(dc{0.0, 2.0 * u}, dcm(emiu, in)));
        }
        // r21 = -i emiu (cos w + (1 + 2iu) xi0 - 3u^2 xi1)
        {
            dc in = {cw + xi0 - 3.0 * u2 * xi1, 2.0 * u * xi0};
            r2[1] = dcm(dc{0.0, -1.0}, dcm(emiu, in));
        }
        // r22 = emiu (xi0 - 3iu xi1)
        r2[2] = dcm(emiu, dc{xi0, -3.0 * u * xi1});
        const double id2 = 0.5 * iden * iden;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            b1[j] = dcs(id2, dcsub(dca(dcs(2.0 * u, r1[j]), dcs(3.0 * u2 - w2, r2[j])),
                                   dcs(2.0 * (15.0 * u2 + w2), f[j])));
            b2[j] = dcs(id2, dcsub(dcsub(r1[j], dcs(3.0 * u, r2[j])), dcs(24.0 * u, f[j])));
        }
    }
    if (neg) {
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const double sg = (j & 1) ? -1.0 : 1.0;
            f[j] = dcs(sg, dcconj(f[j]));
            if (WANT_B) {
                b1[j] = dcs(sg, dcconj(b1[j]));
                b2[j] = dcs(-sg, dcconj(b2[j]));
            }
        }
    }
}

// The same coefficients from the exponential series reduced by Cayley-Hamilton: with
// H^3 = c1 H + c0 (traceless H), H^n = a_n + b_n H + c_n H^2 where (a, b, c)_0 = (1, 0, 0) and
// a_{n+1} = c0 c_n, b_{n+1} = a_n + c1 c_n, c_{n+1} = b_n, so f_j = sum_n i^n / n! (a_n, b_n, c_n)_j;
// b_1j = df_j/dc1 and b_2j = df_j/dc0 follow by differentiating the recursion.  For eigenvalues
// of modulus at most 1 (c1 <= 3/4, since max lambda^2 <= 4 c1 / 3) the terms after N are below
// e / N! in norm: N = 18 for f (1.7e-16), 20 with the derivatives, whose terms carry a factor n
// more (checked against scipy's expm and expm_frechet: 4e-16 and 1.4e-15).  Straight-line FMAs
// instead of acos, three sincos, square roots and divisions; no special case at small c1.
#define GF_CH_SERIES_C1 0.75
template <bool WANT_B, int N = WANT_B ? 20 : 18>
__device__ __forceinline__ void ch_coeffs_series(double c0, double c1, dc f[3], dc b1[3], dc b2[3]) {
    double a = 1.0, b = 0.0, c = 0.0;
    double a0 = 0.0, bb0 = 0.0, cc0 = 0.0, a1 = 0.0, bb1 = 0.0, cc1 = 0.0;  // d/dc0, d/dc1
    double fr[3] = {0.0, 0.0, 0.0}, fi[3] = {0.0, 0.0, 0.0};
    double g1r[3] = {0.0, 0.0, 0.0}, g1i[3] = {0.0, 0.0, 0.0}, g2r[3] = {0.0, 0.0, 0.0}, g2i[3] = {0.0, 0.0, 0.0};
    double w = 1.0;  // 1/n!
#pragma unroll
    for (int n = 0; n < N; n++) {
        const double s = ((n >> 1) & 1) ? -w : w;  // i^n / n! = s (n even) or i s (n odd)
        double* pf = (n & 1) ? fi : fr;
        pf[0] = fma(s, a, pf[0]);
        pf[1] = fma(s, b, pf[1]);
        pf[2] = fma(s, c, pf[2]);
        if (WANT_B) {
            double* p1 = (n & 1) ? g1i : g1r;
            double* p2 = (n & 1) ? g2i : g2r;
            p1[0] = fma(s, a1, p1[0]);
            p1[1] = fma(s, bb1, p1[1]);
            p1[2] = fma(s, cc1, p1[2]);
            p2[0] = fma(s, a0, p2[0]);
            p2[1] = fma(s, bb0, p2[1]);
            p2[2] = fma(s, cc0, p2[2]);
            const double na0 = fma(c0, cc0, c), na1 = c0 * cc1;
            const double nb0 = fma(c1, cc0, a0), nb1 = fma(c1, cc1, a1 + c);
            cc0 = bb0;
            cc1 = bb1;
            a0 = na0;
            a1 = na1;
            bb0 = nb0;
            bb1 = nb1;
        }
        const double na = c0 * c, nb = fma(c1, c, a);
        c = b;
        a = na;
        b = nb;
        w *= 1.0 / (double)(n + 1);
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {
        f[j] = {fr[j], fi[j]};
        if (WANT_B) {
            b1[j] = {g1r[j], g1i[j]};
            b2[j] = {g2r[j], g2i[j]};
        }
    }
}
// Shorter series for the exponential alone when every lane of the wave has small eigenvalues:
// the matrix tail after N terms is at most e^r r^N / N! for eigenvalues of modulus <= r, with
// r^2 <= 4 c1 / 3.  r <= 0.1 (c1 <= 0.0075): 11 terms, tail <= 2.8e-19; r <= 0.25
// (c1 <= 0.046875): 13 terms, tail <= 3.1e-18 (MD steps and flow stages are in these ranges).
// The derivative coefficients keep the full length: their basis 1, H, H^2 does not shrink the
// tail of the individual coefficients.
#define GF_CH_SHORT1_C1 0.0075
#define GF_CH_SHORT2_C1 0.046875

// f_j (and b_1j, b_2j) of exp(iH): the series for c1 <= 3/4, the closed form beyond
template <bool WANT_B>
__device__ __forceinline__ void exp_coeffs(double c0, double c1, dc f[3], dc b1[3], dc b2[3]) {
    if (!WANT_B && wave_all(c1 <= GF_CH_SHORT1_C1))
        ch_coeffs_series<false, 11>(c0, c1, f, b1, b2);
    else if (!WANT_B && wave_all(c1 <= GF_CH_SHORT2_C1))
        ch_coeffs_series<false, 13>(c0, c1, f, b1, b2);
    else if (c1 <= GF_CH_SERIES_C1)
        ch_coeffs_series<WANT_B>(c0, c1, f, b1, b2);
    else
        ch_coeffs<WANT_B>(c0, c1, f, b1, b2);
}


// exp(iH) for Hermitian traceless H
template <typename T>
__device__ __forceinline__ M3<T> expi_herm(const M3<T>& H) {
    double c0, c1;
    herm_invariants(H, c0, c1);
    M3<T> H2 = herm_sq(H);
    M3<T> E;
    dc f[3];
    exp_coeffs<false>(c0, c1, f, nullptr, nullptr);
    const T f0r = T(f[0].r), f0i = T(f[0].i), f1r = T(f[1].r), f1i = T(f[1].i);
    const T f2r = T(f[2].r), f2i = T(f[2].i);
#pragma unroll
    for (int k = 0; k < 9; k++) {
        // f1 H + f2 H^2
        T er = f1r * H.m[k].r - f1i * H.m[k].i + f2r * H2.m[k].r - f2i * H2.m[k].i;
        T ei = f1r * H.m[k].i + f1i * H.m[k].r + f2r * H2.m[k].i + f2i * H2.m[k].r;
        E.m[k].r = er;
        E.m[k].i = ei;
    }
    E.m[0].r += f0r;
    E.m[0].i += f0i;
    E.m[4].r += f0r;
    E.m[4].i += f0i;
    E.m[8].r += f0r;
    E.m[8].i += f0i;
    return E;
}

// exp(Q) for anti-Hermitian traceless Q  (H = -iQ)
template <typename T>
__device__ __forceinline__ M3<T> exp_antiherm(const M3<T>& Q) {
    M3<T> H;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        H.m[k].r = Q.m[k].i;
        H.m[k].i = -Q.m[k].r;
    }
    return expi_herm(H);
}

// L = C star d exp(Q)/dQ:  d tr(C e^Q) = tr(L dQ) for anti-Hermitian traceless Q
// (CdexpQdQ!, stout_fast.jl:888-946, 1031-1081)
template <typename T>
__device__ __forceinline__ M3<T> cdexpqdq(const M3<T>& C, const M3<T>& Q) {
    M3<T> H;  // H = -iQ (Hermitian)
#pragma unroll
    for (int k = 0; k < 9; k++) {
        H.m[k].r = Q.m[k].i;
        H.m[k].i = -Q.m[k].r;
    }
    double c0, c1;
    herm_invariants(H, c0, c1);
    dc f[3], b1[3], b2[3];
    exp_coeffs<true>(c0, c1, f, b1, b2);
    M3<T> H2 = herm_sq(H);
    // traces tr(C), tr(C H), tr(C H^2) in double
    dc t0 = {0.0, 0.0}, t1 = {0.0, 0.0}, t2 = {0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 3; i++) {
        t0.r += double(C.m[4 * i].r);
        t0.i += double(C.m[4 * i].i);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const cx<T> c = C.m[3 * i + k];
            const cx<T> h = H.m[3 * k + i];
            const cx<T> g = H2.m[3 * k + i];
            t1.r += double(c.r) * double(h.r) - double(c.i) * double(h.i);
            t1.i += double(c.r) * double(h.i) + double(c.i) * double(h.r);
            t2.r += double(c.r) * double(g.r) - double(c.i) * double(g.i);
            t2.i += double(c.r) * double(g.i) + double(c.i) * double(g.r);
        }
    }
    const dc tB1 = dca(dca(dcm(b1[0], t0), dcm(b1[1], t1)), dcm(b1[2], t2));
    const dc tB2 = dca(dca(dcm(b2[0], t0), dcm(b2[1], t1)), dcm(b2[2], t2));
    const T a1r = T(tB1.r), a1i = T(tB1.i), a2r = T(tB2.r), a2i = T(tB2.i);
    const T f1r = T(f[1].r), f1i = T(f[1].i), f2r = T(f[2].r), f2i = T(f[2].i);
    M3<T> HC = mul(H, C);
    acc_mul(HC, C, H);  // H C + C H
    M3<T> Lh;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        T r = a1r * H.m[k].r - a1i * H.m[k].i + a2r * H2.m[k].r - a2i * H2.m[k].i + f1r * C.m[k].r -
              f1i * C.m[k].i + f2r * HC.m[k].r - f2i * HC.m[k].i;
        T im = a1r * H.m[k].i + a1i * H.m[k].r + a2r * H2.m[k].i + a2i * H2.m[k].r + f1r * C.m[k].i +
               f1i * C.m[k].r + f2r * HC.m[k].i + f2i * HC.m[k].r;
        // divide by i: (r + i im)/i = im - i r
        Lh.m[k].r = im;
        Lh.m[k].i = -r;
    }
    return Lh;
}

// exp(Q) and L = X star d exp(Q)/dQ together, sharing one evaluation of the invariants and
// Cayley-Hamilton coefficients (stout backward: E = e^{Q_mu}, L = CdexpQdQ(U_mu Dout_mu, Q_mu))
template <typename T>
__device__ __forceinline__ void exp_dexp(const M3<T>& Q, const M3<T>& X, M3<T>& E, M3<T>& L) {
    M3<T> H;  // H = -iQ (Hermitian)
#pragma unroll
    for (int k = 0; k < 9; k++) {
        H.m[k].r = Q.m[k].i;
        H.m[k].i = -Q.m[k].r;
    }
    double c0, c1;
    herm_invariants(H, c0, c1);
    const M3<T> H2 = herm_sq(H);
    dc f[3], b1[3], b2[3];
    exp_coeffs<true>(c0, c1, f, b1, b2);
    const T f0r = T(f[0].r), f0i = T(f[0].i), f1r = T(f[1].r), f1i = T(f[1].i);
    const T f2r = T(f[2].r), f2i = T(f[2].i);
    dc t0 = {0.0, 0.0}, t1 = {0.0, 0.0}, t2 = {0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 3; i++) {
        t0.r += double(X.m[4 * i].r);
        t0.i += double(X.m[4 * i].i);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const cx<T> c = X.m[3 * i + k];
            const cx<T> h = H.m[3 * k + i];
            const cx<T> g = H2.m[3 * k + i];
            t1.r += double(c.r) * double(h.r) - double(c.i) * double(h.i);
            t1.i += double(c.r) * double(h.i) + double(c.i) * double(h.r);
            t2.r += double(c.r) * double(g.r) - double(c.i) * double(g.i);
            t2.i += double(c.r) * double(g.i) + double(c.i) * double(g.r);
        }
    }
    const dc tB1 = dca(dca(dcm(b1[0], t0), dcm(b1[1], t1)), dcm(b1[2], t2));
    const dc tB2 = dca(dca(dcm(b2[0], t0), dcm(b2[1], t1)), dcm(b2[2], t2));
    const T a1r = T(tB1.r), a1i = T(tB1.i), a2r = T(tB2.r), a2i = T(tB2.i);
    M3<T> HC = mul(H, X);
    acc_mul(HC, X, H);
#pragma unroll
    for (int k = 0; k < 9; k++) {
        T r = a1r * H.m[k].r - a1i * H.m[k].i + a2r * H2.m[k].r - a2i * H2.m[k].i + f1r * X.m[k].r -
              f1i * X.m[k].i + f2r * HC.m[k].r - f2i * HC.m[k].i;
        T im = a1r * H.m[k].i + a1i * H.m[k].r + a2r * H2.m[k].i + a2i * H2.m[k].r + f1r * X.m[k].i +
               f1i * X.m[k].r + f2r * HC.m[k].i + f2i * HC.m[k].r;
        L.m[k].r = im;
        L.m[k].i = -r;
    }
    // exp(Q) last: it is not needed while L is formed
#pragma unroll
    for (int k = 0; k < 9; k++) {
        E.m[k].r = f1r * H.m[k].r - f1i * H.m[k].i + f2r * H2.m[k].r - f2i * H2.m[k].i;
        E.m[k].i = f1r * H.m[k].i + f1i * H.m[k].r + f2r * H2.m[k].i + f2i * H2.m[k].r;
    }
    E.m[0].r += f0r;
    E.m[0].i += f0i;
    E.m[4].r += f0r;
    E.m[4].i += f0i;
    E.m[8].r += f0r;
    E.m[8].i += f0i;
}


// ---------------------------------------------------------------------------
// SU(3) link updates U' = exp(X) U.  For float storage the exponential, the product and the
// projection back onto SU(3) all run in double and only the stored result is rounded, so each
// update adds one rounding of the link instead of the accumulated float error of the product.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ M3<T> round_link(const M3<double>& d) {
    M3<T> r;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        r.m[k].r = (T)d.m[k].r;
        r.m[k].i = (T)d.m[k].i;
    }
    return r;
}

// exp(i t sum_a c_a lambda_a / 2) U
template <typename T>
__device__ __forceinline__ M3<T> exp_coeffs_times(const T c[8], double t, const M3<T>& u) {
    if constexpr (std::is_same<T, float>::value) {
        double cd[8];
#pragma unroll
        for (int a = 0; a < 8; a++) cd[a] = (double)c[a];
        M3<double> r = mul(expi_herm(herm_from_coeffs(cd, t)), to_f64(u));
        reunitarize_f64(r);
        return round_link<T>(r);
    } else {
        return mul(expi_herm(herm_from_coeffs(c, t)), u);
    }
}

// exp(Q) U for a traceless anti-Hermitian Q given in double (FP32 links: rounded once)
__device__ __forceinline__ M3<float> exp_antiherm_times(const M3<double>& Q, const M3<float>& u) {
    M3<double> r = mul(exp_antiherm(Q), to_f64(u));
    reunitarize_f64(r);
    return round_link<float>(r);
}
// exp(Q) U for a traceless anti-Hermitian Q
template <typename T>
__device__ __forceinline__ M3<T> exp_antiherm_times(const M3<T>& Q, const M3<T>& u) {
    if constexpr (std::is_same<T, float>::value) {
        M3<double> r = mul(exp_antiherm(to_f64(Q)), to_f64(u));
        reunitarize_f64(r);
        return round_link<T>(r);
    } else {
        return mul(exp_antiherm(Q), u);
    }
}

}  // namespace gf
