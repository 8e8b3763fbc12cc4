// k_flow_stout.hip -- Luscher RK3 Wilson-flow stage, stout smearing forward and backward.
#include "internal.hpp"

namespace gf {

// ---------------------------------------------------------------------------
// K7: one RK3 stage (flow!, gradientflow.jl:171-238, in the low-storage form):
//   F = TA(W_mu V_mu^dag)                  (add_force! plaqonly, AbstractGaugefields.jl:2717-2762)
//   A = alpha F + beta A
//   Wout_mu = exp(A_mu) W_mu               (exp_aF_U!, AbstractGaugefields.jl:2810-2841)
// stages (alpha, beta) = (-eps/4, 0), (-8eps/9, -17/9), (-3eps/4, -1)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_flow_stage(Geo g, const T* __restrict__ W, T* __restrict__ Wout,
                                                       T* __restrict__ A, T alpha, T beta) {
    const SiteMu sm = site_mu_index(g);
    const int s = sm.s, mu = sm.mu;
    if (s >= g.V) return;
    int x[4];
    decode(g, s, x);
    const unsigned V = g.V;
    const M3<T> sig = staple_sigma(g, W, V, s, x, mu);
    const M3<T> w = ldu(W, V, mu, s);
    T c[8];
    ta_coeffs(mul(w, sig), c);
    if (beta != T(0)) {
        T a[8];
        ldp(A, V, mu, s, a);
#pragma unroll
        for (int k = 0; k < 8; k++) c[k] = fma(alpha, c[k], beta * a[k]);
    } else {
#pragma unroll
        for (int k = 0; k < 8; k++) c[k] = alpha * c[k];
    }
    stp<T, GF_ST_NT>(A, V, mu, s, c);
    M3<T> wn = mul(expi_herm(herm_from_coeffs(c, T(1))), w);
    if (kFloatBulk<T>) reunitarize(wn);
    stu<T, GF_ST_NT>(Wout, V, mu, s, wn);
}

// ---------------------------------------------------------------------------
// K5: stout forward level (forward!, stout_fast.jl:250-274; calc_C! :603-624)
//   C_mu = rho V_mu = rho Sigma_mu^dag,  Q_mu = TA(C_mu U_mu^dag),  Uout_mu = exp(Q_mu) U_mu
// STORE keeps C_mu for the backward pass.
// ---------------------------------------------------------------------------
template <typename T, bool STORE>
__global__ __launch_bounds__(kBlock) void k_stout_forward(Geo g, const T* __restrict__ Uin, T* __restrict__ Uout,
                                                          T* __restrict__ Cst, T rho) {
    const SiteMu sm = site_mu_index(g);
    const int s = sm.s, mu = sm.mu;
    if (s >= g.V) return;
    int x[4];
    decode(g, s, x);
    const unsigned V = g.V;
    const M3<T> C = scal(rho, dag(staple_sigma(g, Uin, V, s, x, mu)));
    const M3<T> u = ldu(Uin, V, mu, s);
    const M3<T> E = exp_antiherm(ta_mat(mul_nd(C, u)));
    stu<T, GF_ST_NT>(Uout, V, mu, s, mul(E, u));
    if (STORE) stm<T, GF_ST_NT>(Cst, V, mu, s, C);
}

// ---------------------------------------------------------------------------
// K6a: stout backward, site-local part (backward_dSdUαUβρ_add!, stout_fast.jl:317-407):
//   Din_mu = Dout_mu e^{Q_mu} + (Lam_mu C_mu)^dag,   Lam_mu = TA(CdexpQdQ(U_mu Dout_mu, Q_mu)),
//   G_mu   = U_mu^dag Lam_mu  (= dS/dC_mu)
// STORED reads C_mu from the forward pass, otherwise it is recomputed from the staples.
// ---------------------------------------------------------------------------
template <typename T, bool STORED>
__global__ __launch_bounds__(kBlock) void k_stout_back_local(Geo g, const T* __restrict__ Uin,
                                                             const T* __restrict__ Cst, const T* __restrict__ Dout,
                                                             T* __restrict__ Din, T* __restrict__ G, T rho) {
    const SiteMu sm = site_mu_index(g);
    const int s = sm.s, mu = sm.mu;
    if (s >= g.V) return;
    const unsigned V = g.V;
    M3<T> C;
    if (STORED) {
        C = ldm(Cst, V, mu, s);
    } else {
        int x[4];
        decode(g, s, x);
        C = scal(rho, dag(staple_sigma(g, Uin, V, s, x, mu)));
    }
    M3<T> E, L;
    {
        const M3<T> u = ldu(Uin, V, mu, s);
        exp_dexp(ta_mat(mul_nd(C, u)), mul(u, ldm(Dout, V, mu, s)), E, L);
    }
    const M3<T> Lam = ta_mat(L);
    // re-read U_mu and Dout_mu (L1/L2 hits) instead of keeping them live across exp_dexp
    int s1 = s;
    asm volatile("" : "+v"(s1));
    stm<T, GF_ST_NT>(G, V, mu, s, mul_dn(ldu(Uin, V, mu, s1), Lam));
    M3<T> D = mul(ldm(Dout, V, mu, s1), E);
    acc_dag(D, mul(Lam, C));
    stm<T, GF_ST_NT>(Din, V, mu, s, D);
}

// ---------------------------------------------------------------------------
// K6b: stout backward, staple-derivative part (calc_dSdUν_fromdSCμ_add!, stout_fast.jl:712-785):
// for output direction a at site y and each plane b != a,
//   Din_a(y) += rho [ U_b(y+a) (G_a(y+b) U_b(y)^dag + U_a(y+b)^dag G_b(y))
//                   + U_b(y-b+a)^dag (G_a(y-b) U_b(y-b) + U_a(y-b)^dag G_b(y-b)^dag)
//                   + G_b(y+a)^dag U_a(y+b)^dag U_b(y)^dag
//                   + G_b(y+a-b) U_a(y-b)^dag U_b(y-b) ]
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_stout_back_stencil(Geo g, const T* __restrict__ U,
                                                               const T* __restrict__ G, T* __restrict__ Din, T rho) {
    const SiteMu sm = site_mu_index(g);
    const int y = sm.s, a = sm.mu;
    if (y >= g.V) return;
    int x[4];
    decode(g, y, x);
    const unsigned V = g.V;
    {
        M3<T> acc = zero3<T>();
        const int ypa = y + dlt(g, x, a, 1);
#pragma unroll
        for (int b = 0; b < 4; b++) {
            if (b == a) continue;
            const int ypb0 = y + dlt(g, x, b, 1);
            const int dmb = dlt(g, x, b, -1);
            const int ymb0 = y + dmb, ypamb0 = ypa + dmb;
            // each term group reloads its links through its own opaque index copies, so the
            // compiler issues them group by group instead of hoisting all twelve link loads
            {  // U_b(y+a) [G_a(y+b) U_b(y)^dag + U_a(y+b)^dag G_b(y)]
                int y0 = y, y1 = ypa, ypb = ypb0;
                asm volatile("" : "+v"(y0), "+v"(y1), "+v"(ypb));
                M3<T> t = mul_nd(ldm(G, V, a, ypb), ldu(U, V, b, y0));
                acc_mul_dn(t, ldu(U, V, a, ypb), ldm(G, V, b, y0));
                acc_mul(acc, ldu(U, V, b, y1), t);
            }
            {  // U_b(y-b+a)^dag [G_a(y-b) U_b(y-b) + (G_b(y-b) U_a(y-b))^dag]
                int ymb = ymb0, ypamb = ypamb0;
                asm volatile("" : "+v"(ymb), "+v"(ypamb));
                M3<T> t = mul(ldm(G, V, a, ymb), ldu(U, V, b, ymb));
                acc_dag(t, mul(ldm(G, V, b, ymb), ldu(U, V, a, ymb)));
                acc_mul_dn(acc, ldu(U, V, b, ypamb), t);
            }
            {  // (U_b(y) U_a(y+b) G_b(y+a))^dag
                int y0 = y, y1 = ypa, ypb = ypb0;
                asm volatile("" : "+v"(y0), "+v"(y1), "+v"(ypb));
                acc_dag(acc, mul(mul(ldu(U, V, b, y0), ldu(U, V, a, ypb)), ldm(G, V, b, y1)));
            }
            {  // G_b(y+a-b) U_a(y-b)^dag U_b(y-b)
                int ymb = ymb0, ypamb = ypamb0;
                asm volatile("" : "+v"(ymb), "+v"(ypamb));
                acc_mul(acc, ldm(G, V, b, ypamb), mul_dn(ldu(U, V, a, ymb), ldu(U, V, b, ymb)));
            }
        }
        M3<T> d = ldm(Din, V, a, y);
#pragma unroll
        for (int k = 0; k < 9; k++) {
            d.m[k].r = fma(rho, acc.m[k].r, d.m[k].r);
            d.m[k].i = fma(rho, acc.m[k].i, d.m[k].i);
        }
        stm<T, GF_ST_NT>(Din, V, a, y, d);
    }
}

// ---------------------------------------------------------------------------
template <typename T>
hipError_t launch_flow_stage(const Geo& g, const T* W, T* Wout, T* A, double alpha, double beta, hipStream_t st) {
    k_flow_stage<T><<<grid_sites_mu(g.V), kBlock, 0, st>>>(g, W, Wout, A, (T)alpha, (T)beta);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_stout_forward(const Geo& g, const T* Uin, T* Uout, T* Cst, double rho, hipStream_t st) {
    if (Cst)
        k_stout_forward<T, true><<<grid_sites_mu(g.V), kBlock, 0, st>>>(g, Uin, Uout, Cst, (T)rho);
    else
        k_stout_forward<T, false><<<grid_sites_mu(g.V), kBlock, 0, st>>>(g, Uin, Uout, nullptr, (T)rho);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_stout_backward(const Geo& g, const T* Uin, const T* Cst, const T* Dout, T* Din, T* G, double rho,
                                 hipStream_t st) {
    if (Cst)
        k_stout_back_local<T, true><<<grid_sites_mu(g.V), kBlock, 0, st>>>(g, Uin, Cst, Dout, Din, G, (T)rho);
    else
        k_stout_back_local<T, false><<<grid_sites_mu(g.V), kBlock, 0, st>>>(g, Uin, nullptr, Dout, Din, G, (T)rho);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_stout_back_stencil<T><<<grid_sites_mu(g.V), kBlock, 0, st>>>(g, Uin, G, Din, (T)rho);
    return hipGetLastError();
}

#define GF_INST2(T)                                                                                       \
    template hipError_t launch_flow_stage<T>(const Geo&, const T*, T*, T*, double, double, hipStream_t);  \
    template hipError_t launch_stout_forward<T>(const Geo&, const T*, T*, T*, double, hipStream_t);       \
    template hipError_t launch_stout_backward<T>(const Geo&, const T*, const T*, const T*, T*, T*, double, \
                                                 hipStream_t);

GF_INST2(double)
GF_INST2(float)

}  // namespace gf
