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
// ST: stencil form as in k_force (stencil_form).  At most 168 VGPRs: three waves per SIMD.
template <typename T, int ST>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_flow_stage(Geo g, const T* __restrict__ W, T* __restrict__ Wout,
                                                       T* __restrict__ A, double alpha, double beta) {
    __shared__ T2<T> ol[ST ? 4 : 1][6][64];
    __shared__ T2<T> tp[ST ? 7 : 1][4][64];
    const SiteMu sm = site_mu_index(g);
    const int mu = sm.mu;
    int s;
    if (!stencil_site<(ST > 0)>(g, sm, s)) return;
    int x[4];
    decode(g, s, x);
    const unsigned V = g.V;
    R6<T> own;
    T c[8];
    if constexpr (ST == ST_XROW) {
        ta_force_xrow(g, W, V, s, x, mu, own, ol, tp, c);
    } else if constexpr (ST == ST_SHARED) {
        ta_force_shared(g, W, V, s, x, mu, own, ol, tp, c);
    } else {
        const M3<T> sig = staple_sigma_own(g, W, V, s, x, mu, own);
        ta_coeffs(mul(full_of(own), sig), c);
    }
    const M3<T> w = full_of(own);
    if (beta != T(0)) {
        T a[8];
        ldp<T, GF_LD_NT>(A, V, mu, s, a);
#pragma unroll
        for (int k = 0; k < 8; k++) c[k] = (T)fma(alpha, (double)c[k], beta * (double)a[k]);
    } else {
#pragma unroll
        for (int k = 0; k < 8; k++) c[k] = scale1(alpha, c[k]);
    }
    stp<T, GF_ST_NT>(A, V, mu, s, c);
    const M3<T> wn = exp_coeffs_times(c, 1.0, w);
    stu<T, GF_ST_NT>(Wout, V, mu, s, wn);
}

// ---------------------------------------------------------------------------
// K5: stout forward level (forward!, stout_fast.jl:250-274; calc_C! :603-624)
//   C_mu = rho V_mu = rho Sigma_mu^dag,  Q_mu = TA(C_mu U_mu^dag),  Uout_mu = exp(Q_mu) U_mu
// STORE keeps C_mu for the backward pass.
// ---------------------------------------------------------------------------
template <typename T, bool STORE, bool SHARED>
__global__ __launch_bounds__(kBlock) void k_stout_forward(Geo g, const T* __restrict__ Uin, T* __restrict__ Uout,
                                                          T* __restrict__ Cst, double rho) {
    __shared__ T2<T> xs[SHARED ? 6 : 1][6][64];
    const SiteMu sm = site_mu_index(g);
    const int mu = sm.mu;
    int s;
    if (!stencil_site<SHARED>(g, sm, s)) return;
    int x[4];
    decode(g, s, x);
    const unsigned V = g.V;
    R6<T> own;
    const M3<T> sig = staple_sum_own<T, SHARED>(g, Uin, V, s, x, mu, own, xs);
    const M3<T> u = full_of(own);
    if constexpr (std::is_same<T, float>::value) {
        // C = rho Sigma^dag rounded once (the backward pass reads this C), Q and exp(Q) U in double
        const M3<T> C = round_link<T>(scal(rho, dag(to_f64(sig))));
        stu<T, GF_ST_NT>(Uout, V, mu, s, exp_antiherm_times(ta_mat(mul_nd(to_f64(C), to_f64(u))), u));
        if (STORE) stm<T, GF_ST_NT>(Cst, V, mu, s, C);
    } else {
        const M3<T> C = scal(rho, dag(sig));
        stu<T, GF_ST_NT>(Uout, V, mu, s, exp_antiherm_times(ta_mat(mul_nd(C, u)), u));
        if (STORE) stm<T, GF_ST_NT>(Cst, V, mu, s, C);
    }
}

// ---------------------------------------------------------------------------
// K6a: stout backward, site-local part (backward_dSdUαUβρ_add!, stout_fast.jl:317-407):
//   Din_mu = Dout_mu e^{Q_mu} + (Lam_mu C_mu)^dag,   Lam_mu = TA(CdexpQdQ(U_mu Dout_mu, Q_mu))
// Lam_mu (traceless anti-Hermitian) is stored as its 8 su(3) coefficients; the staple part
// uses G_mu = U_mu^dag Lam_mu (= dS/dC_mu) without storing it.
// STORED reads C_mu from the forward pass, otherwise it is recomputed from the staples.
// ---------------------------------------------------------------------------
// A 3x3 field element kept in the wave's own LDS slice between two uses (each lane reads back what
// it wrote: no barrier), instead of being held in registers or read from memory twice.
template <typename T>
using MStage = T2<T> (*)[9][64];
template <typename T>
__device__ __forceinline__ M3<T> stage_m3(MStage<T> st, int mu, int lane, const M3<T>& m) {
#pragma unroll
    for (int e = 0; e < 9; e++) {
        T2<T> v;
        v.x = m.m[e].r;
        v.y = m.m[e].i;
        st[mu][e][lane] = v;
    }
    return m;
}
template <typename T>
__device__ __forceinline__ M3<T> staged_m3(MStage<T> st, int mu, int lane) {
    // an opaque copy of the index, so that the compiler reads LDS rather than forwarding the
    // stored value (which would keep it live in registers)
    asm volatile("" : "+v"(lane));
    M3<T> m;
#pragma unroll
    for (int e = 0; e < 9; e++) {
        const T2<T> v = st[mu][e][lane];
        m.m[e].r = v.x;
        m.m[e].i = v.y;
    }
    return m;
}
template <typename T, bool STORED, bool SHARED>
__global__ __launch_bounds__(kBlock) void k_stout_back_local(Geo g, const T* __restrict__ Uin,
                                                             const T* __restrict__ Cst, const T* __restrict__ Dout,
                                                             T* __restrict__ Din, T* __restrict__ Lc, double rho) {
    __shared__ T2<T> xs[SHARED ? 6 : 1][6][64];
    // Dout_mu and a stored C_mu are used before and after exp_dexp: kept in LDS in between (the
    // second read from memory missed L2 for a third of the kernel's HBM reads)
    __shared__ T2<T> dst[4][9][64];
    __shared__ T2<T> cst[STORED ? 4 : 1][9][64];
    const SiteMu sm = site_mu_index(g);
    const int mu = sm.mu;
    const int lane = (int)(threadIdx.x & 63u);
    int s;
    if (!stencil_site<SHARED>(g, sm, s)) return;
    const unsigned V = g.V;
    M3<T> C;  // recomputed C; a stored C is staged in LDS instead of kept live
    R6<T> own;
    if (STORED) {
        own = ldr<T, GF_LD_NT>(Uin, V, mu, s);
    } else {  // the same staple evaluation (and rounding) as the forward kernel
        int x[4];
        decode(g, s, x);
        C = round_link<T>(scal(rho, dag(to_f64(staple_sum_own<T, SHARED>(g, Uin, V, s, x, mu, own, xs)))));
    }
    M3<T> E, L;
    {
        const M3<T> u = full_of(own);
        const M3<T> Cq = STORED ? stage_m3<T>(cst, mu, lane, ldm<T, GF_LD_NT>(Cst, V, mu, s)) : C;
        exp_dexp(ta_mat(mul_nd(Cq, u)), mul(u, stage_m3<T>(dst, mu, lane, ldm<T, GF_LD_NT>(Dout, V, mu, s))), E, L);
    }
    const M3<T> Lam = ta_mat(L);
    {
        T c[8];
        ta_coeffs(Lam, c);
        stp<T, GF_ST_NT>(Lc, V, mu, s, c);
    }
    M3<T> D = mul(staged_m3<T>(dst, mu, lane), E);
    acc_dag(D, mul(Lam, STORED ? staged_m3<T>(cst, mu, lane) : C));
    stm<T, GF_ST_NT>(Din, V, mu, s, D);
}

// ---------------------------------------------------------------------------
// K6b: stout backward, staple-derivative part (calc_dSdUν_fromdSCμ_add!, stout_fast.jl:712-785):
// for output direction a at site y and each plane b != a, with G = U^dag Lam,
//   Din_a(y) += rho [ U_b(y+a) (G_a(y+b) U_b(y)^dag + U_a(y+b)^dag G_b(y))
//                   + U_b(y-b+a)^dag (G_a(y-b) U_b(y-b) + U_a(y-b)^dag G_b(y-b)^dag)
//                   + G_b(y+a)^dag U_a(y+b)^dag U_b(y)^dag
//                   + G_b(y+a-b) U_a(y-b)^dag U_b(y-b) ]
// evaluated with Lam directly (Lam^dag = -Lam): with X = U_b(y+a) U_a(y+b)^dag, S = X U_b(y)^dag,
// Z = U_a(y-b)^dag U_b(y-b),
//   upper = X Lam_a(y+b) U_b(y)^dag + S Lam_b(y) - Lam_b(y+a) S
//   lower = U_b(y+a-b)^dag [U_a(y-b)^dag (Lam_a(y-b) - Lam_b(y-b)) U_b(y-b) + Lam_b(y+a-b) Z]
//         = U_b(y+a-b)^dag [U_a(y-b)^dag (Lam_a(y-b) - Lam_b(y-b)) + Lam_b(y+a-b) U_a(y-b)^dag] U_b(y-b)
// Six links and six 8-real Lam per plane instead of ten links and six 18-real G.  Lam = i H with
// H Hermitian (its 8 coefficients): products with Lam are i times Hermitian-structured products.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ Herm<T> herm_at(const T* __restrict__ Lc, unsigned V, int mu, unsigned s) {
    T c[8];
    ldp(Lc, V, mu, s, c);
    return herm_of_coeffs(c);
}
// c += SG i m h  (SG = +1 or -1), each product element accumulated as it is formed
template <typename T, int SG>
__device__ __forceinline__ void acc_i_mh(M3<T>& c, const M3<T>& m, const Herm<T>& h) {
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
            c.m[3 * i + j].r -= T(SG) * t.i;
            c.m[3 * i + j].i += T(SG) * t.r;
        }
}
// c += SG i h m
template <typename T, int SG>
__device__ __forceinline__ void acc_i_hm(M3<T>& c, const Herm<T>& h, const M3<T>& m) {
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
            c.m[3 * i + j].r -= T(SG) * t.i;
            c.m[3 * i + j].i += T(SG) * t.r;
        }
}
// c += h m^dag
template <typename T>
__device__ __forceinline__ void acc_hd(M3<T>& c, const Herm<T>& h, const M3<T>& m) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cx<T>& t = c.m[3 * i + j];
            t.r = fma(h.d[i], m.m[3 * j + i].r, t.r);
            t.i = fma(-h.d[i], m.m[3 * j + i].i, t.i);
#pragma unroll
            for (int k = 0; k < 3; k++)
                if (k != i) cmac_nc(t, herm_off(h, i, k), m.m[3 * j + k]);
        }
}

template <typename T>
__device__ __forceinline__ R6<T> lds_r6(XShare<T> ul, int d, int l) {
    R6<T> r;
#pragma unroll
    for (int e = 0; e < 6; e++) r.v[e] = ul[d][e][l];
    return r;
}
template <typename T>
__device__ __forceinline__ void lds_c8(PairShare<T> hl, int d, int l, T c[8]) {
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const T2<T> v = hl[d][e][l];
        c[2 * e] = v.x;
        c[2 * e + 1] = v.y;
    }
}
// Plane b of K6b added into acc.  SH: U_b(y) and Lam_b(y) from the workgroup's LDS copies
// (ul, hl) instead of memory; A0 (a = 0) and B0 (b = 0): the x-neighbour links and Lam at
// y + 1 (lane lxp) and y - 1 (lane lxm) from LDS as well (workgroups of whole x-rows).
template <typename T, bool SH, bool A0, bool B0>
__device__ __forceinline__ void back_plane(const Geo& g, const T* __restrict__ U, const T* __restrict__ Lc, unsigned V,
                                           int y, const int x[4], int a, int b, int ypa, M3<T>& acc, XShare<T> ul,
                                           PairShare<T> hl, int lane, int lxm, int lxp) {
    const int ypb0 = y + dlt(g, x, b, 1);
    const int dmb = dlt(g, x, b, -1);
    const int ymb0 = y + dmb, ypamb0 = ypa + dmb;
    // each group loads its links through its own opaque index copies, so the compiler
    // issues them group by group instead of hoisting all of them
    {  // upper: i X H_a(y+b) U_b(y)^dag + i [S H_b(y) - H_b(y+a) S],  S = X U_b(y)^dag
        int y0 = y, y1 = ypa, ypb = ypb0;
        asm volatile("" : "+v"(y0), "+v"(y1), "+v"(ypb));
        const R6<T> uby = A0 ? lds_r6<T>(ul, b, lxp) : ldr(U, V, b, y1);                // U_b(y+a)
        const M3<T> X = su3_mul_nd(uby, full_of(B0 ? lds_r6<T>(ul, a, lxp) : ldr(U, V, a, ypb)));  // U_a(y+b)
        T c[8];
        if (B0) lds_c8<T>(hl, a, lxp, c);  // Lam_a(y+b)
        else ldp(Lc, V, a, ypb, c);
        const M3<T> XH = times_i(mul_mh(X, herm_of_coeffs(c)));
        const M3<T> ub = full_of(SH ? lds_r6<T>(ul, b, lane) : ldr(U, V, b, y0));        // U_b(y)
        acc_mul_nd(acc, XH, ub);
        const M3<T> S = su3_mul_nd(rows_of(X), ub);
        if (SH) lds_c8<T>(hl, b, lane, c);  // Lam_b(y)
        else ldp(Lc, V, b, y0, c);
        acc_i_mh<T, 1>(acc, S, herm_of_coeffs(c));
        if (A0) lds_c8<T>(hl, b, lxp, c);  // Lam_b(y+a)
        else ldp(Lc, V, b, y1, c);
        acc_i_hm<T, -1>(acc, herm_of_coeffs(c), S);
    }
    {  // lower: U_b(y+a-b)^dag i [U_a(y-b)^dag H_d + H_b(y+a-b) U_a(y-b)^dag] U_b(y-b),
       // H_d = H_a(y-b) - H_b(y-b)
        int ymb = ymb0, ypamb = ypamb0;
        asm volatile("" : "+v"(ymb), "+v"(ypamb));
        const M3<T> ua = full_of(B0 ? lds_r6<T>(ul, a, lxm) : ldr(U, V, a, ymb));  // U_a(y-b)
        T ca[8], cb[8];
        if (B0) {
            lds_c8<T>(hl, a, lxm, ca);
            lds_c8<T>(hl, 0, lxm, cb);
        } else {
            ldp(Lc, V, a, ymb, ca);
            ldp(Lc, V, b, ymb, cb);
        }
#pragma unroll
        for (int k = 0; k < 8; k++) ca[k] -= cb[k];
        M3<T> A = mul_dh(ua, herm_of_coeffs(ca));
        acc_hd(A, herm_at(Lc, V, b, ypamb), ua);
        const M3<T> ubm = full_of(B0 ? lds_r6<T>(ul, 0, lxm) : ldr(U, V, b, ymb));  // U_b(y-b)
        acc_mul_dn(acc, ldu(U, V, b, ypamb), times_i(mul(A, ubm)));
    }
}

// The staple part of K6b for output direction a at site y (all planes b != a), with the
// stencil form ST: ST_OWN every link and Lam from memory; ST_SHARED U(y) and Lam(y) of the
// workgroup's sites through LDS (per wave 6 link and 3 Lam loads fewer); ST_XROW also the
// x-neighbours (6 + 3 more).  The LDS forms fill ul/hl and pass one workgroup barrier.
template <typename T, int ST>
__device__ __forceinline__ M3<T> back_stencil_acc(const Geo& g, const T* __restrict__ U, const T* __restrict__ Lc,
                                                  int y, const int x[4], int a, int lane, XShare<T> ul,
                                                  PairShare<T> hl) {
    const unsigned V = g.V;
    if constexpr (ST > 0) {
        const R6<T> u = ldr(U, V, a, (unsigned)y);
        T c[8];
        ldp(Lc, V, a, (unsigned)y, c);
#pragma unroll
        for (int e = 0; e < 6; e++) ul[a][e][lane] = u.v[e];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            T2<T> v;
            v.x = c[2 * e];
            v.y = c[2 * e + 1];
            hl[a][e][lane] = v;
        }
        __syncthreads();
    }
    const int nxm1 = g.n[0] - 1;
    const int lxm = xlane(lane, nxm1, -1), lxp = xlane(lane, nxm1, 1);
    M3<T> acc = zero3<T>();
    const int ypa = y + dlt(g, x, a, 1);
    if constexpr (ST == ST_XROW) {
        if (a == 0) {
#pragma unroll 1
            for (int b = 1; b < 4; b++) back_plane<T, true, true, false>(g, U, Lc, V, y, x, a, b, ypa, acc, ul, hl, lane, lxm, lxp);
        } else {
            back_plane<T, true, false, true>(g, U, Lc, V, y, x, a, 0, ypa, acc, ul, hl, lane, lxm, lxp);
#pragma unroll 1
            for (int b = 1; b < 4; b++) {
                if (b == a) continue;
                back_plane<T, true, false, false>(g, U, Lc, V, y, x, a, b, ypa, acc, ul, hl, lane, lxm, lxp);
            }
        }
    } else {
#pragma unroll 1
        for (int b = 0; b < 4; b++) {
            if (b == a) continue;
            back_plane<T, ST == ST_SHARED, false, false>(g, U, Lc, V, y, x, a, b, ypa, acc, ul, hl, lane, lxm, lxp);
        }
    }
    return acc;
}
// Din_a(y) + rho * (staple part): the derivative with respect to the level's input links
template <typename T>
__device__ __forceinline__ M3<T> back_stencil_din(const T* __restrict__ Din, unsigned V, int a, int y, double rho,
                                                  const M3<T>& acc) {
    M3<T> dn = ldm<T, GF_LD_NT>(Din, V, a, y);
#pragma unroll
    for (int k = 0; k < 9; k++) {
        dn.m[k].r = fma1(rho, acc.m[k].r, dn.m[k].r);
        dn.m[k].i = fma1(rho, acc.m[k].i, dn.m[k].i);
    }
    return dn;
}

// K6b.  At most 168 VGPRs (FP64), so that three waves fit per SIMD.
// OUT < 0: Din += rho (staple part).  OUT = FORCE_STORE / FORCE_KICK / FORCE_KICK_EXP: the thin-link
// level of a stout force (test/HMCstout_test_nowing.jl:99-118): Din is not stored but turned into
// the force FP = -(1/3) TA(U_a(y) Din), or FP += eps of that, in the same kernel; FORCE_KICK_EXP
// then also writes Unew = exp(eps_u FP) U_a(y) (update_gaugefields!, molecular_dynamics.jl:513-531)
// to a buffer other than U, which the other threads still read.  OUT = FORCE_KICK_S64 /
// FORCE_KICK_EXP_S64 (FP32 only): FP is the FP64 momenta, kicked by the force as rounded to float;
// U64 the FP64 links, U64_a(y) <- exp(eps_u FP) U64_a(y) in place, and Unew its float rounding.
template <typename T, int ST, int OUT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_stout_back_stencil(
    Geo g, const T* __restrict__ U, const T* __restrict__ Lc, T* __restrict__ Din, double rho,
    T* __restrict__ FP, double eps, T* __restrict__ Unew, double eps_u, double* __restrict__ U64) {
    __shared__ T2<T> ul[ST ? 4 : 1][6][64];
    __shared__ T2<T> hl[ST ? 4 : 1][4][64];
    const SiteMu sm = site_mu_index(g);
    const int a = sm.mu;
    int y;
    if (!stencil_site<(ST > 0)>(g, sm, y)) return;
    int x[4];
    decode(g, y, x);
    const unsigned V = g.V;
    const int lane = (int)(threadIdx.x & 63u);
    const M3<T> dn = back_stencil_din(Din, V, a, y, rho, back_stencil_acc<T, ST>(g, U, Lc, y, x, a, lane, ul, hl));
    if constexpr (OUT < 0) {
        stm<T, GF_ST_NT>(Din, V, a, y, dn);
    } else {  // the force from the thin-link derivative
        constexpr double fac = -1.0 / 3.0;
        R6<T> ur;
        if constexpr (ST > 0)
            ur = lds_r6<T>(ul, a, lane);
        else
            ur = ldr(U, V, a, (unsigned)y);
        const M3<T> ua = full_of(ur);
        T c[8];
        ta_coeffs(mul(ua, dn), c);
        if constexpr (OUT == FORCE_STORE) {
#pragma unroll
            for (int k = 0; k < 8; k++) c[k] = scale1(fac, c[k]);
            stp<T, GF_ST_NT>(FP, V, a, y, c);
        } else if constexpr (OUT == FORCE_KICK_S64 || OUT == FORCE_KICK_EXP_S64) {
            double* P64 = reinterpret_cast<double*>(FP);
            double p[8];
            ldp<double, GF_LD_NT>(P64, V, a, y, p);
#pragma unroll
            for (int k = 0; k < 8; k++) p[k] = fma(eps, (double)scale1(fac, c[k]), p[k]);
            stp<double, GF_ST_NT>(P64, V, a, y, p);
            if constexpr (OUT == FORCE_KICK_EXP_S64) {
                const M3<double> un = exp_coeffs_times(p, eps_u, ldu<double, GF_LD_NT>(U64, V, a, y));
                stu<double, GF_ST_NT>(U64, V, a, y, un);
                stu<T, GF_ST_NT>(Unew, V, a, y, round_link<T>(un));
            }
        } else {
            T p[8];
            ldp<T, GF_LD_NT>(FP, V, a, y, p);
            const double f = eps * fac;
#pragma unroll
            for (int k = 0; k < 8; k++) p[k] = fma1(f, c[k], p[k]);
            stp<T, GF_ST_NT>(FP, V, a, y, p);
            if constexpr (OUT == FORCE_KICK_EXP) stu<T, GF_ST_NT>(Unew, V, a, y, exp_coeffs_times(p, eps_u, ua));
        }
    }
}

// K6a of level l-1 chained after the K6b of level l (stored C): the derivative with respect to
// level l's input links, Dout_{l-1}(y), is formed in registers by the K6b part and is exactly what
// the site-local step of level l-1 needs at the same (y, a), so it is neither stored nor read
// back (1152 B per site fewer per level boundary).  Same arithmetic as k_stout_back_local<T,
// true, false>: Din_{l-1} = Dout e^Q + (Lam C)^dag into Din (each thread reads then writes only its
// own element), Lam_{l-1} into Lout (another buffer than Lc, whose entries at neighbour sites the
// K6b part reads).  Dout and C are staged in the lane's LDS slice between their two uses, in LDS
// that the K6b part's copies occupied before (one workgroup barrier in between).
template <typename T>
__device__ __forceinline__ void back_local_stored(const T* __restrict__ Uin, const T* __restrict__ Cst,
                                                  const M3<T>& dout, T* __restrict__ Din, T* __restrict__ Lout,
                                                  unsigned V, int mu, int s, int lane, MStage<T> dst, MStage<T> cst) {
    M3<T> E, L;
    {
        const M3<T> u = full_of(ldr<T, GF_LD_NT>(Uin, V, mu, s));
        const M3<T> Cq = stage_m3<T>(cst, mu, lane, ldm<T, GF_LD_NT>(Cst, V, mu, s));
        exp_dexp(ta_mat(mul_nd(Cq, u)), mul(u, stage_m3<T>(dst, mu, lane, dout)), E, L);
    }
    const M3<T> Lam = ta_mat(L);
    {
        T c[8];
        ta_coeffs(Lam, c);
        stp<T, GF_ST_NT>(Lout, V, mu, s, c);
    }
    M3<T> D = mul(staged_m3<T>(dst, mu, lane), E);
    acc_dag(D, mul(Lam, staged_m3<T>(cst, mu, lane)));
    stm<T, GF_ST_NT>(Din, V, mu, s, D);
}
template <typename T>
constexpr int kChainLds = (int)(2 * 4 * 9 * 64 * sizeof(T2<T>));  // dst + cst; >= ul + hl
// FP32: three waves per SIMD (at most 168 VGPRs); FP64 holds two (its LDS allows two workgroups per CU)
template <typename T, int ST>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 3 : 1))) void k_stout_back_chain(Geo g, const T* __restrict__ U, const T* __restrict__ Lc,
                                                             T* __restrict__ Din, double rho,
                                                             const T* __restrict__ Uprev, const T* __restrict__ Cprev,
                                                             T* __restrict__ Lout) {
    __shared__ __align__(16) unsigned char smem[kChainLds<T>];
    static_assert(4 * 10 * 64 * sizeof(T2<T>) <= kChainLds<T>, "ul + hl fit the staging space");
    XShare<T> ul = reinterpret_cast<XShare<T>>(smem);
    PairShare<T> hl = reinterpret_cast<PairShare<T>>(smem + 4 * 6 * 64 * sizeof(T2<T>));
    MStage<T> dst = reinterpret_cast<MStage<T>>(smem);
    MStage<T> cst = reinterpret_cast<MStage<T>>(smem + 4 * 9 * 64 * sizeof(T2<T>));
    const SiteMu sm = site_mu_index(g);
    const int a = sm.mu;
    int y;
    if (!stencil_site<(ST > 0)>(g, sm, y)) return;
    int x[4];
    decode(g, y, x);
    const unsigned V = g.V;
    const int lane = (int)(threadIdx.x & 63u);
    const M3<T> dn = back_stencil_din(Din, V, a, y, rho, back_stencil_acc<T, ST>(g, U, Lc, y, x, a, lane, ul, hl));
    if constexpr (ST > 0) __syncthreads();  // every wave is done with ul/hl before the staging reuses them
    back_local_stored<T>(Uprev, Cprev, dn, Din, Lout, V, a, y, lane, dst, cst);
}

// K2 chained into the K6a of the top level (stored C): dS/dU of the top smeared level, coeff
// Sigma_mu(y), is Dout of the top level's site-local step at the same (y, mu)
template <typename T, bool SHARED>
__global__ __launch_bounds__(kBlock) void k_dsdu_chain(Geo g, const T* __restrict__ Utop, double coeff,
                                                       const T* __restrict__ Uin, const T* __restrict__ Cst,
                                                       T* __restrict__ Din, T* __restrict__ Lout) {
    __shared__ __align__(16) unsigned char smem[kChainLds<T>];
    static_assert(6 * 6 * 64 * sizeof(T2<T>) <= kChainLds<T>, "xs fits the staging space");
    XShare<T> xs = reinterpret_cast<XShare<T>>(smem);
    MStage<T> dst = reinterpret_cast<MStage<T>>(smem);
    MStage<T> cst = reinterpret_cast<MStage<T>>(smem + 4 * 9 * 64 * sizeof(T2<T>));
    const SiteMu sm = site_mu_index(g);
    int s;
    if (!stencil_site<SHARED>(g, sm, s)) return;
    int x[4];
    decode(g, s, x);
    M3<T> sig;
    if constexpr (SHARED) {
        R6<T> own;
        sig = staple_sigma_shared(g, Utop, (unsigned)g.V, s, x, sm.mu, own, xs);
    } else {
        sig = staple_sigma(g, Utop, (unsigned)g.V, s, x, sm.mu);
    }
#pragma unroll
    for (int k = 0; k < 9; k++) {
        sig.m[k].r = scale1(coeff, sig.m[k].r);
        sig.m[k].i = scale1(coeff, sig.m[k].i);
    }
    if constexpr (SHARED) __syncthreads();  // every wave is done with xs before the staging reuses it
    back_local_stored<T>(Uin, Cst, sig, Din, Lout, (unsigned)g.V, sm.mu, s, (int)(threadIdx.x & 63u), dst, cst);
}

// ---------------------------------------------------------------------------
template <typename T>
hipError_t launch_flow_stage(const Geo& g, const T* W, T* Wout, T* A, double alpha, double beta, hipStream_t st) {
    const int form = stencil_form(g), nb = grid_sites_mu(g.V);
    if (form == ST_XROW)
        k_flow_stage<T, ST_XROW><<<nb, kBlock, 0, st>>>(g, W, Wout, A, alpha, beta);
    else if (form == ST_SHARED)
        k_flow_stage<T, ST_SHARED><<<nb, kBlock, 0, st>>>(g, W, Wout, A, alpha, beta);
    else
        k_flow_stage<T, ST_OWN><<<nb, kBlock, 0, st>>>(g, W, Wout, A, alpha, beta);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_stout_forward(const Geo& g, const T* Uin, T* Uout, T* Cst, double rho, hipStream_t st) {
    const int nb = grid_sites_mu(g.V);
    const bool sh = g.V % 64 == 0;
    if (Cst && sh)
        k_stout_forward<T, true, true><<<nb, kBlock, 0, st>>>(g, Uin, Uout, Cst, rho);
    else if (Cst)
        k_stout_forward<T, true, false><<<nb, kBlock, 0, st>>>(g, Uin, Uout, Cst, rho);
    else if (sh)
        k_stout_forward<T, false, true><<<nb, kBlock, 0, st>>>(g, Uin, Uout, nullptr, rho);
    else
        k_stout_forward<T, false, false><<<nb, kBlock, 0, st>>>(g, Uin, Uout, nullptr, rho);
    return hipGetLastError();
}
template <typename T, int OUT>
void launch_back_stencil(int form, int nb, hipStream_t st, const Geo& g, const T* Uin, const T* Lc, T* Din,
                         double rho, T* FP, double eps, T* Unew, double eps_u, double* U64) {
    if (form == ST_XROW)
        k_stout_back_stencil<T, ST_XROW, OUT><<<nb, kBlock, 0, st>>>(g, Uin, Lc, Din, rho, FP, eps, Unew, eps_u, U64);
    else if (form == ST_SHARED)
        k_stout_back_stencil<T, ST_SHARED, OUT><<<nb, kBlock, 0, st>>>(g, Uin, Lc, Din, rho, FP, eps, Unew, eps_u, U64);
    else
        k_stout_back_stencil<T, ST_OWN, OUT><<<nb, kBlock, 0, st>>>(g, Uin, Lc, Din, rho, FP, eps, Unew, eps_u, U64);
}
template <typename T>
hipError_t launch_stout_back_chain(const Geo& g, const T* Uin, const T* Lc, T* Din, double rho, const T* Uprev,
                                   const T* Cprev, T* Lout, hipStream_t st) {
    const int nb = grid_sites_mu(g.V), form = stencil_form(g);
    if (form == ST_XROW)
        k_stout_back_chain<T, ST_XROW><<<nb, kBlock, 0, st>>>(g, Uin, Lc, Din, rho, Uprev, Cprev, Lout);
    else if (form == ST_SHARED)
        k_stout_back_chain<T, ST_SHARED><<<nb, kBlock, 0, st>>>(g, Uin, Lc, Din, rho, Uprev, Cprev, Lout);
    else
        k_stout_back_chain<T, ST_OWN><<<nb, kBlock, 0, st>>>(g, Uin, Lc, Din, rho, Uprev, Cprev, Lout);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_dsdu_chain(const Geo& g, const T* Utop, double coeff, const T* Uin, const T* Cst, T* Din, T* Lout,
                             hipStream_t st) {
    const int nb = grid_sites_mu(g.V);
    if (g.V % 64 == 0)
        k_dsdu_chain<T, true><<<nb, kBlock, 0, st>>>(g, Utop, coeff, Uin, Cst, Din, Lout);
    else
        k_dsdu_chain<T, false><<<nb, kBlock, 0, st>>>(g, Utop, coeff, Uin, Cst, Din, Lout);
    return hipGetLastError();
}
// K6a alone (the top level of a backward pass that starts from a given dS/dU)
template <typename T>
hipError_t launch_stout_back_local(const Geo& g, const T* Uin, const T* Cst, const T* Dout, T* Din, T* Lc, double rho,
                                   hipStream_t st) {
    const int nb = grid_sites_mu(g.V);
    if (Cst)
        k_stout_back_local<T, true, false><<<nb, kBlock, 0, st>>>(g, Uin, Cst, Dout, Din, Lc, rho);
    else if (g.V % 64 == 0)  // recompute C exactly as k_stout_forward<T, *, true> formed it
        k_stout_back_local<T, false, true><<<nb, kBlock, 0, st>>>(g, Uin, nullptr, Dout, Din, Lc, rho);
    else
        k_stout_back_local<T, false, false><<<nb, kBlock, 0, st>>>(g, Uin, nullptr, Dout, Din, Lc, rho);
    return hipGetLastError();
}
// K6b alone (the thin-link level, or any level whose lower neighbour is not chained)
template <typename T>
hipError_t launch_stout_back_stencil(const Geo& g, const T* Uin, const T* Lc, T* Din, double rho, hipStream_t st,
                                     int fmode, T* FP, double eps, T* Unew, double eps_u, double* U64) {
    const int nb = grid_sites_mu(g.V), form = stencil_form(g);
    if (fmode == FORCE_STORE)
        launch_back_stencil<T, FORCE_STORE>(form, nb, st, g, Uin, Lc, Din, rho, FP, eps, Unew, eps_u, U64);
    else if (fmode == FORCE_KICK)
        launch_back_stencil<T, FORCE_KICK>(form, nb, st, g, Uin, Lc, Din, rho, FP, eps, Unew, eps_u, U64);
    else if (fmode == FORCE_KICK_EXP)
        launch_back_stencil<T, FORCE_KICK_EXP>(form, nb, st, g, Uin, Lc, Din, rho, FP, eps, Unew, eps_u, U64);
    else if (fmode == FORCE_KICK_S64 || fmode == FORCE_KICK_EXP_S64) {
        if constexpr (sizeof(T) == 4) {  // FP32 force kernels only
            if (!U64 && fmode == FORCE_KICK_EXP_S64) return hipErrorInvalidValue;
            if (fmode == FORCE_KICK_S64)
                launch_back_stencil<T, FORCE_KICK_S64>(form, nb, st, g, Uin, Lc, Din, rho, FP, eps, Unew, eps_u, U64);
            else
                launch_back_stencil<T, FORCE_KICK_EXP_S64>(form, nb, st, g, Uin, Lc, Din, rho, FP, eps, Unew, eps_u,
                                                           U64);
        } else {
            return hipErrorInvalidValue;
        }
    } else
        launch_back_stencil<T, -1>(form, nb, st, g, Uin, Lc, Din, rho, FP, eps, Unew, eps_u, U64);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_stout_backward(const Geo& g, const T* Uin, const T* Cst, const T* Dout, T* Din, T* Lc, double rho,
                                 hipStream_t st, int fmode, T* FP, double eps, T* Unew, double eps_u, double* U64) {
    hipError_t e = launch_stout_back_local<T>(g, Uin, Cst, Dout, Din, Lc, rho, st);
    if (e != hipSuccess) return e;
    return launch_stout_back_stencil<T>(g, Uin, Lc, Din, rho, st, fmode, FP, eps, Unew, eps_u, U64);
}

#define GF_INST2(T)                                                                                       \
    template hipError_t launch_flow_stage<T>(const Geo&, const T*, T*, T*, double, double, hipStream_t);  \
    template hipError_t launch_stout_forward<T>(const Geo&, const T*, T*, T*, double, hipStream_t);       \
    template hipError_t launch_stout_backward<T>(const Geo&, const T*, const T*, const T*, T*, T*, double, \
                                                 hipStream_t, int, T*, double, T*, double, double*);       \
    template hipError_t launch_stout_back_local<T>(const Geo&, const T*, const T*, const T*, T*, T*, double, \
                                                   hipStream_t);                                            \
    template hipError_t launch_stout_back_stencil<T>(const Geo&, const T*, const T*, T*, double, hipStream_t, \
                                                     int, T*, double, T*, double, double*);                 \
    template hipError_t launch_stout_back_chain<T>(const Geo&, const T*, const T*, T*, double, const T*,      \
                                                   const T*, T*, hipStream_t);                              \
    template hipError_t launch_dsdu_chain<T>(const Geo&, const T*, double, const T*, const T*, T*, T*, hipStream_t);

GF_INST2(double)
GF_INST2(float)

}  // namespace gf
