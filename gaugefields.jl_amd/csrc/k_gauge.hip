// k_gauge.hip -- plaquette/action sums, staple derivative, fused force + momentum kick
// (+ exp link update), exp link update, layout conversion.
//
// Stencil kernels run one thread per (site, link direction) (site_mu_index): the four
// directions of a group of 64 sites are the four wavefronts of one workgroup, and groups are
// visited in 4D Morton order per XCD, so neighbour links are re-read from L1/L2 instead of HBM.
#include <algorithm>

#include "internal.hpp"

namespace gf {

// ---------------------------------------------------------------------------
// K1: plaquette sum  sum_{x, mu<nu} Re tr U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
// (calculate_Plaquette, AbstractGaugefields.jl:2684-2699)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_plaquette(Geo g, const T* __restrict__ U, double* __restrict__ partial) {
    const SiteMu sm = site_mu_index(g);
    const int s = sm.s;
    double v = 0.0;
    if (s < g.V) {
        int x[4];
        decode(g, s, x);
        const unsigned V = g.V;
        // the six planes (0,1) (0,2) (0,3) (1,2) (1,3) (2,3): wavefront w takes planes w and w+4
#pragma unroll 1
        for (int pl = sm.mu; pl < 6; pl += 4) {
            const int mu = pl < 3 ? 0 : (pl < 5 ? 1 : 2);
            const int nu = pl < 3 ? pl + 1 : (pl < 5 ? pl - 1 : 3);
            const int spm = s + dlt(g, x, mu, 1), spn = s + dlt(g, x, nu, 1);
            // products in FP64 also for FP32 fields (the stored rows widened, row 2 rebuilt in double):
            // the action/Hamiltonian sums carry no float rounding
            auto ld = [&](int d, int t) { return full_of(r6_f64(ldr(U, V, d, (unsigned)t))); };
            const M3<double> a = mul(ld(mu, s), ld(nu, spm));  // U_mu(x) U_nu(x+mu)
            const M3<double> b = mul(ld(nu, s), ld(mu, spn));  // U_nu(x) U_mu(x+nu)
            v += retr_nd(a, b);
        }
    }
    block_sum_store<kBlock>(v, partial);
}

// K1 for workgroups of whole x-rows (NX divides 64, V a multiple of 64): each wave stores its
// U_w(x) in LDS, so the 13 distinct links the six planes of a site need are loaded once per
// workgroup (4 own links, U_0(x+nu) of the planes (0, nu); the links at x+0 from the x-row)
// instead of 4 per plane (24).  Waves 0..2 take the planes (0, w+1) and (1,2) / (1,3) / (2,3):
// three loaded links each; wave 3 loads only its own link.  Same products in FP64 as k_plaquette.
// 32^4 FP64: 0.190 -> 0.177 ms per plaquette with the host readback.
template <typename T>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_plaquette_xrow(
    Geo g, const T* __restrict__ U, double* __restrict__ partial) {
    __shared__ T2<T> ol[4][6][64];
    const SiteMu sm = site_mu_index(g);
    const int s = sm.s, w = sm.mu;  // V % 64 == 0: every lane is a site
    const int lane = (int)(threadIdx.x & 63u);
    int x[4];
    decode(g, s, x);
    const unsigned V = g.V;
    auto lds_link = [&](int d, int l) {
        R6<T> r;
#pragma unroll
        for (int e = 0; e < 6; e++) r.v[e] = ol[d][e][l];
        return full_of(r6_f64(r));
    };
    {
        const R6<T> o = ldr(U, V, w, (unsigned)s);
#pragma unroll
        for (int e = 0; e < 6; e++) ol[w][e][lane] = o.v[e];
    }
    // the wave's memory links, issued before the barrier: plane (0, n1) needs U_0(x+n1); plane
    // (m2, n2) needs U_n2(x+m2) and U_m2(x+n2)
    const int n1 = w + 1;
    const int m2 = w == 2 ? 2 : 1, n2 = w == 0 ? 2 : 3;
    R6<T> l1;
    if (w < 3) l1 = ldr(U, V, 0, (unsigned)(s + dlt(g, x, n1, 1)));
    __syncthreads();  // every wave's U(x) is in LDS
    double v = 0.0;
    if (w < 3) {
        const int lxp = xlane(lane, g.n[0] - 1, 1);
        // the second plane's links, in flight during the first plane's products
        const R6<T> l2 = ldr(U, V, n2, (unsigned)(s + dlt(g, x, m2, 1)));
        const R6<T> l3 = ldr(U, V, m2, (unsigned)(s + dlt(g, x, n2, 1)));
        // Re tr U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag = Re tr (a b^dag)
        v += retr_nd(mul(lds_link(0, lane), lds_link(n1, lxp)), mul(lds_link(n1, lane), full_of(r6_f64(l1))));
        asm volatile("" : "+v"(v));  // the first plane before the second one's LDS reads
        v += retr_nd(mul(lds_link(m2, lane), full_of(r6_f64(l2))), mul(lds_link(n2, lane), full_of(r6_f64(l3))));
    }
    block_sum_store<kBlock>(v, partial);
}

// untraced plaquette action (evaluate_GaugeAction_untraced!, GaugeActions.jl:190-213, for the loop
// set plaquette u plaquette^dag): coeff sum_{mu<nu} (P_munu(x) + P_munu(x)^dag) per site, written
// in the host layout (3x3 complex, column-major, site-major: element (i, j) of site s at
// ref[18 s + 2 (i + 3 j)]).  Products in FP64 for both storage precisions.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_action_untraced(Geo g, const T* __restrict__ U, double coeff,
                                                            double* __restrict__ ref) {
    const int s = site_index();
    if (s >= g.V) return;
    int x[4];
    decode(g, s, x);
    const unsigned V = g.V;
    auto ld = [&](int d, int t) { return full_of(r6_f64(ldr(U, V, d, (unsigned)t))); };
    M3<double> m = zero3<double>();
#pragma unroll 1
    for (int pl = 0; pl < 6; pl++) {
        const int mu = pl < 3 ? 0 : (pl < 5 ? 1 : 2);
        const int nu = pl < 3 ? pl + 1 : (pl < 5 ? pl - 1 : 3);
        const M3<double> a = mul(ld(mu, s), ld(nu, s + dlt(g, x, mu, 1)));   // U_mu(x) U_nu(x+mu)
        const M3<double> b = mul(ld(nu, s), ld(mu, s + dlt(g, x, nu, 1)));   // U_nu(x) U_mu(x+nu)
        const M3<double> P = mul_nd(a, b);                                   // a b^dag
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                m.m[3 * i + j].r += P.m[3 * i + j].r + P.m[3 * j + i].r;
                m.m[3 * i + j].i += P.m[3 * i + j].i - P.m[3 * j + i].i;
            }
    }
    double* r = ref + (size_t)s * 18;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int c = i + 3 * j;
            r[2 * c] = coeff * m.m[3 * i + j].r;
            r[2 * c + 1] = coeff * m.m[3 * i + j].i;
        }
}

// sum of p_a^2 over all links (p*p, TA_gaugefields_4D_serial.jl:106-129)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_momenta_dot(Geo g, const T* __restrict__ P, double* __restrict__ partial) {
    const SiteMu sm = site_mu_index(g);
    double v = 0.0;
    if (sm.s < g.V) {
        T c[8];
        ldp(P, (unsigned)g.V, sm.mu, (unsigned)sm.s, c);
#pragma unroll
        for (int a = 0; a < 8; a++) v = fma((double)c[a], (double)c[a], v);
    }
    block_sum_store<kBlock>(v, partial);
}

// deterministic final sum of per-block partials (one block)
__global__ __launch_bounds__(1024) void k_sum_partials(const double* __restrict__ partial, int n,
                                                       double* __restrict__ out) {
    __shared__ double red[1024];
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) v += partial[i];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// K2: dSdU_mu = coeff * Sigma_mu
// SHARED: upper-staple products shared between directions through LDS (V % 64 == 0)
template <typename T, bool SHARED>
__global__ __launch_bounds__(kBlock) void k_dsdu(Geo g, const T* __restrict__ U, double coeff, T* __restrict__ D) {
    __shared__ T2<T> xs[SHARED ? 6 : 1][6][64];
    const SiteMu sm = site_mu_index(g);
    int s;
    if (!stencil_site<SHARED>(g, sm, s)) return;
    int x[4];
    decode(g, s, x);
    M3<T> sig;
    if constexpr (SHARED) {
        R6<T> own;
        sig = staple_sigma_shared(g, U, (unsigned)g.V, s, x, sm.mu, own, xs);
    } else {
        sig = staple_sigma(g, U, (unsigned)g.V, s, x, sm.mu);
    }
#pragma unroll
    for (int k = 0; k < 9; k++) {
        sig.m[k].r = scale1(coeff, sig.m[k].r);
        sig.m[k].i = scale1(coeff, sig.m[k].i);
    }
    stm<T, GF_ST_NT>(D, (unsigned)g.V, sm.mu, (unsigned)s, sig);
}

// K3 (+K4): F_mu = -(coeff/3) TA(U_mu Sigma_mu)   (md_force!, molecular_dynamics.jl:251-267)
//   FORCE_STORE    : FP = F
//   FORCE_KICK     : FP += eps_p F                    (update_momenta!, :539-551)
//   FORCE_KICK_EXP : FP += eps_p F; Unew = exp(eps_u FP) U   (then update_gaugefields!, :513-531)
//   FORCE_KICK_S64, FORCE_KICK_EXP_S64 (float only: FP32 forces over an FP64 MD state): the same
//     on the FP64 momenta (FP points to them) and FP64 links U64 (updated in place: a thread reads
//     only its own U64 link); the float force enters as F = float(fac c), as if stored, and Unew
//     receives the updated links rounded to float
// ST (stencil_form): ST_OWN per-thread staples; ST_SHARED upper-staple plaquettes shared between
// directions through LDS (ta_force_shared, V % 64 == 0); ST_XROW the same with the x-neighbour
// links from LDS as well (ta_force_xrow, NX divides 64).  At most 168 VGPRs: three waves per SIMD.
template <typename T, int MODE, int ST>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_force(Geo g, const T* __restrict__ U, T* __restrict__ FP,
                                                  T* __restrict__ Unew, double fac, double eps_u,
                                                  double eps_p, double* __restrict__ U64) {
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
        ta_force_xrow(g, U, V, s, x, mu, own, ol, tp, c);
    } else if constexpr (ST == ST_SHARED) {
        ta_force_shared(g, U, V, s, x, mu, own, ol, tp, c);
    } else {
        const M3<T> sig = staple_sigma_own(g, U, V, s, x, mu, own);
        ta_coeffs(mul(full_of(own), sig), c);
    }
    const M3<T> umu = full_of(own);
    if constexpr (MODE == FORCE_KICK_S64 || MODE == FORCE_KICK_EXP_S64) {
        double* P64 = reinterpret_cast<double*>(FP);
        double p[8];
        ldp<double, GF_LD_NT>(P64, V, mu, s, p);
#pragma unroll
        for (int a = 0; a < 8; a++) p[a] = fma(eps_p, (double)scale1(fac, c[a]), p[a]);
        stp<double, GF_ST_NT>(P64, V, mu, s, p);
        if constexpr (MODE == FORCE_KICK_EXP_S64) {
            const M3<double> un = exp_coeffs_times(p, eps_u, ldu<double, GF_LD_NT>(U64, V, mu, s));
            stu<double, GF_ST_NT>(U64, V, mu, s, un);
            stu<T, GF_ST_NT>(Unew, V, mu, s, round_link<T>(un));
        }
    } else if (MODE == FORCE_STORE) {
#pragma unroll
        for (int a = 0; a < 8; a++) c[a] = scale1(fac, c[a]);
        stp<T, GF_ST_NT>(FP, V, mu, s, c);
    } else {
        T p[8];
        ldp<T, GF_LD_NT>(FP, V, mu, s, p);
#pragma unroll
        for (int a = 0; a < 8; a++) p[a] = fma1(fac, c[a], p[a]);
        stp<T, GF_ST_NT>(FP, V, mu, s, p);
        if (MODE == FORCE_KICK_EXP) {
            const M3<T> un = exp_coeffs_times(p, eps_u, umu);
            stu<T, GF_ST_NT>(Unew, V, mu, s, un);
        }
    }
}

// K4: U_mu <- exp(eps P_mu) U_mu  (exptU! + mul!, molecular_dynamics.jl:513-531)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_exp_update(Geo g, const T* __restrict__ Uin, T* __restrict__ Uout,
                                                       const T* __restrict__ P, double eps) {
    // purely streaming: one thread per site walking its four links keeps more loads in flight
    const int s = site_index();
    if (s >= g.V) return;
    const unsigned V = g.V;
#pragma unroll 1
    for (int mu = 0; mu < 4; mu++) {
        T p[8];
        ldp<T, GF_LD_NT>(P, V, mu, s, p);
        const M3<T> un = exp_coeffs_times(p, eps, ldu<T, GF_LD_NT>(Uin, V, mu, s));
        stu<T, GF_ST_NT>(Uout, V, mu, s, un);
    }
}


// FP32-bulk link field -> FP64 link field (the two stored rows, exactly; row 2 is rebuilt on
// every load in either precision)
__global__ __launch_bounds__(kBlock) void k_links_widen(const float* __restrict__ Uf, double* __restrict__ Ud,
                                                        int V) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= V) return;
#pragma unroll 1
    for (int mu = 0; mu < 4; mu++) {
        const R6<float> r = ldr(Uf, (unsigned)V, mu, (unsigned)s);
        M3<double> m;
#pragma unroll
        for (int k = 0; k < 6; k++) {
            m.m[k].r = (double)r.v[k].x;
            m.m[k].i = (double)r.v[k].y;
        }
        stu(Ud, (unsigned)V, mu, (unsigned)s, m);
    }
}

// FP32 forces over an FP64 MD state: the links U64 and momenta P64 are held and updated in double,
// the force kernels read U32, the links rounded to float (their FP64-state epilogues kick P64 and
// update U64 in the same launch).  The lone steps: MODE 1: U64 <- exp(eps_u P64) U64 (in place:
// only the thread's own link is read) and U32 <- round(U64).  MODE 3: U32 <- round(U64).
// One thread per site walking its four links: a streaming kernel.
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_state64_step(Geo g, const double* __restrict__ P64,
                                                         double* __restrict__ U64, float* __restrict__ U32,
                                                         double eps_u) {
    const int s = site_index();
    if (s >= g.V) return;
    const unsigned V = g.V;
#pragma unroll 1
    for (int mu = 0; mu < 4; mu++) {
        if (MODE == 3) {
            stu<float, GF_ST_NT>(U32, V, mu, s, round_link<float>(ldu<double, GF_LD_NT>(U64, V, mu, s)));
            continue;
        }
        double p[8];
        ldp<double, GF_LD_NT>(P64, V, mu, s, p);
        const M3<double> un = exp_coeffs_times(p, eps_u, ldu<double, GF_LD_NT>(U64, V, mu, s));
        stu<double, GF_ST_NT>(U64, V, mu, s, un);
        stu<float, GF_ST_NT>(U32, V, mu, s, round_link<float>(un));
    }
}

hipError_t launch_state64_step(int mode, const Geo& g, const double* P64, double* U64, float* U32, double eps_u,
                               hipStream_t st) {
    const int nb = grid_for(g.V);
    if (mode == 1)
        k_state64_step<1><<<nb, kBlock, 0, st>>>(g, P64, U64, U32, eps_u);
    else if (mode == 3)
        k_state64_step<3><<<nb, kBlock, 0, st>>>(g, P64, U64, U32, eps_u);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// layout conversion between the reference's (3,3,X,Y,Z,T) complex column-major arrays
// (FP64, per direction) and the SoA planes
// ---------------------------------------------------------------------------
// GEN = false: SU(3) link field (rows 0 and 1 stored); GEN = true: general 3x3 matrix field
template <typename T, bool GEN>
__global__ __launch_bounds__(kBlock) void k_links_from_ref(const double* __restrict__ ref, T* __restrict__ soa, int V) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= V) return;
    for (int mu = 0; mu < 4; mu++) {
        const double* r = ref + ((size_t)mu * V + s) * 18;
        M3<T> m;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int c = i + 3 * j;  // column-major complex index of the reference
                m.m[3 * i + j].r = (T)r[2 * c];
                m.m[3 * i + j].i = (T)r[2 * c + 1];
            }
        if (GEN)
            stm(soa, (unsigned)V, mu, (unsigned)s, m);
        else
            stu(soa, (unsigned)V, mu, (unsigned)s, m);
    }
}
template <typename T, bool GEN>
__global__ __launch_bounds__(kBlock) void k_links_to_ref(const T* __restrict__ soa, double* __restrict__ ref, int V) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= V) return;
    for (int mu = 0; mu < 4; mu++) {
        double* r = ref + ((size_t)mu * V + s) * 18;
        // SU(3) links: the stored rows widened, row 2 rebuilt in double (as the FP64 sums use them)
        const M3<double> m = GEN ? to_f64(ldm(soa, (unsigned)V, mu, (unsigned)s))
                                 : full_of(r6_f64(ldr(soa, (unsigned)V, mu, (unsigned)s)));
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int c = i + 3 * j;
                r[2 * c] = (double)m.m[3 * i + j].r;
                r[2 * c + 1] = (double)m.m[3 * i + j].i;
            }
    }
}
template <typename T>
__global__ __launch_bounds__(kBlock) void k_momenta_from_ref(const double* __restrict__ ref, T* __restrict__ soa, int V) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= V) return;
    for (int mu = 0; mu < 4; mu++) {
        T c[8];
#pragma unroll
        for (int a = 0; a < 8; a++) c[a] = (T)ref[((size_t)mu * V + s) * 8 + a];
        stp(soa, (unsigned)V, mu, (unsigned)s, c);
    }
}
template <typename T>
__global__ __launch_bounds__(kBlock) void k_momenta_to_ref(const T* __restrict__ soa, double* __restrict__ ref, int V) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= V) return;
    for (int mu = 0; mu < 4; mu++) {
        T c[8];
        ldp(soa, (unsigned)V, mu, (unsigned)s, c);
#pragma unroll
        for (int a = 0; a < 8; a++) ref[((size_t)mu * V + s) * 8 + a] = (double)c[a];
    }
}
template <typename T>
__global__ __launch_bounds__(kBlock) void k_set_identity(T* __restrict__ U, int V) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= V) return;
    for (int mu = 0; mu < 4; mu++) stu(U, (unsigned)V, mu, (unsigned)s, ident3<T>());
}
template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill(T* __restrict__ p, size_t n, T v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
template <typename T>
__global__ __launch_bounds__(kBlock) void k_axpby(T* __restrict__ y, const T* __restrict__ x, T a, T b, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i] + b * y[i];
}
// The FP32-bulk storage rounding applied to FP64 fields (precision study): rows 0 and 1 of every
// link, after the double-precision SU(3) projection the FP32 link updates apply when PROJECT,
// rounded to the nearest float in place; momenta: every coefficient.
template <bool PROJECT>
__global__ __launch_bounds__(kBlock) void k_round_links_f32(double* __restrict__ U, int V) {
    const int s = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (s >= V) return;
    for (int mu = 0; mu < 4; mu++) {
        M3<double> m = ldu(U, (unsigned)V, mu, (unsigned)s);
        if (PROJECT) reunitarize_f64(m);
#pragma unroll
        for (int k = 0; k < 6; k++) {
            m.m[k].r = (double)(float)m.m[k].r;
            m.m[k].i = (double)(float)m.m[k].i;
        }
        stu(U, (unsigned)V, mu, (unsigned)s, m);
    }
}
__global__ __launch_bounds__(kBlock) void k_round_f32(double* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = (double)(float)p[i];
}

// Streaming copy of n 16-byte words (the measured HBM ceiling): each thread moves four words
// per pass of a grid-stride loop, loads and stores non-temporal.
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(kBlock) void k_stream_copy(const u32x4* __restrict__ src, u32x4* __restrict__ dst,
                                                        size_t n) {
    const size_t stride = (size_t)gridDim.x * kBlock * 4;
    for (size_t i = (size_t)blockIdx.x * kBlock * 4 + threadIdx.x; i < n; i += stride) {
        u32x4 v[4];
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (i + (size_t)k * kBlock < n) v[k] = __builtin_nontemporal_load(src + i + (size_t)k * kBlock);
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (i + (size_t)k * kBlock < n) __builtin_nontemporal_store(v[k], dst + i + (size_t)k * kBlock);
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_stream_copy(const void* src, void* dst, size_t bytes, hipStream_t st) {
    const size_t n = bytes / 16;
    if (n == 0) return hipSuccess;
    const size_t per_block = (size_t)kBlock * 4;
    const unsigned nb = (unsigned)std::min<size_t>((n + per_block - 1) / per_block, 256 * 16);
    k_stream_copy<<<nb, kBlock, 0, st>>>((const u32x4*)src, (u32x4*)dst, n);
    return hipGetLastError();
}

hipError_t launch_round_f32(double* U, double* P, int V, bool links, bool project, bool momenta, hipStream_t st) {
    if (links) {
        if (project)
            k_round_links_f32<true><<<grid_for(V), kBlock, 0, st>>>(U, V);
        else
            k_round_links_f32<false><<<grid_for(V), kBlock, 0, st>>>(U, V);
    }
    if (momenta) {
        const size_t n = (size_t)32 * V;
        k_round_f32<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, st>>>(P, n);
    }
    return hipGetLastError();
}

hipError_t launch_sum_partials(const double* partial, int n, double* out, hipStream_t st) {
    k_sum_partials<<<1, 1024, 0, st>>>(partial, n, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_plaquette(const Geo& g, const T* U, double* partial, double* out, hipStream_t st) {
    const int nb = grid_sites_mu(g.V);
    if (sizeof(T) == 8 && stencil_form(g) == ST_XROW)  // FP32: 16 B spilled at 168 VGPRs, slower
        k_plaquette_xrow<T><<<nb, kBlock, 0, st>>>(g, U, partial);
    else
        k_plaquette<T><<<nb, kBlock, 0, st>>>(g, U, partial);
    k_sum_partials<<<1, 1024, 0, st>>>(partial, nb, out);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_action_untraced(const Geo& g, const T* U, double coeff, double* ref, hipStream_t st) {
    k_action_untraced<T><<<grid_for(g.V), kBlock, 0, st>>>(g, U, coeff, ref);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_momenta_dot(const Geo& g, const T* P, double* partial, double* out, hipStream_t st) {
    const int nb = grid_sites_mu(g.V);
    k_momenta_dot<T><<<nb, kBlock, 0, st>>>(g, P, partial);
    k_sum_partials<<<1, 1024, 0, st>>>(partial, nb, out);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_dsdu(const Geo& g, const T* U, double coeff, T* D, hipStream_t st) {
    if (g.V % 64 == 0)
        k_dsdu<T, true><<<grid_sites_mu(g.V), kBlock, 0, st>>>(g, U, coeff, D);
    else
        k_dsdu<T, false><<<grid_sites_mu(g.V), kBlock, 0, st>>>(g, U, coeff, D);
    return hipGetLastError();
}
template <typename T, int MODE>
void launch_force_st(int form, int nb, hipStream_t st, const Geo& g, const T* U, T* FP, T* Unew, double fac,
                     double eps_u, double eps_p = 0.0, double* U64 = nullptr) {
    if (form == ST_XROW)
        k_force<T, MODE, ST_XROW><<<nb, kBlock, 0, st>>>(g, U, FP, Unew, fac, eps_u, eps_p, U64);
    else if (form == ST_SHARED)
        k_force<T, MODE, ST_SHARED><<<nb, kBlock, 0, st>>>(g, U, FP, Unew, fac, eps_u, eps_p, U64);
    else
        k_force<T, MODE, ST_OWN><<<nb, kBlock, 0, st>>>(g, U, FP, Unew, fac, eps_u, eps_p, U64);
}
template <typename T>
hipError_t launch_force(int mode, const Geo& g, const T* U, T* FP, T* Unew, double coeff, double eps_p,
                        double eps_u, hipStream_t st, double* U64) {
    const int nb = grid_sites_mu(g.V);
    const double fac = -coeff / 3.0;
    const int form = stencil_form(g);
    switch (mode) {
        case FORCE_STORE:
            launch_force_st<T, FORCE_STORE>(form, nb, st, g, U, FP, Unew, fac, eps_u);
            break;
        case FORCE_KICK:
            launch_force_st<T, FORCE_KICK>(form, nb, st, g, U, FP, Unew, fac * eps_p, eps_u);
            break;
        case FORCE_KICK_EXP:
            launch_force_st<T, FORCE_KICK_EXP>(form, nb, st, g, U, FP, Unew, fac * eps_p, eps_u);
            break;
        case FORCE_KICK_S64:
        case FORCE_KICK_EXP_S64:
            if constexpr (sizeof(T) == 4) {
                if (mode == FORCE_KICK_S64)
                    launch_force_st<T, FORCE_KICK_S64>(form, nb, st, g, U, FP, Unew, fac, 0.0, eps_p, U64);
                else
                    launch_force_st<T, FORCE_KICK_EXP_S64>(form, nb, st, g, U, FP, Unew, fac, eps_u, eps_p, U64);
                break;
            } else {
                return hipErrorInvalidValue;
            }
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
template <typename T>
hipError_t launch_exp_update(const Geo& g, const T* Uin, T* Uout, const T* P, double eps, hipStream_t st) {
    k_exp_update<T><<<grid_for(g.V), kBlock, 0, st>>>(g, Uin, Uout, P, eps);
    return hipGetLastError();
}
hipError_t launch_links_widen(const float* Uf, double* Ud, int V, hipStream_t st) {
    k_links_widen<<<grid_for(V), kBlock, 0, st>>>(Uf, Ud, V);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_links_from_ref(const double* ref, T* soa, int V, bool general, hipStream_t st) {
    if (general)
        k_links_from_ref<T, true><<<grid_for(V), kBlock, 0, st>>>(ref, soa, V);
    else
        k_links_from_ref<T, false><<<grid_for(V), kBlock, 0, st>>>(ref, soa, V);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_links_to_ref(const T* soa, double* ref, int V, bool general, hipStream_t st) {
    if (general)
        k_links_to_ref<T, true><<<grid_for(V), kBlock, 0, st>>>(soa, ref, V);
    else
        k_links_to_ref<T, false><<<grid_for(V), kBlock, 0, st>>>(soa, ref, V);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_momenta_from_ref(const double* ref, T* soa, int V, hipStream_t st) {
    k_momenta_from_ref<T><<<grid_for(V), kBlock, 0, st>>>(ref, soa, V);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_momenta_to_ref(const T* soa, double* ref, int V, hipStream_t st) {
    k_momenta_to_ref<T><<<grid_for(V), kBlock, 0, st>>>(soa, ref, V);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_set_identity(T* U, int V, hipStream_t st) {
    k_set_identity<T><<<grid_for(V), kBlock, 0, st>>>(U, V);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_fill(T* p, size_t n, T v, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_fill<T><<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, st>>>(p, n, v);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_axpby(T* y, const T* x, double a, double b, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_axpby<T><<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, st>>>(y, x, (T)a, (T)b, n);
    return hipGetLastError();
}

#define GF_INST(T)                                                                                       \
    template hipError_t launch_plaquette<T>(const Geo&, const T*, double*, double*, hipStream_t);        \
    template hipError_t launch_action_untraced<T>(const Geo&, const T*, double, double*, hipStream_t);  \
    template hipError_t launch_momenta_dot<T>(const Geo&, const T*, double*, double*, hipStream_t);      \
    template hipError_t launch_dsdu<T>(const Geo&, const T*, double, T*, hipStream_t);                   \
    template hipError_t launch_force<T>(int, const Geo&, const T*, T*, T*, double, double, double,       \
                                        hipStream_t, double*);                                           \
    template hipError_t launch_exp_update<T>(const Geo&, const T*, T*, const T*, double, hipStream_t);   \
    template hipError_t launch_links_from_ref<T>(const double*, T*, int, bool, hipStream_t);             \
    template hipError_t launch_links_to_ref<T>(const T*, double*, int, bool, hipStream_t);               \
    template hipError_t launch_momenta_from_ref<T>(const double*, T*, int, hipStream_t);                 \
    template hipError_t launch_momenta_to_ref<T>(const T*, double*, int, hipStream_t);                   \
    template hipError_t launch_set_identity<T>(T*, int, hipStream_t);                                    \
    template hipError_t launch_fill<T>(T*, size_t, T, hipStream_t);                                      \
    template hipError_t launch_axpby<T>(T*, const T*, double, double, size_t, hipStream_t);

GF_INST(double)
GF_INST(float)

}  // namespace gf
