// k_random.hip -- random fields generated on the device from a counter-based generator, so a
// momentum refresh or a hot start never crosses PCIe and does not depend on the launch shape.
//
// Generator: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Stream of one link (mu, s):
// key = seed, counter = (s, tag << 24 | mu << 16 | block, sweep).
//   tag 1, Gaussian momenta (gaussian_momenta!, API.jl:341-392): block b = 0..3 gives one
//     Box-Muller pair (p_2b, p_2b+1) from two 53-bit uniforms;
//   tag 2, hot start (gaugefields_4D_wing.jl:295-322 + normalize3!, AbstractGaugefields.jl:
//     3378-3457): blocks 0..4 give 18 uniforms u in (0, 1); element (i, j) of the start matrix
//     is (u_2k - 1/2) + i (u_2k+1 - 1/2), k = 3 i + j; rows 0 and 1 are Gram-Schmidt normalized
//     (row 2 is implied by the link storage).
// All arithmetic on the uniforms runs in double for both storage precisions.
#include "internal.hpp"

namespace gf {

namespace {

__device__ __forceinline__ void philox4x32_10(unsigned c[4], unsigned k0, unsigned k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const unsigned hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
        const unsigned hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
        const unsigned n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__device__ __forceinline__ void link_block(unsigned long long seed, unsigned long long sweep, unsigned s, int mu,
                                           int tag, int block, unsigned w[4]) {
    w[0] = s;
    w[1] = ((unsigned)tag << 24) | ((unsigned)mu << 16) | (unsigned)block;
    w[2] = (unsigned)sweep;
    w[3] = (unsigned)(sweep >> 32);
    philox4x32_10(w, (unsigned)seed, (unsigned)(seed >> 32));
}

__device__ __forceinline__ double unit53(unsigned hi, unsigned lo) {
    return ((double)(hi >> 5) * 67108864.0 + (double)(lo >> 6) + 0.5) * (1.0 / 9007199254740992.0);
}
__device__ __forceinline__ double unit32(unsigned w) { return ((double)w + 0.5) * (1.0 / 4294967296.0); }

}  // namespace

template <typename T>
__global__ __launch_bounds__(kBlock) void k_gaussian_momenta(Geo g, T* __restrict__ P, double sigma,
                                                             unsigned long long seed, unsigned long long sweep) {
    const SiteMu sm = site_mu_index(g);
    if (sm.s >= g.V) return;
    T p[8];
#pragma unroll
    for (int b = 0; b < 4; b++) {
        unsigned w[4];
        link_block(seed, sweep, (unsigned)sm.s, sm.mu, 1, b, w);
        const double r = sqrt(-2.0 * log(unit53(w[0], w[1])));
        double sn, cs;
        sincospi(2.0 * unit53(w[2], w[3]), &sn, &cs);
        p[2 * b] = (T)(sigma * r * cs);
        p[2 * b + 1] = (T)(sigma * r * sn);
    }
    stp<T, GF_ST_NT>(P, (unsigned)g.V, sm.mu, (unsigned)sm.s, p);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_hot_links(Geo g, T* __restrict__ U, unsigned long long seed) {
    const SiteMu sm = site_mu_index(g);
    if (sm.s >= g.V) return;
    double u[20];
#pragma unroll
    for (int b = 0; b < 5; b++) {
        unsigned w[4];
        link_block(seed, 0ull, (unsigned)sm.s, sm.mu, 2, b, w);
#pragma unroll
        for (int k = 0; k < 4; k++) u[4 * b + k] = unit32(w[k]);
    }
    M3<double> m;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        m.m[k].r = u[2 * k] - 0.5;
        m.m[k].i = u[2 * k + 1] - 0.5;
    }
    reunitarize_f64(m);
    M3<T> mt;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        mt.m[k].r = (T)m.m[k].r;
        mt.m[k].i = (T)m.m[k].i;
    }
    stu<T, GF_ST_NT>(U, (unsigned)g.V, sm.mu, (unsigned)sm.s, mt);
}

template <typename T>
hipError_t launch_gaussian_momenta(const Geo& g, T* P, double sigma, unsigned long long seed,
                                   unsigned long long sweep, hipStream_t st) {
    k_gaussian_momenta<T><<<grid_sites_mu(g.V), kBlock, 0, st>>>(g, P, sigma, seed, sweep);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_hot_links(const Geo& g, T* U, unsigned long long seed, hipStream_t st) {
    k_hot_links<T><<<grid_sites_mu(g.V), kBlock, 0, st>>>(g, U, seed);
    return hipGetLastError();
}

#define GF_INST_RNG(T)                                                                                     \
    template hipError_t launch_gaussian_momenta<T>(const Geo&, T*, double, unsigned long long,             \
                                                   unsigned long long, hipStream_t);                       \
    template hipError_t launch_hot_links<T>(const Geo&, T*, unsigned long long, hipStream_t);
GF_INST_RNG(double)
GF_INST_RNG(float)
#undef GF_INST_RNG

}  // namespace gf

// ===========================================================================
// Heatbath and overrelaxation (Cabibbo-Marinari over the SU(2) subgroups (1,2), (2,3), (1,3)).
// One thread per site of one checkerboard colour for one direction mu: the staple of U_mu(x)
// only touches U_mu at the other colour and the other directions, so the colour updates in
// place.  Streams (same counter layout as above): tags 3, 4, 5 = heatbath subgroups (block =
// try index, 0xFFFF = the two draws after acceptance), tag 6 = overrelaxation subgroup choice.
// Everything runs in double; FP32 links are rounded once when stored.
// ===========================================================================
namespace gf {

namespace {

struct C2 {
    double r, i;
};
__device__ __forceinline__ C2 c2(double r, double i) { return C2{r, i}; }
__device__ __forceinline__ C2 cmul2(C2 a, C2 b) { return C2{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ C2 cadd2(C2 a, C2 b) { return C2{a.r + b.r, a.i + b.i}; }
struct M2 {
    C2 e[2][2];
};

template <int n, int k>
__device__ __forceinline__ M2 sub2(const M3<double>& m) {
    M2 S;
    S.e[0][0] = c2(m.m[3 * n + n].r, m.m[3 * n + n].i);
    S.e[0][1] = c2(m.m[3 * n + k].r, m.m[3 * n + k].i);
    S.e[1][0] = c2(m.m[3 * k + n].r, m.m[3 * k + n].i);
    S.e[1][1] = c2(m.m[3 * k + k].r, m.m[3 * k + k].i);
    return S;
}

// u <- embed(K at rows/columns n, k) u
template <int n, int k>
__device__ __forceinline__ void su2_left_mul(const M2& K, M3<double>& u) {
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const C2 a = c2(u.m[3 * n + j].r, u.m[3 * n + j].i), b = c2(u.m[3 * k + j].r, u.m[3 * k + j].i);
        const C2 na = cadd2(cmul2(K.e[0][0], a), cmul2(K.e[0][1], b));
        const C2 nb = cadd2(cmul2(K.e[1][0], a), cmul2(K.e[1][1], b));
        u.m[3 * n + j].r = na.r;
        u.m[3 * n + j].i = na.i;
        u.m[3 * k + j].r = nb.r;
        u.m[3 * k + j].i = nb.i;
    }
}

// SU(2) elements (times a real scale) as quaternions: q = (q0, q1, q2, q3) stands for
// X(q) = q0 + i (q1 s1 + q2 s2 + q3 s3) with the Pauli matrices s_k, i.e.
// X = [[q0 + i q3, q2 + i q1], [-q2 + i q1, q0 - i q3]];  det X = |q|^2, X(q)^dag = X(conj q),
// X(p) X(q) = X(p q) with the Hamilton product below.
struct Quat {
    double q0, q1, q2, q3;
};
__device__ __forceinline__ Quat qmul(const Quat& p, const Quat& q) {
    // (p0 q0 - p.q,  p0 q + q0 p - p x q)
    return Quat{p.q0 * q.q0 - (p.q1 * q.q1 + p.q2 * q.q2 + p.q3 * q.q3),
                p.q0 * q.q1 + q.q0 * p.q1 - (p.q2 * q.q3 - p.q3 * q.q2),
                p.q0 * q.q2 + q.q0 * p.q2 - (p.q3 * q.q1 - p.q1 * q.q3),
                p.q0 * q.q3 + q.q0 * p.q3 - (p.q1 * q.q2 - p.q2 * q.q1)};
}
__device__ __forceinline__ Quat qconj(const Quat& q) { return Quat{q.q0, -q.q1, -q.q2, -q.q3}; }
__device__ __forceinline__ double qnorm2(const Quat& q) { return q.q0 * q.q0 + q.q1 * q.q1 + q.q2 * q.q2 + q.q3 * q.q3; }
__device__ __forceinline__ Quat qscale(double s, const Quat& q) { return Quat{s * q.q0, s * q.q1, s * q.q2, s * q.q3}; }
// the quaternion part of a general 2x2 complex block: its projection onto the real span of SU(2)
// (project_onto_SU2!, heatbath/portable/kernels.jl:12-23, forms the same element as a matrix)
__device__ __forceinline__ Quat quat_of(const M2& S) {
    return Quat{0.5 * (S.e[0][0].r + S.e[1][1].r), 0.5 * (S.e[0][1].i + S.e[1][0].i),
                0.5 * (S.e[0][1].r - S.e[1][0].r), 0.5 * (S.e[0][0].i - S.e[1][1].i)};
}
__device__ __forceinline__ M2 mat_of(const Quat& q) {
    M2 m;
    m.e[0][0] = c2(q.q0, q.q3);
    m.e[0][1] = c2(q.q2, q.q1);
    m.e[1][0] = c2(-q.q2, q.q1);
    m.e[1][1] = c2(q.q0, -q.q3);
    return m;
}
// q / |q|: keeps a product of unit quaternions on SU(2) against rounding
__device__ __forceinline__ Quat qunit(const Quat& q) { return qscale(1.0 / sqrt(qnorm2(q)), q); }

// SU(2) heatbath of one Cabibbo-Marinari subgroup (N. Cabibbo, E. Marinari, Phys. Lett. B 119
// (1982) 387), the sampler of A. D. Kennedy, B. J. Pendleton, Phys. Lett. B 156 (1985) 393.
// The action restricted to the subgroup element X is (beta/3) Re tr(X W) with W the subgroup
// block of u Sigma, whose quaternion part is w = |w| v, v in SU(2).  X = X0 v^dag with X0
// drawn from the density sqrt(1 - x0^2) exp(k x0) dx0 dOmega, k = 2 (beta/3) |w|:
// delta = 1 - x0 is Gamma(3/2, 1/k) distributed, sampled as -ln(r1)/k - ln(r2)/k cos^2(2 pi r3)
// and accepted with probability sqrt(1 - delta/2) (test r4^2 <= 1 - delta/2); the direction of
// the vector part is uniform on the sphere.  (The reference's steps are the same published
// algorithm, heatbath/portable/kernels.jl:63-150; they use this stream's draws in this order.)
// Returns false (no subgroup update) if no proposal is accepted within itmax tries.
__device__ __forceinline__ bool kp_update(unsigned long long seed, unsigned long long sweep, unsigned s, int mu,
                                          int tag, double beta, int itmax, const Quat& w, Quat& X) {
    const double wn = sqrt(qnorm2(w));
    const double k = 2.0 * (beta / 3.0) * wn;
    double delta = 0.0;
    bool ok = false;
    for (int t = 0; t < itmax && !ok; t++) {
        unsigned r[4];
        link_block(seed, sweep, s, mu, tag, t, r);
        const double c = cos(2.0 * 3.14159265358979323846 * unit32(r[2]));
        delta = -log(unit32(r[1])) / k + (-log(unit32(r[0])) / k) * c * c;
        const double r4 = unit32(r[3]);
        ok = r4 * r4 <= 1.0 - 0.5 * delta;
    }
    if (!ok) return false;
    unsigned r[4];
    link_block(seed, sweep, s, mu, tag, 0xFFFF, r);
    const double x0 = 1.0 - delta, len = sqrt(1.0 - x0 * x0);
    const double phi = 2.0 * 3.14159265358979323846 * unit32(r[0]);
    const double cth = 2.0 * (unit32(r[1]) - 0.5), sth = sqrt(1.0 - cth * cth);
    const Quat X0{x0, len * sth * cos(phi), len * sth * sin(phi), len * cth};
    X = qunit(qmul(X0, qscale(1.0 / wn, qconj(w))));  // X0 v^dag
    return true;
}

// Overrelaxation of one subgroup (M. Creutz, Phys. Rev. D 36 (1987) 515; F. R. Brown,
// T. J. Woch, Phys. Rev. Lett. 58 (1987) 2394): with v the SU(2) direction of the subgroup block
// of u Sigma, X = (v^dag)^2 maps Re tr(u Sigma) restricted to the subgroup onto itself
// (heatbathmodule.jl:1248-1280 applies the same reflection).
template <int n, int k>
__device__ __forceinline__ void or_subgroup(M3<double>& u, const M3<double>& V) {
    const Quat vd = qunit(qconj(quat_of(sub2<n, k>(mul(u, V)))));
    su2_left_mul<n, k>(mat_of(qunit(qmul(vd, vd))), u);
}

// one heatbath subgroup step; false (u unchanged) if no proposal was accepted
template <int n, int k>
__device__ __forceinline__ bool hb_subgroup(M3<double>& u, const M3<double>& St, unsigned long long seed,
                                            unsigned long long sweep, unsigned s, int mu, int tag, double beta,
                                            int itmax) {
    Quat X;
    if (!kp_update(seed, sweep, s, mu, tag, beta, itmax, quat_of(sub2<n, k>(mul(u, St))), X)) return false;
    su2_left_mul<n, k>(mat_of(X), u);
    return true;
}

// Half-volume index of a heatbath thread.  Half-site h is the pair of sites (2h, 2h+1) (one of
// each colour, NX even), so the 64-site group k holds half-sites 32k .. 32k+31.  With the
// t-sweep tiles of the stencil kernels (V a multiple of 4096: 8 groups per workgroup, a grid
// that is a multiple of 8), lanes 0-31 and 32-63 of wave w of block b take the groups the
// stencil kernels give to blocks 8b + 2w and 8b + 2w + 1: the sites in flight on an XCD then
// sweep its tile slice by slice in t, and the neighbour links a staple re-reads are still in
// that XCD's L2.  Otherwise linear order.
__device__ __forceinline__ int heatbath_half_site(const Geo& g) {
    if (g.tile[0] > 0 && (unsigned)g.V == gridDim.x * 512u && (gridDim.x & 7u) == 0u) {
        const unsigned lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
        const int grp = tsweep_group(g, xcd_block() * 8u + 2u * w + (lane >> 5), gridDim.x * 8u);
        return grp * 32 + (int)(lane & 31u);
    }
    return (int)(xcd_block() * blockDim.x + threadIdx.x);
}

// site of colour `par` with half-volume index h (NX even): x = 2 (h mod NX/2) + ((y+z+t+par) & 1)
__device__ __forceinline__ int colour_site(const Geo& g, int h, int par, int x[4]) {
    const int nxh = g.n[0] >> 1;
    int r = h / nxh;
    x[1] = r % g.n[1];
    r /= g.n[1];
    x[2] = r % g.n[2];
    x[3] = r / g.n[2];
    x[0] = 2 * (h % nxh) + ((x[1] + x[2] + x[3] + par) & 1);
    return x[0] + g.n[0] * (x[1] + g.n[1] * (x[2] + g.n[2] * x[3]));
}

}  // namespace

// MODE 0: heatbath (SU3update_matrix!, heatbath/portable/kernels.jl:199-300; a rejected
// subgroup leaves the link unchanged and counts a failure); MODE 1: overrelaxation
// (SUN_overrelaxation_rng!, heatbathmodule.jl:1295-1321, subgroup i drawn from block i of tag 6:
// n = floor(2 w0 / 2^32), k = n + 1 + floor((2 - n) w1 / 2^32)).
template <typename T, int MODE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? 2 : 3))) void k_heatbath(Geo g, T* __restrict__ U, int mu, int par, double beta,
                                                     unsigned long long seed, unsigned long long sweep, int itmax,
                                                     double* __restrict__ partial) {
    const int h = heatbath_half_site(g);
    double fail = 0.0;
    if (h < (g.V >> 1)) {
        int x[4];
        const int s = colour_site(g, h, par, x);
        const unsigned V = g.V;
        R6<T> own;
        // three staple links in flight ahead of the one in use: the sampler, not the staple, sets
        // the register peak (FP64 at two waves per SIMD, FP32 at three and four); 32^4 heatbath /
        // overrelaxation sweep FP64 1.37 / 1.42 -> 1.33 / 1.31 ms, FP32 0.79 / 0.76 -> 0.76 / 0.67 ms
        M3<double> St = to_f64(staple_sigma_own<T, 3>(g, (const T*)U, V, s, x, mu, own));
        // pin the staple sum: otherwise the compiler interleaves the staple loads with the
        // sampler and needs 256 VGPRs
#pragma unroll
        for (int e = 0; e < 9; e++) asm volatile("" : "+v"(St.m[e].r), "+v"(St.m[e].i));
        __builtin_amdgcn_sched_barrier(0);
        M3<double> u = to_f64(full_of(own));
        bool ok = true;
        if (MODE == 0) {
            ok = hb_subgroup<0, 1>(u, St, seed, sweep, (unsigned)s, mu, 3, beta, itmax) &&
                 hb_subgroup<1, 2>(u, St, seed, sweep, (unsigned)s, mu, 4, beta, itmax) &&
                 hb_subgroup<0, 2>(u, St, seed, sweep, (unsigned)s, mu, 5, beta, itmax);
        } else {
#pragma unroll 1
            for (int i = 0; i < 3; i++) {
                unsigned w[4];
                link_block(seed, sweep, (unsigned)s, mu, 6, i, w);
                const int n = (int)(((unsigned long long)w[0] * 2ull) >> 32);
                const int k = n + 1 + (int)(((unsigned long long)w[1] * (unsigned long long)(2 - n)) >> 32);
                if (n == 1)
                    or_subgroup<1, 2>(u, St);
                else if (k == 1)
                    or_subgroup<0, 1>(u, St);
                else
                    or_subgroup<0, 2>(u, St);
            }
        }
        if (ok) {
            reunitarize_f64(u);
            M3<T> ut;
#pragma unroll
            for (int e = 0; e < 9; e++) {
                ut.m[e].r = (T)u.m[e].r;
                ut.m[e].i = (T)u.m[e].i;
            }
            stu<T>(U, V, mu, (unsigned)s, ut);
        } else {
            fail = 1.0;
        }
    }
    block_sum_store<kBlock>(fail, partial);
}

// Dynamic LDS (unused) of an FP64 heatbath launch: 56 KB per workgroup leaves room for two
// workgroups per CU instead of three.  A heatbath thread owns one site of one colour, so the sites
// in flight per XCD are eight times the stencil kernels'; with a third fewer of them the links their
// staples share stay in L2 between uses (32^4: sweep 1.81 -> 1.52 ms, overrelaxation 1.61 -> 1.42
// ms).  Two waves per SIMD also let the FP64 kernel use up to 256 VGPRs (its amdgpu_waves_per_eu),
// which removes the spills a 168-VGPR budget forced (sweep 1.48 -> 1.37 ms, overrelaxation 1.52 ->
// 1.42 ms, same box).  FP32 links are half the bytes and keep three workgroups.
constexpr size_t kHeatbathOccLds = 56 * 1024;

template <typename T>
hipError_t launch_heatbath(const Geo& g, T* U, int mode, double beta, unsigned long long seed,
                           unsigned long long sweep, int itmax, double* partial, double* failures, hipStream_t st) {
    const int nb = grid_for(g.V / 2);
    const size_t lds = sizeof(T) == 8 ? kHeatbathOccLds : 0;
    for (int mu = 0; mu < 4; mu++)
        for (int par = 0; par < 2; par++) {
            double* part = partial + (size_t)(2 * mu + par) * nb;
            if (mode == 0)
                k_heatbath<T, 0><<<nb, kBlock, lds, st>>>(g, U, mu, par, beta, seed, sweep, itmax, part);
            else
                k_heatbath<T, 1><<<nb, kBlock, lds, st>>>(g, U, mu, par, beta, seed, sweep, itmax, part);
        }
    hipError_t err = hipGetLastError();
    if (err == hipSuccess) err = launch_sum_partials(partial, 8 * nb, failures, st);
    return err;
}

#define GF_INST_HB(T)                                                                                      \
    template hipError_t launch_heatbath<T>(const Geo&, T*, int, double, unsigned long long, unsigned long long, \
                                           int, double*, double*, hipStream_t);
GF_INST_HB(double)
GF_INST_HB(float)
#undef GF_INST_HB

}  // namespace gf
