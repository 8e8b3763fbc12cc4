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
