// lattice.hpp -- site geometry and structure-of-arrays field access for the gfx950 kernels.
//
// HBM layout (one allocation per field, T = double or float):
//   link field     : 4 directions x 18 real planes x V sites; plane (mu*18 + r), r = 2*(3*i + j) + {0:re,1:im}
//   momentum field : 4 directions x  8 real planes x V sites; plane (mu*8 + a), coefficient of i lambda_a/2
// Consecutive sites x, x+1, ... are consecutive in every plane, so one wavefront of 64
// sites reads each plane with fully coalesced 256 B (FP32) / 512 B (FP64) transactions.
// Site order: s = x + NX*(y + NY*(z + NZ*t)), periodic in all four directions.
#pragma once
#include <hip/hip_runtime.h>

#include "su3.hpp"

namespace gf {

struct Geo {
    int n[4];   // NX, NY, NZ, NT
    int st[4];  // site strides 1, NX, NX*NY, NX*NY*NZ
    int V;
};

__device__ __forceinline__ void decode(const Geo& g, int s, int x[4]) {
    int r = s;
    x[0] = r % g.n[0];
    r /= g.n[0];
    x[1] = r % g.n[1];
    r /= g.n[1];
    x[2] = r % g.n[2];
    x[3] = r / g.n[2];
}

// site offset for a step d = +-1 along mu from a site with coordinates x (periodic)
__device__ __forceinline__ int dlt(const Geo& g, const int x[4], int mu, int d) {
    const int xm = x[mu] + d;
    const int wrap = (g.n[mu] - 1) * g.st[mu];
    return xm >= g.n[mu] ? -wrap : (xm < 0 ? wrap : d * g.st[mu]);
}

// Workgroups are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8).  Remap so
// that XCD k walks the contiguous k-th eighth of the lattice: the halo rows a site needs
// (x +- y, z, t neighbours) were then fetched recently by the same XCD and sit in its L2.
__device__ __forceinline__ int site_index() {
    unsigned b = blockIdx.x;
    const unsigned nb = gridDim.x;
    if ((nb & 7u) == 0u) b = (b & 7u) * (nb >> 3) + (b >> 3);
    return (int)(b * blockDim.x + threadIdx.x);
}

template <typename T>
__device__ __forceinline__ M3<T> ldl(const T* __restrict__ p, size_t V, int mu, int s) {
    const T* q = p + (size_t)mu * 18 * V + s;
    M3<T> m;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        m.m[k].r = q[(2 * k) * V];
        m.m[k].i = q[(2 * k + 1) * V];
    }
    return m;
}

template <typename T>
__device__ __forceinline__ void stl(T* __restrict__ p, size_t V, int mu, int s, const M3<T>& m) {
    T* q = p + (size_t)mu * 18 * V + s;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        q[(2 * k) * V] = m.m[k].r;
        q[(2 * k + 1) * V] = m.m[k].i;
    }
}

template <typename T>
__device__ __forceinline__ void ldp(const T* __restrict__ p, size_t V, int mu, int s, T c[8]) {
    const T* q = p + (size_t)mu * 8 * V + s;
#pragma unroll
    for (int a = 0; a < 8; a++) c[a] = q[a * V];
}

template <typename T>
__device__ __forceinline__ void stp(T* __restrict__ p, size_t V, int mu, int s, const T c[8]) {
    T* q = p + (size_t)mu * 8 * V + s;
#pragma unroll
    for (int a = 0; a < 8; a++) q[a * V] = c[a];
}

// ---------------------------------------------------------------------------
// staple sum Sigma_mu(x) = sum_{nu != mu} [U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
//                                         + U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu)]
// so that tr(U_mu Sigma_mu) closes the six plaquettes through U_mu(x)
// (staples of calc_dSdUμ!, GaugeActions.jl:95-123).  The same sum is V_mu^dag for the
// loops_staple_prime staple V_mu of the flow (autostaples/wilsonloops.jl:468-485) and C_mu/rho
// of stout smearing (stout_fast.jl:603-624).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ M3<T> staple_sigma(const Geo& g, const T* __restrict__ U, size_t V, int s,
                                              const int x[4], int mu) {
    M3<T> sig = zero3<T>();
    const int spm = s + dlt(g, x, mu, 1);
#pragma unroll
    for (int nu = 0; nu < 4; nu++) {
        if (nu == mu) continue;
        int s0 = s, s1 = spm, spn = s + dlt(g, x, nu, 1);
        const int dmn = dlt(g, x, nu, -1);
        int smn = s + dmn, spmmn = spm + dmn;
        // opaque copies of the site indices: stops the compiler from merging these link
        // loads with identical ones of other (mu, nu) iterations, which would keep dozens
        // of 3x3 matrices live across the whole kernel and spill
        asm volatile("" : "+v"(s0), "+v"(s1), "+v"(spn), "+v"(smn), "+v"(spmmn));
        // upper: (U_nu(x+mu) U_mu(x+nu)^dag) U_nu(x)^dag
        const M3<T> a = mul_nd(ldl(U, V, nu, s1), ldl(U, V, mu, spn));
        acc_mul_nd(sig, a, ldl(U, V, nu, s0));
        // lower: (U_mu(x-nu) U_nu(x+mu-nu))^dag U_nu(x-nu)
        const M3<T> b = mul(ldl(U, V, mu, smn), ldl(U, V, nu, spmmn));
        acc_mul_dn(sig, b, ldl(U, V, nu, smn));
    }
    return sig;
}

// block sum of a double into partial[blockIdx.x] (fixed order: deterministic)
template <int NT>
__device__ __forceinline__ void block_sum_store(double v, double* __restrict__ partial) {
    __shared__ double red[NT / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < NT / 64; w++) t += red[w];
        partial[blockIdx.x] = t;
    }
}

}  // namespace gf
