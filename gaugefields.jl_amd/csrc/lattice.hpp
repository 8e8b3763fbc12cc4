// lattice.hpp -- site geometry and structure-of-arrays field access for the gfx950 kernels.
//
// HBM layout (one allocation per field, T = double or float, T2 = a (T, T) pair):
//   link field     : 4 directions x 6 complex planes x V sites of T2; plane (mu*6 + 3*i + j)
//                    holds element (i, j), i = 0, 1, of the SU(3) link U_mu as (re, im).  The
//                    third row is not stored: loads rebuild it as conj(row0 x row1), the
//                    unitarity + det 1 identity of SU(3) (12 of 18 reals move through HBM)
//   matrix field   : 4 directions x 9 complex planes (general 3x3 complex: derivatives,
//                    staple sums); plane (mu*9 + 3*i + j) holds element (i, j)
//   momentum field : 4 directions x 4 pair planes x V sites of T2; plane (mu*4 + j) holds the
//                    coefficients (2j, 2j+1) of i lambda_a/2
// FP32 fields pair the planes: "quad plane" q of a direction holds elements 2q and 2q+1 of
// every site as 16 contiguous bytes (a link 3 quad planes, a momentum 2, a general matrix 4 and
// a pair plane for element 8), so that one 16-byte load per lane moves two elements in either
// precision; the bytes per direction are the same as with pair planes.
// Consecutive sites x, x+1, ... are consecutive in every plane, so one wavefront of 64
// sites reads each plane with one fully coalesced 1 KiB access.
// Site order: s = x + NX*(y + NY*(z + NZ*t)), periodic in all four directions.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "su3.hpp"

namespace gf {

struct Geo {
    int n[4];   // NX, NY, NZ, NT
    int st[4];  // site strides 1, NX, NX*NY, NX*NY*NZ
    int V;
    // Morton bit widths of the 64-site group coordinates (xy-group, z, t); mbits[0] < 0
    // selects plain linear group order (extents that are not powers of two)
    int mbits[3];
    // t-sweep tiles (used when tile[0] > 0): the (xy-group, z) plane is cut into 8 tiles of
    // tile[0] x tile[1] columns, one per XCD; each XCD splits its tile along z into `sub`
    // sub-tiles and walks them one after another, each slice by slice in t
    int tile[2];
    int sub;
    // division by n[0], n[1], n[2] (set by set_divisors): q = mulhi(r, dm) >> ds, or r >> ds
    // when dm == 0 (a power of two)
    unsigned dm[3];
    int ds[3];
};

// Divisor constants for r / d, 0 <= r < 2^31: d = 2^s -> shift; otherwise with s = floor(log2 d)
// and m = ceil(2^(32+s) / d) < 2^32, floor(r / d) = floor(m r / 2^(32+s)) exactly, because
// m d - 2^(32+s) < d <= 2^(s+1) makes the error term below 1 for r < 2^31.
inline void set_divisors(Geo& g) {
    for (int m = 0; m < 3; m++) {
        const unsigned long long d = (unsigned long long)g.n[m];
        int s = 0;
        while ((2ULL << s) <= d) s++;  // s = floor(log2 d)
        g.ds[m] = s;
        g.dm[m] = (d == (1ULL << s)) ? 0u : (unsigned)(((1ULL << (32 + s)) + d - 1) / d);
    }
}
__device__ __forceinline__ unsigned divq(const Geo& g, int m, unsigned r) {
    return g.dm[m] == 0u ? (r >> g.ds[m]) : (__umulhi(r, g.dm[m]) >> g.ds[m]);
}

__device__ __forceinline__ void decode(const Geo& g, int s, int x[4]) {
    unsigned r = (unsigned)s;
#pragma unroll
    for (int m = 0; m < 3; m++) {
        const unsigned q = divq(g, m, r);
        x[m] = (int)(r - q * (unsigned)g.n[m]);
        r = q;
    }
    x[3] = (int)r;
}

// site offset for a step d = +-1 along mu from a site with coordinates x (periodic)
__device__ __forceinline__ int dlt(const Geo& g, const int x[4], int mu, int d) {
    const int xm = x[mu] + d;
    const int wrap = (g.n[mu] - 1) * g.st[mu];
    return xm >= g.n[mu] ? -wrap : (xm < 0 ? wrap : d * g.st[mu]);
}

// Workgroups are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8).  xcd_block()
// renumbers them so that XCD k runs the contiguous k-th eighth of the block sequence.
__device__ __forceinline__ unsigned xcd_block() {
    unsigned b = blockIdx.x;
    const unsigned nb = gridDim.x;
    if ((nb & 7u) == 0u) b = (b & 7u) * (nb >> 3) + (b >> 3);
    return b;
}

// one thread per site (legacy site-parallel kernels: reductions, layout conversion)
__device__ __forceinline__ int site_index() { return (int)(xcd_block() * blockDim.x + threadIdx.x); }

// Thread <-> (site, direction) map of the stencil kernels.  A workgroup of 256 threads is
// 4 wavefronts over the same group of 64 consecutive sites; wavefront w handles link
// direction mu = w.  The four directions of a site therefore run on one CU at the same
// time and share their neighbour links through L1/L2.  Group order: when the extents allow
// it, each XCD owns one of 8 tiles of (xy-group, z) columns and sweeps it in t, so the slices
// in flight re-read the previous slice from L2/MALL; otherwise 4D Morton order over
// (xy-group, z, t) (power-of-two extents) or linear order.
__host__ __device__ __forceinline__ int tsweep_group(const Geo& g, unsigned b, unsigned nb) {
    const int ngy = g.n[0] * g.n[1] / 64;
    const unsigned per = nb / 8u, xcd = b / per, loc0 = b % per;
    const int nty = ngy / g.tile[0];
    const int subz = g.tile[1] / g.sub;  // z-planes of a sub-tile
    const int cols = g.tile[0] * subz;
    const unsigned persub = per / (unsigned)g.sub;
    const int si = (int)(loc0 / persub);
    const unsigned loc = loc0 % persub;
    const int t = (int)loc / cols, c = (int)loc % cols;
    const int gy = (int)(xcd % (unsigned)nty) * g.tile[0] + c % g.tile[0];
    const int z = (int)(xcd / (unsigned)nty) * g.tile[1] + si * subz + c / g.tile[0];
    return gy + ngy * (z + g.n[2] * t);
}

__device__ __forceinline__ int group_of_block(const Geo& g) {
    const unsigned b = xcd_block();
    if (g.tile[0] > 0) return tsweep_group(g, b, gridDim.x);
    if (g.mbits[0] < 0) return (int)b;
    unsigned v0 = 0, v1 = 0, v2 = 0, bit = 0;
    const int top = max(g.mbits[0], max(g.mbits[1], g.mbits[2]));
    for (int lv = 0; lv < top; lv++) {
        if (lv < g.mbits[0]) v0 |= ((b >> bit++) & 1u) << lv;
        if (lv < g.mbits[1]) v1 |= ((b >> bit++) & 1u) << lv;
        if (lv < g.mbits[2]) v2 |= ((b >> bit++) & 1u) << lv;
    }
    return (int)(v0 + ((v1 + (v2 << g.mbits[1])) << g.mbits[0]));
}

struct SiteMu {
    int s;   // site
    int mu;  // link direction (wave-uniform)
};
__device__ __forceinline__ SiteMu site_mu_index(const Geo& g) {
    SiteMu r;
    r.s = group_of_block(g) * 64 + (int)(threadIdx.x & 63u);
    r.mu = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    return r;
}

// Field planes hold complex pairs (re, im) of type T2: one 16 B (FP64) or 8 B (FP32) access
// per lane per plane.  Accesses are raw buffer loads/stores: the buffer descriptor (one
// direction of one field) and the plane offset are wave-uniform scalars, the per-lane part is
// the 32-bit byte offset of the site, so no vector address arithmetic is spent per access.
// Offsets past the descriptor's extent read as zero and drop writes.
template <typename T>
struct pair_of;
template <>
struct pair_of<double> {
    using type = double2;
};
template <>
struct pair_of<float> {
    using type = float2;
};
template <typename T>
using T2 = typename pair_of<T>::type;

#define GF_RSRC_WORD3 0x00020000  // raw buffer, 32-bit data format (gfx9 family)

// descriptor over `planes` consecutive planes of V pairs starting at plane `first`
template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t planes_rsrc(const T* p, unsigned V, unsigned first,
                                                             unsigned planes) {
    const char* base = reinterpret_cast<const char*>(p) + (size_t)first * V * sizeof(T2<T>);
    return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)(planes * V * (unsigned)sizeof(T2<T>)),
                                             GF_RSRC_WORD3);
}

// load cache policy: GF_LD_DEFAULT, or GF_LD_NT (non-temporal: inputs each thread reads once,
// so that streaming them through L2 does not evict the neighbour links of the stencils)
#define GF_LD_DEFAULT 0
#define GF_LD_NT 2
template <int AUX = GF_LD_DEFAULT>
__device__ __forceinline__ double2 buf_ld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, const double*) {
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, AUX));
}
template <int AUX = GF_LD_DEFAULT>
__device__ __forceinline__ float2 buf_ld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, const float*) {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, AUX));
}
// store cache policy: GF_ST_DEFAULT, or GF_ST_NT (non-temporal: streamed results that this
// launch does not read back, so they do not evict the neighbour links held in L2)
#define GF_ST_DEFAULT 0
#define GF_ST_NT 2
template <int AUX = GF_ST_DEFAULT>
__device__ __forceinline__ void buf_st(double2 v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(HIP_vector_type<unsigned, 4>::Native_vec_, v), r,
                                           (int)voff, (int)soff, AUX);
}
template <int AUX = GF_ST_DEFAULT>
__device__ __forceinline__ void buf_st(float2 v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(HIP_vector_type<unsigned, 2>::Native_vec_, v), r,
                                          (int)voff, (int)soff, AUX);
}
// FP32 fields keep complex elements (2q, 2q+1) of a site together in one 16-byte slot of "quad
// plane" q (a last odd element in a pair plane after them): one 64-lane load moves two elements,
// half the vector-memory instructions of pair planes (the stencil kernels issue many loads per
// byte they need from HBM).  FP64 elements are 16 bytes already and stay in pair planes.
template <int AUX = GF_LD_DEFAULT>
__device__ __forceinline__ void buf_ld_quad(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, float2& a,
                                            float2& b) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const f4 v = __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, AUX));
    a.x = v[0];
    a.y = v[1];
    b.x = v[2];
    b.y = v[3];
}
template <int AUX = GF_ST_DEFAULT>
__device__ __forceinline__ void buf_st_quad(float2 a, float2 b, __amdgpu_buffer_rsrc_t r, unsigned voff,
                                            unsigned soff) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    f4 v;
    v[0] = a.x;
    v[1] = a.y;
    v[2] = b.x;
    v[3] = b.y;
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(HIP_vector_type<unsigned, 4>::Native_vec_, v), r,
                                           (int)voff, (int)soff, AUX);
}

// SU(3) link field: planes (mu*6 + k), k = 3*i + j < 6, hold rows 0 and 1 of U_mu (FP32: quad
// planes (mu*3 + q) hold elements 2q, 2q+1).
// R6 is a link as loaded (its two stored rows); full_of() rebuilds the third row.
template <typename T>
struct R6 {
    T2<T> v[6];
};
template <typename T, int AUX = GF_LD_DEFAULT>
__device__ __forceinline__ R6<T> ldr(const T* __restrict__ p, unsigned V, int mu, unsigned s) {
    const __amdgpu_buffer_rsrc_t r = planes_rsrc(p, V, 6u * mu, 6u);
    R6<T> o;
    if constexpr (std::is_same<T, float>::value) {
        const unsigned off = s * 16u, ps = V * 16u;
#pragma unroll
        for (int q = 0; q < 3; q++) buf_ld_quad<AUX>(r, off, q * ps, o.v[2 * q], o.v[2 * q + 1]);
    } else {
        const unsigned off = s * (unsigned)sizeof(T2<T>), ps = V * (unsigned)sizeof(T2<T>);
#pragma unroll
        for (int k = 0; k < 6; k++) o.v[k] = buf_ld<AUX>(r, off, k * ps, (const T*)nullptr);
    }
    return o;
}
// row2_k = conj(row0_{k+1} row1_{k+2} - row0_{k+2} row1_{k+1}): exact for SU(3) (unitary, det 1)
template <typename T>
__device__ __forceinline__ void complete_row2(M3<T>& m) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int a = (k + 1) % 3, b = (k + 2) % 3;
        const cx<T> p0 = m.m[a], p1 = m.m[b], q0 = m.m[3 + a], q1 = m.m[3 + b];
        // one product and three FMAs per part (4 instead of 5 operations)
        T re = p0.r * q1.r;
        re = fma(-p0.i, q1.i, re);
        re = fma(-p1.r, q0.r, re);
        re = fma(p1.i, q0.i, re);
        T im = -p0.r * q1.i;
        im = fma(-p0.i, q1.r, im);
        im = fma(p1.r, q0.i, im);
        im = fma(p1.i, q0.r, im);
        m.m[6 + k].r = re;
        m.m[6 + k].i = im;
    }
}
template <typename T>
__device__ __forceinline__ M3<T> full_of(const R6<T>& r) {
    M3<T> m;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        m.m[k].r = r.v[k].x;
        m.m[k].i = r.v[k].y;
    }
    complete_row2(m);
    return m;
}
// the two stored rows in double (exact); full_of() of it rebuilds row 2 in double, the same for
// every kernel that sums in FP64 over FP32-stored links
template <typename T>
__device__ __forceinline__ R6<double> r6_f64(const R6<T>& r) {
    R6<double> o;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        o.v[k].x = (double)r.v[k].x;
        o.v[k].y = (double)r.v[k].y;
    }
    return o;
}
template <typename T, int AUX = GF_LD_DEFAULT>
__device__ __forceinline__ M3<T> ldu(const T* __restrict__ p, unsigned V, int mu, unsigned s) {
    return full_of(ldr<T, AUX>(p, V, mu, s));
}
template <typename T>
__device__ __forceinline__ cx<T> rc(const R6<T>& r, int k) {
    cx<T> c;
    c.r = r.v[k].x;
    c.i = r.v[k].y;
    return c;
}
// the two stored rows of an SU(3) matrix held whole
template <typename T>
__device__ __forceinline__ R6<T> rows_of(const M3<T>& m) {
    R6<T> r;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        r.v[k].x = m.m[k].r;
        r.v[k].y = m.m[k].i;
    }
    return r;
}
// products whose result is SU(3): rows 0 and 1 from the two stored rows of a, row 2 rebuilt
template <typename T>
__device__ __forceinline__ M3<T> su3_mul_nd(const R6<T>& a, const M3<T>& b) {  // a b^dag
    M3<T> c;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cx<T> t;
            t.r = t.i = T(0);
#pragma unroll
            for (int k = 0; k < 3; k++) cmac_nc(t, rc(a, 3 * i + k), b.m[3 * j + k]);
            c.m[3 * i + j] = t;
        }
    complete_row2(c);
    return c;
}
template <typename T>
__device__ __forceinline__ M3<T> su3_mul(const R6<T>& a, const M3<T>& b) {  // a b
    M3<T> c;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cx<T> t;
            t.r = t.i = T(0);
#pragma unroll
            for (int k = 0; k < 3; k++) cmac(t, rc(a, 3 * i + k), b.m[3 * k + j]);
            c.m[3 * i + j] = t;
        }
    complete_row2(c);
    return c;
}

template <typename T>
__device__ __forceinline__ T2<T> t2_of(const cx<T>& c) {
    T2<T> v;
    v.x = c.r;
    v.y = c.i;
    return v;
}
template <typename T, int AUX = GF_ST_DEFAULT>
__device__ __forceinline__ void stu(T* __restrict__ p, unsigned V, int mu, unsigned s, const M3<T>& m) {
    const __amdgpu_buffer_rsrc_t r = planes_rsrc(p, V, 6u * mu, 6u);
    if constexpr (std::is_same<T, float>::value) {
        const unsigned off = s * 16u, ps = V * 16u;
#pragma unroll
        for (int q = 0; q < 3; q++) buf_st_quad<AUX>(t2_of(m.m[2 * q]), t2_of(m.m[2 * q + 1]), r, off, q * ps);
    } else {
        const unsigned off = s * (unsigned)sizeof(T2<T>), ps = V * (unsigned)sizeof(T2<T>);
#pragma unroll
        for (int k = 0; k < 6; k++) buf_st<AUX>(t2_of(m.m[k]), r, off, k * ps);
    }
}

// general matrix field: plane (mu*9 + k) holds complex element k = 3*i + j for all sites (FP32:
// quad planes 0..3 of the direction hold elements 2q, 2q+1, a pair plane after them element 8)
template <typename T, int AUX = GF_LD_DEFAULT>
__device__ __forceinline__ M3<T> ldm(const T* __restrict__ p, unsigned V, int mu, unsigned s) {
    const __amdgpu_buffer_rsrc_t r = planes_rsrc(p, V, 9u * mu, 9u);
    M3<T> m;
    T2<T> v[9];
    if constexpr (std::is_same<T, float>::value) {
        const unsigned ps = V * 16u;
#pragma unroll
        for (int q = 0; q < 4; q++) buf_ld_quad<AUX>(r, s * 16u, q * ps, v[2 * q], v[2 * q + 1]);
        v[8] = buf_ld<AUX>(r, s * 8u, 4u * ps, (const T*)nullptr);
    } else {
        const unsigned off = s * (unsigned)sizeof(T2<T>), ps = V * (unsigned)sizeof(T2<T>);
#pragma unroll
        for (int k = 0; k < 9; k++) v[k] = buf_ld<AUX>(r, off, k * ps, (const T*)nullptr);
    }
#pragma unroll
    for (int k = 0; k < 9; k++) {
        m.m[k].r = v[k].x;
        m.m[k].i = v[k].y;
    }
    return m;
}

template <typename T, int AUX = GF_ST_DEFAULT>
__device__ __forceinline__ void stm(T* __restrict__ p, unsigned V, int mu, unsigned s, const M3<T>& m) {
    const __amdgpu_buffer_rsrc_t r = planes_rsrc(p, V, 9u * mu, 9u);
    if constexpr (std::is_same<T, float>::value) {
        const unsigned ps = V * 16u;
#pragma unroll
        for (int q = 0; q < 4; q++) buf_st_quad<AUX>(t2_of(m.m[2 * q]), t2_of(m.m[2 * q + 1]), r, s * 16u, q * ps);
        buf_st<AUX>(t2_of(m.m[8]), r, s * 8u, 4u * ps);
    } else {
        const unsigned off = s * (unsigned)sizeof(T2<T>), ps = V * (unsigned)sizeof(T2<T>);
#pragma unroll
        for (int k = 0; k < 9; k++) buf_st<AUX>(t2_of(m.m[k]), r, off, k * ps);
    }
}

// momentum field: plane (mu*4 + j) holds coefficients (2j, 2j+1) of P_mu for all sites (FP32:
// quad planes (mu*2 + q) hold coefficients 4q .. 4q+3)
template <typename T, int AUX = GF_LD_DEFAULT>
__device__ __forceinline__ void ldp(const T* __restrict__ p, unsigned V, int mu, unsigned s, T c[8]) {
    const __amdgpu_buffer_rsrc_t r = planes_rsrc(p, V, 4u * mu, 4u);
    T2<T> v[4];
    if constexpr (std::is_same<T, float>::value) {
        const unsigned ps = V * 16u;
#pragma unroll
        for (int q = 0; q < 2; q++) buf_ld_quad<AUX>(r, s * 16u, q * ps, v[2 * q], v[2 * q + 1]);
    } else {
        const unsigned off = s * (unsigned)sizeof(T2<T>), ps = V * (unsigned)sizeof(T2<T>);
#pragma unroll
        for (int j = 0; j < 4; j++) v[j] = buf_ld<AUX>(r, off, j * ps, (const T*)nullptr);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        c[2 * j] = v[j].x;
        c[2 * j + 1] = v[j].y;
    }
}

template <typename T, int AUX = GF_ST_DEFAULT>
__device__ __forceinline__ void stp(T* __restrict__ p, unsigned V, int mu, unsigned s, const T c[8]) {
    const __amdgpu_buffer_rsrc_t r = planes_rsrc(p, V, 4u * mu, 4u);
    T2<T> v[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        v[j].x = c[2 * j];
        v[j].y = c[2 * j + 1];
    }
    if constexpr (std::is_same<T, float>::value) {
        const unsigned ps = V * 16u;
#pragma unroll
        for (int q = 0; q < 2; q++) buf_st_quad<AUX>(v[2 * q], v[2 * q + 1], r, s * 16u, q * ps);
    } else {
        const unsigned off = s * (unsigned)sizeof(T2<T>), ps = V * (unsigned)sizeof(T2<T>);
#pragma unroll
        for (int j = 0; j < 4; j++) buf_st<AUX>(v[j], r, off, j * ps);
    }
}

// ---------------------------------------------------------------------------
// staple sum Sigma_mu(x) = sum_{nu != mu} [U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
//                                         + U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu)]
// so that tr(U_mu Sigma_mu) closes the six plaquettes through U_mu(x)
// (staples of calc_dSdUμ!, GaugeActions.jl:95-123).  The same sum is V_mu^dag for the
// loops_staple_prime staple V_mu of the flow (autostaples/wilsonloops.jl:468-485) and C_mu/rho
// of stout smearing (stout_fast.jl:603-624).
//
// The 18 staple links (and, for staple_sigma_own, U_mu(x) itself as a 19th) are loaded one
// link ahead of their use: the loads of link k+1 are issued before the arithmetic on link k,
// and a scheduling barrier keeps the compiler from hoisting further loads (which would keep
// dozens of matrices live) or sinking these to their use.  The two products of each staple
// have SU(3) intermediates and are formed as su3_mul / su3_mul_nd.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int sel4(int i, int a0, int a1, int a2, int a3) {
    return i == 0 ? a0 : (i == 1 ? a1 : (i == 2 ? a2 : a3));
}
// LOWER: the lower staples only (links 3..5 of blocks J0..2) and, with OWN, U_mu(x).
// DEPTH: links in flight ahead of the one being used (1; 2 where the registers allow)
template <typename T, bool OWN, bool LOWER = false, int J0 = 0, int DEPTH = 1>
__device__ __forceinline__ M3<T> staple_sigma_pipe(const Geo& g, const T* __restrict__ U, unsigned V, int s,
                                                   const int x[4], int mu, R6<T>* own) {
    int fw[4], bw[4];
#pragma unroll
    for (int d = 0; d < 4; d++) {
        fw[d] = dlt(g, x, d, 1);
        bw[d] = dlt(g, x, d, -1);
    }
    const int fmu = sel4(mu, fw[0], fw[1], fw[2], fw[3]);
    // link k < 18: block j = k / 6 of direction nu = j + (j >= mu), slot k % 6:
    //   0 U_nu(x+mu)  1 U_mu(x+nu)  2 U_nu(x)  3 U_mu(x-nu)  4 U_nu(x+mu-nu)  5 U_nu(x-nu)
    // link 18: U_mu(x)
    auto issue = [&](int k) -> R6<T> {
        if (k == 18) return ldr(U, V, mu, (unsigned)s);
        const int j = k / 6, r = k % 6;
        const int nu = j + (j >= mu ? 1 : 0);
        const int fnu = sel4(nu, fw[0], fw[1], fw[2], fw[3]);
        const int bnu = sel4(nu, bw[0], bw[1], bw[2], bw[3]);
        int dir = nu, site = s + bnu;
        if (r == 0) site = s + fmu;
        if (r == 1) dir = mu, site = s + fnu;
        if (r == 2) site = s;
        if (r == 3) dir = mu;
        if (r == 4) site = s + fmu + bnu;
        return ldr(U, V, dir, (unsigned)site);
    };
    constexpr int NLOW = 3 * (3 - J0);
    constexpr int NL = LOWER ? (OWN ? NLOW + 1 : NLOW) : (OWN ? 19 : 18);
    auto slot = [](int i) { return LOWER ? (i < NLOW ? 6 * (i / 3 + J0) + 3 + i % 3 : 18) : i; };
    M3<T> sig = zero3<T>();
    M3<T> X;
    R6<T> A, buf[DEPTH];
#pragma unroll
    for (int d = 0; d < DEPTH; d++)
        if (d < NL) buf[d] = issue(slot(d));
#pragma unroll
    for (int i = 0; i < NL; i++) {
        const int k = slot(i);
        const R6<T> cur = buf[i % DEPTH];
        if (i + DEPTH < NL) buf[i % DEPTH] = issue(slot(i + DEPTH));
        __builtin_amdgcn_sched_barrier(0);
        if (k == 18) {
            *own = cur;
            break;
        }
        switch (k % 6) {
            case 0: A = cur; break;                                  // U_nu(x+mu)
            case 1: X = su3_mul_nd(A, full_of(cur)); break;          // U_nu(x+mu) U_mu(x+nu)^dag
            case 2: acc_mul_nd(sig, X, full_of(cur)); break;         // ... U_nu(x)^dag
            case 3: A = cur; break;                                  // U_mu(x-nu)
            case 4: X = su3_mul(A, full_of(cur)); break;             // U_mu(x-nu) U_nu(x+mu-nu)
            default: acc_mul_dn(sig, X, full_of(cur)); break;        // (...)^dag U_nu(x-nu)
        }
    }
    return sig;
}
// the staple sum and U_mu(x) (own, as loaded): force, flow and stout-forward kernels
template <typename T, int DEPTH = 1>
__device__ __forceinline__ M3<T> staple_sigma_own(const Geo& g, const T* __restrict__ U, unsigned V, int s,
                                                  const int x[4], int mu, R6<T>& own) {
    return staple_sigma_pipe<T, true, false, 0, DEPTH>(g, U, V, s, x, mu, &own);
}

// ---------------------------------------------------------------------------
// Staple sum with the upper-staple products shared between directions.  The product
// X_mu,nu = U_nu(x+mu) U_mu(x+nu)^dag of the (mu, nu) upper staple is the Hermitian conjugate of
// X_nu,mu, and the workgroup holds the four direction-waves of the same 64 sites, so each
// unordered pair is formed once, its rows 0 and 1 are written to LDS, and after one workgroup
// barrier every wave reads the three it needs: per site 12 of 76 link loads and 6 of 24 staple
// products fewer than staple_sigma_own.  Pair (a, b), a < b, is formed by wave
// w0: (0,1), (0,3)  w1: (1,2), (1,3)  w2: (0,2)  w3: (2,3), in a loop of its own ahead of the
// pipelined lower staples.  Requires every lane of the workgroup to reach the barrier: the
// kernels use it only when V is a multiple of 64 (no idle lanes).
// LDS: xs[6][6][64] complex (37 KB FP64, 18 KB FP32).
__device__ __forceinline__ int pair_index(int a, int b) {  // a < b
    return a == 0 ? b - 1 : (a == 1 ? b + 1 : 5);
}
template <typename T>
using XShare = T2<T> (*)[6][64];
template <typename T>
__device__ __forceinline__ M3<T> staple_sigma_shared(const Geo& g, const T* __restrict__ U, unsigned V, int s,
                                                     const int x[4], int mu, R6<T>& own, XShare<T> xs) {
    const int lane = (int)(threadIdx.x & 63u);
    int fw[4];
#pragma unroll
    for (int d = 0; d < 4; d++) fw[d] = dlt(g, x, d, 1);
    const int np = mu < 2 ? 2 : 1;
#pragma unroll 1
    for (int q = 0; q < np; q++) {
        // w0 (0,1),(0,3); w1 (1,2),(1,3); w2 (0,2); w3 (2,3)
        const int a = q == 1 ? mu : (mu == 2 ? 0 : (mu == 3 ? 2 : mu));
        const int b = q == 1 ? 3 : (mu == 0 ? 1 : (mu == 1 ? 2 : (mu == 2 ? 2 : 3)));
        const int fa = sel4(a, fw[0], fw[1], fw[2], fw[3]), fb = sel4(b, fw[0], fw[1], fw[2], fw[3]);
        const M3<T> X = su3_mul_nd(ldr(U, V, b, (unsigned)(s + fa)), ldu(U, V, a, (unsigned)(s + fb)));
        const int p = pair_index(a, b);
#pragma unroll
        for (int e = 0; e < 6; e++) {
            T2<T> v;
            v.x = X.m[e].r;
            v.y = X.m[e].i;
            xs[p][e][lane] = v;
        }
    }
    M3<T> sig = staple_sigma_pipe<T, false, true, 0>(g, U, V, s, x, mu, nullptr);
    __syncthreads();  // every pair's X is in LDS
    // upper staples X_mu,nu U_nu(x)^dag, then U_mu(x); the links loaded one ahead
    auto issue = [&](int k) -> R6<T> {
        if (k == 3) return ldr(U, V, mu, (unsigned)s);
        const int nu = k + (k >= mu ? 1 : 0);
        return ldr(U, V, nu, (unsigned)s);
    };
    R6<T> nxt = issue(0);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const R6<T> cur = nxt;
        if (k + 1 < 4) nxt = issue(k + 1);
        __builtin_amdgcn_sched_barrier(0);
        if (k == 3) {
            own = cur;
            break;
        }
        const int nu = k + (k >= mu ? 1 : 0);
        const int p = pair_index(mu < nu ? mu : nu, mu < nu ? nu : mu);
        M3<T> M;
#pragma unroll
        for (int e = 0; e < 6; e++) {
            const T2<T> v = xs[p][e][lane];
            M.m[e].r = v.x;
            M.m[e].i = v.y;
        }
        complete_row2(M);
        const bool t = mu > nu;  // X_mu,nu = (X_nu,mu)^dag
        M3<T> Mt;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int jj = 0; jj < 3; jj++) {
                Mt.m[3 * i + jj].r = t ? M.m[3 * jj + i].r : M.m[3 * i + jj].r;
                Mt.m[3 * i + jj].i = t ? -M.m[3 * jj + i].i : M.m[3 * i + jj].i;
            }
        acc_mul_nd(sig, Mt, full_of(cur));
    }
    return sig;
}

// ---------------------------------------------------------------------------
// TA(U_mu(x) Sigma_mu(x)) (8 su(3) coefficients, ta_coeffs) with the upper staples taken as
// whole plaquettes shared between directions.  U_mu(x) X_mu,nu U_nu(x)^dag is the plaquette
// P_mu,nu(x), and P_nu,mu = P_mu,nu^dag, so TA(P_nu,mu) = -TA(P_mu,nu).  Each unordered pair's
// plaquette is formed once (by the waves and in the slots of staple_sigma_shared), only its 8
// TA coefficients go to LDS, and the links U_nu(x) reach the other waves through LDS as well:
// per wave 14 link loads (88 vector-memory loads with the momenta) instead of 17 (106).
// Order: lower staples, U_mu(x) (-> LDS, TA of U_mu times the lower staples), the pairs.
// TA is linear, so the result is TA(U_mu Sigma_mu) up to rounding.  Two workgroup barriers
// (every lane must reach them: V a multiple of 64).
// LDS: ol[4][6][64] links and tp[7][4][64] coefficient pairs (52 KB FP64, 26 KB FP32).
template <typename T>
using PairShare = T2<T> (*)[4][64];
template <typename T>
__device__ __forceinline__ void ta_force_shared(const Geo& g, const T* __restrict__ U, unsigned V, int s,
                                                const int x[4], int mu, R6<T>& own, XShare<T> ol,
                                                PairShare<T> tp, T c[8]) {
    const int lane = (int)(threadIdx.x & 63u);
    int fw[4];
#pragma unroll
    for (int d = 0; d < 4; d++) fw[d] = dlt(g, x, d, 1);
    // pairs formed by this wave: w0 (0,1),(0,3); w1 (1,2),(1,3); w2 (0,2); w3 (2,3); waves 2 and 3
    // form (2,3) a second time into the unused slot 6 (same branch-free sequence for all)
    const int pa1 = mu == 2 ? 0 : (mu == 3 ? 2 : mu);
    const int pb1 = mu == 0 ? 1 : (mu == 1 ? 2 : (mu == 2 ? 2 : 3));
    const int pa2 = mu < 2 ? mu : 2;
    const int pb2 = 3;
    const int p1 = pair_index(pa1, pb1), p2 = mu < 2 ? pair_index(pa2, pb2) : 6;
    // TA coefficients of P_a,b = U_a(x) X U_b(x)^dag into slot p
    auto plaq = [&](int a, int b, const M3<T>& X, int p) {
        R6<T> ra, rb;
#pragma unroll
        for (int e = 0; e < 6; e++) {
            ra.v[e] = ol[a][e][lane];
            rb.v[e] = ol[b][e][lane];
        }
        T q[8];
        ta_coeffs(su3_mul_nd(rows_of(su3_mul(ra, X)), full_of(rb)), q);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            T2<T> v;
            v.x = q[2 * e];
            v.y = q[2 * e + 1];
            tp[p][e][lane] = v;
        }
    };
    const M3<T> sig = staple_sigma_pipe<T, true, true>(g, U, V, s, x, mu, &own);
#pragma unroll
    for (int e = 0; e < 6; e++) ol[mu][e][lane] = own.v[e];
    ta_coeffs(mul(full_of(own), sig), c);
    // pin c here: otherwise the staple arithmetic is sunk past the pairs to its use, keeping
    // every loaded staple link live across them
    asm volatile("" : "+v"(c[0]), "+v"(c[1]), "+v"(c[2]), "+v"(c[3]), "+v"(c[4]), "+v"(c[5]), "+v"(c[6]),
                 "+v"(c[7]));
    // the pairs: a loop of its own, so that its loads are not hoisted into the staples above
#pragma unroll 1
    for (int q = 0; q < 2; q++) {
        const int a = q == 0 ? pa1 : pa2, b = q == 0 ? pb1 : pb2;
        const int fa = sel4(a, fw[0], fw[1], fw[2], fw[3]), fb = sel4(b, fw[0], fw[1], fw[2], fw[3]);
        const M3<T> Xp = su3_mul_nd(ldr(U, V, b, (unsigned)(s + fa)), ldu(U, V, a, (unsigned)(s + fb)));
        if (q == 0) __syncthreads();  // every wave's U(x) is in LDS
        plaq(a, b, Xp, q == 0 ? p1 : p2);
    }
    __syncthreads();  // every pair's coefficients are in LDS
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int nu = j + (j >= mu ? 1 : 0);
        const int p = pair_index(mu < nu ? mu : nu, mu < nu ? nu : mu);
        const T sg = mu < nu ? T(1) : T(-1);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const T2<T> v = tp[p][e][lane];
            c[2 * e] = fma(sg, v.x, c[2 * e]);
            c[2 * e + 1] = fma(sg, v.y, c[2 * e + 1]);
        }
    }
}

// ---------------------------------------------------------------------------
// ta_force_shared for workgroups of whole x-rows (NX divides 64, V a multiple of 64): the
// x-neighbours x +- 1 of every site are sites of the same workgroup, so the links at them come
// from the U(x) copies in LDS instead of memory: U_mu(x-1), U_0(x-1) of the nu = 0 lower staple
// and U_b(x+1) of the pairs (0, b).  Pairs: w0 (0,3); w1 (1,2), (0,1); w2 (2,3), (0,2); w3 (1,3)
// (link loads per wave 12, 11, 11, 10 instead of 14).
// LDS: ol[4][6][64] links and tp[6][4][64] coefficient pairs.
__device__ __forceinline__ int xlane(int lane, int nxm1, int d) { return (lane & ~nxm1) | ((lane + d) & nxm1); }
template <typename T>
__device__ __forceinline__ void ta_force_xrow(const Geo& g, const T* __restrict__ U, unsigned V, int s,
                                              const int x[4], int mu, R6<T>& own, XShare<T> ol,
                                              PairShare<T> tp, T c[8]) {
    const int lane = (int)(threadIdx.x & 63u);
    const int nxm1 = g.n[0] - 1;
    const int lxm = xlane(lane, nxm1, -1), lxp = xlane(lane, nxm1, 1);
    int fw[4];
#pragma unroll
    for (int d = 0; d < 4; d++) fw[d] = dlt(g, x, d, 1);
    auto lds_link = [&](int d, int l) {
        R6<T> r;
#pragma unroll
        for (int e = 0; e < 6; e++) r.v[e] = ol[d][e][l];
        return r;
    };
    {
        const R6<T> o = ldr(U, V, mu, (unsigned)s);
#pragma unroll
        for (int e = 0; e < 6; e++) ol[mu][e][lane] = o.v[e];
    }
    __syncthreads();  // every wave's U(x) is in LDS
    M3<T> sig;
    if (mu == 0) {
        sig = staple_sigma_pipe<T, false, true, 0>(g, U, V, s, x, mu, nullptr);
    } else {
        sig = staple_sigma_pipe<T, false, true, 1>(g, U, V, s, x, mu, nullptr);
        // nu = 0: (U_mu(x-1) U_0(x+mu-1))^dag U_0(x-1)
        const M3<T> X = su3_mul(lds_link(mu, lxm), ldu(U, V, 0, (unsigned)(s + fw[mu] + dlt(g, x, 0, -1))));
        acc_mul_dn(sig, X, full_of(lds_link(0, lxm)));
    }
    own = lds_link(mu, lane);  // read back rather than held through the staples
    ta_coeffs(mul(full_of(own), sig), c);
    // pin c here: otherwise the staple arithmetic is sunk past the pairs to its use
    asm volatile("" : "+v"(c[0]), "+v"(c[1]), "+v"(c[2]), "+v"(c[3]), "+v"(c[4]), "+v"(c[5]), "+v"(c[6]),
                 "+v"(c[7]));
    const int np = mu == 0 || mu == 3 ? 1 : 2;
#pragma unroll 1
    for (int q = 0; q < np; q++) {
        // (a, b): w0 (0,3); w1 (1,2), (0,1); w2 (2,3), (0,2); w3 (1,3)
        const int a = q == 1 ? 0 : (mu == 2 ? 2 : (mu == 3 ? 1 : mu));
        const int b = q == 1 ? mu : (mu == 0 ? 3 : (mu == 1 ? 2 : 3));
        // X_ab = U_b(x+a) U_a(x+b)^dag
        const R6<T> ub = a == 0 ? lds_link(b, lxp) : ldr(U, V, b, (unsigned)(s + sel4(a, fw[0], fw[1], fw[2], fw[3])));
        const M3<T> X = su3_mul_nd(ub, ldu(U, V, a, (unsigned)(s + sel4(b, fw[0], fw[1], fw[2], fw[3]))));
        T q8[8];
        ta_coeffs(su3_mul_nd(rows_of(su3_mul(lds_link(a, lane), X)), full_of(lds_link(b, lane))), q8);
        const int p = pair_index(a, b);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            T2<T> v;
            v.x = q8[2 * e];
            v.y = q8[2 * e + 1];
            tp[p][e][lane] = v;
        }
    }
    __syncthreads();  // every pair's coefficients are in LDS
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int nu = j + (j >= mu ? 1 : 0);
        const int p = pair_index(mu < nu ? mu : nu, mu < nu ? nu : mu);
        const T sg = mu < nu ? T(1) : T(-1);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const T2<T> v = tp[p][e][lane];
            c[2 * e] = fma(sg, v.x, c[2 * e]);
            c[2 * e + 1] = fma(sg, v.y, c[2 * e + 1]);
        }
    }
}

// staple_sigma_own or, when SHARED, staple_sigma_shared (xs: the kernel's LDS array)
template <typename T, bool SHARED>
__device__ __forceinline__ M3<T> staple_sum_own(const Geo& g, const T* __restrict__ U, unsigned V, int s,
                                                const int x[4], int mu, R6<T>& own, XShare<T> xs) {
    if constexpr (SHARED)
        return staple_sigma_shared(g, U, V, s, x, mu, own, xs);
    else
        return staple_sigma_own(g, U, V, s, x, mu, own);
}
// Stencil form of the force-type kernels (k_force, k_flow_stage): per-thread staples; upper-
// staple plaquettes shared through LDS (every lane a site: V % 64 == 0); and the same with the
// x-neighbour links from LDS (workgroups of whole x-rows: NX divides 64).
enum { ST_OWN = 0, ST_SHARED = 1, ST_XROW = 2 };
inline int stencil_form(const Geo& g) {
    if (g.V % 64 != 0) return ST_OWN;
    return 64 % g.n[0] == 0 ? ST_XROW : ST_SHARED;
}
// Site of a stencil thread.  The SHARED kernels have a workgroup barrier, so no lane may
// leave early: they run only for V % 64 == 0, where every lane is a site; the clamp keeps
// their indices in range regardless.
template <bool SHARED>
__device__ __forceinline__ bool stencil_site(const Geo& g, const SiteMu& sm, int& s) {
    s = sm.s;
    if constexpr (SHARED) {
        s = min(s, g.V - 1);
        return true;
    } else {
        return s < g.V;
    }
}

// Unpipelined form for kernels with large register state of their own (the stout backward's
// recomputation, dS/dU off the 64-site grid), where the prefetch buffer would cost occupancy: links
// are loaded at their use, through opaque copies of the site indices so that the compiler
// neither merges them with the loads of other (mu, nu) iterations nor hoists them all.
template <typename T>
__device__ __forceinline__ M3<T> staple_sigma(const Geo& g, const T* __restrict__ U, unsigned V, int s,
                                              const int x[4], int mu) {
    M3<T> sig = zero3<T>();
    const int spm = s + dlt(g, x, mu, 1);
#pragma unroll
    for (int nu = 0; nu < 4; nu++) {
        if (nu == mu) continue;
        {  // upper: (U_nu(x+mu) U_mu(x+nu)^dag) U_nu(x)^dag
            int s0 = s, s1 = spm, spn = s + dlt(g, x, nu, 1);
            asm volatile("" : "+v"(s0), "+v"(s1), "+v"(spn));
            const M3<T> a = mul_nd(ldu(U, V, nu, s1), ldu(U, V, mu, spn));
            acc_mul_nd(sig, a, ldu(U, V, nu, s0));
        }
        {  // lower: (U_mu(x-nu) U_nu(x+mu-nu))^dag U_nu(x-nu)
            const int dmn = dlt(g, x, nu, -1);
            int smn = s + dmn, spmmn = spm + dmn;
            asm volatile("" : "+v"(smn), "+v"(spmmn));
            const M3<T> b = mul(ldu(U, V, mu, smn), ldu(U, V, nu, spmmn));
            acc_mul_dn(sig, b, ldu(U, V, nu, smn));
        }
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
