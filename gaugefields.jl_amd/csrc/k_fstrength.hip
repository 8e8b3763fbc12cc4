// k_fstrength.hip -- lattice field strengths F_munu(x) = TA(sum of closed loops at x) and what
// is measured from them: the topological charge density with the plaquette, clover and 2x1
// rectangle loop sets (improved = 5/3 clover - 1/12 rectangle) and the clover energy density.
//
// Reference: _plaquette/_clover/_rectangle_field_strengths and the matching
// *_topological_charge_density (AbstractGaugefields.jl:1195-1397), make_cloverloops
// (autostaples/wilsonloops.jl:166-177), calculate_energy_density
// (samples/measurements/energydensity.jl:32-76).
//
// One workgroup = 64 sites x 4 wavefronts; wavefront w evaluates the planes w and w + 4 of the
// six (mu < nu) (with both clover and rectangle sets: the rectangles of planes 2w, 2w + 1 for
// w < 3, every clover for w = 3) and leaves the 8 su(3) coefficients of each F in LDS;
// wavefront 0 then combines the planes of each site.  With F = sum_a c_a i lambda_a / 2, tr(F F') = -(1/2) c . c', so
//   sum_{mu nu rho sigma} eps tr(F_munu F_rhosigma) = -4 (c01.c23 - c02.c13 + c03.c12)
// and the density -norm Re(...) / (32 pi^2 n_loops^2) = norm (c01.c23 - c02.c13 + c03.c12) /
// (8 pi^2 n_loops^2); the clover energy sum_{mu != nu} -tr(F F)/2 / 16 = sum_planes |c|^2 / 32.
// Products are evaluated in double for both storage precisions.
#include "internal.hpp"

namespace gf {

namespace {

// loop tables: steps (symbol, length), symbol 0 = mu, 1 = nu
constexpr signed char kClover[4][4][2] = {{{0, 1}, {1, 1}, {0, -1}, {1, -1}},
                                          {{1, 1}, {0, -1}, {1, -1}, {0, 1}},
                                          {{1, -1}, {0, 1}, {1, 1}, {0, -1}},
                                          {{0, -1}, {1, -1}, {0, 1}, {1, 1}}};
constexpr signed char kRect[8][4][2] = {{{0, 2}, {1, 1}, {0, -2}, {1, -1}}, {{1, 1}, {0, -2}, {1, -1}, {0, 2}},
                                        {{1, -1}, {0, 2}, {1, 1}, {0, -2}}, {{0, -2}, {1, -1}, {0, 2}, {1, 1}},
                                        {{0, 1}, {1, 2}, {0, -1}, {1, -2}}, {{1, 2}, {0, -1}, {1, -2}, {0, 1}},
                                        {{1, -2}, {0, 1}, {1, 2}, {0, -1}}, {{0, -1}, {1, -2}, {0, 1}, {1, 2}}};

__device__ __forceinline__ unsigned site_of(const Geo& g, const int y[4]) {
    return (unsigned)(y[0] + g.n[0] * (y[1] + g.n[1] * (y[2] + g.n[2] * y[3])));
}

// product of links along one loop from x (evaluate_gaugelinks!, AbstractGaugefields.jl:1797-1860):
// (d, +n): U_d(y), y -> y + d, n times; (d, -n): y -> y - d, U_d(y)^dag, n times.
// FP64 links are SU(3) to double rounding, so every partial product is too: only its rows 0 and
// 1 are formed (72 instead of 108 FMAs per product) and row 2 of the loop is rebuilt once at the
// end; a link entering as the first factor of a positive step needs no row 2 either.  FP32-stored
// links are unitary only to float rounding, which the cross product would carry into every
// partial product: there the full products are formed.  Links are widened to double before
// row 2 is rebuilt, as in every FP64 sum over FP32-stored links.
template <typename T>
__device__ __forceinline__ M3<double> walk(const Geo& g, const T* __restrict__ U, unsigned V, const int x[4],
                                           int mu, int nu, const signed char (&st)[4][2]) {
    constexpr bool ROWS = sizeof(T) == 8;
    int y[4] = {x[0], x[1], x[2], x[3]};
    R6<double> r;   // ROWS: rows 0 and 1 of the product so far
    M3<double> m;   // otherwise: the product so far
    bool first = true;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int d = st[i][0] ? nu : mu;
        const int len = st[i][1];
        const int n = len > 0 ? len : -len;
#pragma unroll
        for (int q = 0; q < 2; q++) {
            if (q < n) {
                if (len > 0) {
                    const R6<double> l = r6_f64(ldr(U, V, d, site_of(g, y)));
                    if constexpr (ROWS)
                        r = first ? l : rows_of(su3_mul(r, full_of(l)));
                    else
                        m = first ? full_of(l) : mul(m, full_of(l));
                    y[d] = y[d] + 1 == g.n[d] ? 0 : y[d] + 1;
                } else {
                    y[d] = y[d] == 0 ? g.n[d] - 1 : y[d] - 1;
                    const M3<double> l = full_of(r6_f64(ldr(U, V, d, site_of(g, y))));
                    if constexpr (ROWS)
                        r = first ? rows_of(dag(l)) : rows_of(su3_mul_nd(r, l));
                    else
                        m = first ? dag(l) : mul_nd(m, l);
                }
                first = false;
            }
        }
    }
    if constexpr (ROWS)
        return full_of(r);
    else
        return m;
}

__device__ __forceinline__ double pair_sum(const double (*c)[8][64], int lane) {
    double a = 0.0, b = 0.0, d = 0.0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        a = fma(c[0][k][lane], c[5][k][lane], a);  // (01)(23)
        b = fma(c[1][k][lane], c[4][k][lane], b);  // (02)(13)
        d = fma(c[2][k][lane], c[3][k][lane], d);  // (03)(12)
    }
    return a - b + d;
}

}  // namespace

// SETS: bit 0 plaquette, bit 1 clover, bit 2 rectangle.  Partial sums per block:
//   partial[0*nb + b] clover energy (sum |c|^2 / 32), [1*nb] plaquette q, [2*nb] clover q,
//   [3*nb] rectangle q.  dens (optional): density of method dens_method per site.
template <typename T, int SETS>
__global__ __launch_bounds__(kBlock) void k_fstrength(Geo g, const T* __restrict__ U, double* __restrict__ partial,
                                                      double* __restrict__ dens, int dens_method) {
    constexpr int NS = ((SETS >> 0) & 1) + ((SETS >> 1) & 1) + ((SETS >> 2) & 1);
    constexpr int IP = 0, IC = (SETS & 1), IR = (SETS & 1) + ((SETS >> 1) & 1);  // slot of each set
    __shared__ double cf[NS][6][8][64];
    const SiteMu sm = site_mu_index(g);
    const int lane = (int)(threadIdx.x & 63u);
    const bool valid = sm.s < g.V;
    const unsigned V = g.V;
    if (valid) {
        int x[4];
        decode(g, sm.s, x);
        // with clover and rectangle sets (a rectangle loop is 6 links, a clover leaf 4): waves
        // 0..2 evaluate the rectangles of two planes each and wave 3 the clovers (and
        // plaquettes) of all six, 96 link steps per wave instead of 128 / 64 (32^4 improved
        // charge 2.45 -> 2.33 ms: the other workgroups on a CU already fill most of the idle
        // waves' time); otherwise wave w takes planes w, w + 4
        constexpr bool SPLIT = (SETS & 2) && (SETS & 4);
        const int pl0 = SPLIT ? (sm.mu < 3 ? 2 * sm.mu : 0) : sm.mu;
        const int pl1 = SPLIT ? (sm.mu < 3 ? 2 * sm.mu + 2 : 6) : 6;
        const int dpl = SPLIT ? 1 : 4;
        const bool do_cl = !SPLIT || sm.mu == 3, do_rect = !SPLIT || sm.mu < 3;
#pragma unroll 1
        for (int pl = pl0; pl < pl1; pl += dpl) {
            const int mu = pl < 3 ? 0 : (pl < 5 ? 1 : 2);
            const int nu = pl < 3 ? pl + 1 : (pl < 5 ? pl - 1 : 3);
            double c[8];
            if ((SETS & 3) && do_cl) {
                // clover leaf 0 is the plaquette
                M3<double> acc = walk(g, U, V, x, mu, nu, kClover[0]);
                if (SETS & 1) {
                    ta_coeffs(acc, c);
#pragma unroll
                    for (int a = 0; a < 8; a++) cf[IP][pl][a][lane] = c[a];
                }
                if (SETS & 2) {
#pragma unroll 1
                    for (int k = 1; k < 4; k++) {
                        const M3<double> l = walk(g, U, V, x, mu, nu, kClover[k]);
#pragma unroll
                        for (int e = 0; e < 9; e++) {
                            acc.m[e].r += l.m[e].r;
                            acc.m[e].i += l.m[e].i;
                        }
                    }
                    ta_coeffs(acc, c);
#pragma unroll
                    for (int a = 0; a < 8; a++) cf[IC][pl][a][lane] = c[a];
                }
            }
            if ((SETS & 4) && do_rect) {
                M3<double> acc = zero3<double>();
#pragma unroll 1
                for (int k = 0; k < 8; k++) {
                    const M3<double> l = walk(g, U, V, x, mu, nu, kRect[k]);
#pragma unroll
                    for (int e = 0; e < 9; e++) {
                        acc.m[e].r += l.m[e].r;
                        acc.m[e].i += l.m[e].i;
                    }
                }
                ta_coeffs(acc, c);
#pragma unroll
                for (int a = 0; a < 8; a++) cf[IR][pl][a][lane] = c[a];
            }
        }
    }
    __syncthreads();
    double e = 0.0, qp = 0.0, qc = 0.0, qr = 0.0;
    if (sm.mu == 0 && valid) {
        const double k8 = 1.0 / (8.0 * 3.14159265358979323846 * 3.14159265358979323846);
        if (SETS & 1) qp = k8 * pair_sum(cf[IP], lane);
        if (SETS & 2) {
            qc = k8 / 16.0 * pair_sum(cf[IC], lane);
#pragma unroll
            for (int p = 0; p < 6; p++)
#pragma unroll
                for (int a = 0; a < 8; a++) e = fma(cf[IC][p][a][lane], cf[IC][p][a][lane], e);
            e *= 1.0 / 32.0;
        }
        if (SETS & 4) qr = 2.0 * k8 / 64.0 * pair_sum(cf[IR], lane);
        if (dens) {
            double q = qp;
            if (dens_method == 1) q = qc;
            if (dens_method == 2) q = (5.0 / 3.0) * qc + (-1.0 / 12.0) * qr;
            if (dens_method == 3) q = qr;
            dens[sm.s] = q;
        }
    }
    const unsigned nb = gridDim.x;
    block_sum_store<kBlock>(e, partial);
    __syncthreads();
    block_sum_store<kBlock>(qp, partial + nb);
    __syncthreads();
    block_sum_store<kBlock>(qc, partial + 2 * nb);
    __syncthreads();
    block_sum_store<kBlock>(qr, partial + 3 * nb);
}

template <typename T>
hipError_t launch_field_strength(const Geo& g, const T* U, int sets, double* partial, double* out4, double* dens,
                                 int dens_method, hipStream_t st) {
    const int nb = grid_sites_mu(g.V);
    switch (sets) {
        case 1:
            k_fstrength<T, 1><<<nb, kBlock, 0, st>>>(g, U, partial, dens, dens_method);
            break;
        case 2:
            k_fstrength<T, 2><<<nb, kBlock, 0, st>>>(g, U, partial, dens, dens_method);
            break;
        case 4:
            k_fstrength<T, 4><<<nb, kBlock, 0, st>>>(g, U, partial, dens, dens_method);
            break;
        case 6:
            k_fstrength<T, 6><<<nb, kBlock, 0, st>>>(g, U, partial, dens, dens_method);
            break;
        case 7:
            k_fstrength<T, 7><<<nb, kBlock, 0, st>>>(g, U, partial, dens, dens_method);
            break;
        default:
            return hipErrorInvalidValue;
    }
    hipError_t err = hipGetLastError();
    for (int k = 0; k < 4 && err == hipSuccess; k++) err = launch_sum_partials(partial + (size_t)k * nb, nb, out4 + k, st);
    return err;
}

#define GF_INST_FS(T)                                                                                     \
    template hipError_t launch_field_strength<T>(const Geo&, const T*, int, double*, double*, double*, int, \
                                                 hipStream_t);
GF_INST_FS(double)
GF_INST_FS(float)
#undef GF_INST_FS

}  // namespace gf
