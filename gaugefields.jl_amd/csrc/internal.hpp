// internal.hpp -- host-side declarations shared by the kernel translation units and the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <type_traits>

#include "lattice.hpp"

namespace gf {

constexpr int kBlock = 256;

// FP32-bulk kernels re-project every link they update onto SU(3)
template <typename T>
inline constexpr bool kFloatBulk = std::is_same<T, float>::value;

inline int grid_for(int V, int block = kBlock) { return (V + block - 1) / block; }
// grid of the (site, direction) kernels: one 256-thread workgroup per 64 sites
inline int grid_sites_mu(int V) { return (V + 63) / 64; }

// force kernel modes
enum ForceMode { FORCE_STORE = 0, FORCE_KICK = 1, FORCE_KICK_EXP = 2, FORCE_KICK_S64 = 3, FORCE_KICK_EXP_S64 = 4 };
// FORCE_KICK_S64 / FORCE_KICK_EXP_S64 (FP32 kernels over an FP64 MD state): the FP32 force kicks the
// FP64 momenta, and the FP64 links are updated in place, their float rounding written to Unew

// ---- k_gauge.hip ----
// dst = src, bytes a multiple of 16 (streaming copy: the measured HBM ceiling)
hipError_t launch_stream_copy(const void* src, void* dst, size_t bytes, hipStream_t st);
// FP64 link rows 0-1 (optionally SU(3)-projected first) and/or momenta rounded to float in place
hipError_t launch_round_f32(double* U, double* P, int V, bool links, bool project, bool momenta, hipStream_t st);
// out[0] = sum of partial[0..n) (one workgroup, fixed order)
hipError_t launch_sum_partials(const double* partial, int n, double* out, hipStream_t st);
template <typename T>
hipError_t launch_plaquette(const Geo& g, const T* U, double* partial, double* out, hipStream_t st);
template <typename T>
hipError_t launch_action_untraced(const Geo& g, const T* U, double coeff, double* ref, hipStream_t st);
template <typename T>
hipError_t launch_momenta_dot(const Geo& g, const T* P, double* partial, double* out, hipStream_t st);
template <typename T>
hipError_t launch_dsdu(const Geo& g, const T* U, double coeff, T* D, hipStream_t st);
// MODE FORCE_STORE: F = -(coeff/3) TA(U Sigma)
// MODE FORCE_KICK : P += eps_p * F
// MODE FORCE_KICK_EXP : P += eps_p * F;  Unew = exp(eps_u P) U
// MODE FORCE_KICK_S64, FORCE_KICK_EXP_S64 (float): the same on the FP64 momenta (FP) and links
//   (U64) of an FP64 state, with Unew the updated links rounded to float
template <typename T>
hipError_t launch_force(int mode, const Geo& g, const T* U, T* FP, T* Unew, double coeff, double eps_p,
                        double eps_u, hipStream_t st, double* U64 = nullptr);
template <typename T>
hipError_t launch_exp_update(const Geo& g, const T* Uin, T* Uout, const T* P, double eps, hipStream_t st);
// reference-layout (column-major 3x3 per link) <-> device field; general = false: SU(3) link
// field (rows 0 and 1 kept, row 2 rebuilt on the way out), true: general matrix field
hipError_t launch_links_widen(const float* Uf, double* Ud, int V, hipStream_t st);
// FP32 forces over an FP64 MD state (k_state64_step): mode 1 exp update of U64 (and its float
// rounding into U32), 3 rounding only
hipError_t launch_state64_step(int mode, const Geo& g, const double* P64, double* U64, float* U32, double eps_u,
                               hipStream_t st);
template <typename T>
hipError_t launch_links_from_ref(const double* ref, T* soa, int V, bool general, hipStream_t st);
template <typename T>
hipError_t launch_links_to_ref(const T* soa, double* ref, int V, bool general, hipStream_t st);
template <typename T>
hipError_t launch_momenta_from_ref(const double* ref, T* soa, int V, hipStream_t st);
template <typename T>
hipError_t launch_momenta_to_ref(const T* soa, double* ref, int V, hipStream_t st);
template <typename T>
hipError_t launch_set_identity(T* U, int V, hipStream_t st);
template <typename T>
hipError_t launch_fill(T* p, size_t n, T v, hipStream_t st);
template <typename T>
hipError_t launch_axpby(T* y, const T* x, double a, double b, size_t n, hipStream_t st);

// ---- k_fstrength.hip ----
// Field strengths of the loop sets in `sets` (bit 0 plaquette, 1 clover, 2 rectangle); out4 =
// {clover energy sum, plaquette Q, clover Q, rectangle Q} (device doubles); dens (device, V) if
// non-null receives the density of dens_method (0 plaquette, 1 clover, 2 improved, 3 rectangle).
template <typename T>
hipError_t launch_field_strength(const Geo& g, const T* U, int sets, double* partial, double* out4, double* dens,
                                 int dens_method, hipStream_t st);

// ---- k_random.hip ----
template <typename T>
hipError_t launch_gaussian_momenta(const Geo& g, T* P, double sigma, unsigned long long seed,
                                   unsigned long long sweep, hipStream_t st);
template <typename T>
hipError_t launch_hot_links(const Geo& g, T* U, unsigned long long seed, hipStream_t st);

// heatbath (mode 0) or overrelaxation (mode 1) sweep: mu = 0..3, colours even then odd;
// failures (device double) = links a heatbath left unchanged.  partial: >= 8 * (V/2 + 255)/256
template <typename T>
hipError_t launch_heatbath(const Geo& g, T* U, int mode, double beta, unsigned long long seed,
                           unsigned long long sweep, int itmax, double* partial, double* failures, hipStream_t st);

// ---- k_flow.hip ----
// A = alpha * TA(W V^dag) + beta * A ; Wout = exp(A) W
template <typename T>
hipError_t launch_flow_stage(const Geo& g, const T* W, T* Wout, T* A, double alpha, double beta,
                             hipStream_t st);

// ---- k_stout.hip ----
template <typename T>
hipError_t launch_stout_forward(const Geo& g, const T* Uin, T* Uout, T* Cst, double rho, hipStream_t st);
template <typename T>
// Lc: scratch momentum-layout field (the 8 su(3) coefficients of Lam per link).
// fmode < 0: Din = dS/dU_in.  fmode = FORCE_STORE / FORCE_KICK / FORCE_KICK_EXP (the thin-link
// level of a stout force, test/HMCstout_test_nowing.jl:99-118): Din is not stored; FP = -(1/3)
// TA(U_in Din), or FP += eps of that, and with FORCE_KICK_EXP also Unew = exp(eps_u FP) U_in.
hipError_t launch_stout_backward(const Geo& g, const T* Uin, const T* Cst, const T* Dout, T* Din, T* Lc,
                                 double rho, hipStream_t st, int fmode = -1, T* FP = nullptr, double eps = 0.0,
                                 T* Unew = nullptr, double eps_u = 0.0, double* U64 = nullptr);
// the site-local half of the stout backward alone (Cst null: C recomputed)
template <typename T>
hipError_t launch_stout_back_local(const Geo& g, const T* Uin, const T* Cst, const T* Dout, T* Din, T* Lc, double rho,
                                   hipStream_t st);
// the stencil half of the stout backward alone (fmode as above)
template <typename T>
hipError_t launch_stout_back_stencil(const Geo& g, const T* Uin, const T* Lc, T* Din, double rho, hipStream_t st,
                                     int fmode, T* FP, double eps, T* Unew, double eps_u, double* U64 = nullptr);
// stored-C mode: the stencil half of level l (Uin, Lc; Din in: its site-local part) chained with
// the site-local half of level l-1 (Uprev, Cprev): Din out = level l-1's site-local part, Lout =
// Lam_{l-1} (a buffer other than Lc)
template <typename T>
hipError_t launch_stout_back_chain(const Geo& g, const T* Uin, const T* Lc, T* Din, double rho, const T* Uprev,
                                   const T* Cprev, T* Lout, hipStream_t st);
// dS/dU of the top level (Utop, coeff) chained with the top level's site-local half (Uin, Cst)
template <typename T>
hipError_t launch_dsdu_chain(const Geo& g, const T* Utop, double coeff, const T* Uin, const T* Cst, T* Din, T* Lout,
                             hipStream_t st);

}  // namespace gf
