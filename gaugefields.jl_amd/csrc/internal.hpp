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
enum ForceMode { FORCE_STORE = 0, FORCE_KICK = 1, FORCE_KICK_EXP = 2 };

// ---- k_gauge.hip ----
template <typename T>
hipError_t launch_plaquette(const Geo& g, const T* U, double* partial, double* out, hipStream_t st);
template <typename T>
hipError_t launch_momenta_dot(const Geo& g, const T* P, double* partial, double* out, hipStream_t st);
template <typename T>
hipError_t launch_dsdu(const Geo& g, const T* U, double coeff, T* D, hipStream_t st);
// MODE FORCE_STORE: F = -(coeff/3) TA(U Sigma)
// MODE FORCE_KICK : P += eps_p * F
// MODE FORCE_KICK_EXP : P += eps_p * F;  Unew = exp(eps_u P) U
template <typename T>
hipError_t launch_force(int mode, const Geo& g, const T* U, T* FP, T* Unew, double coeff, double eps_p,
                        double eps_u, hipStream_t st);
template <typename T>
hipError_t launch_exp_update(const Geo& g, const T* Uin, T* Uout, const T* P, double eps, hipStream_t st);
template <typename T>
hipError_t launch_force_from_deriv(int mode, const Geo& g, const T* U, const T* D, T* FP, double eps,
                                   hipStream_t st);
template <typename T>
hipError_t launch_links_from_ref(const double* ref, T* soa, int V, hipStream_t st);
template <typename T>
hipError_t launch_links_to_ref(const T* soa, double* ref, int V, hipStream_t st);
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

// ---- k_flow.hip ----
// A = alpha * TA(W V^dag) + beta * A ; Wout = exp(A) W
template <typename T>
hipError_t launch_flow_stage(const Geo& g, const T* W, T* Wout, T* A, double alpha, double beta,
                             hipStream_t st);

// ---- k_stout.hip ----
template <typename T>
hipError_t launch_stout_forward(const Geo& g, const T* Uin, T* Uout, T* Cst, double rho, hipStream_t st);
template <typename T>
hipError_t launch_stout_backward(const Geo& g, const T* Uin, const T* Cst, const T* Dout, T* Din, T* G,
                                 double rho, hipStream_t st);

}  // namespace gf
