/*
 * gf_oracle.h -- CPU FP64 restatement of Gaugefields.jl's SU(3) gauge-link hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle that the HIP library is
 * compared against.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py may load it.  The product path (libgfhip) never links, loads or calls
 * anything in oracle/.
 *
 * Field layout: exactly the reference's Julia layout, so that the oracle reads and
 * writes what the reference's arrays hold.
 *   links   : double[4][V][3][3][2]  -- per direction mu (1..4 -> 0..3), the Julia
 *             Array{ComplexF64}(3,3,NX,NY,NZ,NT) in column-major order: element
 *             (i,j,x,y,z,t) at complex offset i + 3*j + 9*s, s = x + NX*(y + NY*(z + NZ*t)),
 *             real part first.  (gaugefields_4D_nowing.jl:18-51, interior only.)
 *   momenta : double[4][V][8]        -- TA_Gaugefields_4D_serial a[k,x,y,z,t], k fastest
 *             (TA_gaugefields_4D_serial.jl:1-30).  Coefficients of sum_a p_a (i lambda_a / 2).
 * dims = {NX, NY, NZ, NT}, periodic boundaries.
 */
#ifndef GF_ORACLE_H
#define GF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* thread count used by the OpenMP loops (0 = library default) */
void gfo_set_threads(int n);
int gfo_get_threads(void);

/* --- single-matrix primitives (3x3 complex, column-major, 18 doubles) --- */
void gfo_ta_coeffs(const double *m, double *c8);                    /* TA projection to 8 coeffs */
void gfo_ta_matrix(const double *m, double *out);                   /* TA projection as a matrix */
void gfo_exp_ta(const double *c8, double t, double *out);           /* exp(t sum c_a i lambda_a/2) */
void gfo_exp_antiherm(const double *q, double *out);                /* exp(Q), Q anti-Hermitian traceless */
void gfo_cdexpqdq(const double *c, const double *q, double *out);  /* C star d exp(Q)/dQ */

/* --- whole-lattice operations --- */
double gfo_plaquette_sum(const int *dims, const double *U);
void gfo_dsdu(const int *dims, const double *U, double coeff, double *out);
void gfo_force(const int *dims, const double *U, double coeff, double *F);
void gfo_kick(const int *dims, double *P, const double *F, double eps);
void gfo_exp_update(const int *dims, double *U, const double *P, double eps);
double gfo_momenta_dot(const int *dims, const double *P);
double gfo_potential(const int *dims, const double *U, double coeff);
/* update_momenta!(eps_p) then update_gaugefields!(eps_u) in the reference's pass structure
 * (whole-lattice passes through full temporaries; the CPU baseline's reference-structured form) */
void gfo_md_update_unfused(const int *dims, double *U, double *P, double coeff, double eps_p, double eps_u);

void gfo_flow_force(const int *dims, const double *U, double *F);
void gfo_flow(const int *dims, double *U, double eps, int nsteps, int every, double *E_out);
double gfo_energy_density(const int *dims, const double *U);

void gfo_stout_forward(const int *dims, const double *Uin, double rho, double *Uout, double *C);
void gfo_stout_backward(const int *dims, const double *Uin, double rho, const double *Dout,
                        double *Din);
void gfo_stout_smear(const int *dims, const double *U, int nlev, double rho, double *Uout);
void gfo_stout_force(const int *dims, const double *U, int nlev, double rho, double coeff,
                     double *F);
double gfo_stout_potential(const int *dims, const double *U, int nlev, double rho, double coeff);

/* topological charge methods (topological_charge(U; method), AbstractGaugefields.jl:1430-1471)
 * plus the rectangle density the improved one is built from */
enum { GFO_TOPO_PLAQUETTE = 0, GFO_TOPO_CLOVER = 1, GFO_TOPO_IMPROVED = 2, GFO_TOPO_RECTANGLE = 3 };
/* density q[V] (may be NULL); returns Q = sum q */
double gfo_topological_charge_density(const int *dims, const double *U, int method, double *q);
/* clover energy density (samples/measurements/energydensity.jl:32-76) */
double gfo_energy_density_clover(const int *dims, const double *U);

/* counter-based random fields: Philox4x32-10 and the per-link streams built on it */
void gfo_philox4x32_10(const uint32_t *ctr, const uint32_t *key, uint32_t *out);
void gfo_gaussian_momenta(const int *dims, double sigma, uint64_t seed, uint64_t sweep, double *P);
void gfo_hot_links(const int *dims, uint64_t seed, double *U);

/* heatbath (mode 0) / overrelaxation (mode 1) sweep over mu and checkerboard colours;
 * returns the number of heatbath links left unchanged (no acceptance within itmax tries) */
long gfo_heatbath_sweep(const int *dims, double *U, double beta, uint64_t seed, uint64_t sweep, int itmax,
                        int mode);
/* Kennedy-Pendleton SU(2) samples: a0 of n elements at coupling k */
void gfo_kp_samples(double k, long n, uint64_t seed, double *a0);

/* integrator: 0 = QPQ, 1 = PQP, 2 = Omelyan 2MN (lambda argument), 3 = 2MN with the adjacent
 * kicks of consecutive steps merged into one (the same map, one force evaluation fewer per step) */
void gfo_md_trajectory(const int *dims, double *U, double *P, double coeff, int nlev,
                       double rho, int integrator, double lambda, int steps, double tau,
                       double *H);

#ifdef __cplusplus
}
#endif
#endif
