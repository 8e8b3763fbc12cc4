/*
 * gfhip.h -- C ABI of libgfhip.so, the MI355X (gfx950) backend for Gaugefields.jl's SU(3)
 * gauge-link hot path.
 *
 * One gf_ctx holds one 4D periodic SU(3) gauge configuration (the reference's
 * Vector of four Gaugefields_4D fields) together with its conjugate momenta, force and
 * work fields, all resident in HBM in a structure-of-arrays layout.  Host arrays cross the
 * boundary in the reference's own Julia layouts (FP64, interior sites only):
 *   links   : double[4][V][9][2]  direction mu, site s = x + NX*(y + NY*(z + NZ*t)),
 *             element (i,j) of the 3x3 matrix at complex offset i + 3*j (column-major),
 *             real part first -- Array{ComplexF64}(3,3,NX,NY,NZ,NT) per direction
 *             (gaugefields_4D_nowing.jl:18-51; LatticeMatrix interior,
 *             gaugefields_4D_MPILattice.jl:34-121).
 *   momenta : double[4][V][8]     coefficients p_a of sum_a p_a (i lambda_a / 2), Gell-Mann
 *             order of TA_gaugefields_4D_serial.jl:243-260 (TA_Gaugefields_4D_serial a[k,...]).
 * Host buffers are borrowed for the duration of the call only.
 *
 * Every call returns GF_OK (0) or a negative status; gf_last_error() describes the last
 * failure.  A context owns one HIP stream; calls on it are stream-ordered and not
 * thread-safe.  Calls that return host scalars synchronize the stream.
 *
 * Precision: GF_F64 stores and computes links/momenta in double.  GF_F32 ("FP32 bulk")
 * stores and computes them in float while every global sum (plaquette, action, p*p,
 * Hamiltonian) and every Cayley-Hamilton coefficient is evaluated in double.  GF_F32_FORCE
 * ("FP32 forces") holds and updates the links and momenta in double (the MD state, the link
 * updates and the Hamiltonian as in GF_F64) and evaluates every force -- staples, stout chain,
 * TA projection -- in float from the links rounded to float (as GF_F32 does); the flow, the
 * heatbath, the field-strength measurements and the explicit stout smearing/backward calls are
 * not available in this form (GF_ESTATE).
 */
#ifndef GFHIP_H
#define GFHIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gf_ctx gf_ctx;

enum { GF_F64 = 0, GF_F32 = 1, GF_F32_FORCE = 2 };
enum { GF_OK = 0, GF_EINVAL = -1, GF_EDEVICE = -2, GF_ENOMEM = -3, GF_ESTATE = -4 };
enum { GF_QPQ = 0, GF_PQP = 1, GF_OMELYAN2MN = 2, GF_FUSED = 0x100 };
enum { GF_STOUT_STORE = 0, GF_STOUT_RECOMPUTE = 1, GF_STOUT_AUTO = 2 };

/* ---- lifecycle ------------------------------------------------------------------- */
int gf_version(void);
/* gauge_configuration (API.jl:178-251) minus the random start: allocates a configuration
 * of extents dims[4] = {NX,NY,NZ,NT} on HIP device `device`, links set to the identity. */
int gf_create(const int *dims, int precision, int device, gf_ctx **out);
int gf_destroy(gf_ctx *ctx);
const char *gf_last_error(const gf_ctx *ctx);
int gf_sync(gf_ctx *ctx);
/* extents, precision and device bytes held by the context */
int gf_info(const gf_ctx *ctx, int *dims, int *precision, size_t *device_bytes);

/* ---- fields ---------------------------------------------------------------------- */
/* cold start: IdentityGauges (AbstractGaugefields.jl:431-548 condition="cold") */
int gf_set_cold(gf_ctx *ctx);
/* substitute_U!(U, host) / copy back (gaugefields_4D_MPILattice.jl substitute_U!) */
int gf_upload_links(gf_ctx *ctx, const double *ref);
int gf_download_links(gf_ctx *ctx, double *ref);
/* momenta P (TA_gaugefields_4D_MPILattice.jl:4-72 storage) */
int gf_upload_momenta(gf_ctx *ctx, const double *mom);
int gf_download_momenta(gf_ctx *ctx, double *mom);
/* clear_U!(P) (TA_gaugefields_4D_MPILattice.jl clear_U!) */
int gf_clear_momenta(gf_ctx *ctx);
/* the force field F written by gf_force / gf_stout_force (md_force! output) */
int gf_download_force(gf_ctx *ctx, double *force);
/* substitute_U!(Uold, U) / substitute_U!(U, Uold): the Metropolis backup slot
 * (test/HMCstout_test_nowing.jl:47-74) */
int gf_save_links(gf_ctx *ctx);
int gf_restore_links(gf_ctx *ctx);

/* ---- random fields on the device (no counterpart layout in the reference: the generator
 * is Philox4x32-10, Salmon et al. SC'11, one counter per link, so results do not depend on
 * the launch shape; key = seed, counter = (site, tag << 24 | mu << 16 | block, sweep)) ---- */
/* gaussian_momenta!(p; sigma) (API.jl:341-392, gauss_distribution!
 * TA_gaugefields_4D_MPILattice.jl:157-193): every p_a ~ N(0, sigma^2) (Box-Muller), stream
 * (seed, sweep) */
int gf_gaussian_momenta(gf_ctx *ctx, double sigma, unsigned long long seed, unsigned long long sweep);
/* gauge_configuration(...; start=:hot) (API.jl:178-251; hot fill gaugefields_4D_wing.jl:295-322
 * + normalize3!, AbstractGaugefields.jl:3378-3457): entries (u - 1/2) + i (u' - 1/2), rows
 * Gram-Schmidt normalized, third row conj(row1 x row2) */
int gf_hot_start(gf_ctx *ctx, unsigned long long seed);

/* heatbath!(U, temps, beta; seed, sweep, ITERATION_MAX) (heatbath/heatbathmodule.jl:2119-2163):
 * one sweep of the plaquette-action SU(3) heatbath, mu = 1..4, even then odd sites, each link by
 * the Cabibbo-Marinari subgroups (1,2), (2,3), (1,3) with Kennedy-Pendleton SU(2) steps
 * (heatbath/portable/kernels.jl:63-300) at weight exp(beta/3 Re tr(U V)), V the staple sum
 * (loops_staple; autostaples/wilsonloops.jl:395-412).  Streams (seed, sweep), tags 3-5; a link
 * with a subgroup not accepted within itmax tries is left unchanged and counted in *failures
 * (the try index is a 16-bit counter field: itmax above 65534 acts as 65534).  All extents
 * must be even. */
int gf_heatbath(gf_ctx *ctx, double beta, unsigned long long seed, unsigned long long sweep, int itmax,
                long long *failures);
/* overrelaxation!(U, ...) (heatbathmodule.jl:2183-2276, SUN_overrelaxation_rng! :1295-1321): same
 * sweep order, three SU(2) reflections per link (subgroup chosen from stream (seed, sweep), tag
 * 6), keeping the action unchanged */
int gf_overrelaxation(gf_ctx *ctx, unsigned long long seed, unsigned long long sweep);

/* ---- measurements ---------------------------------------------------------------- */
/* calculate_Plaquette (AbstractGaugefields.jl:2684-2699): sum_{x,mu<nu} Re tr P_munu.
 * measure_plaquette (API.jl:395-404) divides by 6*V*3. */
int gf_plaquette(gf_ctx *ctx, double *sum);
/* gf_plaquette for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc); no counterpart in the reference.
 * links: device pointer to an SU(3) link field of extents dims in the device layout -- FP64:
 * double[4][6][V][2], plane 6 mu + 3 i + j holding (re, im) of U_mu(x)_ij for rows i = 0, 1;
 * FP32: float[4][3][V][4], quad plane 3 mu + q holding elements 2q, 2q+1 (element 3 i + j)
 * of every site -- site s = x + NX (y + NY (z + NZ t)).  sum: device pointer; receives
 * sum_{x,mu<nu} Re tr P_munu (as gf_plaquette) in stream order on `stream` (a hipStream_t,
 * NULL for the default stream); the call does not synchronize.  No hipMalloc, hipFree or other
 * allocation: the per-workgroup partial sums live in a static 4 MiB device array that each GPU
 * holds until the process ends, so V <= 2^25 sites (larger: GF_EINVAL).  Calls on one GPU
 * share that array: calls that overlap on one GPU (other streams, other host threads) race.
 * Errors are reported by gf_last_error(NULL) on the calling thread. */
int gf_plaquette_links(const void *links, const int *dims, int precision, void *stream, double *sum);
/* real(evaluate_GaugeAction) (GaugeActions.jl:132-142) of push!(action, coeff,
 * plaquette ∪ plaquette^dag): 2*coeff*sum Re tr P.  md_potential = -value/3. */
int gf_action(gf_ctx *ctx, double coeff, double *value);
/* evaluate_GaugeAction_untraced! (GaugeActions.jl:190-213) for the loop set plaquette u
 * plaquette^dag: coeff sum_{mu<nu} (P_munu(x) + P_munu(x)^dag) at every site, into ref as one
 * 3x3 complex matrix per site in the reference's layout (column-major, sites x fastest; 9 V
 * complex).  Its trace summed over sites is gf_action's value. */
int gf_action_untraced(gf_ctx *ctx, double coeff, double *ref);
/* p*p (TA_gaugefields_4D_MPILattice.jl:147-155): sum of all p_a^2 */
int gf_momenta_dot(gf_ctx *ctx, double *value);
/* md_hamiltonian (molecular_dynamics.jl:494-505): md_potential + p*p/2, on the stout-smeared
 * links when gf_stout_config set nlev > 0 */
int gf_hamiltonian(gf_ctx *ctx, double coeff, double *H);

/* ---- derivative, force, MD updates ------------------------------------------------ */
/* calc_dSdUμ! (GaugeActions.jl:95-123) for all four mu; written to ref (link layout) if
 * ref != NULL */
int gf_dsdu(gf_ctx *ctx, double coeff, double *ref);
/* md_force!(F, ::GaugeAction, U) (molecular_dynamics.jl:251-267): F = -(1/3) TA(U dSdU) */
int gf_force(gf_ctx *ctx, double coeff);
/* update_momenta! (molecular_dynamics.jl:539-551): P += eps * F(U), fused in one kernel */
int gf_force_kick(gf_ctx *ctx, double coeff, double eps);
/* update_gaugefields! (molecular_dynamics.jl:513-531): U <- exp(eps P) U */
int gf_exp_update(gf_ctx *ctx, double eps);
/* update_momenta!(eps_p) followed by update_gaugefields!(eps_u) in one kernel
 * (force + kick + exp; the new links go to a second buffer that becomes current) */
int gf_md_update(gf_ctx *ctx, double coeff, double eps_p, double eps_u);
/* md_trajectory! (molecular_dynamics.jl:712-730) with md_step! QPQ/PQP (:604-616) or the
 * Omelyan 2MN step of docs/src/howtoimplement.md:94-112 (integrator | GF_FUSED merges
 * adjacent half steps and uses gf_md_update).  Force: plain Wilson action, or the
 * stout-smeared one when gf_stout_config set nlev > 0.  H[3] = {H0, H1, H1 - H0}. */
int gf_md_trajectory(gf_ctx *ctx, double coeff, int integrator, double lambda, int steps, double tau,
                     double *H);

/* ---- Wilson flow ------------------------------------------------------------------ */
/* flow!(U, ::Gradientflow) (gradientflow.jl:171-238): nsteps Luscher RK3 steps of size eps.
 * If every > 0, E_out[k] = energy density after step (k+1)*every. */
int gf_flow(gf_ctx *ctx, double eps, int nsteps, int every, double *E_out);
/* plaquette energy density E = 2 sum_{mu<nu} Re tr(1 - P_munu) / V = 36 (1 - plaquette)
 * (Luscher JHEP 08 (2010) 071; no built-in counterpart in the reference) */
int gf_energy_density(gf_ctx *ctx, double *E);

/* ---- field strengths: topological charge, clover energy density ------------------- */
enum { GF_TOPO_PLAQUETTE = 0, GF_TOPO_CLOVER = 1, GF_TOPO_IMPROVED = 2, GF_TOPO_RECTANGLE = 3 };
/* topological_charge(U; method) (AbstractGaugefields.jl:1452-1471): Q = sum_x q(x) with
 * q = -norm Re sum eps_{mu nu rho sigma} tr(F_munu F_rhosigma) / (32 pi^2 n^2), F_munu = TA of
 * the sum of the method's n loops at x (plaquette n = 1; clover leaves n = 4; 2x1 rectangles
 * n = 8, norm 2); improved = 5/3 clover - 1/12 rectangle (:1195-1397).  GF_TOPO_RECTANGLE is
 * the reference's internal _rectangle_topological_charge. */
int gf_topological_charge(gf_ctx *ctx, int method, double *Q);
/* topological_charge_density(U; method) (AbstractGaugefields.jl:1430-1446): density[V] on the
 * host, site order x + NX (y + NY (z + NZ t)) */
int gf_topological_charge_density(gf_ctx *ctx, int method, double *density);
/* calculate_energy_density (samples/measurements/energydensity.jl:32-76): clover E =
 * sum_{mu != nu} -tr(F_munu F_munu)/2 / (16 V) with F = TA(sum of the four clover leaves) */
int gf_energy_density_clover(gf_ctx *ctx, double *E);
/* flow!(U, g) steps with the measurements of samples/measurements/topologicalcharge.jl:236-245
 * and energydensity.jl:100-118: after every `every` steps (nsteps / every rows) out receives
 * GF_FLOW_NOBS doubles {t, plaquette / (18 V), E plaquette (as gf_energy_density), E clover,
 * Q plaquette, Q clover, Q improved}. */
enum { GF_FLOW_NOBS = 7 };
int gf_flow_measure(gf_ctx *ctx, double eps, int nsteps, int every, double *out);

/* ---- stout smearing ---------------------------------------------------------------- */
/* nlev stout layers at rho (STOUT_Layer + CovNeuralnet, stout_fast.jl:6-122,
 * Abstractsmearing.jl:61-99); mode GF_STOUT_STORE keeps C_mu per level, GF_STOUT_RECOMPUTE
 * recomputes it in the backward pass, GF_STOUT_AUTO stores when the device has room for the
 * nlev extra link fields (plus a 10 % margin) and recomputes otherwise.  The smeared levels
 * themselves are always kept.  nlev = 0 disables smearing. */
int gf_stout_config(gf_ctx *ctx, int nlev, double rho, int mode);
/* the memory mode in effect after gf_stout_config (GF_STOUT_STORE or GF_STOUT_RECOMPUTE) */
int gf_stout_mode(const gf_ctx *ctx, int *mode);
/* calc_smearedU (Abstractsmearing.jl:237-272): smeared links (all levels kept on device);
 * top level written to ref if ref != NULL */
int gf_stout_smear(gf_ctx *ctx, double *ref);
/* back_prop (Abstractsmearing.jl:352-411): given dS/dU_smeared (link layout) return
 * dS/dU_bare; uses the levels of the last gf_stout_smear */
int gf_stout_backward(gf_ctx *ctx, const double *dout_ref, double *din_ref);
/* stout MD force (test/HMCstout_test_nowing.jl:99-118): smear, calc_dSdUμ! on the smeared
 * links, back_prop, F = -(1/3) TA(U dSdU_bare) */
int gf_stout_force(gf_ctx *ctx, double coeff);
int gf_stout_force_kick(gf_ctx *ctx, double coeff, double eps);
/* the force of one more member of an MD action set added to the force field: F += F(coeff), on
 * the smeared links when levels are configured (md_force! of an MDActionSet: the first member's
 * force written, every further one into a work field and added, _md_add_named_forces!,
 * molecular_dynamics.jl:128-160) */
int gf_stout_force_add(gf_ctx *ctx, double coeff);

/* ---- ILDG / LIME configuration files (host only, no context) ----------------------- */
/* ILDG(filename) header parse (ildg_format.jl:686, read_header :1049-1053): extents lx..lt
 * and precision from the ildg-format record (zeros where the file has none) and the size of
 * the ildg-binary-data payload in bytes. */
int gf_ildg_info(const char *path, int *dims, int *precision_bits, long long *payload_bytes);
/* load_gaugefield! (ildg_format.jl:902-969) + load_binarydata! (:862-887): the SU(3)
 * configuration of extents dims into ref (link layout above).  Precision 32 or 64 bits from
 * the ildg-format record, else from the payload size; *precision_bits reports it. */
int gf_ildg_read(const char *path, const int *dims, double *ref, int *precision_bits);
/* save_binarydata (ildg_format.jl:797-833) + _save_binarydata (:697-746) + pack_ildg_file
 * (:50-65): an ildg-format record and the ildg-binary-data record, precision 32 or 64
 * (anything else: GF_EINVAL, the reference's ArgumentError). */
int gf_ildg_write(const char *path, const int *dims, const double *ref, int precision_bits);
/* message of the last failing gf_ildg_* call on this thread */
const char *gf_ildg_last_error(void);

/* ---- precision study ---------------------------------------------------------------- */
/* No counterpart in the reference: the storage rounding of the FP32-bulk form applied to an FP64
 * context, in place, so that an FP64 trajectory stepped from the host can carry exactly one of the
 * FP32 form's roundings.  GF_ROUND_LINKS: rows 0 and 1 of every link rounded to the nearest float
 * (with GF_ROUND_PROJECT: after the SU(3) projection in double that the FP32 link updates apply);
 * GF_ROUND_MOMENTA: every momentum coefficient.  GF_ESTATE on an FP32 context. */
enum { GF_ROUND_LINKS = 1, GF_ROUND_PROJECT = 2, GF_ROUND_MOMENTA = 4 };
int gf_round_to_float(gf_ctx *ctx, int fields);

/* ---- timing (HIP events on the context's stream) ----------------------------------- */
/* No counterpart in the reference: the device's streaming-copy rate, the measured HBM ceiling
 * quoted beside the 8 TB/s peak.  Copies the context's link field into its second link buffer
 * (the current links are unchanged) reps times after one untimed pass, timed with events
 * GF_NEVENTS-2 and GF_NEVENTS-1 (slots 14, 15) on the context's stream; *bytes_per_s = read +
 * written bytes per second. */
int gf_stream_copy(gf_ctx *ctx, int reps, double *bytes_per_s);
int gf_event_record(gf_ctx *ctx, int slot); /* slot 0..15 */
int gf_event_elapsed(gf_ctx *ctx, int slot_a, int slot_b, float *ms);

#ifdef __cplusplus
}
#endif
#endif
