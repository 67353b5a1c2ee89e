/*
 * su2_mi355x.h -- C ABI of the MI355X (gfx950) pure-gauge HMC molecular-dynamics
 * path of HISKP-LQCD/su2.
 *
 * One opaque handle is one lattice (U(1) or SU(2), d = 2, 3, 4, periodic) held
 * resident in HBM together with its momenta.  Every entry point returns an int
 * status (SU2GPU_OK = 0); on failure su2gpu_last_error() gives a message
 * (thread-local).  Calls on one handle are ordered on the handle's own HIP
 * stream and are not thread-safe per handle; entry points that return host
 * values synchronise that stream.
 *
 * Host arrays use the reference's AoS order (gaugeconfig.hh:211-222): link
 * (t,x,y,z,mu) at ((((t*Lx+x)*Ly+y)*Lz+z)*nd+mu).
 *   SU2GPU_LAYOUT_PACKED : SU(2) 4 doubles (Re a, Im a, Re b, Im b) per link,
 *                          U(1) 1 double (the phase) per link;
 *   SU2GPU_LAYOUT_OBJECT : the reference objects' own bytes, i.e.
 *                          gaugeconfig<_su2>::data() (40 B: size_t N_c = 2, then
 *                          complex a, complex b) or gaugeconfig<_u1>::data()
 *                          (16 B: size_t N_c = 1, then the phase) -- also the
 *                          on-disk format of gaugeconfig::save (gaugeconfig.hh:225-230).
 * Momenta: 3 doubles (a, b, c) per SU(2) link / 1 per U(1) link, the order of
 * adjointfield<double, G>::data (adjointfield.hh:95-198).
 */
#ifndef SU2_MI355X_H
#define SU2_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct su2gpu_lattice su2gpu_lattice;

enum {
  SU2GPU_OK = 0,
  SU2GPU_EINVAL = 1, /* bad argument */
  SU2GPU_EHIP = 2,   /* HIP runtime error (incl. no device) */
  SU2GPU_EIO = 3,    /* file could not be read / written */
  SU2GPU_ESTATE = 4  /* call not valid in the handle's current state */
};
enum { SU2GPU_U1 = 0, SU2GPU_SU2 = 1 };
enum { SU2GPU_FP64 = 0, SU2GPU_FP32 = 1 };
enum { SU2GPU_LAYOUT_PACKED = 0, SU2GPU_LAYOUT_OBJECT = 1 };
/* integrator ids: the reference's enum integrators (integrator.hh:15) */
enum {
  SU2GPU_LEAPFROG = 0,
  SU2GPU_LP_LEAPFROG = 1,
  SU2GPU_OMF4 = 2,
  SU2GPU_LP_OMF4 = 3,
  SU2GPU_EULER = 4,
  SU2GPU_RUTH = 5,
  SU2GPU_OMF2 = 6
};
/* where md_update takes its momenta from */
enum {
  SU2GPU_MOMENTA_HOST_MT19937 = 0, /* std::mt19937(seed) + std::normal_distribution, the
                                      reference's stream (adjointfield.hh:240-258) */
  SU2GPU_MOMENTA_DEVICE_PHILOX = 1 /* counter-based Philox4x32-10 on the GPU (throughput) */
};

/* Result of one HMC trajectory (md_params fields, md_params.hh:3-109, plus the
 * Hamiltonian pieces of the kinetic and gauge monomials, monomial.hh:6-47). */
typedef struct {
  double deltaH;      /* H_new - H_old                                (md_update.hh:47-51) */
  double deltadeltaH; /* reversibility violation, NaN if not tested   (md_update.hh:62-72) */
  int accept;         /* Metropolis decision                          (md_update.hh:54-59) */
  int revtest;        /* 1 if the reversibility test ran */
  double kin_old, gauge_old, kin_new, gauge_new; /* Hold / Hnew of the two monomials */
  double uniform;     /* the accept uniform drawn, -1 if deltaH <= 0 */
} su2gpu_md_result;

/* ---- errors, devices ------------------------------------------------------ */
const char *su2gpu_last_error(void);
int su2gpu_device_count(int *n);

/* ---- lattice handle ---------------------------------------------------------
 * su2gpu_create replaces gaugeconfig<G>(Lx, Ly, Lz, Lt, ndims, beta)
 * (gaugeconfig.hh:49-63) together with the gauge monomial's anisotropy
 * gaugemonomial(timescale, xi) (flat-gaugemonomial.hh:86-92): xi != 1 switches on
 * the anisotropic staples and plaquette weights.  For d < 4 pass Lz = 1 (and
 * Ly = 1 for d = 2), as parse_input_file.cc:57-67 does.  precision selects the
 * plain FP64 form or the FP32 form (FP32 links, staples, forces and momenta,
 * compensated per-site staple sums, reunitarisation every step; FP64 action /
 * kinetic sums and accept/reject).  The links start cold (unit). */
int su2gpu_create(su2gpu_lattice **out, int group, size_t Lx, size_t Ly, size_t Lz, size_t Lt,
                  size_t ndims, double beta, double xi, int precision);
int su2gpu_destroy(su2gpu_lattice *h);
/* gaugeconfig::getBeta / setBeta (gaugeconfig.hh:106-111) */
int su2gpu_set_beta(su2gpu_lattice *h, double beta);
int su2gpu_get_beta(const su2gpu_lattice *h, double *beta);
/* geometry: out[] = {Lx, Ly, Lz, Lt, ndims, volume, group, precision} */
int su2gpu_get_geometry(const su2gpu_lattice *h, size_t out[8]);

/* ---- fields ------------------------------------------------------------------
 * links: gaugeconfig data in, out (operator[], gaugeconfig.hh:179-185)          */
int su2gpu_set_links(su2gpu_lattice *h, const void *src, int layout);
int su2gpu_get_links(su2gpu_lattice *h, void *dst, int layout);
/* momenta: adjointfield data in, out */
int su2gpu_set_momenta(su2gpu_lattice *h, const double *src);
int su2gpu_get_momenta(su2gpu_lattice *h, double *dst);
/* adjointfield::flipsign (adjointfield.hh:121-125) */
int su2gpu_flip_momenta(su2gpu_lattice *h);
/* initnormal(std::mt19937(seed), momenta) (adjointfield.hh:240-258), host stream */
int su2gpu_initnormal(su2gpu_lattice *h, uint64_t seed);
/* FP32 storage emulation of an FP64 handle (no counterpart in the reference;
 * a numerical study in the manner of its rounding integrators lp_leapfrog /
 * round_and_update_momenta, integrator.hh:53-80, update_momenta.hh:35-59):
 * links != 0: every updated link, the start links and the restored final links
 * of md_trajectory / md_update rounded to the nearest FP32 value; momenta != 0:
 * every updated momentum and the start momenta likewise.  Arithmetic stays
 * FP64 and the generic MD kernels run (the specialised step does not round).
 * Both 0 (the default) switch it off.  EINVAL for an FP32 handle. */
int su2gpu_set_fp32_storage_emulation(su2gpu_lattice *h, int links, int momenta);

/* ---- start configurations (host RNG, the reference's exact draw order) ------ */
/* hotstart(U, seed, delta) (gaugeconfig.hh:277-302) */
int su2gpu_hotstart(su2gpu_lattice *h, uint64_t seed, double delta);
/* coldstart (gaugeconfig.hh:254-266) */
int su2gpu_coldstart(su2gpu_lattice *h);
/* random_gauge_trafo(U, seed) (random_gauge_trafo.hh:10-33) */
int su2gpu_random_gauge_trafo(su2gpu_lattice *h, uint64_t seed);
/* Haar-uniform start (no reference counterpart): normalised 4-vectors of
 * std::normal_distribution draws over std::mt19937(seed); U(1): uniform phase. */
int su2gpu_haarstart(su2gpu_lattice *h, uint64_t seed);
/* gaugeconfig::restoreSU (gaugeconfig.hh:115-120) */
int su2gpu_restoreSU(su2gpu_lattice *h);

/* ---- sums (FP64, deterministic, compensated) --------------------------------- */
/* retr_sum_Wplaquettes(U, xi, anisotropic, spatial_only) (flat-gauge_energy.hpp:31-94) */
int su2gpu_retr_sum_plaquettes(su2gpu_lattice *h, double xi, int anisotropic, int spatial_only,
                               double *res);
/* No reference counterpart; for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops).  The sum of su2gpu_retr_sum_plaquettes,
 * written to device memory (*sum_dev, one double) on the handle's stream,
 * without synchronising and without allocating: the block partials live in a
 * static device array of 2 MiB per GPU, shared by every call on that GPU, so
 * calls that overlap on one GPU race.  Lattices of up to 67,108,352 sites;
 * larger ones are rejected with SU2GPU_EINVAL. */
int su2gpu_retr_sum_plaquettes_noalloc(su2gpu_lattice *h, double xi, int anisotropic,
                                       int spatial_only, double *sum_dev);
/* gauge_energy(U, spatial_only) (flat-gauge_energy.hpp:109-137) */
int su2gpu_gauge_energy(su2gpu_lattice *h, int spatial_only, double *res);
/* get_S_G_hmc (flat-gaugemonomial.hh:46-54) = gaugemonomial heatbath/accept value */
int su2gpu_gauge_action(su2gpu_lattice *h, double *S);
/* kineticmonomial heatbath/accept: 0.5 * (P * P) (monomial.hh:33-47, adjointfield.hh:219-238) */
int su2gpu_kinetic_energy(su2gpu_lattice *h, double *K);

/* ---- MD building blocks ------------------------------------------------------- */
/* gaugemonomial::derivative(deriv, h, fac) (flat-gaugemonomial.hh:105-131):
 * deriv += fac * beta/Nc * get_deriv(U K).  deriv is a host adjointfield array
 * (in/out).  su2gpu_derivative_device does the same into the handle's device
 * deriv field, which su2gpu_zero_derivative clears (zeroadjointfield,
 * adjointfield.hh:260-265) and su2gpu_get_derivative downloads. */
int su2gpu_derivative(su2gpu_lattice *h, double fac, double *deriv);
int su2gpu_zero_derivative(su2gpu_lattice *h);
int su2gpu_derivative_device(su2gpu_lattice *h, double fac);
int su2gpu_get_derivative(su2gpu_lattice *h, double *deriv);
/* update_momenta(list{gauge}, deriv, h, dtau) (update_momenta.hh:10-33): P -= dtau * F */
int su2gpu_update_momenta(su2gpu_lattice *h, double dtau);
/* update_gauge(h, dtau) (update_gauge.hh:8-16): U <- exp(dtau P) U */
int su2gpu_update_gauge(su2gpu_lattice *h, double dtau);
/* integrator::integrate(list, h, params, restore) for the integrator id
 * (integrator.hh:18-241; set_integrator :244-312).  exponent is the rounding
 * exponent of lp_leapfrog / lp_omf4. */
int su2gpu_integrate(su2gpu_lattice *h, int integrator, size_t nsteps, double tau, int restore,
                     size_t exponent);

/* ---- one HMC trajectory ------------------------------------------------------
 * md_update(U, engine, params, {kinetic, gauge}, integ) (md_update.hh:18-81)
 * with engine = std::mt19937(seed) (hmc.hpp:94): momenta, H_old, integrate,
 * H_new, Metropolis (one uniform from the same engine iff deltaH > 0), optional
 * reversibility test, restore U_old on reject.  momenta_source picks the
 * reference's host stream or device Philox momenta. */
int su2gpu_md_update(su2gpu_lattice *h, uint64_t seed, int integrator, size_t nsteps, double tau,
                     size_t exponent, int revtest, int momenta_source, su2gpu_md_result *res);
/* Split form for callers that own their engine (the C++ adapter): the
 * caller's momenta (NULL: keep the device momenta) -> trajectory; U_old and
 * U_new both stay resident until su2gpu_md_finish(accept) keeps one. */
int su2gpu_md_trajectory(su2gpu_lattice *h, const double *momenta, int integrator, size_t nsteps,
                         double tau, size_t exponent, int revtest, su2gpu_md_result *res);
int su2gpu_md_finish(su2gpu_lattice *h, int accept);

/* ---- observables ---------------------------------------------------------------- */
/* energy_density(U, E, Q, cloverdef, ss) (flat-energy_density.hh:57-164) */
int su2gpu_energy_density(su2gpu_lattice *h, int clover, int ss, double *E, double *Q);

/* ---- Metropolis ------------------------------------------------------------------ */
/* sweep(U, engines, delta, N_hit, beta, xi, anisotropic) (flat-sweep.hh:64-164):
 * n_hit Metropolis hits on every link, R = random_element(delta)
 * (random_element.hh:9-34), dS = beta/Nc (ReTr(U K) - ReTr(U R K)), accept if
 * dS < 0 or uniform < exp(-dS), U <- restoreSU(U R); anisotropic when xi != 1.
 * The links are visited per direction mu and colour class (the site parity
 * where every extent in use is even, else one of three classes of
 * sum_{nu != mu} h(x_nu) mod 3, a proper colouring for any extents), and the
 * random numbers come from one Philox4x32-10 block per
 * (link, hit) with key = seed and counter (link, hit, sweep): a sweep depends on
 * (seed, sweep) only and is not the reference's per-thread mt19937 stream.
 * *rate = accepted hits / (n_hit * links), *rate_time the same for the
 * temporal links over n_hit * V: the reference's return values. */
int su2gpu_sweep(su2gpu_lattice *h, uint64_t seed, uint64_t sweep, double delta, size_t n_hit,
                 double *rate, double *rate_time);

/* ---- gradient flow (flat-gradient_flow.hh) ------------------------------------- */
/* runge_kutta(h, SW, eps) (flat-gradient_flow.hh:17-50) applied to the handle's links */
int su2gpu_runge_kutta(su2gpu_lattice *h, double eps);
/* gradient_flow(U, path, tmax, eps, xi, tstart, save) (flat-gradient_flow.hh:70-197)
 * on a copy of the links: writes up to max_rows rows of 9 doubles
 * {t, xi_R, P, P_ss, Ep, Ep_ss, Ec, Ec_ss, Q}; *nrows gets the count. */
int su2gpu_gradient_flow(su2gpu_lattice *h, double tmax, double eps, double tstart, double *rows,
                         size_t max_rows, size_t *nrows);

/* ---- gradient-flow adjoint (no reference counterpart: an addition over
 * flat-gradient_flow.hh:17-50) ---------------------------------------------------
 * su2gpu_flow_forward runs nsteps RK3 steps of size eps from the handle's links
 * (left untouched), keeping a checkpoint every ckpt_every steps, and returns the
 * observable at t = nsteps * eps:
 *   SU2GPU_OBS_CLOVER_E      t^2 E(t), E the clover energy density of energy_density
 *                            (flat-energy_density.hh:57-164);
 *   SU2GPU_OBS_WILSON_ACTION S(t) = get_S_G_hmc (flat-gaugemonomial.hh:46-54), whose
 *                            gradient at nsteps = 0 is the HMC force.
 * su2gpu_flow_backward returns dO/dU of the last forward w.r.t. the links it
 * started from, in the momentum layout and frame of gaugemonomial::derivative
 * (U -> exp(X) U, dO = sum grad . x), recomputing the flow states of each
 * checkpoint segment.  SU(2) only. */
enum { SU2GPU_OBS_CLOVER_E = 0, SU2GPU_OBS_WILSON_ACTION = 1 };
int su2gpu_flow_forward(su2gpu_lattice *h, double eps, size_t nsteps, size_t ckpt_every,
                        int observable, double *value);
int su2gpu_flow_backward(su2gpu_lattice *h, double *grad);

/* ---- batched ensemble (no reference counterpart: many independent copies of
 * the md_update of md_update.hh:18-81 advanced together) ---------------------------
 * A batch holds lattices of uneven sizes (U(1) / SU(2), d = 2, 3, 4, own beta,
 * integrator leapfrog or omf2, own step count and tau).  One trajectory of the
 * batch is one MD tick after the other for all lattices at once (per tick one
 * launch for the SU(2) lattices of each of d = 4 and d = 3 and one for the
 * rest; optionally replayed from a captured hipGraph), FP64 action / kinetic
 * sums per lattice, and a Metropolis
 * decision per lattice with the Philox uniform of (seed, traj).  Links start cold. */
typedef struct su2gpu_batch su2gpu_batch;
typedef struct {
  int group;                       /* SU2GPU_U1 or SU2GPU_SU2 */
  size_t Lx, Ly, Lz, Lt, ndims;    /* as su2gpu_create */
  double beta;
  int integrator;                  /* SU2GPU_LEAPFROG or SU2GPU_OMF2 */
  size_t nsteps;
  double tau;
  uint64_t seed;                   /* key of the lattice's device momenta / accept uniform */
} su2gpu_batch_member;
int su2gpu_batch_create(su2gpu_batch **out, const su2gpu_batch_member *members, size_t n,
                        int precision);
int su2gpu_batch_destroy(su2gpu_batch *b);
int su2gpu_batch_size(const su2gpu_batch *b, size_t *n);
/* links / momenta of member i in the packed reference AoS order */
int su2gpu_batch_set_links(su2gpu_batch *b, size_t i, const double *src);
int su2gpu_batch_get_links(su2gpu_batch *b, size_t i, double *dst);
int su2gpu_batch_set_momenta(su2gpu_batch *b, size_t i, const double *src);
/* one trajectory of every member; momenta_source SU2GPU_MOMENTA_DEVICE_PHILOX
 * draws them on the GPU from (member seed, traj), SU2GPU_MOMENTA_HOST_MT19937
 * uses the ones set with su2gpu_batch_set_momenta.  metropolis = 0 accepts
 * every trajectory.  results has one entry per member. */
int su2gpu_batch_trajectory(su2gpu_batch *b, int momenta_source, uint64_t traj, int metropolis,
                            su2gpu_md_result *results);
/* launch shape: threads per tile/workgroup (64, 128, 256) and grid (0 = one pass) */
int su2gpu_batch_set_launch(su2gpu_batch *b, int block, int grid);
int su2gpu_batch_get_launch(const su2gpu_batch *b, int *block, int *grid);
/* autotuner: times the candidate shapes on this batch, applies the fastest and
 * records it in a JSON cache (cache_path, or $SU2GPU_TUNE_CACHE; NULL and unset:
 * no cache); a cached entry is applied without timing */
int su2gpu_batch_autotune(su2gpu_batch *b, const char *cache_path, int *block, int *grid);
/* replay trajectories from a captured hipGraph (on) or launch them eagerly (off) */
int su2gpu_batch_use_graph(su2gpu_batch *b, int on);
/* sum over members of nd * V * nsteps: the link-updates of one batch trajectory */
int su2gpu_batch_link_updates(const su2gpu_batch *b, double *per_trajectory);
/* launch-gap and per-part efficiency (no counterpart in the reference, which
 * has no batch): one eager trajectory of device momenta (with Metropolis, as
 * su2gpu_batch_trajectory) timed with HIP events around every launch group.
 * out[0..7] ms: whole trajectory, momenta, plaquette sums, MD ticks of the
 * SU(2) d = 4 part, of the d = 3 part, of the generic rest, restore, reduction;
 * out[8..10] link updates of the three parts; out[11..13] their active tile
 * fraction (tile-ticks of running lattices over tiles x ticks; the tick
 * launches cover only the running lattices' tiles).  nout >= 14. */
int su2gpu_batch_profile(su2gpu_batch *b, uint64_t traj, double *out, size_t nout);

/* ---- configuration files (gaugeconfig.hh:225-252) ------------------------------ */
int su2gpu_save(su2gpu_lattice *h, const char *path);
int su2gpu_load(su2gpu_lattice *h, const char *path);

/* ---- stream / timing ------------------------------------------------------------ */
int su2gpu_synchronize(su2gpu_lattice *h);
/* Measurement helper (no reference counterpart): times `reps` launches of a
 * plain 16-byte-per-lane copy kernel of nbytes (read nbytes + write nbytes) and
 * returns the average milliseconds per copy -- the attainable HBM bandwidth and
 * the calibration of the hardware byte counters. */
int su2gpu_copy_probe(size_t nbytes, int reps, double *ms_per_copy);
/* HIP-event timing of every kernel launch on the handle's stream (off by
 * default).  su2gpu_profile_read(h, name, &ms, &count) returns the summed
 * duration and launch count of kernel class `name` ("md_fused", "md_p_only",
 * "update_gauge", "plaq", "sumsq", "reduce", "momenta", ...) since the last reset. */
int su2gpu_profile_enable(su2gpu_lattice *h, int on);
int su2gpu_profile_reset(su2gpu_lattice *h);
int su2gpu_profile_read(su2gpu_lattice *h, const char *name, double *ms, long *count);

#ifdef __cplusplus
}
#endif
#endif /* SU2_MI355X_H */
