/*
 * haissinski_mi355x.h -- C ABI of the MI355X (gfx950) longitudinal tracker.
 *
 * Drop-in for the turn loop of HaissinskiJulia (src/evolution.jl:206-471)
 * and the stages it drives (src/rf_kick.jl, src/synchrotron_radiation.jl,
 * src/quantum_excitation.jl, src/wakefield.jl).  Plain C types only: host
 * pointers unless a name says "device"; every call returns an hm_status and
 * leaves a message for hm_last_error() on failure.
 *
 * Arithmetic is IEEE binary64 throughout (no FMA contraction in the
 * per-particle path).  One context owns one GPU and one HIP stream; calls on
 * a context are serialised on that stream and block until their results are
 * on the host.  Different contexts may be used from different host threads.
 */
#ifndef HAISSINSKI_MI355X_H
#define HAISSINSKI_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
typedef int32_t hm_status;
#define HM_OK 0
#define HM_ERR_ARG (-1)          /* bad argument (null pointer, negative size, ...) */
#define HM_ERR_EMPTY_WAKE (-2)   /* wake on with zero particles: log10(0) (evolution.jl:377-379) */
#define HM_ERR_NBINS (-3)        /* "Buffer nbins invalid (must be > 0)." (evolution.jl:264) */
#define HM_ERR_PIPE (-4)         /* "Wake parameter pipe_radius must be positive." (evolution.jl:278) */
#define HM_ERR_DEVICE (-10)      /* HIP runtime error */
#define HM_ERR_OOM (-11)         /* device allocation failed */
#define HM_ERR_LIMIT (-12)       /* size beyond what the kernels support */
#define HM_ERR_BUSY (-13)        /* server: job queue full */
#define HM_ERR_CANCELLED (-14)   /* server: job cancelled or failed by request */

/* ---- reference containers ---------------------------------------------- */

/* SimulationParameters (src/data_structures.jl:81-97), field order kept.
 * Flags are 0/1. */
typedef struct {
    double E0;            /* reference energy [eV] */
    double mass;          /* particle mass [eV] */
    double voltage;       /* RF voltage [V] */
    int64_t harmonic;     /* RF harmonic number */
    double radius;        /* bending radius [m] */
    double pipe_radius;   /* beam pipe radius [m] (reference wake) */
    double alpha_c;       /* momentum compaction */
    double phi_s;         /* synchronous phase [rad] */
    double freq_rf;       /* RF frequency [Hz] */
    int64_t n_turns;
    int32_t use_wakefield, update_eta, update_E0, SR_damping, use_excitation;
} hm_params;

/* Options outside SimulationParameters (all zero = the reference). */
typedef struct {
    const double *normals;   /* injected QE normals, host, turn-major [n_turns][n]; NULL = Philox */
    uint64_t seed;           /* Philox4x32-10 key */
    int64_t turn_offset;     /* absolute index of the first turn (Philox counter) */
    const double *wake_table;   /* W at lag j*step behind the source, host [nbins]; NULL =
                                   wakefield.jl:115-118 at the centres (the reference's index
                                   convolution).  A table by lag is applied causally: bin k
                                   feels the bins k+j ahead of it (larger z is the head,
                                   wakefield.jl:111-113), P_k = I sum_j W_j g_{k+j} */
    double delta_sigma;         /* line-density kernel width; 0 = 0.15 (wakefield.jl:204); <0 = none */
    double bin_half_width_sigmas;  /* 0 = 7.5 (evolution.jl:270) */
    double charge;              /* >0: wake current = charge [C] (replaces evolution.jl:377-386);
                                   the line density is then counts/n_h without the delta kernel
                                   (HM_ERR_ARG with delta_sigma > 0) */
    int32_t n_resonators;       /* >0: W by lag = sum of broadband resonators at s = j*step */
    int32_t timing;             /* 1: time the turn-pass kernel with HIP events */
    const double *resonators;   /* host [n_resonators][3] = (R_s [Ohm], f_r [Hz], Q) */
    int32_t fft_block;          /* 0 = automatic; else the convolution block size (power of 2) */
    int32_t block_size;         /* 0 = automatic; else threads per block of the turn pass (1024) */
    double *monitor;            /* host [n_turns][4] or NULL: per turn the z centroid and sigma_z
                                   after the drift (wake on; 0 otherwise), E0 after the turn and
                                   the binned count -- the moments of the turn-by-turn states of
                                   BeamTurn / assign_to_turn! (data_structures.jl:52-54,
                                   utils.jl:42-50) */
    int32_t serial;             /* 1: use_mpi = false (evolution.jl:245-253,437-447,461-467):
                                   two-pass compute_std (utils.jl:531-544) for the initial and
                                   final spreads, no E0 update of an empty bunch; with the wake
                                   HM_ERR_ARG, as the reference raises a MethodError there
                                   (evolution.jl:267 calls initialize_mpi_buffers! with nothing) */
    int32_t fom_stats;          /* 1: the final two-pass means and Statistics.cor(z, dE) in
                                   hm_result (mean_z, mean_dE, corr_z_dE: two more reads of the
                                   bunch); 0: corr_z_dE is NaN and the means are one-pass */
    int32_t tile_pairs;         /* particle pairs per thread in each tile of the turn pass (a
                                   tile is tile_pairs x block size pairs); 0 = the library's
                                   rule (about 1024 tiles per bunch, at least one pair per
                                   thread).  The per-tile sums follow the tiling, so runs of a
                                   bunch agree bit for bit under the same value */
    int32_t run_pairs;          /* the cost, in pairs, charged to a run of this bunch's tiles
                                   (staging and flushing its LDS) when the turn pass's
                                   workgroup ranges are balanced; 0 = 4 nbins.  Changes
                                   no result */
    int32_t bin_whole;          /* the bin phase of a small bunch (reduction, convolution, E0
                                   update) in one workgroup: 0 = wherever it fits (at most 256
                                   tiles, no interpolation sums, its tables in 144 KB of LDS),
                                   -1 = never (a launch per phase).  Changes no result */
} hm_options;

/* Return of longitudinal_evolve! (evolution.jl:470) plus diagnostics. */
typedef struct {
    double sigma_E;         /* final energy spread */
    double sigma_z;         /* final bunch length */
    double E0;              /* final reference energy */
    int64_t turns_done;
    int32_t drift_skipped;  /* turns where apply_phase_advance! skipped (evolution.jl:151-154) */
    int32_t pad_;
    double ms_total;        /* device time of the turn loop incl. set-up and epilogue [ms] */
    double ms_pass_avg;     /* average turn-pass kernel duration [ms] (timing = 1) */
    double ms_conv_avg;     /* average wake-convolution kernel duration [ms] (timing = 1) */
    int64_t pass_launches;  /* number of turn-pass launches timed */
    double sigma_E0, sigma_z0;  /* initial spreads (evolution.jl:245-253) */
    double bin_half_width;      /* z_width actually used [m] */
    double mean_z, mean_dE;     /* final centroid */
    double corr_z_dE;           /* Statistics.cor of the final z, dE: the emittance figure
                                   of merit (parameter_sensitivity/figures_of_merit.jl:36-47);
                                   NaN unless hm_options.fom_stats */
} hm_result;

typedef struct hm_ctx hm_ctx;       /* one GPU, one stream, cached work buffers */
typedef struct hm_bunch hm_bunch;   /* device-resident particles of one bunch */

/* ---- context ----------------------------------------------------------- */
hm_status hm_create(int32_t device, hm_ctx **out);
hm_status hm_destroy(hm_ctx *ctx);
const char *hm_last_error(void);
/* name (e.g. "gfx950"), CU count, max dynamic LDS per block */
hm_status hm_device_info(hm_ctx *ctx, char *arch, int32_t arch_len, int32_t *cus, int32_t *lds_bytes);

/* ---- whole loop, host arrays ------------------------------------------- */
/* longitudinal_evolve!(particles, params, buffers, comm, use_mpi)
 * (src/evolution.jl:206-212): z, dE are particles.coordinates.z / .ΔE (mutated
 * in place), nbins = length(buffers.λ).  One GPU holds the whole bunch, so
 * the MPI collectives of the reference are on-device reductions. */
hm_status hm_evolve(hm_ctx *ctx, double *z, double *dE, int64_t n, int64_t nbins,
                    const hm_params *params, const hm_options *opts, hm_result *result);

/* ---- device-resident bunches ------------------------------------------- */
hm_status hm_bunch_create(hm_ctx *ctx, int64_t n, hm_bunch **out);
hm_status hm_bunch_destroy(hm_bunch *b);
hm_status hm_bunch_upload(hm_bunch *b, const double *z, const double *dE);
hm_status hm_bunch_download(hm_bunch *b, double *z, double *dE);
/* generate_particles (src/evolution.jl:42-133) for independent Gaussians from
 * the Philox stream: z ~ N(mu_z, sigma_z), dE ~ N(mu_E, sigma_E). */
hm_status hm_bunch_generate(hm_bunch *b, uint64_t seed, double mu_z, double sigma_z,
                            double mu_E, double sigma_E);
/* generate_particles (src/evolution.jl:42-133) into the bunch: for n >= 2 the
 * sample covariance of min(n, 10000) independent draws (variances floored at
 * 1e-18, diagonal lifted to a smallest eigenvalue of 1e-12) and a MvNormal
 * bunch mu + L xi from it; n == 1 draws independent normals; n == 0 draws
 * nothing.  Returns the bunch's compute_std spreads (src/utils.jl:531-544), the
 * input spreads for n == 1 and 0 for n == 0, as the reference does.  Draws
 * come from the Philox stream of `seed`, not Julia's default RNG. */
hm_status hm_bunch_generate_particles(hm_bunch *b, uint64_t seed, double mu_z, double mu_E,
                                      double sigma_z, double sigma_E, double *sigma_E_out,
                                      double *sigma_z_out);
/* longitudinal_evolve! on a resident bunch (no host transfer of particles). */
hm_status hm_bunch_evolve(hm_bunch *b, int64_t nbins, const hm_params *params,
                          const hm_options *opts, hm_result *result);
/* device addresses of the bunch's z and dE arrays (for callers that own kernels) */
hm_status hm_bunch_device_ptrs(hm_bunch *b, double **z_device, double **dE_device);
/* No counterpart in the reference: hm_bunch_evolve on the caller's device
 * arrays z_device, dE_device (mutated in place) with no device allocation, for
 * callers that cannot allocate (code under stream capture, serving loops,
 * harnesses that forbid hipMalloc).  Every work buffer comes from an 8 MiB
 * per-GPU scratch array held until the process ends: at most 32768 bins
 * (HM_ERR_LIMIT beyond); no injected normals, turn monitor or timing
 * (HM_ERR_ARG).  The same kernels as hm_bunch_evolve, so the same results.
 * Calls that overlap on one GPU race on the scratch array. */
hm_status hm_evolve_device_noalloc(hm_ctx *ctx, double *z_device, double *dE_device, int64_t n,
                                   int64_t nbins, const hm_params *params,
                                   const hm_options *opts, hm_result *result);

/* ---- batched bunches: one segmented launch per turn for all of them ----- */
/* Each bunch keeps its own parameters, options, bin count and RNG stream;
 * results are identical to evolving each bunch alone (same per-bunch tiling
 * and reduction order).  Bunches with fewer turns retire early. */
hm_status hm_batch_evolve(hm_ctx *ctx, int32_t nbunches, hm_bunch *const *bunches,
                          const int64_t *nbins, const hm_params *params,
                          const hm_options *opts, hm_result *results);

/* ---- stage entry points (host arrays; unit parity with the reference) --- */
/* rf_kick! (src/rf_kick.jl:18-25,84-87) */
hm_status hm_rf_kick(hm_ctx *ctx, double voltage, double sin_phis, double rf_factor,
                     double phi_s, const double *z, double *dE, int64_t n);
/* quantum_excitation! (src/quantum_excitation.jl:24-72) with given normals r */
hm_status hm_quantum_excitation(hm_ctx *ctx, double E0, double radius, double sigma_E0,
                                const double *r, double *dE, int64_t n);
/* synchrotron_radiation! (src/synchrotron_radiation.jl:12-58) */
hm_status hm_synchrotron_radiation(hm_ctx *ctx, double E0, double radius, double *dE, int64_t n);
/* phase advance: update_eta = 1 -> evolution.jl:331-344, else apply_phase_advance!
 * (evolution.jl:143-167).  Returns 1 in *skipped when the advance is skipped. */
hm_status hm_drift(hm_ctx *ctx, int32_t update_eta, double E0, double mass, double alpha_c,
                   int64_t harmonic, double phi_s, double freq_rf, double *z, const double *dE,
                   int64_t n, int32_t *skipped);
/* calculate_local_histogram! (src/utils.jl:351-401); counts are int64 [nbins] */
hm_status hm_histogram(hm_ctx *ctx, const double *z, int64_t n, double e_first, double e_last,
                       double inv_step, int64_t nbins, int64_t *counts);
/* wakefield.jl:186-319: counts -> potential at the bin centres, by the
 * LDS FFT convolution; wake_tab[j] = W(c_j) (or W by lag), lam_tab[i] = delta(c_i) */
hm_status hm_wake_potential(hm_ctx *ctx, const int64_t *counts, int64_t nbins,
                            const double *wake_tab, const double *lam_tab, double current,
                            int32_t fft_block, double *P);
/* zero_alloc_interpolation! + kick (src/wakefield.jl:22-105,342-344): dE -= interp(P, z) */
hm_status hm_interpolate_kick(hm_ctx *ctx, const double *z, double *dE, int64_t n,
                              const double *centers, int64_t nbins, const double *P);
/* compute_global_std (src/utils.jl:569-606) */
hm_status hm_global_std(hm_ctx *ctx, const double *x, int64_t n, double *out);
/* two-pass statistics of a bunch: out[0..4] = mean z, mean dE, compute_std of
 * z and dE (src/utils.jl:531-544), Statistics.cor(z, dE) as used by the
 * emittance figure of merit (src/parameter_sensitivity/figures_of_merit.jl:36-47) */
hm_status hm_pair_stats(hm_ctx *ctx, const double *z, const double *dE, int64_t n, double *out);
/* the device QE normal stream: normal of particles [i0, i0+n) at a turn */
hm_status hm_normals(hm_ctx *ctx, uint64_t seed, int64_t turn, int64_t i0, int64_t n, double *out);
/* bin edges/centres of range(-zw, zw, length=nbins+1) (evolution.jl:270-274),
 * correctly rounded; host-only helper, needs no GPU */
hm_status hm_bin_grid(double zw, int64_t nbins, double *edges, double *centers, double *step);

/* ---- measurement -------------------------------------------------------- */
/* No counterpart in the reference: the HBM ceiling of this GPU.  Copies
 * `bytes` (rounded down to 16) device-to-device with a plain kernel `iters`
 * times and reports the average kernel time and read+write bandwidth. */
hm_status hm_stream_copy(hm_ctx *ctx, int64_t bytes, int32_t iters, double *ms_avg, double *gb_per_s);

/* ---- serving: continuous batching of tracking jobs ---------------------- */
typedef struct hm_server hm_server;
typedef struct {
    hm_params params;          /* n_turns = turns of this job */
    hm_options opts;           /* seed, extensions; normals must be NULL */
    int64_t n;                 /* particles */
    int64_t nbins;
    uint64_t particle_seed;    /* initial Gaussian bunch from hm_bunch_generate */
    double sigma_z, sigma_E;   /* initial bunch shape */
    int32_t fail_after_turns;  /* >0: the job fails by request after this many turns (test hook) */
    int32_t fail_setup;        /* 1: the job's set-up inside the batch fails by request, after
                                  its slot is taken (test hook of the admission's error path) */
} hm_job;

typedef struct {
    int64_t job_id;
    hm_status status;          /* HM_OK, HM_ERR_CANCELLED, ... */
    int32_t pad_;
    hm_result result;
    double t_submit, t_admit, t_done;   /* host clock [s] */
    double centroid_z, centroid_E;      /* final first moments */
} hm_job_result;

/* budget_bytes: HBM admission cap; max_slots: concurrent resident jobs */
hm_status hm_server_create(hm_ctx *ctx, int64_t budget_bytes, int32_t max_slots, hm_server **out);
hm_status hm_server_destroy(hm_server *s);
hm_status hm_server_submit(hm_server *s, const hm_job *job, int64_t *job_id);
hm_status hm_server_cancel(hm_server *s, int64_t job_id);
/* advance every resident job by up to `turns` turns in segmented launches,
 * admitting queued jobs as memory and slots free up; returns the number of
 * jobs still queued or running */
hm_status hm_server_tick(hm_server *s, int32_t turns, int32_t *active);
/* fetch up to cap finished results (in completion order) */
hm_status hm_server_poll(hm_server *s, hm_job_result *out, int32_t cap, int32_t *count);

#ifdef __cplusplus
}
#endif
#endif
