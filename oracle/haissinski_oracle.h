/*
 * haissinski_oracle.h -- CPU restatement of HaissinskiJulia's longitudinal
 * tracking loop (MPI branch, strict IEEE binary64, no FMA contraction).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
 * load it.  Nothing in the product links or calls it.
 *
 * Every function cites the reference source it restates (paths relative to
 * the reference repository root).  Pins: see the header of
 * haissinski_oracle.c for which test pins which function.
 */
#ifndef HAISSINSKI_ORACLE_H
#define HAISSINSKI_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/constants.jl:5-7 */
#define OR_SPEED_LIGHT 299792458.0
#define OR_MASS_ELECTRON 0.51099895069e6

/* src/data_structures.jl:81-97 -- field order kept. */
typedef struct {
    double E0, mass, voltage;
    int64_t harmonic;
    double radius, pipe_radius, alpha_c, phi_s, freq_rf;
    int64_t n_turns;
    int32_t use_wakefield, update_eta, update_E0, SR_damping, use_excitation;
} or_params;

typedef struct {
    const double *normals;  /* injected QE normals, turn-major [n_turns][n]; NULL -> Philox stream */
    uint64_t seed;          /* Philox key when normals == NULL */
    int64_t turn_offset;    /* absolute turn index of the first turn (Philox counter) */
    int32_t nranks;         /* rank-equivalents (threads) for the MPI decomposition; <=0 -> 1 */
    int32_t pad_;
    /* Extensions (all off = the reference):                                    */
    const double *wake_table;      /* W by lag j*step behind the source, [nbins]; NULL -> W(c_j)
                                      of wakefield.jl:194-196 */
    double delta_sigma;            /* smoothing width; 0 -> 0.15 (wakefield.jl:204); <0 -> no kernel */
    double bin_half_width_sigmas;  /* 0 -> 7.5 (evolution.jl:270) */
    double charge;                 /* >0: current = charge [C] instead of evolution.jl:377-386, and
                                      the line density is counts/n_h without the delta kernel
                                      (delta_sigma > 0 together with charge is an error) */
    int32_t n_resonators;          /* >0: W by lag = sum of broadband resonators at s = j*step */
    int32_t serial;                /* 1: use_mpi = false (evolution.jl:245-253,437-447,461-467):
                                      two-pass compute_std spreads, no E0 update of an empty
                                      bunch; with the wake the reference raises (-5) */
    const double *resonators;      /* [n_resonators][3] = (R_s [Ohm], f_r [Hz], Q) */
    double *monitor;               /* [n_turns][4] or NULL: z centroid and sigma_z after the drift
                                      (wake on), E0 after the turn, binned count */
} or_options;

typedef struct {
    double sigma_E, sigma_z, E0;
    int64_t turns_done;
    int32_t drift_skipped;      /* apply_phase_advance! warnings (evolution.jl:151-154) */
} or_result;

/* utils.jl:119-122 */
int64_t or_next_power_of_two(int64_t n);
/* evolution.jl:270-274: range(-zw, zw, length=nb+1) and its centres, correctly rounded */
void or_bin_grid(double zw, int64_t nb, double *edges, double *centers, double *step_out);
/* utils.jl:569-606 with one rank (one-pass) */
double or_global_std(const double *x, int64_t n);
/* utils.jl:531-544 (two-pass serial helper) */
double or_serial_std(const double *x, int64_t n);
/* utils.jl:484-491 */
double or_calc_rf_factor(double freq_rf, double beta);
/* rf_kick.jl:84-87 */
void or_rf_kick(double voltage, double sin_phis, double rf, double phis,
                const double *z, double *dE, int64_t n);
/* quantum_excitation.jl:58-59,69-71 (normals given) */
void or_quantum_excitation(double E0, double radius, double sigma_E0,
                           const double *r, double *dE, int64_t n);
/* synchrotron_radiation.jl:44-45,54-56 */
void or_synchrotron_radiation(double E0, double radius, double *dE, int64_t n);
/* evolution.jl:331-344 */
void or_drift_update_eta(double *z, const double *dE, int64_t n, double mass, double gamma0,
                         double alpha_c, int64_t harmonic, double beta0, double E0,
                         double rf, double phis);
/* evolution.jl:143-167; returns 1 when the advance is skipped */
int or_apply_phase_advance(double *z, const double *dE, int64_t n, double eta0,
                           int64_t harmonic, double beta0, double E0, double rf, double phis);
/* utils.jl:351-401 */
void or_histogram(const double *x, int64_t n, double e_first, double e_last, double inv_step,
                  int64_t nb, int64_t *counts);
/* wakefield.jl:115-118 */
double or_wake_function(double z, double wake_factor, double wake_sqrt, double inv_ctau);
/* utils.jl:58-66 */
double or_delta(double x, double sigma);
/* extension: broadband resonator wake [V/C] at distance s >= 0 behind the source */
double or_resonator_wake(double s, double Rs, double fr, double Q);
/* FFTW forward DFT / normalised inverse (the transforms wakefield.jl:242-243,307 call) */
void or_fft(double *re, double *im, int64_t L, int inverse);
/* wakefield.jl:186-319 (rank 0 work of the MPI branch): counts -> potential at centres.
 * wake_tab[j] = W(c_j) (wakefield.jl:194-196), lam_tab[i] = delta(c_i, 0.15) (wakefield.jl:219-221). */
void or_wake_potential(const int64_t *counts, int64_t nb, const double *wake_tab,
                       const double *lam_tab, double current, double *P);
/* Lag-indexed (causal) wake: P_k = current * sum_{j>=0} W_j g_{k+j}, the charge at
 * bin k+j (ahead of bin k: larger z is the head, wakefield.jl:111-113) acting on
 * bin k through W at lag j*step.  Same FFT convolution on reversed g and P. */
void or_wake_potential_lag(const int64_t *counts, int64_t nb, const double *wake_tab,
                           const double *lam_tab, double current, double *P);
/* same linear convolution by the direct index sum (independent check) */
void or_wake_potential_direct(const int64_t *counts, int64_t nb, const double *wake_tab,
                              const double *lam_tab, double current, double *P);
/* evolution.jl:373-386 */
double or_wake_current(int64_t n_global, double E0, double radius, double sigma_z,
                       double eta0, double sigma_E0);
/* wakefield.jl:22-105 */
void or_interpolate(const double *z, int64_t n, const double *centers, int64_t nb,
                    const double *P, double *out);
/* Philox4x32-10 (Salmon et al., SC'11) */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* QE normal stream: particle i, turn t (Box-Muller over one Philox block per particle pair) */
void or_normals(uint64_t seed, int64_t turn, int64_t i0, int64_t n, double *out);
/* initial bunch from the Philox stream (independent Gaussians) */
void or_generate_gaussian(uint64_t seed, double mu_z, double sigma_z, double mu_E,
                          double sigma_E, int64_t n, double *z, double *dE);
/* evolution.jl:42-133: sample-covariance MvNormal bunch on the Philox streams */
void or_generate_particles(uint64_t seed, double mu_z, double mu_E, double sigma_z, double sigma_E,
                           int64_t n, double *z, double *dE, double *sigma_E_out,
                           double *sigma_z_out, double *L_out);
/* evolution.jl:206-471, MPI branch with nranks rank-equivalents (serial branch with
 * o->serial).  Returns 0, or -1 bad options, -2 wake with no particles, -3 nbins <= 0,
 * -4 pipe_radius <= 0, -5 wake with the serial branch (MethodError of evolution.jl:267). */
int or_evolve(double *z, double *dE, int64_t n, int64_t nbins, const or_params *p,
              const or_options *o, or_result *res, double *P_last);

#ifdef __cplusplus
}
#endif
#endif
