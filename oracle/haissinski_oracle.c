/*
 * haissinski_oracle.c -- CPU restatement of HaissinskiJulia's multi-turn
 * longitudinal tracking loop, MPI branch (the only branch that runs with the
 * wakefield on: src/evolution.jl:261-267 needs an MPI communicator).
 *
 * TEST INFRASTRUCTURE ONLY -- the checker, never the thing measured or shipped.
 *
 * Semantics: strict IEEE binary64, left-to-right evaluation exactly as the
 * Julia source is written, no FMA contraction (build with -ffp-contract=off).
 * Transcendentals come from the C library.
 *
 * Which test pins which function (tests/test_oracle_*.py):
 *   or_philox4x32_10      Random123 known-answer vectors
 *   or_normals            Box-Muller formula recomputed from raw Philox words, KS test
 *   or_bin_grid           exact rational rounding (Python fractions)
 *   or_global_std         closed form on exact data, numpy two-pass agreement
 *   or_histogram          brute force, edge cases (x = last edge, NaN, +-Inf)
 *   or_fft                numpy.fft, Parseval
 *   or_wake_potential     direct index sum, numpy.convolve
 *   or_interpolate        brute force incl. the z >= c_nb quirk
 *   stage kernels         independent numpy restatement (oracle/numpy_oracle.py)
 *   or_evolve             synchrotron tune, SR damping rate, QE/SR equilibrium,
 *                         rank-count independence, numpy restatement on tiny runs
 */
#include "haissinski_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_TWO_PI 6.283185307179586 /* Julia's 2π (and 2*π) as Float64 */
#define OR_PI 3.141592653589793

/* Julia's max for floats: NaN-propagating. */
static double jl_max(double a, double b) {
    if (a != a) return a;
    if (b != b) return b;
    return (b > a) ? b : a;
}

/* utils.jl:119-122: Int64(2^ceil(Int, log2(n))) */
int64_t or_next_power_of_two(int64_t n) {
    if (n <= 0) return 1;
    int64_t p = 1;
    while (p < n) p <<= 1;
    return p;
}

/* RN(zw * num / den) for integers num, den > 0, computed in binary128 (the
 * product is exact; the quotient is rounded twice, 113 then 53 bits). */
static double rn_scaled_ratio(double zw, int64_t num, int64_t den) {
    __float128 q = ((__float128)zw * (__float128)num) / (__float128)den;
    return (double)q;
}

/* evolution.jl:270-274.  edges[k] = -zw + k*(2zw/nb), k = 0..nb; centres are
 * the midpoints (edges[1:end-1] + edges[2:end]) ./ 2 of the Julia ranges. */
void or_bin_grid(double zw, int64_t nb, double *edges, double *centers, double *step_out) {
    if (edges) {
        for (int64_t k = 0; k <= nb; ++k) edges[k] = rn_scaled_ratio(zw, 2 * k - nb, nb);
        edges[0] = -zw;
        edges[nb] = zw;
    }
    if (centers)
        for (int64_t i = 0; i < nb; ++i) centers[i] = rn_scaled_ratio(zw, 2 * i + 1 - nb, nb);
    if (step_out) *step_out = (2.0 * zw) / (double)nb; /* step(bin_edges), correctly rounded */
}

/* utils.jl:587-605 given the reduced [sum, sum_sq, n]. */
static double gstd_from_sums(double sum, double sumsq, double nd) {
    int64_t n = (int64_t)llround(nd);
    if (n <= 1) return 0.0;
    double mean = sum / (double)n;
    double var = (sumsq / (double)n) - (mean * mean);
    double sample = jl_max(0.0, var * (double)n / (double)(n - 1));
    return sqrt(sample);
}

double or_global_std(const double *x, int64_t n) {
    double s = 0.0, s2 = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double v = x[i];
        s += v;
        s2 += v * v;
    }
    return gstd_from_sums(s, s2, (double)n);
}

/* utils.jl:519-544 */
double or_serial_std(const double *x, int64_t n) {
    if (n <= 1) return 0.0;
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += x[i];
    double mu = s / (double)n;
    double acc = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double d = x[i] - mu;
        acc += d * d;
    }
    double var = acc / (double)(n - 1);
    return sqrt(jl_max(0.0, var));
}

/* utils.jl:484-491 */
double or_calc_rf_factor(double freq_rf, double beta) {
    if (beta <= 0.0) return INFINITY;
    return (freq_rf * OR_TWO_PI) / (beta * OR_SPEED_LIGHT);
}

/* rf_kick.jl:84-87 */
void or_rf_kick(double voltage, double sin_phis, double rf, double phis, const double *z,
                double *dE, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        double phi = ((-z[i]) * rf) + phis;
        dE[i] = dE[i] + voltage * (sin(phi) - sin_phis);
    }
}

/* quantum_excitation.jl:58-59 / synchrotron_radiation.jl:44 (same expression) */
static double sr_dUdE(double E0, double radius) {
    double x = E0 / 1e9;
    return ((4.0 * 8.85e-5) * ((x * x) * x)) / radius;
}

static double qe_excitation(double E0, double radius, double sigma_E0) {
    double d = sr_dUdE(E0, radius);
    double om = 1.0 - d;
    return sqrt(1.0 - (om * om)) * sigma_E0;
}

void or_quantum_excitation(double E0, double radius, double sigma_E0, const double *r,
                           double *dE, int64_t n) {
    double exc = qe_excitation(E0, radius, sigma_E0);
    for (int64_t i = 0; i < n; ++i) dE[i] = dE[i] + exc * r[i];
}

void or_synchrotron_radiation(double E0, double radius, double *dE, int64_t n) {
    double damp = 1.0 - sr_dUdE(E0, radius);
    for (int64_t i = 0; i < n; ++i) dE[i] = dE[i] * damp;
}

/* evolution.jl:331-344 */
void or_drift_update_eta(double *z, const double *dE, int64_t n, double mass, double gamma0,
                         double alpha_c, int64_t harmonic, double beta0, double E0, double rf,
                         double phis) {
    double two_pi_h = OR_TWO_PI * (double)harmonic;
    double denom = (beta0 * beta0) * E0;
    for (int64_t i = 0; i < n; ++i) {
        double dg = dE[i] / mass;
        double gp = gamma0 + dg;
        double eta = alpha_c - 1.0 / (gp * gp);
        double coeff = (two_pi_h * eta) / denom;
        double phi = -((z[i] * rf) - phis);
        phi = phi + coeff * dE[i];
        z[i] = ((-phi) + phis) / rf;
    }
}

/* evolution.jl:143-167 */
int or_apply_phase_advance(double *z, const double *dE, int64_t n, double eta0,
                           int64_t harmonic, double beta0, double E0, double rf, double phis) {
    if (n == 0) return 0;
    if (beta0 <= 0.0 || E0 == 0.0 || rf == 0.0) return 1;
    double coeff = ((OR_TWO_PI * (double)harmonic) * eta0) / ((beta0 * beta0) * E0);
    for (int64_t i = 0; i < n; ++i) {
        double phi = -((z[i] * rf) - phis);
        phi = phi + coeff * dE[i];
        z[i] = ((-phi) + phis) / rf;
    }
    return 0;
}

/* utils.jl:351-401 (bins are 0-based here: b = floor(u) clamped to [0, nb-1]) */
void or_histogram(const double *x, int64_t n, double e_first, double e_last, double inv_step,
                  int64_t nb, int64_t *counts) {
    memset(counts, 0, sizeof(int64_t) * (size_t)nb);
    for (int64_t i = 0; i < n; ++i) {
        double v = x[i];
        int in_range = (v >= e_first) & (v <= e_last);
        if (!in_range) continue;
        double u = (v - e_first) * inv_step;
        int64_t b = (int64_t)floor(u);
        if (b < 0) b = 0;
        if (b > nb - 1) b = nb - 1;
        counts[b] += 1;
    }
}

/* wakefield.jl:115-118 */
double or_wake_function(double z, double wake_factor, double wake_sqrt, double inv_ctau) {
    if (z > 0.0) return 0.0;
    return (wake_factor * exp(z * inv_ctau)) * cos(wake_sqrt * z);
}

/* utils.jl:58-66 */
double or_delta(double x, double sigma) {
    double s_inv = 1.0 / (sqrt(OR_TWO_PI) * sigma);
    double ef = -0.5 / (sigma * sigma);
    return s_inv * exp((x * x) * ef);
}

/* Broadband resonator (Chao, Physics of Collective Beam Instabilities,
 * eq. 2.84): W(s) = 2 a Rs e^{-a s/c} [cos(wb s/c) - (a/wb) sin(wb s/c)],
 * a = w_r/(2Q), wb = sqrt(w_r^2 - a^2); W(0) = a Rs (half the jump).  The
 * overdamped (Q < 1/2) and critical (Q = 1/2) forms use cosh/sinh and the
 * limit sin(x)/wb -> s/c. */
double or_resonator_wake(double s, double Rs, double fr, double Q) {
    double wr = OR_TWO_PI * fr;
    double a = wr / (2.0 * Q);
    if (s < 0.0) return 0.0;
    if (s == 0.0) return a * Rs;
    double tau = s / OR_SPEED_LIGHT;
    double env = 2.0 * a * Rs * exp(-a * tau);
    double d = wr * wr - a * a;
    if (d > 0.0) {
        double wb = sqrt(d);
        return env * (cos(wb * tau) - (a / wb) * sin(wb * tau));
    } else if (d < 0.0) {
        double wb = sqrt(-d);
        return env * (cosh(wb * tau) - (a / wb) * sinh(wb * tau));
    }
    return env * (1.0 - a * tau);
}

/* In-place radix-2 DIT FFT; forward uses exp(-2 pi i jk/L) (FFTW's sign),
 * inverse uses exp(+...) and scales by 1/L (Julia's ifft).  Twiddles are
 * computed in long double. */
void or_fft(double *re, double *im, int64_t L, int inverse) {
    if (L <= 1) return;
    for (int64_t i = 1, j = 0; i < L; ++i) {
        int64_t bit = L >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    const long double two_pi_l = 6.283185307179586476925286766559L;
    double *wr = (double *)malloc(sizeof(double) * (size_t)(L / 2));
    double *wi = (double *)malloc(sizeof(double) * (size_t)(L / 2));
    for (int64_t k = 0; k < L / 2; ++k) {
        long double a = two_pi_l * (long double)k / (long double)L;
        wr[k] = (double)cosl(a);
        wi[k] = inverse ? (double)sinl(a) : -(double)sinl(a);
    }
    for (int64_t len = 2; len <= L; len <<= 1) {
        int64_t half = len >> 1, stride = L / len;
        for (int64_t s = 0; s < L; s += len) {
            for (int64_t k = 0; k < half; ++k) {
                double cr = wr[k * stride], ci = wi[k * stride];
                int64_t a = s + k, b = a + half;
                double tr = re[b] * cr - im[b] * ci;
                double ti = re[b] * ci + im[b] * cr;
                re[b] = re[a] - tr; im[b] = im[a] - ti;
                re[a] = re[a] + tr; im[a] = im[a] + ti;
            }
        }
    }
    free(wr);
    free(wi);
    if (inverse) {
        double sc = 1.0 / (double)L;
        for (int64_t i = 0; i < L; ++i) { re[i] *= sc; im[i] *= sc; }
    }
}

/* wakefield.jl:186-319, the rank-0 work of the MPI branch. */
void or_wake_potential(const int64_t *counts, int64_t nb, const double *wake_tab,
                       const double *lam_tab, double current, double *P) {
    int64_t L = or_next_power_of_two(nb * 2);
    int64_t nh = 0;
    for (int64_t i = 0; i < nb; ++i) nh += counts[i];
    if (nh == 0) { /* wakefield.jl:305-311: the IFFT is skipped and convol zeroed */
        for (int64_t i = 0; i < nb; ++i) P[i] = 0.0;
        return;
    }
    double inv_n = 1.0 / (double)nh;
    double *wr = (double *)calloc((size_t)L, sizeof(double));
    double *wi = (double *)calloc((size_t)L, sizeof(double));
    double *lr = (double *)calloc((size_t)L, sizeof(double));
    double *li = (double *)calloc((size_t)L, sizeof(double));
    for (int64_t i = 0; i < nb; ++i) {
        wr[i] = wake_tab[i];
        double normalized = (double)counts[i] * inv_n;
        lr[i] = lam_tab[i] * normalized;
    }
    or_fft(wr, wi, L, 0);
    or_fft(lr, li, L, 0);
    for (int64_t k = 0; k < L; ++k) { /* (W * L) * current, wakefield.jl:287 */
        double pr = wr[k] * lr[k] - wi[k] * li[k];
        double pi = wr[k] * li[k] + wi[k] * lr[k];
        wr[k] = pr * current;
        wi[k] = pi * current;
    }
    or_fft(wr, wi, L, 1);
    for (int64_t i = 0; i < nb; ++i) P[i] = wr[i];
    free(wr); free(wi); free(lr); free(li);
}

/* Lag-indexed wake tables (the resonator and wake-table extensions) give W at a
 * distance j*step BEHIND the source.  Larger z is the head (wakefield.jl:111-113:
 * "positive z is upstream (arriving earlier)"), so bin k feels the charge of the
 * bins k+j >= k: P_k = I sum_j W_j g_{k+j}.  Reversing g and P turns this into the
 * index convolution of or_wake_potential. */
void or_wake_potential_lag(const int64_t *counts, int64_t nb, const double *wake_tab,
                           const double *lam_tab, double current, double *P) {
    if (nb <= 0) return;
    int64_t *rc = (int64_t *)malloc(sizeof(int64_t) * (size_t)nb);
    double *rl = (double *)malloc(sizeof(double) * (size_t)nb);
    double *rp = (double *)malloc(sizeof(double) * (size_t)nb);
    for (int64_t i = 0; i < nb; ++i) {
        rc[i] = counts[nb - 1 - i];
        rl[i] = lam_tab[nb - 1 - i];
    }
    or_wake_potential(rc, nb, wake_tab, rl, current, rp);
    for (int64_t i = 0; i < nb; ++i) P[i] = rp[nb - 1 - i];
    free(rc); free(rl); free(rp);
}

void or_wake_potential_direct(const int64_t *counts, int64_t nb, const double *wake_tab,
                              const double *lam_tab, double current, double *P) {
    int64_t nh = 0;
    for (int64_t i = 0; i < nb; ++i) nh += counts[i];
    if (nh == 0) {
        for (int64_t i = 0; i < nb; ++i) P[i] = 0.0;
        return;
    }
    double inv_n = 1.0 / (double)nh;
    double *g = (double *)malloc(sizeof(double) * (size_t)nb);
    for (int64_t i = 0; i < nb; ++i) g[i] = lam_tab[i] * ((double)counts[i] * inv_n);
    for (int64_t k = 0; k < nb; ++k) {
        long double acc = 0.0L;
        for (int64_t j = 0; j <= k; ++j) acc += (long double)wake_tab[j] * (long double)g[k - j];
        P[k] = (double)(acc * (long double)current);
    }
    free(g);
}

/* evolution.jl:373-386 */
double or_wake_current(int64_t n_global, double E0, double radius, double sigma_z, double eta0,
                       double sigma_E0) {
    static const double pow10[11] = {1e0, 1e1, 1e2, 1e3, 1e4, 1e5, 1e6, 1e7, 1e8, 1e9, 1e10};
    double log10_N = log10((double)n_global);
    int64_t fl = (log10_N >= 0.0) ? (int64_t)floor(log10_N) : (int64_t)ceil(log10_N - 1.0);
    int64_t pw = fl < -10 ? -10 : (fl > 10 ? 10 : fl);
    double denom = pw >= 0 ? pow10[pw] : 1.0 / pow10[-pw];
    double pf = (1e11 / denom) * (double)n_global;
    return (((pf / E0) / (OR_TWO_PI * radius)) * sigma_z) / (eta0 * (sigma_E0 * sigma_E0));
}

/* wakefield.jl:22-105 (0-based: interval k in [0, nb-2], weight w) */
void or_interpolate(const double *z, int64_t n, const double *centers, int64_t nb,
                    const double *P, double *out) {
    if (nb <= 1) {
        for (int64_t i = 0; i < n; ++i) out[i] = (nb == 1) ? P[0] : 0.0;
        return;
    }
    double c0 = centers[0], c1 = centers[nb - 1];
    double step = (c1 - c0) / (double)(nb - 1);
    double inv = 1.0 / step;
    for (int64_t i = 0; i < n; ++i) {
        double pos = z[i];
        double cl = (pos > c1) ? c1 : ((pos < c0) ? c0 : pos);
        double u = (cl - c0) * inv;
        double fu = floor(u);
        int64_t k = (fu == fu) ? (int64_t)fu : 0; /* NaN: any index, the weight is NaN */
        if (k < 0) k = 0;
        if (k > nb - 2) k = nb - 2;
        double w = u - fu;
        out[i] = P[k] * (1.0 - w) + P[k + 1] * w;
    }
}

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Box-Muller (1958) on one Philox block: u1 in (0,1] from words 0,1 and
 * u2 in [0,1) from words 2,3, 53 bits each. */
static void box_muller(const uint32_t x[4], double *n0, double *n1) {
    double u1 = ((double)(x[0] >> 5) * 67108864.0 + (double)(x[1] >> 6) + 1.0) * 0x1p-53;
    double u2 = ((double)(x[2] >> 5) * 67108864.0 + (double)(x[3] >> 6)) * 0x1p-53;
    double r = sqrt(-2.0 * log(u1));
    double th = OR_TWO_PI * u2;
    *n0 = r * cos(th);
    *n1 = r * sin(th);
}

/* QE stream: the normal of particle i at turn t is output (i & 1) of the
 * block with counter {i>>1 (lo), i>>1 (hi), t, 0} and key {seed lo, seed hi}. */
void or_normals(uint64_t seed, int64_t turn, int64_t i0, int64_t n, double *out) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int64_t i = i0; i < i0 + n; ++i) {
        uint64_t j = (uint64_t)i >> 1;
        uint32_t ctr[4] = {(uint32_t)j, (uint32_t)(j >> 32), (uint32_t)turn, 0u};
        uint32_t x[4];
        or_philox4x32_10(ctr, key, x);
        double a, b;
        box_muller(x, &a, &b);
        out[i - i0] = (i & 1) ? b : a;
    }
}

/* Initial bunch: particle i uses counter {i lo, i hi, 0, 1}: z from the
 * first Box-Muller output, dE from the second. */
void or_generate_gaussian(uint64_t seed, double mu_z, double sigma_z, double mu_E, double sigma_E,
                          int64_t n, double *z, double *dE) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int64_t i = 0; i < n; ++i) {
        uint32_t ctr[4] = {(uint32_t)i, (uint32_t)((uint64_t)i >> 32), 0u, 1u};
        uint32_t x[4];
        or_philox4x32_10(ctr, key, x);
        double a, b;
        box_muller(x, &a, &b);
        z[i] = mu_z + sigma_z * a;
        dE[i] = mu_E + sigma_E * b;
    }
}

/* evolution.jl:42-133 (generate_particles) on the Philox streams.
 *  n <= 0: nothing drawn; spreads 0 (:47-53).
 *  n == 1: independent normals, the particle-1 block of stream 1; the
 *          returned spreads are the inputs (:61-66).
 *  n >= 2: m = min(n, 10000) independent samples (:68-72), sample j from the
 *          block {j lo, j hi, 0, 2}; their sample covariance (Statistics.cov:
 *          mean-centred, m-1 normalisation) with variances floored at 1e-18
 *          (:75-82); the diagonal lifted by 1e-12 - lambda_min when the
 *          smallest eigenvalue is below 1e-12 (:89-94); particle i is
 *          mu + L xi with L the lower Cholesky factor of that matrix and xi
 *          the stream-1 block of particle i (MvNormal, :97-113); the returned
 *          spreads are compute_std of the bunch (utils.jl:531-544).
 * L_out (3 doubles: l11, l21, l22) may be NULL. */
static double seq_mean(const double *x, int64_t n) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += x[i];
    return s / (double)n;
}

static double compute_std(const double *x, int64_t n) {   /* utils.jl:531-544 */
    if (n <= 1) return 0.0;
    const double mu = seq_mean(x, n);
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double d = x[i] - mu;
        s += d * d;
    }
    const double var = s / (double)(n - 1);
    return sqrt(var > 0.0 ? var : 0.0);
}

void or_generate_particles(uint64_t seed, double mu_z, double mu_E, double sigma_z, double sigma_E,
                           int64_t n, double *z, double *dE, double *sigma_E_out,
                           double *sigma_z_out, double *L_out) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    if (n <= 0) {
        *sigma_E_out = 0.0;
        *sigma_z_out = 0.0;
        return;
    }
    if (n == 1) {
        or_generate_gaussian(seed, mu_z, sigma_z, mu_E, sigma_E, 1, z, dE);
        *sigma_E_out = sigma_E;
        *sigma_z_out = sigma_z;
        return;
    }
    const int64_t m = n < 10000 ? n : 10000;
    double *zs = (double *)malloc(sizeof(double) * (size_t)m);
    double *es = (double *)malloc(sizeof(double) * (size_t)m);
    for (int64_t j = 0; j < m; ++j) {
        uint32_t ctr[4] = {(uint32_t)j, (uint32_t)((uint64_t)j >> 32), 0u, 2u};
        uint32_t x[4];
        or_philox4x32_10(ctr, key, x);
        double a, b;
        box_muller(x, &a, &b);
        zs[j] = mu_z + sigma_z * a;
        es[j] = mu_E + sigma_E * b;
    }
    const double mz = seq_mean(zs, m), me = seq_mean(es, m);
    double szz = 0.0, see = 0.0, sze = 0.0;
    for (int64_t j = 0; j < m; ++j) {
        const double dz = zs[j] - mz, de = es[j] - me;
        szz += dz * dz;
        see += de * de;
        sze += dz * de;
    }
    free(zs);
    free(es);
    const double md = (double)(m - 1);
    double a = szz / md, c = see / md;
    const double b = sze / md;
    if (a < 1e-18) a = 1e-18;
    if (c < 1e-18) c = 1e-18;
    /* smallest eigenvalue as det / largest: no cancellation when a << c */
    const double h = (a - c) * 0.5;
    const double lmax = (a + c) * 0.5 + sqrt(h * h + b * b);
    const double lmin = (a * c - b * b) / lmax;
    if (lmin < 1e-12) {
        const double lift = 1e-12 - lmin;
        a += lift;
        c += lift;
    }
    const double l11 = sqrt(a);
    const double l21 = b / l11;
    const double l22 = sqrt(c - l21 * l21);
    if (L_out) {
        L_out[0] = l11;
        L_out[1] = l21;
        L_out[2] = l22;
    }
    for (int64_t i = 0; i < n; ++i) {
        uint32_t ctr[4] = {(uint32_t)i, (uint32_t)((uint64_t)i >> 32), 0u, 1u};
        uint32_t x[4];
        or_philox4x32_10(ctr, key, x);
        double x1, x2;
        box_muller(x, &x1, &x2);
        z[i] = mu_z + l11 * x1;
        dE[i] = mu_E + (l21 * x1 + l22 * x2);
    }
    *sigma_z_out = compute_std(z, n);
    *sigma_E_out = compute_std(dE, n);
}

/* ---------------------------------------------------------------------------
 * evolution.jl:206-471, MPI branch with R rank-equivalents.  Rank r owns the
 * contiguous chunk given by benchmark_total.jl:178-189 (base + remainder to
 * the low ranks).  Allreduce sums are taken in rank order.
 * ------------------------------------------------------------------------- */
typedef struct { int64_t off, cnt; } chunk_t;

static void rank_chunks(int64_t n, int R, chunk_t *ch) {
    int64_t base = n / R, rem = n % R, off = 0;
    for (int r = 0; r < R; ++r) {
        int64_t c = (r < rem) ? base + 1 : base;
        ch[r].off = off;
        ch[r].cnt = c;
        off += c;
    }
}

/* compute_global_std over ranks (utils.jl:569-606). */
static double gstd_ranks(const double *x, const chunk_t *ch, int R, double *scratch) {
#pragma omp parallel for schedule(static, 1) num_threads(R)
    for (int r = 0; r < R; ++r) {
        double s = 0.0, s2 = 0.0;
        const double *p = x + ch[r].off;
        for (int64_t i = 0; i < ch[r].cnt; ++i) {
            double v = p[i];
            s += v;
            s2 += v * v;
        }
        scratch[3 * r] = s;
        scratch[3 * r + 1] = s2;
        scratch[3 * r + 2] = (double)ch[r].cnt;
    }
    double s = 0.0, s2 = 0.0, nn = 0.0;
    for (int r = 0; r < R; ++r) {
        s += scratch[3 * r];
        s2 += scratch[3 * r + 1];
        nn += scratch[3 * r + 2];
    }
    return gstd_from_sums(s, s2, nn);
}

int or_evolve(double *z, double *dE, int64_t n, int64_t nbins, const or_params *p,
              const or_options *o, or_result *res, double *P_last) {
    or_options defaults;
    memset(&defaults, 0, sizeof(defaults));
    if (!o) o = &defaults;
    int R = o->nranks > 0 ? o->nranks : 1;
    if (n < 0) return -1;
    chunk_t *ch = (chunk_t *)malloc(sizeof(chunk_t) * (size_t)R);
    double *scratch = (double *)malloc(sizeof(double) * 3 * (size_t)R);
    rank_chunks(n, R, ch);

    double E0 = p->E0;
    const double mass = p->mass, voltage = p->voltage, radius = p->radius;
    const double alpha_c = p->alpha_c, phis = p->phi_s, freq_rf = p->freq_rf;

    const int serial = o->serial != 0;
    if (serial && p->use_wakefield) {  /* evolution.jl:267: initialize_mpi_buffers! needs an MPI.Comm */
        free(ch); free(scratch); return -5;
    }
    if (o->charge > 0 && o->delta_sigma > 0) { free(ch); free(scratch); return -1; }
    /* evolution.jl:245-253: compute_global_std (MPI) or compute_std (serial) */
    double sigma_E0 = serial ? or_serial_std(dE, n) : gstd_ranks(dE, ch, R, scratch);
    double sigma_z0 = serial ? or_serial_std(z, n) : gstd_ranks(z, ch, R, scratch);

    double *edges = NULL, *centers = NULL, *wtab = NULL, *ltab = NULL, *P = NULL, *pot = NULL;
    int64_t *counts = NULL, *rcounts = NULL;
    double e_first = 0, e_last = 0, inv_step = 0;
    int lag = 0;   /* lag-indexed (causal) wake table */
    if (p->use_wakefield) {
        if (nbins <= 0) { free(ch); free(scratch); return -3; }      /* evolution.jl:264 */
        if (!o->wake_table && o->n_resonators <= 0 && p->pipe_radius <= 0) {
            free(ch); free(scratch); return -4;                           /* :278 */
        }
        if (n == 0) { free(ch); free(scratch); return -2; }          /* log10(0): InexactError */
        double hw = o->bin_half_width_sigmas > 0 ? o->bin_half_width_sigmas : 7.5;
        double zw = hw * sigma_z0;
        if (zw < 1e-9) zw = 1e-6;
        edges = (double *)malloc(sizeof(double) * (size_t)(nbins + 1));
        centers = (double *)malloc(sizeof(double) * (size_t)nbins);
        double step;
        or_bin_grid(zw, nbins, edges, centers, &step);
        e_first = edges[0];
        e_last = edges[nbins];
        inv_step = 1.0 / step;
        wtab = (double *)malloc(sizeof(double) * (size_t)nbins);
        ltab = (double *)malloc(sizeof(double) * (size_t)nbins);
        double wf = (((120.0 * OR_PI) * OR_SPEED_LIGHT) / (OR_PI * p->pipe_radius));
        double ws = sqrt(jl_max(0.0, (2.0 * 30.0) / p->pipe_radius));
        double inv_ctau = 1.0 / 4e-3;
        /* with a charge the line density is counts/n_h: the delta kernel belongs to the
         * reference's dimensionless current heuristic */
        double dsig = o->charge > 0 ? -1.0 : (o->delta_sigma == 0.0 ? 0.15 : o->delta_sigma);
        lag = o->wake_table != NULL || o->n_resonators > 0;
        for (int64_t i = 0; i < nbins; ++i) {
            if (o->wake_table) {
                wtab[i] = o->wake_table[i];
            } else if (o->n_resonators > 0) {
                double s = (double)i * step, w = 0.0;
                for (int q = 0; q < o->n_resonators; ++q)
                    w += or_resonator_wake(s, o->resonators[3 * q], o->resonators[3 * q + 1],
                                           o->resonators[3 * q + 2]);
                wtab[i] = w;
            } else {
                wtab[i] = or_wake_function(centers[i], wf, ws, inv_ctau);
            }
            ltab[i] = dsig > 0 ? or_delta(centers[i], dsig) : 1.0;
        }
        counts = (int64_t *)malloc(sizeof(int64_t) * (size_t)nbins);
        rcounts = (int64_t *)malloc(sizeof(int64_t) * (size_t)nbins * (size_t)R);
        P = (double *)malloc(sizeof(double) * (size_t)nbins);
        pot = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    }
    double *dE_init = p->update_E0 ? (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1)) : NULL;
    double *rnd = p->use_excitation ? (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1)) : NULL;
    double *rsum = (double *)malloc(sizeof(double) * (size_t)R);
    int drift_skipped = 0;
    const double sin_phis = sin(phis);

    for (int64_t t = 0; t < p->n_turns; ++t) {
        /* evolution.jl:300-304 */
        double gamma0 = E0 / mass;
        double beta0 = sqrt(jl_max(0.0, 1.0 - 1.0 / (gamma0 * gamma0)));
        double eta0 = alpha_c - (1.0 / (gamma0 * gamma0));
        double rf = or_calc_rf_factor(freq_rf, beta0);
        double exc_E0 = E0; /* QE and SR use the current E0 */

        if (p->use_excitation && n > 0) {
            if (o->normals) memcpy(rnd, o->normals + (size_t)t * (size_t)n, sizeof(double) * (size_t)n);
        }
        int skip = 0;
        if (!p->update_eta && n > 0)
            skip = (beta0 <= 0.0 || E0 == 0.0 || rf == 0.0);
        if (skip) drift_skipped++;

#pragma omp parallel for schedule(static, 1) num_threads(R)
        for (int r = 0; r < R; ++r) {
            int64_t off = ch[r].off, cnt = ch[r].cnt;
            if (cnt == 0) continue;
            double *zr = z + off, *er = dE + off;
            if (p->update_E0) memcpy(dE_init + off, er, sizeof(double) * (size_t)cnt);
            or_rf_kick(voltage, sin_phis, rf, phis, zr, er, cnt);
            if (p->use_excitation) {
                if (!o->normals) or_normals(o->seed, o->turn_offset + t, off, cnt, rnd + off);
                or_quantum_excitation(exc_E0, radius, sigma_E0, rnd + off, er, cnt);
            }
            if (p->SR_damping) or_synchrotron_radiation(exc_E0, radius, er, cnt);
            if (p->update_eta)
                or_drift_update_eta(zr, er, cnt, mass, gamma0, alpha_c, p->harmonic, beta0, E0, rf, phis);
            else if (!skip)
                or_apply_phase_advance(zr, er, cnt, eta0, p->harmonic, beta0, E0, rf, phis);
        }

        if (p->use_wakefield) {
            double sz = gstd_ranks(z, ch, R, scratch);
            if (o->monitor) {
                double s = 0.0;
                for (int64_t i = 0; i < n; ++i) s += z[i];
                o->monitor[4 * t] = s / (double)n;
                o->monitor[4 * t + 1] = sz;
            }
            double gw = E0 / mass;
            double ew = alpha_c - (1.0 / (gw * gw));
            double current = o->charge > 0 ? o->charge
                                           : or_wake_current(n, E0, radius, sz, ew, sigma_E0);
#pragma omp parallel for schedule(static, 1) num_threads(R)
            for (int r = 0; r < R; ++r)
                or_histogram(z + ch[r].off, ch[r].cnt, e_first, e_last, inv_step, nbins,
                             rcounts + (size_t)r * (size_t)nbins);
            for (int64_t b = 0; b < nbins; ++b) {
                int64_t s = 0;
                for (int r = 0; r < R; ++r) s += rcounts[(size_t)r * (size_t)nbins + b];
                counts[b] = s;
            }
            if (o->monitor) {
                int64_t nh = 0;
                for (int64_t b = 0; b < nbins; ++b) nh += counts[b];
                o->monitor[4 * t + 3] = (double)nh;
            }
            if (lag)
                or_wake_potential_lag(counts, nbins, wtab, ltab, current, P);
            else
                or_wake_potential(counts, nbins, wtab, ltab, current, P);
#pragma omp parallel for schedule(static, 1) num_threads(R)
            for (int r = 0; r < R; ++r) {
                int64_t off = ch[r].off, cnt = ch[r].cnt;
                or_interpolate(z + off, cnt, centers, nbins, P, pot + off);
                for (int64_t i = 0; i < cnt; ++i) dE[off + i] -= pot[off + i];
            }
        }

        if (p->update_E0 && serial) { /* evolution.jl:437-447: compute_mean_diff, n > 0 only */
            if (n > 0) {
                double s = 0.0;
                for (int64_t i = 0; i < n; ++i) s += dE[i] - dE_init[i];
                double mean = s / (double)n;
                E0 = E0 + mean;
                for (int64_t i = 0; i < n; ++i) dE[i] -= mean;
            }
        } else if (p->update_E0) { /* evolution.jl:404-435 */
#pragma omp parallel for schedule(static, 1) num_threads(R)
            for (int r = 0; r < R; ++r) {
                double s = 0.0;
                for (int64_t i = ch[r].off; i < ch[r].off + ch[r].cnt; ++i) s += (dE[i] - dE_init[i]);
                rsum[r] = s;
            }
            double s = 0.0, nn = 0.0;
            for (int r = 0; r < R; ++r) { s += rsum[r]; nn += (double)ch[r].cnt; }
            int64_t ncount = (int64_t)llround(nn);
            double mean = s / (double)ncount;
            E0 = E0 + mean;
            if (n > 0)
                for (int64_t i = 0; i < n; ++i) dE[i] -= mean;
        }
        if (o->monitor) o->monitor[4 * t + 2] = E0;
    }

    /* evolution.jl:461-467 */
    res->sigma_E = serial ? or_serial_std(dE, n) : gstd_ranks(dE, ch, R, scratch);
    res->sigma_z = serial ? or_serial_std(z, n) : gstd_ranks(z, ch, R, scratch);
    res->E0 = E0;
    res->turns_done = p->n_turns;
    res->drift_skipped = drift_skipped;
    if (P_last && P) memcpy(P_last, P, sizeof(double) * (size_t)nbins);

    free(ch); free(scratch); free(edges); free(centers); free(wtab); free(ltab);
    free(counts); free(rcounts); free(P); free(pot); free(dE_init); free(rnd); free(rsum);
    return 0;
}
