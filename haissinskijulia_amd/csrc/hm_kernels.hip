// hm_kernels.hip -- gfx950 kernels of the longitudinal turn loop.
//
// One turn of the reference (src/evolution.jl:295-456) is, per bunch:
//   k_turn_pass   one coalesced read and write of (z, dE) per particle:
//                 the previous turn's wake kick and E0 re-centre, then RF kick
//                 (rf_kick.jl:84-87), quantum excitation (quantum_excitation.jl
//                 :58-71), SR damping (synchrotron_radiation.jl:44-56) and the
//                 drift (evolution.jl:330-347); it bins the new z into an
//                 LDS-privatised histogram (utils.jl:351-401) and writes
//                 deterministic per-wave partial sums of each tile.
//   k_bin_reduce  one workgroup per slice of 1024 of those sums; the last
//                 slice of a bunch to finish adds the slice sums in order:
//                 sigma_z, the wake current (evolution.jl:373-386), the
//                 binned count and the sum of the energy changes.
//   k_conv_*      the line-density convolution with the wake
//                 (wakefield.jl:186-319) by blocked LDS FFTs.
//   k_interp_sum  (update_E0 with the wake) sum of the interpolated potential
//                 over the post-drift z, which the E0 update of
//                 evolution.jl:404-435 needs before the next turn.
//   k_finalize    E0 update and the scalars of the next turn.
//
// All floating-point work is FP64.  Multiplies and adds are never contracted
// (-ffp-contract=off) except inside the elementary functions of hm_math.hpp;
// divides by per-turn constants are multiplies by their reciprocals.  Sums
// are reduced in a fixed order, so a run is bitwise reproducible.
#include <hip/hip_runtime.h>
#include <math.h>

#include "hm_kernels.hpp"

// a*b + k with the polynomial constant k in scalar registers.  Written as
// plain fma, the compiler keeps the ~30 constants of the pass in vector
// registers and copies one into the accumulator before every v_fmac; this
// form needs no copy and leaves the vector registers to the particles.
__device__ __forceinline__ double fma_sk(double a, double b, double k) {
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(k));
    return d;
}
#define HM_QUAL __device__ __forceinline__
#define HM_FMAK(a, b, k) ::fma_sk(a, b, k)
#include "hm_math.hpp"

namespace hm {

struct DevOps {
    static __device__ __forceinline__ double rcp(double a) { return __builtin_amdgcn_rcp(a); }
    static __device__ __forceinline__ double frexp_mant(double u) {
        return __builtin_amdgcn_frexp_mant(u);
    }
    static __device__ __forceinline__ int frexp_exp(double u) { return __builtin_amdgcn_frexp_exp(u); }
};

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

__host__ __device__ inline double jl_max(double a, double b) {
    // Julia's max for floats: NaN wins
    if (a != a) return a;
    if (b != b) return b;
    return (b > a) ? b : a;
}

// compute_global_std from reduced sums (utils.jl:589-605)
__device__ inline double gstd_from_sums(double s, double s2, double nd) {
    const int64_t n = (int64_t)nd;
    if (n <= 1) return 0.0;
    const double mean = s / (double)n;
    const double var = (s2 / (double)n) - (mean * mean);
    return sqrt(jl_max(0.0, var * (double)n / (double)(n - 1)));
}

// Scalars of a turn from the current E0 (evolution.jl:300-304,157;
// synchrotron_radiation.jl:44-45; quantum_excitation.jl:58-59).
__device__ inline void turn_scalars(const BunchDev &B, double E0, TurnScalars &s) {
    s.E0 = E0;
    s.gamma0 = E0 / B.mass;
    const double ig2 = 1.0 / (s.gamma0 * s.gamma0);
    s.beta0 = sqrt(jl_max(0.0, 1.0 - ig2));
    s.eta0 = B.alpha_c - ig2;
    s.rf = (s.beta0 <= 0.0) ? INFINITY : (B.freq_rf * kTwoPi) / (s.beta0 * kSpeedLight);
    s.inv_rf = 1.0 / s.rf;
    s.denom = (s.beta0 * s.beta0) * E0;
    s.kcoef = B.two_pi_h / s.denom;
    s.coeff_c = (B.two_pi_h * s.eta0) / s.denom;
    const double x = E0 / 1e9;
    const double d = ((4.0 * 8.85e-5) * ((x * x) * x)) / B.radius;
    s.damp = 1.0 - d;
    const double om = 1.0 - d;
    s.exc = sqrt(1.0 - (om * om)) * B.sigma_E0;
    s.drift_skip = (!(B.flags & F_ETA)) && (s.beta0 <= 0.0 || E0 == 0.0 || s.rf == 0.0);
}

// sin for the RF kick; arguments far from the bucket (|x| >= 2^19) go to the
// device library, out of line so the common path keeps its registers
__device__ __noinline__ double sin_far(double x) { return sin(x); }

__device__ __forceinline__ double dsin(double x) {
    if (__builtin_expect(!(__builtin_fabs(x) < fm::kReduceMax), 0)) return sin_far(x);
    return fm::sin_reduced(x);
}

// Physics flags of a pass: FX != 0 fixes them at compile time (the common
// all-on configuration), FX == 0 reads them from the bunch.
template <uint32_t FX>
struct Flags {
    uint32_t rt;
    __device__ __forceinline__ bool has(uint32_t f) const {
        return FX ? ((FX & f) != 0) : ((rt & f) != 0);
    }
};

// Philox4x32-10 (Salmon et al., SC'11); the 32x32 -> 64 products give the
// high and low words in one multiply
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t x[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        // three-input xor in one instruction (truth table 0x96 = a ^ b ^ c)
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c1, k0, 0x96);
        const uint32_t n2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c3, k1, 0x96);
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
        // the round keys are uniform: recomputed per round by scalar adds
        // rather than twenty of them held (and spilled) across the pass
        asm volatile("" : "+s"(k0), "+s"(k1));
    }
    x[0] = c0; x[1] = c1; x[2] = c2; x[3] = c3;
}

// sqrt(x) for the Box-Muller radius, 0 <= x <= 74 (u1 >= 2^-53): one
// Goldschmidt step from the hardware reciprocal-square-root estimate and one
// Newton correction, within an ulp of the rounded root.  x = 0 (u1 = 1) gives
// 2^-500 instead of 0.
__device__ __forceinline__ double bm_radius(double x) {
    x = __builtin_fmax(x, 0x1p-1000);
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    const double d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
}

// Box-Muller (1958) on one Philox block: u1 in (0,1], u2 in [0,1), 53 bits each
__device__ __forceinline__ void box_muller(const uint32_t x[4], double &n0, double &n1) {
    const double u1 = ((double)(x[0] >> 5) * 67108864.0 + (double)(x[1] >> 6) + 1.0) * 0x1p-53;
    const double u2 = ((double)(x[2] >> 5) * 67108864.0 + (double)(x[3] >> 6)) * 0x1p-53;
    const double r = bm_radius(-2.0 * fm::log_pos<DevOps>(u1));
    double s, c;
    fm::sincos_reduced(kTwoPi * u2, s, c);
    n0 = r * c;
    n1 = r * s;
}

// QE normals of particle pair p at absolute turn t
__device__ __forceinline__ void qe_pair_normals(uint64_t seed, uint64_t p, uint32_t t, double &n0,
                                                double &n1) {
    uint32_t x[4];
    philox4x32_10((uint32_t)p, (uint32_t)(p >> 32), t, 0u, (uint32_t)seed, (uint32_t)(seed >> 32), x);
    box_muller(x, n0, n1);
}

// zero_alloc_interpolation! index and weight (wakefield.jl:38-42,64-78) for
// nb >= 2.  interp_index clamps z to [c0, c1] by one v_max_f64 and one
// v_min_f64, which map a NaN z to c0; the reference's clamp keeps the NaN, its
// weight u - floor(u) is NaN and so is the interpolated value (cell 0 for the
// index, as the clamp of a converted NaN gives), which interp_index selects
// back.
//
// interp_index_pass, the turn pass's form, clamps the coordinate
// u = (z - c0) * inv to [0, u1], u1 = (c1 - c0) * inv, instead of z to
// [c0, c1]: the same u for every z that is not NaN, since rounding is
// monotonic (z > c1 gives u >= u1, z < c0 gives u <= 0, and the clamped
// z = c0 gives +0).  u is an arithmetic result, so the clamp needs no
// canonicalising of z and the ends, and it is never NaN (max with 0 drops a
// NaN): a NaN z lands on cell 0 with weight 0, and k = min(int(floor(u)),
// nb - 2) lies in [0, nb - 2] without a NaN select or a lower bound.  In the
// turn pass a NaN-z particle's dE turns NaN through the RF term of the same
// pass whatever its kick, so no result depends on that weight (the
// cloud-in-cell deposit counts a NaN z apart).
__device__ __forceinline__ void interp_index_pass(double z, double c0, double u1, double inv,
                                                  int nb, int &k, double &w) {
    const double u = __builtin_fmin(__builtin_fmax((z - c0) * inv, 0.0), u1);
    const double fu = floor(u);
    k = min((int)fu, nb - 2);
    w = u - fu;
}
__device__ __forceinline__ void interp_index(double z, double c0, double c1, double inv, int nb,
                                             int &k, double &w) {
    const double cl = __builtin_fmin(__builtin_fmax(z, c0), c1);
    const double u = (cl - c0) * inv;
    const double fu = floor(u);
    const int kk = (fu == fu) ? (int)fu : 0;
    k = min(max(kk, 0), nb - 2);
    w = (z == z) ? u - fu : z;
}

// the interpolated value (wakefield.jl:94-102)
__device__ __forceinline__ double interp_pot(const double *__restrict__ P, double z, double c0,
                                             double c1, double inv, int nb) {
    if (nb <= 1) return nb == 1 ? P[0] : 0.0;
    int k;
    double w;
    interp_index(z, c0, c1, inv, nb, k, w);
    return P[k] * (1.0 - w) + P[k + 1] * w;
}

// this turn's half of the two-turn histogram buffer
__device__ __forceinline__ uint32_t *hist_of(const BunchDev &B, int64_t turn) {
    return B.hist + (turn & 1) * (int64_t)B.nbins;
}

// histogram bin of calculate_local_histogram! (utils.jl:387-397); -1 = not counted
// (z >= e0 makes z - e0, u and b non-negative: only z = e1 needs the clamp)
__device__ __forceinline__ int hist_bin(double z, double e0, double e1, double inv, int nb) {
    if (!((z >= e0) & (z <= e1))) return -1;
    const double u = (z - e0) * inv;
    const int b = (int)floor(u);
    return min(b, nb - 1);
}

// Deterministic block sum of NV values: wave butterflies, then the waves in
// order.  The result is valid in thread 0.  s_red holds NV * 16 doubles.
template <int NV>
__device__ inline void block_sum(double (&v)[NV], double *s_red) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) s_red[k * 16 + wave] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += s_red[k * 16 + w];
            v[k] = s;
        }
    }
}

// Butterfly sum of NV values over one wave; every lane gets the result.
template <int NV>
__device__ __forceinline__ void wave_sum(double (&v)[NV]) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
    }
}

// One DPP lane move of a double (two 32-bit moves with the same control);
// lanes whose row is not in ROWS receive 0.0.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_move(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWS, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum of NV values over one wave in lane 63, by DPP lane moves (no LDS):
// within each row of 16 lanes by quad permutes (lane ^ 1, lane ^ 2), the
// half-row and the row mirror, which leave the row's sum in all of its lanes
// (each step adds two equal partial sums, in either order the same bits);
// then rows 1 and 3 add rows 0 and 2 (row_bcast:15) and row 3 adds row 1
// (row_bcast:31): lane 63 holds (R3 + R2) + (R1 + R0).  A fixed order.
template <int NV>
__device__ __forceinline__ void wave_total(double (&v)[NV]) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double x = v[k];
        x = x + dpp_move<0xB1, 0xF>(x);    // quad_perm [1,0,3,2]
        x = x + dpp_move<0x4E, 0xF>(x);    // quad_perm [2,3,0,1]
        x = x + dpp_move<0x141, 0xF>(x);   // row_half_mirror
        x = x + dpp_move<0x140, 0xF>(x);   // row_mirror
        x = x + dpp_move<0x142, 0xA>(x);   // row_bcast:15 into rows 1, 3
        x = x + dpp_move<0x143, 0x8>(x);   // row_bcast:31 into row 3
        v[k] = x;
    }
}

// The wave totals of four values, in a fixed order, the total of v[q] in the
// lanes with (lane & 3) == q (lane 60 + q among them).  The quantities are
// transposed across lanes while they are summed: lanes l and l^1 exchange two
// values each and keep the sums of the other two, lanes l and l^2 one, so
// after two steps every lane holds one quantity summed over its quad; four
// more steps (row rotations by 4 and 8, then the rows by lane xor 16 and 32)
// sum it over the wave.  A third of the instructions of four wave_total<1>.
__device__ __forceinline__ double wave_total4(const double (&v)[4]) {
    const int l = (int)threadIdx.x & 63;
    const bool odd = (l & 1) != 0, hi = (l & 2) != 0;
    // quad_perm [1,0,3,2]: quantities 0, 2 (even lanes) or 1, 3 (odd lanes)
    const double ka = (odd ? v[1] : v[0]) + dpp_move<0xB1, 0xF>(odd ? v[0] : v[1]);
    const double kb = (odd ? v[3] : v[2]) + dpp_move<0xB1, 0xF>(odd ? v[2] : v[3]);
    double k = hi ? kb : ka;
    const double s = hi ? ka : kb;
    k = k + dpp_move<0x4E, 0xF>(s);      // quad_perm [2,3,0,1]: quantity l & 3
    k = k + dpp_move<0x124, 0xF>(k);     // row_ror:4
    k = k + dpp_move<0x128, 0xF>(k);     // row_ror:8
    k = k + __shfl_xor(k, 16, 64);
    k = k + __shfl_xor(k, 32, 64);
    return k;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ int bitrev(int m, int logF) {
    return (int)(__brev((unsigned)m) >> (32 - logF));
}

__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
    return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) {
    return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double2 conj2(double2 a) { return make_double2(a.x, -a.y); }

// In-place radix-2 decimation-in-time FFT of X[F] in LDS, bit-reversed input,
// natural-order output, twiddles tw[k] = exp(-2 pi i k / F).  Stage s pairs
// X[i] and X[i + 2^s] (i with bit s clear, k = i mod 2^s) with twiddle
// tw[k << (logF-1-s)].  The stages run in pairs (s, s+1) as radix-2^2
// butterflies: each thread takes the four points base + {0,1,2,3}*2^s
// through both stages in registers, the same additions and products as the
// two radix-2 stages (the same bits) with half the barriers and LDS round
// trips; an odd stage count starts with one radix-2 stage.  nblk blocks of F
// points, X + b*F, are transformed side by side.
__device__ void fft_dit_forward(double2 *X, int F, int logF, const double2 *__restrict__ tw,
                                int nblk = 1) {
    int s = 0;
    if (logF & 1) {   // stage 0 alone: pairs (2b, 2b+1), twiddle tw[0]
        __syncthreads();
        for (int b = threadIdx.x; b < nblk * (F >> 1); b += blockDim.x) {
            const double2 a = X[2 * b];
            const double2 t = cmul(X[2 * b + 1], tw[0]);
            X[2 * b] = cadd(a, t);
            X[2 * b + 1] = csub(a, t);
        }
        s = 1;
    }
    for (; s + 1 < logF; s += 2) {
        const int h = 1 << s, tsh = logF - 1 - s;
        __syncthreads();
        for (int u = threadIdx.x; u < nblk * (F >> 2); u += blockDim.x) {
            const int q = u & ((F >> 2) - 1);
            const int k = q & (h - 1);
            const int i0 = ((u >> (logF - 2)) << logF) + (((q >> s) << (s + 2)) | k);
            const double2 w1 = tw[k << tsh];              // stage s, both pairs
            const double2 wa = tw[k << (tsh - 1)];        // stage s+1, pair (0, 2)
            const double2 wb = tw[(k + h) << (tsh - 1)];  // stage s+1, pair (1, 3)
            const double2 x0 = X[i0], x1 = X[i0 + h], x2 = X[i0 + 2 * h], x3 = X[i0 + 3 * h];
            double2 t = cmul(x1, w1);
            const double2 a0 = cadd(x0, t), a1 = csub(x0, t);
            t = cmul(x3, w1);
            const double2 a2 = cadd(x2, t), a3 = csub(x2, t);
            t = cmul(a2, wa);
            X[i0] = cadd(a0, t);
            X[i0 + 2 * h] = csub(a0, t);
            t = cmul(a3, wb);
            X[i0 + h] = cadd(a1, t);
            X[i0 + 3 * h] = csub(a1, t);
        }
    }
    __syncthreads();
}

// Unnormalised inverse FFT (decimation in frequency, conjugate twiddles):
// natural-order input, output y[m] at X[bitrev(m)].  Stage s (from logF-1
// down to 0) takes a = X[i], c = X[i + 2^s] to a + c and (a - c) conj(w),
// w = tw[k << (logF-1-s)]; stages run in pairs (s, s-1) as radix-2^2
// butterflies on base + {0, 2^(s-1), 2^s, 2^s + 2^(s-1)}, with the same
// operations as the radix-2 stages, and an odd stage count ends with stage 0
// alone.  nblk blocks side by side, as in fft_dit_forward.
__device__ void fft_dif_inverse(double2 *X, int F, int logF, const double2 *__restrict__ tw,
                                int nblk = 1) {
    int s = logF - 1;
    for (; s >= 1; s -= 2) {
        const int h = 1 << s, h2 = h >> 1, tsh = logF - 1 - s;
        __syncthreads();
        for (int u = threadIdx.x; u < nblk * (F >> 2); u += blockDim.x) {
            const int q = u & ((F >> 2) - 1);
            const int k = q & (h2 - 1);
            const int i0 = ((u >> (logF - 2)) << logF) + (((q >> (s - 1)) << (s + 1)) | k);
            const double2 wa = conj2(tw[k << tsh]);          // stage s, pair (0, 2)
            const double2 wb = conj2(tw[(k + h2) << tsh]);   // stage s, pair (1, 3)
            const double2 wc = conj2(tw[k << (tsh + 1)]);    // stage s-1, both pairs
            const double2 x0 = X[i0], x1 = X[i0 + h2], x2 = X[i0 + h], x3 = X[i0 + h + h2];
            const double2 a0 = cadd(x0, x2), a2 = cmul(csub(x0, x2), wa);
            const double2 a1 = cadd(x1, x3), a3 = cmul(csub(x1, x3), wb);
            X[i0] = cadd(a0, a1);
            X[i0 + h2] = cmul(csub(a0, a1), wc);
            X[i0 + h] = cadd(a2, a3);
            X[i0 + h + h2] = cmul(csub(a2, a3), wc);
        }
    }
    if (s == 0) {   // stage 0 alone: pairs (2b, 2b+1), twiddle conj(tw[0])
        __syncthreads();
        const double2 w = conj2(tw[0]);
        for (int b = threadIdx.x; b < nblk * (F >> 1); b += blockDim.x) {
            const double2 a = X[2 * b], c = X[2 * b + 1];
            X[2 * b] = cadd(a, c);
            X[2 * b + 1] = cmul(csub(a, c), w);
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// per-particle turn (fused)
// ---------------------------------------------------------------------------

struct PassCtx {
    double rf, inv_rf, phis, sin_phis, voltage, exc, damp, inv_mass, gamma0, alpha_c, kcoef,
        coeff_c, m_prev;
    int pending_kick, pending_m, drift_skip;
    uint32_t flags;
    const double *P;
    double c0, c1, u1, inv_s, e0, e1, inv_step;
    int nb;
};

__device__ inline void load_ctx(const BunchDev &B, PassCtx &c) {
    const TurnScalars &s = *B.sc;
    c.rf = s.rf;
    c.inv_rf = s.inv_rf;
    c.phis = B.phis;
    c.sin_phis = B.sin_phis;
    c.voltage = B.voltage;
    c.exc = s.exc;
    c.damp = s.damp;
    c.inv_mass = B.inv_mass;
    c.gamma0 = s.gamma0;
    c.alpha_c = B.alpha_c;
    c.kcoef = s.kcoef;
    c.coeff_c = s.coeff_c;
    // 0 when no re-centre is pending: subtracting it then changes nothing
    // (x - 0 = x for every x, -0 and NaN included)
    c.m_prev = s.pending_m ? s.m_prev : 0.0;
    c.pending_kick = s.pending_kick;
    c.pending_m = s.pending_m;
    c.drift_skip = s.drift_skip;
    c.flags = B.flags;
    c.P = B.P;
    c.c0 = B.c_first;
    c.c1 = B.c_last;
    c.u1 = B.u_last;
    c.inv_s = B.inv_s;
    c.e0 = B.e_first;
    c.e1 = B.e_last;
    c.inv_step = B.inv_step;
    c.nb = B.nbins;
}

// The previous turn's wake kick (wakefield.jl:342-344) and E0 re-centre
// (evolution.jl:434), applied at the start of this pass.
__device__ __forceinline__ double apply_pending(const PassCtx &c, double z, double e) {
    if (c.pending_kick) e = e - interp_pot(c.P, z, c.c0, c.c1, c.inv_s, c.nb);
    if (c.pending_m) e = e - c.m_prev;
    return e;
}

// phase advance (evolution.jl:331-344 with update_eta, :159-165 without)
template <uint32_t FX>
__device__ __forceinline__ double drift(const PassCtx &c, Flags<FX> fl, double z, double e) {
    if (fl.has(F_ETA)) {
        const double gp = c.gamma0 + e * c.inv_mass;
        // 1/gp^2 ~ 3e-8 enters eta ~ 1e-3 as a correction: one Newton step
        const double eta = c.alpha_c - fm::rcp1<DevOps>(gp * gp);
        const double coeff = eta * c.kcoef;
        double phi = -((z * c.rf) - c.phis);
        phi = phi + coeff * e;
        return ((-phi) + c.phis) * c.inv_rf;
    } else if (!c.drift_skip) {
        double phi = -((z * c.rf) - c.phis);
        phi = phi + c.coeff_c * e;
        return ((-phi) + c.phis) * c.inv_rf;
    }
    return z;
}

// The RF kick's energy change of a particle at z (rf_kick.jl:85-86); it
// does not depend on the particle's energy.
__device__ __forceinline__ double rf_term(const PassCtx &c, double z) {
    const double phi0 = ((-z) * c.rf) + c.phis;                    // rf_kick.jl:85
    return c.voltage * (dsin(phi0) - c.sin_phis);                  // rf_kick.jl:86
}

// RF (its term precomputed by rf_term), QE, SR and drift of one particle;
// returns dE' - dE_init via de_turn.
template <uint32_t FX>
__device__ __forceinline__ void advance_rest(const PassCtx &c, Flags<FX> fl, double &z, double &e,
                                             double rf, double r, double &de_turn) {
    const double e_init = e;
    e = e + rf;                                                    // rf_kick.jl:86
    if (fl.has(F_QE)) e = e + c.exc * r;                           // quantum_excitation.jl:70
    if (fl.has(F_SR)) e = e * c.damp;                              // synchrotron_radiation.jl:55
    z = drift(c, fl, z, e);
    de_turn = e - e_init;
}

// RF, QE, SR and drift of one particle; returns dE' - dE_init via de_turn.
template <uint32_t FX>
__device__ __forceinline__ void advance(const PassCtx &c, Flags<FX> fl, double &z, double &e,
                                        double r, double &de_turn) {
    advance_rest(c, fl, z, e, rf_term(c, z), r, de_turn);
}

// Explicit address spaces.  Pointers read from a BunchDev are generic to the
// compiler, and a choice between LDS and global memory merges into one
// generic access: both become flat instructions, which wait on the LDS and
// the vector-memory counters together.  The pass therefore names global
// memory (address space 1) and reaches its LDS through the kernel's own
// dynamic-LDS symbol.
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) double g_f64;
typedef __attribute__((address_space(1))) f64x2 g_f64x2;
typedef __attribute__((address_space(1))) uint32_t g_u32;
typedef __attribute__((address_space(1))) unsigned long long g_u64;

__device__ __forceinline__ double2 ld2(const g_f64x2 *a) {
    const f64x2 v = *a;
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ void st2(g_f64x2 *a, double x, double y) {
    f64x2 v;
    v.x = x;
    v.y = y;
    *a = v;
}
// streaming (nontemporal) 16-byte load: for arrays far larger than the caches
__device__ __forceinline__ double2 ldnt(const g_f64x2 *a) {
    const f64x2 v = __builtin_nontemporal_load(a);
    return make_double2(v.x, v.y);
}
// streaming (nontemporal) 16-byte store
__device__ __forceinline__ void stnt(g_f64x2 *a, double x, double y) {
    f64x2 v;
    v.x = x;
    v.y = y;
    __builtin_nontemporal_store(v, a);
}
// P[k]*(1-w) + P[k+1]*w from the neighbour-pair table entry (P[k], P[k+1])
__device__ __forceinline__ double gather_interp(const g_f64x2 *q, double w) {
    const f64x2 v = *q;
    return v.x * (1.0 - w) + v.y * w;
}

__device__ __forceinline__ uint8_t *pass_lds() {
    extern __shared__ __align__(16) uint8_t s_lds[];
    return s_lds;
}

// LH: the histogram is in LDS (else straight to global memory)
template <bool LH>
__device__ __forceinline__ void bin_and_sum(const PassCtx &c, g_u32 *ghist, double z,
                                            double (&v)[4]) {
    v[0] += z;
    v[1] += z * z;
    // hist_bin with its range test as the branch: in range, u = (z - e0) * inv
    // is >= 0, and its conversion to int (truncation) is floor(u)
    if ((z >= c.e0) & (z <= c.e1)) {
        const int b = min((int)((z - c.e0) * c.inv_step), c.nb - 1);
        if (LH)
            atomicAdd(reinterpret_cast<uint32_t *>(pass_lds()) + b, 1u);
        else
            __hip_atomic_fetch_add(ghist + b, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        v[3] += 1.0;
    }
}

// The cloud-in-cell deposit of one post-drift z (F_CIC) into the LDS counters
// at byte offset cb: the cell and weight the next pass's kick takes for it
// (interp_index_pass).  round(w 2^51) is the low 52 bits of the double
// 2^52 + w 2^51, formed by one fused multiply-add (one rounding; w in [0, 1),
// so the sum stays below 2^53).  A NaN z or a non-finite weight is counted in
// n[nb] instead.
__device__ __forceinline__ void cic_deposit(const PassCtx &c, int cb, double z) {
    int k;
    double w;
    interp_index_pass(z, c.c0, c.u1, c.inv_s, c.nb, k, w);
    unsigned long long *sH = reinterpret_cast<unsigned long long *>(pass_lds() + cb);
    unsigned long long *sL = sH + c.nb;
    uint32_t *sN = reinterpret_cast<uint32_t *>(sL + c.nb);
    if ((z == z) & (w == w)) {
        const double y = __builtin_fma(w, kCicScale, 2.0 * kCicScale);
        const unsigned long long W =
            (unsigned long long)__double_as_longlong(y) & ((1ull << 52) - 1);   // <= 2^51
        atomicAdd(sN + k, 1u);
        atomicAdd(sH + k, W >> kCicLoBits);
        atomicAdd(sL + k, W & ((1ull << kCicLoBits) - 1));
    } else {
        atomicAdd(sN + c.nb, 1u);
    }
}

// The fused pass over a run of tiles.  FX = kStdFlags is the common configuration
// with the flags folded at compile time; FX = 0 handles any other.  Both
// evaluate the same expressions, so a bunch gets the same bits from either.
// LH: histogram in LDS at offset 0; the window of P follows at byte hb.
//
// One call covers the run of tiles [lt0, lt1) of bunch B that this workgroup
// owns.  Tiles are whole multiples of the block size, so thread t's pairs in
// tile lt are lt*ppt + t + k*kPassThreads in ascending k, exactly as for a
// tile taken alone: the loop keeps its one-pair prefetch running across the
// tile boundaries and writes each tile's per-wave sums at the tile's end.
// The run's tiles are run[0..cnt) (indices into the tile table, in ascending
// order but not necessarily adjacent: a bunch shared by several workgroups
// has its tiles dealt to them in turn).
// KM, how the kick reads P: kFullWindow (the LDS window holds the whole table:
// no global gathers), kGatherFirst or kPrefetchFirst (a partial window, the
// particles outside it gather from global memory; see the loop body).
// CIC: the bunch's post-drift z are also deposited (cic_deposit) into the LDS
// counters at byte cb.
enum KickMode { kFullWindow = 0, kGatherFirst = 1, kPrefetchFirst = 2 };
template <uint32_t FX, bool LH, int KM, bool CIC>
__device__ __forceinline__ void turn_pass_run(const BunchDev &B, const int32_t *__restrict__ run,
                                              int cnt, int hb, int cb) {
    PassCtx c;
    load_ctx(B, c);
    const Flags<FX> fl{c.flags};
    const bool wake = fl.has(F_WAKE);
    const TurnScalars &sc = *B.sc;
    const int64_t n = B.n;
    const int64_t ppt = B.pairs_per_tile;
    const int64_t full = n >> 1;                    // pairs with two particles
    const int tb = B.tile_begin;
    const uint32_t tabs = (uint32_t)(B.turn_offset + sc.turn);
    const bool inj = fl.has(F_INJ);
    const bool qe = fl.has(F_QE);
    const g_f64 *nrm = inj ? (const g_f64 *)(B.normals + sc.turn * n) : nullptr;
    const uint64_t seed = B.seed;
    const bool kick2 = c.pending_kick && c.nb >= 2;
    const int wlo = window_lo(sc.centre_bin, B.win, c.nb);
    const unsigned wpairs = (unsigned)max(B.win - 1, 0);   // pair starts inside the window
    const double *sP = reinterpret_cast<const double *>(pass_lds() + hb);
    const g_f64x2 *P2 = (const g_f64x2 *)B.P2;
    g_f64 *zp = (g_f64 *)B.z;
    g_f64 *ep = (g_f64 *)B.dE;
    g_u32 *ghist = (g_u32 *)hist_of(B, sc.turn);

    double2 zn = make_double2(0.0, 0.0), en = make_double2(0.0, 0.0);
    // The pair stride is the block size, as the compile-time constant: read
    // from the dispatch packet, blockDim.x was reloaded per tile with a vector
    // load and a full wait (1e8 pass +2%).
    const int64_t step = kPassThreads;
    // per-wave partial sums of a tile: no workgroup barrier per tile, so a wave
    // never waits for the others of its workgroup between tiles
    auto flush = [&](const double (&v)[4], int lt) {
        const double tot = wave_total4(v);
        const int lane = (int)threadIdx.x & 63;
        if (lane >= 60)
            B.part[4 * ((int64_t)lt * kTileWaves + (threadIdx.x >> 6)) + (lane - 60)] = tot;
    };
    // the single particle of an odd bunch sits in the bunch's last tile, the
    // last of this run if the run has it: that tile is flushed after the loop
    const int lt_last = run[cnt - 1] - tb;
    const bool odd_tile = (n & 1) && full >= (int64_t)lt_last * ppt &&
                          full < min((int64_t)(lt_last + 1) * ppt, B.npairs);
    double v[4];   // sum z, sum z^2, sum (dE'-dE_init), in-range count
    bool reload = true;   // this tile's first pair was not prefetched
    for (int t = 0; t < cnt; ++t) {
        const int lt = run[t] - tb;
        const int64_t p0 = (int64_t)lt * ppt;
        const int64_t p1 = min(p0 + ppt, B.npairs);
        const int64_t pf = min(p1, full);
        if (reload) {
            const int64_t p = p0 + threadIdx.x;
            if (p < pf) {
                zn = ldnt((const g_f64x2 *)(zp + 2 * p));
                en = ldnt((const g_f64x2 *)(ep + 2 * p));
            }
            // the tile's first pair completes here (s_waitcnt vmcnt(0)), so that
            // in the loop the pair in hand is never taken for an outstanding
            // load and the waits there count only the loads issued in the loop
            __builtin_amdgcn_s_waitcnt(0x0F70);
        }
        // where the prefetch goes after this tile's last pair: the next tile's
        // first pair of this thread when every thread has one there; else this
        // tile's first pair, read again and not used (and the next tile, the
        // bunch's short last one, loads its first pair itself)
        int64_t nxt = p0;
        reload = false;
        if (t + 1 < cnt) {
            const int64_t q0 = (int64_t)(run[t + 1] - tb) * ppt;
            if (min(q0 + ppt, full) - q0 >= step)
                nxt = q0;
            else
                reload = true;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = 0.0;
        for (int64_t p = p0 + threadIdx.x; p < pf; p += step) {
            const double2 zz = zn, ee = en;
            // The previous turn's potential at (P[k], P[k+1]): read from the LDS
            // window (clamped index, always valid), replaced by one 16-byte
            // global gather for the particles outside a partial window.  With
            // gather_first the gathers are issued before the next pair's
            // prefetch and consumed after the RF and QE terms, whose
            // arithmetic covers their latency.  Otherwise the prefetch goes
            // first.  The order of the loads changes no result.
            constexpr bool gather = KM != kFullWindow;
            constexpr bool gather_first = KM == kGatherFirst;
            const int64_t pn = p + step;
            // the next pair (the next tile's first at a tile end); past the end
            // of the run a pair of this tile is read again and not used, so the
            // load is unconditional and zn, en need no copy of their old value
            auto prefetch = [&]() {
                const int64_t pq = pn < pf ? pn : nxt + threadIdx.x;
                zn = ldnt((const g_f64x2 *)(zp + 2 * pq));
                en = ldnt((const g_f64x2 *)(ep + 2 * pq));
            };
            if (!gather_first) prefetch();
            // set and read only when kick2 (no per-pair clearing)
            double w0, w1;
            double2 q0, q1;
            if (kick2) {
                int k0, k1;
                interp_index_pass(zz.x, c.c0, c.u1, c.inv_s, c.nb, k0, w0);
                interp_index_pass(zz.y, c.c0, c.u1, c.inv_s, c.nb, k1, w1);
                const unsigned j0 = (unsigned)(k0 - wlo), j1 = (unsigned)(k1 - wlo);
                // a full window, and any window of the all-flags pass (its
                // histogram is in LDS, so its window has at least two entries),
                // has wpairs >= 1
                if (KM == kFullWindow || FX != 0 || wpairs) {
                    const unsigned i0 = min(j0, wpairs - 1), i1 = min(j1, wpairs - 1);
                    q0 = make_double2(sP[i0], sP[i0 + 1]);
                    q1 = make_double2(sP[i1], sP[i1 + 1]);
                }
                if (gather) {
                    if (j0 >= wpairs) q0 = ld2(P2 + k0);
                    if (j1 >= wpairs) q1 = ld2(P2 + k1);
                }
            }
            if (gather_first) prefetch();
            double rf0 = 0.0, rf1 = 0.0;
            if (gather_first) {
                rf0 = rf_term(c, zz.x);
                rf1 = rf_term(c, zz.y);
            }
            double r0 = 0.0, r1 = 0.0;
            if (qe) {
                if (inj) {   // turn rows of an odd bunch are only 8-byte aligned
                    r0 = nrm[2 * p];
                    r1 = nrm[2 * p + 1];
                } else {
                    qe_pair_normals(seed, (uint64_t)p, tabs, r0, r1);
                }
            }
            double z0 = zz.x, z1 = zz.y, e0 = ee.x, e1 = ee.y;
            if (c.pending_kick) {
                if (kick2) {
                    e0 = e0 - (q0.x * (1.0 - w0) + q0.y * w0);
                    e1 = e1 - (q1.x * (1.0 - w1) + q1.y * w1);
                } else {
                    e0 = e0 - c.P[0];
                    e1 = e1 - c.P[0];
                }
            }
            // with update_E0 every turn but the first has a re-centre pending,
            // and m_prev is 0 when none is: the all-flags pass subtracts it
            // without a select
            if (fl.has(F_E0) || c.pending_m) {
                e0 = e0 - c.m_prev;
                e1 = e1 - c.m_prev;
            }
            double d0, d1;
            if (gather_first) {
                advance_rest(c, fl, z0, e0, rf0, r0, d0);
                advance_rest(c, fl, z1, e1, rf1, r1, d1);
            } else {
                advance(c, fl, z0, e0, r0, d0);
                advance(c, fl, z1, e1, r1, d1);
            }
            if (wake) {
                bin_and_sum<LH>(c, ghist, z0, v);
                bin_and_sum<LH>(c, ghist, z1, v);
            }
            if (CIC) {
                cic_deposit(c, cb, z0);
                cic_deposit(c, cb, z1);
            }
            v[2] += d0;
            v[2] += d1;
            // The prefetched pair is complete before this pair's stores are
            // issued: a wait for a load behind a store would wait for the store.
            __builtin_amdgcn_s_waitcnt(0x0F70);
            // streaming stores: the bunch is far larger than the caches, which
            // are left to the bin phase's tables
            stnt((g_f64x2 *)(zp + 2 * p), z0, z1);
            stnt((g_f64x2 *)(ep + 2 * p), e0, e1);
        }
        if (odd_tile && t == cnt - 1) break;   // v kept for the single particle
        if (!wake) { v[0] = 0.0; v[1] = 0.0; }
        flush(v, lt);
    }
    // the single particle of an odd bunch (last pair), out of the tile loop so
    // that its inlined turn does not add to the loop's register pressure
    if (odd_tile) {
        if (threadIdx.x == 0) {
            const int64_t i = n - 1;
            double z0 = zp[i], e0 = ep[i], r0 = 0.0, r1 = 0.0;
            if (qe) {
                if (inj) r0 = nrm[i];
                else qe_pair_normals(seed, (uint64_t)full, tabs, r0, r1);
            }
            e0 = apply_pending(c, z0, e0);
            double d0;
            advance(c, fl, z0, e0, r0, d0);
            if (wake) bin_and_sum<LH>(c, ghist, z0, v);
            if (CIC) cic_deposit(c, cb, z0);
            v[2] += d0;
            zp[i] = z0;
            ep[i] = e0;
        }
        if (!wake) { v[0] = 0.0; v[1] = 0.0; }
        flush(v, lt_last);
    }
}

// Adds the LDS histogram of bunch B into this turn's global counts.
__device__ inline void flush_hist(const BunchDev &B) {
    if (!pass_hist_bytes(B)) return;
    __syncthreads();
    const uint32_t *s_hist = reinterpret_cast<const uint32_t *>(pass_lds());
    g_u32 *g = (g_u32 *)hist_of(B, B.sc->turn);
    for (int b = threadIdx.x; b < B.nbins; b += blockDim.x) {
        const uint32_t h = s_hist[b];
        if (h) __hip_atomic_fetch_add(g + b, h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (B.flags & F_CIC) {   // and the deposit into this turn's counters
        const int nb = B.nbins;
        const unsigned long long *sH =
            reinterpret_cast<const unsigned long long *>(pass_lds() + B.cic_off);
        const uint32_t *sN = reinterpret_cast<const uint32_t *>(sH + 2 * nb);
        g_u64 *gH = (g_u64 *)B.cic;
        g_u32 *gN = (g_u32 *)(B.cic + 2 * nb);
        for (int b = threadIdx.x; b <= nb; b += blockDim.x) {
            const uint32_t cnt = sN[b];
            if (!cnt) continue;
            __hip_atomic_fetch_add(gN + b, cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (b < nb) {
                __hip_atomic_fetch_add(gH + b, sH[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_add(gH + nb + b, sH[nb + b], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

// Clears the LDS histogram of bunch B and loads its potential window.
__device__ inline void stage_bunch(const BunchDev &B) {
    __syncthreads();   // the previous bunch's LDS is no longer read
    const int hb = pass_hist_bytes(B);
    uint32_t *h = reinterpret_cast<uint32_t *>(pass_lds());
    for (int b = threadIdx.x; b < hb / 4; b += blockDim.x) h[b] = 0u;
    if (B.flags & F_CIC) {
        uint32_t *sC = reinterpret_cast<uint32_t *>(pass_lds() + B.cic_off);
        for (int b = threadIdx.x; b < cic_bytes(B.nbins) / 4; b += blockDim.x) sC[b] = 0u;
    }
    const TurnScalars &sc = *B.sc;
    if (sc.pending_kick && B.win > 0) {
        double *sP = reinterpret_cast<double *>(pass_lds() + hb);
        const g_f64 *P = (const g_f64 *)(B.P + window_lo(sc.centre_bin, B.win, B.nbins));
        for (int k = threadIdx.x; k < B.win; k += blockDim.x) sP[k] = P[k];
    }
    __syncthreads();
}

// The fused pass over a tile table.  Each workgroup takes a contiguous range
// of entries of the pass's tile order, with about the same number of pairs as
// every other workgroup's (one workgroup per CU), stages a bunch's LDS once
// for its run of the bunch's tiles in the range, streams the run and flushes
// the histogram when it moves on.  The per-tile sums depend on the tile and
// the block size only, not on which workgroup takes the tile.
template <uint32_t FX>
__device__ __forceinline__ void turn_pass_blocks(const BunchDev *__restrict__ bds,
                                                 const int32_t *__restrict__ tile_bunch,
                                                 const int32_t *__restrict__ order,
                                                 const int32_t *__restrict__ ranges) {
    // kStdFlags has the wake on and F_GHIST off: its histogram is in LDS
    constexpr bool kLdsOnly = (FX & F_WAKE) && !(FX & F_GHIST);
    const int e0 = ranges[blockIdx.x];
    const int e1 = ranges[blockIdx.x + 1];
    int staged = -1;
    for (int e = e0; e < e1;) {
        const int bi = tile_bunch[order[e]];
        const BunchDev &B = bds[bi];
        // the run: this workgroup's consecutive entries of the bunch
        int ee = e + 1;
        while (ee < e1 && tile_bunch[order[ee]] == bi) ++ee;
        if (bi != staged) {
            if (staged >= 0) flush_hist(bds[staged]);
            stage_bunch(B);
            staged = bi;
        }
        const int hb = pass_hist_bytes(B);
        const int32_t *run = order + e;
        const int cnt = ee - e;
        // a partial window of P: gather first in the pass with compile-time
        // flags (only it has the registers), prefetch first in the generic one
        constexpr int kPartial = FX != 0 ? kGatherFirst : kPrefetchFirst;
        const bool full = B.win >= B.nbins;
        const int cb = B.cic_off;
        if (kLdsOnly || hb) {
            if (B.flags & F_CIC) {   // the deposit needs the LDS histogram's room
                if (full)
                    turn_pass_run<FX, true, kFullWindow, true>(B, run, cnt, hb, cb);
                else
                    turn_pass_run<FX, true, kPartial, true>(B, run, cnt, hb, cb);
            } else {
                if (full)
                    turn_pass_run<FX, true, kFullWindow, false>(B, run, cnt, hb, cb);
                else
                    turn_pass_run<FX, true, kPartial, false>(B, run, cnt, hb, cb);
            }
        } else {
            if (full)
                turn_pass_run<FX, false, kFullWindow, false>(B, run, cnt, hb, cb);
            else
                turn_pass_run<FX, false, kPrefetchFirst, false>(B, run, cnt, hb, cb);
        }
        e = ee;
    }
    if (staged >= 0) flush_hist(bds[staged]);
}

__global__ __launch_bounds__(kPassThreads) void k_turn_pass(const BunchDev *__restrict__ bds,
                                                            const int32_t *__restrict__ tile_bunch,
                                                            const int32_t *__restrict__ order,
                                                            const int32_t *__restrict__ ranges) {
    turn_pass_blocks<0>(bds, tile_bunch, order, ranges);
}

// the same for tables whose bunches all have exactly kStdFlags
__global__ __launch_bounds__(kPassThreads) void k_turn_pass_std(
    const BunchDev *__restrict__ bds, const int32_t *__restrict__ tile_bunch,
    const int32_t *__restrict__ order, const int32_t *__restrict__ ranges) {
    turn_pass_blocks<kStdFlags>(bds, tile_bunch, order, ranges);
}

// ---------------------------------------------------------------------------
// bin phase: sigma_z, current, line density, blocked FFT convolution
// ---------------------------------------------------------------------------
//
// P_k = I * sum_{j<=k} W_j g_{k-j},  g_i = lam_i * (cnt_i / n_h).
// With block size B and F = 2B, output block o is
//   Re(IFFT(Y_o))[0,B) + Im(IFFT(Y_o))[B,2B),
//   Y_o = sum_j X_{o-j} FFT(W_j),   X_a = FFT(g_a + i g_{a-1}),
// since each block product g_a * W_j is real, has length 2B-1 <= F (no
// wrap), and lands on output blocks a+j (first half) and a+j+1 (second half).

// One slice of a per-bunch reduction: the sum of entries [r*kRedEntries,
// (r+1)*kRedEntries) of a per-wave partial array of n entries, NV doubles
// each (NV = 1 or 4).  Thread t adds entries r*kRedEntries + t + k*kRedThreads
// in ascending k (all loads in flight at once), then block_sum adds the
// threads in its fixed order: the slice sum depends on n and r only.  Valid
// in thread 0.
template <int NV>
__device__ inline void slice_sum(const double *a, int n, int r, double (&v)[NV], double *s_red) {
    const int e0 = r * kRedEntries + (int)threadIdx.x;
    double x[kRedPer][NV];
#pragma unroll
    for (int k = 0; k < kRedPer; ++k) {
        const int e = e0 + k * kRedThreads;
        if constexpr (NV == 4) {
            double2 lo = make_double2(0.0, 0.0), hi = make_double2(0.0, 0.0);
            if (e < n) {
                const g_f64x2 *q = (const g_f64x2 *)a + 2 * (int64_t)e;
                lo = ld2(q);
                hi = ld2(q + 1);
            }
            x[k][0] = lo.x;
            x[k][1] = lo.y;
            x[k][2] = hi.x;
            x[k][3] = hi.y;
        } else {
            x[k][0] = e < n ? ((const g_f64 *)a)[e] : 0.0;
        }
    }
#pragma unroll
    for (int k = 0; k < kRedPer; ++k)
#pragma unroll
        for (int j = 0; j < NV; ++j) v[j] += x[k][j];
    block_sum<NV>(v, s_red);
}

// Counts slice r of bunch B in, its NV sums (thread 0's v) first stored at
// red[r].  Returns true in every thread of the workgroup that counts the
// bunch's last slice in; there thread 0 receives in t the sums of all slices,
// added in slice order, and the count is reset for the next reduction.  The
// count's read-modify-write is acquire-release at device scope, so the last
// slice sees every slice's stores whichever XCD wrote them; its threads read
// the slice sums side by side (one round trip), thread 0 adds them.
template <int NV>
__device__ inline bool slices_done(const BunchDev &B, int r, const double (&v)[NV],
                                   double (&t)[NV]) {
    __shared__ int s_last;
    __shared__ double s_slice[kRedThreads][NV];
    if (B.nred == 1) {   // one slice: nothing to count or to exchange
#pragma unroll
        for (int j = 0; j < NV; ++j) t[j] = v[j];
        return true;
    }
    g_f64 *red = (g_f64 *)B.red;
    g_u32 *cnt = (g_u32 *)B.cnt;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NV; ++j)
            __hip_atomic_store(red + 4 * r + j, v[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t before =
            __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        s_last = before + 1 == (uint32_t)B.nred;
    }
    __syncthreads();
    if (!s_last) return false;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
#pragma unroll
    for (int j = 0; j < NV; ++j) t[j] = 0.0;
    for (int q0 = 0; q0 < B.nred; q0 += kRedThreads) {
        const int q = q0 + (int)threadIdx.x;
        if (q < B.nred) {
            double x[NV];
#pragma unroll
            for (int j = 0; j < NV; ++j)
                x[j] = __hip_atomic_load(red + 4 * q + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
            for (int j = 0; j < NV; ++j) s_slice[threadIdx.x][j] = x[j];
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int k = 0; k < kRedThreads && q0 + k < B.nred; ++k)
#pragma unroll
                for (int j = 0; j < NV; ++j) t[j] += s_slice[k][j];
        __syncthreads();
    }
    if (threadIdx.x == 0) __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

// The turn scalars from the reduced pass sums t = {sum z, sum z^2,
// sum (dE'-dE_init), binned count}.
// bin_scalars_on updates the turn scalars sc in place (the bin phase in one
// workgroup keeps them in registers).
__device__ inline void bin_scalars_on(const BunchDev &B, const double (&t)[4], TurnScalars &sc) {
    sc.sum_dE = t[2];
    if (!(B.flags & F_WAKE)) return;
    const double sz = gstd_from_sums(t[0], t[1], (double)B.n);
    const double mz = t[0] / (double)B.n;
    sc.sigma_z = sz;
    sc.current = (B.flags & F_CHARGE)
                     ? B.charge
                     : (((B.pf / sc.E0) / (kTwoPi * B.radius)) * sz) /
                           (sc.eta0 * (B.sigma_E0 * B.sigma_E0));
    sc.n_h = t[3];
    const double kc = floor((mz - B.c_first) * B.inv_s);
    const double top = (double)(B.nbins - 1);
    sc.centre_bin = (kc == kc) ? (int)(kc < 0.0 ? 0.0 : (kc > top ? top : kc)) : 0;
    if (B.mon) {   // turn monitor: centroid and sigma_z after the drift, binned count
        double *m = B.mon + 4 * sc.turn;
        m[0] = mz;
        m[1] = sz;
        m[3] = t[3];
    }
}

__device__ inline void bin_scalars(const BunchDev &B, const double (&t)[4]) {
    bin_scalars_on(B, t, *B.sc);
}

// The sum over the nred slices of a per-wave partial array by one 1024-thread
// workgroup, with the arithmetic of slice_sum (256 threads) and slices_done:
// the workgroup's four groups of four waves take four slices at a time, each
// group exactly as a 256-thread slice workgroup, and thread 0 adds the slice
// sums in slice order.  Same bits as the sliced form, no cross-workgroup
// count (for batches of many bunches, where the bunches fill the device).
template <int NV>
__device__ inline void whole_sum(const BunchDev &B, const double *a, double (&t)[NV]) {
    __shared__ double s_red[4][NV * 16];
    __shared__ double s_slice[4][NV];
    const int n = B.ntiles * kTileWaves, nred = B.nred;
    const int g = (int)threadIdx.x >> 8, tl = (int)threadIdx.x & 255;
    const int lane = tl & 63, wave = tl >> 6;
#pragma unroll
    for (int j = 0; j < NV; ++j) t[j] = 0.0;
    for (int r0 = 0; r0 < nred; r0 += 4) {
        const int r = r0 + g;
        double v[NV];
#pragma unroll
        for (int j = 0; j < NV; ++j) v[j] = 0.0;
        if (r < nred) {   // slice_sum's loads and order, thread tl of slice r
            const int e0 = r * kRedEntries + tl;
            double x[kRedPer][NV];
#pragma unroll
            for (int k = 0; k < kRedPer; ++k) {
                const int e = e0 + k * kRedThreads;
                if constexpr (NV == 4) {
                    double2 lo = make_double2(0.0, 0.0), hi = make_double2(0.0, 0.0);
                    if (e < n) {
                        const g_f64x2 *q = (const g_f64x2 *)a + 2 * (int64_t)e;
                        lo = ld2(q);
                        hi = ld2(q + 1);
                    }
                    x[k][0] = lo.x;
                    x[k][1] = lo.y;
                    x[k][2] = hi.x;
                    x[k][3] = hi.y;
                } else {
                    x[k][0] = e < n ? ((const g_f64 *)a)[e] : 0.0;
                }
            }
#pragma unroll
            for (int k = 0; k < kRedPer; ++k)
#pragma unroll
                for (int j = 0; j < NV; ++j) v[j] += x[k][j];
        }
        // block_sum's order within the group: wave butterflies, then its four waves
#pragma unroll
        for (int j = 0; j < NV; ++j)
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v[j] += __shfl_xor(v[j], off, 64);
        if (lane == 0)
#pragma unroll
            for (int j = 0; j < NV; ++j) s_red[g][j * 16 + wave] = v[j];
        __syncthreads();
        if (tl == 0)
#pragma unroll
            for (int j = 0; j < NV; ++j) {
                double x = 0.0;
                for (int w = 0; w < 4; ++w) x += s_red[g][j * 16 + w];
                s_slice[g][j] = x;
            }
        __syncthreads();
        if (threadIdx.x == 0)   // slices_done: one slice as it is, else from zero in order
            for (int q = 0; q < 4 && r0 + q < nred; ++q)
#pragma unroll
                for (int j = 0; j < NV; ++j) t[j] = nred == 1 ? s_slice[q][j] : t[j] + s_slice[q][j];
        __syncthreads();
    }
}

// Once per turn and bunch, after the pass: the pass's per-wave sums reduced
// into the turn scalars.  sum (dE'-dE_init) for the E0 update; for a wake
// bunch sigma_z (evolution.jl:357), the wake current (evolution.jl:373-386),
// the binned count, and the interpolation cell of the centroid (where the LDS
// windows of P are placed; the placement changes no result); the turn
// monitor's centroid, sigma_z and binned count.  One workgroup per slice.
__global__ __launch_bounds__(kRedThreads) void k_bin_reduce(const BunchDev *__restrict__ bds,
                                                            const int32_t *__restrict__ items) {
    __shared__ double s_red[4 * 16];
    const BunchDev &B = bds[items[2 * blockIdx.x]];
    const int r = items[2 * blockIdx.x + 1];
    if (!(B.flags & (F_WAKE | F_E0))) return;
    double v[4] = {0.0, 0.0, 0.0, 0.0};   // sum z, sum z^2, sum (dE'-dE_init), binned count
    slice_sum<4>(B.part, B.ntiles * kTileWaves, r, v, s_red);
    double t[4];
    if (!slices_done<4>(B, r, v, t) || threadIdx.x != 0) return;
    bin_scalars(B, t);
}

// k_bin_reduce with one 1024-thread workgroup per bunch (whole_sum)
__global__ __launch_bounds__(1024) void k_bin_reduce_whole(const BunchDev *__restrict__ bds,
                                                           const int32_t *__restrict__ list) {
    const BunchDev &B = bds[list[blockIdx.x]];
    if (!(B.flags & (F_WAKE | F_E0))) return;
    double t[4];
    whole_sum<4>(B, B.part, t);
    if (threadIdx.x == 0) bin_scalars(B, t);
}

// g_i = lam_i * (cnt_i / n_h) (wakefield.jl:210,238).  A wake table by lag
// (F_LAG) is convolved causally, P_k = I sum_j W_j g_{k+j}: the index
// convolution of the reversed density, written back reversed (store_block).
__device__ __forceinline__ double line_density(const BunchDev &B, const uint32_t *h, int i,
                                               double inv_nh) {
    if (B.flags & F_LAG) i = B.nbins - 1 - i;
    return B.lam[i] * ((double)h[i] * inv_nh);
}

// sX <- bit-reversed g_a + i g_{a-1}, zero-padded to F
__device__ inline void load_packed(double2 *sX, const BunchDev &B, int a, double inv_nh) {
    const int nb = B.nbins, Bk = B.B, F = B.F, logF = B.logF;
    const uint32_t *h = hist_of(B, B.sc->turn);
    for (int m = threadIdx.x; m < F; m += blockDim.x) {
        double re = 0.0, im = 0.0;
        if (m < Bk) {
            const int i = a * Bk + m;
            if (i < nb) re = line_density(B, h, i, inv_nh);
            if (a >= 1) im = line_density(B, h, i - Bk, inv_nh);
        }
        sX[bitrev(m, logF)] = make_double2(re, im);
    }
}

// clears bins [aB, aB+B) of the histogram half the next turn fills (nothing
// reads it during this turn)
__device__ inline void clear_next_hist(const BunchDev &B, int a) {
    const int nb = B.nbins, Bk = B.B;
    uint32_t *h = hist_of(B, B.sc->turn + 1);
    for (int m = threadIdx.x; m < Bk; m += blockDim.x)
        if (a * Bk + m < nb) h[a * Bk + m] = 0u;
}

// writes P[i] and its two places in the neighbour-pair table,
// P2[i].x = P[i] and P2[i-1].y = P[i] (P2[nb-1].y stays 0)
__device__ __forceinline__ void put_potential(const BunchDev &B, int i, double v) {
    B.P[i] = v;
    if (B.P2) {
        B.P2[2 * i] = v;
        if (i >= 1) B.P2[2 * i - 1] = v;
    }
}

// P block o from the unnormalised inverse FFT in LDS
__device__ inline void store_block(const BunchDev &B, const double2 *sX, int o, double cur) {
    const int nb = B.nbins, Bk = B.B, logF = B.logF;
    const double scale = cur / (double)B.F;
    for (int m = threadIdx.x; m < Bk; m += blockDim.x) {
        const int i = o * Bk + m;
        if (i >= nb) continue;
        const double2 y0 = sX[bitrev(m, logF)];
        const double2 y1 = sX[bitrev(m + Bk, logF)];
        put_potential(B, (B.flags & F_LAG) ? nb - 1 - i : i, (y0.x + y1.y) * scale);
    }
}

__device__ inline void zero_block(const BunchDev &B, int o) {
    const int nb = B.nbins, Bk = B.B;
    for (int m = threadIdx.x; m < Bk; m += blockDim.x) {
        const int i = o * Bk + m;
        if (i < nb) put_potential(B, (B.flags & F_LAG) ? nb - 1 - i : i, 0.0);
    }
}

// one workgroup per output block, wake support within one block (nJ <= 1)
__global__ __launch_bounds__(kConvThreads) void k_conv_fused(const BunchDev *__restrict__ bds,
                                                             const int32_t *__restrict__ items) {
    extern __shared__ double2 sX[];
    const BunchDev &B = bds[items[2 * blockIdx.x]];
    const int o = items[2 * blockIdx.x + 1];
    const double nh = B.sc->n_h, cur = B.sc->current;
    clear_next_hist(B, o);
    if (nh == 0.0 || B.nJ == 0) {   // wakefield.jl:305-311
        zero_block(B, o);
        return;
    }
    const double2 *tw = reinterpret_cast<const double2 *>(B.tw);
    const double2 *What = reinterpret_cast<const double2 *>(B.What);
    load_packed(sX, B, o, 1.0 / nh);
    fft_dit_forward(sX, B.F, B.logF, tw);
    for (int k = threadIdx.x; k < B.F; k += blockDim.x) sX[k] = cmul(sX[k], What[k]);
    fft_dif_inverse(sX, B.F, B.logF, tw);
    store_block(B, sX, o, cur);
}

// phase A of the blocked convolution: X_a = FFT(g_a + i g_{a-1})
__global__ __launch_bounds__(kConvThreads) void k_conv_fwd(const BunchDev *__restrict__ bds,
                                                           const int32_t *__restrict__ items) {
    extern __shared__ double2 sX[];
    const BunchDev &B = bds[items[2 * blockIdx.x]];
    const int a = items[2 * blockIdx.x + 1];
    const double nh = B.sc->n_h;
    clear_next_hist(B, a);
    if (nh == 0.0) return;
    load_packed(sX, B, a, 1.0 / nh);
    fft_dit_forward(sX, B.F, B.logF, reinterpret_cast<const double2 *>(B.tw));
    double2 *X = reinterpret_cast<double2 *>(B.X) + (int64_t)a * B.F;
    for (int k = threadIdx.x; k < B.F; k += blockDim.x) X[k] = sX[k];
}

// phase B1: Y_o[k] = sum_{j <= min(o, nJ-1)} X_{o-j}[k] FFT(W_j)[k] for every
// output block o, one chunk of spec_chunk(nA, F) frequencies of one bunch per
// workgroup.  The chunk's columns of X and of the wake spectra are staged in
// LDS once (every X_a and W_j is read from memory once per turn, instead of
// once per output block as in k_conv_inv); a thread takes one (frequency,
// output block) pair.  The products are added in ascending j from zero, as in
// k_conv_inv, so Y_o and the potential are the same bits on either path.
__global__ __launch_bounds__(kConvThreads) void k_conv_spec(const BunchDev *__restrict__ bds,
                                                            const int32_t *__restrict__ items) {
    extern __shared__ double2 sS[];
    const BunchDev &B = bds[items[2 * blockIdx.x]];
    const int c = items[2 * blockIdx.x + 1];
    if (B.sc->n_h == 0.0) return;   // k_conv_out writes a zero potential
    const int F = B.F, nA = B.nA, nJ = B.nJ;
    const int kc = spec_chunk(nA, F), k0 = c * kc;
    double2 *sXc = sS;               // [nA][kc]
    double2 *sWc = sS + nA * kc;     // [nJ][kc]
    const double2 *X = reinterpret_cast<const double2 *>(B.X);
    const double2 *What = reinterpret_cast<const double2 *>(B.What);
    for (int i = threadIdx.x; i < nA * kc; i += blockDim.x) {
        const int a = i / kc, kk = i - a * kc;
        sXc[i] = X[(int64_t)a * F + k0 + kk];
    }
    for (int i = threadIdx.x; i < nJ * kc; i += blockDim.x) {
        const int j = i / kc, kk = i - j * kc;
        sWc[i] = What[(int64_t)j * F + k0 + kk];
    }
    __syncthreads();
    double2 *Y = reinterpret_cast<double2 *>(B.Y);
    for (int i = threadIdx.x; i < nA * kc; i += blockDim.x) {
        const int o = i / kc, kk = i - o * kc;
        double2 acc = make_double2(0.0, 0.0);
        const int jmax = min(o, nJ - 1);
#pragma unroll 4
        for (int j = 0; j <= jmax; ++j) {
            const double2 t = cmul(sXc[(o - j) * kc + kk], sWc[j * kc + kk]);
            acc = make_double2(acc.x + t.x, acc.y + t.y);
        }
        Y[(int64_t)o * F + k0 + kk] = acc;
    }
}

// phase B2: inverse FFT of Y_o, P block o
__global__ __launch_bounds__(kConvThreads) void k_conv_out(const BunchDev *__restrict__ bds,
                                                           const int32_t *__restrict__ items) {
    extern __shared__ double2 sX[];
    const BunchDev &B = bds[items[2 * blockIdx.x]];
    const int o = items[2 * blockIdx.x + 1];
    const double nh = B.sc->n_h, cur = B.sc->current;
    if (nh == 0.0) {
        zero_block(B, o);
        return;
    }
    const int F = B.F;
    const double2 *Y = reinterpret_cast<const double2 *>(B.Y) + (int64_t)o * F;
    for (int k = threadIdx.x; k < F; k += blockDim.x) sX[k] = Y[k];
    fft_dif_inverse(sX, F, B.logF, reinterpret_cast<const double2 *>(B.tw));
    store_block(B, sX, o, cur);
}

// phase B (where the spectra of a chunk exceed the LDS): Y_o = sum_j
// X_{o-j} FFT(W_j), inverse FFT, P block o.  The products are summed in j
// order; four j at a time have their loads in flight.
__global__ __launch_bounds__(kConvInvThreads) void k_conv_inv(const BunchDev *__restrict__ bds,
                                                              const int32_t *__restrict__ items) {
    extern __shared__ double2 sX[];
    const BunchDev &B = bds[items[2 * blockIdx.x]];
    const int o = items[2 * blockIdx.x + 1];
    const double nh = B.sc->n_h, cur = B.sc->current;
    if (nh == 0.0) {
        zero_block(B, o);
        return;
    }
    const int F = B.F;
    const int jmax = min(o, B.nJ - 1);
    const double2 *X = reinterpret_cast<const double2 *>(B.X);
    const double2 *What = reinterpret_cast<const double2 *>(B.What);
    for (int k = threadIdx.x; k < F; k += blockDim.x) {
        double2 acc = make_double2(0.0, 0.0);
        int j = 0;
        for (; j + 4 <= jmax + 1; j += 4) {
            double2 xa[4], wa[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                xa[u] = X[(int64_t)(o - j - u) * F + k];
                wa[u] = What[(int64_t)(j + u) * F + k];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const double2 t = cmul(xa[u], wa[u]);
                acc = make_double2(acc.x + t.x, acc.y + t.y);
            }
        }
        for (; j <= jmax; ++j) {
            const double2 t = cmul(X[(int64_t)(o - j) * F + k], What[(int64_t)j * F + k]);
            acc = make_double2(acc.x + t.x, acc.y + t.y);
        }
        sX[k] = acc;
    }
    fft_dif_inverse(sX, F, B.logF, reinterpret_cast<const double2 *>(B.tw));
    store_block(B, sX, o, cur);
}

// Broadband resonator wake [V/C] at distance s >= 0 behind the source
// (Chao, Physics of Collective Beam Instabilities, eq. 2.84):
// W(s) = 2 a Rs e^{-a s/c} [cos(wb s/c) - (a/wb) sin(wb s/c)], a = w_r/(2Q),
// wb = sqrt(w_r^2 - a^2); W(0) = a Rs, half the jump.  Overdamped (Q < 1/2)
// and critical forms by cosh/sinh and the limit.
__device__ inline double resonator_wake(double s, double Rs, double fr, double Q) {
    const double wr = kTwoPi * fr;
    const double a = wr / (2.0 * Q);
    if (s < 0.0) return 0.0;
    if (s == 0.0) return a * Rs;
    const double tau = s / kSpeedLight;
    const double env = 2.0 * a * Rs * exp(-a * tau);
    const double d = wr * wr - a * a;
    if (d > 0.0) {
        const double wb = sqrt(d);
        return env * (cos(wb * tau) - (a / wb) * sin(wb * tau));
    } else if (d < 0.0) {
        const double wb = sqrt(-d);
        return env * (cosh(wb * tau) - (a / wb) * sinh(wb * tau));
    }
    return env * (1.0 - a * tau);
}

// the resonators' wake table by lag (set-up of an evolve call)
__global__ void k_resonator_table(double *__restrict__ W, int nb, double step, ResonatorSet rs) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nb) return;
    const double s = (double)j * step;
    double w = 0.0;
    for (int r = 0; r < rs.n; ++r) w += resonator_wake(s, rs.Rs[r], rs.fr[r], rs.Q[r]);
    W[j] = w;
}

// FFT of one wake block per workgroup (set-up of an evolve call)
__global__ __launch_bounds__(kConvThreads) void k_fft_wake(const BunchDev *__restrict__ bd,
                                                           const double *__restrict__ W) {
    extern __shared__ double2 sX[];
    const BunchDev &B = *bd;
    const int j = blockIdx.x, Bk = B.B, F = B.F, logF = B.logF, nb = B.nbins;
    for (int m = threadIdx.x; m < F; m += blockDim.x) {
        const int i = j * Bk + m;
        const double w = (m < Bk && i < nb) ? W[i] : 0.0;
        sX[bitrev(m, logF)] = make_double2(w, 0.0);
    }
    fft_dit_forward(sX, F, logF, reinterpret_cast<const double2 *>(B.tw));
    double2 *out = reinterpret_cast<double2 *>(const_cast<double *>(B.What)) + (int64_t)j * F;
    for (int k = threadIdx.x; k < F; k += blockDim.x) out[k] = sX[k];
}

// ---------------------------------------------------------------------------
// E0 update support, moments, init, finalize
// ---------------------------------------------------------------------------

// interp_sum_run for tiles of at least G pairs per thread (G = 4 or 8): per
// tile, groups of G pairs per thread.  The sums over a run of tiles [lt0, lt1) of bunch
// B of the interpolated potential at the post-drift z, one sum per tile and
// wave.  The run first stages B.iwin entries of P around the
// centroid in LDS (the whole table unless it exceeds the LDS); particles
// outside the window read the pair table P2.  Each thread takes its pairs
// of a tile in groups of G, in ascending order, with the next group's
// streaming loads (the next tile's first group at a tile's end) issued before
// this group's sums; the group loop has the same trip count in every lane.
// A tile's sums therefore depend on the tile and the block size only (not on G).
template <int G, bool FULLW>
__device__ __forceinline__ void interp_sum_run_tiles(const BunchDev &B, const int32_t *__restrict__ run,
                                                     int nrun) {
    extern __shared__ double sP[];
    const int nb = B.nbins;
    const double *__restrict__ P = B.P;
    const int w = B.iwin;
    const int wlo = window_lo(B.sc->centre_bin, w, nb);
    const unsigned wpairs = (unsigned)(w - 1);
    __syncthreads();   // the previous run's window is no longer read
    for (int k = threadIdx.x; k < w; k += kInterpThreads) sP[k] = P[wlo + k];
    __syncthreads();
    const int64_t n = B.n, ppt = B.pairs_per_tile, full = n >> 1;
    const int64_t bs = kInterpThreads;
    const int K = (int)(ppt / bs);   // pairs of each thread in a full tile
    const double c0 = B.c_first, c1 = B.c_last, inv = B.inv_s;
    const g_f64 *zp = (const g_f64 *)B.z;
    const g_f64x2 *P2 = (const g_f64x2 *)B.P2;
    const g_f64x2 *zpair = (const g_f64x2 *)zp;
    const int tid = (int)threadIdx.x, ibs = (int)bs;
    // pair offsets within a tile are 32-bit; tile bases are uniform
    // The ring's loads are issued unconditionally (a lane without a next pair
    // reads pair 0 and ignores it), so that every wait for a slot is the same
    // count of younger loads.  Needs a two-particle pair: full >= 1.
    const bool ring = nb >= 2 && full >= 1;
    // the j-th tile of the run, and its two-particle pairs
    auto tile_of = [&](int j) { return run[j] - B.tile_begin; };
    auto lim_of = [&](int lt) {
        const int64_t p0 = (int64_t)lt * ppt;
        return (int)(min(min(p0 + ppt, B.npairs), full) - p0);
    };
    double2 zn[G];
    if (ring) {
        const int lt = tile_of(0), lim0 = lim_of(lt);
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const int r = tid + u * ibs;
            zn[u] = ldnt((u < K && r < lim0) ? zpair + (int64_t)lt * ppt + r : zpair);
        }
    }
    for (int j = 0; j < nrun; ++j) {
        const int lt = tile_of(j);
        const int64_t p0 = (int64_t)lt * ppt;
        const int lim = lim_of(lt);   // two-particle pairs of this tile
        double v[1] = {0.0};
        if (ring) {
            for (int k = 0; k < K; k += G) {
                // the next group: the rest of this tile, else the next tile's first
                const g_f64x2 *nbase = zpair + p0;
                int nr = 0, nlim = 0, nk = 0;
                if (k + G < K) {
                    nr = tid + (k + G) * ibs;
                    nlim = lim;
                    nk = K - (k + G);
                } else if (j + 1 < nrun) {
                    const int ltn = tile_of(j + 1);
                    nbase = zpair + (int64_t)ltn * ppt;
                    nr = tid;
                    nlim = lim_of(ltn);
                    nk = K;
                }
                // a ring of G slots: slot u's next load is issued as soon as
                // its current pair is taken
#pragma unroll
                for (int u = 0; u < G; ++u) {
                    const double2 zz = zn[u];
                    const int r = nr + u * ibs;
                    zn[u] = ldnt((u < nk && r < nlim) ? nbase + r : zpair);
                    if (k + u >= K || tid + (k + u) * ibs >= lim) continue;
                    int k0, k1;
                    double w0, w1;
                    interp_index(zz.x, c0, c1, inv, nb, k0, w0);
                    interp_index(zz.y, c0, c1, inv, nb, k1, w1);
                    const unsigned j0 = (unsigned)(k0 - wlo), j1 = (unsigned)(k1 - wlo);
                    const unsigned i0 = min(j0, wpairs - 1), i1 = min(j1, wpairs - 1);
                    double t0 = sP[i0] * (1.0 - w0) + sP[i0 + 1] * w0;
                    double t1 = sP[i1] * (1.0 - w1) + sP[i1 + 1] * w1;
                    // outside the window: the pair table, with the wait in the branch
                    if (!FULLW && j0 >= wpairs) t0 = gather_interp(P2 + k0, w0);
                    if (!FULLW && j1 >= wpairs) t1 = gather_interp(P2 + k1, w1);
                    v[0] += t0;
                    v[0] += t1;
                }
            }
        } else if (nb < 2) {
            for (int64_t p = p0 + threadIdx.x; p < p0 + lim; p += bs) v[0] += P[0] + P[0];
        }
        if ((n & 1) && threadIdx.x == 0 && full >= p0 && full < min(p0 + ppt, B.npairs))
            v[0] += interp_pot(P, zp[n - 1], c0, c1, inv, nb);
        wave_sum<1>(v);   // (the DPP reduction slows this loop's streaming)
        if ((threadIdx.x & 63) == 0) B.ipart[(int64_t)lt * kTileWaves + (threadIdx.x >> 6)] = v[0];
    }
}

// interp_sum_run for tiles of one to three pairs per thread.  The sums over a
// run of tiles [lt0, lt1) of bunch B of the interpolated potential at the
// post-drift z (the E0 update of a wake bunch), one sum per tile and wave.
// The run first stages B.iwin entries of P around the
// centroid in LDS (the whole table unless it exceeds the LDS); particles
// outside the window read the pair table P2.  Each thread takes its pairs of
// the run in order -- pair j in tile lt0 + j/K at offset threadIdx.x +
// (j%K)*blockDim.x, K pairs per thread and tile -- through a ring of four
// streaming loads in flight, so a tile of one pair per thread costs no
// pipeline start.  A tile's sum is written when its last pair is added; it
// depends on the tile and the block size only.
template <bool FULLW>
__device__ __forceinline__ void interp_sum_run_flat(const BunchDev &B, const int32_t *__restrict__ run,
                                                    int nrun) {
    extern __shared__ double sP[];
    const int nb = B.nbins;
    const double *__restrict__ P = B.P;
    const int w = B.iwin;
    const int wlo = window_lo(B.sc->centre_bin, w, nb);
    const unsigned wpairs = (unsigned)(w - 1);
    __syncthreads();   // the previous run's window is no longer read
    for (int k = threadIdx.x; k < w; k += kInterpThreads) sP[k] = P[wlo + k];
    __syncthreads();
    const int64_t n = B.n, ppt = B.pairs_per_tile, full = n >> 1;
    const int tid = (int)threadIdx.x, ibs = kInterpThreads;
    const int K = (int)(ppt / ibs);   // pairs of each thread in a full tile
    const double c0 = B.c_first, c1 = B.c_last, inv = B.inv_s;
    const g_f64 *zp = (const g_f64 *)B.z;
    const g_f64x2 *P2 = (const g_f64x2 *)B.P2;
    const g_f64x2 *zpair = (const g_f64x2 *)zp;
    // two-particle pairs of tile lt (32-bit offsets within a tile)
    auto lim_of = [&](int lt) {
        const int64_t p0 = (int64_t)lt * ppt;
        return (int)(min(min(p0 + ppt, B.npairs), full) - p0);
    };
    // tile lt is complete: the single particle of an odd bunch (thread 0, the
    // bunch's last tile), then the per-wave sums
    auto flush = [&](double &v, int lt) {
        const int64_t p0 = (int64_t)lt * ppt;
        if ((n & 1) && threadIdx.x == 0 && full >= p0 && full < min(p0 + ppt, B.npairs))
            v += interp_pot(P, zp[n - 1], c0, c1, inv, nb);
        double s1[1] = {v};
        wave_total<1>(s1);
        if ((threadIdx.x & 63) == 63) B.ipart[(int64_t)lt * kTileWaves + (threadIdx.x >> 6)] = s1[0];
        v = 0.0;
    };
    auto tile_of = [&](int j) { return run[j] - B.tile_begin; };
    if (nb < 2 || full < 1) {   // a one-bin table, or no two-particle pair
        for (int j = 0; j < nrun; ++j) {
            const int lt = tile_of(j);
            double v = 0.0;
            const int lim = lim_of(lt);
            if (nb < 2)
                for (int r = tid; r < lim; r += ibs) v += P[0] + P[0];
            flush(v, lt);
        }
        return;
    }
    // The ring's loads are issued unconditionally (a lane without a next pair
    // reads pair 0 and ignores it), so that every wait for a slot is the same
    // count of younger loads.  The load and compute cursors are uniform.
    int lj = 0, lk = 0;   // the load cursor: entry lj of the run, pair lk of the thread
    auto load_next = [&]() -> double2 {
        const int r = tid + lk * ibs;
        const int llt = lj < nrun ? tile_of(lj) : 0;
        const bool ok = lj < nrun && r < lim_of(llt);
        const double2 x = ldnt(ok ? zpair + (int64_t)llt * ppt + r : zpair);
        if (++lk == K) {
            lk = 0;
            ++lj;
        }
        return x;
    };
    double2 zn[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) zn[u] = load_next();
    const int J = nrun * K;
    int cj = 0, clt = tile_of(0), ck = 0, clim = lim_of(clt);
    double v = 0.0;
    for (int j = 0; j < J; j += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (j + u >= J) break;
            const double2 zz = zn[u];
            zn[u] = load_next();
            if (tid + ck * ibs < clim) {
                int k0, k1;
                double w0, w1;
                interp_index(zz.x, c0, c1, inv, nb, k0, w0);
                interp_index(zz.y, c0, c1, inv, nb, k1, w1);
                const unsigned j0 = (unsigned)(k0 - wlo), j1 = (unsigned)(k1 - wlo);
                const unsigned i0 = min(j0, wpairs - 1), i1 = min(j1, wpairs - 1);
                double t0 = sP[i0] * (1.0 - w0) + sP[i0 + 1] * w0;
                double t1 = sP[i1] * (1.0 - w1) + sP[i1 + 1] * w1;
                // outside the window: the pair table, with the wait in the branch
                if (!FULLW && j0 >= wpairs) t0 = gather_interp(P2 + k0, w0);
                if (!FULLW && j1 >= wpairs) t1 = gather_interp(P2 + k1, w1);
                v += t0;
                v += t1;
            }
            if (++ck == K) {   // the last pair of tile clt
                flush(v, clt);
                ck = 0;
                if (++cj < nrun) {
                    clt = tile_of(cj);
                    clim = lim_of(clt);
                }
            }
        }
    }
}

// Both forms add each thread's pairs of a tile in ascending order and reduce
// them per wave in a fixed order (the large-tile form by a butterfly through
// LDS permutes, the small-tile form by DPP moves).  Which form a bunch takes
// depends on its size alone.  The per-tile groups cost less bookkeeping for
// large tiles, the run-wide ring no pipeline start per tile for small ones.
// FULLW: the window holds the whole table (no gathers), so the compiler's
// waits for the ring's slots count only the ring's own loads.
template <int G, bool FULLW>
__device__ __forceinline__ void interp_sum_run_w(const BunchDev &B, const int32_t *run, int nrun) {
    if (G > 4 && B.pairs_per_tile >= G * (int64_t)kInterpThreads)
        interp_sum_run_tiles<G, FULLW>(B, run, nrun);
    else if (B.pairs_per_tile >= 4 * (int64_t)kInterpThreads)
        interp_sum_run_tiles<4, FULLW>(B, run, nrun);
    else
        interp_sum_run_flat<FULLW>(B, run, nrun);
}

template <int G>
__device__ __forceinline__ void interp_sum_run(const BunchDev &B, const int32_t *run, int nrun) {
    if (B.iwin >= B.nbins)
        interp_sum_run_w<G, true>(B, run, nrun);
    else
        interp_sum_run_w<G, false>(B, run, nrun);
}

// The interpolation sums of the wake + update_E0 bunches over a tile table:
// each workgroup walks its contiguous range of entries of the sums' tile
// order, one run per bunch (a bunch shared by several workgroups has its
// tiles dealt to them in turn, as in the pass).
template <int G>
__device__ __forceinline__ void interp_sum_blocks(const BunchDev *__restrict__ bds,
                                         const int32_t *__restrict__ tile_bunch,
                                         const int32_t *__restrict__ order,
                                         const int32_t *__restrict__ ranges) {
    const int e0 = ranges[blockIdx.x];
    const int e1 = ranges[blockIdx.x + 1];
    for (int e = e0; e < e1;) {
        const int bi = tile_bunch[order[e]];
        const BunchDev &B = bds[bi];
        int ee = e + 1;
        while (ee < e1 && tile_bunch[order[ee]] == bi) ++ee;
        if ((B.flags & F_WAKE) && (B.flags & F_E0) && !(B.flags & F_CIC))
            interp_sum_run<G>(B, order + e, ee - e);
        e = ee;
    }
}

// At most 64 VGPRs, so that two workgroups fit a CU when both windows fit its
// LDS; rings of four loads.
__global__ __launch_bounds__(kInterpThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_interp_sum(
    const BunchDev *__restrict__ bds, const int32_t *__restrict__ tile_bunch,
    const int32_t *__restrict__ order, const int32_t *__restrict__ ranges) {
    interp_sum_blocks<4>(bds, tile_bunch, order, ranges);
}

// One workgroup per CU (a window of P larger than half the LDS): the
// registers of four waves per SIMD allow rings of eight loads.
__global__ __launch_bounds__(kInterpThreads) void k_interp_sum_deep(
    const BunchDev *__restrict__ bds, const int32_t *__restrict__ tile_bunch,
    const int32_t *__restrict__ order, const int32_t *__restrict__ ranges) {
    interp_sum_blocks<8>(bds, tile_bunch, order, ranges);
}

// Per-tile sums of z, z^2, dE, dE^2 (the initial and final spreads,
// evolution.jl:245-253,461-467); mode 1 first applies the last turn's pending
// wake kick and re-centre (the epilogue) and writes dE back.  The tile is read
// in the turn pass's layout, 16-byte (z, z) and (dE, dE) pairs, thread t taking
// pairs lt*ppt + t + k*1024, two pairs' loads in flight; the single particle
// of an odd bunch goes to thread 0 of the bunch's last tile.  LP: the kick
// reads P from the LDS copy sP (the whole table) instead of global memory.
template <bool LP>
__device__ inline void moments_tile(const BunchDev &B, int lt, int mode, const double *sP,
                                    double *s_red) {
    const int64_t ppt = B.pairs_per_tile;
    const int64_t p0 = (int64_t)lt * ppt;
    const int64_t p1 = min(p0 + ppt, B.npairs);
    const int64_t full = B.n >> 1;                  // pairs with two particles
    const int64_t pf = min(p1, full);
    PassCtx c;
    if (mode == 1) load_ctx(B, c);
    const bool kick = (mode == 1) && c.pending_kick;
    const bool rec = (mode == 1) && c.pending_m;
    g_f64 *zp = (g_f64 *)B.z;
    g_f64 *ep = (g_f64 *)B.dE;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    auto one = [&](double z, double e) -> double {
        if (kick)                                                           // wakefield.jl:342-344
            e = e - (LP ? interp_pot(sP, z, c.c0, c.c1, c.inv_s, c.nb)
                        : interp_pot(c.P, z, c.c0, c.c1, c.inv_s, c.nb));
        if (rec) e = e - c.m_prev;                                          // evolution.jl:434
        v[0] += z;
        v[1] += z * z;
        v[2] += e;
        v[3] += e * e;
        return e;
    };
    constexpr int64_t kStep = kPassThreads;
    for (int64_t p = p0 + threadIdx.x; p < pf; p += 2 * kStep) {
        const int64_t q = p + kStep;
        const bool two = q < pf;
        const double2 za = ldnt((const g_f64x2 *)(zp + 2 * p));
        const double2 ea = ldnt((const g_f64x2 *)(ep + 2 * p));
        double2 zb = make_double2(0.0, 0.0), eb = make_double2(0.0, 0.0);
        if (two) {
            zb = ldnt((const g_f64x2 *)(zp + 2 * q));
            eb = ldnt((const g_f64x2 *)(ep + 2 * q));
        }
        const double a0 = one(za.x, ea.x), a1 = one(za.y, ea.y);
        if (kick || rec) stnt((g_f64x2 *)(ep + 2 * p), a0, a1);
        if (two) {
            const double b0 = one(zb.x, eb.x), b1 = one(zb.y, eb.y);
            if (kick || rec) stnt((g_f64x2 *)(ep + 2 * q), b0, b1);
        }
    }
    if ((B.n & 1) && full >= p0 && full < p1 && threadIdx.x == 0) {
        const int64_t i = B.n - 1;
        const double e = one(zp[i], ep[i]);
        if (kick || rec) ep[i] = e;
    }
    block_sum<4>(v, s_red);
    if (threadIdx.x == 0) {
        double *o = B.moments + 4 * (int64_t)lt;
        o[0] = v[0]; o[1] = v[1]; o[2] = v[2]; o[3] = v[3];
    }
}

// The tiles of a table, tile = blockIdx.x + k*gridDim.x.  With dynamic LDS
// (lds_entries > 0: the epilogue), a workgroup copies a bunch's potential
// into its LDS once for its tiles of that bunch when the bunch has a kick
// pending and its table fits; the per-tile arithmetic is the same either way.
__global__ __launch_bounds__(kPassThreads) void k_moments(const BunchDev *__restrict__ bds,
                                                          const int32_t *__restrict__ tile_bunch,
                                                          int ntiles, int mode, int lds_entries) {
    __shared__ double s_red[4 * 16];
    extern __shared__ __align__(16) double s_pot[];
    int staged = -1;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        __syncthreads();   // s_red and s_pot of the previous tile are no longer read
        const int bi = tile_bunch[tile];
        const BunchDev &B = bds[bi];
        const int lt = tile - B.tile_begin;
        const bool lp = mode == 1 && B.sc->pending_kick && B.nbins >= 2 &&
                        B.nbins <= lds_entries;
        if (lp && staged != bi) {
            for (int k = threadIdx.x; k < B.nbins; k += blockDim.x) s_pot[k] = B.P[k];
            staged = bi;
            __syncthreads();
        }
        if (lp)
            moments_tile<true>(B, lt, mode, s_pot, s_red);
        else
            moments_tile<false>(B, lt, mode, nullptr, s_red);
    }
}

__global__ __launch_bounds__(kFinalThreads) void k_moments_final(const BunchDev *__restrict__ bds,
                                                                 const int32_t *__restrict__ list,
                                                                 int mode) {
    __shared__ double s_red[4 * 16];
    const BunchDev &B = bds[list[blockIdx.x]];
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int t = threadIdx.x; t < B.ntiles; t += blockDim.x) {
        v[0] += B.moments[4 * t];
        v[1] += B.moments[4 * t + 1];
        v[2] += B.moments[4 * t + 2];
        v[3] += B.moments[4 * t + 3];
    }
    block_sum<4>(v, s_red);
    if (threadIdx.x == 0) {
        const double sE = gstd_from_sums(v[2], v[3], (double)B.n);
        const double sz = gstd_from_sums(v[0], v[1], (double)B.n);
        if (mode == 0) {
            B.out[0] = sE;
            B.out[1] = sz;
        } else {
            const TurnScalars &s = *B.sc;
            B.out[2] = sE;
            B.out[3] = sz;
            B.out[4] = s.E0;
            B.out[5] = (double)s.drift_skipped_count;
            B.out[6] = (double)s.turn;
            B.out[7] = (B.n > 0) ? v[0] / (double)B.n : 0.0;   // centroid z
            B.out[8] = (B.n > 0) ? v[2] / (double)B.n : 0.0;   // centroid dE
        }
    }
}

__global__ __launch_bounds__(kFinalThreads) void k_init(const BunchDev *__restrict__ bds,
                                                        const int32_t *__restrict__ list) {
    const BunchDev &B = bds[list[blockIdx.x]];
    if (B.flags & F_WAKE)
        for (int b = threadIdx.x; b < 2 * B.nbins; b += blockDim.x) B.hist[b] = 0u;
    if (threadIdx.x == 0) {
        TurnScalars s = {};
        turn_scalars(B, B.E0_init, s);
        s.m_prev = 0.0;
        s.pending_kick = 0;
        s.pending_m = 0;
        s.drift_skipped_count = 0;
        s.turn = 0;
        s.centre_bin = 0;
        *B.sc = s;
    }
}

// The E0 update (evolution.jl:419-429) from this turn's sums, and the scalars
// of the next turn.  One workgroup per slice of the interpolation sums.
// The E0 update (evolution.jl:419-429) from this turn's sums (psum: the sum of
// the interpolated potential over the bunch), and the scalars of the next turn.
// finish_on updates the turn scalars s in place.
__device__ inline void finish_on(const BunchDev &B, double psum, TurnScalars &s) {
    const uint32_t fl = B.flags;
    double E0 = s.E0;
    if (s.drift_skip) s.drift_skipped_count += 1;
    if (fl & F_E0) {                                  // evolution.jl:419-429
        const double mean = (s.sum_dE - psum) / (double)B.n;
        E0 = E0 + mean;
        s.m_prev = mean;
        s.pending_m = 1;
    } else {
        s.pending_m = 0;
    }
    s.pending_kick = (fl & F_WAKE) ? 1 : 0;
    if (B.mon) B.mon[4 * s.turn + 2] = E0;   // turn monitor: E0 after the turn
    s.turn += 1;
    turn_scalars(B, E0, s);
}

__device__ inline void finish_turn(const BunchDev &B, double psum) {
    TurnScalars s = *B.sc;
    finish_on(B, psum, s);
    *B.sc = s;
}

// The sum over the bunch of the interpolated potential from the cloud-in-cell
// deposit (F_CIC): sum_k (N_k - S_k) P[k] + S_k P[k+1] with
// S_k = (hi_k 2^20 + lo_k) / 2^51, over the cells with particles (a cell
// without particles adds nothing, even where P is not finite); NaN when some
// particle's weight was not finite (as its interpolated value is).  Threads
// [0, kRedThreads) take cells k = t + kRedThreads*j in ascending j and clear
// them for the next turn; the waves' sums are added in wave order.  Valid in
// thread 0; every thread of the workgroup calls it, and the sliced and the
// whole-bunch finalize get the same bits from it.
__device__ inline double cic_psum(const BunchDev &B, double *s_red) {
    const int nb = B.nbins;
    g_u64 *H = (g_u64 *)B.cic;
    g_u32 *N = (g_u32 *)(B.cic + 2 * nb);
    const g_f64 *P = (const g_f64 *)B.P;
    if ((int)threadIdx.x < kRedThreads) {
        // U cells per thread at a time, every load of the group issued before
        // its arithmetic (the loads of empty cells read zeros and are not used)
        constexpr int U = 8;
        double v = 0.0;
        for (int k0 = (int)threadIdx.x; k0 < nb - 1; k0 += U * kRedThreads) {
            uint32_t cnt[U];
            unsigned long long hi[U], lo[U];
            double p0[U], p1[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = k0 + u * kRedThreads;
                const bool in = k < nb - 1;
                cnt[u] = in ? N[k] : 0u;
                hi[u] = in ? H[k] : 0ull;
                lo[u] = in ? H[nb + k] : 0ull;
                p0[u] = in ? P[k] : 0.0;
                p1[u] = in ? P[k + 1] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (cnt[u]) {
                    const double sk = ((double)hi[u] * (double)(1 << kCicLoBits) + (double)lo[u]) *
                                      (1.0 / kCicScale);
                    v += ((double)cnt[u] - sk) * p0[u] + sk * p1[u];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = k0 + u * kRedThreads;
                if (cnt[u]) {
                    N[k] = 0u;
                    H[k] = 0ull;
                    H[nb + k] = 0ull;
                }
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
    }
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < kRedThreads / 64; ++w) t += s_red[w];
        const uint32_t bad = N[nb];
        N[nb] = 0u;
        if (bad) t = __builtin_nan("");
    }
    return t;
}

// finish_turn after the sum of the interpolation sums, one workgroup per slice
// (a cloud-in-cell bunch: its deposit's sum in slice 0, zero in the others)
__global__ __launch_bounds__(kRedThreads) void k_finalize(const BunchDev *__restrict__ bds,
                                                          const int32_t *__restrict__ items) {
    __shared__ double s_red[16];
    const BunchDev &B = bds[items[2 * blockIdx.x]];
    const int r = items[2 * blockIdx.x + 1];
    const uint32_t fl = B.flags;
    double v[1] = {0.0};   // sum of the interpolated P (the sum of dE'-dE_init is in sc)
    if ((fl & F_E0) && (fl & F_WAKE)) {
        if (!(fl & F_CIC))
            slice_sum<1>(B.ipart, B.ntiles * kTileWaves, r, v, s_red);
        else if (r == 0)
            v[0] = cic_psum(B, s_red);
    }
    double t[1];
    if (!slices_done<1>(B, r, v, t) || threadIdx.x != 0) return;
    finish_turn(B, t[0]);
}

// k_finalize with one 1024-thread workgroup per bunch (whole_sum; a
// cloud-in-cell bunch's sum as slices_done adds slice 0's and the zeros)
__global__ __launch_bounds__(1024) void k_finalize_whole(const BunchDev *__restrict__ bds,
                                                         const int32_t *__restrict__ list) {
    __shared__ double s_red[16];
    const BunchDev &B = bds[list[blockIdx.x]];
    const uint32_t fl = B.flags;
    double t[1] = {0.0};
    if ((fl & F_E0) && (fl & F_WAKE)) {
        if (!(fl & F_CIC)) {
            whole_sum<1>(B, B.ipart, t);
        } else {
            const double d = cic_psum(B, s_red);
            t[0] = B.nred == 1 ? d : 0.0 + d;
        }
    }
    if (threadIdx.x == 0) finish_turn(B, t[0]);
}

// cic_psum over the deposit and the potential staged in LDS (sN[nb+1],
// sH[2 nb], sP[nb]): the same cells per thread in the same order, the same
// expressions, so the same bits; the global counters are cleared as there.
__device__ inline double cic_psum_lds(const BunchDev &B, const uint32_t *sN,
                                      const unsigned long long *sH, const double *sP,
                                      double *s_red) {
    const int nb = B.nbins;
    g_u64 *H = (g_u64 *)B.cic;
    g_u32 *N = (g_u32 *)(B.cic + 2 * nb);
    if ((int)threadIdx.x < kRedThreads) {
        double v = 0.0;
        for (int k = (int)threadIdx.x; k < nb - 1; k += kRedThreads) {
            const uint32_t cnt = sN[k];
            if (cnt) {
                const double sk = ((double)sH[k] * (double)(1 << kCicLoBits) + (double)sH[nb + k]) *
                                  (1.0 / kCicScale);
                v += ((double)cnt - sk) * sP[k] + sk * sP[k + 1];
                N[k] = 0u;
                H[k] = 0ull;
                H[nb + k] = 0ull;
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
    }
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < kRedThreads / 64; ++w) t += s_red[w];
        const uint32_t bad = sN[nb];
        N[nb] = 0u;
        if (bad) t = __builtin_nan("");
    }
    return t;
}

// The bin phase of a small bunch in one 1024-thread workgroup, in place of
// k_bin_reduce, the convolution kernels and k_finalize.  Everything it reads
// is staged first, side by side (the turn scalars in thread 0's registers,
// this turn's counts, the line-density kernel, the wake spectra, the FFT
// twiddles, the deposit), so the phase waits on memory about twice instead
// of once per step.  Then: the pass sums reduced as whole_sum does (the sliced form's
// bits) into the turn scalars; the line density of every block packed and
// transformed side by side (nA blocks of F points); the spectral sums Y_o =
// sum_j X_{o-j} FFT(W_j) added in ascending j from zero, as k_conv_inv adds
// them, one (block, frequency) pair per thread; the inverse transforms side
// by side; P and P2 (and P in LDS); the E0 update from the deposit as
// k_finalize_whole.  Every value is the same expression of the same operands
// as on the launch-per-phase path, so the bunch gets the same bits either
// way.  LDS layout: bin_whole_layout.
__device__ inline void bin_phase_whole(const BunchDev &B) {
    extern __shared__ __align__(16) uint8_t s_bw[];
    __shared__ double s_nh, s_cur, s_red[16];
    const uint32_t fl = B.flags;
    const bool wake = (fl & F_WAKE) != 0;
    const bool cic = wake && (fl & F_E0) && (fl & F_CIC);
    const int nb = B.nbins, nA = B.nA, F = B.F, logF = B.logF, Bk = B.B, nJ = B.nJ;
    BinWholeLayout L = bin_whole_layout(wake ? nb : 0, nA, nJ, F, cic);
    double2 *sX = reinterpret_cast<double2 *>(s_bw);
    double2 *sW = sX + nA * F, *sY = sW + nJ * F;
    uint32_t *sHist = reinterpret_cast<uint32_t *>(s_bw + L.hist);
    double *sLam = reinterpret_cast<double *>(s_bw + L.lam);
    double *sP = reinterpret_cast<double *>(s_bw + L.pot);
    double2 *sTw = reinterpret_cast<double2 *>(s_bw + L.tw);
    unsigned long long *sH = reinterpret_cast<unsigned long long *>(s_bw + L.cic);
    uint32_t *sN = reinterpret_cast<uint32_t *>(sH + 2 * nb);
    TurnScalars sc;
    if (threadIdx.x == 0) sc = *B.sc;
    const int64_t turn = B.sc->turn;
    if (wake) {
        const uint32_t *h = hist_of(B, turn);
        for (int i = threadIdx.x; i < nb; i += blockDim.x) {
            sHist[i] = h[i];
            sLam[i] = B.lam[i];
        }
        if (nJ > 1) {
            const double2 *What = reinterpret_cast<const double2 *>(B.What);
            for (int u = threadIdx.x; u < nJ * F; u += blockDim.x) sW[u] = What[u];
        }
        const double2 *twg = reinterpret_cast<const double2 *>(B.tw);
        for (int u = threadIdx.x; u < F / 2; u += blockDim.x) sTw[u] = twg[u];
        if (cic) {
            const unsigned long long *C = B.cic;
            const uint32_t *N = reinterpret_cast<const uint32_t *>(B.cic + 2 * nb);
            for (int i = threadIdx.x; i < 2 * nb; i += blockDim.x) sH[i] = C[i];
            for (int i = threadIdx.x; i <= nb; i += blockDim.x) sN[i] = N[i];
        }
    }
    if (fl & (F_WAKE | F_E0)) {
        double t[4];
        whole_sum<4>(B, B.part, t);   // its barriers also publish the staged LDS
        if (threadIdx.x == 0) {
            bin_scalars_on(B, t, sc);
            s_nh = sc.n_h;
            s_cur = sc.current;
        }
    }
    __syncthreads();
    if (wake) {
        const double nh = s_nh, cur = s_cur;
        const double2 *tw = sTw;   // the twiddles from LDS: no global load per FFT stage
        uint32_t *hn = hist_of(B, turn + 1);   // the next turn's half, cleared
        for (int i = threadIdx.x; i < nb; i += blockDim.x) hn[i] = 0u;
        if (nh == 0.0 || nJ == 0) {   // wakefield.jl:305-311
            for (int i = threadIdx.x; i < nb; i += blockDim.x) {
                put_potential(B, i, 0.0);
                sP[i] = 0.0;
            }
        } else {
            const double inv_nh = 1.0 / nh;
            auto dens = [&](int i) {   // line_density from the staged counts
                if (fl & F_LAG) i = nb - 1 - i;
                return sLam[i] * ((double)sHist[i] * inv_nh);
            };
            for (int u = threadIdx.x; u < nA * F; u += blockDim.x) {   // load_packed, every block
                const int a = u >> logF, m = u & (F - 1);
                double re = 0.0, im = 0.0;
                if (m < Bk) {
                    const int i = a * Bk + m;
                    if (i < nb) re = dens(i);
                    if (a >= 1) im = dens(i - Bk);
                }
                sX[a * F + bitrev(m, logF)] = make_double2(re, im);
            }
            fft_dit_forward(sX, F, logF, tw, nA);
            double2 *sR = sX;   // the spectra to invert
            if (nJ <= 1) {      // k_conv_fused's product
                const double2 *What = reinterpret_cast<const double2 *>(B.What);
                for (int u = threadIdx.x; u < nA * F; u += blockDim.x)
                    sX[u] = cmul(sX[u], What[u & (F - 1)]);
            } else {
                for (int u = threadIdx.x; u < nA * F; u += blockDim.x) {
                    const int o = u >> logF, k = u & (F - 1);
                    const int jmax = min(o, nJ - 1);
                    double2 acc = make_double2(0.0, 0.0);
                    for (int j = 0; j <= jmax; ++j) {
                        const double2 t = cmul(sX[(o - j) * F + k], sW[j * F + k]);
                        acc = make_double2(acc.x + t.x, acc.y + t.y);
                    }
                    sY[u] = acc;
                }
                sR = sY;
            }
            fft_dif_inverse(sR, F, logF, tw, nA);
            const double scale = cur / (double)F;
            for (int u = threadIdx.x; u < nA * Bk; u += blockDim.x) {   // store_block, every block
                const int o = u / Bk, m = u - o * Bk;
                const int i = o * Bk + m;
                if (i >= nb) continue;
                const double2 y0 = sR[o * F + bitrev(m, logF)];
                const double2 y1 = sR[o * F + bitrev(m + Bk, logF)];
                const int k = (fl & F_LAG) ? nb - 1 - i : i;
                const double v = (y0.x + y1.y) * scale;
                put_potential(B, k, v);
                sP[k] = v;
            }
        }
    }
    __syncthreads();   // P in LDS before the deposit's sum reads it
    double t = 0.0;
    if (cic) {
        const double d = cic_psum_lds(B, sN, sH, sP, s_red);
        t = B.nred == 1 ? d : 0.0 + d;
    }
    if (threadIdx.x == 0) {
        finish_on(B, t, sc);
        *B.sc = sc;
    }
}

__global__ __launch_bounds__(1024) void k_bin_whole(const BunchDev *__restrict__ bds,
                                                    const int32_t *__restrict__ list) {
    bin_phase_whole(bds[list[blockIdx.x]]);
}

__global__ void k_turn_scalars(const BunchDev *__restrict__ bd, double E0, TurnScalars *out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        TurnScalars s = {};
        turn_scalars(*bd, E0, s);
        *out = s;
    }
}

// ---------------------------------------------------------------------------
// stage kernels (unit parity); the same device functions as the fused pass
// ---------------------------------------------------------------------------

__global__ void k_stage_rf(double V, double sphis, double rf, double phis,
                           const double *__restrict__ z, double *__restrict__ dE, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double phi = ((-z[i]) * rf) + phis;
        dE[i] = dE[i] + V * (dsin(phi) - sphis);
    }
}

__global__ void k_stage_qe(double exc, const double *__restrict__ r, double *__restrict__ dE,
                           int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dE[i] = dE[i] + exc * r[i];
}

__global__ void k_stage_sr(double damp, double *__restrict__ dE, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dE[i] = dE[i] * damp;
}

__global__ void k_stage_drift(int update_eta, const TurnScalars *__restrict__ scp, double mass,
                              double alpha_c, double phis, double *__restrict__ z,
                              const double *__restrict__ dE, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const TurnScalars &s = *scp;
    PassCtx c = {};
    c.rf = s.rf;
    c.inv_rf = s.inv_rf;
    c.phis = phis;
    c.inv_mass = 1.0 / mass;
    c.gamma0 = s.gamma0;
    c.alpha_c = alpha_c;
    c.kcoef = s.kcoef;
    c.coeff_c = s.coeff_c;
    c.drift_skip = s.drift_skip;
    c.flags = update_eta ? F_ETA : 0u;
    z[i] = drift(c, Flags<0>{c.flags}, z[i], dE[i]);
}

__global__ __launch_bounds__(kPassThreads) void k_stage_hist(const double *__restrict__ z,
                                                             int64_t n, double e0, double e1,
                                                             double inv, int nb,
                                                             uint32_t *__restrict__ ghist,
                                                             int use_lds) {
    extern __shared__ uint32_t s_hist[];
    uint32_t *h = use_lds ? s_hist : ghist;
    if (use_lds) {
        for (int b = threadIdx.x; b < nb; b += blockDim.x) s_hist[b] = 0u;
        __syncthreads();
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int b = hist_bin(z[i], e0, e1, inv, nb);
        if (b >= 0) atomicAdd(&h[b], 1u);
    }
    if (use_lds) {
        __syncthreads();
        for (int b = threadIdx.x; b < nb; b += blockDim.x)
            if (s_hist[b]) atomicAdd(&ghist[b], s_hist[b]);
    }
}

__global__ void k_stage_interp_kick(const double *__restrict__ z, double *__restrict__ dE,
                                    int64_t n, const double *__restrict__ P, double c0, double c1,
                                    double inv, int nb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dE[i] -= interp_pot(P, z[i], c0, c1, inv, nb);
}

__global__ void k_stage_normals(uint64_t seed, int64_t turn, int64_t i0, int64_t n,
                                double *__restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t i = i0 + k;
    double a, b;
    qe_pair_normals(seed, (uint64_t)i >> 1, (uint32_t)turn, a, b);
    out[k] = (i & 1) ? b : a;
}

__global__ void k_generate(uint64_t seed, double mu_z, double sz, double mu_E, double sE,
                           double *__restrict__ z, double *__restrict__ dE, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t x[4];
    philox4x32_10((uint32_t)i, (uint32_t)((uint64_t)i >> 32), 0u, 1u, (uint32_t)seed,
                  (uint32_t)(seed >> 32), x);
    double a, b;
    box_muller(x, a, b);
    z[i] = mu_z + sz * a;
    dE[i] = mu_E + sE * b;
}

// generate_particles (evolution.jl:42-133) for n >= 2, step 1: one workgroup
// forms the sample covariance of m = min(n, 10000) independent samples
// (stream 2, regenerated from their counters for the second pass): mean-
// centred sums with m-1 normalisation (:75-82), variances floored at 1e-18,
// the diagonal lifted when the smallest eigenvalue is below 1e-12 (:89-94),
// then the lower Cholesky factor of the MvNormal (:97-101) -> L[0..2].
__device__ __forceinline__ void gen_sample(uint64_t seed, int64_t j, uint32_t stream, double &a,
                                           double &b) {
    uint32_t x[4];
    philox4x32_10((uint32_t)j, (uint32_t)((uint64_t)j >> 32), 0u, stream, (uint32_t)seed,
                  (uint32_t)(seed >> 32), x);
    box_muller(x, a, b);
}

__global__ __launch_bounds__(1024) void k_gen_cov(uint64_t seed, double mu_z, double mu_E,
                                                  double sz, double sE, int64_t m,
                                                  double *__restrict__ L) {
    __shared__ double s_red[3 * 16];
    __shared__ double s_mean[2];
    double v[3] = {0.0, 0.0, 0.0};
    for (int64_t j = threadIdx.x; j < m; j += blockDim.x) {
        double a, b;
        gen_sample(seed, j, 2u, a, b);
        v[0] += mu_z + sz * a;
        v[1] += mu_E + sE * b;
    }
    double w[2] = {v[0], v[1]};
    block_sum<2>(w, s_red);
    if (threadIdx.x == 0) {
        s_mean[0] = w[0] / (double)m;
        s_mean[1] = w[1] / (double)m;
    }
    __syncthreads();
    const double mz = s_mean[0], me = s_mean[1];
    v[0] = v[1] = v[2] = 0.0;
    for (int64_t j = threadIdx.x; j < m; j += blockDim.x) {
        double a, b;
        gen_sample(seed, j, 2u, a, b);
        const double dz = (mu_z + sz * a) - mz, de = (mu_E + sE * b) - me;
        v[0] += dz * dz;
        v[1] += de * de;
        v[2] += dz * de;
    }
    block_sum<3>(v, s_red);
    if (threadIdx.x == 0) {
        const double md = (double)(m - 1);
        double a = v[0] / md, c = v[1] / md;
        const double b = v[2] / md;
        a = a < 1e-18 ? 1e-18 : a;
        c = c < 1e-18 ? 1e-18 : c;
        // smallest eigenvalue as det / largest: no cancellation when a << c
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
        L[0] = l11;
        L[1] = l21;
        L[2] = sqrt(c - l21 * l21);
    }
}

// step 2: particle i = mu + L xi, xi the stream-1 normals of particle i
__global__ void k_gen_mvn(uint64_t seed, double mu_z, double mu_E, const double *__restrict__ L,
                          double *__restrict__ z, double *__restrict__ dE, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double l11 = L[0], l21 = L[1], l22 = L[2];
    double x1, x2;
    gen_sample(seed, i, 1u, x1, x2);
    z[i] = mu_z + l11 * x1;
    dE[i] = mu_E + (l21 * x1 + l22 * x2);
}

// Two-pass statistics of (z, dE): means, compute_std (utils.jl:531-544) and
// Statistics.cor (the emittance figure of merit, figures_of_merit.jl:36-47).
// Block b sums [b*chunk, (b+1)*chunk); mode 0 sums the values, mode 1 the
// squared deviations and the cross product about stats[0..1].
__global__ __launch_bounds__(1024) void k_pair_sums(int mode, const double *__restrict__ z,
                                                    const double *__restrict__ dE, int64_t n,
                                                    int64_t chunk, const double *__restrict__ stats,
                                                    double *__restrict__ part) {
    __shared__ double s_red[3 * 16];
    const int64_t i0 = (int64_t)blockIdx.x * chunk;
    const int64_t i1 = min(n, i0 + chunk);
    double v[3] = {0.0, 0.0, 0.0};
    const double cz = mode ? stats[0] : 0.0, ce = mode ? stats[1] : 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        if (mode) {
            const double a = z[i] - cz, b = dE[i] - ce;
            v[0] += a * a;
            v[1] += b * b;
            v[2] += a * b;
        } else {
            v[0] += z[i];
            v[1] += dE[i];
        }
    }
    block_sum<3>(v, s_red);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = v[0];
        part[3 * blockIdx.x + 1] = v[1];
        part[3 * blockIdx.x + 2] = v[2];
    }
}

// mode 0: stats[0..1] = means; mode 1: stats[2..3] = sample standard
// deviations of z and dE, stats[4] = their correlation
__global__ __launch_bounds__(256) void k_pair_sums_final(int mode, const double *__restrict__ part,
                                                         int nblocks, int64_t n,
                                                         double *__restrict__ stats) {
    __shared__ double s_red[3 * 16];
    double v[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
        v[0] += part[3 * b];
        v[1] += part[3 * b + 1];
        v[2] += part[3 * b + 2];
    }
    block_sum<3>(v, s_red);
    if (threadIdx.x == 0) {
        if (mode == 0) {
            stats[0] = v[0] / (double)n;
            stats[1] = v[1] / (double)n;
        } else {
            const double vz = v[0] / (double)(n - 1), ve = v[1] / (double)(n - 1);
            stats[2] = sqrt(jl_max(0.0, vz));   // compute_std: sqrt(max(0, variance))
            stats[3] = sqrt(jl_max(0.0, ve));
            double r = v[2] / sqrt(v[0] * v[1]);   // Statistics.cor, clamped to [-1, 1]
            stats[4] = (r != r) ? r : (r > 1.0 ? 1.0 : (r < -1.0 ? -1.0 : r));
        }
    }
}

// Plain device copy (16 B per lane per step, streaming loads and stores,
// 64 workgroups of 1024 per CU): the HBM ceiling of a kernel that reads and
// writes each byte once, and the calibration of the HBM byte counters.
// grid-stride copy with four 16-byte loads of each thread in flight before
// its stores (the fastest form of a sweep over grid, block size and loads in
// flight on MI355X)
__global__ __launch_bounds__(1024) void k_stream_copy(const f64x2 *__restrict__ src,
                                                      f64x2 *__restrict__ dst, int64_t n2) {
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += 4 * st) {
        f64x2 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i + u * st < n2) v[u] = __builtin_nontemporal_load(src + i + u * st);
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i + u * st < n2) __builtin_nontemporal_store(v[u], dst + i + u * st);
    }
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------

static inline unsigned blocks_for(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

hipError_t configure_kernels(int lds) {
    const void *fns[] = {(const void *)k_turn_pass, (const void *)k_turn_pass_std,
                         (const void *)k_conv_fused,
                         (const void *)k_conv_fwd, (const void *)k_conv_inv,
                         (const void *)k_conv_spec, (const void *)k_conv_out,
                         (const void *)k_fft_wake, (const void *)k_stage_hist,
                         (const void *)k_interp_sum, (const void *)k_interp_sum_deep,
                         (const void *)k_bin_whole};
    hipError_t last = hipSuccess;
    for (const void *f : fns) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e != hipSuccess) last = e;
    }
    return last;
}

hipError_t launch_stream_copy(const double *src, double *dst, int64_t n, int blocks, hipStream_t s) {
    if (n <= 1) return hipSuccess;
    hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)blocks), dim3(1024), 0, s,
                       reinterpret_cast<const f64x2 *>(src), reinterpret_cast<f64x2 *>(dst), n / 2);
    return hipGetLastError();
}

hipError_t launch_moments(const BunchDev *bds, const int32_t *tb, int ntiles, int mode,
                          int grid, int lds_entries, hipStream_t s) {
    if (ntiles <= 0) return hipSuccess;
    grid = grid > 0 ? min(grid, ntiles) : ntiles;
    hipLaunchKernelGGL(k_moments, dim3(grid), dim3(kPassThreads), (size_t)lds_entries * 8, s, bds,
                       tb, ntiles, mode, lds_entries);
    return hipGetLastError();
}

hipError_t launch_moments_final(const BunchDev *bds, const int32_t *list, int nb, int mode,
                                hipStream_t s) {
    if (nb <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_moments_final, dim3(nb), dim3(kFinalThreads), 0, s, bds, list, mode);
    return hipGetLastError();
}

hipError_t launch_init(const BunchDev *bds, const int32_t *list, int nb, hipStream_t s) {
    if (nb <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_init, dim3(nb), dim3(kFinalThreads), 0, s, bds, list);
    return hipGetLastError();
}

hipError_t launch_turn_pass(const BunchDev *bds, const int32_t *tb, const int32_t *order,
                            const int32_t *ranges, int grid, size_t lds, int std_only, hipStream_t s) {
    if (grid <= 0) return hipSuccess;
    if (std_only)
        hipLaunchKernelGGL(k_turn_pass_std, dim3((unsigned)grid), dim3(kPassThreads), lds, s, bds,
                           tb, order, ranges);
    else
        hipLaunchKernelGGL(k_turn_pass, dim3((unsigned)grid), dim3(kPassThreads), lds, s, bds, tb,
                           order, ranges);
    return hipGetLastError();
}

hipError_t launch_conv(int kind, const BunchDev *bds, const int32_t *items, int nitems, int Fmax,
                       hipStream_t s) {
    if (nitems <= 0) return hipSuccess;
    const size_t lds = (size_t)Fmax * 16;
    if (kind == CONV_FUSED)
        hipLaunchKernelGGL(k_conv_fused, dim3(nitems), dim3(kConvThreads), lds, s, bds, items);
    else if (kind == CONV_FWD)
        hipLaunchKernelGGL(k_conv_fwd, dim3(nitems), dim3(kConvThreads), lds, s, bds, items);
    else if (kind == CONV_INV)
        hipLaunchKernelGGL(k_conv_inv, dim3(nitems), dim3(kConvInvThreads), lds, s, bds, items);
    else if (kind == CONV_SPEC)
        hipLaunchKernelGGL(k_conv_spec, dim3(nitems), dim3(kConvThreads), (size_t)Fmax, s, bds,
                           items);
    else
        hipLaunchKernelGGL(k_conv_out, dim3(nitems), dim3(kConvThreads), lds, s, bds, items);
    return hipGetLastError();
}

hipError_t launch_interp_sum(const BunchDev *bds, const int32_t *tb, const int32_t *order,
                             const int32_t *ranges, int grid, size_t lds_bytes, bool deep,
                             hipStream_t s) {
    if (grid <= 0) return hipSuccess;
    if (deep)
        hipLaunchKernelGGL(k_interp_sum_deep, dim3((unsigned)grid), dim3(kInterpThreads), lds_bytes,
                           s, bds, tb, order, ranges);
    else
        hipLaunchKernelGGL(k_interp_sum, dim3((unsigned)grid), dim3(kInterpThreads), lds_bytes, s,
                           bds, tb, order, ranges);
    return hipGetLastError();
}

hipError_t launch_finalize(const BunchDev *bds, const int32_t *items, int nitems, hipStream_t s) {
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_finalize, dim3(nitems), dim3(kRedThreads), 0, s, bds, items);
    return hipGetLastError();
}

hipError_t launch_bin_reduce(const BunchDev *bds, const int32_t *items, int nitems, hipStream_t s) {
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bin_reduce, dim3(nitems), dim3(kRedThreads), 0, s, bds, items);
    return hipGetLastError();
}

hipError_t launch_bin_reduce_whole(const BunchDev *bds, const int32_t *list, int nb, hipStream_t s) {
    if (nb <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bin_reduce_whole, dim3(nb), dim3(1024), 0, s, bds, list);
    return hipGetLastError();
}

hipError_t launch_bin_whole(const BunchDev *bds, const int32_t *list, int nb, size_t lds,
                            hipStream_t s) {
    if (nb <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bin_whole, dim3(nb), dim3(1024), lds, s, bds, list);
    return hipGetLastError();
}

hipError_t launch_finalize_whole(const BunchDev *bds, const int32_t *list, int nb, hipStream_t s) {
    if (nb <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_finalize_whole, dim3(nb), dim3(1024), 0, s, bds, list);
    return hipGetLastError();
}

__global__ void k_zero(uint32_t *__restrict__ p, int64_t words) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

__global__ void k_fill(double *__restrict__ p, int64_t n, double v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

hipError_t launch_fill(double *p, size_t n, double v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const int64_t b = ((int64_t)n + 255) / 256;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)(b < 1024 ? b : 1024)), dim3(256), 0, s, p,
                       (int64_t)n, v);
    return hipGetLastError();
}

hipError_t launch_zero(void *p, size_t bytes, hipStream_t s) {
    const int64_t words = (int64_t)(bytes / 4);
    if (words <= 0) return hipSuccess;
    const int64_t b = (words + 255) / 256;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)(b < 1024 ? b : 1024)), dim3(256), 0, s,
                       static_cast<uint32_t *>(p), words);
    return hipGetLastError();
}

hipError_t launch_resonator_table(double *W, int nb, double step, const ResonatorSet &rs,
                                  hipStream_t s) {
    hipLaunchKernelGGL(k_resonator_table, dim3(blocks_for(nb, 256)), dim3(256), 0, s, W, nb, step,
                       rs);
    return hipGetLastError();
}

hipError_t launch_fft_wake(const BunchDev *bd, const double *W, int nJ, int F, hipStream_t s) {
    if (nJ <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fft_wake, dim3(nJ), dim3(kConvThreads), (size_t)F * 16, s, bd, W);
    return hipGetLastError();
}

hipError_t launch_stage_rf(double V, double sphis, double rf, double phis, const double *z,
                           double *dE, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_stage_rf, dim3(blocks_for(n, 256)), dim3(256), 0, s, V, sphis, rf, phis,
                       z, dE, n);
    return hipGetLastError();
}

hipError_t launch_stage_qe(double exc, const double *r, double *dE, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_stage_qe, dim3(blocks_for(n, 256)), dim3(256), 0, s, exc, r, dE, n);
    return hipGetLastError();
}

hipError_t launch_stage_sr(double damp, double *dE, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_stage_sr, dim3(blocks_for(n, 256)), dim3(256), 0, s, damp, dE, n);
    return hipGetLastError();
}

hipError_t launch_stage_drift(int update_eta, const TurnScalars *sc, double mass, double alpha_c,
                              double phis, double *z, const double *dE, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_stage_drift, dim3(blocks_for(n, 256)), dim3(256), 0, s, update_eta, sc,
                       mass, alpha_c, phis, z, dE, n);
    return hipGetLastError();
}

hipError_t launch_stage_hist(const double *z, int64_t n, double e0, double e1, double inv, int nb,
                             uint32_t *hist, int use_lds, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = blocks_for(n, kPassThreads);
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(k_stage_hist, dim3((unsigned)blocks), dim3(kPassThreads),
                       use_lds ? (size_t)nb * 4 : 0, s, z, n, e0, e1, inv, nb, hist, use_lds);
    return hipGetLastError();
}

hipError_t launch_stage_interp_kick(const double *z, double *dE, int64_t n, const double *P,
                                    double c0, double c1, double inv, int nb, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_stage_interp_kick, dim3(blocks_for(n, 256)), dim3(256), 0, s, z, dE, n,
                       P, c0, c1, inv, nb);
    return hipGetLastError();
}

hipError_t launch_stage_normals(uint64_t seed, int64_t turn, int64_t i0, int64_t n, double *out,
                                hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_stage_normals, dim3(blocks_for(n, 256)), dim3(256), 0, s, seed, turn, i0,
                       n, out);
    return hipGetLastError();
}

hipError_t launch_generate(uint64_t seed, double mu_z, double sz, double mu_E, double sE,
                           double *z, double *dE, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_generate, dim3(blocks_for(n, 256)), dim3(256), 0, s, seed, mu_z, sz, mu_E,
                       sE, z, dE, n);
    return hipGetLastError();
}

hipError_t launch_generate_particles(uint64_t seed, double mu_z, double mu_E, double sz, double sE,
                                     double *z, double *dE, int64_t n, double *work,
                                     hipStream_t s) {
    if (n < 2) return hipSuccess;
    double *L = work;
    hipLaunchKernelGGL(k_gen_cov, dim3(1), dim3(1024), 0, s, seed, mu_z, mu_E, sz, sE,
                       n < 10000 ? n : (int64_t)10000, L);
    hipLaunchKernelGGL(k_gen_mvn, dim3(blocks_for(n, 256)), dim3(256), 0, s, seed, mu_z, mu_E,
                       (const double *)L, z, dE, n);
    return launch_pair_stats(z, dE, n, work + 4, work + 12, s);
}

hipError_t launch_pair_stats(const double *z, const double *dE, int64_t n, double *stats,
                             double *part, hipStream_t s) {
    if (n < 2) return hipSuccess;
    const int64_t nbl = (n + 1023) / 1024;
    const int nb = (int)(nbl < kStatBlocks ? nbl : kStatBlocks);
    const int64_t chunk = (n + nb - 1) / nb;
    for (int mode = 0; mode < 2; ++mode) {
        hipLaunchKernelGGL(k_pair_sums, dim3(nb), dim3(1024), 0, s, mode, z, dE, n, chunk,
                           (const double *)stats, part);
        hipLaunchKernelGGL(k_pair_sums_final, dim3(1), dim3(256), 0, s, mode, (const double *)part,
                           nb, n, stats);
    }
    return hipGetLastError();
}

hipError_t launch_turn_scalars(const BunchDev *bd, double E0, TurnScalars *out, hipStream_t s) {
    hipLaunchKernelGGL(k_turn_scalars, dim3(1), dim3(64), 0, s, bd, E0, out);
    return hipGetLastError();
}

}  // namespace hm
