// hm_types.hpp -- layouts shared by the host engine and the gfx950 kernels.
//
// Device memory layout of one bunch (all FP64 unless noted):
//   z[n], dE[n]                 particle state, SoA (data_structures.jl:26-29)
//   hist[2][nbins] (uint32)     line-density counts; turn t fills half t&1 while
//                               the bin phase clears the other half
//   P[nbins]                    wake potential at the bin centres (previous turn)
//   P2[nbins][2]                the same as neighbour pairs, one 16-byte gather per particle
//   lam[nbins]                  delta(c_i, 0.15) line-density kernel (wakefield.jl:219-221)
//   What[nJ][F]   (complex)     FFT of the wake blocks, fixed for one evolve call
//   X[nA][F]      (complex)     FFT of the packed line-density blocks of a turn
//   Y[nA][F]      (complex)     their products with the wake spectra, summed
//                               per output block (before its inverse FFT)
//   part[ntiles][16][4]         per-wave sums of each tile (sum z, sum z^2,
//                               sum dE'-dE_init, in-range count)
//   ipart[ntiles][16]           per-wave sums of each tile of the interpolated
//                               potential (E0 update)
//   moments[ntiles][4]          per-tile sums of z, z^2, dE, dE^2 (initial / final spreads)
//   red[nred][4]                slice sums of the per-bunch reductions, and cnt
//                               (uint32), the slices of the running reduction
//                               counted in so far (zero between reductions)
//   sc                          TurnScalars of the current turn
//   out[8]                      sigma_E0, sigma_z0, sigma_E, sigma_z, E0, ...
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hm {

constexpr double kTwoPi = 6.283185307179586;   // Julia's 2π as Float64
constexpr double kPi = 3.141592653589793;
constexpr double kSpeedLight = 299792458.0;    // src/constants.jl:5

enum BunchFlags : uint32_t {
    F_WAKE = 1u,       // use_wakefield
    F_ETA = 2u,        // update_eta
    F_E0 = 4u,         // update_E0
    F_SR = 8u,         // SR_damping
    F_QE = 16u,        // use_excitation
    F_INJ = 32u,       // injected normals instead of Philox
    F_CHARGE = 64u,    // fixed wake current (charge extension / stage entry)
    F_GHIST = 128u,    // histogram straight to global memory (too many bins for LDS)
    F_LAG = 256u,      // wake table by lag behind the source: bin k feels bins k+j (causal)
    F_CIC = 512u,      // the pass deposits the interpolation weights of the new z (cic)
};

// Cloud-in-cell deposit (F_CIC).  The E0 update needs sum_i interp(P_t, z_i)
// over the post-drift z of turn t, with P_t known only after the histogram of
// those z.  The interpolation is linear in P: with (k_i, w_i) the cell and
// weight of z_i, the sum is sum_k (N_k - S_k) P_t[k] + S_k P_t[k+1], where
// N_k counts the particles of cell k and S_k adds their weights.  The pass
// counts N_k and adds W_i = round(w_i 2^51) into two 64-bit counters per cell
// (W >> 20 and the low 20 bits): integer sums, so they do not depend on the
// order of the additions, and 2^-52 of a weight is as fine as the weight's
// own FP64 rounding.  Layout, in LDS and in global memory: hi[nbins],
// lo[nbins] (64-bit), then n[nbins + 1] (32-bit; n[nbins] counts particles
// whose weight is not finite).
constexpr double kCicScale = 2251799813685248.0;   // 2^51
constexpr int kCicLoBits = 20;

// the all-physics-on configuration the turn pass is specialised for, and the
// flags that select the pass's code path
constexpr uint32_t kStdFlags = F_WAKE | F_ETA | F_E0 | F_SR | F_QE;
constexpr uint32_t kPhysMask = F_WAKE | F_ETA | F_E0 | F_SR | F_QE | F_INJ | F_GHIST;

// Scalars of one turn of one bunch (evolution.jl:300-304 plus the pending
// state of the fused pass).  Written by k_init / k_finalize, read by the
// per-particle kernels.
struct TurnScalars {
    double E0, gamma0, beta0, eta0, rf;
    double inv_rf;     // 1/rf (the phase-to-z conversion, utils.jl:476)
    double denom;      // (beta0*beta0)*E0, drift denominator (evolution.jl:338)
    double kcoef;      // (2 pi h)/denom: the phase coefficient is eta * kcoef
    double coeff_c;    // constant-eta phase coefficient (evolution.jl:157)
    double damp;       // 1 - dU/dE (synchrotron_radiation.jl:44-45)
    double exc;        // sqrt(1-(1-dU/dE)^2) * sigma_E0 (quantum_excitation.jl:58-59)
    double m_prev;     // mean energy change still to subtract (evolution.jl:429-434)
    double sigma_z, current, n_h;   // bin phase of this turn (wake bunches)
    double sum_dE;     // sum over the bunch of dE'-dE_init of this turn (evolution.jl:419-429)
    int32_t pending_kick;   // 1: the previous turn's wake kick is still to be applied
    int32_t pending_m;      // 1: m_prev still to be subtracted
    int32_t drift_skip;     // apply_phase_advance! skip condition of this turn
    int32_t drift_skipped_count;
    int64_t turn;           // turns completed in this call
    int32_t centre_bin;     // interpolation cell of this turn's centroid: the LDS windows
                            // of P are placed around it (placement changes no result)
    int32_t pad_;
};

struct BunchDev {
    double *z, *dE;
    int64_t n, npairs;
    int64_t pairs_per_tile;
    int32_t tile_begin, ntiles;
    uint32_t flags;
    int32_t nbins;
    // physics constants of the call
    double mass, inv_mass, voltage, alpha_c, phis, sin_phis, radius, freq_rf, two_pi_h;
    double sigma_E0;   // initial energy spread (drives QE and the wake current)
    double pf;         // particle factor of evolution.jl:377-383 (depends on n only)
    double charge;     // fixed current when F_CHARGE
    double E0_init;
    uint64_t seed;
    int64_t turn_offset;
    const double *normals;   // device, [turns][n] when F_INJ
    // histogram / interpolation grids (evolution.jl:270-274, wakefield.jl:38-42)
    double e_first, e_last, inv_step;
    double c_first, c_last, inv_s;
    double u_last;        // (c_last - c_first) * inv_s: the interpolation coordinate of the last centre
    uint32_t *hist;       // [2][nbins]: half (turn & 1) is this turn's
    double *P;
    double *P2;           // interleaved pairs (P[k], P[k+1]), k < nbins (P[nbins] := 0)
    const double *lam;
    const double *What;   // interleaved complex [nJ][F][2]
    const double *tw;     // interleaved complex [F/2][2], exp(-2 pi i k / F)
    double *X;            // interleaved complex [nA][F][2]: FFT(g_a + i g_{a-1}) (nJ > 1)
    double *Y;            // interleaved complex [nA][F][2]: sum_j X_{o-j} FFT(W_j), or null
    int32_t B, F, logF, nA, nJ;
    int32_t win;          // entries of P the pass stages in LDS (0: all gathers global)
    int32_t iwin;         // entries of P the interpolation sums stage in LDS
    double *part, *ipart, *moments, *out;
    double *red;          // [nred][4]
    uint32_t *cnt;
    int32_t nred;
    int32_t cic_off;      // F_CIC: byte offset of the pass's LDS deposit
    TurnScalars *sc;
    double *mon;          // device [n_turns][4] turn monitor or null
    unsigned long long *cic;   // F_CIC: this turn's deposit (hi, lo, n; zero between turns)
};

// LDS of a turn-pass workgroup for one bunch: its histogram (nbins counts,
// unless F_GHIST), then a window of `win` entries of the previous turn's
// potential, P[lo .. lo+win), for the wake-kick gathers, then with F_CIC the
// deposit's counters at cic_off.
__host__ __device__ inline int pass_hist_bytes(const BunchDev &B) {
    return ((B.flags & F_WAKE) && !(B.flags & F_GHIST)) ? ((B.nbins * 4 + 15) & ~15) : 0;
}

__host__ __device__ inline int cic_bytes(int nbins) { return (16 * nbins + 4 * (nbins + 1) + 15) & ~15; }
__host__ __device__ inline int pass_cic_bytes(const BunchDev &B) {
    return (B.flags & F_CIC) ? cic_bytes(B.nbins) : 0;
}
// the pass's dynamic LDS for one bunch
__host__ __device__ inline int pass_lds_bytes(const BunchDev &B) {
    return (B.flags & F_CIC) ? B.cic_off + pass_cic_bytes(B) : pass_hist_bytes(B) + 8 * B.win;
}

// first entry of a window of w entries of P centred on `centre`
__host__ __device__ inline int window_lo(int centre, int w, int nbins) {
    const int lo = centre - w / 2;
    return lo < 0 ? 0 : (lo > nbins - w ? nbins - w : lo);
}

}  // namespace hm
