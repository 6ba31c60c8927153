// hm_kernels.hpp -- launch wrappers of the gfx950 kernels (hm_kernels.hip).
// Every wrapper checks nothing itself: the engine validates shapes and
// buffer sizes before it calls one.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hm_types.hpp"

namespace hm {

constexpr int kPassThreads = 1024;     // turn-pass / moments / interp-sum block
constexpr int kConvThreads = 256;      // convolution block
constexpr int kConvInvThreads = 1024;  // phase-B convolution block (spectral sums in flight)
constexpr int kFinalThreads = 256;     // per-bunch set-up and moment blocks
// Per-bunch reductions of the per-wave tile sums: each workgroup sums one
// slice of kRedEntries entries (kRedPer per thread); the last slice of a bunch
// to finish adds the slice sums in order.
constexpr int kRedThreads = 256;
constexpr int kRedPer = 4;
constexpr int kRedEntries = kRedThreads * kRedPer;
constexpr int kInterpThreads = 1024;   // LDS-staged interpolation sums
// partial sums per tile are kept per wave of the 1024-thread pass and
// interpolation-sum workgroups: part[ntiles][kTileWaves][4], ipart[ntiles][kTileWaves]
constexpr int kTileWaves = 16;
static_assert(kPassThreads == 64 * kTileWaves && kInterpThreads == 64 * kTileWaves,
              "per-wave tile sums assume 1024-thread workgroups");
// slices of the reductions of a bunch with ntiles tiles
inline int red_slices(int ntiles) {
    const int e = (ntiles > 0 ? ntiles : 1) * kTileWaves;
    return (e + kRedEntries - 1) / kRedEntries;
}
constexpr int kMaxFFT = 8192;          // largest LDS FFT (128 KiB of complex FP64)
constexpr int kAutoFFTBlock = 1024;    // default convolution block (F = 2048)

// CONV_SPEC + CONV_OUT replace CONV_INV where the spectra of a frequency
// chunk fit the LDS (conv_spec_fits)
enum ConvKind { CONV_FUSED = 0, CONV_FWD = 1, CONV_INV = 2, CONV_SPEC = 3, CONV_OUT = 4 };
constexpr int kConvKinds = 5;
// Frequencies per CONV_SPEC workgroup: one (frequency, output block) pair per
// thread of a 256-thread workgroup, 1 to 16 frequencies (F and the result
// are powers of two, so the chunks tile F).
__host__ __device__ inline int spec_chunk(int nA, int F) {
    int c = 16;
    while (c > 1 && (c * nA > 256 || c > F)) c >>= 1;
    return c;
}
inline size_t conv_spec_lds(int nA, int nJ, int F) {
    return (size_t)(nA + nJ) * (size_t)spec_chunk(nA, F) * 16;
}
inline bool conv_spec_fits(int nA, int nJ, int F, int lds_max) {
    return conv_spec_lds(nA, nJ, F) + 1024 <= (size_t)lds_max;
}

// zero `bytes` (a multiple of 4) at p by a kernel (no memset node: callers may
// capture the stream)
hipError_t launch_zero(void *p, size_t bytes, hipStream_t s);
// p[0..n) = v by a kernel
hipError_t launch_fill(double *p, size_t n, double v, hipStream_t s);
// allow the dynamic-LDS kernels the device's LDS per block (best effort)
hipError_t configure_kernels(int max_dynamic_lds);

// per-tile sums of z, z^2, dE, dE^2; mode 1 first applies the pending wake
// kick and re-centre (the epilogue of the last turn)
hipError_t launch_moments(const BunchDev *bds, const int32_t *tile_bunch, int ntiles, int mode,
                          int grid, int lds_entries, hipStream_t s);
// per-bunch spreads from the tile sums; mode 0 -> out[0..1], mode 1 -> out[2..8]
hipError_t launch_moments_final(const BunchDev *bds, const int32_t *bunch_list, int nb, int mode,
                                hipStream_t s);
hipError_t launch_init(const BunchDev *bds, const int32_t *bunch_list, int nb, hipStream_t s);
// `grid` workgroups over the tile table, workgroup g taking tiles
// [ranges[g], ranges[g+1]); lds_bytes = the largest per-bunch histogram +
// potential window; std_only: every tile's bunch has (flags & kPhysMask) == kStdFlags
hipError_t launch_turn_pass(const BunchDev *bds, const int32_t *tile_bunch, const int32_t *order,
                            const int32_t *ranges, int grid, size_t lds, int std_only, hipStream_t s);
// items are (bunch, block) pairs, for CONV_SPEC (bunch, frequency chunk)
// pairs; Fmax is the largest FFT length among them, for CONV_SPEC the largest
// conv_spec_lds of their bunches in bytes
hipError_t launch_conv(int kind, const BunchDev *bds, const int32_t *items, int nitems, int Fmax,
                       hipStream_t s);
// interpolation sums of the wake + update_E0 bunches: `grid` workgroups over
// the tile table (ranges as for the pass); lds_bytes = 8 x the largest iwin;
// deep: one workgroup per CU, eight streaming loads in flight per thread
hipError_t launch_interp_sum(const BunchDev *bds, const int32_t *tile_bunch, const int32_t *order,
                             const int32_t *ranges, int grid, size_t lds_bytes, bool deep,
                             hipStream_t s);
// per-bunch reduction of the pass's tile sums into the turn scalars (sum of
// dE'-dE_init; sigma_z, current, binned count, centroid cell of a wake bunch),
// and of the interpolation sums into the E0 update.  items holds nitems
// (bunch, slice) pairs, red_slices(ntiles) slices per bunch.
hipError_t launch_bin_reduce(const BunchDev *bds, const int32_t *items, int nitems, hipStream_t s);
hipError_t launch_finalize(const BunchDev *bds, const int32_t *items, int nitems, hipStream_t s);
// the same two reductions with one 1024-thread workgroup per listed bunch and
// no cross-workgroup count: the same bits (for batches of many bunches)
hipError_t launch_bin_reduce_whole(const BunchDev *bds, const int32_t *bunch_list, int nb,
                                   hipStream_t s);
hipError_t launch_finalize_whole(const BunchDev *bds, const int32_t *bunch_list, int nb,
                                 hipStream_t s);
// The whole bin phase (the reduction, the convolution and the finalize) of
// each listed bunch in one 1024-thread workgroup: the same bits as the
// launches of those phases.  For bunches of at most kBinWholeMaxTiles tiles
// without interpolation sums; lds = the largest bin_whole_layout(...).bytes
// among them.
constexpr int kBinWholeMaxTiles = 256;
// LDS of k_bin_whole for one bunch (byte offsets, 16-aligned): the
// convolution's X[nA][F] (and for nJ > 1 FFT(W)[nJ][F] and Y[nA][F]), then
// for a wake bunch this turn's counts[nb], the line-density kernel[nb], the
// potential[nb] and the FFT twiddles[F/2], then for a deposit bunch its
// hi/lo[2 nb] and n[nb + 1].
struct BinWholeLayout {
    size_t hist, lam, pot, tw, cic, bytes;
};
__host__ __device__ inline size_t align16(size_t b) { return (b + 15) & ~(size_t)15; }
__host__ __device__ inline BinWholeLayout bin_whole_layout(int nb, int nA, int nJ, int F,
                                                           bool cic) {
    BinWholeLayout L;
    const size_t conv = nb > 0 ? (size_t)(nJ <= 1 ? nA : 2 * nA + nJ) * (size_t)F * 16 : 0;
    L.hist = conv;
    L.lam = L.hist + align16((size_t)nb * 4);
    L.pot = L.lam + align16((size_t)nb * 8);
    L.tw = L.pot + align16((size_t)nb * 8);   // double2: every region starts 16-aligned
    L.cic = L.tw + (nb > 0 ? (size_t)(F / 2) * 16 : 0);
    L.bytes = L.cic + (cic ? align16((size_t)nb * 16 + ((size_t)nb + 1) * 4) : 0);
    return L;
}
hipError_t launch_bin_whole(const BunchDev *bds, const int32_t *bunch_list, int nb, size_t lds,
                            hipStream_t s);
// Broadband resonators of a lag-indexed wake table, (R_s [Ohm], f_r [Hz], Q)
// each, passed by value.
constexpr int kMaxResonators = 8;
struct ResonatorSet {
    int32_t n;
    double Rs[kMaxResonators], fr[kMaxResonators], Q[kMaxResonators];
};
// W[j] = sum of the resonators' wakes at s = j * step, j < nb
hipError_t launch_resonator_table(double *W, int nb, double step, const ResonatorSet &rs,
                                  hipStream_t s);
// FFT of the wake blocks: What[j] = FFT_F(W[jB .. jB+B) zero-padded)
hipError_t launch_fft_wake(const BunchDev *bd_dev, const double *W, int nJ, int F,
                           hipStream_t s);

// stage kernels (unit parity entry points)
hipError_t launch_stage_rf(double V, double sphis, double rf, double phis, const double *z,
                           double *dE, int64_t n, hipStream_t s);
hipError_t launch_stage_qe(double exc, const double *r, double *dE, int64_t n, hipStream_t s);
hipError_t launch_stage_sr(double damp, double *dE, int64_t n, hipStream_t s);
hipError_t launch_stage_drift(int update_eta, const TurnScalars *sc_dev, double mass,
                              double alpha_c, double phis, double *z, const double *dE, int64_t n,
                              hipStream_t s);
hipError_t launch_stage_hist(const double *z, int64_t n, double e_first, double e_last,
                             double inv_step, int nb, uint32_t *hist, int use_lds, hipStream_t s);
hipError_t launch_stage_interp_kick(const double *z, double *dE, int64_t n, const double *P,
                                    double c0, double c1, double inv, int nb, hipStream_t s);
hipError_t launch_stage_normals(uint64_t seed, int64_t turn, int64_t i0, int64_t n, double *out,
                                hipStream_t s);
hipError_t launch_generate(uint64_t seed, double mu_z, double sz, double mu_E, double sE,
                           double *z, double *dE, int64_t n, hipStream_t s);
hipError_t launch_turn_scalars(const BunchDev *bd_dev, double E0, TurnScalars *out,
                               hipStream_t s);
// two-pass statistics of device arrays z, dE (n >= 2): stats[0..4] = (mean z,
// mean dE, std z, std dE, cor(z, dE)); part holds 3 * kStatBlocks doubles
constexpr int kStatBlocks = 1024;
hipError_t launch_pair_stats(const double *z, const double *dE, int64_t n, double *stats,
                             double *part, hipStream_t s);
// generate_particles for n >= 2 (sample-covariance MvNormal): writes z, dE,
// work[0..2] = Cholesky factor (l11, l21, l22) and the pair statistics of
// the bunch at work[4..8]; work holds kGenWorkDoubles doubles
constexpr int kGenWorkDoubles = 12 + 3 * kStatBlocks;
hipError_t launch_generate_particles(uint64_t seed, double mu_z, double mu_E, double sz, double sE,
                                     double *z, double *dE, int64_t n, double *work,
                                     hipStream_t s);
// device-to-device copy of n doubles (n even) by a plain grid-stride kernel
hipError_t launch_stream_copy(const double *src, double *dst, int64_t n, int blocks, hipStream_t s);

}  // namespace hm
