// hm_stages.hip -- stage entry points of the C ABI (one reference stage per
// call, host arrays in and out), used for unit parity with the reference's
// per-stage functions.  They run the same device code as the fused pass.
#include <hip/hip_runtime.h>

#include <cmath>
#include <memory>

#include "hm_internal.hpp"

using namespace hm;

namespace {

using Lock = std::lock_guard<std::recursive_mutex>;

hm_status stage_in(hm_ctx *ctx, int slot, const double *h, int64_t n, double **d) {
    HM_TRY(ctx->sbuf[slot].ensure((size_t)std::max<int64_t>(n, 1)));
    if (n > 0 && h)
        HIP_TRY(hipMemcpyAsync(ctx->sbuf[slot].p, h, (size_t)n * 8, hipMemcpyHostToDevice, ctx->stream));
    *d = ctx->sbuf[slot].p;
    return HM_OK;
}

hm_status stage_out(hm_ctx *ctx, int slot, double *h, int64_t n) {
    if (n > 0)
        HIP_TRY(hipMemcpyAsync(h, ctx->sbuf[slot].p, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return HM_OK;
}

// quantum_excitation.jl:58 / synchrotron_radiation.jl:44
double dUdE(double E0, double radius) {
    const double x = E0 / 1e9;
    return ((4.0 * 8.85e-5) * ((x * x) * x)) / radius;
}

}  // namespace

extern "C" {

hm_status hm_rf_kick(hm_ctx *ctx, double V, double sphis, double rf, double phis, const double *z,
                     double *dE, int64_t n) {
    if (!ctx || n < 0 || (n && (!z || !dE))) return fail(HM_ERR_ARG, "bad arguments");
    Lock lk(ctx->mu);
    HM_TRY(set_device(ctx));
    double *dz, *de;
    HM_TRY(stage_in(ctx, 0, z, n, &dz));
    HM_TRY(stage_in(ctx, 1, dE, n, &de));
    HIP_TRY(launch_stage_rf(V, sphis, rf, phis, dz, de, n, ctx->stream));
    return stage_out(ctx, 1, dE, n);
}

hm_status hm_quantum_excitation(hm_ctx *ctx, double E0, double radius, double sigma_E0,
                                const double *r, double *dE, int64_t n) {
    if (!ctx || n < 0 || (n && (!r || !dE))) return fail(HM_ERR_ARG, "bad arguments");
    Lock lk(ctx->mu);
    HM_TRY(set_device(ctx));
    const double om = 1.0 - dUdE(E0, radius);
    const double exc = std::sqrt(1.0 - (om * om)) * sigma_E0;      // quantum_excitation.jl:59
    double *dr, *de;
    HM_TRY(stage_in(ctx, 0, r, n, &dr));
    HM_TRY(stage_in(ctx, 1, dE, n, &de));
    HIP_TRY(launch_stage_qe(exc, dr, de, n, ctx->stream));
    return stage_out(ctx, 1, dE, n);
}

hm_status hm_synchrotron_radiation(hm_ctx *ctx, double E0, double radius, double *dE, int64_t n) {
    if (!ctx || n < 0 || (n && !dE)) return fail(HM_ERR_ARG, "bad arguments");
    Lock lk(ctx->mu);
    HM_TRY(set_device(ctx));
    const double damp = 1.0 - dUdE(E0, radius);                     // synchrotron_radiation.jl:45
    double *de;
    HM_TRY(stage_in(ctx, 1, dE, n, &de));
    HIP_TRY(launch_stage_sr(damp, de, n, ctx->stream));
    return stage_out(ctx, 1, dE, n);
}

hm_status hm_drift(hm_ctx *ctx, int32_t update_eta, double E0, double mass, double alpha_c,
                   int64_t harmonic, double phi_s, double freq_rf, double *z, const double *dE,
                   int64_t n, int32_t *skipped) {
    if (!ctx || n < 0 || (n && (!z || !dE))) return fail(HM_ERR_ARG, "bad arguments");
    Lock lk(ctx->mu);
    HM_TRY(set_device(ctx));
    BunchDev bd;
    std::memset(&bd, 0, sizeof(bd));
    bd.mass = mass;
    bd.alpha_c = alpha_c;
    bd.freq_rf = freq_rf;
    bd.two_pi_h = kTwoPi * (double)harmonic;
    bd.radius = 1.0;
    bd.flags = update_eta ? F_ETA : 0u;
    HM_TRY(ctx->sbd.ensure(1));
    HM_TRY(ctx->ssc.ensure(1));
    HIP_TRY(hipMemcpyAsync(ctx->sbd.p, &bd, sizeof(bd), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(launch_turn_scalars(ctx->sbd.p, E0, ctx->ssc.p, ctx->stream));
    double *dz, *de;
    HM_TRY(stage_in(ctx, 0, z, n, &dz));
    HM_TRY(stage_in(ctx, 1, dE, n, &de));
    HIP_TRY(launch_stage_drift(update_eta, ctx->ssc.p, mass, alpha_c, phi_s, dz, de, n, ctx->stream));
    TurnScalars sc;
    HIP_TRY(hipMemcpyAsync(&sc, ctx->ssc.p, sizeof(sc), hipMemcpyDeviceToHost, ctx->stream));
    HM_TRY(stage_out(ctx, 0, z, n));
    if (skipped) *skipped = (n > 0 && sc.drift_skip) ? 1 : 0;
    return HM_OK;
}

hm_status hm_histogram(hm_ctx *ctx, const double *z, int64_t n, double e_first, double e_last,
                       double inv_step, int64_t nbins, int64_t *counts) {
    if (!ctx || n < 0 || nbins <= 0 || nbins > (1 << 22) || (n && !z) || !counts)
        return fail(HM_ERR_ARG, "bad arguments");
    if (n >= (int64_t(1) << 32)) return fail(HM_ERR_LIMIT, "at most 2^32-1 values");
    Lock lk(ctx->mu);
    HM_TRY(set_device(ctx));
    double *dz;
    HM_TRY(stage_in(ctx, 0, z, n, &dz));
    HM_TRY(ctx->shist.ensure((size_t)nbins));
    HIP_TRY(hipMemsetAsync(ctx->shist.p, 0, (size_t)nbins * 4, ctx->stream));
    const int use_lds = (size_t)nbins * 4 <= (size_t)ctx->lds_max - 1024 ? 1 : 0;
    HIP_TRY(launch_stage_hist(dz, n, e_first, e_last, inv_step, (int)nbins, ctx->shist.p, use_lds,
                              ctx->stream));
    std::vector<uint32_t> h((size_t)nbins);
    HIP_TRY(hipMemcpyAsync(h.data(), ctx->shist.p, (size_t)nbins * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int64_t i = 0; i < nbins; ++i) counts[i] = h[(size_t)i];
    return HM_OK;
}

hm_status hm_wake_potential(hm_ctx *ctx, const int64_t *counts, int64_t nbins,
                            const double *wake_tab, const double *lam_tab, double current,
                            int32_t fft_block, double *P) {
    if (!ctx || !counts || !wake_tab || !lam_tab || !P || nbins <= 0 || nbins > (1 << 22))
        return fail(HM_ERR_ARG, "bad arguments");
    if (fft_block && (fft_block < 1 || (fft_block & (fft_block - 1)) || fft_block > kMaxFFT / 2))
        return fail(HM_ERR_ARG, "fft_block must be a power of two <= %d", kMaxFFT / 2);
    Lock lk(ctx->mu);
    HM_TRY(set_device(ctx));
    hipStream_t s = ctx->stream;
    int64_t nh = 0;
    std::vector<uint32_t> h((size_t)nbins);
    for (int64_t i = 0; i < nbins; ++i) {
        if (counts[i] < 0 || counts[i] > 0xffffffffLL) return fail(HM_ERR_ARG, "count out of range");
        h[(size_t)i] = (uint32_t)counts[i];
        nh += counts[i];
    }
    int64_t nw = 0;
    for (int64_t k = nbins - 1; k >= 0; --k)
        if (wake_tab[k] != 0.0) { nw = k + 1; break; }
    const int maxB = max_fft_block(ctx);
    if (fft_block > maxB) return fail(HM_ERR_LIMIT, "fft_block %d exceeds the LDS limit (%d)", fft_block, maxB);
    const int Bk = fft_block ? fft_block : auto_fft_block(nw, nbins, maxB);
    BunchDev d;
    std::memset(&d, 0, sizeof(d));
    d.n = 2;  // sigma_z is not used: the current is given
    d.ntiles = 1;
    d.flags = F_WAKE | F_CHARGE;
    d.charge = current;
    d.nbins = (int32_t)nbins;
    d.B = Bk;
    d.F = 2 * Bk;
    d.logF = ilog2(d.F);
    d.nA = (int32_t)((nbins + Bk - 1) / Bk);
    d.nJ = (int32_t)((nw + Bk - 1) / Bk);
    const size_t what = (size_t)std::max(d.nJ, 1) * (size_t)d.F * 2;
    const size_t xs = (size_t)d.nA * (size_t)d.F * 2;
    HM_TRY(ctx->shist.ensure((size_t)nbins * 2));
    HM_TRY(ctx->sbuf[0].ensure((size_t)nbins));   // W
    HM_TRY(ctx->sbuf[1].ensure((size_t)nbins));   // lam
    HM_TRY(ctx->sbuf[2].ensure((size_t)nbins));   // P
    HM_TRY(ctx->sbuf[3].ensure(what + xs));   // What, X
    HM_TRY(ctx->ssc.ensure(1));
    HM_TRY(ctx->sbd.ensure(1));
    double *What = ctx->sbuf[3].p;
    d.X = What + what;
    const double *tw;
    HM_TRY(twiddles(ctx, d.F, &tw));
    d.hist = ctx->shist.p;   // counts in the first half (turn 0); the kernels clear the second
    d.lam = ctx->sbuf[1].p;
    d.P = ctx->sbuf[2].p;
    d.What = What;
    d.tw = tw;
    d.sc = ctx->ssc.p;
    TurnScalars sc;   // the bin phase reads the binned count and the current
    std::memset(&sc, 0, sizeof(sc));
    sc.E0 = 1.0;
    sc.eta0 = 1.0;
    sc.n_h = (double)nh;
    sc.current = current;
    std::vector<int32_t> items;
    for (int o = 0; o < d.nA; ++o) {
        items.push_back(0);
        items.push_back(o);
    }
    DevBuf<int32_t> di;
    HM_TRY(di.ensure(items.size()));
    HIP_TRY(hipMemcpyAsync(ctx->shist.p, h.data(), (size_t)nbins * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(ctx->sbuf[0].p, wake_tab, (size_t)nbins * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(ctx->sbuf[1].p, lam_tab, (size_t)nbins * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(ctx->ssc.p, &sc, sizeof(sc), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(ctx->sbd.p, &d, sizeof(d), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(di.p, items.data(), items.size() * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(launch_fft_wake(ctx->sbd.p, ctx->sbuf[0].p, d.nJ, d.F, s));
    if (d.nJ <= 1) {
        HIP_TRY(launch_conv(CONV_FUSED, ctx->sbd.p, di.p, d.nA, d.F, s));
    } else {
        HIP_TRY(launch_conv(CONV_FWD, ctx->sbd.p, di.p, d.nA, d.F, s));
        HIP_TRY(launch_conv(CONV_INV, ctx->sbd.p, di.p, d.nA, d.F, s));
    }
    HIP_TRY(hipMemcpyAsync(P, d.P, (size_t)nbins * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    di.release();
    return HM_OK;
}

hm_status hm_interpolate_kick(hm_ctx *ctx, const double *z, double *dE, int64_t n,
                              const double *centers, int64_t nbins, const double *P) {
    if (!ctx || n < 0 || nbins <= 0 || (n && (!z || !dE)) || !centers || !P)
        return fail(HM_ERR_ARG, "bad arguments");
    Lock lk(ctx->mu);
    HM_TRY(set_device(ctx));
    double *dz, *de, *dp;
    HM_TRY(stage_in(ctx, 0, z, n, &dz));
    HM_TRY(stage_in(ctx, 1, dE, n, &de));
    HM_TRY(stage_in(ctx, 2, P, nbins, &dp));
    const double c0 = centers[0], c1 = centers[nbins - 1];
    const double inv = nbins > 1 ? 1.0 / ((c1 - c0) / (double)(nbins - 1)) : 0.0;  // wakefield.jl:39-42
    HIP_TRY(launch_stage_interp_kick(dz, de, n, dp, c0, c1, inv, (int)nbins, ctx->stream));
    return stage_out(ctx, 1, dE, n);
}

hm_status hm_pair_stats(hm_ctx *ctx, const double *z, const double *dE, int64_t n, double *out) {
    if (!ctx || n < 0 || (n && (!z || !dE)) || !out) return fail(HM_ERR_ARG, "bad arguments");
    for (int k = 0; k < 5; ++k) out[k] = 0.0;
    if (n == 1) {
        out[0] = z[0];
        out[1] = dE[0];
    }
    if (n < 2) return HM_OK;
    Lock lk(ctx->mu);
    HM_TRY(set_device(ctx));
    double *dz, *de;
    HM_TRY(stage_in(ctx, 0, z, n, &dz));
    HM_TRY(stage_in(ctx, 1, dE, n, &de));
    HM_TRY(ctx->sbuf[2].ensure((size_t)(8 + 3 * kStatBlocks)));
    double *work = ctx->sbuf[2].p;
    HIP_TRY(launch_pair_stats(dz, de, n, work, work + 8, ctx->stream));
    HIP_TRY(hipMemcpyAsync(out, work, 5 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return HM_OK;
}

hm_status hm_global_std(hm_ctx *ctx, const double *x, int64_t n, double *out) {
    if (!ctx || n < 0 || (n && !x) || !out) return fail(HM_ERR_ARG, "bad arguments");
    if (n == 0) {
        *out = 0.0;
        return HM_OK;
    }
    hm_bunch *b = nullptr;
    HM_TRY(hm_bunch_create(ctx, n, &b));
    std::unique_ptr<hm_bunch, hm_status (*)(hm_bunch *)> guard(b, hm_bunch_destroy);
    HM_TRY(hm_bunch_upload(b, x, x));
    hm_params p;
    std::memset(&p, 0, sizeof(p));
    p.E0 = 1.0;
    p.mass = 1.0;
    p.radius = 1.0;
    hm_result r;
    HM_TRY(hm_bunch_evolve(b, 0, &p, nullptr, &r));
    *out = r.sigma_z0;
    return HM_OK;
}

hm_status hm_stream_copy(hm_ctx *ctx, int64_t bytes, int32_t iters, double *ms_avg, double *gbs) {
    if (!ctx || bytes < 16 || iters <= 0) return fail(HM_ERR_ARG, "bad arguments");
    Lock lk(ctx->mu);
    HM_TRY(set_device(ctx));
    const int64_t n = (bytes / 16) * 2;   // doubles
    DevBuf<double> a, b;
    HM_TRY(a.ensure((size_t)n));
    HM_TRY(b.ensure((size_t)n));
    hipStream_t s = ctx->stream;
    HIP_TRY(hipMemsetAsync(a.p, 0, (size_t)n * 8, s));
    const int blocks = ctx->cus * 64;   // the fastest shape of a sweep on MI355X
    HIP_TRY(launch_stream_copy(a.p, b.p, n, blocks, s));   // warm-up
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, s));
    for (int k = 0; k < iters; ++k) HIP_TRY(launch_stream_copy(a.p, b.p, n, blocks, s));
    HIP_TRY(hipEventRecord(e1, s));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const double avg = (double)ms / iters;
    if (ms_avg) *ms_avg = avg;
    if (gbs) *gbs = 2.0 * (double)n * 8.0 / (avg * 1e-3) / 1e9;
    a.release();
    b.release();
    return HM_OK;
}

hm_status hm_normals(hm_ctx *ctx, uint64_t seed, int64_t turn, int64_t i0, int64_t n, double *out) {
    if (!ctx || n < 0 || i0 < 0 || (n && !out)) return fail(HM_ERR_ARG, "bad arguments");
    Lock lk(ctx->mu);
    HM_TRY(set_device(ctx));
    double *d;
    HM_TRY(stage_in(ctx, 0, nullptr, n, &d));
    HIP_TRY(launch_stage_normals(seed, turn, i0, n, d, ctx->stream));
    return stage_out(ctx, 0, out, n);
}

}  // extern "C"
