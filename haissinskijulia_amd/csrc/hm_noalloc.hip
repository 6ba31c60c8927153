// hm_noalloc.hip -- the allocation-free variant of the resident-bunch evolve.
//
// hm_evolve_device_noalloc runs longitudinal_evolve! (src/evolution.jl:206-471)
// on particles the caller already holds in device memory, with every work
// buffer of the call (histogram halves, potential tables, wake spectra,
// per-tile sums, turn scalars, launch tables, FFT twiddles) carved from one
// file-scope device array, `scratch`, instead of hipMalloc: for callers that
// cannot allocate (code under stream capture, tight serving loops, harnesses
// that forbid it).  The kernels and the order of their work are those of
// hm_bunch_evolve, so the results are the same bits.
//
// Each GPU holds its own copy of `scratch` (kScratchBytes) until the process
// ends.  Its device address is taken once per device with that device
// current (hipGetSymbolAddress, at context creation) and only that address
// is used on the host: the array's name is a host-side stand-in.  Every call
// on a GPU shares that GPU's array and finds the previous call's contents;
// each call clears what must start at zero.  Calls that overlap on one GPU
// race on the array.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "hm_internal.hpp"

namespace {

// 8 MiB: the work set of a bunch of up to 32768 bins (about 32 B per bin for
// each of the three spectra tables, 40 B per bin for the histogram halves and
// the potential tables, ~0.7 MB of per-tile sums for the at most 1025 tiles of
// any bunch size)
constexpr size_t kScratchBytes = size_t(8) << 20;
constexpr int64_t kMaxBins = 32768;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

}  // namespace

namespace hm {

size_t noalloc_scratch_bytes() { return kScratchBytes; }

// called with the context's device current (hm_create)
hm_status noalloc_scratch(void **ptr) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(HM_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    *ptr = p;
    return HM_OK;
}

}  // namespace hm

using namespace hm;

extern "C" hm_status hm_evolve_device_noalloc(hm_ctx *ctx, double *z_device, double *dE_device,
                                              int64_t n, int64_t nbins, const hm_params *params,
                                              const hm_options *opts, hm_result *result) {
    if (!ctx || !params || !result) return fail(HM_ERR_ARG, "ctx/params/result is NULL");
    if (n < 0 || (n > 0 && (!z_device || !dE_device)))
        return fail(HM_ERR_ARG, "bad device particle arrays");
    hm_options o{};
    if (opts) o = *opts;
    if (o.normals || o.monitor || o.timing)
        return fail(HM_ERR_ARG, "hm_evolve_device_noalloc takes no injected normals, turn monitor "
                                "or timing (their buffers scale with the turns)");
    if (params->use_wakefield && nbins > kMaxBins)
        return fail(HM_ERR_LIMIT, "hm_evolve_device_noalloc: at most %lld bins (scratch of %zu bytes)",
                    (long long)kMaxBins, kScratchBytes);
    HM_TRY(validate(params, &o, n, nbins));
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HM_TRY(set_device(ctx));
    if (!ctx->scratch) return fail(HM_ERR_DEVICE, "no scratch address for device %d", ctx->device);
    std::memset(result, 0, sizeof(*result));
    if (n == 0) {   // no wake (validated)
        result->E0 = (params->update_E0 && params->n_turns > 0 && !o.serial) ? NAN : params->E0;
        result->turns_done = params->n_turns;
        result->corr_z_dE = NAN;
        return HM_OK;
    }
    Arena arena;
    arena.base = static_cast<char *>(ctx->scratch);
    arena.cap = kScratchBytes;
    ArenaScope scope(&arena);
    hm_bunch b;   // the caller's particles; work buffers from the arena
    b.ctx = ctx;
    b.n = n;
    b.z = z_device;
    b.dE = dE_device;
    hm_status st;
    {
        Batch B(ctx);
        B.set_graphs(false);   // instantiating a graph allocates
        RunBunch q;
        q.b = &b;
        q.p = *params;
        q.o = o;
        q.nbins = nbins;
        std::vector<int> idx;
        st = B.add({q}, &idx);
        if (st == HM_OK) st = B.advance(params->n_turns);
        if (st == HM_OK) st = B.result(idx[0], result, nullptr, nullptr);
    }
    b.w.release();   // arena memory: nothing to give back
    return st;
}
