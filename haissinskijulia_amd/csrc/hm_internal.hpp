// hm_internal.hpp -- host-side internals shared by the engine, the stage
// entry points and the server (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/haissinski_mi355x.h"
#include "hm_kernels.hpp"
#include "hm_types.hpp"

namespace hm {

hm_status fail(hm_status st, const char *fmt, ...);

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return ::hm::fail(e_ == hipErrorOutOfMemory ? HM_ERR_OOM : HM_ERR_DEVICE,        \
                              "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                              __LINE__);                                                     \
    } while (0)

#define HM_TRY(expr)                \
    do {                            \
        hm_status s_ = (expr);      \
        if (s_ != HM_OK) return s_; \
    } while (0)

// A region of device memory handed out in order, 256-byte aligned, for the
// work buffers of one allocation-free call (hm_evolve_device_noalloc): while
// an ArenaScope is active on a thread, DevBuf::ensure carves from it instead
// of calling hipMalloc, and release() gives nothing back.
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
};
Arena *&current_arena();
struct ArenaScope {
    Arena *prev;
    explicit ArenaScope(Arena *a) : prev(current_arena()) { current_arena() = a; }
    ~ArenaScope() { current_arena() = prev; }
    ArenaScope(const ArenaScope &) = delete;
    ArenaScope &operator=(const ArenaScope &) = delete;
};
// n bytes of the current arena, or null (with the error set) when it is full
void *arena_take(Arena *a, size_t bytes);

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t cap = 0;  // elements
    bool from_arena = false;
    hm_status ensure(size_t n) {
        if (n <= cap && p) return HM_OK;
        release();
        const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
        if (Arena *a = current_arena()) {
            p = static_cast<T *>(arena_take(a, bytes));
            if (!p) return HM_ERR_LIMIT;
            from_arena = true;
            cap = std::max<size_t>(n, 1);
            return HM_OK;
        }
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            p = nullptr;
            return fail(HM_ERR_OOM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
        }
        cap = std::max<size_t>(n, 1);
        return HM_OK;
    }
    void release() {
        if (p && !from_arena) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        from_arena = false;
    }
    size_t bytes() const { return cap * sizeof(T); }
};

// per-bunch work buffers of an evolve call (kept between calls)
struct Work {
    DevBuf<uint32_t> hist;
    DevBuf<double> P, P2, lam, What, X, Wtmp, part, ipart, moments, normals, out, stats, mon;
    DevBuf<TurnScalars> sc;
    DevBuf<double> red, Y;
    DevBuf<uint32_t> cnt;
    DevBuf<unsigned long long> cic;
    void release() {
        hist.release(); P.release(); P2.release(); lam.release(); What.release(); X.release();
        Wtmp.release(); part.release(); ipart.release(); moments.release(); normals.release();
        out.release(); stats.release(); mon.release(); sc.release(); red.release(); cnt.release();
        Y.release(); cic.release();
    }
    size_t bytes() const {
        return hist.bytes() + P.bytes() + P2.bytes() + lam.bytes() + What.bytes() + X.bytes() +
               Wtmp.bytes() + part.bytes() + ipart.bytes() + moments.bytes() + normals.bytes() +
               out.bytes() + stats.bytes() + mon.bytes() + sc.bytes() + red.bytes() + cnt.bytes() +
               Y.bytes() + cic.bytes();
    }
};

// Uploads of small host tables through one pinned staging buffer per
// context: a round packs its pieces into the buffer, enqueues their copies
// without waiting and records an event; the next round waits for that event
// before it packs again.  In an allocation-free call the copies come from the
// caller's pageable memory and the round ends with a stream synchronisation.
struct Staging {
    char *host = nullptr;
    size_t cap = 0, used = 0;
    hipEvent_t done = nullptr;
    bool pending = false;
};

}  // namespace hm

struct hm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int cus = 0;
    int lds_max = 65536;
    char arch[64] = {0};
    std::map<int, hm::DevBuf<double>> twiddles;   // F -> exp(-2 pi i k/F), k < F/2
    hm_bunch *tmp = nullptr;                      // staging bunch of hm_evolve
    hm::DevBuf<double> sbuf[4];                   // stage-entry scratch
    hm::DevBuf<uint32_t> shist;
    hm::DevBuf<hm::TurnScalars> ssc;
    hm::DevBuf<hm::BunchDev> sbd;
    void *scratch = nullptr;      // this device's allocation-free scratch (hm_noalloc.hip)
    hm::Staging stage;            // pinned staging of table uploads
    std::recursive_mutex mu;
};

struct hm_bunch {
    hm_ctx *ctx = nullptr;
    int64_t n = 0;
    double *z = nullptr, *dE = nullptr;
    hm::Work w;
};

namespace hm {

hm_status set_device(hm_ctx *ctx);
// a staging round of at most `bytes` (each piece aligned to 64), its copies
// on ctx's stream, and its end
hm_status stage_begin(hm_ctx *ctx, size_t bytes);
hm_status stage_copy(hm_ctx *ctx, void *dst, const void *src, size_t bytes);
hm_status stage_end(hm_ctx *ctx);
inline size_t stage_size(size_t bytes) { return (bytes + 63) & ~(size_t)63; }
// the device address of the allocation-free scratch on ctx's device (the
// device must be current), and its size
hm_status noalloc_scratch(void **ptr);
size_t noalloc_scratch_bytes();
double rn_ratio(double x, int64_t num, int64_t den);
void bin_grid(double zw, int64_t nb, double *edges, double *centers, double *step);
double particle_factor(int64_t N);
int ilog2(int64_t x);
hm_status twiddles(hm_ctx *ctx, int F, const double **out);
hm_status validate(const hm_params *p, const hm_options *o, int64_t n, int64_t nbins);
int max_fft_block(const hm_ctx *ctx);
int auto_fft_block(int64_t nw, int64_t nbins, int max_block);

// One bunch inside a Batch.
struct RunBunch {
    hm_bunch *b = nullptr;
    hm_params p{};
    hm_options o{};
    std::vector<double> wake_copy, res_copy;   // wake table / resonators until set-up
    int64_t nbins = 0;
    int64_t n_turns = 0;      // turns to run
    int64_t turns_done = 0;
    BunchDev d{};
    bool active = false;      // still turning
    bool finished = false;    // epilogue done, results valid
    bool fail_setup = false;  // test hook: the set-up fails by request after the slot is taken
};

// A set of device-resident bunches that advance together: every turn is one
// turn-pass launch over the tiles of all active bunches, one convolution
// launch over their output blocks, [one interpolation-sum launch] and one
// finalize launch.  Bunches may join (add) and leave (remove) between turns.
class Batch {
  public:
    explicit Batch(hm_ctx *ctx) : ctx_(ctx) {}
    ~Batch();
    // set up new bunches (initial spreads, grids, tables, FFT plan); their
    // slot indices are appended to idx.  All or nothing: on an error every
    // slot this call took is released again and idx is left as it was.
    hm_status add(const std::vector<RunBunch> &specs, std::vector<int> *idx);
    // run up to `turns` turns of every active bunch; bunches that reach their
    // n_turns run the epilogue and become finished
    hm_status advance(int64_t turns);
    // stop a bunch now (epilogue included), e.g. a failed or cancelled job
    hm_status finish_now(int i);
    hm_status result(int i, hm_result *r, double *centroid_z, double *centroid_E);
    void remove(int i);   // forget a finished slot (the bunch itself is the caller's)
    int active_count() const;
    size_t size() const { return rb_.size(); }
    RunBunch &slot(int i) { return rb_[(size_t)i]; }
    void set_timing(bool t) { timing_ = t; }
    void set_graphs(bool g) { graphs_ = g; }   // false: every turn launched directly
    double ms_pass() const { return pass_count_ ? pass_ms_ / (double)pass_count_ : 0.0; }
    double ms_conv() const { return conv_count_ ? conv_ms_ / (double)conv_count_ : 0.0; }
    int64_t pass_count() const { return pass_count_; }

  private:
    hm_status add_slots(const std::vector<RunBunch> &specs, std::vector<int32_t> &fresh);
    hm_status build_tables();
    hm_status upload_descriptors();
    hm_status stage_descriptors();
    hm_status run_subset_moments(const std::vector<int32_t> &which, int mode);
    hm_status epilogue(const std::vector<int32_t> &which);
    hm_status one_turn();
    hm_status flush_timing();
    hm_status capture_graph();
    void drop_graph();

    static constexpr int kGraphTurns = 16;   // turns per captured graph
    // active particles from which turns are launched directly: a turn of
    // 4e6 particles takes ~60 us on the GPU, ten times its launches' cost
    static constexpr int64_t kGraphMaxParticles = 4'000'000;
    hipGraphExec_t graph_exec_ = nullptr;    // kGraphTurns turns of the current tables

    hm_ctx *ctx_;
    std::vector<RunBunch> rb_;
    std::vector<bool> used_;
    DevBuf<BunchDev> d_bd_;
    DevBuf<int32_t> d_tiles_, d_order_, d_tmp_tiles_, d_tmp_list_;
    // per-workgroup tile ranges of the pass and the interpolation sums,
    // [pass_grid_ + 1] then [interp_grid_ + 1] entries, balanced by pairs
    DevBuf<int32_t> d_ranges_;
    int pass_grid_ = 0, interp_grid_ = 0;
    bool interp_deep_ = false;   // one interpolation workgroup per CU: deeper rings
    DevBuf<int32_t> d_conv_[kConvKinds];   // convolution items of each ConvKind
    DevBuf<int32_t> d_red_;       // (bunch, slice) items of the reductions
    DevBuf<int32_t> d_list_;      // the active bunches of the launch-per-phase bin phase
    DevBuf<int32_t> d_bw_;        // the active bunches of k_bin_whole
    int ntiles_ = 0, nred_ = 0, nlist_ = 0, nbw_ = 0;
    size_t bw_lds_ = 0;           // dynamic LDS of k_bin_whole
    // k_bin_whole takes a bunch whose tables fit this much LDS (up to 1024
    // bins at the tuned blocks)
    static constexpr size_t kBinWholeMaxLds = 144 * 1024;
    // one workgroup per bunch for the reductions once the bunches alone fill
    // the device (no cross-workgroup counts, whose device-scope fences cost
    // more than the slices gain there)
    static constexpr int kWholeMinBunches = 16;
    int nconv_[kConvKinds] = {0, 0, 0, 0, 0};
    size_t spec_lds_ = 0;   // LDS of the spectral-sum workgroups
    size_t pass_lds_ = 0;
    int Fmax_ = 2;
    bool any_wake_ = false, any_wake_e0_ = false;   // some bunch with wake / wake + update_E0
    bool any_reduce_ = false;                        // some bunch with wake or update_E0
    bool std_only_ = true;                          // all active bunches have kStdFlags
    size_t interp_lds_ = 0;                         // dynamic LDS of the interpolation sums
    bool dirty_ = true;
    bool timing_ = false;
    bool graphs_ = true;
    std::vector<hipEvent_t> ev_pool_;
    size_t ev_used_ = 0;
    double pass_ms_ = 0.0, conv_ms_ = 0.0;
    int64_t pass_count_ = 0, conv_count_ = 0;
};

}  // namespace hm
