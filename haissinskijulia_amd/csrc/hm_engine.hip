// hm_engine.hip -- host side of the MI355X tracker: contexts, device-resident
// bunches, the batched turn loop and the evolve entry points of the C ABI.
//
// One evolve call (src/evolution.jl:206-471) runs as:
//   set-up   k_moments + k_moments_final   sigma_E0, sigma_z0 (evolution.jl:245-253)
//            host: bin grid, wake and line-density tables (evolution.jl:261-281,
//            wakefield.jl:194-221), FFT block plan; k_fft_wake; k_init
//   turns    k_turn_pass, k_bin_reduce, k_conv, [k_interp_sum], k_finalize   per turn
//   epilogue k_moments(mode 1) + k_moments_final            (evolution.jl:461-470)
// A batch of bunches shares every launch: tiles, convolution items and
// finalize blocks are looked up through per-launch tables.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <memory>

#include "hm_internal.hpp"

namespace hm {

static thread_local std::string g_last_error;

hm_status fail(hm_status st, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return st;
}

Arena *&current_arena() {
    static thread_local Arena *a = nullptr;
    return a;
}

void *arena_take(Arena *a, size_t bytes) {
    const size_t off = (a->used + 255) & ~(size_t)255;
    if (off > a->cap || bytes > a->cap - off) {
        fail(HM_ERR_LIMIT, "the allocation-free scratch (%zu bytes) is too small: %zu more bytes "
                           "needed at offset %zu", a->cap, bytes, off);
        return nullptr;
    }
    a->used = off + bytes;
    return a->base + off;
}

hm_status stage_begin(hm_ctx *ctx, size_t bytes) {
    Staging &g = ctx->stage;
    g.used = 0;
    if (current_arena()) return HM_OK;
    if (g.pending) {
        HIP_TRY(hipEventSynchronize(g.done));
        g.pending = false;
    }
    if (bytes > g.cap) {
        if (g.host) (void)hipHostFree(g.host);
        g.host = nullptr;
        g.cap = 0;
        const size_t c = std::max<size_t>(stage_size(bytes), size_t(1) << 20);
        HIP_TRY(hipHostMalloc((void **)&g.host, c, hipHostMallocDefault));
        g.cap = c;
    }
    if (!g.done) HIP_TRY(hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
    return HM_OK;
}

hm_status stage_copy(hm_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return HM_OK;
    Staging &g = ctx->stage;
    if (current_arena() || !g.host) {
        HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        return HM_OK;
    }
    const size_t off = stage_size(g.used);
    if (off + bytes > g.cap) return fail(HM_ERR_LIMIT, "staging round overflows its size");
    std::memcpy(g.host + off, src, bytes);
    g.used = off + bytes;
    HIP_TRY(hipMemcpyAsync(dst, g.host + off, bytes, hipMemcpyHostToDevice, ctx->stream));
    return HM_OK;
}

hm_status stage_end(hm_ctx *ctx) {
    Staging &g = ctx->stage;
    if (current_arena() || !g.host) {   // pageable sources: done before they go
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        return HM_OK;
    }
    HIP_TRY(hipEventRecord(g.done, ctx->stream));
    g.pending = true;
    return HM_OK;
}

hm_status set_device(hm_ctx *ctx) {
    HIP_TRY(hipSetDevice(ctx->device));
    return HM_OK;
}

// ---------------------------------------------------------------------------
// host helpers: exact bin grid, tables, twiddles
// ---------------------------------------------------------------------------

static int bit_length_u128(unsigned __int128 x) {
    int b = 0;
    while (x) { x >>= 1; ++b; }
    return b;
}

// RN(x * num / den), exact: x = M 2^E with a 53-bit integer M; the quotient
// is formed in 128-bit integers with >= 55 bits and rounded half-to-even
// with a sticky bit.
double rn_ratio(double x, int64_t num, int64_t den) {
    if (x == 0.0 || num == 0) return 0.0;
    const bool neg = (x < 0) != (num < 0);
    int e;
    const double f = std::frexp(std::fabs(x), &e);  // |x| = f 2^e, f in [0.5, 1)
    const uint64_t M = (uint64_t)std::ldexp(f, 53);
    const int E = e - 53;
    const unsigned __int128 N = (unsigned __int128)M * (unsigned __int128)(num < 0 ? -num : num);
    const unsigned __int128 D = (unsigned __int128)den;
    int s = 56 + bit_length_u128(D) - bit_length_u128(N);
    if (s < 0) s = 0;
    const unsigned __int128 Ns = N << s;
    const unsigned __int128 Q = Ns / D;
    const bool sticky = (Ns % D) != 0;
    int r = bit_length_u128(Q) - 53;
    unsigned __int128 mant = Q >> r;
    const unsigned __int128 rem = Q & ((((unsigned __int128)1) << r) - 1);
    const unsigned __int128 half = ((unsigned __int128)1) << (r - 1);
    if (rem > half || (rem == half && (sticky || (mant & 1)))) {
        mant += 1;
        if (mant >> 53) { mant >>= 1; ++r; }
    }
    const double v = std::ldexp((double)(uint64_t)mant, E - s + r);
    return neg ? -v : v;
}

// range(-zw, zw, length=nb+1) and its centres (evolution.jl:273-274),
// correctly rounded; step(bin_edges) = RN(2 zw / nb)
void bin_grid(double zw, int64_t nb, double *edges, double *centers, double *step) {
    if (edges) {
        for (int64_t k = 0; k <= nb; ++k) edges[k] = rn_ratio(zw, 2 * k - nb, nb);
        edges[0] = -zw;
        edges[nb] = zw;
    }
    if (centers)
        for (int64_t i = 0; i < nb; ++i) centers[i] = rn_ratio(zw, 2 * i + 1 - nb, nb);
    if (step) *step = (2.0 * zw) / (double)nb;
}

static double jl_max_h(double a, double b) {
    if (a != a) return a;
    if (b != b) return b;
    return (b > a) ? b : a;
}

// wakefield.jl:115-118
static double ref_wake(double z, double wf, double ws, double inv_ctau) {
    return z > 0 ? 0.0 : (wf * std::exp(z * inv_ctau)) * std::cos(ws * z);
}

// utils.jl:58-66
static double ref_delta(double x, double sigma) {
    const double s_inv = 1.0 / (std::sqrt(kTwoPi) * sigma);
    const double ef = -0.5 / (sigma * sigma);
    return s_inv * std::exp((x * x) * ef);
}

// evolution.jl:377-383
double particle_factor(int64_t N) {
    static const double p10[11] = {1e0, 1e1, 1e2, 1e3, 1e4, 1e5, 1e6, 1e7, 1e8, 1e9, 1e10};
    const double lg = std::log10((double)N);
    const int64_t fl = lg >= 0 ? (int64_t)std::floor(lg) : (int64_t)std::ceil(lg - 1.0);
    const int64_t pw = fl < -10 ? -10 : (fl > 10 ? 10 : fl);
    const double den = pw >= 0 ? p10[pw] : 1.0 / p10[-pw];
    return (1e11 / den) * (double)N;
}

int ilog2(int64_t x) {
    int l = 0;
    while ((int64_t(1) << l) < x) ++l;
    return l;
}

// Largest convolution block whose 2B-point complex FFT fits the LDS of one
// workgroup (kMaxFFT points at most).
int max_fft_block(const hm_ctx *ctx) {
    int Bk = kMaxFFT / 2;
    while (Bk > 64 && (size_t)(2 * Bk) * 16 + 1024 > (size_t)ctx->lds_max) Bk >>= 1;
    return Bk;
}

// Convolution block size: the wake support fits one block when it can, so
// every output block needs one forward and one inverse FFT.
// A block of kAutoFFTBlock keeps the transforms small enough that the
// blocked convolution spreads over many workgroups; the choice depends on the
// bunch's own table only, so results do not depend on the batch.
int auto_fft_block(int64_t nw, int64_t nbins, int max_block) {
    (void)nbins;
    const int cap = std::min(max_block, kAutoFFTBlock);
    int Bk = 64;
    while (Bk < nw && Bk < cap) Bk <<= 1;
    return Bk;
}

hm_status twiddles(hm_ctx *ctx, int F, const double **out) {
    auto it = ctx->twiddles.find(F);
    if (it != ctx->twiddles.end() && it->second.p) {
        *out = it->second.p;
        return HM_OK;
    }
    // an allocation-free call takes a table it has not cached from its scratch
    // (valid for that call), any other call caches one per context
    double *p = nullptr;
    if (Arena *a = current_arena()) {
        p = static_cast<double *>(arena_take(a, sizeof(double) * (size_t)F));
        if (!p) return HM_ERR_LIMIT;
    } else {
        auto &buf = ctx->twiddles[F];
        HM_TRY(buf.ensure((size_t)F));  // F/2 complex
        p = buf.p;
    }
    std::vector<double> h((size_t)F);
    const long double tp = 6.283185307179586476925286766559L;
    for (int k = 0; k < F / 2; ++k) {
        const long double a = tp * (long double)k / (long double)F;
        h[2 * k] = (double)cosl(a);
        h[2 * k + 1] = -(double)sinl(a);
    }
    HIP_TRY(hipMemcpy(p, h.data(), sizeof(double) * (size_t)F, hipMemcpyHostToDevice));
    *out = p;
    return HM_OK;
}

hm_status validate(const hm_params *p, const hm_options *o, int64_t n, int64_t nbins) {
    if (!p) return fail(HM_ERR_ARG, "params is NULL");
    if (n < 0) return fail(HM_ERR_ARG, "negative particle count");
    if (p->n_turns < 0) return fail(HM_ERR_ARG, "n_turns must be >= 0");
    if (n >= (int64_t(1) << 32)) return fail(HM_ERR_LIMIT, "at most 2^32-1 particles per bunch");
    if (p->use_wakefield) {
        if (nbins <= 0) return fail(HM_ERR_NBINS, "Buffer nbins invalid (must be > 0).");
        if (nbins > (int64_t(1) << 22)) return fail(HM_ERR_LIMIT, "at most 2^22 bins");
        const bool ext = o && (o->wake_table || o->n_resonators > 0);
        if (!ext && p->pipe_radius <= 0)
            return fail(HM_ERR_PIPE, "Wake parameter pipe_radius must be positive.");
        if (n == 0)
            return fail(HM_ERR_EMPTY_WAKE, "InexactError: log10 of zero particles (wake current)");
        if (o && o->fft_block) {
            const int fb = o->fft_block;
            if (fb < 1 || (fb & (fb - 1)) || fb > kMaxFFT / 2)
                return fail(HM_ERR_ARG, "fft_block must be a power of two <= %d", kMaxFFT / 2);
        }
    }
    if (o && o->serial && p->use_wakefield)   // evolution.jl:267 with comm = nothing
        return fail(HM_ERR_ARG,
                    "MethodError: no method matching initialize_mpi_buffers!(::SimulationBuffers{"
                    "Float64}, ::Int64, ::Nothing, ::Int64) (the wakefield needs use_mpi = true)");
    if (o && o->charge > 0 && o->delta_sigma > 0)
        return fail(HM_ERR_ARG, "charge sets the current in coulombs: the line density is "
                                "counts/n_h without the delta kernel (delta_sigma must be <= 0)");
    if (o && o->n_resonators > 0 && !o->resonators) return fail(HM_ERR_ARG, "resonators is NULL");
    if (o && o->block_size && o->block_size != kPassThreads)
        return fail(HM_ERR_ARG, "block_size must be 0 or %d", kPassThreads);
    if (o && (o->tile_pairs < 0 || o->tile_pairs > (1 << 20) || o->run_pairs < 0))
        return fail(HM_ERR_ARG, "tile_pairs must be in [0, 2^20] and run_pairs >= 0");
    if (o && (o->bin_whole < -1 || o->bin_whole > 0))
        return fail(HM_ERR_ARG, "bin_whole must be 0 or -1");
    return HM_OK;
}

// ---------------------------------------------------------------------------
// Batch
// ---------------------------------------------------------------------------

static int64_t pairs_per_tile_for(int64_t npairs, int threads, int32_t tile_pairs) {
    // tiling depends on the bunch size (and the option) alone, so a bunch is
    // reduced in the same order alone and in any batch: by default about 1024
    // tiles of a large bunch (four per CU), at least one pair per thread.  A
    // workgroup streams its run of consecutive tiles of a bunch as one loop,
    // so small tiles cost no extra pipeline start.
    if (tile_pairs > 0) return (int64_t)tile_pairs * threads;
    const int64_t target_tiles = 1024;
    int64_t ppt = (npairs + target_tiles - 1) / target_tiles;
    ppt = ((ppt + threads - 1) / threads) * threads;
    return std::max<int64_t>(ppt, (int64_t)threads);
}

// The turn pass's order of the tile table.  The balanced ranges give each
// workgroup a contiguous block of entries; a bunch whose tiles fall in the
// blocks of several workgroups has its tiles dealt to them in turn instead
// (tile k of the bunch to the next of them that still needs one, each keeping
// its count), so each workgroup keeps its bunches and its pairs but streams
// tiles spread over the bunch.  Measured on MI355X: 256 workgroups each
// streaming its own contiguous block of two read and two written FP64 arrays
// sustain 3.8 TB/s, the same blocks dealt as 8192-pair tiles 5.2 TB/s.  Each
// workgroup's entries stay in ascending order.
static void deal_tiles(const std::vector<int32_t> &tiles, const int32_t *ranges, int grid,
                       std::vector<int32_t> &out) {
    const int nt = (int)tiles.size();
    const size_t base = out.size();
    out.resize(base + (size_t)nt);
    int32_t *order = out.data() + base;
    for (int t = 0; t < nt; ++t) order[t] = t;
    if (grid <= 1) return;
    std::vector<int> ws;
    std::vector<int32_t> start, need, fill;
    for (int b0 = 0; b0 < nt;) {   // the bunch's tiles: [b0, b1) of the table
        int b1 = b0 + 1;
        while (b1 < nt && tiles[(size_t)b1] == tiles[(size_t)b0]) ++b1;
        ws.clear(); start.clear(); need.clear(); fill.clear();
        for (int w = 0; w < grid; ++w) {
            const int lo = std::max(ranges[(size_t)w], b0), hi = std::min(ranges[(size_t)w + 1], b1);
            if (hi > lo) {
                ws.push_back(w);
                start.push_back(lo);
                need.push_back(hi - lo);
                fill.push_back(0);
            }
        }
        if (ws.size() > 1) {
            int t = b0;
            while (t < b1)
                for (size_t k = 0; k < ws.size() && t < b1; ++k)
                    if (fill[k] < need[k]) order[start[k] + fill[k]++] = t++;
        }
        b0 = b1;
    }
}

// Contiguous ranges of the tile table for `grid` workgroups, [out[g],
// out[g+1]), that make the largest range's load as small as it can be: a
// range's load is its tiles' weights plus, where run costs are given, the
// cost of every run it starts (its first tile, and each tile whose bunch
// differs from the previous tile's: the workgroup stages and flushes that
// bunch's LDS once per run).  The smallest feasible bound is found by
// bisection, each bound checked by filling ranges greedily in order; the fill
// at that bound gives the ranges, and workgroups beyond the last range get
// empty ones.  A workgroup's time is about its load, and the launch lasts as
// long as the largest.
static void balanced_ranges(const std::vector<int64_t> &w, const std::vector<int64_t> *run_cost,
                            const std::vector<int32_t> &bunch, int grid, std::vector<int32_t> &out) {
    if (grid <= 0) return;
    const int nt = (int)w.size();
    auto cost = [&](int t, bool starts) {
        return w[(size_t)t] + (run_cost && starts ? (*run_cost)[(size_t)t] : 0);
    };
    // the ranges the greedy fill needs under bound M (grid + 1: infeasible);
    // with b, their starts after the first
    auto fill = [&](int64_t M, std::vector<int32_t> *b) {
        int count = 0;
        int64_t load = 0;
        bool open = false;
        for (int t = 0; t < nt; ++t) {
            int64_t c = cost(t, !open || bunch[(size_t)t] != bunch[(size_t)t - 1]);
            if (open && load + c > M) {
                if (b) b->push_back((int32_t)t);
                ++count;
                load = 0;
                c = cost(t, true);
            }
            open = true;
            if (c > M || count >= grid) return grid + 1;
            load += c;
        }
        return count + (open ? 1 : 0);
    };
    int64_t lo = 0, hi = 0;
    for (int t = 0; t < nt; ++t) {
        lo = std::max(lo, cost(t, true));
        hi += cost(t, true);
    }
    while (lo < hi) {   // the smallest M with fill(M) <= grid (hi always is)
        const int64_t mid = lo + (hi - lo) / 2;
        if (fill(mid, nullptr) <= grid) hi = mid;
        else lo = mid + 1;
    }
    const size_t base = out.size();
    out.push_back(0);
    std::vector<int32_t> b;
    fill(lo, &b);
    for (int32_t x : b) out.push_back(x);
    while (out.size() < base + (size_t)grid + 1) out.push_back((int32_t)nt);
}

Batch::~Batch() {
    drop_graph();
    for (auto e : ev_pool_) (void)hipEventDestroy(e);
    d_bd_.release(); d_tiles_.release(); d_order_.release();
    for (auto &c : d_conv_) c.release();
    d_red_.release(); d_list_.release(); d_bw_.release();
    d_tmp_tiles_.release(); d_tmp_list_.release(); d_ranges_.release();
}

int Batch::active_count() const {
    int c = 0;
    for (size_t i = 0; i < rb_.size(); ++i)
        if (used_[i] && rb_[i].active) ++c;
    return c;
}

hm_status Batch::upload_descriptors() {
    HM_TRY(stage_begin(ctx_, stage_size(rb_.size() * sizeof(BunchDev))));
    HM_TRY(stage_descriptors());
    return stage_end(ctx_);
}

// the descriptors' copy within a staging round
hm_status Batch::stage_descriptors() {
    HM_TRY(d_bd_.ensure(std::max<size_t>(rb_.size(), 1)));
    std::vector<BunchDev> h(rb_.size());
    for (size_t i = 0; i < rb_.size(); ++i) h[i] = rb_[i].d;
    return stage_copy(ctx_, d_bd_.p, h.data(), h.size() * sizeof(BunchDev));
}

// launch tables of the active bunches
hm_status Batch::build_tables() {
    std::vector<int32_t> tiles, conv[kConvKinds], red, list, bw;
    pass_lds_ = 0;
    bw_lds_ = 0;
    spec_lds_ = 0;
    Fmax_ = 2;
    any_wake_ = any_wake_e0_ = any_reduce_ = false;
    std_only_ = true;
    interp_lds_ = 0;
    // The inverse phase in two kernels (spectral sums per frequency chunk, then
    // the inverse FFTs) pays an extra launch; it is used when the one-kernel
    // form would re-read more than kSpecMinBytes of spectra (nA^2 F / 2 complex
    // pairs per bunch), e.g. for many bunches.  Both forms give the same bits.
    constexpr double kSpecMinBytes = 24e6;
    double reread = 0.0;
    for (size_t i = 0; i < rb_.size(); ++i) {
        const RunBunch &q = rb_[i];
        if (used_[i] && q.active && (q.d.flags & F_WAKE) && q.d.nJ > 1)
            reread += (double)q.d.nA * (double)q.d.nA * (double)q.d.F * 16.0;
    }
    const bool use_spec = reread > kSpecMinBytes;
    for (size_t i = 0; i < rb_.size(); ++i) {
        RunBunch &q = rb_[i];
        if (!used_[i] || !q.active) continue;
        q.d.tile_begin = (int32_t)tiles.size();
        for (int t = 0; t < q.d.ntiles; ++t) tiles.push_back((int32_t)i);
        if ((q.d.flags & kPhysMask) != kStdFlags) std_only_ = false;
        if (q.d.flags & F_WAKE) pass_lds_ = std::max(pass_lds_, (size_t)pass_lds_bytes(q.d));
        // the whole bin phase in one workgroup (k_bin_whole) for a small bunch
        // without interpolation sums; the other bunches take a launch per phase
        {
            const uint32_t f = q.d.flags;
            const bool wake = (f & F_WAKE) != 0;
            const bool dep = wake && (f & F_E0) && (f & F_CIC);
            const size_t lds =
                bin_whole_layout(wake ? q.d.nbins : 0, q.d.nA, q.d.nJ, q.d.F, dep).bytes;
            if (q.o.bin_whole == 0 && q.d.ntiles <= kBinWholeMaxTiles && lds <= kBinWholeMaxLds &&
                (!wake || !(f & F_E0) || (f & F_CIC))) {
                bw.push_back((int32_t)i);
                bw_lds_ = std::max(bw_lds_, lds);
                continue;
            }
        }
        list.push_back((int32_t)i);
        for (int r = 0; r < q.d.nred; ++r) {
            red.push_back((int32_t)i);
            red.push_back(r);
        }
        if (q.d.flags & (F_WAKE | F_E0)) any_reduce_ = true;
        if (q.d.flags & F_WAKE) {
            any_wake_ = true;
            if ((q.d.flags & F_E0) && !(q.d.flags & F_CIC)) {   // the interpolation sums' bunches
                any_wake_e0_ = true;
                interp_lds_ = std::max(interp_lds_, (size_t)q.d.iwin * 8);
            }
            // the inverse phase: spectral sums per frequency chunk, then one
            // inverse FFT per output block (CONV_SPEC, CONV_OUT), or both per
            // output block (CONV_INV: a small re-read volume, or a chunk's
            // spectra larger than the LDS)
            const bool spec = use_spec && q.d.Y;
            const int binv = spec ? CONV_OUT : CONV_INV;
            for (int o = 0; o < q.d.nA; ++o) {
                if (q.d.nJ <= 1) {
                    conv[CONV_FUSED].push_back((int32_t)i);
                    conv[CONV_FUSED].push_back(o);
                } else {
                    conv[CONV_FWD].push_back((int32_t)i);
                    conv[CONV_FWD].push_back(o);
                    conv[binv].push_back((int32_t)i);
                    conv[binv].push_back(o);
                }
            }
            if (q.d.nJ > 1 && spec) {
                for (int c = 0; c * spec_chunk(q.d.nA, q.d.F) < q.d.F; ++c) {
                    conv[CONV_SPEC].push_back((int32_t)i);
                    conv[CONV_SPEC].push_back(c);
                }
                spec_lds_ = std::max(spec_lds_, conv_spec_lds(q.d.nA, q.d.nJ, q.d.F));
            }
            Fmax_ = std::max(Fmax_, q.d.F);
        }
    }
    ntiles_ = (int)tiles.size();
    nred_ = (int)red.size() / 2;
    nlist_ = (int)list.size();
    nbw_ = (int)bw.size();
    // Workgroup ranges: contiguous runs of the tile table with equal numbers
    // of pairs (tiles of different bunches differ in size), for the pass and
    // for the interpolation sums (which only count wake + update_E0 bunches).
    std::vector<int64_t> wpass((size_t)ntiles_), winterp((size_t)ntiles_), rcost((size_t)ntiles_);
    for (int t = 0; t < ntiles_; ++t) {
        const BunchDev &d = rb_[(size_t)tiles[(size_t)t]].d;
        const int64_t lt = t - d.tile_begin;
        const int64_t pairs = std::min(d.pairs_per_tile, d.npairs - lt * d.pairs_per_tile);
        // a run of a bunch's tiles first stages the bunch's LDS (histogram,
        // deposit, window of P) and ends by flushing it: about as long as
        // streaming 4 pairs per bin (the 256-bunch scan at 4096 bins: best at
        // 12288-16384), charged to each range once per run
        const RunBunch &q = rb_[(size_t)tiles[(size_t)t]];
        const int64_t cost = q.o.run_pairs > 0 ? q.o.run_pairs : 4 * (int64_t)d.nbins;
        rcost[(size_t)t] = (d.flags & F_WAKE) ? cost : 0;
        wpass[(size_t)t] = pairs;
        winterp[(size_t)t] = ((d.flags & F_WAKE) && (d.flags & F_E0) && !(d.flags & F_CIC)) ? pairs : 0;
    }
    const int per_cu = interp_lds_ + 1024 <= (size_t)ctx_->lds_max / 2 ? 2 : 1;
    pass_grid_ = std::min(ntiles_, ctx_->cus);
    interp_grid_ = any_wake_e0_ ? std::min(ntiles_, per_cu * ctx_->cus) : 0;
    interp_deep_ = per_cu == 1;
    // ranges: [pass_grid_ + 1] then [interp_grid_ + 1] boundaries; order: the
    // pass's tile order, then the interpolation sums'
    std::vector<int32_t> ranges, order;
    balanced_ranges(wpass, &rcost, tiles, pass_grid_, ranges);
    balanced_ranges(winterp, nullptr, tiles, interp_grid_, ranges);
    deal_tiles(tiles, ranges.data(), pass_grid_, order);
    deal_tiles(tiles, ranges.data() + pass_grid_ + 1, interp_grid_, order);
    // one staging round for every table and the descriptors
    size_t bytes = stage_size(ranges.size() * 4) + stage_size(list.size() * 4) +
                   stage_size(bw.size() * 4) +
                   stage_size(red.size() * 4) + stage_size(tiles.size() * 4) +
                   stage_size(order.size() * 4) +
                   stage_size(rb_.size() * sizeof(BunchDev));
    for (int k = 0; k < kConvKinds; ++k) bytes += stage_size(conv[k].size() * 4);
    HM_TRY(d_ranges_.ensure(std::max<size_t>(ranges.size(), 1)));
    HM_TRY(d_tiles_.ensure(std::max<size_t>(tiles.size(), 1)));
    HM_TRY(d_order_.ensure(std::max<size_t>(order.size(), 1)));
    HM_TRY(d_red_.ensure(std::max<size_t>(red.size(), 2)));
    HM_TRY(d_list_.ensure(std::max<size_t>(list.size(), 1)));
    HM_TRY(d_bw_.ensure(std::max<size_t>(bw.size(), 1)));
    for (int k = 0; k < kConvKinds; ++k) {
        nconv_[k] = (int)conv[k].size() / 2;
        HM_TRY(d_conv_[k].ensure(std::max<size_t>(conv[k].size(), 2)));
    }
    HM_TRY(stage_begin(ctx_, bytes));
    HM_TRY(stage_copy(ctx_, d_ranges_.p, ranges.data(), ranges.size() * 4));
    HM_TRY(stage_copy(ctx_, d_list_.p, list.data(), list.size() * 4));
    HM_TRY(stage_copy(ctx_, d_bw_.p, bw.data(), bw.size() * 4));
    HM_TRY(stage_copy(ctx_, d_red_.p, red.data(), red.size() * 4));
    for (int k = 0; k < kConvKinds; ++k)
        HM_TRY(stage_copy(ctx_, d_conv_[k].p, conv[k].data(), conv[k].size() * 4));
    HM_TRY(stage_copy(ctx_, d_tiles_.p, tiles.data(), tiles.size() * 4));
    HM_TRY(stage_copy(ctx_, d_order_.p, order.data(), order.size() * 4));
    HM_TRY(stage_descriptors());
    HM_TRY(stage_end(ctx_));
    dirty_ = false;
    return HM_OK;
}

// k_moments over the tiles of `which` only (tile_begin relative to that table)
hm_status Batch::run_subset_moments(const std::vector<int32_t> &which, int mode) {
    if (which.empty()) return HM_OK;
    hipStream_t s = ctx_->stream;
    std::vector<int32_t> tiles;
    std::vector<int32_t> saved;
    for (int32_t i : which) {
        saved.push_back(rb_[(size_t)i].d.tile_begin);
        rb_[(size_t)i].d.tile_begin = (int32_t)tiles.size();
        for (int t = 0; t < rb_[(size_t)i].d.ntiles; ++t) tiles.push_back(i);
    }
    HM_TRY(d_tmp_tiles_.ensure(std::max<size_t>(tiles.size(), 1)));
    HM_TRY(d_tmp_list_.ensure(which.size()));
    HM_TRY(stage_begin(ctx_, stage_size(tiles.size() * 4) + stage_size(which.size() * 4) +
                                 stage_size(rb_.size() * sizeof(BunchDev))));
    HM_TRY(stage_copy(ctx_, d_tmp_tiles_.p, tiles.data(), tiles.size() * 4));
    HM_TRY(stage_copy(ctx_, d_tmp_list_.p, which.data(), which.size() * 4));
    HM_TRY(stage_descriptors());
    HM_TRY(stage_end(ctx_));
    // the epilogue stages each bunch's potential in LDS (one workgroup per
    // CU, two where both tables fit) when the largest table fits
    int lds_entries = 0, grid = 0;
    if (mode == 1) {
        int64_t nbmax = 0;
        for (int32_t i : which)
            if (rb_[(size_t)i].d.flags & F_WAKE) nbmax = std::max<int64_t>(nbmax, rb_[(size_t)i].d.nbins);
        const int64_t cap = ((int64_t)ctx_->lds_max - 1024) / 8;
        if (nbmax >= 2) {
            lds_entries = (int)std::min(nbmax, cap);
            grid = ctx_->cus * (2 * ((int64_t)lds_entries * 8 + 1024) <= (int64_t)ctx_->lds_max ? 2 : 1);
        }
    }
    HIP_TRY(launch_moments(d_bd_.p, d_tmp_tiles_.p, (int)tiles.size(), mode, grid, lds_entries, s));
    HIP_TRY(launch_moments_final(d_bd_.p, d_tmp_list_.p, (int)which.size(), mode, s));
    for (size_t k = 0; k < which.size(); ++k) rb_[(size_t)which[k]].d.tile_begin = saved[k];
    dirty_ = true;  // descriptors on the device carry the temporary tile_begin
    return HM_OK;
}

hm_status Batch::epilogue(const std::vector<int32_t> &which) {
    if (which.empty()) return HM_OK;
    HM_TRY(run_subset_moments(which, 1));
    hipStream_t s = ctx_->stream;
    for (int32_t i : which) {
        RunBunch &q = rb_[(size_t)i];
        // two-pass means, spreads and cor(z, dE) of the final bunch -> out[9..13]:
        // the emittance figure of merit (figures_of_merit.jl:36-47) and the
        // compute_std spreads of the serial branch (evolution.jl:464-467)
        if (q.o.fom_stats || q.o.serial) {
            HM_TRY(q.b->w.stats.ensure((size_t)3 * kStatBlocks));
            HIP_TRY(launch_zero(q.d.out + 9, 5 * sizeof(double), s));
            HIP_TRY(launch_pair_stats(q.b->z, q.b->dE, q.b->n, q.d.out + 9, q.b->w.stats.p, s));
        }
        q.active = false;
        q.finished = true;
    }
    dirty_ = true;
    return HM_OK;
}

hm_status Batch::add(const std::vector<RunBunch> &specs, std::vector<int> *idx) {
    std::vector<int32_t> fresh;
    const hm_status st = add_slots(specs, fresh);
    if (st != HM_OK) {
        // no copy of this call may still read the caller's memory, and every
        // slot taken here is released: nothing may launch on the bunches the
        // caller frees after this failure
        (void)hipStreamSynchronize(ctx_->stream);
        for (int32_t i : fresh) {
            used_[(size_t)i] = false;
            rb_[(size_t)i].active = false;
            rb_[(size_t)i].finished = false;
            rb_[(size_t)i].b = nullptr;
        }
        dirty_ = true;
        drop_graph();
        return st;
    }
    if (idx)
        for (int32_t i : fresh) idx->push_back(i);
    return HM_OK;
}

hm_status Batch::add_slots(const std::vector<RunBunch> &specs, std::vector<int32_t> &fresh) {
    hipStream_t s = ctx_->stream;
    for (const RunBunch &spec : specs) {
        HM_TRY(validate(&spec.p, &spec.o, spec.b->n, spec.nbins));
        int slot = -1;
        for (size_t i = 0; i < rb_.size(); ++i)
            if (!used_[i]) { slot = (int)i; break; }
        if (slot < 0) {
            slot = (int)rb_.size();
            rb_.emplace_back();
            used_.push_back(false);
        }
        RunBunch &q = rb_[(size_t)slot];
        q = spec;
        used_[(size_t)slot] = true;
        fresh.push_back((int32_t)slot);
        if (q.fail_setup) return fail(HM_ERR_CANCELLED, "set-up failed by request (test hook)");
        q.turns_done = 0;
        q.n_turns = q.p.n_turns;
        q.finished = false;
        q.active = true;
        hm_bunch *b = q.b;
        BunchDev &d = q.d;
        std::memset(&d, 0, sizeof(d));
        const hm_params *p = &q.p;
        d.z = b->z;
        d.dE = b->dE;
        d.n = b->n;
        d.npairs = (b->n + 1) / 2;
        d.pairs_per_tile = pairs_per_tile_for(d.npairs, kPassThreads, q.o.tile_pairs);
        d.ntiles = (int32_t)((d.npairs + d.pairs_per_tile - 1) / d.pairs_per_tile);
        uint32_t fl = 0;
        if (p->use_wakefield) fl |= F_WAKE;
        if (p->update_eta) fl |= F_ETA;
        if (p->update_E0) fl |= F_E0;
        if (p->SR_damping) fl |= F_SR;
        if (p->use_excitation) fl |= F_QE;
        if (q.o.normals && p->use_excitation) fl |= F_INJ;
        if (q.o.charge > 0) fl |= F_CHARGE;
        if (p->use_wakefield && (q.o.wake_table || q.o.n_resonators > 0)) fl |= F_LAG;
        d.flags = fl;
        d.mass = p->mass;
        d.inv_mass = 1.0 / p->mass;
        d.voltage = p->voltage;
        d.alpha_c = p->alpha_c;
        d.phis = p->phi_s;
        d.sin_phis = std::sin(p->phi_s);
        d.radius = p->radius;
        d.freq_rf = p->freq_rf;
        d.two_pi_h = kTwoPi * (double)p->harmonic;
        d.pf = b->n > 0 ? particle_factor(b->n) : 0.0;
        d.charge = q.o.charge;
        d.E0_init = p->E0;
        d.seed = q.o.seed;
        d.turn_offset = q.o.turn_offset;
        d.nbins = (int32_t)(p->use_wakefield ? q.nbins : 0);
        Work &w = b->w;
        HM_TRY(w.part.ensure((size_t)std::max(d.ntiles, 1) * kTileWaves * 4));
        HM_TRY(w.ipart.ensure((size_t)std::max(d.ntiles, 1) * kTileWaves));
        HM_TRY(w.moments.ensure((size_t)std::max(d.ntiles, 1) * 4));
        HM_TRY(w.sc.ensure(1));
        HM_TRY(w.out.ensure(16));
        HIP_TRY(launch_zero(w.out.p, 16 * sizeof(double), s));
        d.nred = red_slices(d.ntiles);
        HM_TRY(w.red.ensure((size_t)d.nred * 4));
        HM_TRY(w.cnt.ensure(1));
        HIP_TRY(launch_zero(w.cnt.p, sizeof(uint32_t), s));
        d.red = w.red.p;
        d.cnt = w.cnt.p;
        d.part = w.part.p;
        d.ipart = w.ipart.p;
        d.moments = w.moments.p;
        d.sc = w.sc.p;
        d.out = w.out.p;
        if (fl & F_INJ) {
            const size_t cnt = (size_t)q.n_turns * (size_t)b->n;
            HM_TRY(w.normals.ensure(cnt));
            HIP_TRY(hipMemcpyAsync(w.normals.p, q.o.normals, cnt * sizeof(double),
                                   hipMemcpyHostToDevice, s));
            d.normals = w.normals.p;
        }
        if (q.o.monitor && q.n_turns > 0) {   // copied to q.o.monitor by run_batch
            HM_TRY(w.mon.ensure((size_t)q.n_turns * 4));
            HIP_TRY(launch_zero(w.mon.p, (size_t)q.n_turns * 4 * sizeof(double), s));
            d.mon = w.mon.p;
        }
        q.o.normals = nullptr;  // host memory of the caller: not kept
        q.o.wake_table = nullptr;
        q.o.resonators = nullptr;
        q.o.monitor = nullptr;
        // keep the wake description: the bin grid needs sigma_z0 first
        q.wake_copy.clear();
        q.res_copy.clear();
        if (p->use_wakefield) {
            if (spec.o.wake_table)
                q.wake_copy.assign(spec.o.wake_table, spec.o.wake_table + spec.nbins);
            else if (spec.o.n_resonators > 0)
                q.res_copy.assign(spec.o.resonators, spec.o.resonators + 3 * spec.o.n_resonators);
        }
    }
    bool injected = false;
    for (int32_t i : fresh) injected = injected || (rb_[(size_t)i].d.flags & F_INJ);
    if (injected) HIP_TRY(hipStreamSynchronize(s));  // the normals' copies read the caller's memory
    // initial spreads (evolution.jl:245-253): compute_global_std, or for the
    // serial branch compute_std (two passes) into out[0..1]
    HM_TRY(run_subset_moments(fresh, 0));
    for (int32_t i : fresh) {
        RunBunch &q = rb_[(size_t)i];
        if (!q.o.serial || q.b->n < 2) continue;   // n < 2: both forms give 0
        HM_TRY(q.b->w.stats.ensure((size_t)3 * kStatBlocks));
        HIP_TRY(launch_pair_stats(q.b->z, q.b->dE, q.b->n, q.d.out + 9, q.b->w.stats.p, s));
        HIP_TRY(hipMemcpyAsync(q.d.out, q.d.out + 12, sizeof(double), hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(q.d.out + 1, q.d.out + 11, sizeof(double), hipMemcpyDeviceToDevice,
                               s));
    }
    // host tables, uploaded in one staging round after the loop
    std::vector<std::pair<double *, std::vector<double>>> keep;
    for (int32_t i : fresh) {
        RunBunch &q = rb_[(size_t)i];
        BunchDev &d = q.d;
        double o2[2];
        HIP_TRY(hipMemcpyAsync(o2, d.out, sizeof(o2), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        d.sigma_E0 = o2[0];
        const double sz0 = o2[1];
        if (!(d.flags & F_WAKE)) continue;
        // evolution.jl:263-281
        const int64_t nb = q.nbins;
        const double hw = q.o.bin_half_width_sigmas > 0 ? q.o.bin_half_width_sigmas : 7.5;
        double zw = hw * sz0;
        if (zw < 1e-9) zw = 1e-6;
        // with a charge the line density is counts/n_h: the delta kernel
        // belongs to the reference's dimensionless current heuristic
        const double dsig = q.o.charge > 0 ? -1.0 : (q.o.delta_sigma == 0.0 ? 0.15 : q.o.delta_sigma);
        // every centre only where a table is a function of the centres (the
        // reference's wake W(c_j), the delta kernel); else the two ends
        const bool lag_table = !q.wake_copy.empty() || !q.res_copy.empty();
        const bool need_centres = !lag_table || dsig > 0;
        std::vector<double> centers(need_centres ? (size_t)nb : 0);
        double step;
        if (need_centres) {
            bin_grid(zw, nb, nullptr, centers.data(), &step);
            d.c_first = centers[0];
            d.c_last = centers[(size_t)nb - 1];
        } else {
            bin_grid(zw, nb, nullptr, nullptr, &step);
            d.c_first = rn_ratio(zw, 1 - nb, nb);
            d.c_last = rn_ratio(zw, nb - 1, nb);
        }
        d.e_first = -zw;
        d.e_last = zw;
        d.inv_step = 1.0 / step;                                              // utils.jl:378
        d.inv_s = nb > 1 ? 1.0 / ((d.c_last - d.c_first) / (double)(nb - 1)) : 0.0;  // wakefield.jl:39-42
        d.u_last = (d.c_last - d.c_first) * d.inv_s;
        std::vector<double> W, lam;
        const double wf = ((120.0 * kPi) * kSpeedLight) / (kPi * q.p.pipe_radius);   // evolution.jl:279
        const double ws = std::sqrt(jl_max_h(0.0, (2.0 * 30.0) / q.p.pipe_radius));  // evolution.jl:280
        const double inv_ctau = 1.0 / 4e-3;                                          // wakefield.jl:158
        if (dsig > 0) {   // else lam = 1, filled on the device
            lam.resize((size_t)nb);
            for (int64_t k = 0; k < nb; ++k) lam[(size_t)k] = ref_delta(centers[(size_t)k], dsig);
        }
        // the wake table: the caller's, the reference's W(c_j) on the host, or
        // the resonators' W(j step) computed on the device (k_resonator_table;
        // a resonator's wake has no zero tail on any grid here: support nb)
        int64_t nw = nb;
        ResonatorSet rset{};
        if (!q.res_copy.empty()) {
            rset.n = (int32_t)(q.res_copy.size() / 3);
            if (rset.n > kMaxResonators)
                return fail(HM_ERR_LIMIT, "at most %d resonators", kMaxResonators);
            for (int r = 0; r < rset.n; ++r) {
                rset.Rs[r] = q.res_copy[(size_t)(3 * r)];
                rset.fr[r] = q.res_copy[(size_t)(3 * r + 1)];
                rset.Q[r] = q.res_copy[(size_t)(3 * r + 2)];
            }
        } else {
            W.assign((size_t)nb, 0.0);
            for (int64_t k = 0; k < nb; ++k)
                W[(size_t)k] = q.wake_copy.empty() ? ref_wake(centers[(size_t)k], wf, ws, inv_ctau)
                                                   : q.wake_copy[(size_t)k];
            nw = 0;
            for (int64_t k = nb - 1; k >= 0; --k)
                if (W[(size_t)k] != 0.0) { nw = k + 1; break; }
        }
        q.wake_copy.clear();
        q.res_copy.clear();
        const int maxB = max_fft_block(ctx_);
        if (q.o.fft_block > maxB)
            return fail(HM_ERR_LIMIT, "fft_block %d exceeds the LDS limit (%d)", q.o.fft_block, maxB);
        const int Bk = q.o.fft_block ? q.o.fft_block : auto_fft_block(nw, nb, maxB);
        d.B = Bk;
        d.F = 2 * Bk;
        d.logF = ilog2(d.F);
        d.nA = (int32_t)((nb + Bk - 1) / Bk);
        d.nJ = (int32_t)((nw + Bk - 1) / Bk);
        // the histogram in LDS only with room for a window of at least two
        // entries of P beside it (the all-flags pass relies on that)
        if ((size_t)nb * 4 + 16 > (size_t)ctx_->lds_max - 1024) d.flags |= F_GHIST;
        // With update_E0 the pass deposits the interpolation weights of its
        // post-drift z (cloud in cell) where its LDS holds the counters beside
        // the histogram and the whole table of P, instead of the interpolation
        // sums' second read of z.  The counters hold any bunch of fewer than
        // 2^32 particles; a single bin has no cells.
        d.flags &= ~F_CIC;
        if ((d.flags & F_E0) && !(d.flags & F_GHIST) && nb >= 2 && d.n < ((int64_t)1 << 32) &&
            (int64_t)pass_hist_bytes(d) + 8 * nb + cic_bytes((int)nb) <=
                (int64_t)ctx_->lds_max - 1024)
            d.flags |= F_CIC;
        {   // the pass's LDS window of P: what the histogram (and deposit) leave of the LDS
            const int64_t cap = ((int64_t)ctx_->lds_max - 1024 - (int64_t)pass_hist_bytes(d) -
                                 (int64_t)pass_cic_bytes(d)) / 8;
            d.win = cap >= 2 ? (int32_t)std::min<int64_t>(nb, cap) : 0;
            d.cic_off = (d.flags & F_CIC) ? pass_hist_bytes(d) + ((d.win * 8 + 15) & ~15) : 0;
            // the interpolation sums' window: the whole table when it fits
            d.iwin = (int32_t)std::min<int64_t>(nb, ((int64_t)ctx_->lds_max - 1024) / 8);
        }
        Work &w = q.b->w;
        HM_TRY(w.hist.ensure((size_t)nb * 2));
        HM_TRY(w.P.ensure((size_t)nb));
        HM_TRY(w.P2.ensure((size_t)nb * 2));
        HM_TRY(w.lam.ensure((size_t)nb));
        HM_TRY(w.Wtmp.ensure((size_t)nb));
        HM_TRY(w.What.ensure((size_t)std::max(d.nJ, 1) * (size_t)d.F * 2));
        d.Y = nullptr;
        if (d.nJ > 1) {
            HM_TRY(w.X.ensure((size_t)d.nA * (size_t)d.F * 2));
            d.X = w.X.p;
            if (conv_spec_fits(d.nA, d.nJ, d.F, ctx_->lds_max)) {
                HM_TRY(w.Y.ensure((size_t)d.nA * (size_t)d.F * 2));
                d.Y = w.Y.p;
            }
        }
        if (lam.empty())
            HIP_TRY(launch_fill(w.lam.p, (size_t)nb, 1.0, s));
        else
            keep.emplace_back(w.lam.p, std::move(lam));
        if (rset.n > 0)
            HIP_TRY(launch_resonator_table(w.Wtmp.p, (int)nb, step, rset, s));
        else
            keep.emplace_back(w.Wtmp.p, std::move(W));
        HIP_TRY(launch_zero(w.P.p, (size_t)nb * 8, s));
        HIP_TRY(launch_zero(w.P2.p, (size_t)nb * 16, s));
        d.cic = nullptr;
        if (d.flags & F_CIC) {   // cleared again by each turn's finalize
            const size_t words = 2 * (size_t)nb + ((size_t)nb + 2) / 2;   // hi, lo, n
            HM_TRY(w.cic.ensure(words));
            HIP_TRY(launch_zero(w.cic.p, words * 8, s));
            d.cic = w.cic.p;
        }
        d.P2 = w.P2.p;
        d.hist = w.hist.p;
        d.P = w.P.p;
        d.lam = w.lam.p;
        d.What = w.What.p;
        const double *tw;
        HM_TRY(twiddles(ctx_, d.F, &tw));
        d.tw = tw;
    }
    {   // the host tables, the descriptors and the new bunches' list: one round
        HM_TRY(d_tmp_list_.ensure(fresh.size()));
        size_t bytes = stage_size(rb_.size() * sizeof(BunchDev)) + stage_size(fresh.size() * 4);
        for (auto &kv : keep) bytes += stage_size(kv.second.size() * 8);
        HM_TRY(stage_begin(ctx_, bytes));
        for (auto &kv : keep) HM_TRY(stage_copy(ctx_, kv.first, kv.second.data(), kv.second.size() * 8));
        HM_TRY(stage_descriptors());
        HM_TRY(stage_copy(ctx_, d_tmp_list_.p, fresh.data(), fresh.size() * 4));
        HM_TRY(stage_end(ctx_));
        keep.clear();
    }
    for (int32_t i : fresh) {
        RunBunch &q = rb_[(size_t)i];
        if (q.d.flags & F_WAKE)
            HIP_TRY(launch_fft_wake(d_bd_.p + i, q.b->w.Wtmp.p, q.d.nJ, q.d.F, s));
    }
    HIP_TRY(launch_init(d_bd_.p, d_tmp_list_.p, (int)fresh.size(), s));
    std::vector<int32_t> zero_turn;
    for (int32_t i : fresh)
        if (rb_[(size_t)i].n_turns == 0) zero_turn.push_back(i);
    HM_TRY(epilogue(zero_turn));
    dirty_ = true;
    return HM_OK;
}

hm_status Batch::one_turn() {
    hipStream_t s = ctx_->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    if (timing_) {
        if (ev_used_ + 3 > ev_pool_.size()) {
            if (ev_pool_.size() >= 3 * 4096) HM_TRY(flush_timing());
            while (ev_used_ + 3 > ev_pool_.size()) {
                hipEvent_t e;
                HIP_TRY(hipEventCreate(&e));
                ev_pool_.push_back(e);
            }
        }
        e0 = ev_pool_[ev_used_];
        e1 = ev_pool_[ev_used_ + 1];
        e2 = ev_pool_[ev_used_ + 2];
        ev_used_ += 3;
        HIP_TRY(hipEventRecord(e0, s));
    }
    HIP_TRY(launch_turn_pass(d_bd_.p, d_tiles_.p, d_order_.p, d_ranges_.p, pass_grid_, pass_lds_,
                             std_only_ ? 1 : 0, s));
    if (timing_) HIP_TRY(hipEventRecord(e1, s));
    HIP_TRY(launch_bin_whole(d_bd_.p, d_bw_.p, nbw_, bw_lds_, s));
    const bool whole = nlist_ >= kWholeMinBunches;
    if (any_reduce_)
        HIP_TRY(whole ? launch_bin_reduce_whole(d_bd_.p, d_list_.p, nlist_, s)
                      : launch_bin_reduce(d_bd_.p, d_red_.p, nred_, s));
    if (any_wake_) {
        HIP_TRY(launch_conv(CONV_FUSED, d_bd_.p, d_conv_[CONV_FUSED].p, nconv_[CONV_FUSED], Fmax_, s));
        HIP_TRY(launch_conv(CONV_FWD, d_bd_.p, d_conv_[CONV_FWD].p, nconv_[CONV_FWD], Fmax_, s));
        HIP_TRY(launch_conv(CONV_SPEC, d_bd_.p, d_conv_[CONV_SPEC].p, nconv_[CONV_SPEC],
                            (int)spec_lds_, s));
        HIP_TRY(launch_conv(CONV_OUT, d_bd_.p, d_conv_[CONV_OUT].p, nconv_[CONV_OUT], Fmax_, s));
        HIP_TRY(launch_conv(CONV_INV, d_bd_.p, d_conv_[CONV_INV].p, nconv_[CONV_INV], Fmax_, s));
    }
    if (timing_) HIP_TRY(hipEventRecord(e2, s));
    if (any_wake_e0_)   // two workgroups per CU when both windows fit the LDS
        HIP_TRY(launch_interp_sum(d_bd_.p, d_tiles_.p, d_order_.p + ntiles_, d_ranges_.p + pass_grid_ + 1,
                                  interp_grid_,
                                  interp_lds_, interp_deep_, s));
    HIP_TRY(whole ? launch_finalize_whole(d_bd_.p, d_list_.p, nlist_, s)
                  : launch_finalize(d_bd_.p, d_red_.p, nred_, s));
    return HM_OK;
}

hm_status Batch::flush_timing() {
    if (ev_used_ == 0) return HM_OK;
    HIP_TRY(hipEventSynchronize(ev_pool_[ev_used_ - 1]));
    for (size_t k = 0; k + 3 <= ev_used_; k += 3) {
        float a = 0.f, b = 0.f;
        HIP_TRY(hipEventElapsedTime(&a, ev_pool_[k], ev_pool_[k + 1]));
        HIP_TRY(hipEventElapsedTime(&b, ev_pool_[k + 1], ev_pool_[k + 2]));
        pass_ms_ += a;
        ++pass_count_;
        if (any_wake_ || nbw_ > 0) {
            conv_ms_ += b;
            ++conv_count_;
        }
    }
    ev_used_ = 0;
    return HM_OK;
}

void Batch::drop_graph() {
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    graph_exec_ = nullptr;
}

// Captures kGraphTurns turns of the current tables into one executable
// graph (kernel nodes only: the turn reads all of its state from device
// memory, so replays advance the bunches like direct launches).
hm_status Batch::capture_graph() {
    drop_graph();
    hipStream_t s = ctx_->stream;
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hm_status st = HM_OK;
    for (int t = 0; t < kGraphTurns && st == HM_OK; ++t) st = one_turn();
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (st != HM_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess)
        return fail(HM_ERR_DEVICE, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&graph_exec_, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        graph_exec_ = nullptr;
        return fail(HM_ERR_DEVICE, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    }
    return HM_OK;
}

hm_status Batch::advance(int64_t turns) {
    int64_t step = 0;
    for (;;) {
        std::vector<int32_t> retiring;
        for (size_t i = 0; i < rb_.size(); ++i)
            if (used_[i] && rb_[i].active && rb_[i].turns_done >= rb_[i].n_turns)
                retiring.push_back((int32_t)i);
        if (!retiring.empty()) {
            HM_TRY(flush_timing());
            HM_TRY(epilogue(retiring));
        }
        if (step >= turns || active_count() == 0) break;
        if (dirty_) {
            drop_graph();
            HM_TRY(build_tables());
        }
        // turns every active bunch still runs, within this call's budget
        int64_t left = turns - step;
        for (size_t i = 0; i < rb_.size(); ++i)
            if (used_[i] && rb_[i].active) left = std::min(left, rb_[i].n_turns - rb_[i].turns_done);
        // Graphs pay off where the turn is short enough for the launches to
        // matter; a large bunch's turn hides them, and its call skips the
        // capture and instantiation
        int64_t active_particles = 0;
        for (size_t i = 0; i < rb_.size(); ++i)
            if (used_[i] && rb_[i].active) active_particles += rb_[i].d.n;
        int64_t did = 1;
        if (graphs_ && !timing_ && left >= kGraphTurns && active_particles < kGraphMaxParticles) {
            if (!graph_exec_) HM_TRY(capture_graph());
            HIP_TRY(hipGraphLaunch(graph_exec_, ctx_->stream));
            did = kGraphTurns;
        } else {
            HM_TRY(one_turn());
        }
        step += did;
        for (size_t i = 0; i < rb_.size(); ++i)
            if (used_[i] && rb_[i].active) rb_[i].turns_done += did;
    }
    HM_TRY(flush_timing());
    HIP_TRY(hipStreamSynchronize(ctx_->stream));
    return HM_OK;
}

hm_status Batch::finish_now(int i) {
    if (i < 0 || (size_t)i >= rb_.size() || !used_[(size_t)i]) return fail(HM_ERR_ARG, "bad slot");
    if (!rb_[(size_t)i].active) return HM_OK;
    HM_TRY(flush_timing());
    return epilogue({(int32_t)i});
}

hm_status Batch::result(int i, hm_result *r, double *cz, double *cE) {
    if (i < 0 || (size_t)i >= rb_.size() || !used_[(size_t)i]) return fail(HM_ERR_ARG, "bad slot");
    RunBunch &q = rb_[(size_t)i];
    double o[16];
    HIP_TRY(hipMemcpyAsync(o, q.d.out, sizeof(o), hipMemcpyDeviceToHost, ctx_->stream));
    HIP_TRY(hipStreamSynchronize(ctx_->stream));
    std::memset(r, 0, sizeof(*r));
    r->sigma_E0 = o[0];
    r->sigma_z0 = o[1];
    r->sigma_E = o[2];
    r->sigma_z = o[3];
    r->E0 = o[4];
    r->drift_skipped = (int32_t)o[5];
    r->turns_done = (int64_t)o[6];
    r->ms_pass_avg = ms_pass();
    r->ms_conv_avg = ms_conv();
    r->pass_launches = pass_count_;
    r->bin_half_width = (q.d.flags & F_WAKE) ? q.d.e_last : 0.0;
    const bool two_pass = (q.o.fom_stats || q.o.serial) && q.b->n >= 2;
    if (q.o.serial && two_pass) {   // compute_std of the final bunch (evolution.jl:464-467)
        r->sigma_E = o[12];
        r->sigma_z = o[11];
    }
    r->mean_z = two_pass ? o[9] : o[7];
    r->mean_dE = two_pass ? o[10] : o[8];
    r->corr_z_dE = (q.o.fom_stats && q.b->n >= 2) ? o[13] : NAN;
    if (cz) *cz = o[7];
    if (cE) *cE = o[8];
    return HM_OK;
}

void Batch::remove(int i) {
    if (i < 0 || (size_t)i >= rb_.size()) return;
    used_[(size_t)i] = false;
    rb_[(size_t)i].active = false;
    dirty_ = true;
}

}  // namespace hm

using namespace hm;

// ---------------------------------------------------------------------------
// C ABI: context, bunches, evolve
// ---------------------------------------------------------------------------

static hm_status run_batch(hm_ctx *ctx, int32_t nb, hm_bunch *const *bunches, const int64_t *nbins,
                           const hm_params *params, const hm_options *opts, hm_result *results) {
    if (!ctx || nb <= 0 || !bunches || !params || !results) return fail(HM_ERR_ARG, "bad batch arguments");
    HM_TRY(set_device(ctx));
    bool timing = false;
    std::vector<RunBunch> specs;
    std::vector<int32_t> orig;
    for (int32_t i = 0; i < nb; ++i) {
        if (!bunches[i] || bunches[i]->ctx != ctx)
            return fail(HM_ERR_ARG, "bunch %d is NULL or of another context", i);
        RunBunch q;
        q.b = bunches[i];
        q.p = params[i];
        if (opts) q.o = opts[i];
        q.nbins = nbins ? nbins[i] : 0;
        HM_TRY(validate(&q.p, &q.o, q.b->n, q.nbins));
        if (q.o.timing) timing = true;
        if (q.b->n == 0) {   // no wake (validated)
            hm_result &r = results[i];
            std::memset(&r, 0, sizeof(r));
            // MPI branch: E0 becomes 0/0 after an update (evolution.jl:426); the
            // serial branch updates only a non-empty bunch (evolution.jl:437-447)
            r.E0 = (q.p.update_E0 && q.p.n_turns > 0 && !q.o.serial) ? NAN : q.p.E0;
            r.corr_z_dE = NAN;
            r.turns_done = q.p.n_turns;
            continue;
        }
        specs.push_back(q);
        orig.push_back(i);
    }
    if (specs.empty()) return HM_OK;
    Batch B(ctx);
    B.set_timing(timing);
    hipEvent_t t0, t1;
    HIP_TRY(hipEventCreate(&t0));
    HIP_TRY(hipEventCreate(&t1));
    HIP_TRY(hipEventRecord(t0, ctx->stream));
    std::vector<int> idx;
    HM_TRY(B.add(specs, &idx));
    int64_t max_turns = 0;
    for (auto &q : specs) max_turns = std::max<int64_t>(max_turns, q.p.n_turns);
    HM_TRY(B.advance(max_turns));
    HIP_TRY(hipEventRecord(t1, ctx->stream));
    HIP_TRY(hipEventSynchronize(t1));
    float total_ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&total_ms, t0, t1));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    for (size_t k = 0; k < idx.size(); ++k) {
        HM_TRY(B.result(idx[k], &results[orig[k]], nullptr, nullptr));
        results[orig[k]].ms_total = total_ms;
        const RunBunch &q = specs[k];
        if (q.o.monitor && q.p.n_turns > 0)
            HIP_TRY(hipMemcpyAsync(q.o.monitor, q.b->w.mon.p,
                                   (size_t)q.p.n_turns * 4 * sizeof(double), hipMemcpyDeviceToHost,
                                   ctx->stream));
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return HM_OK;
}

extern "C" {

const char *hm_last_error(void) { return g_last_error.c_str(); }

hm_status hm_create(int32_t device, hm_ctx **out) {
    if (!out) return fail(HM_ERR_ARG, "out is NULL");
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(HM_ERR_ARG, "no HIP device %d (%d present)", device, count);
    auto ctx = std::make_unique<hm_ctx>();
    ctx->device = device;
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    std::snprintf(ctx->arch, sizeof(ctx->arch), "%s", prop.gcnArchName);
    if (std::strncmp(ctx->arch, "gfx950", 6) != 0)
        return fail(HM_ERR_DEVICE, "device %d is %s; this library is built for gfx950", device, ctx->arch);
    ctx->cus = prop.multiProcessorCount;
    int lds = 0;
    HIP_TRY(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    ctx->lds_max = lds;
    // best effort: dynamic LDS up to the device limit needs no attribute on
    // this platform, and the runtime may reject the request
    (void)configure_kernels(lds);
    (void)hipGetLastError();
    HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    HM_TRY(noalloc_scratch(&ctx->scratch));   // this device's copy, with the device current
    *out = ctx.release();
    return HM_OK;
}

hm_status hm_destroy(hm_ctx *ctx) {
    if (!ctx) return HM_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->tmp) hm_bunch_destroy(ctx->tmp);
    for (auto &kv : ctx->twiddles) kv.second.release();
    for (auto &b : ctx->sbuf) b.release();
    ctx->shist.release();
    ctx->ssc.release();
    ctx->sbd.release();
    if (ctx->stage.host) (void)hipHostFree(ctx->stage.host);
    if (ctx->stage.done) (void)hipEventDestroy(ctx->stage.done);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return HM_OK;
}

hm_status hm_device_info(hm_ctx *ctx, char *arch, int32_t arch_len, int32_t *cus, int32_t *lds) {
    if (!ctx) return fail(HM_ERR_ARG, "ctx is NULL");
    if (arch && arch_len > 0) std::snprintf(arch, (size_t)arch_len, "%s", ctx->arch);
    if (cus) *cus = ctx->cus;
    if (lds) *lds = ctx->lds_max;
    return HM_OK;
}

hm_status hm_bunch_create(hm_ctx *ctx, int64_t n, hm_bunch **out) {
    if (!ctx || !out) return fail(HM_ERR_ARG, "ctx/out is NULL");
    if (n < 0) return fail(HM_ERR_ARG, "negative particle count");
    HM_TRY(set_device(ctx));
    auto b = std::make_unique<hm_bunch>();
    b->ctx = ctx;
    b->n = n;
    const size_t bytes = (size_t)std::max<int64_t>(n, 1) * sizeof(double);
    hipError_t e = hipMalloc(&b->z, bytes);
    if (e != hipSuccess) return fail(HM_ERR_OOM, "hipMalloc z (%zu bytes): %s", bytes, hipGetErrorString(e));
    e = hipMalloc(&b->dE, bytes);
    if (e != hipSuccess) {
        (void)hipFree(b->z);
        return fail(HM_ERR_OOM, "hipMalloc dE (%zu bytes): %s", bytes, hipGetErrorString(e));
    }
    *out = b.release();
    return HM_OK;
}

hm_status hm_bunch_destroy(hm_bunch *b) {
    if (!b) return HM_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    (void)hipFree(b->z);
    (void)hipFree(b->dE);
    b->w.release();
    delete b;
    return HM_OK;
}

hm_status hm_bunch_upload(hm_bunch *b, const double *z, const double *dE) {
    if (!b || (b->n > 0 && (!z || !dE))) return fail(HM_ERR_ARG, "bad upload arguments");
    HM_TRY(set_device(b->ctx));
    if (b->n == 0) return HM_OK;
    HIP_TRY(hipMemcpyAsync(b->z, z, (size_t)b->n * 8, hipMemcpyHostToDevice, b->ctx->stream));
    HIP_TRY(hipMemcpyAsync(b->dE, dE, (size_t)b->n * 8, hipMemcpyHostToDevice, b->ctx->stream));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    return HM_OK;
}

hm_status hm_bunch_download(hm_bunch *b, double *z, double *dE) {
    if (!b || (b->n > 0 && (!z || !dE))) return fail(HM_ERR_ARG, "bad download arguments");
    HM_TRY(set_device(b->ctx));
    if (b->n == 0) return HM_OK;
    HIP_TRY(hipMemcpyAsync(z, b->z, (size_t)b->n * 8, hipMemcpyDeviceToHost, b->ctx->stream));
    HIP_TRY(hipMemcpyAsync(dE, b->dE, (size_t)b->n * 8, hipMemcpyDeviceToHost, b->ctx->stream));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    return HM_OK;
}

hm_status hm_bunch_device_ptrs(hm_bunch *b, double **zd, double **ed) {
    if (!b) return fail(HM_ERR_ARG, "bunch is NULL");
    if (zd) *zd = b->z;
    if (ed) *ed = b->dE;
    return HM_OK;
}

hm_status hm_bunch_generate(hm_bunch *b, uint64_t seed, double mu_z, double sigma_z, double mu_E,
                            double sigma_E) {
    if (!b) return fail(HM_ERR_ARG, "bunch is NULL");
    HM_TRY(set_device(b->ctx));
    HIP_TRY(launch_generate(seed, mu_z, sigma_z, mu_E, sigma_E, b->z, b->dE, b->n, b->ctx->stream));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    return HM_OK;
}

hm_status hm_bunch_generate_particles(hm_bunch *b, uint64_t seed, double mu_z, double mu_E,
                                     double sigma_z, double sigma_E, double *sigma_E_out,
                                     double *sigma_z_out) {
    if (!b || !sigma_E_out || !sigma_z_out) return fail(HM_ERR_ARG, "bad arguments");
    std::lock_guard<std::recursive_mutex> lk(b->ctx->mu);
    HM_TRY(set_device(b->ctx));
    hipStream_t s = b->ctx->stream;
    if (b->n <= 0) {                        // evolution.jl:47-53
        *sigma_E_out = 0.0;
        *sigma_z_out = 0.0;
        return HM_OK;
    }
    if (b->n == 1) {                        // evolution.jl:61-66: the input spreads
        HIP_TRY(launch_generate(seed, mu_z, sigma_z, mu_E, sigma_E, b->z, b->dE, 1, s));
        HIP_TRY(hipStreamSynchronize(s));
        *sigma_E_out = sigma_E;
        *sigma_z_out = sigma_z;
        return HM_OK;
    }
    HM_TRY(b->w.moments.ensure((size_t)kGenWorkDoubles));
    double *work = b->w.moments.p;
    HIP_TRY(launch_generate_particles(seed, mu_z, mu_E, sigma_z, sigma_E, b->z, b->dE, b->n, work,
                                      s));
    double stats[4];
    HIP_TRY(hipMemcpyAsync(stats, work + 4, sizeof(stats), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *sigma_z_out = stats[2];
    *sigma_E_out = stats[3];
    return HM_OK;
}

hm_status hm_bunch_evolve(hm_bunch *b, int64_t nbins, const hm_params *params,
                          const hm_options *opts, hm_result *result) {
    if (!b) return fail(HM_ERR_ARG, "bunch is NULL");
    std::lock_guard<std::recursive_mutex> lk(b->ctx->mu);
    hm_bunch *arr[1] = {b};
    return run_batch(b->ctx, 1, arr, &nbins, params, opts, result);
}

hm_status hm_batch_evolve(hm_ctx *ctx, int32_t nbunches, hm_bunch *const *bunches,
                          const int64_t *nbins, const hm_params *params, const hm_options *opts,
                          hm_result *results) {
    if (!ctx) return fail(HM_ERR_ARG, "ctx is NULL");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    return run_batch(ctx, nbunches, bunches, nbins, params, opts, results);
}

hm_status hm_evolve(hm_ctx *ctx, double *z, double *dE, int64_t n, int64_t nbins,
                    const hm_params *params, const hm_options *opts, hm_result *result) {
    if (!ctx || !params || !result) return fail(HM_ERR_ARG, "ctx/params/result is NULL");
    if (n < 0 || (n > 0 && (!z || !dE))) return fail(HM_ERR_ARG, "bad particle arrays");
    HM_TRY(validate(params, opts, n, nbins));
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (ctx->tmp && ctx->tmp->n != n) {
        hm_bunch_destroy(ctx->tmp);
        ctx->tmp = nullptr;
    }
    if (!ctx->tmp) HM_TRY(hm_bunch_create(ctx, n, &ctx->tmp));
    HM_TRY(hm_bunch_upload(ctx->tmp, z, dE));
    hm_bunch *arr[1] = {ctx->tmp};
    HM_TRY(run_batch(ctx, 1, arr, &nbins, params, opts, result));
    return hm_bunch_download(ctx->tmp, z, dE);
}

hm_status hm_bin_grid(double zw, int64_t nbins, double *edges, double *centers, double *step) {
    if (nbins <= 0 || !(zw > 0)) return fail(HM_ERR_ARG, "bad arguments");
    bin_grid(zw, nbins, edges, centers, step);
    return HM_OK;
}

}  // extern "C"
