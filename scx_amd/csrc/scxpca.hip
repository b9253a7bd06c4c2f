// Host side of the MI355X randomized sparse PCA: device memory, the power loop
// and the C ABI declared in include/scxpca.h.
//
// Pipeline for one call (reference order: scx-accel/src/pca/cpu.rs:535-605,
// GPU reference scx-gpu/src/gpu_pca.rs:419-697):
//   stats pass   -> mu, sum_sq, col_nnz            (k_col_stats)
//   Omega (host, overlapped with the stats pass)   (omega.cpp)
//   Y = X~ Omega - 1 (mu^T Omega)                  (k_spmm_fwd)
//   Q = CholeskyQR(Y)                              (k_gram, k_chol, k_apply)
//   repeat q times:
//     B = X~^T Q - mu (1^T Q)                      (k_spmm_t + k_spmm_t_reduce)
//     Q_B = CholeskyQR(B) in f64                   (k_gram64, k_chol, k_apply64)
//     Y = X~ Q_B - 1 (mu^T Q_B); Q = CholeskyQR(Y)
//   B = X~^T Q - mu (1^T Q)
//   Rayleigh-Ritz: B = Q_B R (f64 CholeskyQR3), R = U_R S V^T (Jacobi, f64),
//   components = (Q_B U_R)^T, embeddings = Q (V S)
// Every kernel runs on the matrix's stream; the call synchronizes once at the end.
//
// Deviation from the CPU loop, recorded because it changes rounding only:
// the CPU skips the QR of B when n_power_iterations <= 2 (cpu.rs:574-584);
// here B is always orthonormalized before the next forward pass (as the
// reference GPU loop does, pca_operator.rs:100-123). In exact arithmetic both
// span the same subspace; the orthonormal panel keeps the FP32 Gram of the
// next Y well conditioned.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <queue>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/scxpca.h"
#include "kernels.h"

namespace scxpca {
void gaussian_matrix(uint64_t rows, uint64_t cols, uint64_t seed, double *out);
void omega_key(uint64_t seed, uint32_t key[8]);
void omega_tables(double *x, double *f);
bool omega_walk(const uint8_t *c, uint64_t n_words, uint64_t n_samples, uint64_t seed, uint32_t *start,
                std::vector<std::pair<uint32_t, double>> *overrides);
}

using namespace scxpca;

namespace {

thread_local std::string g_last_error;

struct Error : std::runtime_error {
    scxpca_status status;
    Error(scxpca_status s, const std::string &m) : std::runtime_error(m), status(s) {}
};

void hip_check(hipError_t e, const char *what) {
    if (e == hipSuccess) return;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        throw Error(SCXPCA_ERR_OOM, std::string("GPU out of memory: ") + what);
    throw Error(SCXPCA_ERR_DEVICE, std::string("HIP error '") + hipGetErrorString(e) + "' in " + what);
}
#define HIPC(x) hip_check((x), #x)

template <typename F>
scxpca_status guarded(F &&f) {
    try {
        f();
        return SCXPCA_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        return e.status;
    } catch (const std::bad_alloc &) {
        g_last_error = "host out of memory";
        return SCXPCA_ERR_OOM;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return SCXPCA_ERR_DEVICE;
    }
}

// Owned device allocation.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void ensure(size_t n) {
        if (n <= bytes && p) return;
        release();
        if (n == 0) n = 16;
        HIPC(hipMalloc(&p, n));
        bytes = n;
    }
    template <typename T>
    T *as() const { return static_cast<T *>(p); }
};

inline unsigned blocks_for(int64_t n, int per = 256) {
    const int64_t b = (n + per - 1) / per;
    return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(b, 1 << 30)));
}

constexpr int kCB = 128;       // gene width of the row split offsets
// split offsets, the per-(tile, chunk) tile-CSC build with its per-tile column
// partials, and the chunked forward SpMM while their metadata stays under this
// size; the general build and the row-parallel forward beyond
// (SCXPCA_SPLIT_MAX_BYTES, read when a matrix is created, overrides it)
constexpr double kMaxSplitBytes = 32.0 * (1ull << 30);

double split_max_bytes() {
    const char *e = std::getenv("SCXPCA_SPLIT_MAX_BYTES");
    if (e && *e) {
        char *end = nullptr;
        const double v = std::strtod(e, &end);
        if (end != e && *end == '\0' && v >= 0.0) return v;
        throw Error(SCXPCA_ERR_INVALID, "SCXPCA_SPLIT_MAX_BYTES must be a non-negative number");
    }
    return kMaxSplitBytes;
}
constexpr int kMaxColBlocks = 1024;  // partial rows of the column-sum kernels

struct DeviceGuard {
    int prev = 0;
    explicit DeviceGuard(int dev) {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
            throw Error(SCXPCA_ERR_NO_DEVICE, "no ROCm GPU visible");
        if (dev < 0 || dev >= n)
            throw Error(SCXPCA_ERR_NO_DEVICE, "device=" + std::to_string(dev) + " but only " + std::to_string(n) +
                                                  " GPU" + (n == 1 ? " is" : "s are") + " visible");
        HIPC(hipGetDevice(&prev));
        HIPC(hipSetDevice(dev));
    }
    ~DeviceGuard() { (void)hipSetDevice(prev); }
};

}  // namespace

// ---------------------------------------------------------------------------
// Resident matrix
// ---------------------------------------------------------------------------

// Buffers of the device Omega generation. The host vectors stay alive with
// their owner, so the uploads from them need no synchronization.
// Pinned host buffer (hipHostMalloc): copies to and from it are DMA transfers
// that need no staging through a runtime bounce buffer. Reused across calls;
// callers synchronise the stream before rewriting or freeing it.
struct PinnedBuf {
    void *p = nullptr;
    size_t bytes = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf &) = delete;
    PinnedBuf &operator=(const PinnedBuf &) = delete;
    ~PinnedBuf() { release(); }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
    void ensure(size_t n) {
        if (n <= bytes && p) return;
        release();
        if (n == 0) n = 16;
        HIPC(hipHostMalloc(&p, n, hipHostMallocDefault));
        bytes = n;
    }
    template <typename T>
    T *as() const { return static_cast<T *>(p); }
};

struct OmegaScratch {
    DevBuf w, c, v, start, tab, ov;
    PinnedBuf hc, hstart;  // draw outcomes in, draw starts out
    std::vector<uint32_t> hpos;
    std::vector<double> htab, hval;
    // a first attempt issued ahead of other work (omega_begin): its
    // parameters, and an event after the outcomes' copy to the host
    hipEvent_t ev = nullptr;
    bool started = false;
    uint64_t s_rows = 0, s_seed = 0, s_words = 0;
    int s_k = 0;
    OmegaScratch() = default;
    OmegaScratch(const OmegaScratch &) = delete;
    OmegaScratch &operator=(const OmegaScratch &) = delete;
    ~OmegaScratch() {
        if (ev) (void)hipEventDestroy(ev);
    }
};

struct scxpca_matrix {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t n_obs = 0, n_vars = 0, nnz = 0;
    DevBuf own_indptr, own_idx, own_val;
    const int64_t *indptr = nullptr;
    const uint32_t *idx = nullptr;
    const float *val = nullptr;
    double target = 0.0;
    int flags = 0;
    DevBuf rowf;
    // PFlog (v4) matrix: X~ is delta = log1p(4 alpha x) and the operator is
    // delta + b 1^T with the per-cell baseline b (device, f64)
    bool pflog = false;
    scxpca_pflog_params pf_params{};
    scxpca_pflog_fit pf_fit{};
    DevBuf pfb, pfstat;          // baseline; its sum and sum of squares
    PinnedBuf hpf;               // host copy of pfstat
    double pf_bsum = 0.0, pf_b2 = 0.0;
    DevBuf spl;
    int nsplit = 0;
    DevBuf xt;  // X~ values in CSR order (f32), computed by prepare()
    DevBuf lg;    // gene mod kChunkMod per entry (u16), the forward SpMM's LDS chunk index
    // the BF16 sketch form's packed entries: xpk (CSR order, bf16(x~) | gene
    // mod kChunkMod; split layout only) and tpk (tile-CSC order)
    DevBuf xpk;
    bool xpk_ready = false, xpk_in_prepare = false;
    // lg of the current prepare is written (a prepare that writes xpk leaves
    // it out; the FP32 forward unpacks it from xpk when it runs)
    bool lg_ready = false;
    bool tpk_ready = false;     // tpk holds the packed bf16 entries of the current prepare
    bool tpk_in_build = false;  // the BF16 form ran on this matrix: the tile-CSC build writes tpk too
    // row-tiled CSC (tiles of kTileRows rows, gene-major inside a tile)
    DevBuf tptr, trow, tval, tpk;
    int n_tiles = 0;
    // entries per gene (host copy) and the load-balanced gene order of the
    // transposed SpMM (GW genes x 64 16-lane groups x nch blocks)
    DevBuf gcnt;
    std::vector<uint32_t> host_gcnt;
    struct GeneOrder {
        int gw = 0, nch = 0, tw = 0;
        bool valid = false;
        PinnedBuf host;  // kept alive for the asynchronous upload
        DevBuf dev;
    } gorder;
    bool use_split = false;  // split offsets, k_build_tcsc4 (+ per-tile column partials), k_spmm_fwd4
    // cached column statistics of X~
    bool stats_ready = false;
    DevBuf tsum, tsq;  // per-tile column-statistics partials (written by the tile-CSC build)
    DevBuf sum, sq, cnt, psum, psq, pcnt;
    // PCA workspace
    DevBuf y, q, panel, panel2, panel64, b64, t64a, t64b, tpart, gpart, cpart, small, status;
    PinnedBuf hgcnt;  // host side of gcnt: per-gene counts + the validation word
    DevBuf dres;      // packed end-of-call results (k_pack_results)
    PinnedBuf hres;   // their host copy
    OmegaScratch om;  // device Omega scratch (device and host buffers)
    // stage profiler (events on the matrix stream; off unless enabled)
    bool profiling = false;
    uint32_t prof_mask = ~0u;  // stages the profiler times (bit s = scxpca_stage s)
    struct Mark { int stage; hipEvent_t a, b; };
    std::vector<Mark> marks;

    void clear_marks() {
        for (auto &mk : marks) { (void)hipEventDestroy(mk.a); (void)hipEventDestroy(mk.b); }
        marks.clear();
    }
    // covariance route: Gram partials, pair tables, covariance / eigenvectors,
    // eigenvalues and their workspace, the solver handle (created on first use)
    DevBuf covpart, covtab, cov, evw, eve, evinfo, covvar;
    rocblas_handle rb = nullptr;
    ~scxpca_matrix() {
        clear_marks();
        if (rb) (void)rocblas_destroy_handle(rb);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// Records a start/stop event pair around one stage launch when profiling.
struct StageScope {
    scxpca_matrix *m;
    hipStream_t s;
    int idx = -1;
    StageScope(scxpca_matrix *mm, hipStream_t ss, int stage) : m(mm), s(ss) {
        if (!m->profiling || !((m->prof_mask >> stage) & 1u)) return;
        scxpca_matrix::Mark mk{stage, nullptr, nullptr};
        HIPC(hipEventCreate(&mk.a));
        HIPC(hipEventCreate(&mk.b));
        HIPC(hipEventRecord(mk.a, s));
        m->marks.push_back(mk);
        idx = static_cast<int>(m->marks.size()) - 1;
    }
    ~StageScope() {
        if (idx >= 0) (void)hipEventRecord(m->marks[idx].b, s);
    }
};


// Row factors, X~, chunk offsets and the tile-CSC of the current CSR, with
// the matrix validated in the same pass. Returns the validation bits of
// k_prepare_rows (0: valid); with any bit set the builds did nothing and no
// later stage may run on this matrix.
int prepare(scxpca_matrix *m) {
    m->stats_ready = false;
    m->tpk_ready = false;
    m->xpk_ready = false;
    if (m->n_obs == 0) return 0;
    StageScope prof(m, m->stream, SCXPCA_STAGE_PREPARE);
    // gcnt holds the per-gene entry counts, at [n_vars] the validation bits of
    // k_prepare_rows and at the next even word the complemented position of
    // the first entry that failed a value check: one copy brings all back
    const int64_t epw = (m->n_vars + 2) / 2 * 2, gwords = epw + 2;
    m->gcnt.ensure(sizeof(uint32_t) * gwords);
    int *vflag = reinterpret_cast<int *>(m->gcnt.as<uint32_t>() + m->n_vars);
    unsigned long long *errpos = reinterpret_cast<unsigned long long *>(m->gcnt.as<uint32_t>() + epw);
    HIPC(hipMemsetAsync(m->gcnt.p, 0, sizeof(uint32_t) * gwords, m->stream));
    const bool want_b = (m->flags & kXfScale) != 0;
    constexpr int kSumParts = 256;  // blocks of the baseline sum
    if (want_b) {
        m->pfb.ensure(sizeof(double) * m->n_obs);
        m->pfstat.ensure(sizeof(double) * (2 + 2 * kSumParts));
        m->hpf.ensure(sizeof(double) * 2);
    }
    const unsigned row_blocks = std::min(blocks_for(m->n_obs, 4), 65536u);  // one row per wave
    const bool want_xpk = m->xpk_in_prepare && m->use_split;
    static_assert(kCB == 128, "k_prepare_rows counts 128-gene split boundaries");
    auto prep = (m->flags & kXfLogF64) ? k_prepare_rows<true, true>
                : (m->flags & (kXfCheck | kXfScale)) ? k_prepare_rows<true, false> : k_prepare_rows<false, false>;
    prep<<<row_blocks, 256, 0, m->stream>>>(m->indptr, m->idx, m->val, m->n_obs, m->target, m->flags,
                                            m->use_split ? m->nsplit : 0,
                                            (m->flags & kXfNorm) ? m->rowf.as<double>() : nullptr,
                                            m->xt.as<float>(), want_xpk ? nullptr : m->lg.as<uint16_t>(),
                                            want_xpk ? m->xpk.as<uint32_t>() : nullptr,
                                            m->use_split ? m->spl.as<uint32_t>() : nullptr, m->n_vars,
                                            vflag, want_b ? m->pfb.as<double>() : nullptr, errpos);
    HIPC(hipGetLastError());
    m->xpk_ready = want_xpk;
    m->lg_ready = !want_xpk;
    if (m->use_split)
    {
        const int nck = static_cast<int>((m->n_vars + kB4Genes - 1) / kB4Genes);
        auto build = m->tpk_in_build ? k_build_tcsc4<true> : k_build_tcsc4<false>;
        build<<<static_cast<unsigned>((m->n_tiles + 7) / 8 * 8) * nck, 512, 0, m->stream>>>(
            m->indptr, m->idx, m->xt.as<float>(), m->spl.as<uint32_t>(), m->nsplit, nck, m->n_obs, m->n_vars,
            m->tptr.as<uint32_t>(), m->trow.as<uint8_t>(), m->tval.as<float>(), m->tsum.as<double>(),
            m->tsq.as<double>(), vflag, m->tpk_in_build ? m->tpk.as<uint32_t>() : nullptr);
        m->tpk_ready = m->tpk_in_build;
    }
    else
        k_build_tcsc<<<m->n_tiles, 512, 0, m->stream>>>(m->indptr, m->idx, m->xt.as<float>(), m->n_obs,
                                                         static_cast<int>(m->n_vars), m->tptr.as<uint32_t>(),
                                                         m->trow.as<uint8_t>(), m->tval.as<float>(),
                                                         vflag);
    HIPC(hipGetLastError());
    // entries per gene (counted into the zeroed gcnt) -> host: the transposed
    // SpMM's gene order is built from them
    if (m->n_vars > 0) {
        const int tpb = 64;
        const dim3 grid(static_cast<unsigned>((m->n_vars + 255) / 256), static_cast<unsigned>((m->n_tiles + tpb - 1) / tpb));
        k_gene_counts_tcsc<<<grid, 256, 0, m->stream>>>(m->tptr.as<uint32_t>(), m->n_tiles, tpb, m->n_vars,
                                                         m->gcnt.as<uint32_t>());
        HIPC(hipGetLastError());
    }
    if (want_b) {  // sum and sum of squares of the baseline (fixed order)
        double *parts = m->pfstat.as<double>() + 2;
        k_sum_sq_f64<<<kSumParts, 256, 0, m->stream>>>(m->pfb.as<double>(), m->n_obs, parts, nullptr);
        HIPC(hipGetLastError());
        k_sum_sq_f64<<<1, 256, 0, m->stream>>>(parts, kSumParts, nullptr, m->pfstat.as<double>());
        HIPC(hipGetLastError());
        HIPC(hipMemcpyAsync(m->hpf.p, m->pfstat.p, sizeof(double) * 2, hipMemcpyDeviceToHost, m->stream));
    }
    m->hgcnt.ensure(sizeof(uint32_t) * gwords);
    HIPC(hipMemcpyAsync(m->hgcnt.p, m->gcnt.p, sizeof(uint32_t) * gwords, hipMemcpyDeviceToHost, m->stream));
    HIPC(hipStreamSynchronize(m->stream));
    const uint32_t *hg = m->hgcnt.as<uint32_t>();
    m->host_gcnt.assign(hg, hg + m->n_vars);
    m->gorder.valid = false;
    if (want_b) {
        m->pf_bsum = m->hpf.as<double>()[0];
        m->pf_b2 = m->hpf.as<double>()[1];
    }
    return static_cast<int>(hg[m->n_vars]);
}

// Position of the first entry that failed k_prepare_rows' value check.
int64_t error_position(const scxpca_matrix *m) {
    const int64_t epw = (m->n_vars + 2) / 2 * 2;
    unsigned long long w = 0;
    std::memcpy(&w, m->hgcnt.as<uint32_t>() + epw, sizeof(w));
    return static_cast<int64_t>(~w);
}

// Rust's Display of an f32 for the common cases (shortest round-trip digits).
std::string fmt_f32(float v) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    char buf[64];
    for (int prec = 1; prec <= 9; ++prec) {
        std::snprintf(buf, sizeof(buf), "%.*g", prec, static_cast<double>(v));
        if (std::strtof(buf, nullptr) == v) break;
    }
    return buf;
}

// The rejection for the validation bits of prepare(). The value checks
// report the first offending entry as the reference does: estimate_alpha's
// finite guard (finite.rs:22-32), delta_from_raw_csr (pyscx/src/accel/
// pflog.rs:375-397) and pflog_baseline_from_delta's guard (pflog.rs:46-56).
// row_base / entry_base: the matrix's first row and entry inside the caller's
// CSR (a streamed shard reports positions in the whole matrix).
void throw_invalid(const scxpca_matrix *m, int h, int64_t row_base = 0, int64_t entry_base = 0) {
    if (h & 4) throw Error(SCXPCA_ERR_INVALID, "indptr must be non-decreasing");
    if (h & 2) throw Error(SCXPCA_ERR_INVALID, "column index out of range");
    if (h & 1) throw Error(SCXPCA_ERR_INVALID, "device CSR rows must have sorted column indices");
    if (h & 8)
        throw Error(SCXPCA_ERR_INVALID, "a CSR row repeats a column index (sum duplicate entries before calling pca)");
    if (h & (16 | 32)) {
        const int64_t pos = error_position(m);
        float v = 0.f;
        HIPC(hipMemcpy(&v, m->val + pos, sizeof(float), hipMemcpyDefault));
        if (!(m->flags & kXfScale))
            throw Error(SCXPCA_ERR_INVALID, "HVG input contains a non-finite value (" + fmt_f32(v) +
                                                ") at nonzero index " + std::to_string(pos + entry_base) +
                                                "; this accelerator requires finite input \u2014 filter/QC NaN "
                                                "and Inf before computing variance");
        if (m->flags & kXfLogF64) {
            // the cell: last row whose start is <= pos
            int64_t lo = 0, hi = m->n_obs;
            while (hi - lo > 1) {
                const int64_t mid = lo + (hi - lo) / 2;
                int64_t a = 0;
                HIPC(hipMemcpy(&a, m->indptr + mid, sizeof(int64_t), hipMemcpyDefault));
                if (a <= pos) lo = mid; else hi = mid;
            }
            if (!std::isfinite(v))
                throw Error(SCXPCA_ERR_INVALID, "pflog: non-finite count " + fmt_f32(v) + " at cell " +
                                                    std::to_string(lo + row_base) + "; counts must be finite");
            throw Error(SCXPCA_ERR_INVALID, "pflog: negative count " + fmt_f32(v) + " at cell " +
                                                std::to_string(lo + row_base) + "; counts must be non-negative");
        }
        const float d = std::log1p(static_cast<float>(static_cast<double>(v) * m->target));
        throw Error(SCXPCA_ERR_INVALID, "PFlog input contains a non-finite value (" + fmt_f32(d) +
                                            ") at nonzero index " + std::to_string(pos + entry_base) +
                                            "; PFlog requires finite raw counts \u2014 filter/QC NaN and Inf first");
    }
}

// Gene order for the transposed SpMM (k_spmm_t7: tw gene owners per block,
// the 16-lane groups of its waves, gw slots each). Genes sorted by entry
// count (descending, ties by index) are cut into units of four consecutive
// genes, one per group of a wave (the four run in lockstep, so a unit's genes
// have neighbouring counts). Units go, largest first (cost: the unit's
// largest count), to the wave with the least total so far among the
// nch * tw / 4 waves with a free slot (ties: the lowest wave), a wave's k-th
// unit filling its slot k: per-wave totals, and so the blocks' totals, come
// out within about one unit of each other, and slot 0 holds each wave's
// densest genes. Owner o = wave * 4 + group takes h[o * gw + slot]; unused
// slots hold n_vars. The dealt order into h[0, max(S, 1)), S = nch * tw * gw.
void deal_gene_order(const scxpca_matrix *m, int gw, int nch, int tw, uint32_t *h) {
    constexpr int quad = 4;
    const int64_t nv = m->n_vars;
    const int64_t S = static_cast<int64_t>(nch) * tw * gw;
    const int64_t waves = static_cast<int64_t>(nch) * tw / quad;
    const int64_t nc = static_cast<int64_t>(m->host_gcnt.size());
    auto count = [&](int64_t c) { return c < nc ? m->host_gcnt[c] : 0u; };
    // genes by count descending, ties by index: a stable LSD radix sort of the
    // indices on the complemented counts (11-bit digits, only as many passes
    // as the largest count needs)
    std::vector<uint32_t> key(nv), tmp(nv);
    uint32_t cmax = 0;
    for (int64_t c = 0; c < nv; ++c) {
        key[c] = static_cast<uint32_t>(c);
        cmax = std::max(cmax, count(c));
    }
    constexpr int kDigit = 11;
    for (int shift = 0; shift < 32 && (shift == 0 || (cmax >> shift) != 0); shift += kDigit) {
        uint32_t hist[(1 << kDigit) + 1] = {};
        auto digit = [&](uint32_t g) { return ((~count(g)) >> shift) & ((1u << kDigit) - 1u); };
        for (int64_t i = 0; i < nv; ++i) ++hist[digit(key[i]) + 1];
        for (int d = 0; d < (1 << kDigit); ++d) hist[d + 1] += hist[d];
        for (int64_t i = 0; i < nv; ++i) tmp[hist[digit(key[i])]++] = key[i];
        key.swap(tmp);
    }
    for (int64_t i = 0; i < std::max<int64_t>(S, 1); ++i) h[i] = static_cast<uint32_t>(nv);
    // min-heap of (load, wave): waves with a free slot
    using Item = std::pair<uint64_t, int64_t>;
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> heap;
    for (int64_t w = 0; w < waves; ++w) heap.push({0u, w});
    std::vector<int> used(static_cast<size_t>(waves), 0);
    for (int64_t u = 0; u * quad < nv && !heap.empty(); ++u) {
        const Item top = heap.top();
        heap.pop();
        const int64_t w = top.second;
        const int slot = used[static_cast<size_t>(w)]++;
        const uint32_t first = static_cast<uint32_t>(key[u * quad]);
        for (int g = 0; g < quad && u * quad + g < nv; ++g)
            h[(w * quad + g) * gw + slot] = static_cast<uint32_t>(key[u * quad + g]);
        if (used[static_cast<size_t>(w)] < gw) heap.push({top.first + count(first), w});
    }
}

bool gene_order_valid(const scxpca_matrix *m, int gw, int nch, int tw) {
    const auto &go = m->gorder;
    return go.valid && go.gw == gw && go.nch == nch && go.tw == tw;
}

const uint32_t *ensure_gene_order(scxpca_matrix *m, int gw, int nch, int tw) {
    auto &go = m->gorder;
    if (gene_order_valid(m, gw, nch, tw)) return go.dev.as<uint32_t>();
    const int64_t S = static_cast<int64_t>(nch) * tw * gw;
    // the previous upload from go.host (if any) was drained by prepare's sync
    go.host.ensure(sizeof(uint32_t) * std::max<int64_t>(S, 1));
    deal_gene_order(m, gw, nch, tw, go.host.as<uint32_t>());
    uint32_t *h = go.host.as<uint32_t>();
    go.dev.ensure(sizeof(uint32_t) * std::max<int64_t>(S, 1));
    HIPC(hipMemcpyAsync(go.dev.p, h, sizeof(uint32_t) * std::max<int64_t>(S, 1), hipMemcpyHostToDevice, m->stream));
    go.gw = gw;
    go.nch = nch;
    go.tw = tw;
    go.valid = true;
    return go.dev.as<uint32_t>();
}

// Layout choice and (grow-only) buffers of the derived data for the current
// n_obs / n_vars / nnz of m, before a prepare.
void size_derived(scxpca_matrix *m) {
    m->nsplit = static_cast<int>((m->n_vars + kCB - 1) / kCB);
    m->n_tiles = static_cast<int>((m->n_obs + kTileRows - 1) / kTileRows);
    const double split_bytes = 4.0 * static_cast<double>(m->n_obs) * (m->nsplit + 1) +
                               16.0 * static_cast<double>(m->n_tiles) * static_cast<double>(m->n_vars);
    m->use_split = split_bytes <= split_max_bytes();
    if (m->flags & kXfNorm) m->rowf.ensure(sizeof(double) * m->n_obs);
    if (m->use_split) m->spl.ensure(sizeof(uint32_t) * m->n_obs * (m->nsplit + 1));
    m->tptr.ensure(sizeof(uint32_t) * static_cast<size_t>(m->n_tiles) * (m->n_vars + 1));
    m->trow.ensure(m->nnz + kEntryPad);
    m->tval.ensure(sizeof(float) * (m->nnz + kEntryPad));
    if (m->tpk_in_build) m->tpk.ensure(sizeof(uint32_t) * (m->nnz + kEntryPad));
    if (m->xpk_in_prepare && m->use_split) m->xpk.ensure(sizeof(uint32_t) * std::max<int64_t>(m->nnz, 1));
    m->xt.ensure(sizeof(float) * std::max<int64_t>(m->nnz, 1));
    m->lg.ensure(sizeof(uint16_t) * std::max<int64_t>(m->nnz, 1));
    if (m->use_split) {
        m->tsum.ensure(sizeof(double) * static_cast<size_t>(m->n_tiles) * std::max<int64_t>(m->n_vars, 1));
        m->tsq.ensure(sizeof(double) * static_cast<size_t>(m->n_tiles) * std::max<int64_t>(m->n_vars, 1));
    }
}

int transform_flags(double target, int log1p) { return (target > 0.0 ? kXfNorm : 0) | (log1p ? kXfLog1p : 0); }

// SCXPCA_FAIL_RESIDENT_BUILD=1 makes every resident build of a host CSR fail
// as if HBM ran out while its derived structures were allocated: a host-side
// switch for exercising the decline to the streamed operator.
bool forced_resident_failure() {
    const char *e = std::getenv("SCXPCA_FAIL_RESIDENT_BUILD");
    return e && e[0] == '1' && e[1] == '\0';
}

// Upload (or borrow) the CSR and prepare it under the transform (target,
// flags: the kXf* bits).
void build_matrix(scxpca_matrix *m, const scxpca_csr *x, double target, int flags) {
    if (x->n_obs < 0 || x->n_vars < 0) throw Error(SCXPCA_ERR_INVALID, "matrix shape must be non-negative");
    if (x->n_vars > 0xffffffffll) throw Error(SCXPCA_ERR_INVALID, "n_vars exceeds the u32 index range");
    m->n_obs = x->n_obs;
    m->n_vars = x->n_vars;
    m->target = target;
    m->flags = flags;
    HIPC(hipStreamCreate(&m->stream));
    if (x->on_device) {
        m->indptr = x->indptr;
        m->idx = x->indices;
        m->val = x->data;
        HIPC(hipMemcpy(&m->nnz, x->indptr + x->n_obs, sizeof(int64_t), hipMemcpyDeviceToHost));
        if (m->n_obs > 0) {
            // indptr first: the row pass dereferences entries through it
            DevBuf flag;
            flag.ensure(sizeof(int));
            HIPC(hipMemsetAsync(flag.p, 0, sizeof(int), m->stream));
            k_check_indptr<<<blocks_for(m->n_obs, 256), 256, 0, m->stream>>>(m->indptr, m->n_obs, flag.as<int>());
            HIPC(hipGetLastError());
            int h = 0;
            HIPC(hipMemcpyAsync(&h, flag.p, sizeof(int), hipMemcpyDeviceToHost, m->stream));
            HIPC(hipStreamSynchronize(m->stream));
            if (h) throw Error(SCXPCA_ERR_INVALID, "indptr must be non-decreasing");
        }
    } else {
        if (x->indptr[0] != 0) throw Error(SCXPCA_ERR_INVALID, "indptr must be non-decreasing");
        for (int64_t r = 0; r < x->n_obs; ++r)
            if (x->indptr[r + 1] < x->indptr[r]) throw Error(SCXPCA_ERR_INVALID, "indptr must be non-decreasing");
        m->nnz = x->indptr[x->n_obs];
        // the entries are uploaded as given and checked on the device by the
        // row pass; rows found unsorted are sorted here and uploaded again
        m->own_indptr.ensure(sizeof(int64_t) * (x->n_obs + 1));
        m->own_idx.ensure(sizeof(uint32_t) * std::max<int64_t>(m->nnz, 1));
        m->own_val.ensure(sizeof(float) * std::max<int64_t>(m->nnz, 1));
        HIPC(hipMemcpyAsync(m->own_indptr.p, x->indptr, sizeof(int64_t) * (x->n_obs + 1), hipMemcpyHostToDevice, m->stream));
        if (m->nnz) {
            HIPC(hipMemcpyAsync(m->own_idx.p, x->indices, sizeof(uint32_t) * m->nnz, hipMemcpyHostToDevice, m->stream));
            HIPC(hipMemcpyAsync(m->own_val.p, x->data, sizeof(float) * m->nnz, hipMemcpyHostToDevice, m->stream));
        }
        m->indptr = m->own_indptr.as<int64_t>();
        m->idx = m->own_idx.as<uint32_t>();
        m->val = m->own_val.as<float>();
    }
    if (m->n_obs == 0) return;
    if (!x->on_device && forced_resident_failure())
        throw Error(SCXPCA_ERR_OOM, "GPU out of memory: resident CSR build (forced by SCXPCA_FAIL_RESIDENT_BUILD)");
    size_derived(m);
    int h = prepare(m);
    if ((h & 1) && !x->on_device) {
        // canonicalize host rows (stable sort by index) and upload again
        std::vector<uint32_t> sidx(x->indices, x->indices + m->nnz);
        std::vector<float> sval(x->data, x->data + m->nnz);
        std::vector<int64_t> perm;
        for (int64_t r = 0; r < x->n_obs; ++r) {
            const int64_t a = x->indptr[r], b = x->indptr[r + 1];
            perm.resize(b - a);
            for (int64_t i = 0; i < b - a; ++i) perm[i] = a + i;
            std::stable_sort(perm.begin(), perm.end(),
                             [&](int64_t u, int64_t v) { return x->indices[u] < x->indices[v]; });
            for (int64_t i = 0; i < b - a; ++i) {
                sidx[a + i] = x->indices[perm[i]];
                sval[a + i] = x->data[perm[i]];
            }
        }
        HIPC(hipMemcpyAsync(m->own_idx.p, sidx.data(), sizeof(uint32_t) * m->nnz, hipMemcpyHostToDevice, m->stream));
        HIPC(hipMemcpyAsync(m->own_val.p, sval.data(), sizeof(float) * m->nnz, hipMemcpyHostToDevice, m->stream));
        HIPC(hipStreamSynchronize(m->stream));  // staging vectors go out of scope
        h = prepare(m);
    }
    throw_invalid(m, h);
}

// The BF16 sketch passes' X~: the tile-CSC entries packed as bf16(x~) | row
// << 16 | (gene mod 64) << 24 (one 4-byte load per entry, the gene without
// a search over the run starts)
void ensure_tpk(scxpca_matrix *m) {
    if (m->tpk_ready) return;
    m->tpk.ensure(sizeof(uint32_t) * (m->nnz + kEntryPad));
    if (m->n_obs > 0 && m->n_vars > 0) {
        const int64_t groups = static_cast<int64_t>(m->n_tiles) * ((m->n_vars + 63) / 64);
        k_pack_tile_entries<<<blocks_for(groups, 4), 256, 0, m->stream>>>(
            m->indptr, m->tptr.as<uint32_t>(), m->trow.as<uint8_t>(), m->tval.as<float>(), m->n_vars, m->n_tiles,
            m->tpk.as<uint32_t>());
    }
    // later prepares of this matrix (the split layout) write the packed entries in the tile-CSC build
    m->tpk_in_build = true;
    HIPC(hipGetLastError());
    m->tpk_ready = true;
}

// The BF16 sketch form's forward entries on the split layout: xpk =
// bf16(x~) << 16 | gene mod kChunkMod, packed from xt / lg once, written by
// later prepares directly
void ensure_xpk(scxpca_matrix *m) {
    if (!m->use_split || m->xpk_ready) return;
    m->xpk.ensure(sizeof(uint32_t) * std::max<int64_t>(m->nnz, 1));
    if (m->nnz > 0) {
        k_pack_row_entries<<<blocks_for(m->nnz), 256, 0, m->stream>>>(m->xt.as<float>(), m->lg.as<uint16_t>(),
                                                                      m->nnz, m->xpk.as<uint32_t>());
        HIPC(hipGetLastError());
    }
    m->xpk_in_prepare = true;
    m->xpk_ready = true;
}

// lg for the FP32 forward after a prepare that wrote xpk instead
void ensure_lg(scxpca_matrix *m) {
    if (m->lg_ready || !m->use_split) return;
    if (m->nnz > 0) {
        k_unpack_row_genes<<<blocks_for(m->nnz), 256, 0, m->stream>>>(m->xpk.as<uint32_t>(), m->nnz,
                                                                      m->lg.as<uint16_t>());
        HIPC(hipGetLastError());
    }
    m->lg_ready = true;
}

// both packed entry forms of the BF16 sketch passes
void ensure_bf16_entries(scxpca_matrix *m) {
    ensure_tpk(m);
    ensure_xpk(m);
}

// Column statistics of X~ (sum, sum_sq, nnz per gene) from the tile-CSC,
// cached on the matrix.
void ensure_stats(scxpca_matrix *m) {
    if (m->stats_ready) return;
    m->sum.ensure(sizeof(double) * m->n_vars);
    m->sq.ensure(sizeof(double) * m->n_vars);
    m->cnt.ensure(sizeof(uint32_t) * m->n_vars);
    if (m->n_obs == 0) {  // no rows: all statistics are zero
        HIPC(hipMemsetAsync(m->sum.p, 0, sizeof(double) * m->n_vars, m->stream));
        HIPC(hipMemsetAsync(m->sq.p, 0, sizeof(double) * m->n_vars, m->stream));
        HIPC(hipMemsetAsync(m->cnt.p, 0, sizeof(uint32_t) * m->n_vars, m->stream));
        m->stats_ready = true;
        return;
    }
    if (m->use_split) {
        // the tile-CSC build wrote per-tile partials: tiles -> ranges -> totals
        StageScope prof(m, m->stream, SCXPCA_STAGE_STATS);
        const int nrg = std::max(1, std::min(m->n_tiles, 64));
        const int tpr = (m->n_tiles + nrg - 1) / nrg;
        m->psum.ensure(sizeof(double) * nrg * m->n_vars);
        m->psq.ensure(sizeof(double) * nrg * m->n_vars);
        const unsigned gx = static_cast<unsigned>((m->n_vars + 255) / 256);
        k_stats_ranges<<<dim3(gx, nrg), 256, 0, m->stream>>>(m->tsum.as<double>(), m->tsq.as<double>(), m->n_tiles,
                                                             tpr, m->n_vars, m->psum.as<double>(), m->psq.as<double>());
        HIPC(hipGetLastError());
        k_stats_ranges<<<dim3(gx, 1), 256, 0, m->stream>>>(m->psum.as<double>(), m->psq.as<double>(), nrg, nrg,
                                                           m->n_vars, m->sum.as<double>(), m->sq.as<double>());
        HIPC(hipGetLastError());
        HIPC(hipMemcpyAsync(m->cnt.p, m->gcnt.p, sizeof(uint32_t) * m->n_vars, hipMemcpyDeviceToDevice, m->stream));
        m->stats_ready = true;
        return;
    }
    const int nch = static_cast<int>((m->n_vars + 127) / 128);  // k_col_stats4: 16 waves x 8 genes
    const int ntr = std::max(1, std::min(m->n_tiles, std::max(1, 512 / nch)));
    const int tiles_per_tr = (m->n_tiles + ntr - 1) / ntr;
    m->psum.ensure(sizeof(double) * ntr * m->n_vars);
    m->psq.ensure(sizeof(double) * ntr * m->n_vars);
    m->pcnt.ensure(sizeof(uint32_t) * ntr * m->n_vars);
    {
        StageScope prof(m, m->stream, SCXPCA_STAGE_STATS);
        k_col_stats4<<<ntr * nch, 1024, 0, m->stream>>>(m->indptr, m->tptr.as<uint32_t>(), m->trow.as<uint8_t>(),
                                                       m->tval.as<float>(), nch, tiles_per_tr,
                                                       m->n_tiles, m->n_obs, m->n_vars, m->psum.as<double>(),
                                                       m->psq.as<double>(), m->pcnt.as<uint32_t>());
    }
    HIPC(hipGetLastError());
    k_col_stats_reduce<<<blocks_for(m->n_vars), 256, 0, m->stream>>>(m->psum.as<double>(), m->psq.as<double>(),
                                                                     m->pcnt.as<uint32_t>(), ntr, m->n_vars,
                                                                     m->sum.as<double>(), m->sq.as<double>(),
                                                                     m->cnt.as<uint32_t>());
    HIPC(hipGetLastError());
    m->stats_ready = true;
}

// The dense PFlog transform Z (pyscx/src/accel/pflog.rs:506-533): every row
// filled with f32(b_i), then delta added at the stored entries (f32, as the
// reference); two launches, so the fill is complete before the adds, and
// each (row, column) is stored once, so the adds do not race.
__global__ void k_pflog_fill(const double *b, int64_t n_obs, int64_t n_vars, float *z) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = t; i < n_obs * n_vars; i += stride) z[i] = static_cast<float>(b[i / n_vars]);
}

__global__ void k_pflog_scatter(const int64_t *indptr, const uint32_t *idx, const float *delta, int64_t n_obs,
                                int64_t n_vars, float *z) {
    const int lane = threadIdx.x & 63;
    const int64_t wstride = static_cast<int64_t>(gridDim.x) * (blockDim.x / 64);
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * (blockDim.x / 64) + threadIdx.x / 64; r < n_obs; r += wstride)
        for (int64_t p = indptr[r] + lane; p < indptr[r + 1]; p += 64) z[r * n_vars + idx[p]] += delta[p];
}

// mu = sum / n (+ shift: the PFlog baseline mean, cpu.rs:838-843)
__global__ void k_mean(const double *sum, int64_t n_vars, double n, double shift, double *mu) {
    const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v < n_vars) mu[v] = sum[v] / n + shift;
}

// ---------------------------------------------------------------------------
// The PCA engine, specialised on the padded panel width KP.
// ---------------------------------------------------------------------------

template <int KP>
void device_omega(hipStream_t s, uint64_t rows, int k, uint64_t seed, float *panel, double *out64, OmegaScratch &sc,
                  int kpw = 0);

template <int KP>
struct Engine {
    // pivot floors (relative to the column's diagonal) of the rank-tolerant
    // Cholesky: the f32-Gram tall passes, the f64 Rayleigh-Ritz passes
    static constexpr double kStableFloor = 1e-6;
    static constexpr double kSmallFloor = 1e-13;
    // transposed SpMM k_spmm_t7 (genes owned per 16-lane group) and forward
    // SpMM k_spmm_fwd4 (rows owned per 16-lane group), over the KP / 64
    // 64-column halves of the panel
    static constexpr int GW = 8;                    // k_spmm_t7: genes per owner (Z in registers)
    static constexpr int P0 = 4;                    // k_spmm_t7: slot-0 prefetch, 16-entry components per group
    static constexpr int P1 = 3;                    // k_spmm_t7: slot-1 prefetch
    static constexpr int RC = 2;                    // k_spmm_t7: 16-entry components per ring load
    static constexpr int TW = 16;                   // k_spmm_t7: waves per block
    static constexpr int TU = 4 * TW;               // k_spmm_t7: gene owners (16-lane groups) per block
    static constexpr int CG = TU * GW;              // k_spmm_t7: genes per block
    static constexpr int64_t kFwd4MinRows = 256 * 1024;  // k_spmm_fwd4: one 1024-row block per CU
    static constexpr int F4PF = 2;                  // k_spmm_fwd4: ring depth, slots
    static constexpr int F4PC = 2;                  // k_spmm_fwd4: prefetched 16-entry components per window
    static constexpr int F4RC = 1;                  // k_spmm_fwd4: components per overflow load
    // the BF16 form's packed entries take one register per prefetched
    // component instead of two: deeper prefetch in the same budget
    static constexpr int P1B = 4, RCB = 4;          // k_spmm_t7<..., BFE>
    static constexpr int F4PCB = 3, F4RCB = 2;      // k_spmm_fwd4<..., BFE>
    scxpca_matrix *m;
    hipStream_t s;
    int k = 0;
    // small KP x KP f64 scratch layout inside m->small
    double *G, *R1, *R2, *R3, *Rt, *W64, *U, *V, *Acol, *Vcol, *sig, *mu, *sumq, *mc64, *Rfin, *cs64, *bq64;
    float *W32, *mc32, *cs32;
    // PFlog operator delta + b 1^T: the rank-1 baseline terms of both SpMMs
    const double *rowb = nullptr;
    int *status;
    int nrb_t = 0, nch_t = 0, ncol_blk = 0, gram_waves = 0, gram_blocks = 0, tiles_per_tr = 0;
    const uint32_t *gperm = nullptr;  // wave-slot gene order of the transposed SpMM
    int64_t rows_per_wave = 0, rows_per_colblk = 0, rows_per_gblk = 0;

    // Grid of the transposed SpMM: nch gene chunks x nrb tile ranges of tpr
    // tiles (about 1024 / TW blocks in total for 256 CUs, nrb a multiple of 8
    // for the XCD-aware block order).
    static void t_geometry(const scxpca_matrix *mm, int &nch, int &nrb, int &tpr) {
        nch = static_cast<int>((mm->n_vars + CG - 1) / CG);
        int ntr = std::max(1, (256 * 16 / TW) / std::max(nch, 1));
        ntr = std::min(ntr, std::max(1, mm->n_tiles));
        nrb = (ntr + 7) / 8 * 8;
        tpr = (mm->n_tiles + nrb - 1) / nrb;
    }

    explicit Engine(scxpca_matrix *mm, int kk) : m(mm), s(mm->stream), k(kk) {
        const int64_t n = m->n_obs, nv = m->n_vars;
        t_geometry(m, nch_t, nrb_t, tiles_per_tr);
        gperm = ensure_gene_order(m, GW, nch_t, TU);
        gram_waves = static_cast<int>(std::max<int64_t>(4, std::min<int64_t>((n + 255) / 256, KP == 64 ? 1024 : 512)));
        gram_waves = (gram_waves + 3) / 4 * 4;
        rows_per_wave = (n + gram_waves - 1) / gram_waves;
        rows_per_wave += rows_per_wave & 1;
        gram_blocks = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>((n + 511) / 512, KP == 64 ? 1024 : 512)));
        rows_per_gblk = (n + gram_blocks - 1) / gram_blocks;
        rows_per_gblk = (rows_per_gblk + 15) / 16 * 16;
        ncol_blk = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>((std::max(n, nv) + 1023) / 1024, 1024)));
        rows_per_colblk = (std::max(n, nv) + ncol_blk - 1) / ncol_blk;

        const size_t panel_rows = static_cast<size_t>(std::max<int64_t>(n, 1));
        m->y.ensure(sizeof(float) * panel_rows * KP);
        m->q.ensure(sizeof(float) * panel_rows * KP);
        m->panel.ensure(sizeof(float) * std::max<int64_t>(nv, 1) * KP);
        m->panel2.ensure(sizeof(float) * std::max<int64_t>(nv, 1) * KP);
        m->panel64.ensure(sizeof(double) * std::max<int64_t>(nv, 1) * KP);
        m->b64.ensure(sizeof(double) * std::max<int64_t>(nv, 1) * KP);
        m->t64a.ensure(sizeof(double) * std::max<int64_t>(nv, 1) * KP);
        m->t64b.ensure(sizeof(double) * std::max<int64_t>(nv, 1) * KP);
        m->tpart.ensure(sizeof(float) * static_cast<size_t>(nrb_t) * std::max<int64_t>(nv, 1) * KP);
        const int nb64 = gram64_blocks();
        m->gpart.ensure(sizeof(double) * static_cast<size_t>(std::max(std::max(gram_waves, gram_blocks), nb64)) * KP * KP);
        m->cpart.ensure(sizeof(double) * static_cast<size_t>(std::max(kMaxColBlocks, nrb_t * nch_t)) * KP);
        constexpr size_t KK = static_cast<size_t>(KP) * KP;
        m->small.ensure(sizeof(double) * (KK * 11 + 6 * KP + std::max<int64_t>(nv, 1)) + sizeof(float) * (KK + 2 * KP));
        double *d = m->small.as<double>();
        G = d; R1 = G + KK; R2 = R1 + KK; R3 = R2 + KK; Rt = R3 + KK; W64 = Rt + KK; U = W64 + KK; V = U + KK;
        Acol = V + KK; Vcol = Acol + KK; Rfin = Vcol + KK;  // Rfin: R^-1 of the final basis's last pass
        sig = Rfin + KK; sumq = sig + KP; mc64 = sumq + KP; cs64 = mc64 + KP + KP; bq64 = cs64 + KP;
        mu = bq64 + KP;
        W32 = reinterpret_cast<float *>(mu + std::max<int64_t>(nv, 1));
        mc32 = W32 + KK;
        cs32 = mc32 + KP;
        if (m->pflog && m->n_obs > 0) rowb = m->pfb.as<double>();
        m->status.ensure(sizeof(int));
        status = m->status.as<int>();
    }

    int gram64_blocks() const {
        return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>((m->n_vars + 31) / 32, 256)));
    }

    void check() { HIPC(hipGetLastError()); }

    // mean(b) = sum(b) / max(n, 1) (cpu.rs:838), 0 for a plain matrix
    double baseline_mean() const {
        return m->pflog ? m->pf_bsum / std::max(static_cast<double>(m->n_obs), 1.0) : 0.0;
    }

    // out = w^T A over the rows of a tall f32 panel A (nrows x KP), f64;
    // w == null: plain column sums. cpart: kMaxColBlocks x KP f64 partials.
    static void colsum_into(hipStream_t st, const float *A, const double *w, int64_t nrows, double *cpart,
                            double *out64, float *out32) {
        const int nb = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>((nrows + 63) / 64, kMaxColBlocks)));
        const int64_t rpb = (nrows + nb - 1) / nb;
        k_wcolsum4<KP><<<nb, 256, 0, st>>>(A, w, nrows, rpb, cpart);
        HIPC(hipGetLastError());
        k_colsum_reduce<KP><<<KP, 256, 0, st>>>(cpart, nb, out64, out32);
        HIPC(hipGetLastError());
    }
    void weighted_colsum(const float *A, const double *w, int64_t nrows, double *out64, float *out32) {
        colsum_into(s, A, w, nrows, m->cpart.as<double>(), out64, out32);
    }

    // mc = mu^T panel (f32 for the forward epilogue)
    void mean_correction(const float *panel) {
        const int64_t nv = m->n_vars;
        const int nb = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>((nv + 15) / 16, kMaxColBlocks)));
        const int64_t rpb = (nv + nb - 1) / nb;
        k_colsum<KP, float><<<nb, 256, 0, s>>>(panel, mu, nv, rpb, m->cpart.as<double>());
        check();
        k_colsum_reduce<KP><<<KP, 256, 0, s>>>(m->cpart.as<double>(), nb, mc64, mc32);
        check();
    }

    // bf16: the BF16 sketch form, x~ rounded to bf16 (the packed entries xpk
    // on the split layout)
    template <bool OFF, bool BF>
    void launch_fwd(const float *panel, const float *mcp, float *Y) {
        // rows owned per 16-lane group (64 rows per wave, 1024 per block) once
        // the grid fills the 256 CUs; below that the row-parallel kernel (one
        // wave per row, the panel from L2) has the parallelism
        if (m->use_split && m->n_obs >= kFwd4MinRows) {
            constexpr int NWF = 16;
            constexpr int CG4 = 512;  // genes per 128 KB chunk of a 64-column half
            const int nch4 = static_cast<int>((m->n_vars + CG4 - 1) / CG4);
            const unsigned grid = static_cast<unsigned>((m->n_obs + NWF * 64 - 1) / (NWF * 64));
            k_spmm_fwd4<64, KP, 16, NWF, CG4, F4PF, BF ? F4PCB : F4PC, BF ? F4RCB : F4RC, OFF, BF>
                <<<dim3(grid, KP / 64), NWF * 64, 0, s>>>(
                m->indptr, m->lg.as<uint16_t>(), m->xt.as<float>(), BF ? m->xpk.as<uint32_t>() : nullptr, panel,
                mcp, Y, m->spl.as<uint32_t>(), m->nsplit, CG4 / kCB, nch4, m->n_obs, m->n_vars, m->nnz, rowb, cs32);
        } else {
            const unsigned grid = std::min<unsigned>(blocks_for(m->n_obs, 4), 16384u);
            k_spmm_fwd<KP, OFF, BF><<<grid, 256, 0, s>>>(m->indptr, m->idx, m->xt.as<float>(), panel, mcp, Y,
                                                           m->n_obs, rowb, cs32);
        }
    }

    // Y = X~ panel - 1 mc^T (+ b (1^T panel)^T for a PFlog matrix, cpu.rs:617-647)
    void spmm_fwd(const float *panel, bool centered, bool bf16, float *Y) {
        if (m->n_obs == 0) return;
        if (!bf16) ensure_lg(m);
        if (rowb) weighted_colsum(panel, nullptr, m->n_vars, cs64, cs32);
        StageScope prof(m, s, bf16 ? SCXPCA_STAGE_SPMM_FWD_BF16 : SCXPCA_STAGE_SPMM_FWD);
        const float *mcp = centered ? mc32 : nullptr;
        if (rowb) {
            if (bf16) launch_fwd<true, true>(panel, mcp, Y);
            else launch_fwd<true, false>(panel, mcp, Y);
        } else {
            if (bf16) launch_fwd<false, true>(panel, mcp, Y);
            else launch_fwd<false, false>(panel, mcp, Y);
        }
        check();
    }

    // Scratch of one transposed SpMM (transposed() below): the tile-range
    // partials of Z (nrb x n_vars x KP f32), of 1^T Q (nrb x nch x KP f64),
    // the column-sum partials (kMaxColBlocks x KP f64) and 1^T Q, b^T Q (KP f64).
    struct TScratch {
        float *tpart;
        double *qsp, *cpart, *sumq, *bq;
    };

    // B64 (and/or B32) = X~^T Q - mu sumq^T (+ 1 (b^T Q) for a PFlog matrix,
    // rowb its baseline; cpu.rs:653-687) over the buffers in sc. mu null: no
    // centering. Every partial is written by the launch (the grid covers all
    // nrb x nch slots), so sc needs no clearing between calls. bf16: the BF16
    // sketch form (x~ rounded to bf16, the packed entries tpk).
    static void transposed(scxpca_matrix *mm, hipStream_t st, const uint32_t *gp, int nch, int nrb, int tpr,
                           const float *Q, const double *mu, const double *rowb, const TScratch &sc,
                           double *B64, float *B32, bool bf16 = false) {
        const int64_t nv = mm->n_vars;
        double *qsp = mu ? sc.qsp : nullptr;
        if (mm->n_obs > 0) {
            const dim3 grid(static_cast<unsigned>(nrb * nch), KP / 64);
            StageScope prof(mm, st, bf16 ? SCXPCA_STAGE_SPMM_T_BF16 : SCXPCA_STAGE_SPMM_T);
            if (bf16)
                k_spmm_t7<64, KP, GW, P0, P1B, RCB, TW, true><<<grid, TW * 64, 0, st>>>(
                    mm->indptr, mm->tptr.as<uint32_t>(), nullptr, nullptr, mm->tpk.as<uint32_t>(),
                    mm->nnz + kEntryPad, Q, gp, nch, nrb, tpr, mm->n_obs, nv, mm->n_tiles, sc.tpart, qsp);
            else
                k_spmm_t7<64, KP, GW, P0, P1, RC, TW, false><<<grid, TW * 64, 0, st>>>(
                    mm->indptr, mm->tptr.as<uint32_t>(), mm->trow.as<uint8_t>(), mm->tval.as<float>(), nullptr,
                    mm->nnz + kEntryPad, Q, gp, nch, nrb, tpr, mm->n_obs, nv, mm->n_tiles, sc.tpart, qsp);
            HIPC(hipGetLastError());
        } else {  // no rows: zero partials (a fill kernel, no memset node)
            k_zero_words<<<blocks_for(static_cast<int64_t>(nrb) * nv * KP), 256, 0, st>>>(
                reinterpret_cast<uint32_t *>(sc.tpart), static_cast<int64_t>(nrb) * nv * KP);
            HIPC(hipGetLastError());
            if (qsp) {
                k_zero_words<<<blocks_for(2 * static_cast<int64_t>(nrb) * nch * KP), 256, 0, st>>>(
                    reinterpret_cast<uint32_t *>(qsp), 2 * static_cast<int64_t>(nrb) * nch * KP);
                HIPC(hipGetLastError());
            }
        }
        if (mu) {
            k_colsum_reduce<KP><<<KP, 256, 0, st>>>(qsp, nrb * nch, sc.sumq, nullptr);
            HIPC(hipGetLastError());
        }
        if (rowb) colsum_into(st, Q, rowb, mm->n_obs, sc.cpart, sc.bq, nullptr);
        StageScope prof(mm, st, SCXPCA_STAGE_SPMM_T_REDUCE);
        k_spmm_t_reduce<KP><<<blocks_for(nv * KP), 256, 0, st>>>(sc.tpart, nrb, nv, mu, sc.sumq, B64, B32,
                                                                  rowb ? sc.bq : nullptr);
        HIPC(hipGetLastError());
    }

    // B64 (and/or B32) = X~^T Q - mu sumq^T (+ 1 (b^T Q) for a PFlog matrix)
    // on the matrix's own scratch
    void spmm_t(const float *Q, bool centered, bool bf16, double *B64, float *B32) {
        // 1^T Q partials in cpart (nrb_t x nch_t x KP), then b^T Q's in the
        // same buffer: colsum_into runs after the 1^T Q reduction
        const TScratch sc{m->tpart.as<float>(), m->cpart.as<double>(), m->cpart.as<double>(), sumq, bq64};
        transposed(m, s, gperm, nch_t, nrb_t, tiles_per_tr, Q, centered ? mu : nullptr, rowb, sc, B64, B32, bf16);
    }

    // One Cholesky step on the current Gram G: W32/W64 = R^-1, R -> Rdst.
    // floor_tau: 0 strict (non-SPD is an error), > 0 the rank-tolerant pivot floor
    void chol(double shift, double *Rdst, double floor_tau) {
        StageScope prof(m, s, SCXPCA_STAGE_SOLVE);
        if constexpr (KP == 64) {
            if (floor_tau > 0.0)
                k_chol_w<true><<<1, 256, 0, s>>>(G, k, shift, Rdst, W64, W32, status, floor_tau);
            else
                k_chol_w<false><<<1, 256, 0, s>>>(G, k, shift, Rdst, W64, W32, status, 0.0);
        } else {
            k_chol2<KP><<<1, 256, 0, s>>>(G, k, shift, Rdst, W64, W32, status, floor_tau);
        }
        check();
    }

    // out[:, :out_cols] = Y W32 for a tall f32 panel Y (n x KP); upper: W32 is
    // upper triangular (an inverse Cholesky factor).
    void apply(const float *Y, float *out, int64_t n, int out_cols, int64_t out_ld, bool upper) {
        if constexpr (KP == 64) {
            const unsigned grid = std::max(std::min<unsigned>(blocks_for((n + 15) / 16, 4), 2048u), 1u);
            if (upper)
                k_apply4<KP, true><<<grid, 256, 0, s>>>(Y, W32, out, n, out_cols, out_ld);
            else
                k_apply4<KP, false><<<grid, 256, 0, s>>>(Y, W32, out, n, out_cols, out_ld);
        } else {
            const unsigned grid = std::min<unsigned>(blocks_for((n + 31) / 32, 4), 8192u);
            k_apply2<KP><<<std::max(grid, 1u), 256, 0, s>>>(Y, W32, out, n, out_cols, out_ld);
        }
        check();
    }

    // Tall f32 panel CholeskyQR: CQR2, or shifted CQR3 (stable) for Householder mode.
    // Returns the buffer that holds Q (one of src / tmp).
    // With final == false (a basis that only feeds the next product of the
    // power iteration) one shifted pass: Q = Y R^-1 spans Y exactly for any
    // nonsingular R, and the shift keeps the f32 Gram positive definite; the
    // final basis gets the requested method. The stable mode floors pivots
    // (kStableFloor) instead of failing, so an exactly rank-deficient Y (k
    // above the rank, where the reference's Householder QR still returns an
    // orthonormal basis) gets one: the floored directions are normalised by
    // the following passes. CholeskyQR2 keeps the strict non-SPD error.
    // With apply_last == false the last pass forms no Q: W32 holds its R^-1
    // and src is returned unchanged, for a caller that applies R^-1 to a
    // product linear in src.
    float *cqr_tall(float *src, float *tmp, int qr_method, int64_t n = -1, bool final = true, bool apply_last = true) {
        if (n < 0) n = m->n_obs;
        const int passes = !final ? 1 : (qr_method == SCXPCA_QR_HOUSEHOLDER ? 3 : 2);
        const int gblocks = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>((n + 127) / 128, gram_blocks)));
        int64_t rpg = (n + gblocks - 1) / gblocks;
        rpg = (rpg + 15) / 16 * 16;
        for (int pass = 0; pass < passes; ++pass) {
            {
                StageScope prof(m, s, SCXPCA_STAGE_GRAM);
                if constexpr (KP == 64)
                    k_gram4<KP><<<gblocks, 256, 0, s>>>(src, n, rpg, m->gpart.as<double>());
                else
                    k_gram2<KP><<<gblocks, 256, 0, s>>>(src, n, rpg, m->gpart.as<double>());
                check();
                if constexpr (KP == 64)
                    k_sum_parts16<KP><<<(KP * KP + 15) / 16, 256, 0, s>>>(m->gpart.as<double>(), gblocks, G);
                else
                    k_gram_reduce<KP><<<(KP * KP + 63) / 64, 256, 0, s>>>(m->gpart.as<double>(), gblocks, G);
            }
            check();
            const double shift = !final ? 1e-7 : ((qr_method == SCXPCA_QR_HOUSEHOLDER && pass == 0) ? 1e-5 : 0.0);
            chol(shift, R1, qr_method == SCXPCA_QR_HOUSEHOLDER ? kStableFloor : 0.0);
            if (!apply_last && pass + 1 == passes) break;
            {
                StageScope prof(m, s, SCXPCA_STAGE_APPLY);
                apply(src, tmp, n, KP, KP, true);
            }
            check();
            std::swap(src, tmp);
        }
        return src;
    }

    // Small f64 CholeskyQR3 (shifted first pass) of A (n_vars x KP) in place
    // into Qout; R = R3 R2 R1 written to Rt. Pivots are floored (kSmallFloor),
    // never an error: B of rank below k (the reference's thin SVD accepts it)
    // still gives an orthonormal Q_B with A = Q_B R.
    void cqr_small64(const double *A, double *Qout) {
        const int64_t nv = m->n_vars;
        const int nb = gram64_blocks();
        const int64_t rpb = (nv + nb - 1) / nb;
        const double *src = A;
        double *bufs[3] = {m->t64a.as<double>(), m->t64b.as<double>(), Qout};
        double *Rs[3] = {R1, R2, R3};
        StageScope prof(m, s, SCXPCA_STAGE_SMALL);
        for (int pass = 0; pass < 3; ++pass) {
            k_gram64<KP><<<nb, 256, 0, s>>>(src, nv, rpb, m->gpart.as<double>());
            check();
            k_sum_parts16<KP><<<(KP * KP + 15) / 16, 256, 0, s>>>(m->gpart.as<double>(), nb, G);
            check();
            chol(pass == 0 ? 1e-13 : 0.0, Rs[pass], kSmallFloor);
            k_apply64<KP><<<blocks_for(nv * KP), 256, 0, s>>>(src, W64, bufs[pass], nullptr, nv);
            check();
            src = bufs[pass];
        }
        // Rt = R3 * (R2 * R1)
        k_small_gemm64<KP><<<KP / 16, 256, 0, s>>>(R2, R1, Acol);
        check();
        k_small_gemm64<KP><<<KP / 16, 256, 0, s>>>(R3, Acol, Rt);
        check();
    }

    // Total variance with the cancellation guard (guarded_total_variance_*,
    // cpu.rs:1029-1092) from the column sums and sums of squares; hmu gets
    // the column means.
    // stored_ss() returns sum over the stored entries of (x~ - mu)^2 (the
    // stable recompute's device pass)
    template <class StoredSS>
    double total_variance(bool centered, const std::vector<double> &hsum, const std::vector<double> &hsq,
                          std::vector<double> &hmu, int &recomputed, StoredSS &&stored_ss) {
        const int64_t n = m->n_obs, nv = m->n_vars;
        hmu.assign(nv, 0.0);
        for (int64_t v = 0; v < nv; ++v) hmu[v] = hsum[v] / static_cast<double>(n);
        const double denom = std::max(static_cast<double>(n) - 1.0, 1.0);
        double tv = 0.0;
        recomputed = 0;
        if (m->pflog) {
            // pflog_total_variance (cpu.rs:702-769): S_dd - D B2 (uncentered),
            // S_dd - D B2 - n sum_j mu_j^2 with mu_j = colmean_j + mean(b)
            // (centered); never recomputed, clamped at 0
            double s_dd = 0.0;
            for (int64_t v = 0; v < nv; ++v) s_dd += hsq[v];
            double tss = s_dd - static_cast<double>(nv) * m->pf_b2;
            if (centered) {
                const double bm = baseline_mean();
                double m2 = 0.0;
                for (int64_t v = 0; v < nv; ++v) {
                    hmu[v] += bm;
                    m2 += hmu[v] * hmu[v];
                }
                tss -= static_cast<double>(n) * m2;
            }
            return std::max(tss / denom, 0.0);
        }
        if (!centered) {
            for (int64_t v = 0; v < nv; ++v) tv += hsq[v];
            tv /= denom;
        } else {
            double tot = 0.0, sumsq = 0.0, msq = 0.0;
            for (int64_t v = 0; v < nv; ++v) tot += hsq[v] - static_cast<double>(n) * hmu[v] * hmu[v];
            tv = tot / denom;
            for (int64_t v = 0; v < nv; ++v) sumsq += hsq[v];
            for (int64_t v = 0; v < nv; ++v) msq += hmu[v] * hmu[v];
            msq *= static_cast<double>(n);
            if (sumsq > 0.0 && (sumsq - msq) <= 1e-7 * sumsq) {
                // cancelled: stable centered recompute on the device
                recomputed = 1;
                double t = stored_ss();
                std::vector<uint32_t> hcnt(nv);
                HIPC(hipMemcpyAsync(hcnt.data(), m->cnt.p, sizeof(uint32_t) * nv, hipMemcpyDeviceToHost, s));
                HIPC(hipStreamSynchronize(s));
                for (int64_t v = 0; v < nv; ++v) {
                    const double zeros = static_cast<double>(std::max<int64_t>(n - static_cast<int64_t>(hcnt[v]), 0));
                    t += zeros * hmu[v] * hmu[v];
                }
                tv = t / denom;
            }
        }
        return tv;
    }

    // sum over the stored entries of (x~ - mu)^2 on a resident matrix mm (f64
    // block partials in a fixed order)
    double stored_centered_ss(scxpca_matrix *mm, const double *mu_dev) {
        const int nb = 1024;
        DevBuf part;
        part.ensure(sizeof(double) * nb);
        k_centered_ss<<<nb, 256, 0, s>>>(mm->indptr, mm->idx, mm->xt.as<float>(), mu_dev, mm->n_obs,
                                         part.as<double>());
        check();
        std::vector<double> hp(nb);
        HIPC(hipMemcpyAsync(hp.data(), part.p, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
        double t = 0.0;
        for (int i = 0; i < nb; ++i) t += hp[i];
        return t;
    }

    // The operator of a resident matrix: its statistics, SpMMs and stable
    // recompute on the matrix itself.
    struct ResidentOps {
        Engine *e;
        void stats() { ensure_stats(e->m); }
        void fwd(const float *panel, bool centered, bool bf16, float *Y) { e->spmm_fwd(panel, centered, bf16, Y); }
        void t(const float *Q, bool centered, bool bf16, double *B64, float *B32) {
            e->spmm_t(Q, centered, bf16, B64, B32);
        }
        double stored_ss() { return e->stored_centered_ss(e->m, e->mu); }
    };

    void run(const scxpca_params *p, scxpca_result *out) {
        ResidentOps ops{this};
        run_with(ops, p, out);
    }

    // The power loop, Rayleigh-Ritz and results over an operator (ResidentOps,
    // or the streamed shards of a StreamOps).
    template <class Ops>
    void run_with(Ops &ops, const scxpca_params *p, scxpca_result *out) {
        const int64_t n = m->n_obs, nv = m->n_vars;
        const bool centered = p->zero_center != 0;
        const bool sketch_bf16 = p->precision == SCXPCA_PRECISION_BF16_SKETCH;
        const int nc = static_cast<int>(p->n_comps);
        HIPC(hipMemsetAsync(status, 0, sizeof(int), s));
        if (sketch_bf16) ensure_bf16_entries(m);
        ops.stats();
        k_mean<<<blocks_for(nv), 256, 0, s>>>(m->sum.as<double>(), nv, static_cast<double>(n), baseline_mean(), mu);
        check();
        // Omega -> f32 panel, generated on the device
        device_omega<KP>(s, static_cast<uint64_t>(nv), k, p->seed, m->panel.as<float>(), nullptr, m->om);
        float *Y = m->y.as<float>(), *T = m->q.as<float>();
        if (centered) mean_correction(m->panel.as<float>());
        ops.fwd(m->panel.as<float>(), centered, sketch_bf16, Y);
        const int64_t n_it = p->n_power_iterations;
        // The last CholeskyQR pass of every tall basis is never applied: the
        // panel stays as it is with R^-1 in W32 / W64, and since the
        // transposed product, its centering term and the PFlog term are
        // linear in the panel, X~^T (Y R^-1) = (X~^T Y) R^-1 is applied to the
        // n_vars x KP result instead of the n_obs x KP panel (for the final
        // basis also to the embedding weights).
        float *Q = cqr_tall(Y, T, p->qr_method, -1, n_it == 0, false);
        float *Qother = (Q == Y) ? T : Y;
        for (int64_t it = 0; it < n_it; ++it) {
            // B (f32 n_vars x KP) -> basis by the tall MFMA CholeskyQR
            ops.t(Q, centered, sketch_bf16, nullptr, m->panel2.as<float>());
            {
                StageScope prof(m, s, SCXPCA_STAGE_APPLY);
                apply(m->panel2.as<float>(), m->panel.as<float>(), nv, KP, KP, true);
            }
            float *qb = cqr_tall(m->panel.as<float>(), m->panel2.as<float>(), p->qr_method, nv, false);
            if (centered) mean_correction(qb);
            ops.fwd(qb, centered, sketch_bf16, Qother);
            const bool last = it + 1 == n_it;
            float *Qn = cqr_tall(Qother, Q, p->qr_method, -1, last, false);
            Qother = (Qn == Q) ? Qother : Q;
            Q = Qn;
        }
        // Final B pass in FP32 (also in the BF16 sketch form), then B R^-1 in
        // f64; R^-1 kept (Rfin) for the embedding weights
        ops.t(Q, centered, false, m->b64.as<double>(), nullptr);
        {
            StageScope prof(m, s, SCXPCA_STAGE_APPLY);
            k_apply64<KP><<<blocks_for(nv * KP), 256, 0, s>>>(m->b64.as<double>(), W64, m->t64b.as<double>(), nullptr,
                                                           nv);
            check();
            HIPC(hipMemcpyAsync(Rfin, W64, sizeof(double) * KP * KP, hipMemcpyDeviceToDevice, s));
        }
        // Rayleigh-Ritz: B = Q_B Rt, Rt = U_R S V^T
        cqr_small64(m->t64b.as<double>(), m->panel64.as<double>());
        {
            StageScope prof(m, s, SCXPCA_STAGE_SOLVE);
            if constexpr (KP == 64)
                k_jacobi_w<KP><<<1, 512, 0, s>>>(Rt, k, U, V, sig, status);
            else
                k_jacobi_lds<KP><<<1, 256, 0, s>>>(Rt, k, Vcol, U, V, sig, status);
        }
        check();
        // Uhat = Q_B U_R (n_vars x KP, f64) -> components
        k_apply64<KP><<<blocks_for(nv * KP), 256, 0, s>>>(m->panel64.as<double>(), U, m->t64a.as<double>(), nullptr, nv);
        check();
        float *comps_dev = nullptr;
        DevBuf comps_tmp, emb_tmp;
        if (out->on_device) {
            comps_dev = out->components;
        } else {
            comps_tmp.ensure(sizeof(float) * std::max<int64_t>(nv * nc, 1));
            comps_dev = comps_tmp.as<float>();
        }
        k_components<KP><<<blocks_for(nv * nc), 256, 0, s>>>(m->t64a.as<double>(), nc, nv, comps_dev);
        check();
        k_small_gemm64<KP><<<KP / 16, 256, 0, s>>>(Rfin, V, Acol);  // Acol is free after the Rayleigh-Ritz QR
        check();
        k_emb_weights<KP><<<1, 256, 0, s>>>(Acol, sig, nc, W32);
        check();
        float *emb_dev = nullptr;
        if (out->on_device) {
            emb_dev = out->embeddings;
        } else {
            emb_dev = Qother;  // n_obs x KP scratch, embeddings written with ld = nc
        }
        apply(Q, emb_dev, n, nc, nc, false);
        check();
        // results back
        // sig, column sums, sums of squares and status: one packed copy into pinned memory
        const int64_t nres = KP + 2 * nv + 1;
        m->dres.ensure(sizeof(double) * nres);
        m->hres.ensure(sizeof(double) * nres);  // the previous call's copy was drained by its sync
        k_pack_results<<<blocks_for(nres), 256, 0, s>>>(sig, KP, m->sum.as<double>(), m->sq.as<double>(), nv, status,
                                                       m->dres.as<double>());
        check();
        HIPC(hipMemcpyAsync(m->hres.p, m->dres.p, sizeof(double) * nres, hipMemcpyDeviceToHost, s));
        if (!out->on_device) {
            HIPC(hipMemcpyAsync(out->embeddings, emb_dev, sizeof(float) * n * nc, hipMemcpyDeviceToHost, s));
            HIPC(hipMemcpyAsync(out->components, comps_dev, sizeof(float) * nv * nc, hipMemcpyDeviceToHost, s));
        }
        HIPC(hipStreamSynchronize(s));
        const double *hres = m->hres.as<double>();
        const std::vector<double> hsig(hres, hres + KP), hsum(hres + KP, hres + KP + nv),
            hsq(hres + KP + nv, hres + KP + 2 * nv);
        const int hstatus = static_cast<int>(hres[KP + 2 * nv]);
        if (hstatus & 1) {
            if (p->qr_method == SCXPCA_QR_CHOLESKY)
                throw Error(SCXPCA_ERR_LINALG,
                            "CholeskyQR2 failed: non-SPD; retry with qr_method='householder'");
            throw Error(SCXPCA_ERR_LINALG, "shifted CholeskyQR failed: Gram matrix not positive definite");
        }
        if (hstatus & 2) throw Error(SCXPCA_ERR_LINALG, "SVD failed: Jacobi sweeps did not converge");
        // total variance with the cancellation guard (cpu.rs:1029-1092)
        std::vector<double> hmu;
        int recomputed = 0;
        const double tv = total_variance(centered, hsum, hsq, hmu, recomputed, [&] { return ops.stored_ss(); });
        const double denom = std::max(static_cast<double>(n) - 1.0, 1.0);
        for (int j = 0; j < nc; ++j) {
            const double var = hsig[j] * hsig[j] / denom;
            if (out->variance) out->variance[j] = var;
            if (out->variance_ratio) out->variance_ratio[j] = tv > 0.0 ? var / tv : 0.0;
        }
        if (out->sigma)
            for (int j = 0; j < k; ++j) out->sigma[j] = hsig[j];
        if (out->mean && centered) std::memcpy(out->mean, hmu.data(), sizeof(double) * nv);
        out->k = k;
        out->total_variance = tv;
        out->tv_recomputed = recomputed;
    }

    // Covariance route (covariance_pca, cpu.rs:1663-1854): G = X~^T X~ by
    // k_cov_gram, cov = (G - n mu mu^T) / (n - 1), the symmetric EVD
    // (rocSOLVER dsyevd, f64), the top n_comps eigenpairs, and the embeddings
    // (X~ - mu) V by the forward SpMM. k = n_comps (no oversampling).
    void run_covariance(const scxpca_params *p, scxpca_result *out) {
        const int64_t n = m->n_obs, nv = m->n_vars;
        const bool centered = p->zero_center != 0;
        const int nc = static_cast<int>(p->n_comps);
        HIPC(hipMemsetAsync(status, 0, sizeof(int), s));
        ensure_stats(m);
        k_mean<<<blocks_for(nv), 256, 0, s>>>(m->sum.as<double>(), nv, static_cast<double>(n), 0.0, mu);
        check();
        // block pairs (I <= J) of 64-gene blocks and the pair index of each
        const int nblk = static_cast<int>((nv + kCovGB - 1) / kCovGB);
        const int npairs = nblk * (nblk + 1) / 2;
        std::vector<int> tab(2 * static_cast<size_t>(npairs) + static_cast<size_t>(nblk) * nblk, -1);
        {
            int q = 0;
            for (int I = 0; I < nblk; ++I)
                for (int J = I; J < nblk; ++J, ++q) {
                    tab[2 * q] = I;
                    tab[2 * q + 1] = J;
                    tab[2 * static_cast<size_t>(npairs) + static_cast<size_t>(I) * nblk + J] = q;
                }
        }
        m->covtab.ensure(sizeof(int) * tab.size());
        HIPC(hipMemcpyAsync(m->covtab.p, tab.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice, s));
        const int2 *pairs = m->covtab.as<int2>();
        const int *pob = m->covtab.as<int>() + 2 * npairs;
        // tile ranges: enough blocks to fill the GPU several times over
        int ntr = std::max(1, std::min(m->n_tiles, (2048 + npairs - 1) / npairs));
        const int tpr = (m->n_tiles + ntr - 1) / ntr;
        ntr = (m->n_tiles + tpr - 1) / tpr;
        m->covpart.ensure(sizeof(double) * static_cast<size_t>(ntr) * npairs * kCovGB * kCovGB);
        m->cov.ensure(sizeof(double) * static_cast<size_t>(nv) * nv);
        {
            StageScope prof(m, s, SCXPCA_STAGE_GRAM);
            k_cov_gram<<<dim3(npairs, ntr), 1024, 0, s>>>(m->indptr, m->tptr.as<uint32_t>(), m->trow.as<uint8_t>(),
                                                         m->tval.as<float>(), pairs, npairs, tpr, m->n_tiles, n, nv,
                                                         m->covpart.as<double>());
            check();
            k_cov_finalize<<<blocks_for(nv * nv), 256, 0, s>>>(m->covpart.as<double>(), npairs, ntr, pob, nblk,
                                                                centered ? mu : nullptr, n, nv,
                                                                m->cov.as<double>());
            check();
        }
        // all eigenpairs, ascending; eigenvectors overwrite cov (column-major)
        m->evw.ensure(sizeof(double) * nv);
        m->eve.ensure(sizeof(double) * nv);
        m->evinfo.ensure(sizeof(rocblas_int));
        if (!m->rb) {
            if (rocblas_create_handle(&m->rb) != rocblas_status_success)
                throw Error(SCXPCA_ERR_DEVICE, "rocblas_create_handle failed");
        }
        if (rocblas_set_stream(m->rb, s) != rocblas_status_success) throw Error(SCXPCA_ERR_DEVICE, "rocblas_set_stream failed");
        {
            StageScope prof(m, s, SCXPCA_STAGE_SOLVE);
            const rocblas_int rn = static_cast<rocblas_int>(nv);
            const rocblas_status st = rocsolver_dsyevd(m->rb, rocblas_evect_original, rocblas_fill_lower, rn,
                                                       m->cov.as<double>(), rn, m->evw.as<double>(),
                                                       m->eve.as<double>(), m->evinfo.as<rocblas_int>());
            if (st != rocblas_status_success)
                throw Error(SCXPCA_ERR_DEVICE, "rocsolver_dsyevd failed (status " + std::to_string(static_cast<int>(st)) + ")");
        }
        // top nc: variance, components, the f32 panel V (nv x KP)
        float *comps_dev = nullptr;
        DevBuf comps_tmp;
        if (out->on_device) {
            comps_dev = out->components;
        } else {
            comps_tmp.ensure(sizeof(float) * std::max<int64_t>(nv * nc, 1));
            comps_dev = comps_tmp.as<float>();
        }
        m->covvar.ensure(sizeof(double) * std::max(nc, 1));
        // embeddings (X~ - mu) V, KP components at a time (more than KP
        // components: one forward pass per KP-wide slice of V)
        float *Y = m->y.as<float>();
        DevBuf emb_tmp;
        float *emb_dev = nullptr;
        if (out->on_device) {
            emb_dev = out->embeddings;
        } else if (nc <= KP) {
            emb_dev = m->q.as<float>();  // n x KP scratch holds the packed n x nc embeddings
        } else {
            emb_tmp.ensure(sizeof(float) * n * nc);
            emb_dev = emb_tmp.as<float>();
        }
        for (int c0 = 0; c0 < nc; c0 += KP) {
            const int ncb = std::min(KP, nc - c0);
            k_cov_topk<KP><<<blocks_for(nv * KP), 256, 0, s>>>(m->cov.as<double>(), m->evw.as<double>(), nv, nv, nc,
                                                               comps_dev, m->panel.as<float>(),
                                                               m->covvar.as<double>(), c0);
            check();
            if (centered) mean_correction(m->panel.as<float>());
            spmm_fwd(m->panel.as<float>(), centered, false, Y);
            k_unpack_cols<KP><<<blocks_for(n * ncb), 256, 0, s>>>(Y, n, ncb, emb_dev, nc, c0);
            check();
        }
        // results back
        std::vector<double> hvar(nc), hsum(nv), hsq(nv);
        rocblas_int hinfo = 0;
        HIPC(hipMemcpyAsync(hvar.data(), m->covvar.p, sizeof(double) * nc, hipMemcpyDeviceToHost, s));
        HIPC(hipMemcpyAsync(hsum.data(), m->sum.p, sizeof(double) * nv, hipMemcpyDeviceToHost, s));
        HIPC(hipMemcpyAsync(hsq.data(), m->sq.p, sizeof(double) * nv, hipMemcpyDeviceToHost, s));
        HIPC(hipMemcpyAsync(&hinfo, m->evinfo.p, sizeof(rocblas_int), hipMemcpyDeviceToHost, s));
        if (!out->on_device) {
            HIPC(hipMemcpyAsync(out->embeddings, emb_dev, sizeof(float) * n * nc, hipMemcpyDeviceToHost, s));
            HIPC(hipMemcpyAsync(out->components, comps_dev, sizeof(float) * nv * nc, hipMemcpyDeviceToHost, s));
        }
        HIPC(hipStreamSynchronize(s));
        if (hinfo != 0) throw Error(SCXPCA_ERR_LINALG, "Eigendecomposition failed: the symmetric eigensolver did not converge");
        std::vector<double> hmu;
        int recomputed = 0;
        const double tv = total_variance(centered, hsum, hsq, hmu, recomputed,
                                         [&] { return stored_centered_ss(m, mu); });
        const double denom = std::max(static_cast<double>(n) - 1.0, 1.0);
        for (int j = 0; j < nc; ++j) {
            if (out->variance) out->variance[j] = hvar[j];
            if (out->variance_ratio) out->variance_ratio[j] = tv > 0.0 ? hvar[j] / tv : 0.0;
        }
        if (out->sigma)  // singular values of the centred X~: sqrt((n - 1) lambda)
            for (int j = 0; j < nc; ++j) out->sigma[j] = std::sqrt(hvar[j] * denom);
        if (out->mean && centered) std::memcpy(out->mean, hmu.data(), sizeof(double) * nv);
        out->k = nc;
        out->total_variance = tv;
        out->tv_recomputed = recomputed;
    }
};

void validate(const scxpca_matrix *m, const scxpca_params *p) {
    // validate_inputs (cpu.rs:979-997)
    if (p->n_comps <= 0) throw Error(SCXPCA_ERR_INVALID, "n_components must be > 0");
    if (m->n_obs == 0 || m->n_vars == 0) throw Error(SCXPCA_ERR_INVALID, "matrix must be non-empty");
    const int64_t mn = std::min(m->n_obs, m->n_vars);
    if (p->n_comps > mn)
        throw Error(SCXPCA_ERR_INVALID, "n_components (" + std::to_string(p->n_comps) +
                                            ") exceeds matrix rank bound min(n_obs, n_vars) = " + std::to_string(mn));
    if (p->n_oversamples < 0 || p->n_power_iterations < 0)
        throw Error(SCXPCA_ERR_INVALID, "n_oversamples and n_power_iterations must be >= 0");
    if (p->qr_method != SCXPCA_QR_HOUSEHOLDER && p->qr_method != SCXPCA_QR_CHOLESKY)
        throw Error(SCXPCA_ERR_INVALID, "Invalid qr_method; expected householder or cholesky");
    if (p->precision != SCXPCA_PRECISION_FP32 && p->precision != SCXPCA_PRECISION_BF16_SKETCH)
        throw Error(SCXPCA_ERR_INVALID, "Invalid precision");
}

int64_t pca_k(const scxpca_matrix *m, const scxpca_params *p) {
    return std::min<int64_t>(std::min<int64_t>(p->n_comps + p->n_oversamples, m->n_vars), m->n_obs);
}

// One attempt's keystream and draw outcomes for n_words draw positions, and
// the outcomes' copy to the host, on s.
void omega_launch(hipStream_t s, uint64_t n_words, uint64_t seed, OmegaScratch &sc) {
    sc.w.ensure(sizeof(uint64_t) * n_words);
    sc.c.ensure(n_words);
    sc.v.ensure(sizeof(double) * n_words);
    if (!sc.tab.p) {  // the ziggurat tables are constants: built and uploaded once
        sc.tab.ensure(sizeof(double) * 2 * 257);
        sc.htab.resize(2 * 257);
        omega_tables(sc.htab.data(), sc.htab.data() + 257);
        HIPC(hipMemcpyAsync(sc.tab.p, sc.htab.data(), sizeof(double) * sc.htab.size(), hipMemcpyHostToDevice, s));
    }
    ChaChaKey key;
    omega_key(seed, key.k);
    k_chacha8_u64<<<blocks_for((n_words + 7) / 8), 256, 0, s>>>(key, n_words, sc.w.as<uint64_t>());
    HIPC(hipGetLastError());
    k_zig_outcome<<<std::min(blocks_for(n_words), 65536u), 256, 0, s>>>(
        sc.w.as<uint64_t>(), n_words, sc.tab.as<double>(), sc.tab.as<double>() + 257, sc.c.as<uint8_t>(),
        sc.v.as<double>());
    HIPC(hipGetLastError());
    sc.hc.ensure(n_words);
    HIPC(hipMemcpyAsync(sc.hc.p, sc.c.p, n_words, hipMemcpyDeviceToHost, s));
}

// device_omega's first attempt issued ahead: the keystream and outcomes run
// (and come back to the host) while the caller queues other work on s, and
// device_omega with the same rows, k and seed then waits only for the copy.
void omega_begin(hipStream_t s, uint64_t rows, int k, uint64_t seed, OmegaScratch &sc) {
    sc.started = false;
    const uint64_t n = rows * static_cast<uint64_t>(k);
    const uint64_t n_words = n + n / 16 + 4096;
    if (n == 0 || n_words >= (1ull << 32)) return;
    omega_launch(s, n_words, seed, sc);
    if (!sc.ev) HIPC(hipEventCreateWithFlags(&sc.ev, hipEventDisableTiming));
    HIPC(hipEventRecord(sc.ev, s));
    sc.started = true;
    sc.s_rows = rows;
    sc.s_k = k;
    sc.s_seed = seed;
    sc.s_words = n_words;
}

// Omega (rows x k, the reference stream) on the device: keystream and
// per-position draw outcomes in parallel, the chain of draw starts walked on
// the host, values gathered into the f32 panel (KP columns) and/or f64 out64.
// Scratch comes from the DevBufs given (the matrix's, or the caller's).
// kpw > 0: panel is a block-major wide panel of kpw columns (k > 128).
template <int KP>
void device_omega(hipStream_t s, uint64_t rows, int k, uint64_t seed, float *panel, double *out64, OmegaScratch &sc,
                  int kpw) {
    auto gather = [&]() {
        if (kpw > 0)
            k_omega_blocks<<<blocks_for(static_cast<int64_t>(rows) * kpw), 256, 0, s>>>(
                sc.v.as<double>(), sc.start.as<uint32_t>(), static_cast<int64_t>(rows), k, kpw, panel);
        else
            k_omega_panel<KP><<<blocks_for(rows * KP), 256, 0, s>>>(sc.v.as<double>(), sc.start.as<uint32_t>(),
                                                                     static_cast<int64_t>(rows), k, panel, out64);
        HIPC(hipGetLastError());
    };
    const uint64_t n = rows * static_cast<uint64_t>(k);
    if (n == 0) return;
    uint64_t n_words = n + n / 16 + 4096;
    const bool pre = sc.started && sc.s_rows == rows && sc.s_k == k && sc.s_seed == seed && sc.s_words == n_words;
    sc.started = false;
    // the previous call's uploads from the host vectors were drained by the
    // wait that precedes every walk (same stream, issued before the copy)
    for (int attempt = 0; attempt < 3; ++attempt, n_words *= 2) {
        if (n_words >= (1ull << 32)) break;
        if (attempt == 0 && pre) {
            HIPC(hipEventSynchronize(sc.ev));  // only the outcomes' copy: later work on s keeps running
        } else {
            omega_launch(s, n_words, seed, sc);
            HIPC(hipStreamSynchronize(s));
        }
        sc.hstart.ensure(sizeof(uint32_t) * n);
        std::vector<std::pair<uint32_t, double>> ov;
        if (!omega_walk(sc.hc.as<uint8_t>(), n_words, n, seed, sc.hstart.as<uint32_t>(), &ov)) continue;  // more words
        sc.start.ensure(sizeof(uint32_t) * n);
        HIPC(hipMemcpyAsync(sc.start.p, sc.hstart.p, sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
        if (!ov.empty()) {
            sc.hpos.resize(ov.size());
            sc.hval.resize(ov.size());
            for (size_t i = 0; i < ov.size(); ++i) {
                sc.hpos[i] = ov[i].first;
                sc.hval[i] = ov[i].second;
            }
            sc.ov.ensure((sizeof(uint32_t) + sizeof(double)) * ov.size());
            double *dval = sc.ov.as<double>();
            uint32_t *dpos = reinterpret_cast<uint32_t *>(dval + ov.size());
            HIPC(hipMemcpyAsync(dval, sc.hval.data(), sizeof(double) * ov.size(), hipMemcpyHostToDevice, s));
            HIPC(hipMemcpyAsync(dpos, sc.hpos.data(), sizeof(uint32_t) * ov.size(), hipMemcpyHostToDevice, s));
            k_omega_override<<<blocks_for(ov.size()), 256, 0, s>>>(sc.v.as<double>(), dpos, dval,
                                                                    static_cast<int>(ov.size()));
            HIPC(hipGetLastError());
        }
        gather();
        return;
    }
    // the chain needed far more draws than expected: the host stream
    HIPC(hipStreamSynchronize(s));
    sc.hval.resize(n);
    gaussian_matrix(rows, static_cast<uint64_t>(k), seed, sc.hval.data());
    sc.v.ensure(sizeof(double) * n);
    HIPC(hipMemcpyAsync(sc.v.p, sc.hval.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    sc.start.ensure(sizeof(uint32_t) * n);
    sc.hstart.ensure(sizeof(uint32_t) * n);
    for (uint64_t i = 0; i < n; ++i) sc.hstart.as<uint32_t>()[i] = static_cast<uint32_t>(i);
    HIPC(hipMemcpyAsync(sc.start.p, sc.hstart.p, sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
    gather();
}

// ---------------------------------------------------------------------------
// k' = n_comps + n_oversamples > 128: the power loop over wide panels. A panel
// is nb blocks of 128 columns (block-major, kernels.h), and every product runs
// block by block through the KP = 128 engine's SpMMs, on the resident or the
// streamed operator alike (the CSR is read nb times per pass). The dense
// algebra is the reference GPU path's (scx-gpu/src/cusolver.rs:87-409,
// gpu_pca.rs:592-697) on rocBLAS / rocSOLVER: the tall f32 panels are
// orthonormalized by Householder QR (geqrf + orgqr; qr_method householder) or
// by CholeskyQR2 (Gram by sgemm, potrf, trsm, twice; a failed potrf is the
// non-SPD error; qr_method cholesky) in column-major copies, and the
// Rayleigh-Ritz SVD of B (n_vars x k', f64) is gesvd. Every basis gets the
// requested QR (no shifted single pass for intermediate bases).
// ---------------------------------------------------------------------------
struct WideEngine {
    scxpca_matrix *m;
    Engine<128> &eg;
    hipStream_t s;
    int k = 0, nb = 0, kpw = 0;
    DevBuf col, col64, tau, U, Vt, S, E, W, G, info;

    WideEngine(scxpca_matrix *mm, Engine<128> &e, int kk) : m(mm), eg(e), s(mm->stream), k(kk) {
        if (m->n_obs > INT32_MAX || m->n_vars > INT32_MAX)
            throw Error(SCXPCA_ERR_INVALID, "n_comps + n_oversamples > 128 needs n_obs and n_vars below 2^31");
        nb = (k + kWideBlock - 1) / kWideBlock;
        kpw = nb * kWideBlock;
        const int64_t n = std::max<int64_t>(m->n_obs, 1), nv = std::max<int64_t>(m->n_vars, 1);
        m->y.ensure(sizeof(float) * n * kpw);
        m->q.ensure(sizeof(float) * n * kpw);
        m->panel.ensure(sizeof(float) * nv * kpw);   // M
        m->panel2.ensure(sizeof(float) * nv * kpw);  // B (f32, sketch passes)
        m->b64.ensure(sizeof(double) * nv * kpw);    // B (f64, final pass)
        col.ensure(sizeof(float) * std::max(n, nv) * k);
        col64.ensure(sizeof(double) * nv * k);
        tau.ensure(sizeof(float) * k);
        U.ensure(sizeof(double) * nv * k);
        Vt.ensure(sizeof(double) * k * k);
        S.ensure(sizeof(double) * k);
        E.ensure(sizeof(double) * k);
        W.ensure(sizeof(float) * k * k);
        G.ensure(sizeof(float) * k * k);
        info.ensure(sizeof(rocblas_int) * 2);
        if (!m->rb && rocblas_create_handle(&m->rb) != rocblas_status_success)
            throw Error(SCXPCA_ERR_DEVICE, "rocblas_create_handle failed");
        if (rocblas_set_stream(m->rb, s) != rocblas_status_success)
            throw Error(SCXPCA_ERR_DEVICE, "rocblas_set_stream failed");
    }

    void check() { HIPC(hipGetLastError()); }
    static void rbcheck(rocblas_status st, const char *what) {
        if (st != rocblas_status_success)
            throw Error(SCXPCA_ERR_DEVICE, std::string(what) + " failed (status " + std::to_string(static_cast<int>(st)) + ")");
    }

    // Qout (block-major rows x kpw) <- an orthonormal basis of A (block-major);
    // col keeps the basis column-major (rows x k)
    void qr(const float *A, int64_t rows, float *Qout, int qr_method) {
        const rocblas_int r = static_cast<rocblas_int>(rows), kk = k;
        float *c = col.as<float>();
        const unsigned gr = static_cast<unsigned>((rows + 63) / 64);
        k_blocks_to_cols<float, float><<<dim3(gr, (k + 63) / 64), 256, 0, s>>>(A, rows, k, c);
        check();
        if (qr_method == SCXPCA_QR_HOUSEHOLDER) {
            rbcheck(rocsolver_sgeqrf(m->rb, r, kk, c, r, tau.as<float>()), "rocsolver_sgeqrf");
            rbcheck(rocsolver_sorgqr(m->rb, r, kk, kk, c, r, tau.as<float>()), "rocsolver_sorgqr");
        } else {
            const float one = 1.f, zero = 0.f;
            for (int pass = 0; pass < 2; ++pass) {
                rbcheck(rocblas_sgemm(m->rb, rocblas_operation_transpose, rocblas_operation_none, kk, kk, r, &one, c, r,
                                      c, r, &zero, G.as<float>(), kk),
                        "rocblas_sgemm");
                rbcheck(rocsolver_spotrf(m->rb, rocblas_fill_upper, kk, G.as<float>(), kk, info.as<rocblas_int>()),
                        "rocsolver_spotrf");
                k_info_status<<<1, 64, 0, s>>>(info.as<rocblas_int>(), 1, eg.status);
                check();
                rbcheck(rocblas_strsm(m->rb, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                                      rocblas_diagonal_non_unit, r, kk, &one, G.as<float>(), kk, c, r),
                        "rocblas_strsm");
            }
        }
        k_cols_to_blocks<float><<<dim3(gr, kpw / 64), 256, 0, s>>>(c, rows, k, kpw, Qout);
        check();
    }

    template <class Ops>
    void run_with(Ops &ops, const scxpca_params *p, scxpca_result *out) {
        const int64_t n = m->n_obs, nv = m->n_vars;
        const bool centered = p->zero_center != 0;
        const bool bf16 = p->precision == SCXPCA_PRECISION_BF16_SKETCH;
        const int nc = static_cast<int>(p->n_comps);
        HIPC(hipMemsetAsync(eg.status, 0, sizeof(int), s));
        if (bf16) ensure_bf16_entries(m);
        ops.stats();
        k_mean<<<blocks_for(nv), 256, 0, s>>>(m->sum.as<double>(), nv, static_cast<double>(n), eg.baseline_mean(), eg.mu);
        check();
        float *M = m->panel.as<float>(), *B32 = m->panel2.as<float>(), *Y = m->y.as<float>(), *Q = m->q.as<float>();
        double *B64 = m->b64.as<double>();
        const int64_t bn = n * kWideBlock, bv = nv * kWideBlock;  // block strides
        device_omega<128>(s, static_cast<uint64_t>(nv), k, p->seed, M, nullptr, m->om, kpw);
        auto fwd_all = [&](const float *Mp, float *Yp) {
            for (int b = 0; b < nb; ++b) {
                if (centered) eg.mean_correction(Mp + b * bv);
                ops.fwd(Mp + b * bv, centered, bf16, Yp + b * bn);
            }
        };
        auto t_all = [&](const float *Qp, bool lowp, double *Bd, float *Bf) {
            for (int b = 0; b < nb; ++b)
                ops.t(Qp + b * bn, centered, lowp, Bd ? Bd + b * bv : nullptr, Bf ? Bf + b * bv : nullptr);
        };
        fwd_all(M, Y);
        qr(Y, n, Q, p->qr_method);
        for (int64_t it = 0; it < p->n_power_iterations; ++it) {
            t_all(Q, bf16, nullptr, B32);
            qr(B32, nv, M, p->qr_method);
            fwd_all(M, Y);
            qr(Y, n, Q, p->qr_method);
        }
        // final B pass in FP32 (also in the BF16 sketch form); col still holds Q
        t_all(Q, false, B64, nullptr);
        double *b = col64.as<double>();
        k_blocks_to_cols<double, double><<<dim3(static_cast<unsigned>((nv + 63) / 64), (k + 63) / 64), 256, 0, s>>>(
            B64, nv, k, b);
        check();
        rbcheck(rocsolver_dgesvd(m->rb, rocblas_svect_singular, rocblas_svect_singular, static_cast<rocblas_int>(nv), k,
                                 b, static_cast<rocblas_int>(nv), S.as<double>(), U.as<double>(),
                                 static_cast<rocblas_int>(nv), Vt.as<double>(), k, E.as<double>(), rocblas_outofplace,
                                 info.as<rocblas_int>() + 1),
                "rocsolver_dgesvd");
        k_info_status<<<1, 64, 0, s>>>(info.as<rocblas_int>() + 1, 2, eg.status);
        check();
        // components = U[:, :nc]^T; embeddings = Q (V S)[:, :nc]
        DevBuf comps_tmp, emb_tmp;
        float *comps_dev = out->components, *emb_dev = out->embeddings;
        if (!out->on_device) {
            comps_tmp.ensure(sizeof(float) * nv * nc);
            emb_tmp.ensure(sizeof(float) * n * nc);
            comps_dev = comps_tmp.as<float>();
            emb_dev = emb_tmp.as<float>();
        }
        k_f64_to_f32<<<blocks_for(nv * nc), 256, 0, s>>>(U.as<double>(), comps_dev, nv * nc);
        check();
        k_wide_emb_weights<<<blocks_for(static_cast<int64_t>(k) * nc), 256, 0, s>>>(Vt.as<double>(), S.as<double>(), k,
                                                                                   nc, W.as<float>());
        check();
        {
            const float one = 1.f, zero = 0.f;
            rbcheck(rocblas_sgemm(m->rb, rocblas_operation_transpose, rocblas_operation_transpose, nc,
                                  static_cast<rocblas_int>(n), k, &one, W.as<float>(), k, col.as<float>(),
                                  static_cast<rocblas_int>(n), &zero, emb_dev, nc),
                    "rocblas_sgemm");
        }
        std::vector<double> hsig(k), hsum(nv), hsq(nv);
        int hstatus = 0;
        HIPC(hipMemcpyAsync(hsig.data(), S.p, sizeof(double) * k, hipMemcpyDeviceToHost, s));
        HIPC(hipMemcpyAsync(hsum.data(), m->sum.p, sizeof(double) * nv, hipMemcpyDeviceToHost, s));
        HIPC(hipMemcpyAsync(hsq.data(), m->sq.p, sizeof(double) * nv, hipMemcpyDeviceToHost, s));
        HIPC(hipMemcpyAsync(&hstatus, eg.status, sizeof(int), hipMemcpyDeviceToHost, s));
        if (!out->on_device) {
            HIPC(hipMemcpyAsync(out->embeddings, emb_dev, sizeof(float) * n * nc, hipMemcpyDeviceToHost, s));
            HIPC(hipMemcpyAsync(out->components, comps_dev, sizeof(float) * nv * nc, hipMemcpyDeviceToHost, s));
        }
        HIPC(hipStreamSynchronize(s));
        if (hstatus & 1) {
            if (p->qr_method == SCXPCA_QR_CHOLESKY)
                throw Error(SCXPCA_ERR_LINALG, "CholeskyQR2 failed: non-SPD; retry with qr_method='householder'");
            throw Error(SCXPCA_ERR_LINALG, "Householder QR failed");
        }
        if (hstatus & 2) throw Error(SCXPCA_ERR_LINALG, "SVD failed: the bidiagonal iteration did not converge");
        std::vector<double> hmu;
        int recomputed = 0;
        const double tv = eg.total_variance(centered, hsum, hsq, hmu, recomputed, [&] { return ops.stored_ss(); });
        const double denom = std::max(static_cast<double>(n) - 1.0, 1.0);
        for (int j = 0; j < nc; ++j) {
            const double var = hsig[j] * hsig[j] / denom;
            if (out->variance) out->variance[j] = var;
            if (out->variance_ratio) out->variance_ratio[j] = tv > 0.0 ? var / tv : 0.0;
        }
        if (out->sigma)
            for (int j = 0; j < k; ++j) out->sigma[j] = hsig[j];
        if (out->mean && centered) std::memcpy(out->mean, hmu.data(), sizeof(double) * nv);
        out->k = k;
        out->total_variance = tv;
        out->tv_recomputed = recomputed;
    }

    void run(const scxpca_params *p, scxpca_result *out) {
        typename Engine<128>::ResidentOps ops{&eg};
        run_with(ops, p, out);
    }
};

void run_pca(scxpca_matrix *m, const scxpca_params *p, scxpca_result *out) {
    validate(m, p);
    const int64_t k = pca_k(m, p);
    DeviceGuard dg(m->device);
    hipEvent_t e0, e1;
    HIPC(hipEventCreate(&e0));
    HIPC(hipEventCreate(&e1));
    HIPC(hipEventRecord(e0, m->stream));
    // Omega's keystream and outcomes first: the host walk of its draw chain
    // then overlaps the statistics and set-up work queued below
    if (k <= 128) omega_begin(m->stream, static_cast<uint64_t>(m->n_vars), static_cast<int>(k), p->seed, m->om);
    if (p->precision == SCXPCA_PRECISION_BF16_SKETCH) ensure_bf16_entries(m);
    ensure_stats(m);
    if (k <= 64) {
        Engine<64> eng(m, static_cast<int>(k));
        eng.run(p, out);
    } else if (k <= 128) {
        Engine<128> eng(m, static_cast<int>(k));
        eng.run(p, out);
    } else {
        Engine<128> eng(m, 128);
        WideEngine wide(m, eng, static_cast<int>(k));
        wide.run(p, out);
    }
    HIPC(hipEventRecord(e1, m->stream));
    HIPC(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPC(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    out->seconds_device = ms * 1e-3;
    out->resident = 1;
}

// The covariance genes limit: the n_vars x n_vars f64 covariance and the
// Gram partials stay modest (1.7 GiB at the limit).
constexpr int64_t kMaxCovVars = 16384;

void run_pca_covariance(scxpca_matrix *m, const scxpca_params *p, scxpca_result *out) {
    if (m->pflog) throw Error(SCXPCA_ERR_INVALID, "the covariance route is not defined for a PFlog matrix");
    validate(m, p);
    if (m->n_vars > kMaxCovVars)
        throw Error(SCXPCA_ERR_INVALID, "the covariance route supports up to " + std::to_string(kMaxCovVars) +
                                            " genes (n_vars = " + std::to_string(m->n_vars) + "); use method='randomized'");
    DeviceGuard dg(m->device);
    hipEvent_t e0, e1;
    HIPC(hipEventCreate(&e0));
    HIPC(hipEventCreate(&e1));
    HIPC(hipEventRecord(e0, m->stream));
    if (p->n_comps <= 64) {
        Engine<64> eng(m, static_cast<int>(p->n_comps));
        eng.run_covariance(p, out);
    } else {
        Engine<128> eng(m, static_cast<int>(p->n_comps));
        eng.run_covariance(p, out);
    }
    HIPC(hipEventRecord(e1, m->stream));
    HIPC(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPC(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    out->seconds_device = ms * 1e-3;
    out->resident = 1;
}

template <int KP>
void stage_spmm(scxpca_matrix *m, int transpose, int64_t k, const double *in, const double *means, int bf16,
                double *out) {
    Engine<KP> eng(m, static_cast<int>(k));
    const int64_t n = m->n_obs, nv = m->n_vars;
    hipStream_t s = m->stream;
    if (bf16) ensure_bf16_entries(m);
    const bool centered = means != nullptr;
    if (centered) HIPC(hipMemcpyAsync(eng.mu, means, sizeof(double) * nv, hipMemcpyHostToDevice, s));
    if (!transpose) {
        HIPC(hipMemcpyAsync(m->panel64.p, in, sizeof(double) * nv * k, hipMemcpyHostToDevice, s));
        k_pack_panel<KP><<<blocks_for(nv * KP), 256, 0, s>>>(m->panel64.as<double>(), nv, static_cast<int>(k),
                                                             m->panel.as<float>());
        HIPC(hipGetLastError());
        if (centered) eng.mean_correction(m->panel.as<float>());
        eng.spmm_fwd(m->panel.as<float>(), centered, bf16 != 0, m->y.as<float>());
        std::vector<float> h(static_cast<size_t>(n) * KP);
        HIPC(hipMemcpyAsync(h.data(), m->y.p, sizeof(float) * n * KP, hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
        for (int64_t r = 0; r < n; ++r)
            for (int64_t j = 0; j < k; ++j) out[r * k + j] = h[r * KP + j];
    } else {
        std::vector<float> h(static_cast<size_t>(std::max<int64_t>(n, 1)) * KP, 0.f);
        for (int64_t r = 0; r < n; ++r)
            for (int64_t j = 0; j < k; ++j) h[r * KP + j] = static_cast<float>(in[r * k + j]);
        HIPC(hipMemcpyAsync(m->q.p, h.data(), sizeof(float) * n * KP, hipMemcpyHostToDevice, s));
        eng.spmm_t(m->q.as<float>(), centered, bf16 != 0, m->b64.as<double>(), nullptr);
        std::vector<double> hb(static_cast<size_t>(nv) * KP);
        HIPC(hipMemcpyAsync(hb.data(), m->b64.p, sizeof(double) * nv * KP, hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
        for (int64_t v = 0; v < nv; ++v)
            for (int64_t j = 0; j < k; ++j) out[v * k + j] = hb[v * KP + j];
    }
}

template <int KP>
void stage_qr(int64_t rows, int64_t k, const double *a, int qr_method, double *qout) {
    // A throwaway matrix shell provides the stream and workspace.
    scxpca_matrix shell;
    HIPC(hipStreamCreate(&shell.stream));
    shell.n_obs = rows;
    shell.n_vars = 1;
    Engine<KP> eng(&shell, static_cast<int>(k));
    hipStream_t s = shell.stream;
    std::vector<float> h(static_cast<size_t>(rows) * KP, 0.f);
    for (int64_t r = 0; r < rows; ++r)
        for (int64_t j = 0; j < k; ++j) h[r * KP + j] = static_cast<float>(a[r * k + j]);
    HIPC(hipMemsetAsync(eng.status, 0, sizeof(int), s));
    HIPC(hipMemcpyAsync(shell.y.p, h.data(), sizeof(float) * rows * KP, hipMemcpyHostToDevice, s));
    float *Q = eng.cqr_tall(shell.y.as<float>(), shell.q.as<float>(), qr_method);
    int st = 0;
    HIPC(hipMemcpyAsync(h.data(), Q, sizeof(float) * rows * KP, hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(&st, eng.status, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    if (st & 1) {
        if (qr_method == SCXPCA_QR_CHOLESKY)
            throw Error(SCXPCA_ERR_LINALG, "CholeskyQR2 failed: non-SPD; retry with qr_method='householder'");
        throw Error(SCXPCA_ERR_LINALG, "shifted CholeskyQR failed: Gram matrix not positive definite");
    }
    for (int64_t r = 0; r < rows; ++r)
        for (int64_t j = 0; j < k; ++j) qout[r * k + j] = h[r * KP + j];
}


// Projection Z = (X~ - 1 mu^T) W (forward) and dW = (X~ - 1 mu^T)^T dZ
// (backward) on a caller stream; W / dW are n_vars x k, Z / dZ n_obs x k, f32
// row-major device buffers.
template <int KP>
void project(scxpca_matrix *m, bool backward, const float *in, int64_t k, const float *mean, float *out,
             hipStream_t user) {
    hipStream_t saved = m->stream;
    if (user) m->stream = user;
    struct Restore {
        scxpca_matrix *m;
        hipStream_t s;
        ~Restore() { m->stream = s; }
    } restore{m, saved};
    Engine<KP> eng(m, static_cast<int>(k));
    hipStream_t s = m->stream;
    const int64_t n = m->n_obs, nv = m->n_vars;
    const bool centered = mean != nullptr;
    if (centered) {
        k_f32_to_f64<<<blocks_for(nv), 256, 0, s>>>(mean, eng.mu, nv);
        HIPC(hipGetLastError());
    }
    if (!backward) {
        const float *panel = in;
        if (k != KP) {
            k_pack_panel32<KP><<<blocks_for(nv * KP), 256, 0, s>>>(in, nv, static_cast<int>(k), m->panel.as<float>());
            HIPC(hipGetLastError());
            panel = m->panel.as<float>();
        }
        if (centered) eng.mean_correction(panel);
        float *dst = (k == KP) ? out : m->y.as<float>();
        eng.spmm_fwd(panel, centered, false, dst);
        if (k != KP) {
            k_unpack_panel32<KP><<<blocks_for(n * k), 256, 0, s>>>(dst, n, static_cast<int>(k), out);
            HIPC(hipGetLastError());
        }
    } else {
        const float *qp = in;
        if (k != KP) {
            k_pack_panel32<KP><<<blocks_for(n * KP), 256, 0, s>>>(in, n, static_cast<int>(k), m->q.as<float>());
            HIPC(hipGetLastError());
            qp = m->q.as<float>();
        }
        float *dst = (k == KP) ? out : m->panel.as<float>();
        eng.spmm_t(qp, centered, false, nullptr, dst);
        if (k != KP) {
            k_unpack_panel32<KP><<<blocks_for(nv * k), 256, 0, s>>>(dst, nv, static_cast<int>(k), out);
            HIPC(hipGetLastError());
        }
    }
    if (!user) HIPC(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// Allocation-free projection backward (scxpca_project_backward_noalloc): the
// same kernels and results as project<KP>(backward), with every buffer
// beyond the caller's inputs and outputs (the transposed SpMM's tile-range
// partials, the 1^T dZ / b^T dZ sums, the f64 mean, the packed dZ / dW of a
// k below the panel width, the wave-slot gene order when the matrix has none
// yet) carved from one file-scope device array. Each GPU holds its own copy
// of the array (at its own address, found once per device by
// hipGetSymbolAddress); calls on one GPU share it, so calls that overlap on
// one GPU race. Every region a call reads is written earlier in that call.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxScratchDevices] = {};

// The device address of `scratch` on device dev (the current device).
void *noalloc_scratch(int dev) {
    if (dev < 0 || dev >= kMaxScratchDevices)
        throw Error(SCXPCA_ERR_NO_DEVICE, "device ordinal beyond the allocation-free entry point's table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            throw Error(SCXPCA_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

template <int KP>
void project_backward_noalloc(scxpca_matrix *m, const float *dz, int64_t k, const float *mean, float *dw,
                              hipStream_t user) {
    using E = Engine<KP>;
    hipStream_t s = user ? user : m->stream;
    const int64_t n = m->n_obs, nv = m->n_vars;
    int nch = 0, nrb = 0, tpr = 0;
    E::t_geometry(m, nch, nrb, tpr);
    const bool pack = k != KP;
    const bool have_order = gene_order_valid(m, E::GW, nch, E::TU);
    const int64_t slots = std::max<int64_t>(static_cast<int64_t>(nch) * E::TU * E::GW, 1);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) / 256 * 256;
        return o;
    };
    const size_t o_tpart = take(sizeof(float) * nrb * nv * KP);
    const size_t o_qsp = take(sizeof(double) * nrb * nch * KP);
    const size_t o_cpart = take(sizeof(double) * kMaxColBlocks * KP);
    const size_t o_sumq = take(sizeof(double) * KP), o_bq = take(sizeof(double) * KP);
    const size_t o_mu = mean ? take(sizeof(double) * std::max<int64_t>(nv, 1)) : 0;
    const size_t o_in = pack ? take(sizeof(float) * std::max<int64_t>(n, 1) * KP) : 0;
    const size_t o_out = pack ? take(sizeof(float) * std::max<int64_t>(nv, 1) * KP) : 0;
    const size_t o_perm = have_order ? 0 : take(sizeof(uint32_t) * slots);
    if (off > kNoAllocScratchBytes)
        throw Error(SCXPCA_ERR_INVALID, "scxpca_project_backward_noalloc: this matrix and k need " +
                                            std::to_string(off) + " bytes of scratch, above the entry point's fixed " +
                                            std::to_string(kNoAllocScratchBytes) +
                                            " bytes; use scxpca_project_backward");
    char *base = static_cast<char *>(noalloc_scratch(m->device));
    const uint32_t *gp = nullptr;
    if (have_order) {
        gp = m->gorder.dev.template as<uint32_t>();
    } else {  // dealt on the host, staged through pageable memory (no allocation on the device)
        std::vector<uint32_t> h(slots);
        deal_gene_order(m, E::GW, nch, E::TU, h.data());
        HIPC(hipMemcpyAsync(base + o_perm, h.data(), sizeof(uint32_t) * slots, hipMemcpyHostToDevice, s));
        HIPC(hipStreamSynchronize(s));
        gp = reinterpret_cast<const uint32_t *>(base + o_perm);
    }
    double *mu = nullptr;
    if (mean) {
        mu = reinterpret_cast<double *>(base + o_mu);
        k_f32_to_f64<<<blocks_for(nv), 256, 0, s>>>(mean, mu, nv);
        HIPC(hipGetLastError());
    }
    const float *qp = dz;
    float *dst = dw;
    if (pack) {
        float *packed = reinterpret_cast<float *>(base + o_in);
        k_pack_panel32<KP><<<blocks_for(n * KP), 256, 0, s>>>(dz, n, static_cast<int>(k), packed);
        HIPC(hipGetLastError());
        qp = packed;
        dst = reinterpret_cast<float *>(base + o_out);
    }
    const typename E::TScratch sc{reinterpret_cast<float *>(base + o_tpart), reinterpret_cast<double *>(base + o_qsp),
                                  reinterpret_cast<double *>(base + o_cpart), reinterpret_cast<double *>(base + o_sumq),
                                  reinterpret_cast<double *>(base + o_bq)};
    const double *rowb = (m->pflog && n > 0) ? m->pfb.as<double>() : nullptr;
    E::transposed(m, s, gp, nch, nrb, tpr, qp, mu, rowb, sc, nullptr, dst);
    if (pack) {
        k_unpack_panel32<KP><<<blocks_for(nv * k), 256, 0, s>>>(dst, nv, static_cast<int>(k), dw);
        HIPC(hipGetLastError());
    }
    if (!user) HIPC(hipStreamSynchronize(s));
}

// PFlog flags for the delta pass.
int pflog_flags(const scxpca_pflog_params &pp) {
    return kXfScale | kXfLog1p | kXfCheck | (pp.delta_form == SCXPCA_PFLOG_DELTA_F64 ? kXfLogF64 : 0);
}

// estimate_alpha (scx-accel/src/pflog.rs:340-398) from the raw column sums
// and sums of squares: finalize_column_moments (scx-sparse/src/moments.rs:
// 193-233) per gene, alpha_g = (var - mean) / mean^2 for every gene with
// mean > mu_min (finite values only), the median of the pool (:403-415), the
// fallback when the pool is empty or its median is not positive.
#pragma clang fp contract(off)
scxpca_pflog_fit estimate_alpha(const std::vector<double> &sum, const std::vector<double> &sq, int64_t n,
                                const scxpca_pflog_params &pp) {
    std::vector<double> cand;
    if (n > 0) {
        const double nf = static_cast<double>(n), denom = std::max(nf - 1.0, 1.0);
        for (size_t j = 0; j < sum.size(); ++j) {
            const double mean = sum[j] / nf;
            const double residual = sq[j] - nf * mean * mean;
            double var = residual / denom;
            if (var < 0.0) var = 0.0;
            if (!(mean > pp.mu_min)) continue;
            const double a = (var - mean) / (mean * mean);
            if (std::isfinite(a)) cand.push_back(a);
        }
    }
    scxpca_pflog_fit f{};
    f.estimated = 1;
    f.n_genes_used = static_cast<int64_t>(cand.size());
    double alpha = 0.0;
    if (!cand.empty()) {
        std::sort(cand.begin(), cand.end());
        const size_t h = cand.size() / 2;
        alpha = cand.size() % 2 ? cand[h] : 0.5 * (cand[h - 1] + cand[h]);
    }
    if (cand.empty() || !std::isfinite(alpha) || alpha <= 0.0) {
        f.alpha = pp.fallback_alpha;
        f.fell_back = 1;
    } else {
        f.alpha = alpha;
        f.fell_back = 0;
    }
    f.pseudocount = 1.0 / (4.0 * f.alpha);
    return f;
}
#pragma clang fp contract(on)

// The PFlog preparation of an uploaded matrix: with an estimated alpha a raw
// pass (X~ = x, finite check) gives the column moments, then the delta pass
// (X~ = log1p(4 alpha x), baseline) replaces it.
void prepare_pflog(scxpca_matrix *m, bool raw_ready) {
    if (m->pf_fit.estimated) {
        if (!raw_ready) {
            m->flags = kXfCheck;
            m->target = 0.0;
            throw_invalid(m, prepare(m));
        }
        ensure_stats(m);
        std::vector<double> hs(m->n_vars), hq(m->n_vars);
        if (m->n_vars > 0) {
            HIPC(hipMemcpyAsync(hs.data(), m->sum.p, sizeof(double) * m->n_vars, hipMemcpyDeviceToHost, m->stream));
            HIPC(hipMemcpyAsync(hq.data(), m->sq.p, sizeof(double) * m->n_vars, hipMemcpyDeviceToHost, m->stream));
        }
        HIPC(hipStreamSynchronize(m->stream));
        m->pf_fit = estimate_alpha(hs, hq, m->n_obs, m->pf_params);
    }
    m->flags = pflog_flags(m->pf_params);
    m->target = 4.0 * m->pf_fit.alpha;
    throw_invalid(m, prepare(m));
}

void build_pflog(scxpca_matrix *m, const scxpca_csr *x, const scxpca_pflog_params &pp) {
    const double a = pp.alpha;
    if (!(a == 0.0 || (a > 0.0 && std::isfinite(a))))
        throw Error(SCXPCA_ERR_INVALID, "pflog: alpha must be positive and finite, got " + fmt_f32(static_cast<float>(a)));
    if (pp.delta_form != SCXPCA_PFLOG_DELTA_LAZY && pp.delta_form != SCXPCA_PFLOG_DELTA_F64)
        throw Error(SCXPCA_ERR_INVALID, "unknown PFlog delta form");
    if (!(pp.fallback_alpha > 0.0 && std::isfinite(pp.fallback_alpha)) || !std::isfinite(pp.mu_min))
        throw Error(SCXPCA_ERR_INVALID, "PFlog alpha options must be finite with a positive fallback");
    if (x->n_vars == 0) throw Error(SCXPCA_ERR_INVALID, "PFlog requires n_vars > 0 for centering");
    m->pflog = true;
    m->pf_params = pp;
    m->pf_fit = scxpca_pflog_fit{};
    if (a == 0.0) {
        m->pf_fit.estimated = 1;
        build_matrix(m, x, 0.0, kXfCheck);  // the raw pass
    } else {
        m->pf_fit.alpha = a;
        m->pf_fit.pseudocount = 1.0 / (4.0 * a);
        m->pf_fit.n_genes_used = -1;
        build_matrix(m, x, 4.0 * a, pflog_flags(pp));
        return;
    }
    prepare_pflog(m, true);
}

void refresh(scxpca_matrix *m) {
    if (m->pflog) prepare_pflog(m, false);
    else throw_invalid(m, prepare(m));
}
}  // namespace

// ---------------------------------------------------------------------------
// Streaming operator for a CSR larger than HBM (StreamingPcaOperator,
// scx-gpu/src/linear_operator.rs:61-266; PinnedCsrSlot / GpuCsrSlot,
// scx-gpu/src/staging.rs). The CSR stays in host memory; every pass of the
// power loop (the statistics, each forward and transposed product, the
// stable total-variance recompute) drives it shard by shard through two
// device slots. A shard's rows go host -> pinned staging (a host copy) ->
// HBM (an asynchronous copy on a copy stream), so the next shard's transfer
// runs while the current shard is prepared and multiplied on the compute
// stream. Each shard is a matrix of its own (the resident prepare: X~, the
// split offsets, the tile-CSC) whose products write the shard's rows of the
// global Y, read the shard's rows of the global Q, and add their n_vars x KP
// f64 transposed contribution into one accumulator; the centering
// B = sum_s X~_s^T Q_s - mu (1^T Q) is applied once, after the last shard.
// Nothing is cached between passes: every pass moves the whole CSR over PCIe,
// as the reference's streaming operator re-uploads on every multiply.
// ---------------------------------------------------------------------------

__global__ void k_rebase(int64_t *ip, int64_t n, int64_t base) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) ip[i] -= base;
}

__global__ void k_add_f64(double *acc, const double *x, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) acc[i] += x[i];
}

__global__ void k_add_u32(uint32_t *acc, const uint32_t *x, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) acc[i] += x[i];
}

// B[v][j] = acc[v][j] - mu[v] sumq[j] (mu null: no centering); B64 / B32 optional
template <int KP>
__global__ void k_center_b(const double *acc, const double *mu, const double *sumq, int64_t n_vars, double *B64,
                           float *B32) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n_vars * KP) return;
    double v = acc[t];
    if (mu) v -= mu[t / KP] * sumq[t % KP];
    if (B64) B64[t] = v;
    if (B32) B32[t] = static_cast<float>(v);
}

// dst <- src, n bytes, split over up to nt host threads (page-aligned pieces)
void par_memcpy(void *dst, const void *src, size_t n, int nt) {
    if (nt <= 1 || n < (size_t(8) << 20)) {
        std::memcpy(dst, src, n);
        return;
    }
    const size_t per = ((n + nt - 1) / nt + 4095) & ~size_t(4095);
    std::vector<std::thread> th;
    for (size_t off = 0; off < n; off += per) {
        const size_t len = std::min(per, n - off);
        th.emplace_back([=] { std::memcpy(static_cast<char *>(dst) + off, static_cast<const char *>(src) + off, len); });
    }
    for (auto &t : th) t.join();
}

// true when p is page-locked host memory (hipHostMalloc'd): copies can then
// go to the device directly; pageable memory (an error or an unregistered
// type here) is staged through pinned buffers
bool host_pinned(const void *p) {
    hipPointerAttribute_t a;
    const hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // clear the query's error
        return false;
    }
    return a.type == hipMemoryTypeHost;
}

struct StreamSlot {
    DevBuf ip, ix, dv;        // the shard's CSR in HBM (indptr rebased to 0)
    PinnedBuf hip, hix, hdv;  // pinned staging
    hipEvent_t copied = nullptr;
    scxpca_matrix m;          // the shard matrix over ip / ix / dv (stream borrowed)
    int64_t r0 = 0;
    ~StreamSlot() {
        if (copied) (void)hipEventDestroy(copied);
        m.stream = nullptr;  // the run's compute stream, not the slot's
    }
};

template <int KP>
struct StreamOps {
    Engine<KP> &eg;  // the engine of the shell matrix: global panels, QR, results
    const scxpca_csr *x;
    int64_t shard_rows, n, nv;
    int nshards, k;
    hipStream_t cs, copy = nullptr;
    StreamSlot slot[2];
    DevBuf acc;
    bool pinned_ix = false, pinned_dv = false;  // the caller's arrays are page-locked: no staging
    // shard transform: the shell's (normalize / log1p), PFlog with the shell's
    // alpha (pflog), or the raw pass of the alpha estimate (raw)
    bool pflog = false, raw = false;
    double *baseline_out = nullptr;  // host n_obs: the PFlog baseline, gathered by stats()
    int host_threads = 1;
    static constexpr size_t kChunk = size_t(64) << 20;  // staging piece: its transfer overlaps the next copy

    StreamOps(Engine<KP> &e, const scxpca_csr *xx, int64_t rows_per_shard)
        : eg(e), x(xx), shard_rows(rows_per_shard), n(xx->n_obs), nv(xx->n_vars), k(e.k), cs(e.s) {
        nshards = static_cast<int>((n + shard_rows - 1) / shard_rows);
        HIPC(hipStreamCreate(&copy));
        pinned_ix = x->indptr[n] > 0 && host_pinned(x->indices);
        pinned_dv = x->indptr[n] > 0 && host_pinned(x->data);
        host_threads = static_cast<int>(std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
        for (auto &sl : slot) {
            HIPC(hipEventCreateWithFlags(&sl.copied, hipEventDisableTiming));
            sl.m.stream = cs;
            sl.m.device = eg.m->device;
        }
    }
    ~StreamOps() {
        if (copy) {
            (void)hipStreamSynchronize(copy);
            (void)hipStreamDestroy(copy);
        }
        (void)hipStreamSynchronize(cs);
    }

    // host bytes -> HBM on the copy stream: directly from page-locked memory,
    // else through the pinned staging buffer in kChunk pieces (threaded host
    // copy of piece i + 1 while piece i transfers)
    void put(void *ddev, PinnedBuf &pin, const void *src, size_t bytes, bool src_pinned) {
        if (bytes == 0) return;
        if (src_pinned) {
            HIPC(hipMemcpyAsync(ddev, src, bytes, hipMemcpyHostToDevice, copy));
            return;
        }
        pin.ensure(bytes);
        for (size_t off = 0; off < bytes; off += kChunk) {
            const size_t len = std::min(kChunk, bytes - off);
            par_memcpy(static_cast<char *>(pin.p) + off, static_cast<const char *>(src) + off, len, host_threads);
            HIPC(hipMemcpyAsync(static_cast<char *>(ddev) + off, static_cast<char *>(pin.p) + off, len,
                                hipMemcpyHostToDevice, copy));
        }
    }

    // host rows of shard s -> the slot's device CSR (copy stream)
    void stage(int s) {
        StreamSlot &sl = slot[s & 1];
        const int64_t r0 = static_cast<int64_t>(s) * shard_rows, r1 = std::min(n, r0 + shard_rows);
        const int64_t a = x->indptr[r0], nnz = x->indptr[r1] - a, rows = r1 - r0;
        sl.ip.ensure(sizeof(int64_t) * (rows + 1));
        sl.ix.ensure(sizeof(uint32_t) * (nnz + kEntryPad));
        sl.dv.ensure(sizeof(float) * (nnz + kEntryPad));
        put(sl.ip.p, sl.hip, x->indptr + r0, sizeof(int64_t) * (rows + 1), false);
        put(sl.ix.p, sl.hix, x->indices + a, sizeof(uint32_t) * nnz, pinned_ix);
        put(sl.dv.p, sl.hdv, x->data + a, sizeof(float) * nnz, pinned_dv);
        HIPC(hipEventRecord(sl.copied, copy));
    }

    // the staged shard s as a prepared matrix (the prepare waits on the compute stream)
    StreamSlot &load(int s) {
        StreamSlot &sl = slot[s & 1];
        const int64_t r0 = static_cast<int64_t>(s) * shard_rows, r1 = std::min(n, r0 + shard_rows);
        const int64_t a = x->indptr[r0];
        HIPC(hipStreamWaitEvent(cs, sl.copied, 0));
        k_rebase<<<blocks_for(r1 - r0 + 1), 256, 0, cs>>>(sl.ip.as<int64_t>(), r1 - r0 + 1, a);
        HIPC(hipGetLastError());
        scxpca_matrix *m = &sl.m;
        m->n_obs = r1 - r0;
        m->n_vars = nv;
        m->nnz = x->indptr[r1] - a;
        m->indptr = sl.ip.as<int64_t>();
        m->idx = sl.ix.as<uint32_t>();
        m->val = sl.dv.as<float>();
        if (raw) {
            m->pflog = false;
            m->target = 0.0;
            m->flags = kXfCheck;
        } else if (pflog) {
            m->pflog = true;
            m->pf_params = eg.m->pf_params;
            m->pf_fit = eg.m->pf_fit;
            m->target = eg.m->target;
            m->flags = pflog_flags(m->pf_params);
        } else {
            m->target = eg.m->target;
            m->flags = eg.m->flags;
        }
        size_derived(m);
        throw_invalid(m, prepare(m), r0, a);
        sl.r0 = r0;
        return sl;
    }

    // f(slot) for every shard in order; shard s + 1 is staged while s computes
    template <class F>
    void each_shard(F &&f) {
        stage(0);
        for (int sh = 0; sh < nshards; ++sh) {
            StreamSlot &sl = load(sh);
            f(sl);
            if (sh + 1 < nshards) stage(sh + 1);
        }
    }

    void stats() {
        scxpca_matrix *g = eg.m;
        g->sum.ensure(sizeof(double) * nv);
        g->sq.ensure(sizeof(double) * nv);
        g->cnt.ensure(sizeof(uint32_t) * nv);
        HIPC(hipMemsetAsync(g->sum.p, 0, sizeof(double) * nv, cs));
        HIPC(hipMemsetAsync(g->sq.p, 0, sizeof(double) * nv, cs));
        HIPC(hipMemsetAsync(g->cnt.p, 0, sizeof(uint32_t) * nv, cs));
        if (pflog) g->pf_bsum = g->pf_b2 = 0.0;
        each_shard([&](StreamSlot &sl) {
            ensure_stats(&sl.m);
            if (pflog) {  // the shards' baseline sums, in shard order
                g->pf_bsum += sl.m.pf_bsum;
                g->pf_b2 += sl.m.pf_b2;
                if (baseline_out)
                    HIPC(hipMemcpyAsync(baseline_out + sl.r0, sl.m.pfb.p, sizeof(double) * sl.m.n_obs,
                                        hipMemcpyDeviceToHost, cs));
            }
            k_add_f64<<<blocks_for(nv), 256, 0, cs>>>(g->sum.as<double>(), sl.m.sum.as<double>(), nv);
            k_add_f64<<<blocks_for(nv), 256, 0, cs>>>(g->sq.as<double>(), sl.m.sq.as<double>(), nv);
            k_add_u32<<<blocks_for(nv), 256, 0, cs>>>(g->cnt.as<uint32_t>(), sl.m.cnt.as<uint32_t>(), nv);
            HIPC(hipGetLastError());
        });
        g->stats_ready = true;
    }

    // Y = X~ panel - 1 mc^T, shard by shard into the shard's rows of Y
    void fwd(const float *panel, bool centered, bool bf16, float *Y) {
        (void)bf16;
        each_shard([&](StreamSlot &sl) {
            Engine<KP> es(&sl.m, k);
            if (centered) HIPC(hipMemcpyAsync(es.mc32, eg.mc32, sizeof(float) * KP, hipMemcpyDeviceToDevice, cs));
            es.spmm_fwd(panel, centered, false, Y + sl.r0 * KP);
        });
    }

    // B = sum_s X~_s^T Q_s - mu (1^T Q)
    void t(const float *Q, bool centered, bool bf16, double *B64, float *B32) {
        (void)bf16;
        acc.ensure(sizeof(double) * nv * KP);
        HIPC(hipMemsetAsync(acc.p, 0, sizeof(double) * nv * KP, cs));
        each_shard([&](StreamSlot &sl) {
            Engine<KP> es(&sl.m, k);
            es.spmm_t(Q + sl.r0 * KP, false, false, sl.m.b64.as<double>(), nullptr);
            k_add_f64<<<blocks_for(nv * KP), 256, 0, cs>>>(acc.as<double>(), sl.m.b64.as<double>(), nv * KP);
            HIPC(hipGetLastError());
        });
        if (centered) eg.weighted_colsum(Q, nullptr, n, eg.sumq, nullptr);
        k_center_b<KP><<<blocks_for(nv * KP), 256, 0, cs>>>(acc.as<double>(), centered ? eg.mu : nullptr, eg.sumq, nv,
                                                            B64, B32);
        HIPC(hipGetLastError());
    }

    double stored_ss() {
        double t = 0.0;
        each_shard([&](StreamSlot &sl) { t += eg.stored_centered_ss(&sl.m, eg.mu); });
        return t;
    }
};

// pflog_pca on a streamed host CSR (pflog_pca_with_depth, cpu.rs:815-866,
// which streams its delta source): the alpha estimate by a streamed raw pass
// (estimate_alpha, pflog.rs:340-398), then the PCA over PFlog shards (each
// shard's delta and baseline from its own rows; the shards' b (1^T M) terms
// fill their rows of Y and their 1 (b_s^T Q_s) terms sum to 1 (b^T Q)).
void run_pflog_streamed(const scxpca_csr *x, const scxpca_pflog_params *pp, const scxpca_params *p,
                        int64_t shard_rows, scxpca_result *out, scxpca_pflog_fit *fit, double *baseline,
                        int device);

void run_pca_streamed(const scxpca_csr *x, const scxpca_params *p, int64_t shard_rows, scxpca_result *out,
                      int device) {
    run_pflog_streamed(x, nullptr, p, shard_rows, out, nullptr, nullptr, device);
}

void run_pflog_streamed(const scxpca_csr *x, const scxpca_pflog_params *pp, const scxpca_params *p,
                        int64_t shard_rows, scxpca_result *out, scxpca_pflog_fit *fit, double *baseline,
                        int device) {
    if (x->on_device) throw Error(SCXPCA_ERR_INVALID, "the streamed PCA reads a host CSR (on_device = 0)");
    if (shard_rows <= 0) throw Error(SCXPCA_ERR_INVALID, "shard_rows must be positive");
    if (p->precision != SCXPCA_PRECISION_FP32)
        throw Error(SCXPCA_ERR_INVALID, "the streamed PCA computes in FP32 (precision must be FP32)");
    if (x->n_obs < 0 || x->n_vars < 0) throw Error(SCXPCA_ERR_INVALID, "matrix shape must be non-negative");
    if (x->n_vars > 0xffffffffll) throw Error(SCXPCA_ERR_INVALID, "n_vars exceeds the u32 index range");
    if (x->n_obs > 0 && x->indptr[0] != 0) throw Error(SCXPCA_ERR_INVALID, "indptr must be non-decreasing");
    for (int64_t r = 0; r < x->n_obs; ++r)
        if (x->indptr[r + 1] < x->indptr[r]) throw Error(SCXPCA_ERR_INVALID, "indptr must be non-decreasing");
    DeviceGuard dg(device);
    scxpca_matrix shell;  // global panels and statistics; no CSR of its own
    shell.device = device;
    HIPC(hipStreamCreate(&shell.stream));
    shell.n_obs = x->n_obs;
    shell.n_vars = x->n_vars;
    shell.nnz = x->n_obs > 0 ? x->indptr[x->n_obs] : 0;
    shell.target = p->target_sum;
    shell.flags = transform_flags(p->target_sum, p->log1p);
    if (pp) {
        const double a = pp->alpha;
        if (!(a == 0.0 || (a > 0.0 && std::isfinite(a))))
            throw Error(SCXPCA_ERR_INVALID, "pflog: alpha must be positive and finite, got " + fmt_f32(static_cast<float>(a)));
        if (pp->delta_form != SCXPCA_PFLOG_DELTA_LAZY && pp->delta_form != SCXPCA_PFLOG_DELTA_F64)
            throw Error(SCXPCA_ERR_INVALID, "unknown PFlog delta form");
        if (!(pp->fallback_alpha > 0.0 && std::isfinite(pp->fallback_alpha)) || !std::isfinite(pp->mu_min))
            throw Error(SCXPCA_ERR_INVALID, "PFlog alpha options must be finite with a positive fallback");
        if (x->n_vars == 0) throw Error(SCXPCA_ERR_INVALID, "PFlog requires n_vars > 0 for centering");
        shell.pflog = true;
        shell.pf_params = *pp;
        shell.pf_fit = scxpca_pflog_fit{};
        if (a > 0.0) {
            shell.pf_fit.alpha = a;
            shell.pf_fit.pseudocount = 1.0 / (4.0 * a);
            shell.pf_fit.n_genes_used = -1;
        } else {
            shell.pf_fit.estimated = 1;
        }
    }
    validate(&shell, p);
    const int64_t k = pca_k(&shell, p);
    hipEvent_t e0, e1;
    HIPC(hipEventCreate(&e0));
    HIPC(hipEventCreate(&e1));
    HIPC(hipEventRecord(e0, shell.stream));
    auto drive = [&](auto &eg, auto &ops) {
        if (pp && shell.pf_fit.estimated) {  // the raw pass: column moments of the counts -> alpha
            ops.raw = true;
            ops.stats();
            std::vector<double> hs(x->n_vars), hq(x->n_vars);
            HIPC(hipMemcpyAsync(hs.data(), shell.sum.p, sizeof(double) * x->n_vars, hipMemcpyDeviceToHost, shell.stream));
            HIPC(hipMemcpyAsync(hq.data(), shell.sq.p, sizeof(double) * x->n_vars, hipMemcpyDeviceToHost, shell.stream));
            HIPC(hipStreamSynchronize(shell.stream));
            shell.pf_fit = estimate_alpha(hs, hq, x->n_obs, *pp);
            ops.raw = false;
        }
        if (pp) {
            shell.target = 4.0 * shell.pf_fit.alpha;
            ops.pflog = true;
            ops.baseline_out = baseline;
        }
        if (out) eg.run_with(ops, p, out);
        else ops.stats();  // fit and baseline only: one streamed pass over the PFlog shards
    };
    if (k <= 64) {
        Engine<64> eg(&shell, static_cast<int>(k));
        StreamOps<64> ops(eg, x, shard_rows);
        drive(eg, ops);
    } else if (k <= 128) {
        Engine<128> eg(&shell, static_cast<int>(k));
        StreamOps<128> ops(eg, x, shard_rows);
        drive(eg, ops);
    } else {
        Engine<128> eg(&shell, 128);
        StreamOps<128> ops(eg, x, shard_rows);
        WideEngine wide(&shell, eg, static_cast<int>(k));
        drive(wide, ops);
    }
    if (fit) *fit = shell.pf_fit;
    HIPC(hipEventRecord(e1, shell.stream));
    HIPC(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPC(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (out) {
        out->seconds_device = ms * 1e-3;
        out->resident = 0;
    }
}

// HBM a resident matrix and its PCA workspace take in this build: the CSR
// copy, row factors, X~ and its chunk index, the tile-CSC, the split offsets
// and per-tile partials while they stay under the metadata cap, the n x KP
// panels and the n_vars x KP workspace (scx_amd.accel._resident_bytes is the
// same model).
double resident_footprint(int64_t n, int64_t nv, int64_t nnz, int64_t k) {
    const double kp = k <= 64 ? 64.0 : static_cast<double>((k + 127) / 128 * 128);
    const double tiles = static_cast<double>((n + kTileRows - 1) / kTileRows);
    const double nsplit = static_cast<double>((nv + kCB - 1) / kCB);
    const double dn = static_cast<double>(n), dv = static_cast<double>(nv), dz = static_cast<double>(nnz);
    double b = 8.0 * (dn + 1) + 8.0 * dz;
    b += 8.0 * dn + 4.0 * dz + 2.0 * dz + 5.0 * (dz + kEntryPad);
    b += 4.0 * tiles * (dv + 1);
    const double split = 4.0 * dn * (nsplit + 1) + 16.0 * tiles * dv;
    if (split <= kMaxSplitBytes) b += split;
    b += 2 * 4.0 * dn * kp + 2 * 4.0 * dv * kp + 4 * 8.0 * dv * kp + 4 * 256.0 * dv * kp + 8.0 * 1024 * kp * kp;
    if (k > 128) b += 4.0 * std::max(dn, dv) * k + 16.0 * dv * k;  // column-major copies of the wide route
    return b;
}

// try_build_resident_csr's pre-flight (scx-gpu/src/gpu_pca_resident.rs:
// 54,71-74,87-135): resident when the footprint fits free HBM / 1.2 and
// SCX_GPU_PCA_RESIDENT is not "0".
bool resident_preflight(int64_t n, int64_t nv, int64_t nnz, int64_t k, size_t free_bytes) {
    const char *e = std::getenv("SCX_GPU_PCA_RESIDENT");
    if (e && std::strcmp(e, "0") == 0) return false;
    return resident_footprint(n, nv, nnz, k) <= static_cast<double>(free_bytes) / 1.2;
}

// Rows per shard of the streamed operator: two shards' structures within an
// eighth of the usable free HBM, a multiple of the tile height, at least one
// tile (scx_amd.accel.stream_shard_rows is the same rule).
int64_t streamed_shard_rows(int64_t n, int64_t nv, int64_t nnz, int64_t k, size_t free_bytes) {
    const double budget = static_cast<double>(free_bytes) / 1.2 / 8.0;
    const double per_row = 8.0 + 8.0 * static_cast<double>(k) +
                           21.0 * static_cast<double>(nnz) / static_cast<double>(std::max<int64_t>(n, 1)) +
                           4.0 * static_cast<double>((nv + kCB - 1) / kCB + 1) + 20.0 * static_cast<double>(nv + 1) / 256.0;
    int64_t rows = static_cast<int64_t>(budget / (2.0 * per_row)) / kTileRows * kTileRows;
    return std::max<int64_t>(kTileRows, std::min<int64_t>(rows, std::max<int64_t>(n, 1)));
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

scxpca_status scxpca_matrix_create(const scxpca_csr *x, double target_sum, int32_t log1p, int device,
                                   scxpca_matrix **out) {
    return guarded([&] {
        if (!x || !out) throw Error(SCXPCA_ERR_INVALID, "null argument");
        *out = nullptr;
        DeviceGuard dg(device);
        std::unique_ptr<scxpca_matrix> m(new scxpca_matrix());
        m->device = device;
        build_matrix(m.get(), x, target_sum, transform_flags(target_sum, log1p));
        *out = m.release();
    });
}

scxpca_status scxpca_matrix_create_pflog(const scxpca_csr *x, const scxpca_pflog_params *pp, int device,
                                         scxpca_matrix **out) {
    return guarded([&] {
        if (!x || !pp || !out) throw Error(SCXPCA_ERR_INVALID, "null argument");
        *out = nullptr;
        DeviceGuard dg(device);
        std::unique_ptr<scxpca_matrix> m(new scxpca_matrix());
        m->device = device;
        build_pflog(m.get(), x, *pp);
        *out = m.release();
    });
}

scxpca_status scxpca_matrix_pflog_fit(scxpca_matrix *m, scxpca_pflog_fit *fit, double *baseline) {
    return guarded([&] {
        if (!m) throw Error(SCXPCA_ERR_INVALID, "null matrix");
        if (!m->pflog) throw Error(SCXPCA_ERR_INVALID, "not a PFlog matrix");
        DeviceGuard dg(m->device);
        if (fit) *fit = m->pf_fit;
        if (baseline && m->n_obs > 0) {
            HIPC(hipMemcpyAsync(baseline, m->pfb.p, sizeof(double) * m->n_obs, hipMemcpyDeviceToHost, m->stream));
            HIPC(hipStreamSynchronize(m->stream));
        }
    });
}

scxpca_status scxpca_pflog_pca_streamed(const scxpca_csr *x, const scxpca_pflog_params *pp, const scxpca_params *p,
                                        int64_t shard_rows, scxpca_result *out, scxpca_pflog_fit *fit,
                                        double *baseline, int device) {
    return guarded([&] {
        if (!x || !pp || !p) throw Error(SCXPCA_ERR_INVALID, "null argument");
        run_pflog_streamed(x, pp, p, shard_rows, out, fit, baseline, device);
    });
}

scxpca_status scxpca_randomized_pca_streamed(const scxpca_csr *x, const scxpca_params *p, int64_t shard_rows,
                                             scxpca_result *out, int device) {
    return guarded([&] {
        if (!x || !p || !out) throw Error(SCXPCA_ERR_INVALID, "null argument");
        run_pca_streamed(x, p, shard_rows, out, device);
    });
}

scxpca_status scxpca_matrix_pflog_dense(scxpca_matrix *m, float *z, int32_t on_device) {
    return guarded([&] {
        if (!m || !z) throw Error(SCXPCA_ERR_INVALID, "null argument");
        if (!m->pflog) throw Error(SCXPCA_ERR_INVALID, "not a PFlog matrix");
        DeviceGuard dg(m->device);
        const int64_t n = m->n_obs, nv = m->n_vars;
        if (n == 0) return;
        DevBuf tmp;
        float *dz = z;
        if (!on_device) {
            tmp.ensure(sizeof(float) * n * nv);
            dz = tmp.as<float>();
        }
        k_pflog_fill<<<std::min(blocks_for(n * nv), 65536u), 256, 0, m->stream>>>(m->pfb.as<double>(), n, nv, dz);
        HIPC(hipGetLastError());
        k_pflog_scatter<<<std::min(blocks_for(n, 4), 65536u), 256, 0, m->stream>>>(m->indptr, m->idx, m->xt.as<float>(),
                                                                                 n, nv, dz);
        HIPC(hipGetLastError());
        if (!on_device) HIPC(hipMemcpyAsync(z, dz, sizeof(float) * n * nv, hipMemcpyDeviceToHost, m->stream));
        HIPC(hipStreamSynchronize(m->stream));
    });
}

scxpca_status scxpca_matrix_destroy(scxpca_matrix *m) {
    return guarded([&] {
        if (!m) return;
        DeviceGuard dg(m->device);
        (void)hipStreamSynchronize(m->stream);
        delete m;
    });
}

scxpca_status scxpca_matrix_pca(scxpca_matrix *m, const scxpca_params *p, scxpca_result *out) {
    return guarded([&] {
        if (!m || !p || !out) throw Error(SCXPCA_ERR_INVALID, "null argument");
        run_pca(m, p, out);
    });
}

scxpca_status scxpca_randomized_pca(const scxpca_csr *x, const scxpca_params *p, scxpca_result *out,
                                    int device) {
    return guarded([&] {
        if (!x || !p || !out) throw Error(SCXPCA_ERR_INVALID, "null argument");
        if (p->n_comps <= 0) throw Error(SCXPCA_ERR_INVALID, "n_components must be > 0");
        if (x->n_obs == 0 || x->n_vars == 0) throw Error(SCXPCA_ERR_INVALID, "matrix must be non-empty");
        DeviceGuard dg(device);
        const bool host = !x->on_device;
        const int64_t nnz = host ? x->indptr[x->n_obs] : 0;
        const int64_t k = std::min<int64_t>(std::min<int64_t>(p->n_comps + p->n_oversamples, x->n_vars), x->n_obs);
        size_t free_b = 0, total_b = 0;
        HIPC(hipMemGetInfo(&free_b, &total_b));
        if (host && p->precision == SCXPCA_PRECISION_FP32 &&
            !resident_preflight(x->n_obs, x->n_vars, nnz, k, free_b)) {
            // declined by the pre-flight: the streamed operator (gpu_pca.rs:566-590)
            run_pca_streamed(x, p, streamed_shard_rows(x->n_obs, x->n_vars, nnz, k, free_b), out, device);
            out->resident = 0;
            return;
        }
        std::unique_ptr<scxpca_matrix> m(new scxpca_matrix());
        m->device = device;
        try {
            build_matrix(m.get(), x, p->target_sum, transform_flags(p->target_sum, p->log1p));
        } catch (const Error &e) {
            // a runtime failure while building the resident matrix is a decline
            // (decline_on_runtime_failure, scx-gpu/src/error.rs:141-162); input
            // defects propagate, as does anything the streamed route cannot take
            const bool runtime = e.status == SCXPCA_ERR_OOM || e.status == SCXPCA_ERR_DEVICE;
            if (!runtime || !host || p->precision != SCXPCA_PRECISION_FP32) throw;
            m.reset();
            (void)hipGetLastError();  // the failed allocation's error must not surface at a later launch check
            HIPC(hipMemGetInfo(&free_b, &total_b));
            run_pca_streamed(x, p, streamed_shard_rows(x->n_obs, x->n_vars, nnz, k, free_b), out, device);
            out->resident = 0;
            return;
        }
        run_pca(m.get(), p, out);
        out->resident = 1;
    });
}

scxpca_status scxpca_matrix_stats(scxpca_matrix *m, double *sums, double *sum_sq, uint32_t *col_nnz,
                                  double *row_factor) {
    return guarded([&] {
        if (!m) throw Error(SCXPCA_ERR_INVALID, "null matrix");
        DeviceGuard dg(m->device);
        if (m->n_obs == 0 || m->n_vars == 0) throw Error(SCXPCA_ERR_INVALID, "matrix must be non-empty");
        ensure_stats(m);
        if (sums) HIPC(hipMemcpyAsync(sums, m->sum.p, sizeof(double) * m->n_vars, hipMemcpyDeviceToHost, m->stream));
        if (sum_sq) HIPC(hipMemcpyAsync(sum_sq, m->sq.p, sizeof(double) * m->n_vars, hipMemcpyDeviceToHost, m->stream));
        if (col_nnz) HIPC(hipMemcpyAsync(col_nnz, m->cnt.p, sizeof(uint32_t) * m->n_vars, hipMemcpyDeviceToHost, m->stream));
        if (row_factor) {
            if (m->flags & kXfScale)
                for (int64_t r = 0; r < m->n_obs; ++r) row_factor[r] = m->target;
            else if (m->flags & kXfNorm)
                HIPC(hipMemcpyAsync(row_factor, m->rowf.p, sizeof(double) * m->n_obs, hipMemcpyDeviceToHost, m->stream));
            else
                for (int64_t r = 0; r < m->n_obs; ++r) row_factor[r] = 1.0;
        }
        HIPC(hipStreamSynchronize(m->stream));
    });
}

scxpca_status scxpca_matrix_spmm(scxpca_matrix *m, int32_t transpose, int64_t k, const double *in,
                                 const double *means, int32_t bf16, double *out) {
    return guarded([&] {
        if (!m || !in || !out) throw Error(SCXPCA_ERR_INVALID, "null argument");
        if (k <= 0 || k > 128) throw Error(SCXPCA_ERR_INVALID, "k must be in 1..128");
        if (m->n_obs == 0 || m->n_vars == 0) throw Error(SCXPCA_ERR_INVALID, "matrix must be non-empty");
        DeviceGuard dg(m->device);
        if (k <= 64) stage_spmm<64>(m, transpose, k, in, means, bf16, out);
        else stage_spmm<128>(m, transpose, k, in, means, bf16, out);
    });
}

scxpca_status scxpca_qr(const double *a, int64_t rows, int64_t k, int32_t qr_method, int device,
                        double *q_out) {
    return guarded([&] {
        if (!a || !q_out) throw Error(SCXPCA_ERR_INVALID, "null argument");
        if (k <= 0 || k > 128 || rows < k) throw Error(SCXPCA_ERR_INVALID, "need rows >= k and 1 <= k <= 128");
        DeviceGuard dg(device);
        if (k <= 64) stage_qr<64>(rows, k, a, qr_method, q_out);
        else stage_qr<128>(rows, k, a, qr_method, q_out);
    });
}

scxpca_status scxpca_gaussian(uint64_t rows, uint64_t cols, uint64_t seed, double *out) {
    return guarded([&] {
        if (!out && rows * cols) throw Error(SCXPCA_ERR_INVALID, "null output");
        gaussian_matrix(rows, cols, seed, out);
    });
}

scxpca_status scxpca_matrix_refresh(scxpca_matrix *m) {
    return guarded([&] {
        if (!m) throw Error(SCXPCA_ERR_INVALID, "null matrix");
        DeviceGuard dg(m->device);
        refresh(m);
    });
}

scxpca_status scxpca_gaussian_device(uint64_t rows, uint64_t cols, uint64_t seed, int device, double *out) {
    return guarded([&] {
        if (!out && rows * cols > 0) throw Error(SCXPCA_ERR_INVALID, "null output");
        if (cols == 0 || cols > 128) throw Error(SCXPCA_ERR_INVALID, "cols must be in 1..128");
        DeviceGuard dg(device);
        hipStream_t s;
        HIPC(hipStreamCreate(&s));
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamDestroy(s); }
        } sg{s};
        OmegaScratch sc;
        DevBuf b64;
        b64.ensure(sizeof(double) * std::max<uint64_t>(rows * cols, 1));
        device_omega<128>(s, rows, static_cast<int>(cols), seed, nullptr, b64.as<double>(), sc);
        HIPC(hipMemcpyAsync(out, b64.p, sizeof(double) * rows * cols, hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
    });
}

scxpca_status scxpca_project_forward(scxpca_matrix *m, const float *w, int64_t k, const float *mean, float *z,
                                     void *stream) {
    return guarded([&] {
        if (!m || !w || !z) throw Error(SCXPCA_ERR_INVALID, "null argument");
        if (k <= 0 || k > 128) throw Error(SCXPCA_ERR_INVALID, "k must be in 1..128");
        DeviceGuard dg(m->device);
        if (k <= 64) project<64>(m, false, w, k, mean, z, static_cast<hipStream_t>(stream));
        else project<128>(m, false, w, k, mean, z, static_cast<hipStream_t>(stream));
    });
}

scxpca_status scxpca_project_backward(scxpca_matrix *m, const float *dz, int64_t k, const float *mean, float *dw,
                                      void *stream) {
    return guarded([&] {
        if (!m || !dz || !dw) throw Error(SCXPCA_ERR_INVALID, "null argument");
        if (k <= 0 || k > 128) throw Error(SCXPCA_ERR_INVALID, "k must be in 1..128");
        DeviceGuard dg(m->device);
        if (k <= 64) project<64>(m, true, dz, k, mean, dw, static_cast<hipStream_t>(stream));
        else project<128>(m, true, dz, k, mean, dw, static_cast<hipStream_t>(stream));
    });
}

scxpca_status scxpca_project_backward_noalloc(scxpca_matrix *m, const float *dz, int64_t k, const float *mean,
                                              float *dw, void *stream) {
    return guarded([&] {
        if (!m || !dz || !dw) throw Error(SCXPCA_ERR_INVALID, "null argument");
        if (k <= 0 || k > 128) throw Error(SCXPCA_ERR_INVALID, "k must be in 1..128");
        DeviceGuard dg(m->device);
        if (k <= 64) project_backward_noalloc<64>(m, dz, k, mean, dw, static_cast<hipStream_t>(stream));
        else project_backward_noalloc<128>(m, dz, k, mean, dw, static_cast<hipStream_t>(stream));
    });
}

scxpca_status scxpca_matrix_profile(scxpca_matrix *m, int32_t enable, double *ms_total, int64_t *launches,
                                    int32_t n_stages) {
    return guarded([&] {
        if (!m) throw Error(SCXPCA_ERR_INVALID, "null matrix");
        DeviceGuard dg(m->device);
        if (ms_total || launches) {
            HIPC(hipStreamSynchronize(m->stream));
            for (int i = 0; i < n_stages; ++i) {
                if (ms_total) ms_total[i] = 0.0;
                if (launches) launches[i] = 0;
            }
            for (auto &mk : m->marks) {
                if (mk.stage < 0 || mk.stage >= n_stages) continue;
                float ms = 0.f;
                HIPC(hipEventElapsedTime(&ms, mk.a, mk.b));
                if (ms_total) ms_total[mk.stage] += ms;
                if (launches) launches[mk.stage] += 1;
            }
        }
        m->clear_marks();
        m->profiling = enable != 0;
    });
}

scxpca_status scxpca_matrix_pca_covariance(scxpca_matrix *m, const scxpca_params *p, scxpca_result *out) {
    return guarded([&] {
        if (!m || !p || !out) throw Error(SCXPCA_ERR_INVALID, "null argument");
        run_pca_covariance(m, p, out);
    });
}

scxpca_status scxpca_matrix_profile_stages(scxpca_matrix *m, uint32_t stage_mask) {
    return guarded([&] {
        if (!m) throw Error(SCXPCA_ERR_INVALID, "null matrix");
        m->prof_mask = stage_mask ? stage_mask : ~0u;
    });
}

int scxpca_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

scxpca_status scxpca_device_info(int device, char *name, size_t name_len, int64_t *total_mem, int32_t *n_cu) {
    return guarded([&] {
        DeviceGuard dg(device);
        hipDeviceProp_t prop;
        HIPC(hipGetDeviceProperties(&prop, device));
        if (name && name_len) {
            std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
        }
        if (total_mem) *total_mem = static_cast<int64_t>(prop.totalGlobalMem);
        if (n_cu) *n_cu = prop.multiProcessorCount;
    });
}

size_t scxpca_last_error(char *buf, size_t len) {
    const size_t n = g_last_error.size();
    if (buf && len) {
        const size_t c = std::min(n, len - 1);
        std::memcpy(buf, g_last_error.data(), c);
        buf[c] = '\0';
    }
    return n;
}

}  // extern "C"

// On-GPU shard decode (include/scxshard.h).
#include "shard_decode.h"
#include "shard_encode.h"
