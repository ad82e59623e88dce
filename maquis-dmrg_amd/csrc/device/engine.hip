// C-ABI implementation: contexts, device-resident boundaries, sigma plans, boundary steps,
// reshapes and the device-resident Jacobi-Davidson driver.
//
// A schedule (the reference's ScheduleNew tables) is lowered into three device task tables:
//   T stage  : grouped GEMM, T_ti = psi_slab * R_cohort (or L_b * ket slab for the left step)
//   S stage  : gather/scale regions, S = sum alpha * T slab
//   final    : grouped K-segmented GEMM, sigma[lb_out] = sum_cohorts L_cohort * S (owner-computes,
//              no atomics), or the prop_l / prop_r products of the boundary steps.
// All operand addresses are checked against their buffers on the host before any launch.
#include <hip/hip_runtime.h>
#include <hipblaslt/hipblaslt.h>
#include <hipblaslt/hipblaslt-ext.hpp>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../host/host_types.hpp"
#include "kernels.hpp"

using namespace mq;

namespace {
thread_local std::string g_err;

const bool g_trace = trace_enabled();

// The block-sparse S stage (FP64 / FP32) is on by default; MQ_S_DENSE=1 makes plans created while
// it is set run every group's S stage as one dense library GEMM instead.
bool s_sparse_enabled() {
    const char* e = std::getenv("MQ_S_DENSE");
    return !(e && e[0] == '1');
}

#define HIPCHK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess)                                                                      \
            ::mq::fail(MQ_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));               \
    } while (0)

template <class F>
int guarded(F&& f) {
    try {
        f();
        return MQ_OK;
    } catch (const Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return MQ_ERR_NOMEM;
    } catch (const std::exception& e) {
        g_err = e.what();
        return MQ_ERR_INTERNAL;
    }
}

// Device allocation owned by RAII.
// Releases the workspace pools of every context (set up with the contexts): an allocation that
// fails retries once after it.
static void (*g_release_pools)() = nullptr;

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t n) {
        release();
        if (n == 0) return;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess && g_release_pools) {
            (void)hipGetLastError();
            g_release_pools();
            e = hipMalloc(&p, n);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            fail(MQ_ERR_NOMEM, "hipMalloc of " + std::to_string(n) + " bytes failed");
        }
        bytes = n;
    }
    void ensure(size_t n) {
        if (n > bytes) alloc(n);
    }
    void swap(DevBuf& o) {
        std::swap(p, o.p);
        std::swap(bytes, o.bytes);
    }
    // allocation cleared to zero bits: library GEMMs called with beta = 0 may still read their C
    // operand (0 * NaN of stale memory is NaN), so workspaces never hold non-finite garbage
    void alloc_zeroed(size_t n, hipStream_t st) {
        alloc((n + 7) / 8 * 8);
        if (!p) return;
        HIPCHK(launch_fill(st, static_cast<double*>(p), bytes / 8, 0.0));
        HIPCHK(hipStreamSynchronize(st));
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

template <class T>
void upload(DevBuf& b, const std::vector<T>& v, hipStream_t st) {
    b.alloc(v.size() * sizeof(T));
    if (!v.empty()) {
        HIPCHK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
    }
}

inline uint64_t align32(uint64_t x) { return round_up(x, 32); }

}  // namespace

// ============================================================================ context
struct mq_plan;

// A captured batch of applications (mq_sigma_apply_batched): one graph over the operand slots of
// its plans, with one branch per plan.
struct BatchGraph {
    hipGraphExec_t exec = nullptr;
    bool seen = false;
};

// One library GEMM of the reduced-precision S stage through hipBLASLt: descriptors and the
// algorithm chosen for its shape (the library's first heuristic, or the fastest of its candidates
// after mq_plan_autotune).
struct LtGemm {
    bool ready = false, ok = false;
    bool chosen = false;   // algorithm from the exact-shape or shape-class cache (no retiming)
    hipblasLtMatmulDesc_t op = nullptr;
    hipblasLtMatrixLayout_t a = nullptr, b = nullptr, c = nullptr;
    hipblasLtMatmulAlgo_t algo{};
    void release() {
        if (a) (void)hipblasLtMatrixLayoutDestroy(a);
        if (b) (void)hipblasLtMatrixLayoutDestroy(b);
        if (c) (void)hipblasLtMatrixLayoutDestroy(c);
        if (op) (void)hipblasLtMatmulDescDestroy(op);
        a = b = c = nullptr;
        op = nullptr;
        ready = ok = false;
    }
};
constexpr size_t LT_WORKSPACE = size_t(64) << 20;

struct mq_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr;   // boundary prefetches (host -> device), overlapping `stream`
    rocblas_handle blas = nullptr;   // plain dense GEMMs of the grouped S stage, on `stream`
    hipblasLtHandle_t lt = nullptr;  // reduced-precision S stage GEMMs
    DevBuf lt_ws;
    // one hipBLASLt workspace per side stream: the instances of a batch run their library GEMMs
    // concurrently, and a shared workspace would be a race for algorithms that use it
    std::vector<std::unique_ptr<DevBuf>> lt_ws_side;
    void* lt_workspace(hipStream_t s) const {
        for (size_t k = 0; k < side.size() && k < lt_ws_side.size(); ++k)
            if (side[k] == s) return lt_ws_side[k]->p;
        return lt_ws.p;
    }
    // hipBLASLt algorithms timed by mq_plan_autotune, keyed by (precision, M, N, K), and the
    // library solution index of each winner keyed by its shape class (precision and half-octave
    // buckets of M, N, K): a later GEMM of another shape in the same class adopts that solution
    // when the library confirms it supports the shape
    std::map<std::array<uint64_t, 4>, hipblasLtMatmulAlgo_t> lt_tuned;
    std::map<std::array<int, 4>, int> lt_class;
    static std::array<int, 4> lt_class_key(int mi, uint64_t M, uint64_t N, uint64_t K) {
        auto h = [](uint64_t x) { return int(std::floor(2.0 * std::log2(double(std::max<uint64_t>(x, 1))))); };
        return {mi, h(M), h(N), h(K)};
    }
    // workspaces of destroyed plans, reused by the next plans, one slot per kind: a sweep creates
    // one plan per site, and allocating and releasing tens of GB per site costs more than the
    // site's products.  A slot keeps the largest buffer returned to it; a request it cannot serve
    // allocates 25% headroom, so that the slightly larger plans of the next sites still fit.
    enum PoolKind { POOL_WS, POOL_PART, POOL_WS32, POOL_BF, POOL_KINDS };
    DevBuf pool[POOL_KINDS];
    // pinned host staging of the dense MPO matrices W_D of the plan (or step chunk) being lowered;
    // grows, and is reused by the next lowering once the previous upload has been synchronized
    double* wd_host = nullptr;
    size_t wd_cap = 0;
    double* wd_stage(size_t n, bool zero = true) {
        if (n > wd_cap) {
            if (wd_host) (void)hipHostFree(wd_host);
            wd_host = nullptr;
            wd_cap = 0;
            const size_t cap = (n + (size_t(1) << 20) - 1) & ~((size_t(1) << 20) - 1);
            HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&wd_host), cap * sizeof(double), hipHostMallocDefault));
            wd_cap = cap;
        }
        if (!zero) return wd_host;
        const int64_t chunk = int64_t(1) << 18;
#pragma omp parallel for schedule(static)
        for (int64_t o = 0; o < int64_t(n); o += chunk)
            std::memset(wd_host + o, 0, size_t(std::min<int64_t>(chunk, int64_t(n) - o)) * sizeof(double));
        return wd_host;
    }
    void pool_put(PoolKind kind, DevBuf& b) {
        if (!b.p) return;
        if (pool[kind].bytes < b.bytes) pool[kind].swap(b);
        b.release();
    }
    void pool_clear() {
        for (auto& b : pool) b.release();
    }
    // a workspace of at least n bytes: the slot's buffer if it is large enough, else a new
    // allocation (with headroom; exactly n bytes after the pools are released when that fails).
    // The first n bytes are cleared to zero, except that a reused buffer is taken as it is when
    // zero_reused is false: callers that pass false read no workspace element before writing it
    // in the same application other than through masks (grouped sigma plans and steps with the
    // block-sparse S stage: T_D slabs no T task writes are cleared per application by zero
    // chunks; tile rows and keys past the data feed only outputs that are not stored or are
    // zeroed).  MQ_POISON_WS=1 fills such reused buffers with NaN instead (tests).
    void take(PoolKind kind, DevBuf& dst, size_t n, bool zero_reused = true) {
        n = (n + 7) / 8 * 8;
        dst.release();
        if (pool[kind].p && pool[kind].bytes >= n) {
            dst.swap(pool[kind]);
            if (!zero_reused) {
                if (std::getenv("MQ_POISON_WS"))
                    HIPCHK(launch_fill(stream, static_cast<double*>(dst.p), n / 8, std::numeric_limits<double>::quiet_NaN()));
                return;
            }
        } else {
            pool[kind].release();
            const size_t want = (n + n / 4 + (size_t(2) << 20) - 1) & ~((size_t(2) << 20) - 1);
            if (trace_enabled()) std::fprintf(stderr, "[mq] workspace pool miss: %.2f GB allocated\n", 1e-9 * double(want));
            if (hipMalloc(&dst.p, want) == hipSuccess) {
                dst.bytes = want;
            } else {
                (void)hipGetLastError();
                dst.p = nullptr;
                pool_clear();
                dst.alloc(n);
            }
        }
        HIPCHK(launch_fill(stream, static_cast<double*>(dst.p), n / 8, 0.0));
    }
    DevBuf red_partial, red_result;
    // batched reductions (mq_vec_multi_dot, Jacobi-Davidson): k * RED_BLOCKS partial sums, k results
    DevBuf mdot_partial, mdot_result;
    int mdot_cap = 0;
    void mdot_reserve(int k) {
        if (k <= mdot_cap) return;
        mdot_partial.alloc(size_t(k) * RED_BLOCKS * sizeof(double));
        mdot_result.alloc(size_t(k) * sizeof(double));
        mdot_cap = k;
    }
    double* noalloc_scratch = nullptr;   // this device's static scratch (mq_vec_multi_dot_noalloc)
    DevBuf step_ws;   // grows; reused by boundary steps
    // batched applications: side streams (default priority) the instances of a batch fork onto,
    // and the captured graphs keyed by (plans, precision)
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> side_ev;
    // one rocBLAS handle per side stream, bound to it once: a handle owns one device workspace,
    // which library GEMMs of concurrent instances would otherwise share
    std::vector<rocblas_handle> side_blas;
    rocblas_handle blas_for(hipStream_t s) const {
        for (size_t k = 0; k < side.size() && k < side_blas.size(); ++k)
            if (side[k] == s) return side_blas[k];
        return blas;
    }
    hipEvent_t fork_ev = nullptr;
    std::map<std::pair<std::vector<const mq_plan*>, int>, BatchGraph> batches;
    // autotuned tile policies, keyed by (stage, precision, grouped, log2 flop histogram over size classes)
    std::map<std::vector<int>, std::array<int, 5>> tuned;
    void drop_batches() {
        for (auto& kv : batches)
            if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
        batches.clear();
    }
};

#define LTCHK(x)                                                                                              \
    do {                                                                                                      \
        hipblasStatus_t s_ = (x);                                                                             \
        if (s_ != HIPBLAS_STATUS_SUCCESS)                                                                     \
            fail(MQ_ERR_HIP, std::string("hipBLASLt: status ") + std::to_string(int(s_)) + " at " #x);        \
    } while (0)

#define BLASCHK(x)                                                                                            \
    do {                                                                                                      \
        rocblas_status s_ = (x);                                                                              \
        if (s_ != rocblas_status_success)                                                                     \
            fail(MQ_ERR_HIP, std::string("rocBLAS: ") + rocblas_status_to_string(s_) + " at " #x);         \
    } while (0)

struct mq_boundary {
    mq_ctx* ctx = nullptr;
    BoundaryIndex index;
    std::vector<uint64_t> cohort_off;
    uint64_t total = 0;
    DevBuf data;
    // Residency (the reference's storage::gpu / storage::disk, utils/storage.h:87-200,271-539).  A
    // boundary's data never changes after it is made, so a host or disk copy, once written, stays
    // valid: evicting again only frees the device copy.
    double* host = nullptr;               // pinned host copy (total doubles), or null
    std::string disk;                     // spill file holding the data, or empty
    bool resident = true;                 // the device copy is allocated and (after `arrived`) valid
    hipEvent_t arrived = nullptr;         // end of the last prefetch on ctx->copy_stream
    mutable int plans = 0;                // live sigma plans that read the boundary (no eviction)
    // storage for exported descriptors
    mutable std::vector<mq_charge> bra_c, ket_c;
    double* ptr() const {
        MQ_REQUIRE(resident, "boundary not resident on the device (mq_boundary_prefetch)");
        return data.as<double>();
    }
    ~mq_boundary() {
        if (arrived) (void)hipEventSynchronize(arrived), (void)hipEventDestroy(arrived);
        if (host) (void)hipHostFree(host);
        if (!disk.empty()) std::remove(disk.c_str());
    }
    void layout() {
        cohort_off.resize(index.n_cohorts());
        total = 0;
        for (uint32_t ci = 0; ci < index.n_cohorts(); ++ci) {
            cohort_off[ci] = total;
            total += align32(index.cohort_size(ci));
        }
        if (total == 0) total = 32;
    }
};

struct mq_schedule {
    Schedule s;
    uint64_t flops = 0;
    mutable std::vector<mq_charge> bra_c, ket_c;
};

// Lowered stage tables (device) + the host copies used for validation.
// One GEMM stage on the device: the lowering's tasks split by tile configuration (one launch
// each), sharing one segment table.
// Row-size classes of GEMM tasks and the tile configurations the autotuner tries for each.  The
// class of a task is fixed by its M; -1 in a policy means the default choice of gemm_pick_cfg.
constexpr int N_SIZE_CLASS = 5;
inline int size_class(uint32_t M) { return M <= 16 ? 0 : M <= 32 ? 1 : M <= 64 ? 2 : M <= 128 ? 3 : 4; }
const std::vector<int> CLASS_CANDIDATES[N_SIZE_CLASS] = {
    {GEMM_CFG_16x128, GEMM_CFG_64x64},
    {GEMM_CFG_32x128, GEMM_CFG_64x64, GEMM_CFG_16x128},
    {GEMM_CFG_64x64, GEMM_CFG_32x128, GEMM_CFG_128x128},
    {GEMM_CFG_64x64, GEMM_CFG_128x128, GEMM_CFG_32x128},
    {GEMM_CFG_128x128, GEMM_CFG_64x64}};
struct TilePolicy {
    int cls[N_SIZE_CLASS] = {-1, -1, -1, -1, -1};
};

struct GemmStage {
    std::vector<GemmTask> all;                     // the lowering's tasks (rebuilt per policy)
    TilePolicy policy;
    std::vector<GemmTask> cfg[GEMM_N_CFG];
    uint32_t tiles[GEMM_N_CFG] = {};
    std::vector<uint32_t> tile_task[GEMM_N_CFG];   // launches with many multi-tile tasks: task of each tile
    DevBuf d_tasks[GEMM_N_CFG], d_tile_task[GEMM_N_CFG];
    DevBuf d_segs;
    // per tile configuration (128x128 and 64x64): every task of it suits the lean stacked T-stage
    // kernel (t_stage.hip) / the lean column-unit final-stage kernel (f_stage.hip)
    bool stacked[GEMM_N_CFG] = {}, units[GEMM_N_CFG] = {};

    // Flops per size class: the block-size histogram the autotuner's decisions are keyed on.
    std::vector<uint64_t> class_flops(const std::vector<GemmSeg>& segs) const {
        std::vector<uint64_t> f(N_SIZE_CLASS, 0);
        for (const auto& t : all)
            for (uint32_t s = t.seg0; s < t.seg1; ++s) f[size_class(t.M)] += 2ull * t.M * t.N * segs[s].K;
        return f;
    }
    void build(const std::vector<GemmTask>& tasks) {
        all = tasks;
        rebuild();
    }
    void rebuild() {
        for (auto& c : cfg) c.clear();
        for (const auto& t : all) {
            int c = policy.cls[size_class(t.M)];
            if (c < 0) c = gemm_pick_cfg(t.M, t.N);
            // the split-K 16x16 kernel covers one 16x16 tile; column units, row units and remapped
            // A are implemented in the LDS-tiled kernels only
            if (c == GEMM_CFG_16x16 && (t.M > 16 || t.N > 16 || t.colunit != MAP_NONE || t.rowunit != MAP_NONE))
                c = GEMM_CFG_16x128;
            cfg[c].push_back(t);
        }
        for (int c = 0; c < GEMM_N_CFG; ++c) {
            uint64_t tile = 0;
            for (auto& t : cfg[c]) {
                t.tiles_n = (t.N + GEMM_CFG_BN[c] - 1) / GEMM_CFG_BN[c];
                t.tile_begin = uint32_t(tile);
                tile += uint64_t((t.M + GEMM_CFG_BM[c] - 1) / GEMM_CFG_BM[c]) * t.tiles_n;
                MQ_REQUIRE(tile < (1ull << 31), "too many GEMM tiles in one launch");
            }
            tiles[c] = uint32_t(tile);
            tile_task[c].clear();
            if (cfg[c].size() > 64 && tile != cfg[c].size() && tile < (1ull << 28)) {
                tile_task[c].reserve(tile);
                for (uint32_t t = 0; t < cfg[c].size(); ++t) {
                    const GemmTask& k = cfg[c][t];
                    const uint64_t n = uint64_t((k.M + GEMM_CFG_BM[c] - 1) / GEMM_CFG_BM[c]) * k.tiles_n;
                    tile_task[c].insert(tile_task[c].end(), n, t);
                }
            }
        }
    }
    void upload_all(const std::vector<GemmSeg>& segs, hipStream_t st) {
        static const bool generic = std::getenv("MQ_T_GENERIC") != nullptr;
        static const bool generic_f = std::getenv("MQ_F_GENERIC") != nullptr;
        for (int c = 0; c < GEMM_N_CFG; ++c) {
            const bool lean = c == GEMM_CFG_128x128 || c == GEMM_CFG_64x64;
            units[c] = lean && !generic_f && !cfg[c].empty();
            stacked[c] = lean && !generic && !cfg[c].empty();
            for (const auto& t : cfg[c]) {
                bool ok = t.rowunit == MAP_NONE && t.colunit != MAP_NONE && !t.beta && t.m_off == 0 &&
                          t.cmap == MAP_NONE && t.c_base != BASE_WS && t.M > 0 && t.N > 0 && t.seg1 > t.seg0;
                for (uint32_t k = t.seg0; ok && k < t.seg1; ++k) {
                    const GemmSeg& g = segs[k];
                    ok = g.units != MAP_NONE && g.bmap == MAP_NONE && !g.trans_b && g.bbh > 0 &&
                         (g.amap == MAP_NONE ? !g.trans_a : g.trans_a && g.ascale != MAP_NONE);
                }
                units[c] = units[c] && ok;
            }
            for (const auto& t : cfg[c]) {
                bool ok = t.rowunit != MAP_NONE && t.seg1 == t.seg0 + 1 && !t.beta && t.M > 0 && t.N > 0;
                if (ok) {
                    const GemmSeg& g = segs[t.seg0];
                    ok = g.amap == MAP_NONE && g.bmap == MAP_NONE && g.units == MAP_NONE && !g.trans_a;
                }
                stacked[c] = stacked[c] && ok;
            }
        }
        upload(d_segs, segs, st);
        for (int c = 0; c < GEMM_N_CFG; ++c) {
            upload(d_tasks[c], cfg[c], st);
            upload(d_tile_task[c], tile_task[c], st);
        }
    }
    void run(hipStream_t st, int mode, const char* types, Bases b, GemmTables tb) const {
        for (int c = 0; c < GEMM_N_CFG; ++c) {
            if (!tiles[c]) continue;
            const uint32_t* tt = tile_task[c].empty() ? nullptr : d_tile_task[c].as<uint32_t>();
            // MODE_BF16X3 reaches the lean kernels (split products); the generic kernel runs it in FP32
            const int gmode = mode == MODE_BF16X3 ? MODE_F32 : mode;
            const bool f64 = mode == MODE_F64, ddf = !f64 && std::strcmp(types, "ddf") == 0;
            const bool dfd = !f64 && std::strcmp(types, "dfd") == 0;
            if (stacked[c] && (f64 || ddf))
                HIPCHK(launch_t_stack(st, c, mode, ddf, d_tasks[c].as<GemmTask>(), uint32_t(cfg[c].size()),
                                      d_segs.as<GemmSeg>(), tiles[c], b, tb, tt));
            else if (units[c] && (f64 || dfd))
                HIPCHK(launch_f_unit(st, c, mode, d_tasks[c].as<GemmTask>(), uint32_t(cfg[c].size()),
                                     d_segs.as<GemmSeg>(), tiles[c], b, tb, tt));
            else
                HIPCHK(launch_gemm(st, c, gmode, types, d_tasks[c].as<GemmTask>(), uint32_t(cfg[c].size()),
                                   d_segs.as<GemmSeg>(), tiles[c], b, tb, tt));
        }
    }
    uint64_t device_bytes() const {
        uint64_t n = d_segs.bytes;
        for (const auto& d : d_tasks) n += d.bytes;
        for (const auto& d : d_tile_task) n += d.bytes;
        return n;
    }
};

// A list of gather regions with its own source table: region dst = sum_t alpha_t src_t.
struct GatherSet {
    std::vector<GatherRegion> regions, small;
    std::vector<uint64_t> src_off;
    std::vector<double> alpha;
    uint32_t chunks = 0;
    DevBuf d_regions, d_small, d_src_off, d_alpha;

    bool empty() const { return regions.empty() && small.empty(); }
    void finish() {
        std::vector<GatherRegion> big;
        for (const auto& r : regions) (uint64_t(r.R) * r.C <= uint64_t(GATHER_SMALL) ? small : big).push_back(r);
        regions.swap(big);
        uint32_t ch = 0;
        for (auto& r : regions) {
            r.chunk_begin = ch;
            uint64_t c = std::max<uint64_t>(1, (uint64_t(r.R) * r.C + GATHER_CHUNK - 1) / GATHER_CHUNK);
            MQ_REQUIRE(uint64_t(ch) + c < (1ull << 31), "too many gather chunks");
            ch += uint32_t(c);
        }
        chunks = ch;
    }
    void upload_all(hipStream_t st) {
        upload(d_regions, regions, st);
        upload(d_small, small, st);
        upload(d_src_off, src_off, st);
        upload(d_alpha, alpha, st);
    }
    void run(hipStream_t st, double* src, double* dst) const {
        if (empty()) return;
        HIPCHK(launch_gather(st, d_regions.as<GatherRegion>(), uint32_t(regions.size()), d_src_off.as<uint64_t>(),
                             d_alpha.as<double>(), chunks, src, dst));
        HIPCHK(launch_gather_small(st, d_small.as<GatherRegion>(), uint32_t(small.size()), d_src_off.as<uint64_t>(),
                                   d_alpha.as<double>(), src, dst));
    }
    uint64_t device_bytes() const { return d_regions.bytes + d_small.bytes + d_src_off.bytes + d_alpha.bytes; }
};

struct StagePlan {
    std::vector<GemmTask> t_tasks, f_tasks;
    std::vector<GemmSeg> t_segs, f_segs;
    std::vector<GatherRegion> regions;      // large regions, chunked
    std::vector<GatherRegion> small;        // regions of <= GATHER_SMALL elements, one wavefront each
    std::vector<uint64_t> src_off;
    std::vector<double> alpha;
    std::vector<uint32_t> maps;             // column / row block remaps of the GEMM stages
    std::vector<GemmUnit> units;            // column units of B (grouped final stage)
    std::vector<double> scales;             // block scales of remapped transposed A
    std::vector<GemmRow> rows;              // row units of stacked tasks (grouped T stage)
    // grouped path: T_D ranges no T task writes, cleared at the start of an application (in the
    // workspace, or in the three A' copies of the bf16 buffer for BF16x3)
    std::vector<ZeroChunk> zero_ws, zero_bf;
    bool split_t = false;       // all T tasks stacked: BF16x3 T stage writes A' directly
    uint64_t bf_a_elems = 0;    // A' part of the bf16 buffer
    GemmStage gT, gF;   // device form of the T and final GEMM stages
    // adjoint (plans created with MQ_SIGMA_ADJOINT): U_D = L^T psi in the S_D layout, as stacked tasks
    std::vector<GemmTask> u_tasks;
    std::vector<GemmSeg> u_segs;
    GemmStage gU;
    // grouped path: split-K partial sums of the final stage (FP64, base BASE_BRA) and their
    // reduction into sigma
    GatherSet red;
    uint64_t part_elems = 0;
    // grouped path: the S stage S_D = T_D W_D as plain dense GEMMs (rocBLAS), one per group, on the
    // workspace (element offsets).  BF16x3 uses one bf16 GEMM per group with K' = 3K and FP32
    // accumulation: A' = [A_hi A_lo A_hi] (split per application from T_D) and B' = [B_lo; B_hi; B_hi]
    // (split once from W_D), at bfa_off / bfb_off of the bf16 buffer.
    struct LibGemm {
        uint64_t a_off, b_off, c_off;
        uint32_t M, N, K, lda, ldb, ldc;
        uint64_t bfa_off, bfb_off;
        bool sparse = false;   // the group's S stage runs in the block-sparse kernel (bsr) instead
    };
    std::vector<LibGemm> lib;
    // grouped path: the S stage as the block-sparse MFMA kernel (s_stage.hip) over the nonzero 16-key
    // tiles of every group's W_D, in all three precisions; the library GEMMs above then serve only
    // plans made with MQ_S_DENSE=1 and the adjoint
    struct Bsr {
        std::vector<BsrTask> tasks;
        std::vector<BsrRun> runs;
        // the nonzeros of the compact W tiles: positions in the tile array (tiles x 128 columns x 16
        // keys) and values, n_ent of each, in the context's pinned staging buffer (valid until the
        // upload, which the lowering's caller runs before the next lowering); the device builds the
        // tiles from them
        uint64_t* ent_pos = nullptr;
        double* ent_val = nullptr;
        uint64_t n_ent = 0;
        bool uploaded = false;
        uint64_t tiles = 0;
        uint64_t flops = 0;             // MFMA flops of the visited tiles (unpadded)
        bool narrow = true;             // every task's T_D step window fits 32-bit byte offsets
        // per tile: its group, first key and first column: the dense W_D can be rebuilt from the
        // tiles (the flop model reads W_D; the device keeps only the tiles)
        struct TileAt { uint32_t group, k0, n0; };
        std::vector<TileAt> tile_at;
        DevBuf d_tasks, d_runs, d_w64, d_w32, d_whi, d_wlo;
    } bsr;
    bool bsr_on = false;
    rocblas_handle blas = nullptr;
    // hipBLASLt form of the FP32 (index 0) and BF16x3 (index 1) library GEMMs, prepared on first
    // use outside graph capture; a GEMM without a hipBLASLt algorithm runs on rocBLAS
    mq_ctx* lt_ctx = nullptr;
    mutable std::vector<LtGemm> ltg[2];
    StagePlan() = default;
    StagePlan(const StagePlan&) = delete;
    StagePlan& operator=(const StagePlan&) = delete;
    ~StagePlan() {
        for (auto& v : ltg)
            for (auto& g : v) g.release();
    }
    // descriptors of lib GEMM i in mode mi; returns whether hipBLASLt runs it
    bool lt_prepare(int mi, size_t i, hipStream_t st) const {
        if (!lt_ctx || !lt_ctx->lt) return false;
        if (ltg[mi].size() != lib.size()) ltg[mi].resize(lib.size());
        LtGemm& L = ltg[mi][i];
        if (L.ready) return L.ok;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        HIPCHK(hipStreamIsCapturing(st, &cs));
        if (cs != hipStreamCaptureStatusNone) return false;   // never query the library inside a capture
        const LibGemm& g = lib[i];
        const hipDataType in = mi == 0 ? HIP_R_32F : HIP_R_16BF;
        const uint64_t K = mi == 0 ? g.K : 3ull * g.K;
        const uint64_t lda = g.lda, ldb = mi == 0 ? g.ldb : 3ull * g.K;
        LTCHK(hipblasLtMatmulDescCreate(&L.op, HIPBLAS_COMPUTE_32F, HIP_R_32F));
        hipblasOperation_t nn = HIPBLAS_OP_N;
        LTCHK(hipblasLtMatmulDescSetAttribute(L.op, HIPBLASLT_MATMUL_DESC_TRANSA, &nn, sizeof(nn)));
        LTCHK(hipblasLtMatmulDescSetAttribute(L.op, HIPBLASLT_MATMUL_DESC_TRANSB, &nn, sizeof(nn)));
        LTCHK(hipblasLtMatrixLayoutCreate(&L.a, in, g.M, K, lda));
        LTCHK(hipblasLtMatrixLayoutCreate(&L.b, in, K, g.N, ldb));
        LTCHK(hipblasLtMatrixLayoutCreate(&L.c, HIP_R_32F, g.M, g.N, g.ldc));
        std::array<uint64_t, 4> key{uint64_t(mi), g.M, g.N, K};
        auto it = lt_ctx->lt_tuned.find(key);
        if (it != lt_ctx->lt_tuned.end()) {
            L.algo = it->second;
            L.ok = L.chosen = true;
        } else if (lt_adopt_class(mi, g.M, g.N, K, L)) {
            L.ok = L.chosen = true;
        } else {
            hipblasLtMatmulPreference_t pref;
            LTCHK(hipblasLtMatmulPreferenceCreate(&pref));
            size_t wsz = LT_WORKSPACE;
            LTCHK(hipblasLtMatmulPreferenceSetAttribute(pref, HIPBLASLT_MATMUL_PREF_MAX_WORKSPACE_BYTES, &wsz, sizeof(wsz)));
            hipblasLtMatmulHeuristicResult_t res[1];
            int n = 0;
            hipblasStatus_t hs = hipblasLtMatmulAlgoGetHeuristic(lt_ctx->lt, L.op, L.a, L.b, L.c, L.c, pref, 1, res, &n);
            (void)hipblasLtMatmulPreferenceDestroy(pref);
            L.ok = hs == HIPBLAS_STATUS_SUCCESS && n > 0;
            if (L.ok) L.algo = res[0].algo;
        }
        L.ready = true;
        return L.ok;
    }
    // the solution tuned for this GEMM's shape class, if the library supports it for this shape
    bool lt_adopt_class(int mi, uint64_t M, uint64_t N, uint64_t K, LtGemm& L) const {
        auto c = lt_ctx->lt_class.find(mq_ctx::lt_class_key(mi, M, N, K));
        if (c == lt_ctx->lt_class.end()) return false;
        std::vector<int> idx{c->second};
        std::vector<hipblasLtMatmulHeuristicResult_t> res;
        if (hipblaslt_ext::getAlgosFromIndex(lt_ctx->lt, idx, res) != HIPBLAS_STATUS_SUCCESS || res.empty()) return false;
        const float one = 1.f, zero = 0.f;
        size_t need = 0;
        if (hipblaslt_ext::matmulIsAlgoSupported(lt_ctx->lt, L.op, &one, L.a, L.b, &zero, L.c, L.c, res[0].algo, need) !=
                HIPBLAS_STATUS_SUCCESS ||
            need > LT_WORKSPACE)
            return false;
        L.algo = res[0].algo;
        return true;
    }
    // operands of lib GEMM i in mode mi (0: FP32 workspace, 1: bf16 split buffer)
    void lt_operands(int mi, size_t i, Bases b, const void*& A, const void*& B, float*& C) const {
        float* ws = reinterpret_cast<float*>(b.p[BASE_WS]);
        const LibGemm& g = lib[i];
        A = mi == 0 ? static_cast<const void*>(ws + g.a_off) : static_cast<const void*>(bf + g.bfa_off);
        B = mi == 0 ? static_cast<const void*>(ws + g.b_off) : static_cast<const void*>(bf + g.bfb_off);
        C = ws + g.c_off;
    }
    // time up to 8 of the library's candidate algorithms for lib GEMM i and keep the fastest
    // (cached in the context by shape and by shape class); returns the number of timings.  A GEMM
    // whose algorithm came from either cache is not timed again.
    int lt_tune(int mi, size_t i, hipStream_t st, Bases b, hipEvent_t e0, hipEvent_t e1) const {
        if (lib[i].sparse) return 0;   // the group runs in the block-sparse kernel
        if (!lt_prepare(mi, i, st)) return 0;
        LtGemm& L = ltg[mi][i];
        if (L.chosen) return 0;
        const LibGemm& g = lib[i];
        const uint64_t K = mi == 0 ? g.K : 3ull * g.K;
        std::array<uint64_t, 4> key{uint64_t(mi), g.M, g.N, K};
        auto it = lt_ctx->lt_tuned.find(key);
        if (it != lt_ctx->lt_tuned.end()) {
            L.algo = it->second;
            return 0;
        }
        hipblasLtMatmulPreference_t pref;
        LTCHK(hipblasLtMatmulPreferenceCreate(&pref));
        size_t wsz = LT_WORKSPACE;
        LTCHK(hipblasLtMatmulPreferenceSetAttribute(pref, HIPBLASLT_MATMUL_PREF_MAX_WORKSPACE_BYTES, &wsz, sizeof(wsz)));
        hipblasLtMatmulHeuristicResult_t res[8];
        int n = 0;
        hipblasStatus_t hs = hipblasLtMatmulAlgoGetHeuristic(lt_ctx->lt, L.op, L.a, L.b, L.c, L.c, pref, 8, res, &n);
        (void)hipblasLtMatmulPreferenceDestroy(pref);
        if (hs != HIPBLAS_STATUS_SUCCESS || n <= 0) return 0;
        const void *A, *B;
        float* C;
        lt_operands(mi, i, b, A, B, C);
        const float one = 1.f, zero = 0.f;
        float best = 0.f;
        int timed = 0;
        for (int c = 0; c < n; ++c) {
            auto run = [&]() {
                return hipblasLtMatmul(lt_ctx->lt, L.op, &one, A, L.a, B, L.b, &zero, C, L.c, C, L.c, &res[c].algo,
                                       lt_ctx->lt_workspace(st), LT_WORKSPACE, st);
            };
            if (run() != HIPBLAS_STATUS_SUCCESS) continue;   // warm-up; a failing candidate is skipped
            HIPCHK(hipEventRecord(e0, st));
            for (int r = 0; r < 2; ++r) LTCHK(run());
            HIPCHK(hipEventRecord(e1, st));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            ++timed;
            if (timed == 1 || ms < best) best = ms, L.algo = res[c].algo;
        }
        if (timed) {
            lt_ctx->lt_tuned[key] = L.algo;
            const int sol = hipblaslt_ext::getIndexFromAlgo(L.algo);
            if (sol >= 0) lt_ctx->lt_class[mq_ctx::lt_class_key(mi, g.M, g.N, K)] = sol;
            L.chosen = true;
        }
        return timed;
    }
    void lt_run(int mi, size_t i, hipStream_t st, const void* A, const void* B, float* C) const {
        const LtGemm& L = ltg[mi][i];
        const float one = 1.f, zero = 0.f;
        LTCHK(hipblasLtMatmul(lt_ctx->lt, L.op, &one, A, L.a, B, L.b, &zero, C, L.c, C, L.c, &L.algo,
                              lt_ctx->lt_workspace(st), LT_WORKSPACE, st));
    }
    uint64_t t_end = 0;     // T_D occupies workspace elements [0, t_end)
    uint64_t bf_elems = 0;  // size of the bf16 buffer
    uint16_t* bf = nullptr;
    uint32_t g_chunks = 0;
    uint64_t ws_elems = 0;
    uint64_t s_begin = 0, s_elems = 0;   // workspace range zero-filled at the start of each apply
    DevBuf d_maps, d_units, d_scales, d_rows, d_zero_ws, d_zero_bf, d_regions, d_small, d_src_off, d_alpha;
    uint64_t sizes[N_BASES] = {0, 0, 0, 0, 0, 0};   // element counts of each base for validation

    void finish_tables() {
        gT.build(t_tasks);
        gF.build(f_tasks);
        gU.build(u_tasks);
        red.finish();
        std::vector<GatherRegion> big;
        for (const auto& r : regions) {
            if (uint64_t(r.R) * r.C <= uint64_t(GATHER_SMALL)) small.push_back(r);
            else big.push_back(r);
        }
        regions.swap(big);
        uint32_t ch = 0;
        for (auto& r : regions) {
            r.chunk_begin = ch;
            uint64_t n = uint64_t(r.R) * r.C;
            uint64_t c = (n + GATHER_CHUNK - 1) / GATHER_CHUNK;
            if (c == 0) c = 1;
            MQ_REQUIRE(uint64_t(ch) + c < (1ull << 31), "too many gather chunks");
            ch += uint32_t(c);
        }
        g_chunks = ch;
    }

    // Every element touched by a GEMM segment / gather region must lie inside its buffer.
    static uint64_t extent(uint32_t rows_contig, uint32_t ncols, uint32_t ld) {
        if (rows_contig == 0 || ncols == 0) return 0;
        return uint64_t(ncols - 1) * ld + rows_contig;
    }
    void check_in(uint8_t base, uint64_t off, uint64_t ext, const char* what) const {
        MQ_REQUIRE(base < N_BASES, "bad base");
        if (ext == 0) return;
        if (off + ext > sizes[base])
            fail(MQ_ERR_INTERNAL, std::string("operand out of bounds: ") + what + " base " + std::to_string(base) +
                                      " off " + std::to_string(off) + " ext " + std::to_string(ext) + " size " +
                                      std::to_string(sizes[base]));
    }
    // A stacked task: its rows are partitioned into contiguous row units; every unit's A block and
    // C blocks lie inside their buffers; one plain (non-transposed, unmapped) A segment.
    void check_stacked(const GemmTask& t, const std::vector<GemmSeg>& segs, bool split) const {
        MQ_REQUIRE(uint64_t(t.rowunit) + t.M / 16 <= maps.size(), "row unit table range");
        MQ_REQUIRE(t.seg0 < t.seg1 && t.seg1 <= segs.size(), "stacked task segment range");
        MQ_REQUIRE(t.colunit == MAP_NONE && t.cbw > 0, "stacked task shape");
        for (uint32_t q = t.seg0; q < t.seg1; ++q) {
            const GemmSeg& g = segs[q];
            MQ_REQUIRE(g.units == MAP_NONE && g.bmap == MAP_NONE, "stacked task with remapped B");
            MQ_REQUIRE(g.amap != MAP_NONE || !g.trans_a, "stacked task with a transposed A outside a block map");
            MQ_REQUIRE(g.amap == MAP_NONE || t.seg1 - t.seg0 >= 1, "stacked block map");
            check_in(g.b_base, g.b_off, g.trans_b ? extent(t.N, g.K, g.ldb) : extent(g.K, t.N, g.ldb), "stacked B");
        }
        MQ_REQUIRE(t.M % 16 == 0, "stacked task rows not a multiple of 16");
        uint32_t m = 0;
        while (m < t.M) {
            const uint32_t u = maps[t.rowunit + m / 16];
            MQ_REQUIRE(u < rows.size(), "row unit out of range");
            const GemmRow& R = rows[u];
            MQ_REQUIRE(R.row0 == m && m % 16 == 0, "row unit does not start at its first row");
            uint32_t e = m;
            while (e < t.M && maps[t.rowunit + e / 16] == u) e += 16;
            MQ_REQUIRE((e - m) % 16 == 0 && R.nrows <= e - m && e - m < R.nrows + 16, "row unit span");
            const uint32_t nr = R.nrows;
            MQ_REQUIRE(R.ldc >= 1 && R.c_si >= 1, "row unit output strides");
            for (uint32_t q = t.seg0; q < t.seg1 && nr; ++q) {
                const GemmSeg& g = segs[q];
                if (!g.K) continue;
                if (g.amap == MAP_NONE) {
                    check_in(g.a_base, g.a_off + R.a_src, uint64_t(nr - 1) * R.a_si + uint64_t(g.K - 1) * R.a_sk + 1,
                             "stacked A block");
                } else {
                    MQ_REQUIRE(uint64_t(g.amap) + R.jblk < maps.size(), "stacked A block map range");
                    const uint64_t si = g.trans_a ? g.lda : 1, sk = g.trans_a ? 1 : g.lda;
                    check_in(g.a_base, g.a_off + uint64_t(maps[g.amap + R.jblk]) * g.bldk,
                             uint64_t(nr - 1) * si + uint64_t(g.K - 1) * sk + 1, "stacked A unit block");
                }
            }
            // C blocks: [off, off + ext) of the workspace, and of the A' copies when split
            auto c_block = [&](uint64_t rel, uint32_t ncols) {
                if (!nr || !ncols) return;
                const uint64_t ext = uint64_t(nr - 1) * R.c_si + uint64_t(ncols - 1) * R.ldc + 1;
                check_in(t.c_base, t.c_off + R.c_src + rel, ext, "stacked C block");
                if (split && t.c_off + R.c_split + rel + 2 * R.split_stride + ext > bf_a_elems)
                    fail(MQ_ERR_INTERNAL, "stacked split output out of bounds");
            };
            if (split) MQ_REQUIRE(R.c_si == 1, "split output with a row stride");
            if (R.cmap == MAP_NONE) {
                c_block(0, t.N);
            } else {
                MQ_REQUIRE(R.c_si == 1, "column-remapped stacked output with a row stride");
                const uint32_t nblk = (t.N + t.cbw - 1) / t.cbw;
                MQ_REQUIRE(uint64_t(R.cmap) + nblk <= maps.size(), "row unit column map range");
                for (uint32_t j = 0; j < nblk; ++j)
                    c_block(uint64_t(maps[R.cmap + j]) * R.cldk, std::min(t.cbw, t.N - j * t.cbw));
            }
            m = e;
        }
    }
    void validate() const {
        auto gemm = [&](const std::vector<GemmTask>& tasks, const std::vector<GemmSeg>& segs, bool split) {
            for (const auto& t : tasks) {
                if (t.rowunit != MAP_NONE) {
                    check_stacked(t, segs, split);
                    continue;
                }
                MQ_REQUIRE(t.ldc >= t.M || t.N <= 1, "ldc < M");
                if (t.cmap == MAP_NONE) {
                    check_in(t.c_base, t.c_off + t.m_off, extent(t.M, t.N, t.ldc), "C");
                } else {
                    MQ_REQUIRE(t.cbw > 0 && t.ldc >= t.M, "column remap without block width");
                    uint32_t nblk = (t.N + t.cbw - 1) / t.cbw;
                    MQ_REQUIRE(uint64_t(t.cmap) + nblk <= maps.size(), "column map out of range");
                    for (uint32_t j = 0; j < nblk; ++j) {
                        uint32_t w = std::min(t.cbw, t.N - j * t.cbw);
                        check_in(t.c_base, t.c_off + t.m_off + uint64_t(maps[t.cmap + j]) * t.cldk,
                                 extent(t.M, w, t.ldc), "C block");
                    }
                }
                MQ_REQUIRE(t.seg0 <= t.seg1 && t.seg1 <= segs.size(), "segment range");
                // column units: unit of every column, and the column count of each unit
                std::vector<uint32_t> ucols;
                if (t.colunit != MAP_NONE) {
                    MQ_REQUIRE(uint64_t(t.colunit) + t.N <= maps.size() && t.nunits > 0, "column unit table range");
                    ucols.assign(t.nunits, 0);
                    for (uint32_t n = 0; n < t.N; ++n) {
                        MQ_REQUIRE(maps[t.colunit + n] < t.nunits, "column unit out of range");
                        ++ucols[maps[t.colunit + n]];
                    }
                }
                for (uint32_t s = t.seg0; s < t.seg1; ++s) {
                    const GemmSeg& g = segs[s];
                    if (g.amap != MAP_NONE) {
                        MQ_REQUIRE(g.trans_a && g.bbh > 0 && g.K % g.bbh == 0, "remapped A shape");
                        const uint32_t nblk = g.K / g.bbh;
                        MQ_REQUIRE(uint64_t(g.amap) + nblk <= maps.size() && uint64_t(g.ascale) + nblk <= scales.size(),
                                   "A block map range");
                        for (uint32_t j = 0; j < nblk; ++j)
                            check_in(g.a_base, g.a_off + maps[g.amap + j] + uint64_t(t.m_off) * g.lda,
                                     extent(g.bbh, t.M, g.lda), "A^T block");
                    } else if (!g.trans_a) {
                        check_in(g.a_base, g.a_off + t.m_off, extent(t.M, g.K, g.lda), "A");
                    } else {
                        check_in(g.a_base, g.a_off + uint64_t(t.m_off) * g.lda, extent(g.K, t.M, g.lda), "A^T");
                    }
                    if (g.units != MAP_NONE) {
                        MQ_REQUIRE(t.colunit != MAP_NONE && g.bmap == MAP_NONE, "unit segment without column units");
                        MQ_REQUIRE(g.bbh > 0 && g.K % g.bbh == 0, "unit segment shape");
                        MQ_REQUIRE(uint64_t(g.units) + t.nunits <= units.size(), "unit table range");
                        const uint32_t nblk = g.K / g.bbh;
                        for (uint32_t u = 0; u < t.nunits; ++u) {
                            const GemmUnit& U = units[g.units + u];
                            if (!U.valid) continue;
                            MQ_REQUIRE(U.col0 + ucols[u] <= t.N, "unit columns");
                            for (uint32_t n = U.col0; n < U.col0 + ucols[u]; ++n)
                                MQ_REQUIRE(maps[t.colunit + n] == u, "unit columns are not contiguous");
                            MQ_REQUIRE(uint64_t(U.map) + nblk <= maps.size(), "unit map range");
                            for (uint32_t j = 0; j < nblk; ++j)
                                check_in(g.b_base, g.b_off + U.src + uint64_t(maps[U.map + j]) * U.ld,
                                         uint64_t(g.bbh) * ucols[u], "B unit block");
                        }
                    } else if (g.bmap != MAP_NONE) {
                        MQ_REQUIRE(!g.trans_b && g.bbh > 0 && g.K % g.bbh == 0, "row remap shape");
                        uint32_t nblk = g.K / g.bbh;
                        MQ_REQUIRE(uint64_t(g.bmap) + nblk <= maps.size(), "row map out of range");
                        for (uint32_t j = 0; j < nblk; ++j)
                            check_in(g.b_base, g.b_off + uint64_t(maps[g.bmap + j]) * g.bldk,
                                     extent(g.bbh, t.N, g.ldb), "B block");
                    } else if (!g.trans_b) {
                        check_in(g.b_base, g.b_off, extent(g.K, t.N, g.ldb), "B");
                    } else {
                        check_in(g.b_base, g.b_off, extent(t.N, g.K, g.ldb), "B^T");
                    }
                }
            }
        };
        gemm(t_tasks, t_segs, split_t);
        gemm(f_tasks, f_segs, false);
        gemm(u_tasks, u_segs, false);
        for (const auto& t : bsr.tasks) {   // every T_D read (128 rows of each key column) and S_D write
            MQ_REQUIRE(!bsr.narrow || (32ull * t.ld + S_BSR_TILE_M) * sizeof(double) < (1ull << 32),
                       "block-sparse task too wide for 32-bit offsets");
            MQ_REQUIRE(t.M <= S_BSR_TILE_M && t.N <= S_BSR_TILE_N && t.run0 <= t.run1 && t.run1 <= bsr.runs.size(),
                       "block-sparse task shape");
            uint32_t kmax = 0;
            for (uint32_t r = t.run0; r < t.run1; ++r) {
                const BsrRun& R = bsr.runs[r];
                MQ_REQUIRE(R.k0 % S_BSR_TILE_K == 0 && R.nk > 0, "block-sparse run");
                MQ_REQUIRE(uint64_t(R.tile) + (R.nk + S_BSR_TILE_K - 1) / S_BSR_TILE_K <= bsr.tiles, "block-sparse tile");
                kmax = std::max(kmax, R.k0 + R.nk);
            }
            if (kmax) check_in(BASE_WS, t.a_off, uint64_t(kmax - 1) * t.ld + S_BSR_TILE_M, "block-sparse T_D rows");
            check_in(BASE_WS, t.c_off, extent(t.M, t.N, t.ld), "block-sparse S_D tile");
        }
        if (split_t)
            for (const auto& t : t_tasks)
                MQ_REQUIRE(t.rowunit != MAP_NONE && t.c_base == BASE_WS && !t.beta, "split T stage needs stacked tasks");
        for (const auto& z : zero_ws) check_in(BASE_WS, z.off, z.len, "zero chunk");
        for (const auto& z : zero_bf)
            if (z.off + z.len > bf_a_elems) fail(MQ_ERR_INTERNAL, "bf16 zero chunk out of bounds");
        for (const auto& g : lib) {
            MQ_REQUIRE(g.lda >= g.M && g.ldb >= g.K && g.ldc >= g.M, "library GEMM leading dimensions");
            MQ_REQUIRE(g.a_off + extent(g.M, g.K, g.lda) <= t_end, "library GEMM A outside T_D");
            check_in(BASE_WS, g.a_off, extent(g.M, g.K, g.lda), "library GEMM A");
            check_in(BASE_WS, g.b_off, extent(g.K, g.N, g.ldb), "library GEMM B");
            check_in(BASE_WS, g.c_off, extent(g.M, g.N, g.ldc), "library GEMM C");
        }
        for (const auto* list : {&red.regions, &red.small})
            for (const auto& r : *list) {
                check_in(BASE_OUT, r.dst_off, extent(r.R, r.C, r.ld_dst), "reduced block");
                MQ_REQUIRE(r.t0 <= r.t1 && r.t1 <= red.src_off.size(), "reduction source range");
                for (uint32_t t = r.t0; t < r.t1; ++t)
                    check_in(BASE_BRA, red.src_off[t], uint64_t(r.R) * r.C, "partial sum");
            }
        for (const auto* list : {&regions, &small})
            for (const auto& r : *list) {
                check_in(BASE_WS, r.dst_off, extent(r.R, r.C, r.ld_dst), "S region");
                MQ_REQUIRE(r.t0 <= r.t1 && r.t1 <= src_off.size(), "region task range");
                for (uint32_t t = r.t0; t < r.t1; ++t) check_in(BASE_WS, src_off[t], uint64_t(r.R) * r.C, "T slab");
            }
    }
    void upload_all(hipStream_t st) {
        gT.upload_all(t_segs, st);
        gF.upload_all(f_segs, st);
        gU.upload_all(u_segs, st);
        if (bsr_on && !bsr.uploaded) {   // once: the host copy of the nonzeros is released after it
            upload(bsr.d_tasks, bsr.tasks, st);
            upload(bsr.d_runs, bsr.runs, st);
            // the W tiles in the three arithmetics: FP64 zero-filled and the nonzeros scattered into
            // it; FP32 and the bf16 split of the FP32 value converted on the device
            const uint64_t nw = bsr.tiles * uint64_t(S_BSR_TILE_N) * S_BSR_TILE_K;
            bsr.d_w64.alloc(std::max<uint64_t>(nw, 1) * sizeof(double));
            HIPCHK(launch_fill(st, bsr.d_w64.as<double>(), nw, 0.0));
            if (bsr.n_ent) {
                DevBuf d_pos, d_val;
                d_pos.alloc(bsr.n_ent * sizeof(uint64_t));
                d_val.alloc(bsr.n_ent * sizeof(double));
                HIPCHK(hipMemcpyAsync(d_pos.p, bsr.ent_pos, bsr.n_ent * sizeof(uint64_t), hipMemcpyHostToDevice, st));
                HIPCHK(hipMemcpyAsync(d_val.p, bsr.ent_val, bsr.n_ent * sizeof(double), hipMemcpyHostToDevice, st));
                HIPCHK(launch_scatter(st, bsr.n_ent, d_pos.as<uint64_t>(), d_val.as<double>(), bsr.d_w64.as<double>()));
                HIPCHK(hipStreamSynchronize(st));   // before the entry buffers are freed and the staging reused
            }
            bsr.d_w32.alloc(std::max<uint64_t>(nw, 1) * sizeof(float));
            bsr.d_whi.alloc(std::max<uint64_t>(nw, 1) * sizeof(uint16_t));
            bsr.d_wlo.alloc(std::max<uint64_t>(nw, 1) * sizeof(uint16_t));
            HIPCHK(launch_split_tiles(st, bsr.d_w64.as<double>(), nw, bsr.d_w32.as<float>(), bsr.d_whi.as<uint16_t>(),
                                      bsr.d_wlo.as<uint16_t>()));
            HIPCHK(hipStreamSynchronize(st));   // the host copy is released below
            bsr.ent_pos = nullptr;
            bsr.ent_val = nullptr;
            bsr.n_ent = 0;
            bsr.uploaded = true;
        }
        upload(d_regions, regions, st);
        upload(d_small, small, st);
        upload(d_src_off, src_off, st);
        upload(d_alpha, alpha, st);
        upload(d_maps, maps, st);
        upload(d_units, units, st);
        upload(d_scales, scales, st);
        upload(d_rows, rows, st);
        upload(d_zero_ws, zero_ws, st);
        upload(d_zero_bf, zero_bf, st);
        red.upload_all(st);
    }
    // the four stages of an application: zero-fill, T GEMM, middle (gather or S GEMM), final GEMM.
    // mode != MODE_F64 (grouped plans only): T_D, S_D and W_D are FP32 in a float workspace.
    void stage(int k, hipStream_t st, Bases b, int mode = MODE_F64) const {
        const GemmTables tb{d_maps.as<uint32_t>(), d_units.as<GemmUnit>(), d_rows.as<GemmRow>(), d_scales.as<double>()};
        const bool mixed = mode != MODE_F64;
        switch (k) {
            case 0:
                if (s_elems) {
                    if (!mixed) HIPCHK(launch_fill(st, b.p[BASE_WS] + s_begin, s_elems, 0.0));
                    else   // zero bits: the float range [s_begin, s_begin + s_elems) as doubles (both multiples of 32)
                        HIPCHK(launch_fill(st, reinterpret_cast<double*>(reinterpret_cast<float*>(b.p[BASE_WS]) + s_begin),
                                           s_elems / 2, 0.0));
                }
                if (mode == MODE_BF16X3 && split_t && !bsr_on)
                    HIPCHK(launch_zero_chunks(st, d_zero_bf.as<ZeroChunk>(), uint32_t(zero_bf.size()), bf, 2));
                else
                    HIPCHK(launch_zero_chunks(st, d_zero_ws.as<ZeroChunk>(), uint32_t(zero_ws.size()), b.p[BASE_WS],
                                              mixed ? 4 : 8));
                break;
            case 1:
                // BF16x3: split products in the lean 128 x 128 T tasks (t_stage.hip), FP32 elsewhere
                if (mode == MODE_BF16X3 && split_t && !bsr_on) {
                    Bases bs = b;   // C of the stacked T tasks: the A' copies in the bf16 buffer
                    bs.p[BASE_WS] = reinterpret_cast<double*>(bf);
                    gT.run(st, MODE_F32, "dds", bs, tb);
                } else {
                    gT.run(st, mode, "ddf", b, tb);
                }
                break;
            case 2:
                if (bsr_on) {
                    const void* b0 = mode == MODE_F64 ? bsr.d_w64.p : mode == MODE_F32 ? bsr.d_w32.p : bsr.d_whi.p;
                    const void* b1 = mode == MODE_BF16X3 ? bsr.d_wlo.p : nullptr;
                    HIPCHK(launch_s_bsr(st, mode, bsr.d_tasks.as<BsrTask>(), uint32_t(bsr.tasks.size()),
                                        bsr.d_runs.as<BsrRun>(), b0, b1, b.p[BASE_WS], bsr.narrow));
                }
                run_lib(st, b, mode);
                if (!mixed) run_gather(st, b.p[BASE_WS], b.p[BASE_WS]);
                break;
            default:
                gF.run(st, mode, "dfd", b, tb);
                red.run(st, b.p[BASE_BRA], b.p[BASE_OUT]);
        }
    }
    void run(hipStream_t st, Bases b, int mode = MODE_F64) const {
        for (int k = 0; k < 4; ++k) stage(k, st, b, mode);
    }
    // U stage of the adjoint (FP64)
    void run_u(hipStream_t st, Bases b) const {
        const GemmTables tb{d_maps.as<uint32_t>(), d_units.as<GemmUnit>(), d_rows.as<GemmRow>(), d_scales.as<double>()};
        gU.run(st, MODE_F64, "ddd", b, tb);
    }
    void run_lib(hipStream_t st, Bases b, int mode) const {
        if (lib.empty()) return;
        const rocblas_operation nn = rocblas_operation_none;
        const rocblas_handle blas = lt_ctx ? lt_ctx->blas_for(st) : this->blas;
        if (mode == MODE_F64) {
            const double one = 1.0, zero = 0.0;
            const double* ws = b.p[BASE_WS];
            for (const auto& g : lib)
                if (!g.sparse)
                    BLASCHK(rocblas_dgemm(blas, nn, nn, g.M, g.N, g.K, &one, ws + g.a_off, g.lda, ws + g.b_off, g.ldb, &zero,
                                      const_cast<double*>(ws) + g.c_off, g.ldc));
        } else if (mode == MODE_F32) {
            const float one = 1.f, zero = 0.f;
            float* ws = reinterpret_cast<float*>(b.p[BASE_WS]);
            for (size_t i = 0; i < lib.size(); ++i) {
                const auto& g = lib[i];
                if (g.sparse) continue;
                if (lt_prepare(0, i, st))
                    lt_run(0, i, st, ws + g.a_off, ws + g.b_off, ws + g.c_off);
                else
                    BLASCHK(rocblas_sgemm(blas, nn, nn, g.M, g.N, g.K, &one, ws + g.a_off, g.lda, ws + g.b_off, g.ldb,
                                          &zero, ws + g.c_off, g.ldc));
            }
        } else {
            float* ws = reinterpret_cast<float*>(b.p[BASE_WS]);
            const float one = 1.f, zero = 0.f;
            const rocblas_datatype b16 = rocblas_datatype_bf16_r, f32 = rocblas_datatype_f32_r;
            for (size_t i = 0; i < lib.size(); ++i) {
                const auto& g = lib[i];
                if (g.sparse) continue;
                MQ_REQUIRE(bf, "bf16 split buffer not allocated");
                if (!split_t) HIPCHK(launch_split3_bf16(st, ws + g.a_off, bf + g.bfa_off, uint64_t(g.M) * g.K));
                if (lt_prepare(1, i, st)) {
                    lt_run(1, i, st, bf + g.bfa_off, bf + g.bfb_off, ws + g.c_off);
                    continue;
                }
                BLASCHK(rocblas_gemm_ex(blas, nn, nn, g.M, g.N, 3 * g.K, &one, bf + g.bfa_off, b16, g.M,
                                        bf + g.bfb_off, b16, 3 * g.K, &zero, ws + g.c_off, f32, g.ldc, ws + g.c_off,
                                        f32, g.ldc, f32, rocblas_gemm_algo_standard, 0, 0));
            }
        }
    }
    // gather of both region lists: src offsets are relative to src, destinations to dst
    void run_gather(hipStream_t st, double* src, double* dst) const {
        HIPCHK(launch_gather(st, d_regions.as<GatherRegion>(), uint32_t(regions.size()), d_src_off.as<uint64_t>(),
                             d_alpha.as<double>(), g_chunks, src, dst));
        HIPCHK(launch_gather_small(st, d_small.as<GatherRegion>(), uint32_t(small.size()), d_src_off.as<uint64_t>(),
                                   d_alpha.as<double>(), src, dst));
    }
    template <class F>
    void for_regions(F&& f) const {
        for (const auto& r : regions) f(r);
        for (const auto& r : small) f(r);
    }
};

static_assert(int(MQ_PREC_F64) == int(MODE_F64) && int(MQ_PREC_F32) == int(MODE_F32) &&
                  int(MQ_PREC_BF16X3) == int(MODE_BF16X3),
              "precision enum and GEMM modes agree");

struct mq_plan {
    mq_ctx* ctx = nullptr;
    StagePlan st;
    DevBuf ws;
    const mq_boundary* left = nullptr;
    const mq_boundary* right = nullptr;
    uint64_t vec_size = 0;
    mutable uint64_t flops = 0;
    mutable std::function<uint64_t()> s_flops;   // grouped plans: the S-stage flops, evaluated on demand
    std::vector<uint64_t> out_offsets;
    // grouped plans: the dense MPO matrices W_D (ws_init_elems of them) at ws_init_off of the
    // workspace, and the FP32 workspace of the reduced-precision modes (taken on first use)
    bool grouped = false;
    uint64_t ws_init_elems = 0;
    uint64_t ws_init_off = 0;
    DevBuf ws32;
    DevBuf part;   // split-K partial sums of the final stage (FP64)
    DevBuf bf;     // bf16 operands of the BF16x3 S stage
    // Jacobi-Davidson vectors (allocated on first use)
    DevBuf jd_buf;
    int jd_cap = 0;
    // replayable HIP graphs of whole applications, one per precision, over the plan's own operand
    // slots: every Davidson iteration replays it instead of re-issuing every launch
    bool use_graphs = true;
    hipGraphExec_t graphs[3] = {nullptr, nullptr, nullptr};   // per precision
    bool seen[3] = {false, false, false};                     // precision applied once eagerly
    DevBuf io;                                                // operand slots of the graphs: in, out
    // adjoint (MQ_SIGMA_ADJOINT): per MPO term t, the entries [adj_ptr[t], adj_ptr[t+1]) of
    // (offset in the W_D region, coefficient) through which the term's scale enters W_D
    bool adjoint = false;
    uint32_t n_terms = 0;
    std::vector<uint32_t> adj_ptr;
    std::vector<uint64_t> adj_idx;
    std::vector<double> adj_coef;
    uint64_t s_region0 = 0, s_region1 = 0;   // the S_D groups' workspace range
    DevBuf adj_g, d_adj_ptr, d_adj_idx, d_adj_coef, d_grad;

    void drop_graphs() {
        for (auto& g : graphs)
            if (g) (void)hipGraphExecDestroy(g), g = nullptr;
        for (auto& b : seen) b = false;
    }
    ~mq_plan() {
        drop_graphs();
        if (left) --left->plans;
        if (right) --right->plans;
    }
    // the boundaries the plan reads: made resident (device copies complete in `ctx->stream` order)
    // and pinned there until the plan is destroyed
    void bind(const mq_boundary* l, const mq_boundary* r);

    // the base pointers of one application in the given mode
    Bases bases(const double* psi, double* sigma, int mode) {
        Bases b{};
        if (mode == MODE_F64) {
            b.p[BASE_WS] = ws.as<double>();
        } else {
            MQ_REQUIRE(grouped, "reduced-precision sigma needs a GROUPED plan");
            if (!ws32.p) {
                // the FP32 workspace; W_D converted on the device from the FP64 copy
                ctx->take(mq_ctx::POOL_WS32, ws32, std::max<uint64_t>(st.ws_elems, 64) * sizeof(float));
                if (ws_init_elems && !st.bsr_on)
                    HIPCHK(launch_to_float(ctx->stream, ws.as<double>() + ws_init_off, ws32.as<float>() + ws_init_off,
                                           ws_init_elems));
            }
            b.p[BASE_WS] = reinterpret_cast<double*>(ws32.as<float>());
            if (mode == MODE_BF16X3 && st.bf_elems && !bf.p && !st.bsr_on) {
                // bf16 operands of the one-GEMM BF16x3 form: A' regions (written per application) and
                // B' = [W_lo; W_hi; W_hi] per group (constant, split here on the device from W_D)
                ctx->take(mq_ctx::POOL_BF, bf, std::max<uint64_t>(st.bf_elems, 64) * sizeof(uint16_t));
                st.bf = bf.as<uint16_t>();
                for (const auto& g : st.lib)
                    HIPCHK(launch_split_w_bf16(ctx->stream, ws.as<double>() + g.b_off, g.ldb, g.K, g.N,
                                               st.bf + g.bfb_off));
            }
        }
        b.p[BASE_IN] = const_cast<double*>(psi);
        b.p[BASE_OUT] = sigma;
        b.p[BASE_LEFT] = left->ptr();
        b.p[BASE_RIGHT] = right->ptr();
        b.p[BASE_BRA] = part.as<double>();
        return b;
    }
};

// ============================================================================ lowering
namespace {

GemmTask& add_gemm(std::vector<GemmTask>& tasks, std::vector<GemmSeg>& segs, uint8_t c_base, uint64_t c_off,
                   uint32_t M, uint32_t N, uint32_t ldc, const std::vector<GemmSeg>& s, uint8_t beta = 0) {
    static GemmTask dummy;
    if (M == 0 || N == 0 || s.empty()) return dummy;
    GemmTask t{};
    t.c_off = c_off;
    t.M = M;
    t.N = N;
    t.ldc = ldc;
    t.seg0 = uint32_t(segs.size());
    segs.insert(segs.end(), s.begin(), s.end());
    t.seg1 = uint32_t(segs.size());
    t.c_base = c_base;
    t.beta = beta;
    t.cmap = MAP_NONE;
    t.cbw = N;
    t.cldk = 0;
    t.colunit = MAP_NONE;
    t.nunits = 0;
    t.m_off = 0;
    t.rowunit = MAP_NONE;
    tasks.push_back(t);
    return tasks.back();
}

// A task over an existing segment range (tasks of one slice share their segments), rows
// [m_off, m_off + M) of op(A) and C.
void add_gemm_range(std::vector<GemmTask>& tasks, uint8_t c_base, uint64_t c_off, uint32_t M, uint32_t N, uint32_t ldc,
                    uint32_t seg0, uint32_t seg1, uint32_t m_off) {
    if (M == 0 || N == 0 || seg0 == seg1) return;
    GemmTask t{};
    t.c_off = c_off;
    t.M = M;
    t.N = N;
    t.ldc = ldc;
    t.seg0 = seg0;
    t.seg1 = seg1;
    t.c_base = c_base;
    t.cmap = MAP_NONE;
    t.cbw = N;
    t.colunit = MAP_NONE;
    t.m_off = m_off;
    t.rowunit = MAP_NONE;
    tasks.push_back(t);
}

GemmSeg seg(uint8_t ab, uint64_t ao, uint32_t lda, bool ta, uint8_t bb, uint64_t bo, uint32_t ldb, bool tb, uint32_t K,
            double alpha = 1.0) {
    GemmSeg g{};
    g.a_off = ao;
    g.b_off = bo;
    g.K = K;
    g.lda = lda;
    g.ldb = ldb;
    g.a_base = ab;
    g.b_base = bb;
    g.trans_a = ta;
    g.trans_b = tb;
    g.alpha = alpha;
    g.bmap = MAP_NONE;
    g.bbh = 1;
    g.bldk = 0;
    g.units = MAP_NONE;
    g.amap = MAP_NONE;
    g.ascale = MAP_NONE;
    return g;
}

// T stage of sigma and of the right step: T_ti = ket_slab (ls x rs_ket) * [R_b] (rs_ket x nb*rs_bra)
// (tasks.cpp:650-752).  Returns the workspace offset of every T.
std::vector<uint64_t> lower_T_right(StagePlan& P, const Schedule& S, const BoundaryIndex& R,
                                    const std::vector<uint64_t>& r_off, const std::vector<uint32_t>& ket_rows,
                                    const std::vector<uint64_t>& ket_off, uint64_t& ws) {
    std::vector<uint64_t> t_off(S.t_ci.size());
    for (size_t ti = 0; ti < S.t_ci.size(); ++ti) {
        uint32_t ci = S.t_ci[ti], cie = S.t_ci_eff[ti], lbk = S.t_lb_ket[ti];
        MQ_REQUIRE(ci < R.n_cohorts() && cie < R.n_cohorts() && lbk < ket_rows.size(), "T tuple out of range");
        uint32_t M = ket_rows[lbk];
        uint32_t K = uint32_t(R.left_size(ci)), brs = uint32_t(R.right_size(ci));
        uint32_t nb = R.n_blocks[cie];
        MQ_REQUIRE(R.left_size(cie) * R.right_size(cie) == uint64_t(K) * brs, "cohort shape mismatch");
        t_off[ti] = ws;
        ws += align32(uint64_t(M) * nb * brs);
        uint64_t a_off = ket_off[lbk] + uint64_t(S.t_mps_offset[ti]) * M;
        if (ci == cie) {
            add_gemm(P.t_tasks, P.t_segs, BASE_WS, t_off[ti], M, nb * brs, M,
                     {seg(BASE_IN, a_off, M, false, BASE_RIGHT, r_off[cie], K, false, K)});
        } else {
            for (uint32_t b = 0; b < nb; ++b)
                add_gemm(P.t_tasks, P.t_segs, BASE_WS, t_off[ti] + uint64_t(b) * M * brs, M, brs, M,
                         {seg(BASE_IN, a_off, M, false, BASE_RIGHT, r_off[cie] + uint64_t(b) * K * brs, brs, true, K)});
        }
    }
    return t_off;
}

// S stage of sigma / right step (create_s_r, tasks.cpp:466-494): S is (nSrows*ls) x stripe.
std::vector<uint64_t> lower_S_right(StagePlan& P, const Schedule& S, const std::vector<uint64_t>& t_off, uint64_t& ws) {
    std::vector<uint64_t> s_off(S.n_cohorts());
    P.s_begin = ws;
    for (uint32_t m = 0; m < S.n_blocks; ++m)
        for (uint64_t c = S.coh_begin[m]; c < S.coh_begin[m + 1]; ++c) {
            uint32_t ls = S.coh_ls[c], nS = S.coh_nSrows[c], stripe = S.coh_stripe[c];
            uint64_t lda = uint64_t(ls) * nS;
            MQ_REQUIRE(lda < (1ull << 32), "S leading dimension overflow");
            s_off[c] = ws;
            ws += align32(lda * stripe);
            for (uint64_t u = S.coh_unit_begin[c]; u < S.coh_unit_begin[c + 1]; ++u) {
                uint32_t ms = S.unit_ms[u], off = S.unit_offset[u];
                uint64_t task = S.unit_task_begin[u];
                for (uint64_t l = S.unit_line_begin[u]; l < S.unit_line_begin[u + 1]; ++l) {
                    uint32_t b = S.line_b1[l], cnt = S.line_count[l];
                    MQ_REQUIRE(b < nS && uint64_t(off) + ms <= stripe, "S region outside S");
                    GatherRegion r{};
                    r.dst_off = s_off[c] + uint64_t(off) * lda + uint64_t(b) * ls;
                    r.R = ls;
                    r.C = ms;
                    r.ld_dst = uint32_t(lda);
                    r.t0 = uint32_t(P.src_off.size());
                    for (uint32_t k = 0; k < cnt; ++k, ++task) {
                        uint64_t ti = S.t_begin[m] + S.task_ti[task];
                        MQ_REQUIRE(ti < S.t_begin[m + 1], "task ti out of range");
                        P.src_off.push_back(t_off[ti] + uint64_t(S.task_col[task]) * ls);
                        P.alpha.push_back(S.task_alpha[task]);
                    }
                    r.t1 = uint32_t(P.src_off.size());
                    if (uint64_t(r.R) * r.C) P.regions.push_back(r);
                }
                MQ_REQUIRE(task == S.unit_task_begin[u + 1], "line counts do not match the task list");
            }
        }
    P.s_elems = ws - P.s_begin;
    return s_off;
}

}  // namespace

// ============================================================================ C-ABI
extern "C" {

const char* mq_last_error(void) { return g_err.c_str(); }

// Scratch of the allocation-free batched dot (mq_vec_multi_dot_noalloc): the per-workgroup partial
// sums of up to MQ_NOALLOC_MAXK vectors, one copy per GPU for the life of the process.  Host code
// never uses the symbol itself: each device's address comes from hipGetSymbolAddress at context
// creation (g_scratch_dev) and is what kernels receive.
static __device__ __align__(256) double scratch[MQ_NOALLOC_MAXK * RED_BLOCKS];
static std::mutex g_scratch_mu;
static std::map<int, double*> g_scratch_dev;   // device ordinal -> address of its scratch

static std::vector<mq_ctx*> g_contexts;
static void release_all_pools() {
    for (mq_ctx* c : g_contexts) {
        (void)hipStreamSynchronize(c->stream);
        c->pool_clear();
    }
}

int mq_ctx_create(int device, mq_ctx** out) {
    return guarded([&] {
        MQ_REQUIRE(out, "null output");
        auto c = std::make_unique<mq_ctx>();
        c->device = device;
        HIPCHK(hipSetDevice(device));
        {
            // once per device, with that device current
            std::lock_guard<std::mutex> lk(g_scratch_mu);
            auto it = g_scratch_dev.find(device);
            if (it == g_scratch_dev.end()) {
                void* ptr = nullptr;
                HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
                MQ_REQUIRE(ptr, "hipGetSymbolAddress returned no address for the static scratch");
                it = g_scratch_dev.emplace(device, static_cast<double*>(ptr)).first;
            }
            c->noalloc_scratch = it->second;
        }
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIPCHK(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
        BLASCHK(rocblas_create_handle(&c->blas));
        BLASCHK(rocblas_set_stream(c->blas, c->stream));
        LTCHK(hipblasLtCreate(&c->lt));
        c->lt_ws.alloc(LT_WORKSPACE);
        c->red_partial.alloc(RED_BLOCKS * sizeof(double));
        c->red_result.alloc(64 * sizeof(double));
        g_release_pools = release_all_pools;
        g_contexts.push_back(c.get());
        *out = c.release();
    });
}

int mq_ctx_destroy(mq_ctx* ctx) {
    return guarded([&] {
        if (!ctx) return;
        (void)hipStreamSynchronize(ctx->stream);
        ctx->drop_batches();
        for (auto h : ctx->side_blas) (void)rocblas_destroy_handle(h);
        for (auto sd : ctx->side) (void)hipStreamDestroy(sd);
        for (auto e : ctx->side_ev) (void)hipEventDestroy(e);
        if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
        ctx->pool_clear();
        if (ctx->wd_host) (void)hipHostFree(ctx->wd_host);
        g_contexts.erase(std::remove(g_contexts.begin(), g_contexts.end(), ctx), g_contexts.end());
        if (ctx->blas) (void)rocblas_destroy_handle(ctx->blas);
        if (ctx->lt) (void)hipblasLtDestroy(ctx->lt);
        (void)hipStreamSynchronize(ctx->copy_stream);
        (void)hipStreamDestroy(ctx->copy_stream);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
    });
}

int mq_ctx_synchronize(mq_ctx* ctx) {
    return guarded([&] {
        MQ_REQUIRE(ctx, "null ctx");
        HIPCHK(hipStreamSynchronize(ctx->stream));
    });
}

void* mq_ctx_stream(mq_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

// ---------------------------------------------------------------------------- layouts
int mq_site_vector_layout(const mq_site_desc* site, uint64_t* n, uint64_t* offs) {
    return guarded([&] {
        MQ_REQUIRE(site && n, "null argument");
        SiteLayout L(*site);
        *n = L.rp_total;
        if (offs) std::copy(L.rp_offset.begin(), L.rp_offset.end(), offs);
    });
}

int mq_site_left_layout(const mq_site_desc* site, uint64_t* n, uint64_t* offs) {
    return guarded([&] {
        MQ_REQUIRE(site && n, "null argument");
        SiteLayout L(*site);
        *n = L.lp_total;
        if (offs) std::copy(L.lp_offset.begin(), L.lp_offset.end(), offs);
    });
}

}  // extern "C"

// Reshape right-paired <-> left-paired.  Per (left sector lb, physical sector s) slab the
// right-paired elements are contiguous; the left-paired destination is a strided 3-d copy.
// The device kernel is generated from the slab list through the gather machinery: a slab is
// d_s*rs regions of ls contiguous elements with alpha = 1.
namespace {
int reshape(mq_ctx* ctx, const mq_site_desc* site, const double* src, double* dst, bool to_left) {
    return guarded([&] {
        MQ_REQUIRE(ctx && site && src && dst, "null argument");
        SiteLayout L(*site);
        ProductBasis rpb(L.phys_i, L.right_i, -1), lpb(L.phys_i, L.left_i, +1);
        // One "region" per (lb, s, ss): R = ls rows, C = rs columns.
        // to_left : dst = left-paired (column stride lp_rows), src = right-paired contiguous slab
        // to_right: dst = right-paired (column stride ls),     src = left-paired strided -> use R=ls,
        //           C=1 regions per column (contiguous in both layouts).
        std::vector<GatherRegion> regs;
        std::vector<uint64_t> so;
        std::vector<double> al;
        for (uint32_t lb = 0; lb < L.left_i.size(); ++lb)
            for (uint32_t s = 0; s < L.phys_i.size(); ++s) {
                uint32_t rb = L.right_i.position(L.left_i.charges[lb] + L.phys_i.charges[s]);
                if (rb == L.right_i.size()) continue;
                uint32_t ls = L.left_i.sizes[lb], rs = L.right_i.sizes[rb], d = L.phys_i.sizes[s];
                uint64_t col0 = rpb(s, rb), row0 = lpb(s, lb);
                for (uint32_t ss = 0; ss < d; ++ss) {
                    uint64_t rp = L.rp_offset[lb] + (col0 + uint64_t(ss) * rs) * ls;   // ls x rs, ld ls
                    uint64_t lp = L.lp_offset[rb] + row0 + uint64_t(ss) * ls;          // ls x rs, ld lp_rows
                    if (to_left) {
                        GatherRegion r{};
                        r.dst_off = lp;
                        r.R = ls;
                        r.C = rs;
                        r.ld_dst = L.lp_rows[rb];
                        r.t0 = uint32_t(so.size());
                        so.push_back(rp);
                        al.push_back(1.0);
                        r.t1 = uint32_t(so.size());
                        if (ls && rs) regs.push_back(r);
                    } else {
                        for (uint32_t c = 0; c < rs; ++c) {
                            GatherRegion r{};
                            r.dst_off = rp + uint64_t(c) * ls;
                            r.R = ls;
                            r.C = 1;
                            r.ld_dst = ls;
                            r.t0 = uint32_t(so.size());
                            so.push_back(lp + uint64_t(c) * L.lp_rows[rb]);
                            al.push_back(1.0);
                            r.t1 = uint32_t(so.size());
                            if (ls) regs.push_back(r);
                        }
                    }
                }
            }
        StagePlan P;
        P.regions = std::move(regs);
        P.src_off = std::move(so);
        P.alpha = std::move(al);
        P.finish_tables();
        // validation: gather reads from src, writes to dst
        uint64_t n_src = to_left ? L.rp_total : L.lp_total, n_dst = to_left ? L.lp_total : L.rp_total;
        P.for_regions([&](const GatherRegion& r) {
            MQ_REQUIRE(StagePlan::extent(r.R, r.C, r.ld_dst) + r.dst_off <= n_dst, "reshape dst out of range");
            for (uint32_t t = r.t0; t < r.t1; ++t)
                MQ_REQUIRE(P.src_off[t] + uint64_t(r.R) * r.C <= n_src, "reshape src out of range");
        });
        P.upload_all(ctx->stream);
        HIPCHK(launch_fill(ctx->stream, dst, n_dst, 0.0));
        P.run_gather(ctx->stream, const_cast<double*>(src), dst);
        HIPCHK(hipStreamSynchronize(ctx->stream));   // tables are freed on return
    });
}
}  // namespace

extern "C" {

int mq_reshape_right_to_left(mq_ctx* ctx, const mq_site_desc* site, const double* src, double* dst) {
    return reshape(ctx, site, src, dst, true);
}
int mq_reshape_left_to_right(mq_ctx* ctx, const mq_site_desc* site, const double* src, double* dst) {
    return reshape(ctx, site, src, dst, false);
}

// ---------------------------------------------------------------------------- boundaries
// Residency of a boundary: bring an evicted boundary back to the device (allocate, copy its host
// copy, read from the spill file first if there is none, on ctx->copy_stream, record `arrived`).
static void boundary_prefetch(mq_boundary* b) {
    if (b->resident) return;
    mq_ctx* ctx = b->ctx;
    const size_t bytes = b->total * sizeof(double);
    if (!b->host) {
        MQ_REQUIRE(!b->disk.empty(), "evicted boundary without a host or disk copy");
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&b->host), bytes, hipHostMallocDefault));
        FILE* f = std::fopen(b->disk.c_str(), "rb");
        MQ_REQUIRE(f != nullptr, "cannot open the boundary spill file " + b->disk);
        const size_t got = std::fread(b->host, 1, bytes, f);
        std::fclose(f);
        MQ_REQUIRE(got == bytes, "short read of the boundary spill file " + b->disk);
    }
    b->data.alloc(bytes);
    if (!b->arrived) HIPCHK(hipEventCreateWithFlags(&b->arrived, hipEventDisableTiming));
    HIPCHK(hipMemcpyAsync(b->data.p, b->host, bytes, hipMemcpyHostToDevice, ctx->copy_stream));
    HIPCHK(hipEventRecord(b->arrived, ctx->copy_stream));
    b->resident = true;
}

// Make the work queued next on `st` see the boundary's device data, prefetching it first if it was
// evicted (the reference fetches a boundary the same way before use, storage::gpu::fetch).  Only
// residency changes: the data itself is immutable.
static void boundary_ready(const mq_boundary* bc, hipStream_t st) {
    mq_boundary* b = const_cast<mq_boundary*>(bc);
    boundary_prefetch(b);
    if (b->arrived) HIPCHK(hipStreamWaitEvent(st, b->arrived, 0));
}

void mq_plan::bind(const mq_boundary* l, const mq_boundary* r) {
    if (l) {
        boundary_ready(l, ctx->stream);
        left = l;
        ++l->plans;
    }
    if (r) {
        boundary_ready(r, ctx->stream);
        right = r;
        ++r->plans;
    }
}

int mq_boundary_evict(mq_ctx* ctx, mq_boundary* b, const char* path) {
    return guarded([&] {
        MQ_REQUIRE(ctx && b && b->ctx == ctx, "bad argument");
        MQ_REQUIRE(b->plans == 0, "boundary read by a live sigma plan: destroy the plan first");
        const size_t bytes = b->total * sizeof(double);
        if (b->resident) {
            if (b->arrived) HIPCHK(hipEventSynchronize(b->arrived));
            if (!b->host && b->disk.empty()) {
                HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&b->host), bytes, hipHostMallocDefault));
                HIPCHK(hipMemcpyAsync(b->host, b->data.p, bytes, hipMemcpyDeviceToHost, ctx->stream));
            }
            HIPCHK(hipStreamSynchronize(ctx->stream));   // queued work that reads the boundary, and the copy
            b->data.release();
            b->resident = false;
        }
        if (path) {
            if (b->disk.empty()) {
                MQ_REQUIRE(b->host != nullptr, "boundary without a host copy to spill");
                FILE* f = std::fopen(path, "wb");
                MQ_REQUIRE(f != nullptr, std::string("cannot create the boundary spill file ") + path);
                const size_t put = std::fwrite(b->host, 1, bytes, f);
                const int rc = std::fclose(f);
                MQ_REQUIRE(put == bytes && rc == 0, std::string("write of the boundary spill file failed: ") + path);
                b->disk = path;
            }
            if (b->host) (void)hipHostFree(b->host), b->host = nullptr;
        }
    });
}

int mq_boundary_prefetch(mq_ctx* ctx, mq_boundary* b) {
    return guarded([&] {
        MQ_REQUIRE(ctx && b && b->ctx == ctx, "bad argument");
        boundary_prefetch(b);
    });
}

int mq_boundary_residency(const mq_boundary* b, int* where, uint64_t* device_bytes, uint64_t* host_bytes) {
    return guarded([&] {
        MQ_REQUIRE(b && where, "bad argument");
        *where = b->resident ? MQ_RESIDENT_DEVICE : b->host ? MQ_RESIDENT_HOST : MQ_RESIDENT_DISK;
        if (device_bytes) *device_bytes = b->data.bytes;
        if (host_bytes) *host_bytes = b->host ? b->total * sizeof(double) : 0;
    });
}

int mq_boundary_create_trivial(mq_ctx* ctx, const mq_index_desc* idx, uint32_t aux_dim, mq_boundary** out) {
    return guarded([&] {
        MQ_REQUIRE(ctx && idx && out && aux_dim > 0, "bad argument");
        auto b = std::make_unique<mq_boundary>();
        b->ctx = ctx;
        Index I(*idx);
        b->index = BoundaryIndex(I, I, aux_dim);
        // Boundary(ud, ld, ad): one diagonal cohort per sector, blocks b*block_size (boundary.h:258-283)
        for (uint32_t i = 0; i < I.size(); ++i) {
            std::vector<int64_t> off(aux_dim);
            uint64_t bs = uint64_t(I.sizes[i]) * I.sizes[i];
            for (uint32_t k = 0; k < aux_dim; ++k) off[k] = int64_t(k * bs);
            b->index.add_cohort(i, i, off);
        }
        b->layout();
        b->data.alloc(b->total * sizeof(double));
        HIPCHK(launch_fill(ctx->stream, b->ptr(), b->total, 1.0));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        *out = b.release();
    });
}

int mq_boundary_create(mq_ctx* ctx, const mq_bindex_desc* idx, const double* const* host, mq_boundary** out) {
    return guarded([&] {
        MQ_REQUIRE(ctx && idx && out, "bad argument");
        auto b = std::make_unique<mq_boundary>();
        b->ctx = ctx;
        b->index = BoundaryIndex(*idx);
        for (uint32_t ci = 0; ci < b->index.n_cohorts(); ++ci)
            for (uint32_t k = 0; k < b->index.aux_dim; ++k) {
                int64_t o = b->index.offset(ci, k);
                uint32_t owner = b->index.tr[ci] ? b->index.cohort_index(b->index.ket.charges[b->index.ket_pos[ci]],
                                                                         b->index.bra.charges[b->index.bra_pos[ci]])
                                                 : ci;
                if (o > -1 && owner < b->index.n_cohorts())
                    MQ_REQUIRE(uint64_t(o) + b->index.block_size(ci) <= b->index.cohort_size(owner) ||
                                   b->index.transposes[size_t(ci) * b->index.aux_dim + k],
                               "boundary offset beyond cohort data");
            }
        b->layout();
        b->data.alloc(b->total * sizeof(double));
        HIPCHK(launch_fill(ctx->stream, b->ptr(), b->total, 0.0));
        if (host)
            for (uint32_t ci = 0; ci < b->index.n_cohorts(); ++ci) {
                uint64_t n = b->index.cohort_size(ci);
                if (n && host[ci])
                    HIPCHK(hipMemcpyAsync(b->ptr() + b->cohort_off[ci], host[ci], n * sizeof(double),
                                          hipMemcpyHostToDevice, ctx->stream));
            }
        HIPCHK(hipStreamSynchronize(ctx->stream));
        *out = b.release();
    });
}

int mq_boundary_destroy(mq_boundary* b) {
    return guarded([&] { delete b; });
}

int mq_boundary_index(const mq_boundary* b, mq_bindex_desc* out) {
    return guarded([&] {
        MQ_REQUIRE(b && out, "bad argument");
        b->index.fill_desc(out, b->bra_c, b->ket_c);
    });
}

int mq_boundary_download(const mq_boundary* b, uint32_t ci, double* host) {
    return guarded([&] {
        MQ_REQUIRE(b && host && ci < b->index.n_cohorts(), "bad argument");
        boundary_ready(b, b->ctx->stream);
        uint64_t n = b->index.cohort_size(ci);
        if (n)
            HIPCHK(hipMemcpyAsync(host, b->ptr() + b->cohort_off[ci], n * sizeof(double), hipMemcpyDeviceToHost,
                                  b->ctx->stream));
        HIPCHK(hipStreamSynchronize(b->ctx->stream));
    });
}

int mq_boundary_cohort_ptr(const mq_boundary* b, uint32_t ci, const double** dev) {
    return guarded([&] {
        MQ_REQUIRE(b && dev && ci < b->index.n_cohorts(), "bad argument");
        boundary_ready(b, b->ctx->stream);
        *dev = b->ptr() + b->cohort_off[ci];
    });
}

int mq_boundary_traces(const mq_boundary* b, double* out) {
    return guarded([&] {
        MQ_REQUIRE(b && out, "bad argument");
        boundary_ready(b, b->ctx->stream);
        const BoundaryIndex& I = b->index;
        std::vector<uint64_t> off;
        std::vector<uint32_t> len, bb;
        for (uint32_t ci = 0; ci < I.n_cohorts(); ++ci) {
            if (I.tr[ci]) continue;
            for (uint32_t k = 0; k < I.aux_dim; ++k) {
                int64_t o = I.offset(ci, k);
                if (o < 0 || I.transposes[size_t(ci) * I.aux_dim + k]) continue;
                off.push_back(b->cohort_off[ci] + uint64_t(o));
                len.push_back(uint32_t(I.block_size(ci)));
                bb.push_back(k);
            }
        }
        std::fill(out, out + I.aux_dim, 0.0);
        if (off.empty()) return;
        DevBuf d_off, d_len, d_out;
        upload(d_off, off, b->ctx->stream);
        upload(d_len, len, b->ctx->stream);
        d_out.alloc(off.size() * sizeof(double));
        HIPCHK(launch_block_sums(b->ctx->stream, d_off.as<uint64_t>(), d_len.as<uint32_t>(), nullptr,
                                 uint32_t(off.size()), b->ptr(), d_out.as<double>()));
        std::vector<double> part(off.size());
        HIPCHK(hipMemcpyAsync(part.data(), d_out.p, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                              b->ctx->stream));
        HIPCHK(hipStreamSynchronize(b->ctx->stream));
        for (size_t i = 0; i < part.size(); ++i) out[bb[i]] += part[i];
    });
}

// ---------------------------------------------------------------------------- SU2
int mq_su2_coupling(int a, int b, int c, int d, int e, int f, int g, int h, int i, double init, double* out4) {
    return guarded([&] {
        MQ_REQUIRE(out4, "null output");
        su2_set_coupling(a, b, c, d, e, f, g, h, i, init, out4);
    });
}

double mq_wigner9j(int a, int b, int c, int d, int e, int f, int g, int h, int i) {
    return wigner9j(a, b, c, d, e, f, g, h, i);
}

// ---------------------------------------------------------------------------- schedules
int mq_schedule_build_sigma(const mq_site_desc* site, const mq_mpo_desc* mpo, const mq_bindex_desc* left,
                            const mq_bindex_desc* right, mq_schedule** out) {
    return guarded([&] {
        MQ_REQUIRE(site && mpo && left && right && out, "bad argument");
        SiteLayout L(*site);
        Mpo W(*mpo);
        BoundaryIndex LI(*left), RI(*right);
        auto s = std::make_unique<mq_schedule>();
        s->s = build_sigma_schedule(L, W, LI, RI);
        s->flops = schedule_flops(s->s, RI, L.left_i.sizes);
        *out = s.release();
    });
}

int mq_schedule_build_left(const mq_site_desc* site, const mq_mpo_desc* mpo, const mq_bindex_desc* left,
                           int symmetric, mq_schedule** out) {
    return guarded([&] {
        MQ_REQUIRE(site && mpo && left && out, "bad argument");
        SiteLayout L(*site);
        Mpo W(*mpo);
        BoundaryIndex LI(*left);
        auto s = std::make_unique<mq_schedule>();
        s->s = build_left_schedule(L, W, LI, symmetric != 0);
        s->flops = schedule_flops(s->s, LI, L.right_i.sizes);
        *out = s.release();
    });
}

int mq_schedule_build_right(const mq_site_desc* site, const mq_mpo_desc* mpo, const mq_bindex_desc* right,
                            int symmetric, mq_schedule** out) {
    return guarded([&] {
        MQ_REQUIRE(site && mpo && right && out, "bad argument");
        SiteLayout L(*site);
        Mpo W(*mpo);
        BoundaryIndex RI(*right);
        auto s = std::make_unique<mq_schedule>();
        s->s = build_right_schedule(L, W, RI, symmetric != 0);
        s->flops = schedule_flops(s->s, RI, L.left_i.sizes);
        *out = s.release();
    });
}

int mq_schedule_build_left_index(const mq_site_desc* site, const mq_mpo_desc* mpo, const mq_bindex_desc* left,
                                 int symmetric, mq_schedule** out) {
    return guarded([&] {
        MQ_REQUIRE(site && mpo && left && out, "bad argument");
        SiteLayout L(*site);
        Mpo W(*mpo);
        BoundaryIndex LI(*left);
        auto s = std::make_unique<mq_schedule>();
        s->s.kind = Schedule::LEFT;
        s->s.aux_dim = W.col_dim;
        s->s.out_index = left_step_index(L, W, LI, symmetric != 0);
        *out = s.release();
    });
}

int mq_schedule_build_right_index(const mq_site_desc* site, const mq_mpo_desc* mpo, const mq_bindex_desc* right,
                                  int symmetric, mq_schedule** out) {
    return guarded([&] {
        MQ_REQUIRE(site && mpo && right && out, "bad argument");
        SiteLayout L(*site);
        Mpo W(*mpo);
        BoundaryIndex RI(*right);
        auto s = std::make_unique<mq_schedule>();
        s->s.kind = Schedule::RIGHT;
        s->s.aux_dim = W.row_dim;
        s->s.out_index = right_step_index(L, W, RI, symmetric != 0);
        *out = s.release();
    });
}

int mq_schedule_out_index(const mq_schedule* s, mq_bindex_desc* out) {
    return guarded([&] {
        MQ_REQUIRE(s && out, "bad argument");
        MQ_REQUIRE(s->s.kind != Schedule::SIGMA, "sigma schedules produce no boundary");
        s->s.out_index.fill_desc(out, s->bra_c, s->ket_c);
    });
}

int mq_schedule_export(const mq_schedule* s, mq_schedule_desc* out) {
    return guarded([&] {
        MQ_REQUIRE(s && out, "bad argument");
        s->s.export_desc(out);
    });
}

int mq_schedule_flops(const mq_schedule* s, uint64_t* flops) {
    return guarded([&] {
        MQ_REQUIRE(s && flops, "bad argument");
        *flops = s->flops;
    });
}

int mq_schedule_destroy(mq_schedule* s) {
    return guarded([&] { delete s; });
}

}  // extern "C"

// ============================================================================ sigma plans
namespace {

// Lower a sigma schedule: T stage, S stage, then per output block one K-segmented GEMM over
// every cohort that writes it (Cohort::contract, tasks.cpp:286-344).
void lower_sigma(mq_plan* p, const Schedule& S, const std::vector<uint32_t>& ket_rows,
                 const std::vector<uint32_t>& ket_cols, const std::vector<uint64_t>& ket_off) {
    const BoundaryIndex& LI = p->left->index;
    const BoundaryIndex& RI = p->right->index;
    StagePlan& P = p->st;
    uint64_t ws = 0;
    std::vector<uint64_t> t_off = lower_T_right(P, S, RI, p->right->cohort_off, ket_rows, ket_off, ws);
    std::vector<uint64_t> s_off = lower_S_right(P, S, t_off, ws);
    // group cohorts by output block
    std::vector<std::vector<uint32_t>> by_out(ket_rows.size());
    for (uint32_t c = 0; c < S.n_cohorts(); ++c) {
        MQ_REQUIRE(S.coh_rb[c] < ket_rows.size(), "cohort output block out of range");
        by_out[S.coh_rb[c]].push_back(c);
    }
    for (uint32_t lbo = 0; lbo < ket_rows.size(); ++lbo) {
        if (by_out[lbo].empty()) continue;
        std::vector<GemmSeg> segs;
        uint32_t M = ket_rows[lbo], N = ket_cols[lbo];
        for (uint32_t c : by_out[lbo]) {
            uint32_t ci = S.coh_ci[c], cie = S.coh_ci_eff[c];
            uint32_t ls_in = S.coh_ls[c], ls_out = S.coh_rs[c], nS = S.coh_nSrows[c];
            MQ_REQUIRE(ls_out == M && S.coh_stripe[c] == N, "cohort shape does not match the output block");
            MQ_REQUIRE(cie < LI.n_cohorts() && uint64_t(nS) * ls_in * ls_out <= LI.cohort_size(cie),
                       "left cohort smaller than the schedule assumes");
            uint64_t ldS = uint64_t(nS) * ls_in;
            if (ci == cie) {
                segs.push_back(seg(BASE_LEFT, p->left->cohort_off[cie], ls_out, false, BASE_WS, s_off[c],
                                   uint32_t(ldS), false, uint32_t(ldS)));
            } else {
                for (uint32_t b = 0; b < nS; ++b)
                    segs.push_back(seg(BASE_LEFT, p->left->cohort_off[cie] + uint64_t(b) * ls_in * ls_out, ls_in, true,
                                       BASE_WS, s_off[c] + uint64_t(b) * ls_in, uint32_t(ldS), false, ls_in));
            }
        }
        add_gemm(P.f_tasks, P.f_segs, BASE_OUT, ket_off[lbo], M, N, M, segs);
    }
    P.ws_elems = ws;
}

// ---------------------------------------------------------------------------- grouped sigma
// sigma(lc_out; s_out, rc_out) = sum_{b1, lc_in} L_b1(lc_out, lc_in) S_b1(lc_in; s_out, rc_out)
// S_b1(lc_in; s_out, rc_out)  = sum_{b2, s_in} W_b1b2(s_out <- s_in) T_b2(lc_in; s_in, rc_out)
// For abelian groups the MPO coefficient matrix W[(b2, s_in, ss1), (b1, s_out, ss2)] does not
// depend on the sector pair (lc_in, rc_out) beyond its charge transfer D = rc_out - lc_in.  Pairs
// are grouped by D; per group T_D (rows: pair slab elements, columns k = (b2, s_in, ss1)) is written
// by the T GEMMs through a column-block remap, S_D = T_D W_D is one dense MFMA GEMM, and the final
// GEMMs read S_D through a row-block remap.  The conjugation scales of transposed boundary blocks
// are applied as per-segment alpha.  Same result as the reference's T / S / L stages
// (solver/tasks.cpp:286-344, 466-494, 650-752) with the MPO-term sum done as a dense product.
struct GroupedLowering {
    // keys (b, s, ss) (MPO index, physical sector, state in the sector) numbered densely as
    // b * nst + soff[s] + ss, which is ascending in (b, s, ss)
    uint32_t nst = 0, nkeys = 0;
    std::vector<uint32_t> soff;
    uint32_t key(uint32_t b, uint32_t s, uint32_t ss) const { return b * nst + soff[s] + ss; }
    struct Pair {
        uint32_t lb_in, rb_out, g;
        uint64_t m_off;
        std::vector<uint32_t> kk, nn;   // indices into the group's K / N lists
    };
    struct Group {
        Charge delta;
        int spin = 0;   // SU2: spin of the pairs' left sector (one value per group)
        std::vector<uint32_t> K, N;          // the group's keys, ascending
        std::vector<int32_t> kpos, npos;     // key -> position in K / N, -1 if absent
        uint64_t M = 0, t_off = 0, s_off = 0, w_off = 0, bf_off = 0;
        std::vector<uint32_t> pairs;
        uint32_t kid(uint32_t k) const {
            MQ_REQUIRE(kpos[k] >= 0, "key not in its group");
            return uint32_t(kpos[k]);
        }
        uint32_t nid(uint32_t k) const {
            MQ_REQUIRE(npos[k] >= 0, "key not in its group");
            return uint32_t(npos[k]);
        }
    };
    std::vector<Pair> pairs;
    std::vector<Group> groups;
    std::unordered_map<uint64_t, uint32_t> pair_of;   // (lb_in, rb_out) -> pair
};

// The S-stage part of the reference's flop model of a grouped sigma (tasks.cpp:388-398, 777-796):
// one alpha (2 flops per element of the pair's slab) per nonzero MPO coupling (k, n) between a
// key k the pair multiplies in the T stage and a key n it contracts in the final stage.
uint64_t grouped_s_flops(const GroupedLowering& G, const double* wd, uint64_t wd_off, const std::vector<uint32_t>& lsz,
                         const std::vector<uint32_t>& rsz) {
    uint64_t f = 0;
    for (const auto& g : G.groups) {
        const uint64_t Kg = g.K.size(), Ng = g.N.size();
        const double* Wg = wd + (g.w_off - wd_off);
        std::vector<std::vector<uint32_t>> col_nz(Ng);
        for (uint64_t n = 0; n < Ng; ++n)
            for (uint64_t k = 0; k < Kg; ++k)
                if (Wg[n * Kg + k] != 0.0) col_nz[n].push_back(uint32_t(k));
        std::vector<char> kin(Kg);
        for (uint32_t pi : g.pairs) {
            const auto& pr = G.pairs[pi];
            std::fill(kin.begin(), kin.end(), 0);
            for (auto k : pr.kk) kin[k] = 1;
            uint64_t na = 0;
            for (auto n : pr.nn)
                for (auto k : col_nz[n]) na += kin[k];
            f += 2ull * lsz[pr.lb_in] * rsz[pr.rb_out] * na;
        }
    }
    return f;
}

// A boundary step lowered by the grouped path: kind Schedule::LEFT (T = L psi) or RIGHT (T = psi R),
// the new boundary's index and cohort offsets (the final stage's output).
struct GroupedStep {
    int kind;
    const BoundaryIndex* out;
    const std::vector<uint64_t>* out_off;
    // only sector pairs whose ket sector (right step: left sector of the pair; left step: right
    // sector) lies in [lo, hi): a step is lowered in such chunks to bound its workspace
    uint32_t lo = 0, hi = 0xffffffffu;
};

// One nonzero of a group's W_D: column n (N key position), row k (K key position), value.
struct WEntry {
    uint32_t n, k;
    double v;
};
// Allocator whose resize() leaves trivial elements uninitialised (the nonzero lists are filled
// completely right after they are sized; clearing hundreds of MB first would cost as much)
template <class T> struct UninitAlloc : std::allocator<T> {
    template <class U> struct rebind { using other = UninitAlloc<U>; };
    UninitAlloc() = default;
    template <class U> UninitAlloc(const UninitAlloc<U>&) {}
    template <class U> void construct(U* q) { ::new (static_cast<void*>(q)) U; }
    template <class U, class... A> void construct(U* q, A&&... a) { ::new (static_cast<void*>(q)) U(std::forward<A>(a)...); }
};
typedef std::vector<WEntry, UninitAlloc<WEntry>> WList;

// Pattern-class order of one group's keys (see lower_sigma_grouped): rows k and columns n of W_D
// are classed by a hash of their nonzero index lists; classes are ordered by size (largest first),
// then by their first key; the all-zero class goes last.  W holds the group's nonzeros in an order
// in which the columns n of every row k, and the rows k of every column n, ascend (as in a
// column-major scan of the dense matrix, which fixes the hashes); their positions, the group's
// K / N lists, kpos / npos and its pairs' kk / nn position lists are renumbered consistently (W is
// left in the old order of its elements).
void reorder_group_keys(GroupedLowering& G, GroupedLowering::Group& g, WList& W) {
    const uint32_t Kg = uint32_t(g.K.size()), Ng = uint32_t(g.N.size());
    if (Kg < 2 && Ng < 2) return;
    const uint64_t seed = 0x9E3779B97F4A7C15ull;
    auto mix = [](uint64_t h, uint64_t v) {
        h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
        return h * 0xFF51AFD7ED558CCDull;
    };
    std::vector<uint64_t> hk(Kg, seed), hn(Ng, seed);
    std::vector<uint32_t> ck(Kg, 0), cn(Ng, 0);   // nonzero counts (the all-zero class)
    for (const WEntry& e : W) {
        hk[e.k] = mix(hk[e.k], e.n + 1);
        hn[e.n] = mix(hn[e.n], e.k + 1);
        ++ck[e.k];
        ++cn[e.n];
    }
    auto order = [](const std::vector<uint64_t>& h, const std::vector<uint32_t>& cnt) {
        const uint32_t n = uint32_t(h.size());
        std::unordered_map<uint64_t, std::pair<uint32_t, uint32_t>> cls;   // hash -> (size, first index)
        for (uint32_t i = 0; i < n; ++i) {
            auto it = cls.find(h[i]);
            if (it == cls.end()) cls.emplace(h[i], std::make_pair(1u, i));
            else ++it->second.first;
        }
        std::vector<uint32_t> perm(n);
        for (uint32_t i = 0; i < n; ++i) perm[i] = i;
        std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) {
            const bool za = cnt[a] == 0, zb = cnt[b] == 0;
            if (za != zb) return zb;
            const auto& ca = cls.at(h[a]);
            const auto& cb = cls.at(h[b]);
            if (ca.first != cb.first) return ca.first > cb.first;
            return ca.second < cb.second;
        });
        return perm;   // new position -> old position
    };
    const std::vector<uint32_t> pk = order(hk, ck), pn = order(hn, cn);
    std::vector<uint32_t> ik(Kg), in(Ng);
    for (uint32_t i = 0; i < Kg; ++i) ik[pk[i]] = i;
    for (uint32_t i = 0; i < Ng; ++i) in[pn[i]] = i;
    for (WEntry& e : W) {
        e.n = in[e.n];
        e.k = ik[e.k];
    }
    std::vector<uint32_t> K2(Kg), N2(Ng);
    for (uint32_t i = 0; i < Kg; ++i) K2[i] = g.K[pk[i]], g.kpos[K2[i]] = int32_t(i);
    for (uint32_t i = 0; i < Ng; ++i) N2[i] = g.N[pn[i]], g.npos[N2[i]] = int32_t(i);
    g.K.swap(K2);
    g.N.swap(N2);
    for (uint32_t pi : g.pairs) {
        for (auto& k : G.pairs[pi].kk) k = ik[k];
        for (auto& n : G.pairs[pi].nn) n = in[n];
    }
}

// Block-sparse S stage (the nonzeros of each group's W_D, after reorder_group_keys): the
// columns are cut into panels of 128, the keys into tiles of 16; each panel keeps the tiles holding
// a nonzero, stored compactly ([n][k], zero padded) and grouped into runs of consecutive tiles; one
// task per 128-row panel of T_D and column panel, in blocks of 8 x 8 panels (below).  A panel
// without any nonzero gets no run: its tasks write zeros.  Groups are scanned and their tiles
// copied in parallel (each group owns a contiguous tile range); tasks and runs are appended in group
// order.  info receives per group (W_D nonzeros, visited fraction).
struct BsrGroupScan {
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> panel_runs;   // (first key tile, tiles) per panel
    uint64_t tiles = 0, nnz = 0, visited = 0;
};
void add_bsr_groups(StagePlan::Bsr& B, const std::vector<GroupedLowering::Group>& groups,
                    const std::vector<WList>& wg, std::vector<std::pair<uint64_t, double>>* info) {
    constexpr uint32_t TK = S_BSR_TILE_K, TN = S_BSR_TILE_N, TM = S_BSR_TILE_M;
    const size_t ng = groups.size();
    std::vector<BsrGroupScan> sc(ng);
    // per group and (panel, key tile): the tile's number within the group (in panel, then key
    // order), ~0u for a tile without a nonzero
    std::vector<std::vector<uint32_t>> at(ng);
#pragma omp parallel for schedule(dynamic, 1)
    for (int gi = 0; gi < int(ng); ++gi) {
        const auto& g = groups[gi];
        if (!g.M || g.K.empty() || g.N.empty()) continue;
        const uint32_t Kg = uint32_t(g.K.size()), Ng = uint32_t(g.N.size());
        const uint32_t nkt = (Kg + TK - 1) / TK, npn = (Ng + TN - 1) / TN;
        BsrGroupScan& S = sc[gi];
        S.panel_runs.resize(npn);
        std::vector<uint32_t>& A = at[gi];
        A.assign(uint64_t(npn) * nkt, ~0u);
        for (const WEntry& e : wg[gi]) A[uint64_t(e.n / TN) * nkt + e.k / TK] = 0;
        S.nnz = wg[gi].size();
        uint32_t local = 0;
        for (uint32_t j = 0; j < npn; ++j) {
            const uint32_t n0 = j * TN, n1 = std::min(Ng, n0 + TN);
            uint32_t* z = A.data() + uint64_t(j) * nkt;
            for (uint32_t t = 0; t < nkt;) {
                if (z[t] == ~0u) {
                    ++t;
                    continue;
                }
                uint32_t t1 = t;
                while (t1 < nkt && z[t1] != ~0u) z[t1++] = local++;
                S.panel_runs[j].push_back({t, t1 - t});
                S.tiles += t1 - t;
                S.visited += uint64_t(std::min(Kg, t1 * TK) - t * TK) * (n1 - n0);
                t = t1;
            }
        }
    }
    // tile ranges, then each nonzero's place in the compact tiles ([n][k] per tile, zero padded):
    // the device fills the tiles from these (position, value) pairs
    std::vector<uint64_t> tile0(ng + 1, B.tiles);
    for (size_t gi = 0; gi < ng; ++gi) tile0[gi + 1] = tile0[gi] + sc[gi].tiles;
    MQ_REQUIRE(tile0[ng] < (1ull << 32), "too many block-sparse tiles");
    MQ_REQUIRE(B.n_ent == 0 && B.ent_pos && B.ent_val, "block-sparse entries: one lowering per staging");
    std::vector<uint64_t> e0(ng + 1, 0);
    for (size_t gi = 0; gi < ng; ++gi) e0[gi + 1] = e0[gi] + (sc[gi].tiles ? wg[gi].size() : 0);
    B.n_ent = e0[ng];
#pragma omp parallel for schedule(dynamic, 1)
    for (int gi = 0; gi < int(ng); ++gi) {
        if (!sc[gi].tiles) continue;
        const uint32_t nkt = (uint32_t(groups[gi].K.size()) + TK - 1) / TK;
        const std::vector<uint32_t>& A = at[gi];
        uint64_t q = e0[gi];
        for (const WEntry& e : wg[gi]) {
            const uint32_t j = e.n / TN, t = e.k / TK;
            const uint64_t tile = tile0[gi] + A[uint64_t(j) * nkt + t];
            B.ent_pos[q] = tile * (uint64_t(TN) * TK) + uint64_t(e.n - j * TN) * TK + (e.k - t * TK);
            B.ent_val[q] = e.v;
            ++q;
        }
    }
    // runs and tasks in group order
    for (size_t gi = 0; gi < ng; ++gi) {
        const auto& g = groups[gi];
        const BsrGroupScan& S = sc[gi];
        if (info) info->push_back({S.nnz, g.K.empty() || g.N.empty() ? 0.0 :
                                              double(S.visited) / (double(g.K.size()) * double(g.N.size()))});
        if (!g.M || g.K.empty() || g.N.empty()) continue;
        MQ_REQUIRE(g.M < (1ull << 32), "group too large for the block-sparse S stage");
        const uint32_t Kg = uint32_t(g.K.size()), Ng = uint32_t(g.N.size()), M = uint32_t(g.M);
        const uint32_t npn = uint32_t(S.panel_runs.size());
        uint64_t tile = tile0[gi];
        std::vector<std::pair<uint32_t, uint32_t>> pr(npn);
        for (uint32_t j = 0; j < npn; ++j) {
            pr[j].first = uint32_t(B.runs.size());
            for (const auto& r : S.panel_runs[j]) {
                const uint32_t k0 = r.first * TK, k1 = std::min(Kg, (r.first + r.second) * TK);
                B.runs.push_back({k0, k1 - k0, uint32_t(tile), 0u});
                for (uint32_t t = 0; t < r.second; ++t)
                    B.tile_at.push_back({uint32_t(gi), (r.first + t) * TK, j * TN});
                tile += r.second;
            }
            pr[j].second = uint32_t(B.runs.size());
        }
        // tasks in blocks of 8 row panels x 8 column panels, row panel fastest inside a block: the ~64
        // workgroups an XCD runs at once then share 8 T_D row panels and 8 W panels in its L2, instead
        // of one row panel with every W panel of the group (or the reverse)
        constexpr uint32_t SB = 8;
        const uint32_t npm = (M + TM - 1) / TM;
        B.narrow = B.narrow && (32ull * M + TM) * sizeof(double) < (1ull << 32);
        for (uint32_t jb = 0; jb < npn; jb += SB)
            for (uint32_t mb = 0; mb < npm; mb += SB)
                for (uint32_t j = jb; j < std::min(npn, jb + SB); ++j)
                    for (uint32_t mp = mb; mp < std::min(npm, mb + SB); ++mp) {
                        const uint32_t m0 = mp * TM, mr = std::min(TM, M - m0);
                        const uint32_t n0 = j * TN, nw = std::min(Ng, n0 + TN) - n0;
                        B.tasks.push_back({g.t_off + m0, g.s_off + uint64_t(n0) * g.M + m0, mr, nw, M, pr[j].first,
                                           pr[j].second, 0u});
                    }
        B.flops += 2ull * g.M * S.visited;
    }
    B.tiles = tile0[ng];
}

// The dense W_D go to p->ctx's pinned staging buffer (wd_elems of them, at ws_init_off of the
// workspace), which the caller uploads and synchronizes before the next lowering.
void lower_sigma_grouped(mq_plan* p, const SiteLayout& site, const Mpo& W, uint64_t& wd_elems,
                         uint64_t& ws_init_off, bool adjoint, const GroupedStep* step = nullptr) {
    const Index &left_i = site.left_i, &right_i = site.right_i, &phys_i = site.phys_i;
    const int kind = step ? step->kind : Schedule::SIGMA;
    const bool sigma = kind == Schedule::SIGMA, rstep = kind == Schedule::RIGHT, lstep = kind == Schedule::LEFT;
    const Symm& sy = site.symm;
    MQ_REQUIRE(!(adjoint && !sigma), "the adjoint is a sigma-plan option");
    static const BoundaryIndex none;
    const BoundaryIndex& LI = p->left ? p->left->index : none;
    const BoundaryIndex& RI = p->right ? p->right->index : none;
    const BoundaryIndex& NI = step ? *step->out : none;   // the new boundary of a step
    if (!rstep) {
        MQ_REQUIRE(p->left && LI.aux_dim == W.row_dim, "left boundary aux_dim != mpo row dimension");
        require_bond(LI, left_i, "left");
    }
    if (!lstep) {
        MQ_REQUIRE(p->right && RI.aux_dim == W.col_dim, "right boundary aux_dim != mpo column dimension");
        require_bond(RI, right_i, "right");
    }
    if (rstep) MQ_REQUIRE(NI.aux_dim == W.row_dim, "new right boundary aux_dim != mpo row dimension");
    if (lstep) MQ_REQUIRE(NI.aux_dim == W.col_dim, "new left boundary aux_dim != mpo column dimension");
    ProductBasis rpb(phys_i, right_i, -1);
    GroupedLowering G;
    G.soff.resize(phys_i.size());
    for (uint32_t s = 0; s < phys_i.size(); ++s) {
        G.soff[s] = G.nst;
        G.nst += phys_i.sizes[s];
    }
    {
        const uint64_t nk = uint64_t(std::max(W.row_dim, W.col_dim)) * G.nst;
        MQ_REQUIRE(nk < (1ull << 31), "MPO bond dimension x physical dimension too large");
        G.nkeys = uint32_t(nk);
    }
    // Pairs are grouped by the charge transfer D = rc_out - lc_in and, for SU2, by the spin of lc_in:
    // the spins of all four sectors a coupling of sigma or a step involves then follow from the
    // group and the physical sectors of the (k, n) keys, so every SU2 9j factor is a constant of
    // the group's W_D.  (delta key, spin) -> group, ordered: deterministic numbering.
    std::map<std::pair<uint64_t, int>, uint32_t> group_of;
    // stored block j of a (non-transposed) cohort -> MPO index b
    auto block_order = [](const BoundaryIndex& B, uint32_t ci) {
        std::vector<uint32_t> jb(B.n_blocks[ci], MQ_NONE);
        uint64_t bs = B.block_size(ci);
        for (uint32_t b = 0; b < B.aux_dim; ++b) {
            int64_t o = B.offset(ci, b);
            if (o < 0) continue;
            MQ_REQUIRE(!B.transposes[size_t(ci) * B.aux_dim + b], "transposed block inside a stored cohort");
            MQ_REQUIRE(B.conjugate_scale(ci, b) == 1.0, "scaled block inside a stored cohort");
            uint64_t j = uint64_t(o) / bs;
            MQ_REQUIRE(j < jb.size() && uint64_t(o) % bs == 0 && jb[j] == MQ_NONE, "cohort block offsets");
            jb[j] = b;
        }
        for (auto b : jb) MQ_REQUIRE(b != MQ_NONE, "cohort with unmapped stored block");
        return jb;
    };
    // ---- pairs and groups
    for (uint32_t lb = 0; lb < left_i.size(); ++lb) {
        Charge lc_in = left_i.charges[lb];
        if (rstep && (lb < step->lo || lb >= step->hi)) continue;
        for (uint32_t rb = 0; rb < right_i.size(); ++rb) {
            Charge rc_out = right_i.charges[rb];
            if (lstep && (rb < step->lo || rb >= step->hi)) continue;
            // keys of the pair: K = (b, s_in, ss) of the boundary multiplied in the T stage, N = (b,
            // s_out, ss) of the boundary the final stage contracts with (sigma: L; steps: the new
            // boundary's stored cohorts).  For the left step the pair is (bra left sector, ket right
            // sector) and T = L psi; otherwise (ket left sector, output right sector) and T = psi R.
            std::vector<uint32_t> K, N;
            auto add_keys = [&](std::vector<uint32_t>& out, const BoundaryIndex& B, uint32_t c, uint32_t s,
                                bool stored_only) {
                if (c >= B.n_cohorts() || (stored_only && B.tr[c])) return;
                for (uint32_t b = 0; b < B.aux_dim; ++b)
                    if (B.has_block(c, b))
                        for (uint32_t ss = 0; ss < phys_i.sizes[s]; ++ss) out.push_back(G.key(b, s, ss));
            };
            for (uint32_t s = 0; s < phys_i.size(); ++s) {
                if (lstep) {
                    // lc_in is the bra's left sector, rc_out the ket's right sector
                    add_keys(K, LI, LI.cohort_index(lc_in, rc_out - phys_i.charges[s]), s, false);
                    add_keys(N, NI, NI.cohort_index(lc_in + phys_i.charges[s], rc_out), s, true);
                    continue;
                }
                add_keys(K, RI, RI.cohort_index(lc_in + phys_i.charges[s], rc_out), s, false);
                if (sigma)
                    add_keys(N, LI, LI.cohort_index(rc_out - phys_i.charges[s], lc_in), s, false);
                else
                    add_keys(N, NI, NI.cohort_index(lc_in, rc_out - phys_i.charges[s]), s, true);
            }
            if (K.empty() || N.empty()) continue;
            Charge d = rc_out - lc_in;
            auto ins = group_of.emplace(std::make_pair(d.key(), sy.spin(lc_in)), uint32_t(G.groups.size()));
            if (ins.second) {
                G.groups.emplace_back();
                G.groups.back().delta = d;
                G.groups.back().spin = sy.spin(lc_in);
                G.groups.back().kpos.assign(G.nkeys, -1);
                G.groups.back().npos.assign(G.nkeys, -1);
            }
            GroupedLowering::Group& g = G.groups[ins.first->second];
            GroupedLowering::Pair pr;
            pr.lb_in = lb;
            pr.rb_out = rb;
            pr.g = ins.first->second;
            pr.m_off = g.M;
            g.M += uint64_t(left_i.sizes[lb]) * right_i.sizes[rb];
            // keys for now; numbered within the group once every pair is known
            for (auto k : K) g.kpos[k] = 0;
            for (auto n : N) g.npos[n] = 0;
            pr.kk = std::move(K);
            pr.nn = std::move(N);
            g.pairs.push_back(uint32_t(G.pairs.size()));
            G.pair_of[uint64_t(lb) << 32 | rb] = uint32_t(G.pairs.size());
            G.pairs.push_back(std::move(pr));
        }
    }
    trace_phase("lower: pairs and keys");
    // ---- number each group's keys in ascending (b, s, ss) order, so that the rows and columns of
    // W_D follow the MPO's block structure
    for (auto& g : G.groups) {
        auto number = [&](std::vector<int32_t>& pos, std::vector<uint32_t>& keys) {
            keys.clear();
            for (uint32_t k = 0; k < G.nkeys; ++k)
                if (pos[k] >= 0) {
                    pos[k] = int32_t(keys.size());
                    keys.push_back(k);
                }
        };
        number(g.kpos, g.K);
        number(g.npos, g.N);
        for (uint32_t pi : g.pairs) {
            for (auto& k : G.pairs[pi].kk) k = uint32_t(g.kpos[k]);
            for (auto& n : G.pairs[pi].nn) n = uint32_t(g.npos[n]);
        }
    }
    // ---- workspace: T_D, S_D, W_D per group
    uint64_t ws = 0;
    for (auto& g : G.groups) {
        MQ_REQUIRE(g.M < (1ull << 32) && g.K.size() < (1u << 31) && g.N.size() < (1u << 31), "group too large");
        g.t_off = ws;
        ws += align32(g.M * g.K.size());
    }
    // T_D columns that no pair writes are cleared per application by zero chunks (below), not by
    // a dense fill
    p->st.s_begin = 0;
    p->st.s_elems = 0;
    p->st.t_end = ws;
    // bf16 buffer: per group A' = [T_hi | T_lo | T_hi] (M x 3K), written by the T stage of BF16x3
    // applications; the B' regions follow (S stage)
    for (auto& g : G.groups) {
        g.bf_off = p->st.bf_elems;
        p->st.bf_elems += align32(3ull * g.M * g.K.size());
    }
    p->st.bf_a_elems = p->st.bf_elems;
    p->s_region0 = ws;
    for (auto& g : G.groups) {
        g.s_off = ws;
        ws += align32(g.M * g.N.size());
    }
    p->s_region1 = ws;
    ws_init_off = ws;
    for (auto& g : G.groups) {
        g.w_off = ws;
        ws += align32(uint64_t(g.K.size()) * g.N.size());
    }
    wd_elems = ws - ws_init_off;
    // W_D is assembled as its nonzeros per group (wg) for every plan but the adjoint's; the dense
    // matrices (the pinned staging, then the workspace) serve only the adjoint and the library S
    // stage (MQ_S_DENSE=1), while the block-sparse S stage builds its tiles from the nonzeros
    const bool sparse_w = !adjoint;
    const bool bsr = !adjoint && s_sparse_enabled();
    double* const wd = bsr ? nullptr : p->ctx->wd_stage(wd_elems);
    std::vector<WList> wg(sparse_w ? G.groups.size() : 0);
    trace_phase("lower: key order, layout");
    if (g_trace) {
        uint64_t tk = 0, tn = 0, tm = 0;
        for (const auto& g : G.groups) tk += g.K.size(), tn += g.N.size(), tm += g.M;
        std::fprintf(stderr, "[mq] groups %zu pairs %zu keys K %llu N %llu rows %llu, W_D %.1f MB, workspace %.1f GB\n",
                     G.groups.size(), G.pairs.size(), (unsigned long long)tk, (unsigned long long)tn,
                     (unsigned long long)tm, 8e-6 * double(ws - ws_init_off), 8e-9 * double(ws));
    }
    struct AdjEntry { uint32_t t; uint64_t w; double c; };
    std::vector<AdjEntry> adj;
    // one nonzero of W_D of group g at (column n, row k), from one MPO row
    struct RawW { uint32_t g, n, k; double v; };
    std::vector<std::vector<RawW>> raw(sparse_w ? W.row_dim : 0);
    // ---- W_D: for every MPO entry, its (k, n) in each group holding both keys.  The groups of
    // a K key are listed per key (CSR over the dense key numbering)
    std::vector<uint32_t> kg_off(size_t(G.nkeys) + 1, 0);
    for (const auto& g : G.groups)
        for (uint32_t k : g.K) ++kg_off[k + 1];
    for (uint32_t k = 0; k < G.nkeys; ++k) kg_off[k + 1] += kg_off[k];
    std::vector<std::pair<uint32_t, uint32_t>> kg_ent(kg_off[G.nkeys]);
    {
        std::vector<uint32_t> fill(kg_off.begin(), kg_off.end() - 1);
        for (uint32_t gi = 0; gi < G.groups.size(); ++gi)
            for (uint32_t i = 0; i < G.groups[gi].K.size(); ++i) kg_ent[fill[G.groups[gi].K[i]]++] = {gi, i};
    }
    // Rows b1 of the MPO write disjoint W_D entries (b1 is part of every N key of sigma and the
    // right step, and of every K key of the left step), so they are assembled in parallel; each
    // entry still sums its contributions in MPO order.  The adjoint's entry list keeps MPO order.
    std::atomic<int> bad{0};
    // SU2 coupling of an MPO entry inside group g (site_hamil_tasks.hpp:93-125, left_tasks.hpp /
    // right_tasks.hpp): spins J (ket left), J' (ket right), I (bra left), I' (bra right) of the
    // sectors it connects.  sigma and the right step pair a ket left sector with a bra right
    // sector; the left step pairs a bra left sector with a ket right sector.
    auto su2_unit = [&](CouplingCache& cc, const GroupedLowering::Group& g, uint32_t s_in, uint32_t s_out,
                        int A, int K, int Ap) -> const double* {
        const int din = sy.spin(phys_i.charges[s_in]), dout = sy.spin(phys_i.charges[s_out]);
        int J, Jp, I, Ip;
        if (lstep) {
            I = g.spin;
            Jp = g.spin + sy.spin(g.delta);
            J = Jp - din;
            Ip = I + dout;
        } else {
            J = g.spin;
            Ip = g.spin + sy.spin(g.delta);
            Jp = J + din;
            I = Ip - dout;
        }
        if (J < 0 || Jp < 0 || I < 0 || Ip < 0) return nullptr;
        return cc.get(J, std::abs(J - Jp), Jp, A, K, Ap, I, std::abs(I - Ip), Ip);
    };
    // the coupling of entry e of op (unit couplings u, null: none): its (row spin, col spin) case
    auto su2_case = [](const double* u, uint32_t e, const Op& op) {
        if (!u) return 0.0;
        const unsigned rs = op.row_spin[e], cs = op.col_spin[e];
        const int casenr = (rs == 2 && cs == 2) ? 3 : rs == 2 ? 1 : cs == 2 ? 2 : 0;   // op_iterate.hpp:56-60
        return u[casenr];
    };
    // raw_out: the row's W_D nonzeros.  All contributions to one entry come from one MPO cell
    // (b1, b2) (its K and N keys carry b1 and b2), which touches at most nst x nst key pairs: they
    // are summed per cell in a small accumulator, in the order the cell makes them (as the dense
    // accumulation adds them), and the cell's nonzero sums are emitted in (k, n) order.
    auto mpo_row = [&](uint32_t b1, std::vector<AdjEntry>* adj_out, std::vector<RawW>* raw_out) {
        thread_local CouplingCache ccache;
        thread_local std::vector<double> cacc;   // per (k key, n key, group of the k key) of a cell
        thread_local std::vector<char> cset;
        // SU2: the unit couplings of a group depend on the operator block, not on its entries: kept
        // per group for the block being expanded (ustamp: the block it was computed for)
        thread_local std::vector<const double*> ucache;
        thread_local std::vector<uint64_t> ustamp;
        thread_local uint64_t stamp = 0;
        if (ustamp.size() < G.groups.size()) {
            ucache.assign(G.groups.size(), nullptr);
            ustamp.assign(G.groups.size(), 0);
        }
        const uint32_t nst = G.nst;
        for (const auto& cell : W.rows[b1]) {
            uint32_t b2 = cell.other;
            const uint32_t bk = lstep ? b1 : b2, bn = lstep ? b2 : b1;
            uint32_t qmax = 0;
            if (raw_out) {
                for (uint32_t kl = 0; kl < nst; ++kl)
                    qmax = std::max(qmax, kg_off[bk * nst + kl + 1] - kg_off[bk * nst + kl]);
                cacc.assign(size_t(nst) * nst * qmax, 0.0);
                cset.assign(size_t(nst) * nst * qmax, 0);
            }
            for (uint32_t t = cell.t0; t < cell.t1; ++t) {
                const Op& op = W.ops[W.terms[t].op];
                double sc = W.terms[t].scale;
                for (const OpBlock& wb : op.blocks) {
                    ++stamp;
                    uint32_t s_in = phys_i.position(wb.in), s_out = phys_i.position(wb.out);
                    if (s_in >= phys_i.size() || s_out >= phys_i.size()) {
                        bad = 1;
                        return;
                    }
                    for (uint32_t e = wb.e0; e < wb.e1; ++e) {
                        if (op.row[e] >= phys_i.sizes[s_in] || op.col[e] >= phys_i.sizes[s_out]) {
                            bad = 2;
                            return;
                        }
                        // K keys carry the MPO index of the T-stage boundary (b2 for psi R, b1 for L psi)
                        const uint32_t kd = G.key(lstep ? b1 : b2, s_in, op.row[e]);
                        const uint32_t nd = G.key(lstep ? b2 : b1, s_out, op.col[e]);
                        for (uint32_t q = kg_off[kd]; q < kg_off[kd + 1]; ++q) {
                            const uint32_t gi = kg_ent[q].first;
                            const auto& g = G.groups[gi];
                            const int32_t ni = g.npos[nd];
                            if (ni < 0) continue;
                            double c = op.coeff[e];
                            if (sy.su2()) {
                                if (ustamp[gi] != stamp) {
                                    ucache[gi] = su2_unit(ccache, g, s_in, s_out, W.left_spin[b1], op.spin, W.right_spin[b2]);
                                    ustamp[gi] = stamp;
                                }
                                c = op.coeff[e] * su2_case(ucache[gi], e, op);
                            }
                            if (raw_out) {
                                const size_t slot = (size_t(kd - bk * nst) * nst + (nd - bn * nst)) * qmax + (q - kg_off[kd]);
                                cacc[slot] += c * sc;
                                cset[slot] = 1;
                                continue;
                            }
                            const uint64_t w = g.w_off - ws_init_off + uint64_t(ni) * g.K.size() + kg_ent[q].second;
                            wd[w] += c * sc;
                            if (adj_out) adj_out->push_back({t, w, c});
                        }
                    }
                }
            }
            if (raw_out)
                for (uint32_t kl = 0; kl < nst; ++kl) {
                    const uint32_t kd = bk * nst + kl;
                    for (uint32_t nl = 0; nl < nst; ++nl) {
                        const uint32_t nd = bn * nst + nl;
                        for (uint32_t q = kg_off[kd]; q < kg_off[kd + 1]; ++q) {
                            const size_t slot = (size_t(kl) * nst + nl) * qmax + (q - kg_off[kd]);
                            if (!cset[slot] || cacc[slot] == 0.0) continue;
                            raw_out->push_back({kg_ent[q].first, uint32_t(G.groups[kg_ent[q].first].npos[nd]),
                                                kg_ent[q].second, cacc[slot]});
                        }
                    }
                }
        }
    };
    if (adjoint) {
        for (uint32_t b1 = 0; b1 < W.row_dim; ++b1) mpo_row(b1, &adj, nullptr);
    } else {
#pragma omp parallel for schedule(dynamic, 4)
        for (int b1 = 0; b1 < int(W.row_dim); ++b1) mpo_row(uint32_t(b1), nullptr, &raw[b1]);
    }
    MQ_REQUIRE(bad != 1, "operator charge not in physical index");
    MQ_REQUIRE(bad != 2, "operator element outside its physical sector");
    trace_phase("lower: W_D rows");
    if (sparse_w) {
        // A group's entries are the rows' entries in row order.  Key positions grow with b1 and
        // b2 (keys are b-major; a row's cells ascend in b2; a cell emits in (k, n) order), so for
        // every k its columns n, and for every n its rows k, come in ascending order, the order
        // reorder_group_keys hashes them in.
        const size_t ng = G.groups.size(), nb = raw.size();
        std::vector<uint64_t> off(nb * ng, 0);   // per (b1, g): entries, then first slot
#pragma omp parallel for schedule(dynamic, 16)
        for (int b1 = 0; b1 < int(nb); ++b1)
            for (const RawW& e : raw[b1]) ++off[size_t(b1) * ng + e.g];
        for (size_t g = 0; g < ng; ++g) {
            uint64_t n = 0;
            for (size_t b1 = 0; b1 < nb; ++b1) {
                const uint64_t c = off[b1 * ng + g];
                off[b1 * ng + g] = n;
                n += c;
            }
            wg[g].resize(n);
        }
#pragma omp parallel for schedule(dynamic, 16)
        for (int b1 = 0; b1 < int(nb); ++b1) {
            for (const RawW& e : raw[b1]) wg[e.g][off[size_t(b1) * ng + e.g]++] = {e.n, e.k, e.v};
            std::vector<RawW>().swap(raw[b1]);
        }
    }
    trace_phase("lower: W_D");
    // ---- key order for the block-sparse S stage: within each group, keys whose rows (K) or
    // columns (N) of W_D have the same nonzero pattern are made adjacent, larger classes first, so
    // that W_D's nonzeros form dense rectangles and few of its 16-row tiles are empty and visited.
    // Only positions change: every later consumer reads them through kpos / npos and the pairs'
    // kk / nn lists.  Adjoint plans keep the ascending order their term map was built in.
    if (!adjoint) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int gi = 0; gi < int(G.groups.size()); ++gi) {
            auto& g = G.groups[gi];
            reorder_group_keys(G, g, wg[gi]);
            if (!bsr) {   // the library S stage: the dense W_D in the new order
                double* Wg = wd + (g.w_off - ws_init_off);
                const uint64_t Kg = g.K.size();
                for (const WEntry& e : wg[gi]) Wg[uint64_t(e.n) * Kg + e.k] = e.v;
            }
        }
        trace_phase("lower: key classes");
    }
    StagePlan& P = p->st;
    if (!lstep) {
        // ---- T stage.  The psi blocks that meet one right cohort (every (lb_in, s, ss) with
        // lc_in + q_s = rc_in, for the cohort's rc_out) are stacked into the rows of one task, so the
        // cohort is read once per column tile for all of them; each block's rows are written to its own
        // pair slab and keys through a row unit.  A stored cohort is one task over all its blocks
        // (columns (j, rs) of block j); a transposed cohort is one task per block: the partner's block,
        // transposed and scaled by its conjugation factor.
        struct RowBlock { uint32_t pair, s, ss; };
        struct Stack { uint32_t K = 0, brs = 0, rbi = 0; std::vector<RowBlock> blocks; };
        std::map<uint32_t, Stack> stacks;   // right cohort -> its psi blocks (deterministic order)
        for (uint32_t pi = 0; pi < G.pairs.size(); ++pi) {
            const auto& pr = G.pairs[pi];
            uint32_t lb = pr.lb_in, rb = pr.rb_out;
            Charge lc_in = left_i.charges[lb], rc_out = right_i.charges[rb];
            for (uint32_t s = 0; s < phys_i.size(); ++s) {
                Charge rc_in = lc_in + phys_i.charges[s];
                uint32_t rbi = right_i.position(rc_in);
                uint32_t ci = RI.cohort_index(rc_in, rc_out);
                if (rbi == right_i.size() || ci >= RI.n_cohorts()) continue;
                Stack& st = stacks[ci];
                st.K = right_i.sizes[rbi];
                st.brs = right_i.sizes[rb];
                st.rbi = rbi;
                for (uint32_t ss = 0; ss < phys_i.sizes[s]; ++ss) st.blocks.push_back({pi, s, ss});
            }
        }
        // one stacked task over the given psi blocks: row units with the column keys of `b2s`
        auto stacked_task = [&](const Stack& st, const std::vector<uint32_t>& b2s, const GemmSeg& sg) {
            std::vector<uint32_t> unit_of_blk;
            for (const RowBlock& rbk : st.blocks) {
                const auto& pr = G.pairs[rbk.pair];
                const auto& g = G.groups[pr.g];
                const uint32_t lb = pr.lb_in, M = left_i.sizes[lb];
                if (!M) continue;
                MQ_REQUIRE(g.M < (1ull << 32), "group too large for a row unit");
                GemmRow R{};
                R.a_src = site.rp_offset[lb] + (rpb(rbk.s, st.rbi) + uint64_t(rbk.ss) * st.K) * M;
                R.c_si = 1;
                R.a_si = 1;
                R.a_sk = M;
                R.a_scale = 1.0;
                R.row0 = uint32_t(unit_of_blk.size()) * 16;
                R.nrows = M;
                R.c_src = g.t_off + pr.m_off;
                R.c_split = g.bf_off + pr.m_off;
                R.split_stride = g.M * g.K.size();
                R.cldk = uint32_t(g.M);
                R.ldc = M;
                R.cmap = uint32_t(P.maps.size());
                for (uint32_t b2 : b2s) P.maps.push_back(g.kid(G.key(b2, rbk.s, rbk.ss)));
                unit_of_blk.insert(unit_of_blk.end(), (M + 15) / 16, uint32_t(P.rows.size()));
                P.rows.push_back(R);
            }
            if (unit_of_blk.empty() || !st.brs) return;
            const uint32_t Mt = uint32_t(unit_of_blk.size()) * 16;
            GemmTask& t = add_gemm(P.t_tasks, P.t_segs, BASE_WS, 0, Mt, uint32_t(b2s.size()) * st.brs, Mt, {sg});
            t.cbw = st.brs;
            t.rowunit = uint32_t(P.maps.size());
            P.maps.insert(P.maps.end(), unit_of_blk.begin(), unit_of_blk.end());
        };
        for (const auto& kv : stacks) {
            const uint32_t ci = kv.first;
            const Stack& st = kv.second;
            if (st.blocks.empty()) continue;
            if (!RI.tr[ci]) {
                std::vector<uint32_t> jb = block_order(RI, ci);
                if (jb.empty()) continue;
                stacked_task(st, jb, seg(BASE_IN, 0, 1, false, BASE_RIGHT, p->right->cohort_off[ci], st.K, false, st.K));
            } else {
                const uint32_t cie = RI.cohort_index(RI.ket.charges[RI.ket_pos[ci]], RI.bra.charges[RI.bra_pos[ci]]);
                MQ_REQUIRE(cie < RI.n_cohorts(), "transposed right cohort without partner");
                for (uint32_t b2 = 0; b2 < RI.aux_dim; ++b2) {
                    if (!RI.has_block(ci, b2)) continue;
                    const uint64_t o = uint64_t(RI.offset(ci, b2));
                    stacked_task(st, {b2}, seg(BASE_IN, 0, 1, false, BASE_RIGHT, p->right->cohort_off[cie] + o, st.brs, true,
                                               st.K, RI.conjugate_scale(ci, b2)));
                }
            }
        }
    } else {
        // ---- T stage of the left step: X(pair (lb_bra, rb_ket), key (b1, s, ss)) = L_b1(lb_bra, lb_ket)
        // ket(lb_ket)[:, columns (s, ss, rb_ket)], stacked per left cohort and (s, ss): its blocks b1
        // are row units sharing the ket columns
        for (uint32_t lb = 0; lb < left_i.size(); ++lb)
            for (uint32_t lbk = 0; lbk < left_i.size(); ++lbk) {
                const uint32_t cl = LI.cohort_index(left_i.charges[lb], left_i.charges[lbk]);
                const uint32_t lsb = left_i.sizes[lb], lsk = left_i.sizes[lbk];
                if (cl >= LI.n_cohorts() || !lsb || !lsk) continue;
                // the cohort's blocks: (b1, A address, strides, scale)
                struct LBlk { uint32_t b1; uint64_t a; uint32_t si, sk; double sc; };
                std::vector<LBlk> blks;
                if (!LI.tr[cl]) {
                    std::vector<uint32_t> jb = block_order(LI, cl);
                    for (uint32_t j = 0; j < jb.size(); ++j)
                        blks.push_back({jb[j], p->left->cohort_off[cl] + uint64_t(j) * lsb * lsk, 1, lsb, 1.0});
                } else {
                    const uint32_t cle = LI.cohort_index(left_i.charges[lbk], left_i.charges[lb]);
                    MQ_REQUIRE(cle < LI.n_cohorts(), "transposed left cohort without partner");
                    for (uint32_t b1 = 0; b1 < LI.aux_dim; ++b1)
                        if (LI.has_block(cl, b1))
                            blks.push_back({b1, p->left->cohort_off[cle] + uint64_t(LI.offset(cl, b1)), lsk, 1,
                                            LI.conjugate_scale(cl, b1)});
                }
                if (blks.empty()) continue;
                for (uint32_t s = 0; s < phys_i.size(); ++s) {
                    const uint32_t rb = right_i.position(left_i.charges[lbk] + phys_i.charges[s]);
                    if (rb == right_i.size() || !right_i.sizes[rb]) continue;
                    auto pit = G.pair_of.find(uint64_t(lb) << 32 | rb);
                    if (pit == G.pair_of.end()) continue;
                    const auto& pr = G.pairs[pit->second];
                    const auto& g = G.groups[pr.g];
                    const uint32_t rs = right_i.sizes[rb];
                    for (uint32_t ss = 0; ss < phys_i.sizes[s]; ++ss) {
                        std::vector<uint32_t> unit_of_blk;
                        for (const LBlk& bk : blks) {
                            GemmRow R{};
                            R.a_src = bk.a;
                            R.a_si = bk.si;
                            R.a_sk = bk.sk;
                            R.a_scale = bk.sc;
                            R.row0 = uint32_t(unit_of_blk.size()) * 16;
                            R.nrows = lsb;
                            R.c_src = g.t_off + pr.m_off + uint64_t(g.kid(G.key(bk.b1, s, ss))) * g.M;
                            R.ldc = lsb;
                            R.c_si = 1;
                            R.cmap = MAP_NONE;
                            unit_of_blk.insert(unit_of_blk.end(), (lsb + 15) / 16, uint32_t(P.rows.size()));
                            P.rows.push_back(R);
                        }
                        const uint32_t Mt = uint32_t(unit_of_blk.size()) * 16;
                        const uint64_t col0 = rpb(s, rb) + uint64_t(ss) * rs;
                        GemmTask& t = add_gemm(P.t_tasks, P.t_segs, BASE_WS, 0, Mt, rs, Mt,
                                               {seg(BASE_LEFT, 0, 1, false, BASE_IN, site.rp_offset[lbk] + col0 * lsk,
                                                    lsk, false, lsk)});
                        t.cbw = rs;
                        t.rowunit = uint32_t(P.maps.size());
                        P.maps.insert(P.maps.end(), unit_of_blk.begin(), unit_of_blk.end());
                    }
                }
            }
    }
    trace_phase("lower: T stage");
    if (g_trace)
        std::fprintf(stderr, "[mq] T tasks %zu segs %zu rows %zu maps %zu\n", P.t_tasks.size(), P.t_segs.size(),
                     P.rows.size(), P.maps.size());
    P.split_t = true;   // every T task is stacked: BF16x3 writes A' from the T epilogue
    // ---- zero chunks: the (pair, key) slabs of T_D that no T task writes, merged along the rows
    // of each key column (consecutive pairs of a group are consecutive rows)
    auto add_zero = [&](std::vector<ZeroChunk>& list, uint64_t off, uint64_t len) {
        for (uint64_t o = 0; o < len; o += ZERO_CHUNK)
            list.push_back({off + o, uint32_t(std::min<uint64_t>(ZERO_CHUNK, len - o)), 0});
    };
    for (const auto& g : G.groups) {
        const uint64_t Kg = g.K.size();
        std::vector<std::vector<char>> has(g.pairs.size(), std::vector<char>(Kg, 0));
        for (size_t q = 0; q < g.pairs.size(); ++q)
            for (auto k : G.pairs[g.pairs[q]].kk) has[q][k] = 1;
        for (uint64_t k = 0; k < Kg; ++k) {
            uint64_t run0 = 0, run = 0;
            auto flush = [&]() {
                if (!run) return;
                add_zero(P.zero_ws, g.t_off + k * g.M + run0, run);
                for (uint64_t c = 0; c < 3; ++c) add_zero(P.zero_bf, g.bf_off + (k + c * Kg) * g.M + run0, run);
                run = 0;
            };
            for (size_t q = 0; q < g.pairs.size(); ++q) {
                const auto& pr = G.pairs[g.pairs[q]];
                const uint64_t slab = uint64_t(left_i.sizes[pr.lb_in]) * right_i.sizes[pr.rb_out];
                if (has[q][k] || !slab) {
                    if (slab) flush();
                    continue;
                }
                if (!run) run0 = pr.m_off;
                MQ_REQUIRE(run0 + run == pr.m_off, "pair slabs of a group are not consecutive");
                run += slab;
            }
            flush();
        }
    }
    trace_phase("lower: zero chunks");
    // ---- S stage: S_D = T_D W_D, one plain dense GEMM per group (rocBLAS; the MPO coefficient
    // matrices are 10-30% dense, and the library GEMM runs at near the MFMA peak)
    for (auto& g : G.groups) {
        if (!g.M || g.K.empty() || g.N.empty()) continue;
        MQ_REQUIRE(g.M < (1ull << 31), "group too large for one library GEMM");
        const uint64_t Kg = g.K.size(), Ng = g.N.size();
        MQ_REQUIRE(3 * Kg < (1ull << 31), "group too large for one library GEMM");
        P.lib.push_back({g.t_off, g.w_off, g.s_off, uint32_t(g.M), uint32_t(Ng), uint32_t(Kg), uint32_t(g.M),
                         uint32_t(Kg), uint32_t(g.M), g.bf_off, 0});
        // the block-sparse kernel runs every group (below)
        P.lib.back().sparse = bsr;
    }
    if (bsr) {   // the block-sparse kernel over the visited tiles of W_D
        std::vector<std::pair<uint64_t, double>> info;
        uint64_t cap = 1;
        for (const auto& w : wg) cap += w.size();
        double* stage = p->ctx->wd_stage(2 * cap, false);   // positions, then values
        P.bsr.ent_pos = reinterpret_cast<uint64_t*>(stage);
        P.bsr.ent_val = stage + cap;
        add_bsr_groups(P.bsr, G.groups, wg, g_trace ? &info : nullptr);
        std::vector<WList>().swap(wg);
        P.bsr_on = true;
        for (size_t gi = 0; gi < info.size(); ++gi) {
            const auto& g = G.groups[gi];
            const double dense = double(g.K.size()) * double(g.N.size());
            std::fprintf(stderr, "[mq] S group M %llu K %zu N %zu: W_D density %.3f, visited tiles %.3f\n",
                         (unsigned long long)g.M, g.K.size(), g.N.size(), dense ? double(info[gi].first) / dense : 0.0,
                         info[gi].second);
        }
        trace_phase("lower: S tiles");
    }
    for (auto& L : P.lib) {   // B' regions after the A' regions
        L.bfb_off = P.bf_elems;
        P.bf_elems += align32(3ull * L.K * L.N);
    }
    if (sigma) {
        // ---- final stage: one task per output block lb_out (ls_out x all its columns), K segments over
        // the left cohorts (lc_out, lc_in).  The block's columns fall into units (s, ss2) of rs columns;
        // unit u of segment lb_in reads its B from S_D of the pair (lb_in, rb(s)) through the unit table,
        // so each left cohort is read once per column tile (not once per physical state).
        for (uint32_t lbo = 0; lbo < left_i.size(); ++lbo) {
            Charge lc_out = left_i.charges[lbo];
            uint32_t ls_out = left_i.sizes[lbo];
            const uint32_t Nb = site.rp_cols[lbo];
            if (!ls_out || !Nb) continue;
            struct Unit { uint32_t s, ss, rb, col0, rs; };
            std::vector<Unit> us;
            for (uint32_t s = 0; s < phys_i.size(); ++s) {
                Charge rc_out = lc_out + phys_i.charges[s];
                uint32_t rb = right_i.position(rc_out);
                if (rb == right_i.size()) continue;
                uint32_t rs = right_i.sizes[rb];
                for (uint32_t ss = 0; ss < phys_i.sizes[s]; ++ss)
                    us.push_back({s, ss, rb, uint32_t(rpb(s, rb)) + ss * rs, rs});
            }
            const uint32_t colunit = uint32_t(P.maps.size());
            P.maps.resize(P.maps.size() + Nb, MQ_NONE);
            for (uint32_t u = 0; u < us.size(); ++u)
                for (uint32_t r = 0; r < us[u].rs; ++r) {
                    MQ_REQUIRE(us[u].col0 + r < Nb && P.maps[colunit + us[u].col0 + r] == MQ_NONE, "output column units");
                    P.maps[colunit + us[u].col0 + r] = u;
                }
            for (uint32_t n = 0; n < Nb; ++n) MQ_REQUIRE(P.maps[colunit + n] != MQ_NONE, "output column without unit");
            // unit descriptors of one segment: the S_D block of the pair (lb_in, rb(s)) and the S_D column
            // of every b1 block of the segment's K range
            auto add_units = [&](uint32_t lb_in, const std::vector<uint32_t>& b1s) {
                const uint32_t first = uint32_t(P.units.size());
                for (const Unit& U : us) {
                    GemmUnit d{};
                    d.col0 = U.col0;
                    auto pit = G.pair_of.find(uint64_t(lb_in) << 32 | U.rb);
                    if (pit != G.pair_of.end()) {
                        const auto& pr = G.pairs[pit->second];
                        const auto& g = G.groups[pr.g];
                        d.src = g.s_off + pr.m_off;
                        d.ld = uint32_t(g.M);
                        d.map = uint32_t(P.maps.size());
                        d.valid = 1;
                        for (uint32_t b1 : b1s) P.maps.push_back(g.nid(G.key(b1, U.s, U.ss)));
                    }
                    P.units.push_back(d);
                }
                return first;
            };
            // segments: the left cohorts (lc_out, lc_in), each cut into pieces of whole blocks of about
            // KPART columns of op(A), so that split-K parts below can be balanced however the cohort
            // sizes are distributed
            const uint64_t KPART = 2048;
            std::vector<GemmSeg> segs;
            for (uint32_t lb = 0; lb < left_i.size(); ++lb) {
                Charge lc_in = left_i.charges[lb];
                uint32_t cl = LI.cohort_index(lc_out, lc_in);
                if (cl >= LI.n_cohorts()) continue;
                uint32_t ls_in = left_i.sizes[lb];
                if (!ls_in) continue;
                const uint32_t per = uint32_t(std::max<uint64_t>(1, KPART / ls_in));   // blocks per piece
                if (!LI.tr[cl]) {
                    std::vector<uint32_t> jb = block_order(LI, cl);
                    for (uint32_t j0 = 0; j0 < jb.size(); j0 += per) {
                        const uint32_t j1 = std::min<uint32_t>(uint32_t(jb.size()), j0 + per);
                        GemmSeg sg = seg(BASE_LEFT, p->left->cohort_off[cl] + uint64_t(j0) * ls_in * ls_out, ls_out, false,
                                         BASE_WS, 0, ls_in, false, (j1 - j0) * ls_in);
                        sg.bbh = ls_in;
                        sg.units = add_units(lb, std::vector<uint32_t>(jb.begin() + j0, jb.begin() + j1));
                        segs.push_back(sg);
                    }
                } else {
                    // Hermitian partner: the blocks are the partner cohort's, transposed, each with its
                    // conjugation scale (A block map + block scales)
                    uint32_t cle = LI.cohort_index(lc_in, lc_out);
                    MQ_REQUIRE(cle < LI.n_cohorts(), "transposed left cohort without partner");
                    std::vector<uint32_t> b1s;
                    const uint32_t amap = uint32_t(P.maps.size()), ascale = uint32_t(P.scales.size());
                    for (uint32_t b1 = 0; b1 < LI.aux_dim; ++b1) {
                        if (!LI.has_block(cl, b1)) continue;
                        const int64_t o = LI.offset(cl, b1);
                        MQ_REQUIRE(o >= 0 && o < (int64_t(1) << 32), "partner block offset");
                        b1s.push_back(b1);
                        P.maps.push_back(uint32_t(o));
                        P.scales.push_back(LI.conjugate_scale(cl, b1));
                    }
                    for (uint32_t j0 = 0; j0 < b1s.size(); j0 += per) {
                        const uint32_t j1 = std::min<uint32_t>(uint32_t(b1s.size()), j0 + per);
                        GemmSeg sg = seg(BASE_LEFT, p->left->cohort_off[cle], ls_in, true, BASE_WS, 0, ls_in, false,
                                         (j1 - j0) * ls_in);
                        sg.bbh = ls_in;
                        sg.amap = amap + j0;
                        sg.ascale = ascale + j0;
                        sg.units = add_units(lb, std::vector<uint32_t>(b1s.begin() + j0, b1s.begin() + j1));
                        segs.push_back(sg);
                    }
                }
            }
            if (segs.empty()) continue;
            // split K over several tasks when it is long, so that the few output blocks fill the GPU:
            // each part writes its own FP64 partial block, and one gather region sums the parts into
            // sigma in a fixed order (deterministic, no atomics)
            uint64_t ktot = 0;
            for (const auto& sg : segs) ktot += sg.K;
            const uint32_t nparts = uint32_t(std::min<uint64_t>({uint64_t(segs.size()), 64, (ktot + KPART - 1) / KPART}));
            auto emit = [&](uint8_t c_base, uint64_t c_off, size_t a0, size_t a1) {
                std::vector<GemmSeg> part(segs.begin() + a0, segs.begin() + a1);
                GemmTask& t = add_gemm(P.f_tasks, P.f_segs, c_base, c_off, ls_out, Nb, ls_out, part);
                t.colunit = colunit;
                t.nunits = uint32_t(us.size());
            };
            if (nparts <= 1) {
                emit(BASE_OUT, site.rp_offset[lbo], 0, segs.size());
                continue;
            }
            GatherRegion r{};
            r.dst_off = site.rp_offset[lbo];
            r.R = ls_out;
            r.C = Nb;
            r.ld_dst = ls_out;
            r.t0 = uint32_t(P.red.src_off.size());
            size_t a0 = 0;
            uint64_t acc = 0;
            for (uint32_t q = 0; q < nparts && a0 < segs.size(); ++q) {
                // parts of about ktot / nparts each; the last takes the rest
                const uint64_t target = (ktot * (q + 1)) / nparts;
                size_t a1 = a0;
                while (a1 < segs.size() && (acc < target || a1 == a0)) acc += segs[a1++].K;
                if (q + 1 == nparts) {
                    while (a1 < segs.size()) acc += segs[a1++].K;
                }
                const uint64_t off = P.part_elems;
                P.part_elems += align32(uint64_t(ls_out) * Nb);
                emit(BASE_BRA, off, a0, a1);
                P.red.src_off.push_back(off);
                P.red.alpha.push_back(1.0);
                a0 = a1;
            }
            r.t1 = uint32_t(P.red.src_off.size());
            P.red.regions.push_back(r);
        }
    } else {
        // ---- final stage of a step: one stacked task per stored cohort of the new boundary, its
        // blocks as row units, one K segment per (bra sector, s, ss) unit of columns; each segment
        // reads the S_D slab of every block's key through its A block map.
        //   right step: R'_b1(ket, bra) = sum S_b1(ket; s, ss, r) bra(bra; s, ss, r)
        //   left step:  L'_b2(bra, ket)^T = sum S_b2(l; s, ss, ket)^T bra(l; s, ss, bra), written
        //               through the units' row stride
        for (uint32_t cn = 0; cn < NI.n_cohorts(); ++cn) {
            if (NI.tr[cn]) continue;
            const uint32_t a_sec = NI.bra_pos[cn], b_sec = NI.ket_pos[cn];
            const uint32_t ket_sec = rstep ? a_sec : b_sec;
            if (ket_sec < step->lo || ket_sec >= step->hi) continue;
            // rows of each block (ket sector size for both steps), and the output's other dimension
            const uint32_t rows_u = rstep ? left_i.sizes[a_sec] : right_i.sizes[b_sec];
            const uint32_t ncol = rstep ? left_i.sizes[b_sec] : right_i.sizes[a_sec];
            if (!rows_u || !ncol) continue;
            const uint64_t bs = NI.block_size(cn);
            std::vector<uint32_t> bidx;   // MPO index of each stored block, in storage order
            std::vector<uint64_t> boff;
            for (uint32_t b = 0; b < NI.aux_dim; ++b)
                if (NI.has_block(cn, b)) bidx.push_back(b), boff.push_back(uint64_t(NI.offset(cn, b)));
            if (bidx.empty()) continue;
            MQ_REQUIRE(bs == uint64_t(rows_u) * ncol, "new cohort block shape");
            std::vector<uint32_t> unit_of_blk;
            for (uint32_t j = 0; j < bidx.size(); ++j) {
                GemmRow R{};
                R.a_scale = 1.0;
                R.row0 = uint32_t(unit_of_blk.size()) * 16;
                R.nrows = rows_u;
                R.jblk = j;
                R.c_src = (*step->out_off)[cn] + boff[j];
                R.ldc = rstep ? rows_u : 1;      // right: (ket x bra) column-major; left: the transpose
                R.c_si = rstep ? 1 : ncol;
                R.cmap = MAP_NONE;
                unit_of_blk.insert(unit_of_blk.end(), (rows_u + 15) / 16, uint32_t(P.rows.size()));
                P.rows.push_back(R);
            }
            std::vector<GemmSeg> segs;
            for (uint32_t s = 0; s < phys_i.size(); ++s) {
                // the pair whose S_D slab this unit of columns reads, and the bra block
                uint32_t pl, pr_sec, bra_blk;
                uint64_t col0;
                if (rstep) {
                    const uint32_t rb = right_i.position(left_i.charges[b_sec] + phys_i.charges[s]);
                    if (rb == right_i.size()) continue;
                    pl = a_sec, pr_sec = rb, bra_blk = b_sec;
                    col0 = rpb(s, rb);
                } else {
                    const uint32_t lb = left_i.position(right_i.charges[a_sec] - phys_i.charges[s]);
                    if (lb == left_i.size()) continue;
                    pl = lb, pr_sec = b_sec, bra_blk = lb;
                    col0 = rpb(s, a_sec);
                }
                auto pit = G.pair_of.find(uint64_t(pl) << 32 | pr_sec);
                if (pit == G.pair_of.end()) continue;
                const auto& pr = G.pairs[pit->second];
                const auto& g = G.groups[pr.g];
                const uint32_t lsb = left_i.sizes[bra_blk];
                const uint32_t width = rstep ? right_i.sizes[pr_sec] : right_i.sizes[a_sec];   // unit columns
                for (uint32_t ss = 0; ss < phys_i.sizes[s]; ++ss) {
                    const uint32_t amap = uint32_t(P.maps.size());
                    for (uint32_t b : bidx) P.maps.push_back(g.nid(G.key(b, s, ss)));
                    const uint64_t c0 = col0 + uint64_t(ss) * width;
                    GemmSeg sg;
                    if (rstep)   // A = S slab (ket rows x r columns), B = bra columns transposed
                        sg = seg(BASE_WS, g.s_off + pr.m_off, rows_u, false, BASE_BRA, site.rp_offset[bra_blk] + c0 * lsb,
                                 lsb, true, width);
                    else         // A = S slab transposed (ket columns x l rows), B = bra columns
                        sg = seg(BASE_WS, g.s_off + pr.m_off, lsb, true, BASE_BRA, site.rp_offset[bra_blk] + c0 * lsb,
                                 lsb, false, lsb);
                    sg.amap = amap;
                    sg.bldk = g.M;
                    segs.push_back(sg);
                }
            }
            if (segs.empty()) continue;
            const uint32_t Mt = uint32_t(unit_of_blk.size()) * 16;
            GemmTask& t = add_gemm(P.f_tasks, P.f_segs, BASE_OUT, 0, Mt, ncol, Mt, segs);
            t.cbw = ncol;
            t.rowunit = uint32_t(P.maps.size());
            P.maps.insert(P.maps.end(), unit_of_blk.begin(), unit_of_blk.end());
        }
    }
    if (adjoint) {
        // ---- term map: CSR over the MPO terms, entries in their accumulation order
        std::stable_sort(adj.begin(), adj.end(), [](const AdjEntry& a, const AdjEntry& b) { return a.t < b.t; });
        p->adjoint = true;
        p->n_terms = uint32_t(W.terms.size());
        p->adj_ptr.assign(p->n_terms + 1, 0);
        for (const auto& a : adj) ++p->adj_ptr[a.t + 1];
        for (uint32_t t = 0; t < p->n_terms; ++t) p->adj_ptr[t + 1] += p->adj_ptr[t];
        p->adj_idx.resize(adj.size());
        p->adj_coef.resize(adj.size());
        for (size_t i = 0; i < adj.size(); ++i) p->adj_idx[i] = adj[i].w, p->adj_coef[i] = adj[i].c;
        // ---- U stage: for every output block lb_out and unit u = (s, ss2) of its columns, the
        // slabs U(pair(lb_in, rb(s)), key(b1, s, ss2)) = L_b1^T psi(lb_out)[:, unit u] of all left
        // cohorts (lc_out, lc_in) and their blocks b1, stacked as row units (the adjoint of the
        // final stage: psi . sigma = sum_g <U_g, T_g W_g>)
        for (uint32_t lbo = 0; lbo < left_i.size(); ++lbo) {
            Charge lc_out = left_i.charges[lbo];
            const uint32_t ls_out = left_i.sizes[lbo];
            if (!ls_out || !site.rp_cols[lbo]) continue;
            for (uint32_t s = 0; s < phys_i.size(); ++s) {
                Charge rc_out = lc_out + phys_i.charges[s];
                const uint32_t rb = right_i.position(rc_out);
                if (rb == right_i.size()) continue;
                const uint32_t rs = right_i.sizes[rb];
                if (!rs) continue;
                for (uint32_t ss = 0; ss < phys_i.sizes[s]; ++ss) {
                    const uint32_t col0 = uint32_t(rpb(s, rb)) + ss * rs;
                    std::vector<uint32_t> unit_of_blk;
                    for (uint32_t lb = 0; lb < left_i.size(); ++lb) {
                        const uint32_t cl = LI.cohort_index(lc_out, left_i.charges[lb]);
                        const uint32_t ls_in = left_i.sizes[lb];
                        if (cl >= LI.n_cohorts() || !ls_in) continue;
                        auto pit = G.pair_of.find(uint64_t(lb) << 32 | rb);
                        if (pit == G.pair_of.end()) continue;
                        const auto& pr = G.pairs[pit->second];
                        const auto& g = G.groups[pr.g];
                        auto add_row = [&](uint32_t b1, uint64_t a_src, uint32_t a_si, uint32_t a_sk, double scale) {
                            GemmRow R{};
                            R.a_src = a_src;
                            R.a_si = a_si;
                            R.a_sk = a_sk;
                            R.a_scale = scale;
                            R.row0 = uint32_t(unit_of_blk.size()) * 16;
                            R.nrows = ls_in;
                            R.c_src = g.s_off + pr.m_off + uint64_t(g.nid(G.key(b1, s, ss))) * g.M;
                            R.ldc = ls_in;
                            R.c_si = 1;
                            R.cmap = MAP_NONE;
                            unit_of_blk.insert(unit_of_blk.end(), (ls_in + 15) / 16, uint32_t(P.rows.size()));
                            P.rows.push_back(R);
                        };
                        if (!LI.tr[cl]) {
                            std::vector<uint32_t> jb = block_order(LI, cl);
                            for (uint32_t j = 0; j < jb.size(); ++j)
                                add_row(jb[j], p->left->cohort_off[cl] + uint64_t(j) * ls_out * ls_in, ls_out, 1, 1.0);
                        } else {
                            const uint32_t cle = LI.cohort_index(left_i.charges[lb], lc_out);
                            MQ_REQUIRE(cle < LI.n_cohorts(), "transposed left cohort without partner");
                            for (uint32_t b1 = 0; b1 < LI.aux_dim; ++b1) {
                                if (!LI.has_block(cl, b1)) continue;
                                add_row(b1, p->left->cohort_off[cle] + uint64_t(LI.offset(cl, b1)), 1, ls_in,
                                        LI.conjugate_scale(cl, b1));
                            }
                        }
                    }
                    if (unit_of_blk.empty()) continue;
                    const uint32_t Mt = uint32_t(unit_of_blk.size()) * 16;
                    GemmTask& t = add_gemm(P.u_tasks, P.u_segs, BASE_WS, 0, Mt, rs, Mt,
                                           {seg(BASE_LEFT, 0, 1, false, BASE_IN,
                                                site.rp_offset[lbo] + uint64_t(col0) * ls_out, ls_out, false, ls_out)});
                    t.cbw = rs;
                    t.rowunit = uint32_t(P.maps.size());
                    P.maps.insert(P.maps.end(), unit_of_blk.begin(), unit_of_blk.end());
                }
            }
        }
    }
    // the block-sparse S stage reads 128 rows of every key column it visits (the rows of a short
    // panel past M are zeroed after loading): room for those past the last region
    P.ws_elems = ws + (P.bsr_on ? S_BSR_TILE_M : 0);
    if (!sigma) return;
    // T and L reference flops now; the S part (a scan of W_D per pair) when first asked for
    p->flops = 0;
    for (auto& pr : G.pairs) {
        uint32_t lb = pr.lb_in, rb = pr.rb_out;
        for (uint32_t s = 0; s < phys_i.size(); ++s) {
            Charge rc_in = left_i.charges[lb] + phys_i.charges[s];
            uint32_t ci = RI.cohort_index(rc_in, right_i.charges[rb]);
            if (ci >= RI.n_cohorts()) continue;
            uint32_t cie = RI.tr[ci] ? RI.cohort_index(right_i.charges[rb], rc_in) : ci;
            p->flops += 2ull * left_i.sizes[lb] * RI.cohort_size(cie) * phys_i.sizes[s];
        }
    }
    for (uint32_t lbo = 0; lbo < left_i.size(); ++lbo)
        for (uint32_t lb = 0; lb < left_i.size(); ++lb) {
            uint32_t cl = LI.cohort_index(left_i.charges[lbo], left_i.charges[lb]);
            if (cl >= LI.n_cohorts()) continue;
            uint32_t cle = LI.tr[cl] ? LI.cohort_index(left_i.charges[lb], left_i.charges[lbo]) : cl;
            p->flops += 2ull * left_i.sizes[lbo] * left_i.sizes[lb] * LI.n_blocks[cle] * site.rp_cols[lbo];
        }
    for (auto& g : G.groups) {
        std::vector<int32_t>().swap(g.kpos);
        std::vector<int32_t>().swap(g.npos);
    }
    auto lowered = std::make_shared<const GroupedLowering>(std::move(G));
    p->s_flops = [lowered, p, lsz = left_i.sizes, rsz = right_i.sizes] {
        std::vector<double> wd(p->ws_init_elems);
        const StagePlan::Bsr& B = p->st.bsr;
        if (!wd.empty() && p->st.bsr_on) {
            // the block-sparse plan keeps only W_D's nonzero tiles on the device: rebuild W_D from them
            constexpr uint32_t TK = S_BSR_TILE_K, TN = S_BSR_TILE_N;
            std::vector<double> tiles(B.tiles * TN * TK);
            if (!tiles.empty()) {
                HIPCHK(hipMemcpyAsync(tiles.data(), B.d_w64.p, tiles.size() * sizeof(double), hipMemcpyDeviceToHost,
                                      p->ctx->stream));
                HIPCHK(hipStreamSynchronize(p->ctx->stream));
            }
            for (uint64_t t = 0; t < B.tile_at.size(); ++t) {
                const auto& at = B.tile_at[t];
                const auto& g = lowered->groups[at.group];
                const uint64_t Kg = g.K.size(), Ng = g.N.size();
                double* Wg = wd.data() + (g.w_off - p->ws_init_off);
                for (uint32_t n = 0; n < TN && at.n0 + n < Ng; ++n)
                    for (uint32_t k = 0; k < TK && at.k0 + k < Kg; ++k)
                        Wg[uint64_t(at.n0 + n) * Kg + at.k0 + k] = tiles[t * TN * TK + n * TK + k];
            }
        } else if (!wd.empty()) {
            // W_D from the plan's workspace (no application writes that region)
            HIPCHK(hipMemcpyAsync(wd.data(), p->ws.as<double>() + p->ws_init_off, wd.size() * sizeof(double),
                                  hipMemcpyDeviceToHost, p->ctx->stream));
            HIPCHK(hipStreamSynchronize(p->ctx->stream));
        }
        return grouped_s_flops(*lowered, wd.data(), p->ws_init_off, lsz, rsz);
    };
}

}  // namespace

extern "C" {

int mq_sigma_plan_create_from_schedule(mq_ctx* ctx, const mq_schedule_desc* sched, uint32_t n_ket_blocks,
                                       const uint32_t* ket_rows, const uint32_t* ket_cols, const mq_boundary* left,
                                       const mq_boundary* right, mq_plan** out) {
    return guarded([&] {
        MQ_REQUIRE(ctx && sched && ket_rows && ket_cols && left && right && out, "bad argument");
        MQ_REQUIRE(sched->n_blocks == n_ket_blocks, "schedule block count != psi block count");
        // sizes of the flattened arrays
        uint32_t nc = uint32_t(sched->coh_begin[sched->n_blocks]);
        uint32_t nu = uint32_t(sched->coh_unit_begin[nc]);
        uint64_t nl = sched->unit_line_begin[nu], nt = sched->unit_task_begin[nu];
        Schedule S;
        S.import_desc(*sched, nc, nu, nl, nt);
        auto p = std::make_unique<mq_plan>();
        p->ctx = ctx;
        p->bind(left, right);
        std::vector<uint32_t> rows(ket_rows, ket_rows + n_ket_blocks), cols(ket_cols, ket_cols + n_ket_blocks);
        std::vector<uint64_t> off(n_ket_blocks);
        uint64_t tot = 0;
        for (uint32_t b = 0; b < n_ket_blocks; ++b) off[b] = tot, tot += round_up(uint64_t(rows[b]) * cols[b], MQ_BUFFER_ALIGNMENT);
        p->vec_size = tot;
        p->out_offsets = off;
        lower_sigma(p.get(), S, rows, cols, off);
        p->flops = schedule_flops(S, right->index, rows);
        StagePlan& P = p->st;
        P.finish_tables();
        P.sizes[BASE_WS] = P.ws_elems;
        P.sizes[BASE_IN] = P.sizes[BASE_OUT] = tot;
        P.sizes[BASE_LEFT] = left->total;
        P.sizes[BASE_RIGHT] = right->total;
        P.validate();
        p->ws.alloc_zeroed(std::max<uint64_t>(P.ws_elems, 32) * sizeof(double), ctx->stream);
        P.upload_all(ctx->stream);
        HIPCHK(hipStreamSynchronize(ctx->stream));
        *out = p.release();
    });
}

int mq_sigma_plan_create_ex(mq_ctx* ctx, const mq_site_desc* site, const mq_mpo_desc* mpo, const mq_boundary* left,
                            const mq_boundary* right, int algorithm, mq_plan** out) {
    return guarded([&] {
        MQ_REQUIRE(ctx && site && mpo && left && right && out, "bad argument");
        trace_phase("(caller)");
        SiteLayout L(*site);
        if (algorithm == MQ_SIGMA_AUTO) algorithm = MQ_SIGMA_GROUPED;
        if (algorithm == (MQ_SIGMA_AUTO | MQ_SIGMA_ADJOINT)) algorithm = MQ_SIGMA_GROUPED | MQ_SIGMA_ADJOINT;
        if (algorithm == MQ_SIGMA_SCHEDULE) {
            int rc = mq_sigma_plan_create(ctx, site, mpo, left, right, out);
            if (rc != MQ_OK) fail(rc, g_err);
            return;
        }
        const bool adjoint = (algorithm & MQ_SIGMA_ADJOINT) != 0;
        algorithm &= ~int(MQ_SIGMA_ADJOINT);
        MQ_REQUIRE(algorithm == MQ_SIGMA_GROUPED,
                   adjoint ? "the sigma adjoint needs a GROUPED plan" : "unknown sigma algorithm");
        Mpo W(*mpo);
        auto p = std::make_unique<mq_plan>();
        p->ctx = ctx;
        p->bind(left, right);
        p->vec_size = L.rp_total;
        uint64_t wd_elems = 0, ws_init_off = 0;
        trace_phase("plan: layout, mpo");
        lower_sigma_grouped(p.get(), L, W, wd_elems, ws_init_off, adjoint);
        trace_phase("lower: final stage");
        StagePlan& P = p->st;
        P.blas = ctx->blas;
        P.lt_ctx = ctx;
        P.finish_tables();
        trace_phase("plan: tables");
        P.sizes[BASE_WS] = P.ws_elems;
        P.sizes[BASE_IN] = P.sizes[BASE_OUT] = L.rp_total;
        P.sizes[BASE_LEFT] = left->total;
        P.sizes[BASE_RIGHT] = right->total;
        P.sizes[BASE_BRA] = P.part_elems;
        P.validate();
        // a reused workspace is cleared for library S stages (the adjoint, MQ_S_DENSE=1): library
        // GEMMs called with beta = 0 were seen to propagate NaN from stale memory in their C operand
        const bool zero_reused = adjoint || !P.bsr_on;
        if (P.part_elems) ctx->take(mq_ctx::POOL_PART, p->part, P.part_elems * sizeof(double), zero_reused);
        ctx->take(mq_ctx::POOL_WS, p->ws, std::max<uint64_t>(P.ws_elems, 32) * sizeof(double), zero_reused);
        // the dense W_D serve the library S stage and the adjoint; the block-sparse S stage reads tiles
        if (wd_elems && !P.bsr_on)
            HIPCHK(hipMemcpyAsync(p->ws.as<double>() + ws_init_off, ctx->wd_host, wd_elems * sizeof(double),
                                  hipMemcpyHostToDevice, ctx->stream));
        trace_phase("plan: workspaces");
        P.upload_all(ctx->stream);
        HIPCHK(hipStreamSynchronize(ctx->stream));
        trace_phase("plan: upload");
        p->grouped = true;
        p->ws_init_elems = wd_elems;
        p->ws_init_off = ws_init_off;
        *out = p.release();
    });
}

int mq_sigma_plan_create(mq_ctx* ctx, const mq_site_desc* site, const mq_mpo_desc* mpo, const mq_boundary* left,
                         const mq_boundary* right, mq_plan** out) {
    return guarded([&] {
        MQ_REQUIRE(ctx && site && mpo && left && right && out, "bad argument");
        SiteLayout L(*site);
        Mpo W(*mpo);
        Schedule S = build_sigma_schedule(L, W, left->index, right->index);
        mq_schedule_desc d{};
        S.export_desc(&d);
        int rc = mq_sigma_plan_create_from_schedule(ctx, &d, L.left_i.size(), L.left_i.sizes.data(), L.rp_cols.data(),
                                                    left, right, out);
        if (rc != MQ_OK) fail(rc, g_err);
    });
}

int mq_sigma_apply_ex(mq_plan* p, const double* psi, double* sigma, int precision) {
    return guarded([&] {
        MQ_REQUIRE(p && psi && sigma, "bad argument");
        MQ_REQUIRE(psi != sigma, "psi and sigma must not alias");
        MQ_REQUIRE(precision == MQ_PREC_F64 || precision == MQ_PREC_F32 || precision == MQ_PREC_BF16X3,
                   "unknown precision");
        hipStream_t st = p->ctx->stream;
        if (!p->use_graphs) {
            Bases b = p->bases(psi, sigma, precision);
            HIPCHK(launch_fill(st, sigma, p->vec_size, 0.0));
            p->st.run(st, b, precision);
            return;
        }
        // Graph replay: one graph per precision over plan-owned operand slots, so that every
        // application (any psi / sigma, e.g. each Jacobi-Davidson iteration) replays it; the
        // operands are copied in and out around the replay.  The first application of a precision
        // runs eagerly (libraries load code and workspaces on first use, which a capture must not
        // contain), the second is captured.
        if (!p->io.p) p->io.alloc(2 * std::max<uint64_t>(p->vec_size, 1) * sizeof(double));
        double* in = p->io.as<double>();
        double* out = in + std::max<uint64_t>(p->vec_size, 1);
        Bases b = p->bases(in, out, precision);   // allocates the FP32 / bf16 workspaces before any capture
        auto issue = [&] {
            HIPCHK(launch_fill(st, out, p->vec_size, 0.0));
            p->st.run(st, b, precision);
        };
        HIPCHK(hipMemcpyAsync(in, psi, p->vec_size * sizeof(double), hipMemcpyDeviceToDevice, st));
        hipGraphExec_t& ge = p->graphs[precision];
        if (!ge && !p->seen[precision]) {
            p->seen[precision] = true;
            issue();
        } else {
            if (!ge) {
                hipGraph_t g = nullptr;
                HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                try {
                    issue();
                } catch (...) {
                    (void)hipStreamEndCapture(st, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                HIPCHK(hipStreamEndCapture(st, &g));
                hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                HIPCHK(e);
            }
            HIPCHK(hipGraphLaunch(ge, st));
        }
        HIPCHK(hipMemcpyAsync(sigma, out, p->vec_size * sizeof(double), hipMemcpyDeviceToDevice, st));
    });
}

int mq_sigma_apply_batched(mq_ctx* ctx, int n, mq_plan* const* plans, const double* const* psi,
                           double* const* sigma, int precision) {
    return guarded([&] {
        MQ_REQUIRE(ctx && n > 0 && plans && psi && sigma, "bad argument");
        MQ_REQUIRE(precision == MQ_PREC_F64 || precision == MQ_PREC_F32 || precision == MQ_PREC_BF16X3,
                   "unknown precision");
        std::vector<const mq_plan*> key(plans, plans + n);
        for (int i = 0; i < n; ++i) {
            MQ_REQUIRE(plans[i] && plans[i]->ctx == ctx && psi[i] && sigma[i] && psi[i] != sigma[i],
                       "batch operand " + std::to_string(i));
            for (int j = 0; j < i; ++j) MQ_REQUIRE(plans[j] != plans[i], "a plan appears twice in the batch");
        }
        hipStream_t st = ctx->stream;
        constexpr int NSIDE = 4;
        if (ctx->side.empty()) {
            ctx->side.resize(NSIDE);
            ctx->side_ev.resize(NSIDE);
            for (int k = 0; k < NSIDE; ++k) {
                HIPCHK(hipStreamCreateWithFlags(&ctx->side[k], hipStreamNonBlocking));
                HIPCHK(hipEventCreateWithFlags(&ctx->side_ev[k], hipEventDisableTiming));
                ctx->lt_ws_side.push_back(std::make_unique<DevBuf>());
                ctx->lt_ws_side.back()->alloc(LT_WORKSPACE);
                rocblas_handle h = nullptr;
                BLASCHK(rocblas_create_handle(&h));
                ctx->side_blas.push_back(h);
                BLASCHK(rocblas_set_stream(h, ctx->side[k]));
            }
            HIPCHK(hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming));
        }
        // operand slots and workspaces of every plan exist before any capture
        std::vector<Bases> bs(n);
        std::vector<double*> ins(n), outs(n);
        for (int i = 0; i < n; ++i) {
            mq_plan* p = plans[i];
            if (!p->io.p) p->io.alloc(2 * std::max<uint64_t>(p->vec_size, 1) * sizeof(double));
            ins[i] = p->io.as<double>();
            outs[i] = ins[i] + std::max<uint64_t>(p->vec_size, 1);
            bs[i] = p->bases(ins[i], outs[i], precision);
        }
        // the batch: fork the instances over the side streams, join back on the context's stream
        auto issue = [&] {
            HIPCHK(hipEventRecord(ctx->fork_ev, st));
            for (int k = 0; k < NSIDE; ++k) HIPCHK(hipStreamWaitEvent(ctx->side[k], ctx->fork_ev, 0));
            for (int i = 0; i < n; ++i) {
                hipStream_t s = ctx->side[i % NSIDE];
                HIPCHK(launch_fill(s, outs[i], plans[i]->vec_size, 0.0));
                plans[i]->st.run(s, bs[i], precision);   // library GEMMs take the side stream's own handle
            }
            for (int k = 0; k < NSIDE; ++k) {
                HIPCHK(hipEventRecord(ctx->side_ev[k], ctx->side[k]));
                HIPCHK(hipStreamWaitEvent(st, ctx->side_ev[k], 0));
            }
        };
        for (int i = 0; i < n; ++i)
            HIPCHK(hipMemcpyAsync(ins[i], psi[i], plans[i]->vec_size * sizeof(double), hipMemcpyDeviceToDevice, st));
        BatchGraph& bg = ctx->batches[std::make_pair(key, precision)];
        if (!bg.exec && !bg.seen) {
            bg.seen = true;   // first use runs eagerly (library code loads on first use)
            issue();
        } else {
            if (!bg.exec) {
                hipGraph_t g = nullptr;
                HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                try {
                    issue();
                } catch (...) {
                    (void)hipStreamEndCapture(st, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                HIPCHK(hipStreamEndCapture(st, &g));
                hipError_t e = hipGraphInstantiate(&bg.exec, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                HIPCHK(e);
            }
            HIPCHK(hipGraphLaunch(bg.exec, st));
        }
        for (int i = 0; i < n; ++i)
            HIPCHK(hipMemcpyAsync(sigma[i], outs[i], plans[i]->vec_size * sizeof(double), hipMemcpyDeviceToDevice, st));
    });
}

int mq_plan_autotune(mq_plan* p, int precision, int* timed) {
    return guarded([&] {
        MQ_REQUIRE(p, "bad argument");
        const bool library_only = (precision & MQ_AUTOTUNE_LIBRARY_ONLY) != 0;
        precision &= ~int(MQ_AUTOTUNE_LIBRARY_ONLY);
        MQ_REQUIRE(precision == MQ_PREC_F64 || precision == MQ_PREC_F32 || precision == MQ_PREC_BF16X3,
                   "unknown precision");
        mq_ctx* ctx = p->ctx;
        hipStream_t st = ctx->stream;
        StagePlan& P = p->st;
        if (library_only && (precision == MQ_PREC_F64 || !p->grouped)) {   // nothing to tune
            if (timed) *timed = 0;
            return;
        }
        if (library_only) {   // only GEMMs whose algorithm is not cached need operands and timing
            bool any = false;
            for (size_t i = 0; i < P.lib.size() && !any; ++i)
                any = !P.lib[i].sparse && P.lt_prepare(precision == MQ_PREC_F32 ? 0 : 1, i, st) &&
                      !P.ltg[precision == MQ_PREC_F32 ? 0 : 1][i].chosen;
            if (!any) {
                if (timed) *timed = 0;
                return;
            }
        }
        if (!p->io.p) p->io.alloc(2 * std::max<uint64_t>(p->vec_size, 1) * sizeof(double));
        double* in = p->io.as<double>();
        double* out = in + std::max<uint64_t>(p->vec_size, 1);
        Bases b = p->bases(in, out, precision);
        HIPCHK(launch_fill(st, in, p->vec_size, 0.0));   // values do not change the timing
        HIPCHK(launch_fill(st, out, p->vec_size, 0.0));
        P.run(st, b, precision);   // a full application: the final stage then reads a valid S
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        int n_timed = 0;
        auto time_stage = [&](int k) {
            P.stage(k, st, b, precision);   // warm-up
            HIPCHK(hipEventRecord(e0, st));
            for (int r = 0; r < 3; ++r) P.stage(k, st, b, precision);
            HIPCHK(hipEventRecord(e1, st));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            ++n_timed;
            return ms;
        };
        struct StageRef { int k; GemmStage* g; const std::vector<GemmSeg>* segs; };
        for (StageRef sr : {StageRef{1, &P.gT, &P.t_segs}, StageRef{3, &P.gF, &P.f_segs}}) {
            if (library_only) break;
            GemmStage& G = *sr.g;
            std::vector<uint64_t> cf = G.class_flops(*sr.segs);
            std::vector<int> key = {sr.k, precision, int(p->grouped)};
            for (auto f : cf) key.push_back(f ? int(std::log2(double(f))) : -1);
            auto apply = [&](const TilePolicy& pol) {
                G.policy = pol;
                G.rebuild();
                G.upload_all(*sr.segs, st);
            };
            auto it = ctx->tuned.find(key);
            if (it != ctx->tuned.end()) {
                TilePolicy pol;
                for (int c = 0; c < N_SIZE_CLASS; ++c) pol.cls[c] = it->second[c];
                apply(pol);
                continue;
            }
            TilePolicy pol;
            for (int c = 0; c < N_SIZE_CLASS; ++c) {
                if (!cf[c]) continue;
                apply(pol);
                float best = time_stage(sr.k);
                int best_cfg = -1;
                for (int cand : CLASS_CANDIDATES[c]) {
                    TilePolicy trial = pol;
                    trial.cls[c] = cand;
                    apply(trial);
                    float t = time_stage(sr.k);
                    if (t < 0.97f * best) best = t, best_cfg = cand;
                }
                pol.cls[c] = best_cfg;
            }
            apply(pol);
            std::array<int, 5> rec;
            for (int c = 0; c < N_SIZE_CLASS; ++c) rec[c] = pol.cls[c];
            ctx->tuned[key] = rec;
        }
        if (precision != MQ_PREC_F64)   // the library GEMMs of the S stage: fastest hipBLASLt candidate
            for (size_t i = 0; i < P.lib.size(); ++i)
                n_timed += P.lt_tune(precision == MQ_PREC_F32 ? 0 : 1, i, st, b, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        HIPCHK(hipStreamSynchronize(st));
        // the device task tables were re-uploaded: graphs that captured the old ones are stale
        p->drop_graphs();
        auto& B = ctx->batches;
        for (auto it = B.begin(); it != B.end();) {
            const auto& ps = it->first.first;
            if (std::find(ps.begin(), ps.end(), p) != ps.end()) {
                if (it->second.exec) (void)hipGraphExecDestroy(it->second.exec);
                it = B.erase(it);
            } else {
                ++it;
            }
        }
        if (timed) *timed = n_timed;
    });
}

int mq_sigma_adjoint(mq_plan* p, const double* psi, double* grad) {
    return guarded([&] {
        MQ_REQUIRE(p && psi && grad, "bad argument");
        MQ_REQUIRE(p->adjoint, "plan created without MQ_SIGMA_ADJOINT");
        mq_ctx* ctx = p->ctx;
        hipStream_t st = ctx->stream;
        StagePlan& P = p->st;
        if (!p->d_grad.p) {
            upload(p->d_adj_ptr, p->adj_ptr, st);
            upload(p->d_adj_idx, p->adj_idx, st);
            upload(p->d_adj_coef, p->adj_coef, st);
            p->d_grad.alloc(std::max<uint32_t>(p->n_terms, 1) * sizeof(double));
            p->adj_g.alloc_zeroed(std::max<uint64_t>(P.ws_elems - p->ws_init_off, 1) * sizeof(double), st);
        }
        Bases b = p->bases(psi, nullptr, MODE_F64);
        double* ws = b.p[BASE_WS];
        // T_D = psi R (FP64), then U_D = L^T psi in the S_D region (cleared first: keys without a
        // left block stay zero)
        P.stage(0, st, b, MODE_F64);
        P.stage(1, st, b, MODE_F64);
        HIPCHK(launch_fill(st, ws + p->s_region0, p->s_region1 - p->s_region0, 0.0));
        P.run_u(st, b);
        // G_g = T_g^T U_g in the W_D layout: the derivative of psi . sigma with respect to W_D
        BLASCHK(rocblas_set_stream(P.blas, st));
        const double one = 1.0, zero = 0.0;
        double* G = p->adj_g.as<double>();
        for (const auto& g : P.lib)
            BLASCHK(rocblas_dgemm(P.blas, rocblas_operation_transpose, rocblas_operation_none, g.K, g.N, g.M, &one,
                                  ws + g.a_off, g.lda, ws + g.c_off, g.ldc, &zero, G + (g.b_off - p->ws_init_off), g.K));
        // per term: the coefficients of its op entries against G
        HIPCHK(launch_term_grad(st, p->n_terms, p->d_adj_ptr.as<uint32_t>(), p->d_adj_idx.as<uint64_t>(),
                                p->d_adj_coef.as<double>(), G, p->d_grad.as<double>()));
        HIPCHK(hipMemcpyAsync(grad, p->d_grad.p, p->n_terms * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    });
}

int mq_plan_set_graphs(mq_plan* p, int enable) {
    return guarded([&] {
        MQ_REQUIRE(p, "bad argument");
        HIPCHK(hipStreamSynchronize(p->ctx->stream));
        p->use_graphs = enable != 0;
        if (!p->use_graphs) p->drop_graphs();
    });
}

int mq_sigma_apply(mq_plan* p, const double* psi, double* sigma) {
    return mq_sigma_apply_ex(p, psi, sigma, MQ_PREC_F64);
}

int mq_plan_profile(mq_plan* p, const double* psi, double* sigma, int reps, int precision, double* stage_ms,
                    uint64_t* flops_out, uint64_t* gather_bytes) {
    return guarded([&] {
        MQ_REQUIRE(p && psi && sigma && stage_ms && reps > 0, "bad argument");
        MQ_REQUIRE(psi != sigma, "psi and sigma must not alias");
        MQ_REQUIRE(precision == MQ_PREC_F64 || precision == MQ_PREC_F32 || precision == MQ_PREC_BF16X3,
                   "unknown precision");
        const StagePlan& P = p->st;
        hipStream_t st = p->ctx->stream;
        Bases b = p->bases(psi, sigma, precision);
        hipEvent_t ev[5];
        for (auto& e : ev) HIPCHK(hipEventCreate(&e));
        double acc[4] = {0, 0, 0, 0};
        for (int r = 0; r < reps; ++r) {
            HIPCHK(hipEventRecord(ev[0], st));
            HIPCHK(launch_fill(st, sigma, p->vec_size, 0.0));
            for (int k = 0; k < 4; ++k) {
                P.stage(k, st, b, precision);
                HIPCHK(hipEventRecord(ev[k + 1], st));
            }
            HIPCHK(hipEventRecord(ev[4], st));
            HIPCHK(hipEventSynchronize(ev[4]));
            for (int k = 0; k < 4; ++k) {
                float ms = 0;
                HIPCHK(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
                acc[k] += ms;
            }
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
        for (int k = 0; k < 4; ++k) stage_ms[k] = acc[k] / reps;
        if (flops_out) {
            uint64_t ft = 0, fg = 0, ff = 0;
            auto gemm_flops = [](const std::vector<GemmTask>& tasks, const std::vector<GemmSeg>& segs) {
                uint64_t f = 0;
                for (const auto& t : tasks)
                    for (uint32_t s = t.seg0; s < t.seg1; ++s) f += 2ull * t.M * t.N * segs[s].K;
                return f;
            };
            ft = gemm_flops(P.t_tasks, P.t_segs);
            ff = gemm_flops(P.f_tasks, P.f_segs);
            P.for_regions([&](const GatherRegion& r) { fg += 2ull * r.R * r.C * (r.t1 - r.t0); });
            // S stage: the block-sparse kernel's visited tiles, and the dense library GEMMs of the rest
            for (const auto& g : P.lib)
                if (!g.sparse) fg += 2ull * g.M * g.N * g.K;
            fg += P.bsr.flops;
            flops_out[0] = ft;
            flops_out[1] = fg;
            flops_out[2] = ff;
        }
        if (gather_bytes) {
            uint64_t by = 0;
            P.for_regions([&](const GatherRegion& r) { by += 8ull * r.R * r.C * (uint64_t(r.t1 - r.t0) + 1); });
            *gather_bytes = by;
        }
    });
}

int mq_plan_vector_size(const mq_plan* p, uint64_t* n) {
    return guarded([&] {
        MQ_REQUIRE(p && n, "bad argument");
        *n = p->vec_size;
    });
}

int mq_plan_flops(const mq_plan* p, uint64_t* f) {
    return guarded([&] {
        MQ_REQUIRE(p && f, "bad argument");
        if (p->s_flops) {
            p->flops += p->s_flops();
            p->s_flops = nullptr;
        }
        *f = p->flops;
    });
}

int mq_plan_workspace_bytes(const mq_plan* p, uint64_t* bytes) {
    return guarded([&] {
        MQ_REQUIRE(p && bytes, "bad argument");
        const StagePlan& P = p->st;
        *bytes = p->ws.bytes + p->ws32.bytes + p->part.bytes + P.gT.device_bytes() + P.gF.device_bytes() + P.bsr.d_tasks.bytes + P.bsr.d_runs.bytes + P.bsr.d_w64.bytes + P.bsr.d_w32.bytes + P.bsr.d_whi.bytes + P.bsr.d_wlo.bytes + P.red.device_bytes() + P.d_maps.bytes + P.d_units.bytes + P.d_rows.bytes + P.d_regions.bytes +
                 P.d_small.bytes + P.d_src_off.bytes + P.d_alpha.bytes;
    });
}

int mq_plan_destroy(mq_plan* p) {
    return guarded([&] {
        if (p && p->ctx) {
            (void)hipStreamSynchronize(p->ctx->stream);
            if (p->grouped) {   // the large workspaces go back to the context for the next plans
                mq_ctx* c = p->ctx;
                c->pool_put(mq_ctx::POOL_WS, p->ws);
                c->pool_put(mq_ctx::POOL_WS32, p->ws32);
                c->pool_put(mq_ctx::POOL_BF, p->bf);
                c->pool_put(mq_ctx::POOL_PART, p->part);
            }
            // batch graphs that contain the plan refer to its buffers
            auto& B = p->ctx->batches;
            for (auto it = B.begin(); it != B.end();) {
                const auto& ps = it->first.first;
                if (std::find(ps.begin(), ps.end(), p) != ps.end()) {
                    if (it->second.exec) (void)hipGraphExecDestroy(it->second.exec);
                    it = B.erase(it);
                } else {
                    ++it;
                }
            }
        }
        delete p;
    });
}

}  // extern "C"

// ============================================================================ boundary steps
namespace {

// New boundary storage for the cohorts of a step: schedule cohort c is new cohort c
// (add_cohort in schedule order, move_boundary.hpp:202-205, 331-336).
std::unique_ptr<mq_boundary> new_boundary(mq_ctx* ctx, const Schedule& S) {
    auto b = std::make_unique<mq_boundary>();
    b->ctx = ctx;
    b->index = S.out_index;
    b->layout();
    b->data.alloc(b->total * sizeof(double));
    HIPCHK(launch_fill(ctx->stream, b->ptr(), b->total, 0.0));
    return b;
}

}  // namespace

extern "C" {

// overlap_mpo_left_step (move_boundary.hpp:147-279): T_ti = L_b * ket slab (create_T_left,
// tasks.cpp:545-648), S_left gather (create_s, tasks.cpp:425-464), L'(ci) = bra_lp^T * S (prop_l,
// tasks.cpp:198-241).
static int left_step_schedule(mq_ctx* ctx, const mq_site_desc* site, const mq_mpo_desc* mpo, const double* bra_lp,
                              const double* ket_rp, const mq_boundary* left, int symmetric, mq_boundary** out) {
    return guarded([&] {
        MQ_REQUIRE(ctx && site && mpo && bra_lp && ket_rp && left && out, "bad argument");
        SiteLayout L(*site);
        Mpo W(*mpo);
        Schedule S = build_left_schedule(L, W, left->index, symmetric != 0);
        const BoundaryIndex& LI = left->index;
        auto nb = new_boundary(ctx, S);
        StagePlan P;
        uint64_t ws = 0;
        std::vector<uint64_t> t_off(S.t_ci.size());
        for (uint32_t m = 0; m < S.n_blocks; ++m) {   // m = rb_ket
            uint32_t N = L.right_i.sizes[m];
            for (uint64_t ti = S.t_begin[m]; ti < S.t_begin[m + 1]; ++ti) {
                uint32_t ci = S.t_ci[ti], cie = S.t_ci_eff[ti], lbk = S.t_lb_ket[ti];
                uint32_t M = uint32_t(LI.left_size(ci)), K = uint32_t(LI.right_size(ci)), nbk = LI.n_blocks[cie];
                MQ_REQUIRE(K == L.left_i.sizes[lbk], "left cohort ket size != ket block rows");
                t_off[ti] = ws;
                ws += align32(uint64_t(nbk) * M * N);
                uint64_t b_off = L.rp_offset[lbk] + uint64_t(S.t_mps_offset[ti]) * K;
                for (uint32_t b = 0; b < nbk; ++b) {
                    uint64_t a_off = left->cohort_off[cie] + uint64_t(b) * M * K;
                    bool tr = (ci != cie);
                    add_gemm(P.t_tasks, P.t_segs, BASE_WS, t_off[ti] + uint64_t(b) * M * N, M, N, M,
                             {seg(BASE_LEFT, a_off, tr ? K : M, tr, BASE_IN, b_off, K, false, K)});
                }
            }
        }
        // S (stripe x nSrows*rs), column stride stripe (create_s, tasks.cpp:425-454)
        std::vector<uint64_t> s_off(S.n_cohorts());
        P.s_begin = ws;
        for (uint32_t m = 0; m < S.n_blocks; ++m)
            for (uint64_t c = S.coh_begin[m]; c < S.coh_begin[m + 1]; ++c) {
                uint32_t rs = S.coh_rs[c], nS = S.coh_nSrows[c], stripe = S.coh_stripe[c];
                s_off[c] = ws;
                ws += align32(uint64_t(stripe) * nS * rs);
                for (uint64_t u = S.coh_unit_begin[c]; u < S.coh_unit_begin[c + 1]; ++u) {
                    uint32_t ms = S.unit_ms[u], off = S.unit_offset[u];
                    uint64_t task = S.unit_task_begin[u];
                    for (uint64_t l = S.unit_line_begin[u]; l < S.unit_line_begin[u + 1]; ++l) {
                        uint32_t b = S.line_b1[l], cnt = S.line_count[l];
                        MQ_REQUIRE(b < nS && uint64_t(off) + ms <= stripe, "S region outside S");
                        GatherRegion r{};
                        r.dst_off = s_off[c] + uint64_t(b) * rs * stripe + off;
                        r.R = ms;
                        r.C = rs;
                        r.ld_dst = stripe;
                        r.t0 = uint32_t(P.src_off.size());
                        for (uint32_t k = 0; k < cnt; ++k, ++task) {
                            uint64_t ti = S.t_begin[m] + S.task_ti[task];
                            MQ_REQUIRE(ti < S.t_begin[m + 1], "task ti out of range");
                            P.src_off.push_back(t_off[ti] + uint64_t(S.task_col[task]) * rs);
                            P.alpha.push_back(S.task_alpha[task]);
                        }
                        r.t1 = uint32_t(P.src_off.size());
                        if (uint64_t(r.R) * r.C) P.regions.push_back(r);
                    }
                }
            }
        P.s_elems = ws - P.s_begin;
        // prop_l: new cohort c (rs_bra x nSrows*rs_ket) = bra_lp(rb_bra)^T * S
        for (uint32_t c = 0; c < S.n_cohorts(); ++c) {
            uint32_t rbb = S.coh_lb[c], ls = S.coh_ls[c], rs = S.coh_rs[c], nS = S.coh_nSrows[c], stripe = S.coh_stripe[c];
            MQ_REQUIRE(stripe == L.lp_rows[rbb] && ls == L.right_i.sizes[rbb], "prop_l shape mismatch");
            MQ_REQUIRE(nb->index.cohort_size(c) == uint64_t(nS) * ls * rs, "new cohort size mismatch");
            add_gemm(P.f_tasks, P.f_segs, BASE_OUT, nb->cohort_off[c], ls, nS * rs, ls,
                     {seg(BASE_BRA, L.lp_offset[rbb], stripe, true, BASE_WS, s_off[c], stripe, false, stripe)});
        }
        P.ws_elems = ws;
        P.finish_tables();
        P.sizes[BASE_WS] = ws;
        P.sizes[BASE_IN] = L.rp_total;
        P.sizes[BASE_BRA] = L.lp_total;
        P.sizes[BASE_OUT] = nb->total;
        P.sizes[BASE_LEFT] = left->total;
        P.validate();
        P.upload_all(ctx->stream);
        ctx->step_ws.ensure(std::max<uint64_t>(ws, 32) * sizeof(double));
        Bases bs{};
        bs.p[BASE_WS] = ctx->step_ws.as<double>();
        bs.p[BASE_IN] = const_cast<double*>(ket_rp);
        bs.p[BASE_OUT] = nb->ptr();
        bs.p[BASE_LEFT] = left->ptr();
        bs.p[BASE_BRA] = const_cast<double*>(bra_lp);
        P.run(ctx->stream, bs);
        HIPCHK(hipStreamSynchronize(ctx->stream));
        *out = nb.release();
    });
}

// overlap_mpo_right_step (move_boundary.hpp:281-404): T as in sigma, S = create_s_r, then
// R'_b = S_b * bra_rp^T for every b-row (prop_r, tasks.cpp:243-284).
static int right_step_schedule(mq_ctx* ctx, const mq_site_desc* site, const mq_mpo_desc* mpo, const double* bra_rp,
                               const double* ket_rp, const mq_boundary* right, int symmetric, mq_boundary** out) {
    return guarded([&] {
        MQ_REQUIRE(ctx && site && mpo && bra_rp && ket_rp && right && out, "bad argument");
        boundary_ready(right, ctx->stream);
        SiteLayout L(*site);
        Mpo W(*mpo);
        Schedule S = build_right_schedule(L, W, right->index, symmetric != 0);
        auto nb = new_boundary(ctx, S);
        StagePlan P;
        uint64_t ws = 0;
        std::vector<uint64_t> t_off = lower_T_right(P, S, right->index, right->cohort_off, L.left_i.sizes, L.rp_offset, ws);
        std::vector<uint64_t> s_off = lower_S_right(P, S, t_off, ws);
        for (uint32_t c = 0; c < S.n_cohorts(); ++c) {
            uint32_t lbb = S.coh_rb[c], ls = S.coh_ls[c], rs = S.coh_rs[c], nS = S.coh_nSrows[c], stripe = S.coh_stripe[c];
            MQ_REQUIRE(stripe == L.rp_cols[lbb] && rs == L.left_i.sizes[lbb], "prop_r shape mismatch");
            MQ_REQUIRE(nb->index.cohort_size(c) == uint64_t(nS) * ls * rs, "new cohort size mismatch");
            uint32_t ldS = nS * ls;
            for (uint32_t b = 0; b < nS; ++b)
                add_gemm(P.f_tasks, P.f_segs, BASE_OUT, nb->cohort_off[c] + uint64_t(b) * ls * rs, ls, rs, ls,
                         {seg(BASE_WS, s_off[c] + uint64_t(b) * ls, ldS, false, BASE_BRA, L.rp_offset[lbb], rs, true, stripe)});
        }
        P.ws_elems = ws;
        P.finish_tables();
        P.sizes[BASE_WS] = ws;
        P.sizes[BASE_IN] = L.rp_total;
        P.sizes[BASE_BRA] = L.rp_total;
        P.sizes[BASE_OUT] = nb->total;
        P.sizes[BASE_RIGHT] = right->total;
        P.validate();
        P.upload_all(ctx->stream);
        ctx->step_ws.ensure(std::max<uint64_t>(ws, 32) * sizeof(double));
        Bases bs{};
        bs.p[BASE_WS] = ctx->step_ws.as<double>();
        bs.p[BASE_IN] = const_cast<double*>(ket_rp);
        bs.p[BASE_OUT] = nb->ptr();
        bs.p[BASE_RIGHT] = right->ptr();
        bs.p[BASE_BRA] = const_cast<double*>(bra_rp);
        P.run(ctx->stream, bs);
        HIPCHK(hipStreamSynchronize(ctx->stream));
        *out = nb.release();
    });
}

// Grouped boundary step (abelian groups): the new boundary's index from the index-only builder,
// then T (psi R or L psi), S_D = T_D W_D and the final contraction with the bra as stacked tasks.
// bra_rp: the bra in the right-paired layout.
static int step_grouped(mq_ctx* ctx, int kind, const mq_site_desc* site, const mq_mpo_desc* mpo, const double* bra_rp,
                        const double* ket_rp, const mq_boundary* bnd, int symmetric, mq_boundary** out) {
    return guarded([&] {
        trace_phase("(caller)");
        SiteLayout L(*site);
        Mpo W(*mpo);
        trace_phase("step: layout, mpo");
        auto nb = std::make_unique<mq_boundary>();
        nb->ctx = ctx;
        nb->index = kind == Schedule::LEFT ? left_step_index(L, W, bnd->index, symmetric != 0)
                                           : right_step_index(L, W, bnd->index, symmetric != 0);
        nb->layout();
        trace_phase("step: new index");
        nb->data.alloc(std::max<uint64_t>(nb->total, 1) * sizeof(double));
        HIPCHK(launch_fill(ctx->stream, nb->ptr(), std::max<uint64_t>(nb->total, 1), 0.0));
        // The step's intermediates (T_D, S_D) are several times the new boundary.  The ket sectors
        // are processed in chunks whose workspace fits a budget: 40% of the free device memory
        // (at least 1 GiB); a chunk that does not fit is halved.  MQ_STEP_BUDGET_ELEMS overrides the
        // budget (FP64 elements, no floor): a budget of 1 runs every ket sector as its own chunk.
        size_t free_b = 0, total_b = 0;
        HIPCHK(hipMemGetInfo(&free_b, &total_b));
        uint64_t budget = std::max<uint64_t>(uint64_t(1) << 27, uint64_t(0.4 * double(free_b + ctx->step_ws.bytes)) / 8);
        if (const char* env = std::getenv("MQ_STEP_BUDGET_ELEMS")) {
            const unsigned long long v = std::strtoull(env, nullptr, 10);
            MQ_REQUIRE(v > 0, "MQ_STEP_BUDGET_ELEMS must be a positive element count");
            budget = uint64_t(v);
        }
        const uint32_t n_ket = kind == Schedule::LEFT ? L.right_i.size() : L.left_i.size();
        std::vector<std::pair<uint32_t, uint32_t>> todo{{0u, n_ket}};
        DevBuf sws;   // the chunks' workspace
        while (!todo.empty()) {
            const auto range = todo.back();
            todo.pop_back();
            if (range.first >= range.second) continue;
            mq_plan p;
            p.ctx = ctx;
            p.bind(kind == Schedule::LEFT ? bnd : nullptr, kind == Schedule::RIGHT ? bnd : nullptr);
            uint64_t wd_elems = 0, ws_init_off = 0;
            GroupedStep spec{kind, &nb->index, &nb->cohort_off};
            spec.lo = range.first;
            spec.hi = range.second;
            lower_sigma_grouped(&p, L, W, wd_elems, ws_init_off, false, &spec);
            StagePlan& P = p.st;
            if (P.ws_elems > budget && range.second - range.first > 1) {
                const uint32_t mid = range.first + (range.second - range.first) / 2;
                todo.push_back({mid, range.second});
                todo.push_back({range.first, mid});
                continue;
            }
            P.blas = ctx->blas;
            P.finish_tables();
            P.sizes[BASE_WS] = P.ws_elems;
            P.sizes[BASE_IN] = P.sizes[BASE_BRA] = L.rp_total;
            P.sizes[BASE_OUT] = nb->total;
            P.sizes[kind == Schedule::LEFT ? BASE_LEFT : BASE_RIGHT] = bnd->total;
            P.validate();
            P.upload_all(ctx->stream);
            // workspace: the context's pooled workspace (the site's sigma plan has returned it), taken
            // as it is with the block-sparse S stage (the grouped stages clear what they read
            // unwritten; cleared for library GEMMs) and grown between chunks when a chunk needs
            // more; W_D copied in
            trace_phase("step: tables");
            const size_t need = std::max<uint64_t>(P.ws_elems, 32) * sizeof(double);
            if (need > sws.bytes) {
                ctx->pool_put(mq_ctx::POOL_WS, sws);
                ctx->take(mq_ctx::POOL_WS, sws, need, !P.bsr_on);
            }
            double* ws = sws.as<double>();
            if (wd_elems && !P.bsr_on)
                HIPCHK(hipMemcpyAsync(ws + ws_init_off, ctx->wd_host, wd_elems * sizeof(double), hipMemcpyHostToDevice,
                                      ctx->stream));
            Bases bs{};
            bs.p[BASE_WS] = ws;
            bs.p[BASE_IN] = const_cast<double*>(ket_rp);
            bs.p[BASE_OUT] = nb->ptr();
            bs.p[BASE_BRA] = const_cast<double*>(bra_rp);
            bs.p[kind == Schedule::LEFT ? BASE_LEFT : BASE_RIGHT] = const_cast<double*>(bnd->ptr());
            BLASCHK(rocblas_set_stream(ctx->blas, ctx->stream));
            if (g_trace) HIPCHK(hipStreamSynchronize(ctx->stream));
            trace_phase("step: workspace");
            P.run(ctx->stream, bs, MODE_F64);
            HIPCHK(hipStreamSynchronize(ctx->stream));   // the chunk's tables and the W_D staging are reused next
            trace_phase("step: run");
        }
        ctx->pool_put(mq_ctx::POOL_WS, sws);   // back to the pool for the next site's sigma plan
        *out = nb.release();
    });
}

int mq_boundary_left_step_ex(mq_ctx* ctx, const mq_site_desc* site, const mq_mpo_desc* mpo, const double* bra_lp,
                             const double* ket_rp, const mq_boundary* left, int symmetric, int algorithm,
                             mq_boundary** out) {
    return guarded([&] {
        MQ_REQUIRE(ctx && site && mpo && bra_lp && ket_rp && left && out, "bad argument");
        boundary_ready(left, ctx->stream);
        SiteLayout L(*site);
        if (algorithm == MQ_SIGMA_AUTO) algorithm = MQ_SIGMA_GROUPED;
        int rc;
        if (algorithm == MQ_SIGMA_SCHEDULE) {
            rc = left_step_schedule(ctx, site, mpo, bra_lp, ket_rp, left, symmetric, out);
        } else {
            MQ_REQUIRE(algorithm == MQ_SIGMA_GROUPED, "unknown step algorithm");
            // the grouped step reads the bra right-paired
            DevBuf bra_rp;
            bra_rp.alloc(std::max<uint64_t>(L.rp_total, 1) * sizeof(double));
            rc = mq_reshape_left_to_right(ctx, site, bra_lp, bra_rp.as<double>());
            if (rc == MQ_OK) rc = step_grouped(ctx, Schedule::LEFT, site, mpo, bra_rp.as<double>(), ket_rp, left, symmetric, out);
        }
        if (rc != MQ_OK) fail(rc, g_err);
    });
}

int mq_boundary_left_step(mq_ctx* ctx, const mq_site_desc* site, const mq_mpo_desc* mpo, const double* bra_lp,
                          const double* ket_rp, const mq_boundary* left, int symmetric, mq_boundary** out) {
    return mq_boundary_left_step_ex(ctx, site, mpo, bra_lp, ket_rp, left, symmetric, MQ_SIGMA_AUTO, out);
}

int mq_boundary_right_step_ex(mq_ctx* ctx, const mq_site_desc* site, const mq_mpo_desc* mpo, const double* bra_rp,
                              const double* ket_rp, const mq_boundary* right, int symmetric, int algorithm,
                              mq_boundary** out) {
    return guarded([&] {
        MQ_REQUIRE(ctx && site && mpo && bra_rp && ket_rp && right && out, "bad argument");
        boundary_ready(right, ctx->stream);
        SiteLayout L(*site);
        if (algorithm == MQ_SIGMA_AUTO) algorithm = MQ_SIGMA_GROUPED;
        int rc;
        if (algorithm == MQ_SIGMA_SCHEDULE) {
            rc = right_step_schedule(ctx, site, mpo, bra_rp, ket_rp, right, symmetric, out);
        } else {
            MQ_REQUIRE(algorithm == MQ_SIGMA_GROUPED, "unknown step algorithm");
            rc = step_grouped(ctx, Schedule::RIGHT, site, mpo, bra_rp, ket_rp, right, symmetric, out);
        }
        if (rc != MQ_OK) fail(rc, g_err);
    });
}

int mq_boundary_right_step(mq_ctx* ctx, const mq_site_desc* site, const mq_mpo_desc* mpo, const double* bra_rp,
                           const double* ket_rp, const mq_boundary* right, int symmetric, mq_boundary** out) {
    return mq_boundary_right_step_ex(ctx, site, mpo, bra_rp, ket_rp, right, symmetric, MQ_SIGMA_AUTO, out);
}

// ---------------------------------------------------------------------------- vectors
int mq_vec_dot(mq_ctx* ctx, uint64_t n, const double* x, const double* y, double* result) {
    return guarded([&] {
        MQ_REQUIRE(ctx && x && y && result, "bad argument");
        HIPCHK(launch_dot(ctx->stream, n, x, y, ctx->red_partial.as<double>(), ctx->red_result.as<double>()));
        HIPCHK(hipMemcpyAsync(result, ctx->red_result.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    });
}

int mq_vec_multi_dot(mq_ctx* ctx, uint64_t n, int k, const double* V, uint64_t ldv, const double* x, double* result) {
    return guarded([&] {
        MQ_REQUIRE(ctx && V && x && result && k >= 1, "bad argument");
        MQ_REQUIRE(k == 1 || ldv >= n, "ldv < n");
        ctx->mdot_reserve(k);
        HIPCHK(launch_multi_dot(ctx->stream, n, k, V, ldv, x, ctx->mdot_partial.as<double>(),
                                ctx->mdot_result.as<double>()));
        HIPCHK(hipMemcpyAsync(result, ctx->mdot_result.p, size_t(k) * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    });
}

int mq_vec_multi_dot_noalloc(mq_ctx* ctx, uint64_t n, int k, const double* V, uint64_t ldv, const double* x,
                             double* result) {
    return guarded([&] {
        MQ_REQUIRE(ctx && V && x && result && k >= 1, "bad argument");
        MQ_REQUIRE(k <= MQ_NOALLOC_MAXK, "mq_vec_multi_dot_noalloc: k exceeds the static scratch (MQ_NOALLOC_MAXK vectors)");
        MQ_REQUIRE(k == 1 || ldv >= n, "ldv < n");
        MQ_REQUIRE(ctx->noalloc_scratch, "no static scratch address for this device");
        // Every partial sum a call reads is written by the same call: nothing needs clearing.
        HIPCHK(launch_multi_dot(ctx->stream, n, k, V, ldv, x, ctx->noalloc_scratch, result));
    });
}

int mq_vec_nrm2(mq_ctx* ctx, uint64_t n, const double* x, double* result) {
    return guarded([&] {
        double d = 0;
        int rc = mq_vec_dot(ctx, n, x, x, &d);
        if (rc) fail(rc, g_err);
        *result = std::sqrt(d);
    });
}

int mq_vec_axpy(mq_ctx* ctx, uint64_t n, double a, const double* x, double* y) {
    return guarded([&] {
        MQ_REQUIRE(ctx && x && y, "bad argument");
        HIPCHK(launch_axpy(ctx->stream, n, a, x, y));
    });
}

int mq_vec_scal(mq_ctx* ctx, uint64_t n, double a, double* x) {
    return guarded([&] {
        MQ_REQUIRE(ctx && x, "bad argument");
        HIPCHK(launch_scal(ctx->stream, n, a, x));
    });
}

int mq_vec_copy(mq_ctx* ctx, uint64_t n, const double* x, double* y) {
    return guarded([&] {
        MQ_REQUIRE(ctx && x && y, "bad argument");
        if (n) HIPCHK(hipMemcpyAsync(y, x, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    });
}

}  // extern "C"

// ============================================================================ Jacobi-Davidson
namespace {

// Smallest eigenpair of the symmetric dim x dim matrix M (column-major, upper triangle used) by
// cyclic Jacobi rotations; replaces LAPACK dsyevx with range 'I', il = iu = 1 (ietl/jacobi.h:476-516).
double smallest_eigenpair(const std::vector<double>& Mfull, int ld, int dim, std::vector<double>& s) {
    std::vector<double> A(dim * dim), V(dim * dim, 0.0);
    for (int j = 0; j < dim; ++j)
        for (int i = 0; i < dim; ++i) A[i + j * dim] = (i <= j) ? Mfull[i + j * ld] : Mfull[j + i * ld];
    for (int i = 0; i < dim; ++i) V[i + i * dim] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0;
        for (int j = 0; j < dim; ++j)
            for (int i = 0; i < j; ++i) off += A[i + j * dim] * A[i + j * dim];
        if (off < 1e-30) break;
        for (int p = 0; p < dim; ++p)
            for (int q = p + 1; q < dim; ++q) {
                double apq = A[p + q * dim];
                if (std::fabs(apq) < 1e-300) continue;
                double app = A[p + p * dim], aqq = A[q + q * dim];
                double theta = (aqq - app) / (2 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
                double c = 1 / std::sqrt(t * t + 1), sn = t * c;
                for (int k = 0; k < dim; ++k) {
                    double akp = A[k + p * dim], akq = A[k + q * dim];
                    A[k + p * dim] = c * akp - sn * akq;
                    A[k + q * dim] = sn * akp + c * akq;
                }
                for (int k = 0; k < dim; ++k) {
                    double apk = A[p + k * dim], aqk = A[q + k * dim];
                    A[p + k * dim] = c * apk - sn * aqk;
                    A[q + k * dim] = sn * apk + c * aqk;
                }
                for (int k = 0; k < dim; ++k) {
                    double vkp = V[k + p * dim], vkq = V[k + q * dim];
                    V[k + p * dim] = c * vkp - sn * vkq;
                    V[k + q * dim] = sn * vkp + c * vkq;
                }
            }
    }
    int imin = 0;
    for (int i = 1; i < dim; ++i)
        if (A[i + i * dim] < A[imin + imin * dim]) imin = i;
    s.assign(dim, 0.0);
    for (int i = 0; i < dim; ++i) s[i] = V[i + imin * dim];
    return A[imin + imin * dim];
}

}  // namespace

extern "C" {

// ietl::jacobi_davidson::calculate_eigenvalue (ietl/jacobi.h:386-474) with
// basic_iteration(max_iter, tol, tol) (ietl/iteration.h:40-78, solver/ietl_jacobi_davidson.h:58-64).
}  // extern "C"

namespace {

// Device-side Jacobi-Davidson state of one plan: V[0..max_iter], VA[0..max_iter-1], u, uA and two
// spare vectors (x, xA) used by the refinement loop.
struct JdSpace {
    mq_plan* p;
    mq_ctx* ctx;
    uint64_t n, stride;
    int max_iter;
    int gmres_iter = 0;   // Krylov vectors of the GMRES correction (0: the plain correction)
    int n_ortho = 0;      // vectors the search space is kept orthogonal to
    double* base;
    std::vector<double> hres;   // host copy of the last batched reduction
    JdSpace(mq_plan* p_, int max_iter_, int gmres_iter_ = 0, int n_ortho_ = 0)
        : p(p_), ctx(p_->ctx), n(p_->vec_size), stride(align32(p_->vec_size)), max_iter(max_iter_),
          gmres_iter(gmres_iter_), n_ortho(n_ortho_) {
        // V, VA, u, uA, x, xA; GMRES: its basis, w and the operator's scratch; the ortho vectors
        int cap = 2 * max_iter + 5 + (gmres_iter > 0 ? gmres_iter + 3 : 0) + n_ortho;
        if (p->jd_cap < cap) {
            p->jd_buf.alloc(uint64_t(cap) * stride * sizeof(double));
            p->jd_cap = cap;
        }
        base = p->jd_buf.as<double>();
        ctx->mdot_reserve(max_iter + 1);
    }
    // r[j] = W_j . x for j < k with W_j = first + j * stride (contiguous basis vectors), one batched
    // reduction and one host round trip; the results also stay on the device (ctx->mdot_result)
    const double* mdot(const double* first, int k, const double* x) {
        HIPCHK(launch_multi_dot(ctx->stream, n, k, first, stride, x, ctx->mdot_partial.as<double>(),
                                ctx->mdot_result.as<double>()));
        hres.resize(k);
        HIPCHK(hipMemcpyAsync(hres.data(), ctx->mdot_result.p, k * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        return hres.data();
    }
    // y -= sum_{j<k} r_j W_j with the coefficients of the last mdot, still on the device
    void subtract_last(const double* first, int k, double* y) {
        HIPCHK(launch_multi_axpy(ctx->stream, n, k, -1.0, ctx->mdot_result.as<double>(), nullptr, first, stride, y));
    }
    // y = sum_{j<k} c_j W_j (host coefficients)
    void combine(const double* first, int k, const double* c, double* y) {
        HIPCHK(launch_fill(ctx->stream, y, n, 0.0));
        HIPCHK(launch_multi_axpy(ctx->stream, n, k, 1.0, nullptr, c, first, stride, y));
    }
    double* V(int i) { return base + uint64_t(i) * stride; }
    double* VA(int i) { return base + uint64_t(max_iter + 1 + i) * stride; }
    double* u() { return base + uint64_t(2 * max_iter + 1) * stride; }
    double* uA() { return base + uint64_t(2 * max_iter + 2) * stride; }
    double* x() { return base + uint64_t(2 * max_iter + 3) * stride; }
    double* xA() { return base + uint64_t(2 * max_iter + 4) * stride; }
    double* Gv(int i) { return base + uint64_t(2 * max_iter + 5 + i) * stride; }   // GMRES basis, i <= gmres_iter
    double* Gw() { return Gv(gmres_iter + 1); }
    double* Gt() { return Gv(gmres_iter + 2); }
    double* O(int j) { return base + uint64_t(2 * max_iter + 5 + (gmres_iter > 0 ? gmres_iter + 3 : 0) + j) * stride; }
    std::vector<double> o_norm2;   // o_j . o_j

    // The ortho vectors, Gram-Schmidt'ed in order as the caller's copies are in
    // solver/ietl_jacobi_davidson.h:41-47
    void set_ortho(const double* const* vecs) {
        o_norm2.assign(n_ortho, 0.0);
        for (int j = 0; j < n_ortho; ++j) {
            copy(vecs[j], O(j));
            for (int j0 = 0; j0 < j; ++j0) axpy(-dot(O(j0), O(j)) / o_norm2[j0], O(j0), O(j));
            o_norm2[j] = dot(O(j), O(j));
            MQ_REQUIRE(o_norm2[j] > 0, "Jacobi-Davidson: an ortho vector is zero or dependent on the others");
        }
    }
    // t -= (o.t)/(o.o) o for each ortho vector in turn (DavidsonVS::project, solver/ietl_interface_dv.h:94-99)
    void project(double* t) {
        for (int j = 0; j < n_ortho; ++j) axpy(-dot(O(j), t) / o_norm2[j], O(j), t);
    }
    // y = (1 - u u^T)(H - theta)(1 - u u^T) x  (jcd_solver_operator, ietl/jacobi.h:266-284)
    void jcd_op(const double* x, double* y, double theta, int precision, int* n_sigma) {
        double* t2 = Gt();
        copy(x, t2);
        axpy(-dot(u(), x), u(), t2);
        sigma(t2, y, precision);
        ++*n_sigma;
        axpy(-theta, t2, y);
        axpy(-dot(u(), y), u(), y);
    }
    // GMRES on the correction equation from the initial guess t (ietl_gmres, ietl/gmres.h:78-140,
    // with its bookkeeping: at most gmres_iter - 1 Arnoldi steps, and on convergence at step i the
    // update uses the first i basis vectors).  b = -r; the result overwrites t.
    void gmres(double* t, const double* r, double theta, double abs_tol, int precision, int* n_sigma) {
        const int m = gmres_iter;
        std::vector<double> H(size_t(m + 1) * (m + 1), 0.0), sv(m + 1, 0.0), cs(m + 1, 0.0), sn(m + 1, 0.0);
        auto h = [&](int i, int j) -> double& { return H[size_t(i) + size_t(j) * (m + 1)]; };
        jcd_op(t, Gv(0), theta, precision, n_sigma);
        // v0 = b - A t = -r - A t
        scal(-1.0, Gv(0));
        axpy(-1.0, r, Gv(0));
        sv[0] = std::sqrt(dot(Gv(0), Gv(0)));
        if (std::fabs(sv[0]) < abs_tol) return;
        scal(1.0 / sv[0], Gv(0));
        auto rot = [](double& dx, double& dy, double c, double sn_) {
            const double r0 = c * dx + sn_ * dy, r1 = -sn_ * dx + c * dy;
            dx = r0, dy = r1;
        };
        int i = 0;
        for (; i < m - 1; ++i) {
            jcd_op(Gv(i), Gw(), theta, precision, n_sigma);
            for (int k = 0; k <= i; ++k) {
                h(k, i) = dot(Gw(), Gv(k));
                axpy(-h(k, i), Gv(k), Gw());
            }
            h(i + 1, i) = std::sqrt(dot(Gw(), Gw()));
            copy(Gw(), Gv(i + 1));
            scal(1.0 / h(i + 1, i), Gv(i + 1));
            for (int k = 0; k < i; ++k) rot(h(k, i), h(k + 1, i), cs[k], sn[k]);
            const double dx = h(i, i), dy = h(i + 1, i);
            if (dy == 0.0) {
                cs[i] = 1, sn[i] = 0;
            } else if (std::fabs(dy) > std::fabs(dx)) {
                const double tq = dx / dy;
                sn[i] = 1.0 / std::sqrt(1.0 + tq * tq);
                cs[i] = tq * sn[i];
            } else {
                const double tq = dy / dx;
                cs[i] = 1.0 / std::sqrt(1.0 + tq * tq);
                sn[i] = tq * cs[i];
            }
            rot(h(i, i), h(i + 1, i), cs[i], sn[i]);
            rot(sv[i], sv[i + 1], cs[i], sn[i]);
            if (std::fabs(sv[i + 1]) < abs_tol) break;
        }
        // back substitution over the first i columns, t += sum_k y_k v_k
        std::vector<double> y(sv.begin(), sv.begin() + i);
        for (int a = i - 1; a >= 0; --a) {
            y[a] /= h(a, a);
            for (int b = a - 1; b >= 0; --b) y[b] -= h(b, a) * y[a];
        }
        for (int k = 0; k < i; ++k) axpy(y[k], Gv(k), t);
    }
    double dot(const double* a, const double* b) {
        double r;
        int rc = mq_vec_dot(ctx, n, a, b, &r);
        if (rc) fail(rc, g_err);
        return r;
    }
    void axpy(double a, const double* x, double* y) { HIPCHK(launch_axpy(ctx->stream, n, a, x, y)); }
    void scal(double a, double* x) { HIPCHK(launch_scal(ctx->stream, n, a, x)); }
    void copy(const double* x, double* y) {
        if (x != y) HIPCHK(hipMemcpyAsync(y, x, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    }
    void sigma(const double* x, double* y, int precision) {
        int rc = mq_sigma_apply_ex(p, x, y, precision);
        if (rc) fail(rc, g_err);
    }

    // One Jacobi-Davidson run for the smallest eigenpair (ietl/jacobi.h:386-474) with sigma in
    // `precision`.  If init_A is given it is the product H*init, used for the first search vector
    // in place of a sigma call.  Stops when ||r|| <= max(tol |theta|, stop_abs) or ||r|| < tol
    // (ietl/iteration.h:56-58), or after max_iter search vectors.  On return u() holds the Ritz
    // vector, uA() its residual uA - theta u; returns theta.
    double run(const double* init, const double* init_A, double tol, double stop_abs, int precision, int* n_sigma,
               int* iters, double* rnorm) {
        std::vector<double> M(size_t(max_iter) * max_iter, 0.0), s;
        const double kappa = 0.25;
        copy(init, V(0));
        project(V(0));
        int it = 0;
        double theta = 0, rn = 0;
        while (true) {
            double* t = V(it);
            // Gram-Schmidt with refinement (ietl/jacobi.h:408-415) in its classical form: the overlaps
            // of a pass come from one batched reduction together with t.t (V(it) is t itself), and
            // the projection uses them on the device.  A second pass runs under the reference's
            // condition ||t|| < kappa tau.
            const double* r = mdot(V(0), it + 1, t);
            const double tau = std::sqrt(r[it]);
            double nt2 = r[it];
            if (it > 0) {
                subtract_last(V(0), it, t);
                r = mdot(V(0), it + 1, t);
                nt2 = r[it];
                if (std::sqrt(nt2) < kappa * tau) {
                    subtract_last(V(0), it, t);
                    nt2 = mdot(V(0), it + 1, t)[it];
                }
            }
            if (n_ortho) {   // project out the ortho space (ietl/jacobi.h:417-418)
                project(t);
                nt2 = dot(t, t);
            }
            double nt = std::sqrt(nt2);
            MQ_REQUIRE(nt > 0, "Jacobi-Davidson: zero search vector");
            scal(1.0 / nt, t);
            if (it == 0 && init_A) {
                copy(init_A, VA(0));
                scal(1.0 / nt, VA(0));
            } else {
                sigma(t, VA(it), precision);
                ++*n_sigma;
            }
            {
                const double* mc = mdot(V(0), it + 1, VA(it));   // M(i, it) = V(i) . VA(it), i <= it
                for (int i = 0; i <= it; ++i) M[i + size_t(it) * max_iter] = mc[i];
            }
            // with an exact first product, the coupling of the first vector to the new one is taken
            // from it: V(it).(H V(0)) is exact where V(0).(H~ V(it)) carries the reduced-precision
            // error, and it is this coupling that sets the size of the update of V(0)
            if (init_A && it > 0) M[size_t(it) * max_iter] = dot(V(it), VA(0));
            theta = smallest_eigenpair(M, max_iter, it + 1, s);
            combine(V(0), it + 1, s.data(), u());     // u = V s
            combine(VA(0), it + 1, s.data(), uA());   // uA = VA s
            project(uA());
            axpy(-theta, u(), uA());   // r = uA - theta u
            ++it;
            // u and r are consecutive vectors: [u . r, r . r] in one reduction
            const double* ur = mdot(u(), 2, uA());
            const double dru = ur[0];
            rn = std::sqrt(ur[1]);
            bool conv = (rn <= std::max(tol * std::fabs(theta), stop_abs) || rn < tol);
            if (conv || it >= max_iter) break;
            // t = -r + (r.u)/(u.u) u, then GMRES on (1 - uu^T)(H - theta)(1 - uu^T) t = -r to the
            // tolerance 2^-(it+1) when gmres_iter > 0 (jcd_gmres_solver, ietl/jacobi.h:141-157, 466-468)
            const double duu = dot(u(), u());
            double* tn = V(it);
            copy(uA(), tn);
            scal(-1.0, tn);
            axpy(dru / duu, u(), tn);
            if (gmres_iter > 0) gmres(tn, uA(), theta, 1.0 / std::pow(2.0, double(it + 1)), precision, n_sigma);
        }
        if (iters) *iters = it;
        if (rnorm) *rnorm = rn;
        return theta;
    }
};

}  // namespace

extern "C" {

int mq_jd_solve(mq_plan* p, const double* init, double* out, double tol, int max_iter, int gmres_iter,
                double* energy, int* iterations) {
    return mq_jd_solve_ex(p, init, out, tol, max_iter, gmres_iter, 0, nullptr, energy, iterations);
}

int mq_jd_solve_ex(mq_plan* p, const double* init, double* out, double tol, int max_iter, int gmres_iter, int n_ortho,
                   const double* const* ortho, double* energy, int* iterations) {
    return guarded([&] {
        MQ_REQUIRE(p && init && out && energy && max_iter > 0, "bad argument");
        MQ_REQUIRE(gmres_iter >= 0 && gmres_iter <= 64, "gmres_iter must lie in [0, 64]");
        MQ_REQUIRE(n_ortho >= 0 && (n_ortho == 0 || ortho), "bad ortho vectors");
        for (int j = 0; j < n_ortho; ++j) MQ_REQUIRE(ortho[j], "null ortho vector");
        // as the reference, no more ortho vectors than the space can hold (ietl_jacobi_davidson.h:41-42)
        if (uint64_t(n_ortho) >= p->vec_size) n_ortho = int(p->vec_size) - 1;
        JdSpace J(p, max_iter, gmres_iter, n_ortho);
        if (n_ortho) J.set_ortho(ortho);
        int ns = 0;
        double theta = J.run(init, nullptr, tol, 0.0, MQ_PREC_F64, &ns, iterations, nullptr);
        J.copy(J.u(), out);
        HIPCHK(hipStreamSynchronize(p->ctx->stream));
        *energy = theta;
    });
}

int mq_jd_solve_refined(mq_plan* p, const double* init, double* out, double tol, int max_iter, int precision,
                        int max_outer, double* energy, int* counts, double* residual) {
    return guarded([&] {
        MQ_REQUIRE(p && init && out && energy && max_iter > 0 && max_outer > 0, "bad argument");
        MQ_REQUIRE(precision == MQ_PREC_F64 || precision == MQ_PREC_F32 || precision == MQ_PREC_BF16X3,
                   "unknown precision");
        JdSpace J(p, max_iter);
        int n_red = 0, n64 = 0, outer = 0;
        double theta = 0, rn = 0;
        const double* x_A = nullptr;
        J.copy(init, J.x());
        while (outer < max_outer) {
            ++outer;
            // inner run: stop once its residual is a decade below the last FP64 one; the FP64 step
            // then measures the true residual (the reduced-precision products put a floor under the
            // inner residual, below which inner steps gain nothing)
            double stop_abs = (outer > 1) ? 1e-1 * rn : 0.0;
            J.run(J.x(), x_A, tol, stop_abs, precision, &n_red, nullptr, nullptr);
            // FP64 refinement step: exact product, Rayleigh quotient and residual of the Ritz vector
            J.copy(J.u(), J.x());
            double nx = std::sqrt(J.dot(J.x(), J.x()));
            J.scal(1.0 / nx, J.x());
            J.sigma(J.x(), J.xA(), MQ_PREC_F64);
            ++n64;
            x_A = J.xA();
            theta = J.dot(J.x(), J.xA());
            J.copy(J.xA(), J.uA());
            J.axpy(-theta, J.x(), J.uA());
            rn = std::sqrt(J.dot(J.uA(), J.uA()));
            if (rn <= tol * std::fabs(theta) || rn < tol) break;
        }
        J.copy(J.x(), out);
        HIPCHK(hipStreamSynchronize(p->ctx->stream));
        *energy = theta;
        if (counts) counts[0] = n_red, counts[1] = n64, counts[2] = outer;
        if (residual) *residual = rn;
    });
}

}  // extern "C"
