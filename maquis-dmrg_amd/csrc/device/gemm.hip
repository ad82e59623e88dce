// Grouped, K-segmented GEMM on gfx950 MFMA, templated on operand storage types, compute mode and
// tile shape.
//
//   MODE_F64    v_mfma_f64_16x16x4f64, FP64 operands and accumulation (the plain path)
//   MODE_F32    v_mfma_f32_16x16x4f32, FP32 operands and accumulation (the reduced-precision
//               paths: their T and final stages are bandwidth-bound, and the BF16x3 products
//               of the S stage are plain dense GEMMs done by the library)
//
// Operands are read in their storage type (TA, TB: double or float) and converted when staged;
// C is written in TC.  One launch covers every tile of every task of one tile configuration (the
// reference issues one cuBLAS call per block product, solver/tasks.cpp:226-229, 277-280,
// 335-336, 630-639, 738-748).  The symmetry blocks of a contraction range from a few rows to
// thousands, so each task is run with the tile configuration that fits its (M, N):
//
//   GEMM_CFG_128x128  4 waves (2x2), each 64x64  = 4x4 MFMA blocks   large blocks
//   GEMM_CFG_64x64    4 waves (2x2), each 32x32  = 2x2 MFMA blocks
//   GEMM_CFG_32x128   4 waves (1x4), each 32x32                      M <= 32
//   GEMM_CFG_16x128   4 waves (1x4), each 16x32  = 1x2 MFMA blocks   M <= 16
//   GEMM_CFG_16x16    4 waves split K of one 16x16 tile, operands read straight from memory
//                     in the MFMA lane layout, partial sums reduced through LDS    M, N <= 16
//
// MFMA lane layouts on gfx950 (lane l = 0..63, lk = l/16), measured with one-hot operands:
//   f64 16x16x4 : A row l%16, k lk;  B col l%16, k lk;  D[r] row 4r + lk, col l%16
//   f32 16x16x4 : A row l%16, k lk;  B col l%16, k lk;  D[r] row 4lk + r, col l%16
#include <hip/hip_runtime.h>

#include <type_traits>

#include "kernels.hpp"

namespace mq {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <int MODE> struct ModeTraits;
template <> struct ModeTraits<MODE_F64> { typedef double T; typedef d4 Acc; static constexpr int BK = 16, KS = 4; };
template <> struct ModeTraits<MODE_F32> { typedef float T; typedef f4 Acc; static constexpr int BK = 16, KS = 4; };

// output row of accumulator register r for lanes of k-group lk = l/16
template <int MODE> __device__ inline int d_row(int r, int lk) { return MODE == MODE_F64 ? 4 * r + lk : 4 * lk + r; }

__device__ inline d4 mfma_step(double a, double b, d4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ inline f4 mfma_step(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
// ------------------------------------------------------------------------------ operand access
// Element loads are branch-free: an out-of-range element reads offset 0 of its operand (the first
// element of a stored block: a valid address holding finite data) and is scaled by zero.  The value
// is always used, so the compiler keeps the load unconditional and the loads of a step stay in
// flight together instead of each sitting in a divergent branch behind a full wait.
template <typename TA>
__device__ inline double a_load(const TA* A, const GemmSeg& sg, bool ok, uint32_t m, uint32_t k) {
    const uint64_t sm = sg.trans_a ? sg.lda : 1u, sk = sg.trans_a ? 1u : sg.lda;
    const double v = double(A[(uint64_t(m) * sm + uint64_t(k) * sk) & (0ull - uint64_t(ok))]);
    return v * (ok ? sg.alpha : 0.0);
}
template <typename TB>
__device__ inline double b_load(const TB* B, const GemmSeg& sg, const uint32_t* maps, bool ok, uint32_t k,
                                uint32_t n) {
    uint64_t off;
    if (sg.bmap != MAP_NONE) {
        const uint32_t blk = min(k, sg.K - 1) / sg.bbh;
        const uint64_t row = maps[sg.bmap + blk];
        off = row * sg.bldk + (k - blk * sg.bbh) + uint64_t(n) * sg.ldb;
    } else {
        const uint64_t sk = sg.trans_b ? sg.ldb : 1u, sn = sg.trans_b ? 1u : sg.ldb;
        off = uint64_t(k) * sk + uint64_t(n) * sn;
    }
    const double v = double(B[off & (0ull - uint64_t(ok))]);
    return v * (ok ? 1.0 : 0.0);
}

template <typename TC>
__device__ inline TC* c_column(const GemmTask& tk, const Bases& bases, const uint32_t* maps, uint32_t gn) {
    TC* C = reinterpret_cast<TC*>(bases.p[tk.c_base]) + tk.c_off;
    if (tk.cmap != MAP_NONE) {
        const uint32_t blk = gn / tk.cbw;
        return C + uint64_t(maps[tk.cmap + blk]) * tk.cldk + uint64_t(gn - blk * tk.cbw) * tk.ldc;
    }
    return C + uint64_t(gn) * tk.ldc;
}


// Task of a tile: a direct table where the launch has one (many small tasks), else the identity
// (one tile per task) or a binary search over the tile prefixes.
__device__ inline uint32_t find_task(const GemmTask* tasks, uint32_t n_tasks, const uint32_t* tile_task, uint32_t tile) {
    if (tile_task) return tile_task[tile];
    if (n_tasks == gridDim.x) return tile;
    uint32_t lo = 0, hi = n_tasks;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (tasks[mid].tile_begin <= tile) lo = mid; else hi = mid;
    }
    return lo;
}

// ------------------------------------------------------------------------------ LDS-tiled kernel
// Software pipelined: the K steps of all segments form one sequence; while the MFMAs consume
// step t from one LDS buffer, the global loads of step t+1 are in flight into registers, which
// are then converted and stored into the other buffer (one barrier per step).
template <int MODE, int BM, int BN>
struct Smem {
    typedef typename ModeTraits<MODE>::T T;
    static constexpr int BK = ModeTraits<MODE>::BK;
    static constexpr size_t buf = (sizeof(T) * size_t(BK) * (BM + 1 + BN + 1) + 15) / 16 * 16;
};

template <typename TA, typename TB, typename TC, int MODE, int PM, int PN, int WM, int WN>
__global__ __launch_bounds__(GEMM_THREADS) void gemm_kernel(const GemmTask* __restrict__ tasks, uint32_t n_tasks,
                                                            const GemmSeg* __restrict__ segs, Bases bases,
                                                            GemmTables tables,
                                                            const uint32_t* __restrict__ tile_task) {
    const uint32_t* __restrict__ maps = tables.maps;
    const GemmUnit* __restrict__ units = tables.units;
    const GemmRow* __restrict__ rows = tables.rows;
    const double* __restrict__ scales = tables.scales;
    static_assert(PM * PN * 64 == GEMM_THREADS, "four waves");
    constexpr bool SPLIT = std::is_same<TC, SplitBf16>::value;
    typedef typename std::conditional<SPLIT, uint16_t, TC>::type CT;   // C element storage
    typedef typename ModeTraits<MODE>::T T;
    typedef typename ModeTraits<MODE>::Acc Acc;
    constexpr int BK = ModeTraits<MODE>::BK;
    constexpr int BM = 16 * WM * PM, BN = 16 * WN * PN;
    constexpr int NA = BK * BM / GEMM_THREADS, NB = BK * BN / GEMM_THREADS;
    static_assert(NA * GEMM_THREADS == BK * BM && NB * GEMM_THREADS == BK * BN, "whole staging passes");
    typedef Smem<MODE, BM, BN> SM;
    __shared__ __align__(16) char smem[2 * SM::buf];

    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const GemmTask tk = tasks[find_task(tasks, n_tasks, tile_task, tile)];
    const uint32_t local = tile - tk.tile_begin;
    const bool stacked = tk.rowunit != MAP_NONE;
    const uint32_t M = tk.M, N = tk.N;
    const uint32_t tiles_m = (M + BM - 1) / BM;
    const uint32_t m0 = (stacked ? local % tiles_m : local / tk.tiles_n) * BM;
    const uint32_t n0 = (stacked ? local / tiles_m : local % tk.tiles_n) * BN;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = (wave % PM) * 16 * WM, wn = (wave / PM) * 16 * WN;
    const int l16 = lane & 15, lk = lane >> 4;

    Acc acc[WM][WN];
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j) acc[i][j] = Acc{0, 0, 0, 0};

    // stacked task: the A row of this thread (m = tid % BM for every staged element of a
    // non-transposed A) and its unit's column stride
    uint64_t a_row = 0;
    uint32_t a_ld = 0, a_ok = 0, a_i = 0, a_j = 0;
    double a_rscale = 1.0;
    if (stacked) {
        const uint32_t m = m0 + uint32_t(tid % BM);
        const GemmRow R = rows[maps[tk.rowunit + min(m, M - 1) / 16]];
        const uint32_t i = m - R.row0;
        a_ok = uint32_t(m < M) & uint32_t(i < R.nrows);
        a_row = R.a_src + uint64_t(i) * R.a_si;
        a_ld = R.a_sk;
        a_rscale = R.a_scale;
        a_i = i;
        a_j = R.jblk;
    }
    // cursor over the K steps of all segments: (sgi, k0) with k0 < segs[sgi].K
    uint32_t sgi = tk.seg0, k0 = 0;
    GemmSeg sg;
    auto settle = [&]() {
        while (sgi < tk.seg1) {
            sg = segs[sgi];
            if (k0 < sg.K) return true;
            ++sgi;
            k0 = 0;
        }
        return false;
    };
    // staged operands of one step, in registers as loaded (storage type), with their in-range masks
    // and the index order they were loaded in.  Nothing is computed on the loaded values until
    // lstore, after the MFMAs of the current step: the loads stay in flight across them.
    TA va[NA];
    TB vb[NB];
    uint32_t oka = 0, okb = 0;
    double r_alpha = 1.0;
    bool r_trans_a = false, r_kfast_b = true;
    double a_scale = 1.0;   // block scale of a remapped transposed A (all of a thread's elements share k)
    auto load_a = [&](const TA* __restrict__ A) {
        if (sg.amap != MAP_NONE && !stacked) {
            const uint32_t gk = k0 + uint32_t(tid % BK);   // trans_a: k = e % BK = tid % BK
            const uint32_t blk = min(gk, sg.K - 1) / sg.bbh;
            const uint64_t boff = uint64_t(maps[sg.amap + blk]) + (gk - blk * sg.bbh);
            a_scale = scales[sg.ascale + blk];
#pragma unroll
            for (int i = 0; i < NA; ++i) {
                const int e = tid + GEMM_THREADS * i;
                const uint32_t gm = m0 + e / BK;
                const uint32_t ok = uint32_t(gm < M) & uint32_t(gk < sg.K);
                oka |= uint32_t(gk < sg.K) << i;
                va[i] = A[(boff + uint64_t(gm + tk.m_off) * sg.lda) & (0ull - uint64_t(ok))];
            }
            return;
        }
        a_scale = 1.0;
        if (stacked) {
            a_scale = a_rscale;
            // this thread's row: unit strides, or the segment's block of the unit (A block map)
            uint64_t row = a_row, ldk = a_ld;
            if (sg.amap != MAP_NONE) {
                row = uint64_t(maps[sg.amap + a_j]) * sg.bldk + uint64_t(a_i) * (sg.trans_a ? sg.lda : 1u);
                ldk = sg.trans_a ? 1u : sg.lda;
            }
#pragma unroll
            for (int i = 0; i < NA; ++i) {
                const int e = tid + GEMM_THREADS * i;
                const uint32_t gk = k0 + e / BM;
                const uint32_t ok = a_ok & uint32_t(gk < sg.K);
                oka |= uint32_t(gk < sg.K) << i;
                va[i] = A[(row + uint64_t(gk) * ldk) & (0ull - uint64_t(ok))];
            }
            return;
        }
        const uint64_t sm = sg.trans_a ? sg.lda : 1u, sk = sg.trans_a ? 1u : sg.lda;
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            const int e = tid + GEMM_THREADS * i;
            const int m = r_trans_a ? e / BK : e % BM;
            const int k = r_trans_a ? e % BK : e / BM;
            const uint32_t gm = m0 + m, gk = k0 + k;
            const uint32_t ok = uint32_t(gm < M) & uint32_t(gk < sg.K);
            oka |= uint32_t(gk < sg.K) << i;
            va[i] = A[(uint64_t(gm + tk.m_off) * sm + uint64_t(gk) * sk) & (0ull - uint64_t(ok))];
        }
    };
    // column units: the B source of this thread's columns for the current segment (the columns of
    // a thread's elements, n0 + (tid + 256 i) / BK, are fixed across the K steps)
    uint64_t u_src[NB];
    uint32_t u_ld[NB], u_map[NB];
    uint32_t u_ok = 0, u_sgi = 0xffffffffu;
    auto resolve_units = [&]() {
        u_ok = 0;
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int e = tid + GEMM_THREADS * i;
            const uint32_t n = n0 + e / BK;
            const uint32_t nc = min(n, N - 1);
            const GemmUnit U = units[sg.units + maps[tk.colunit + nc]];
            const uint32_t ok = uint32_t(n < N) & uint32_t(U.valid != 0);
            u_ok |= ok << i;
            u_src[i] = U.src + uint64_t(sg.bbh) * (nc - U.col0);
            u_ld[i] = U.ld;
            u_map[i] = ok ? U.map : 0u;
        }
        u_sgi = sgi;
    };
    auto gload = [&]() {
        const TA* __restrict__ A = reinterpret_cast<const TA*>(bases.p[sg.a_base]) + sg.a_off;
        const TB* __restrict__ B = reinterpret_cast<const TB*>(bases.p[sg.b_base]) + sg.b_off;
        r_trans_a = sg.trans_a && !stacked;   // stacked tasks stage A rows m-fast in every case
        r_kfast_b = sg.bmap != MAP_NONE || sg.units != MAP_NONE || !sg.trans_b;
        r_alpha = sg.alpha;
        oka = okb = 0;
        if (sg.units != MAP_NONE) {
            // every element of this thread has k = k0 + tid % BK: one block row for all of them
            if (u_sgi != sgi) resolve_units();
            const uint32_t gk = k0 + uint32_t(tid % BK);
            const uint32_t okk = uint32_t(gk < sg.K);
            const uint32_t blk = min(gk, sg.K - 1) / sg.bbh;
            const uint32_t kin = gk - blk * sg.bbh;
            uint32_t row[NB];
#pragma unroll
            for (int i = 0; i < NB; ++i) row[i] = maps[(u_map[i] + blk) & (0u - ((u_ok >> i) & 1u))];
            load_a(A);
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                const uint32_t ok = ((u_ok >> i) & 1u) & okk;
                okb |= ok << i;
                vb[i] = B[(u_src[i] + uint64_t(row[i]) * u_ld[i] + kin) & (0ull - uint64_t(ok))];
            }
        } else if (sg.bmap != MAP_NONE) {
            // row-block remapped B: the map entries of all elements first, then the A loads, then
            // the B loads (only these wait for the map entries)
            uint32_t row[NB], kin[NB];
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                const int e = tid + GEMM_THREADS * i;
                const uint32_t gk = k0 + e % BK;
                const uint32_t blk = min(gk, sg.K - 1) / sg.bbh;
                kin[i] = gk - blk * sg.bbh;
                row[i] = maps[sg.bmap + blk];
            }
            load_a(A);
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                const int e = tid + GEMM_THREADS * i;
                const uint32_t gn = n0 + e / BK, gk = k0 + e % BK;
                const uint32_t ok = uint32_t(gn < N) & uint32_t(gk < sg.K);
                okb |= uint32_t(gk < sg.K) << i;
                vb[i] = B[(uint64_t(row[i]) * sg.bldk + kin[i] + uint64_t(gn) * sg.ldb) & (0ull - uint64_t(ok))];
            }
        } else {
            const uint64_t bsk = sg.trans_b ? sg.ldb : 1u, bsn = sg.trans_b ? 1u : sg.ldb;
            load_a(A);
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                const int e = tid + GEMM_THREADS * i;
                const int n = r_kfast_b ? e / BK : e % BN;
                const int k = r_kfast_b ? e % BK : e / BN;
                const uint32_t gn = n0 + n, gk = k0 + k;
                const uint32_t ok = uint32_t(gn < N) & uint32_t(gk < sg.K);
                okb |= uint32_t(gk < sg.K) << i;
                vb[i] = B[(uint64_t(gk) * bsk + uint64_t(gn) * bsn) & (0ull - uint64_t(ok))];
            }
        }
    };
    // Staged element i of thread tid is e = tid + 256 i.  m-fast A: m = e % BM = tid % BM for every i
    // (BM divides 256), k = tid / BM + i (256 / BM); k-fast A: k = tid % BK, m = tid / BK + 16 i; B
    // likewise with n and BN.  Both give LDS offsets base + i * constant, so the two orders are two
    // uniform branches with immediate offsets.  Values are masked only where they must read as zero
    // (the mask bits: k beyond the segment, B of an invalid column unit); out-of-range rows and
    // columns carry the finite value at their clamped address into outputs that are not stored.
    static_assert(GEMM_THREADS % BM == 0 && GEMM_THREADS % BN == 0 && GEMM_THREADS % BK == 0, "staging maps");
    auto lstore = [&](int buf) {
        char* base = smem + buf * SM::buf;
        T* As = reinterpret_cast<T*>(base);
        T* Bs = reinterpret_cast<T*>(base + sizeof(T) * BK * (BM + 1));
        const T sa = T(r_alpha * a_scale);
        const bool mask_a = oka != (1u << NA) - 1u, mask_b = okb != (1u << NB) - 1u;
        T xa[NA], xb[NB];
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            xa[i] = T(va[i]) * sa;
            if (mask_a && !((oka >> i) & 1u)) xa[i] = T(0);
        }
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            xb[i] = T(vb[i]);
            if (mask_b && !((okb >> i) & 1u)) xb[i] = T(0);
        }
        if (r_trans_a) {
            T* p = As + (tid % BK) * (BM + 1) + tid / BK;
#pragma unroll
            for (int i = 0; i < NA; ++i) p[i * (GEMM_THREADS / BK)] = xa[i];
        } else {
            T* p = As + (tid / BM) * (BM + 1) + tid % BM;
#pragma unroll
            for (int i = 0; i < NA; ++i) p[i * (GEMM_THREADS / BM) * (BM + 1)] = xa[i];
        }
        if (r_kfast_b) {
            T* p = Bs + (tid % BK) * (BN + 1) + tid / BK;
#pragma unroll
            for (int i = 0; i < NB; ++i) p[i * (GEMM_THREADS / BK)] = xb[i];
        } else {
            T* p = Bs + (tid / BN) * (BN + 1) + tid % BN;
#pragma unroll
            for (int i = 0; i < NB; ++i) p[i * (GEMM_THREADS / BN) * (BN + 1)] = xb[i];
        }
    };
    auto compute = [&](int buf) {
        const char* base = smem + buf * SM::buf;
        const T(*As)[BM + 1] = reinterpret_cast<const T(*)[BM + 1]>(base);
        const T(*Bs)[BN + 1] = reinterpret_cast<const T(*)[BN + 1]>(base + sizeof(T) * BK * (BM + 1));
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            const int kr = kk + lk;
            T a[WM], b[WN];
#pragma unroll
            for (int i = 0; i < WM; ++i) a[i] = As[kr][wm + 16 * i + l16];
#pragma unroll
            for (int j = 0; j < WN; ++j) b[j] = Bs[kr][wn + 16 * j + l16];
#pragma unroll
            for (int i = 0; i < WM; ++i)
#pragma unroll
                for (int j = 0; j < WN; ++j) acc[i][j] = mfma_step(a[i], b[j], acc[i][j]);
        }
    };

    if (settle()) {
        gload();
        lstore(0);
        __syncthreads();
        int buf = 0;
        while (true) {
            k0 += BK;
            const bool more = settle();
            if (more) gload();   // next step's loads overlap this step's MFMAs
            compute(buf);
            if (!more) break;
            lstore(buf ^ 1);
            __syncthreads();
            buf ^= 1;
        }
    }

    if (stacked) {
        // row units: each 16-row block of this wave lies in one unit, with its own C block and
        // column remap.  The unit ids, then the descriptors, then the column offsets are loaded
        // for all blocks before the first store, so the lookups overlap.
        CT* __restrict__ C = reinterpret_cast<CT*>(bases.p[tk.c_base]) + tk.c_off;
        uint32_t uid[WM];
#pragma unroll
        for (int i = 0; i < WM; ++i) uid[i] = maps[tk.rowunit + min(m0 + wm + 16 * i, M - 1) / 16];
        GemmRow R[WM];
#pragma unroll
        for (int i = 0; i < WM; ++i) R[i] = rows[uid[i]];
#pragma unroll
        for (int i = 0; i < WM; ++i) {
            uint64_t coff[WN];
#pragma unroll
            for (int j = 0; j < WN; ++j) {
                const uint32_t gn = min(n0 + wn + 16 * j + l16, N - 1);
                if (R[i].cmap != MAP_NONE) {
                    const uint32_t blk = gn / tk.cbw;
                    coff[j] = uint64_t(maps[R[i].cmap + blk]) * R[i].cldk + uint64_t(gn - blk * tk.cbw) * R[i].ldc;
                } else {
                    coff[j] = uint64_t(gn) * R[i].ldc;
                }
            }
            // FP32 accumulators: registers r = 0..3 of a lane are 4 consecutive rows of one column.
            // With all four inside the unit, each output copy is one 8-byte (bf16) or 16-byte
            // (float) store, so a wave writes whole 32/64-byte row runs per column instead of
            // scattered 2/4-byte pieces (global memory accepts element-aligned vector stores)
            if constexpr (MODE == MODE_F32 && (SPLIT || std::is_same<CT, float>::value)) {
                const uint32_t gm0 = m0 + wm + 16 * i + 4 * lk;
                const uint32_t li0 = gm0 - R[i].row0;
                if (!tk.beta && R[i].c_si == 1 && gm0 + 3 < M && li0 + 3 < R[i].nrows) {
#pragma unroll
                    for (int j = 0; j < WN; ++j) {
                        if (n0 + wn + 16 * j + l16 >= N) continue;
                        if constexpr (SPLIT) {
                            typedef uint16_t u16x4 __attribute__((ext_vector_type(4)));
                            typedef u16x4 __attribute__((aligned(2))) u16x4_a2;
                            u16x4 hi, lo;
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const float v = float(acc[i][j][r]);
                                hi[r] = bf16_rne(v);
                                lo[r] = bf16_rne(v - __uint_as_float(uint32_t(hi[r]) << 16));
                            }
                            uint16_t* __restrict__ Cm = C + R[i].c_split + li0 + coff[j];
                            const uint64_t sd = R[i].split_stride;
                            *reinterpret_cast<u16x4_a2*>(Cm) = hi;
                            *reinterpret_cast<u16x4_a2*>(Cm + sd) = lo;
                            *reinterpret_cast<u16x4_a2*>(Cm + 2 * sd) = hi;
                        } else {
                            typedef float __attribute__((ext_vector_type(4), aligned(4))) f4_a4;
                            *reinterpret_cast<f4_a4*>(C + R[i].c_src + li0 + coff[j]) =
                                f4_a4{float(acc[i][j][0]), float(acc[i][j][1]), float(acc[i][j][2]), float(acc[i][j][3])};
                        }
                    }
                    continue;
                }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t gm = m0 + wm + 16 * i + d_row<MODE>(r, lk);
                const uint32_t li = gm - R[i].row0;
                if (gm >= M || li >= R[i].nrows) continue;
                if constexpr (SPLIT) {
                    uint16_t* __restrict__ Cm = C + R[i].c_split + li;
                    const uint64_t sd = R[i].split_stride;
#pragma unroll
                    for (int j = 0; j < WN; ++j) {
                        if (n0 + wn + 16 * j + l16 >= N) continue;
                        const float v = float(acc[i][j][r]);
                        const uint16_t hi = bf16_rne(v);
                        const uint16_t lo = bf16_rne(v - __uint_as_float(uint32_t(hi) << 16));
                        Cm[coff[j]] = hi;
                        Cm[coff[j] + sd] = lo;
                        Cm[coff[j] + 2 * sd] = hi;
                    }
                } else {
                    CT* __restrict__ Cm = C + R[i].c_src + uint64_t(li) * R[i].c_si;
#pragma unroll
                    for (int j = 0; j < WN; ++j) {
                        if (n0 + wn + 16 * j + l16 >= N) continue;
                        CT v = CT(acc[i][j][r]);
                        if (tk.beta) v += Cm[coff[j]];
                        Cm[coff[j]] = v;
                    }
                }
            }
        }
        return;
    }
    if constexpr (SPLIT) return;   // split output: stacked tasks only
    else {
#pragma unroll
    for (int j = 0; j < WN; ++j) {
        const uint32_t gn = n0 + wn + 16 * j + l16;
        if (gn >= N) continue;
        TC* __restrict__ Cn = c_column<TC>(tk, bases, maps, gn);
#pragma unroll
        for (int i = 0; i < WM; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t gm = m0 + wm + 16 * i + d_row<MODE>(r, lk);
                if (gm < M) {
                    TC v = TC(acc[i][j][r]);
                    if (tk.beta) v += Cn[gm + tk.m_off];
                    Cn[gm + tk.m_off] = v;
                }
            }
    }
    }
}

// ------------------------------------------------------------------------------ split-K 16x16 kernel
// One workgroup per task (M, N <= 16).  The K steps of all segments are dealt round-robin to the
// four waves; each wave reads its operand fragments directly from memory in the MFMA lane layout.
template <typename TA, typename TB, typename TC, int MODE>
__global__ __launch_bounds__(GEMM_THREADS) void gemm_tiny_kernel(const GemmTask* __restrict__ tasks, uint32_t n_tasks,
                                                                 const GemmSeg* __restrict__ segs, Bases bases,
                                                                 GemmTables tables,
                                                                 const uint32_t* __restrict__ tile_task) {
    const uint32_t* __restrict__ maps = tables.maps;
    typedef typename ModeTraits<MODE>::T T;
    typedef typename ModeTraits<MODE>::Acc Acc;
    constexpr int KS = ModeTraits<MODE>::KS;
    constexpr int NW = GEMM_THREADS / 64;
    __shared__ T red[NW - 1][64][4];

    const GemmTask tk = tasks[find_task(tasks, n_tasks, tile_task, xcd_tile(blockIdx.x, gridDim.x))];
    const uint32_t M = tk.M, N = tk.N;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int l16 = lane & 15, lk = lane >> 4;
    const uint32_t row = l16, col = l16;

    Acc acc = {0, 0, 0, 0};
    uint32_t step = 0;   // K steps of the previous segments: deals continue across segments
    for (uint32_t sgi = tk.seg0; sgi < tk.seg1; ++sgi) {
        const GemmSeg sg = segs[sgi];
        const TA* __restrict__ A = reinterpret_cast<const TA*>(bases.p[sg.a_base]) + sg.a_off;
        const TB* __restrict__ B = reinterpret_cast<const TB*>(bases.p[sg.b_base]) + sg.b_off;
        const uint32_t K = sg.K;
        const uint32_t nsteps = (K + KS - 1) / KS;
        const uint32_t first = (uint32_t(wave) + NW - step % NW) % NW;
        // U steps per pass: all their loads are issued before the MFMAs consume them
        constexpr int U = 8;
        for (uint32_t s = first; s < nsteps; s += NW * U) {
            T a[U], b[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t k = (s + NW * u) * KS + lk;
                a[u] = T(a_load(A, sg, (row < M) & (k < K), row + tk.m_off, k));
                b[u] = T(b_load(B, sg, maps, (col < N) & (k < K), k, col));
            }
#pragma unroll
            for (int u = 0; u < U; ++u) acc = mfma_step(a[u], b[u], acc);
        }
        step += nsteps;
    }
    if (wave > 0)
#pragma unroll
        for (int r = 0; r < 4; ++r) red[wave - 1][lane][r] = acc[r];
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int w = 0; w < NW - 1; ++w)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[r] += red[w][lane][r];
        if (col < N) {
            TC* __restrict__ Cn = c_column<TC>(tk, bases, maps, col);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t gm = d_row<MODE>(r, lk);
                if (gm < M) {
                    TC v = TC(acc[r]);
                    if (tk.beta) v += Cn[gm + tk.m_off];
                    Cn[gm + tk.m_off] = v;
                }
            }
        }
    }
}

// ------------------------------------------------------------------------------ dispatch
template <typename TA, typename TB, typename TC, int MODE>
static hipError_t launch_typed(hipStream_t st, int cfg, const GemmTask* tasks, uint32_t n_tasks, const GemmSeg* segs,
                               uint32_t n_tiles, Bases bases, GemmTables tb, const uint32_t* tile_task) {
    const dim3 grid(n_tiles), block(GEMM_THREADS);
    switch (cfg) {
        case GEMM_CFG_128x128:
            hipLaunchKernelGGL((gemm_kernel<TA, TB, TC, MODE, 2, 2, 4, 4>), grid, block, 0, st, tasks, n_tasks, segs,
                               bases, tb, tile_task);
            break;
        case GEMM_CFG_64x64:
            hipLaunchKernelGGL((gemm_kernel<TA, TB, TC, MODE, 2, 2, 2, 2>), grid, block, 0, st, tasks, n_tasks, segs,
                               bases, tb, tile_task);
            break;
        case GEMM_CFG_32x128:
            hipLaunchKernelGGL((gemm_kernel<TA, TB, TC, MODE, 1, 4, 2, 2>), grid, block, 0, st, tasks, n_tasks, segs,
                               bases, tb, tile_task);
            break;
        case GEMM_CFG_16x128:
            hipLaunchKernelGGL((gemm_kernel<TA, TB, TC, MODE, 1, 4, 1, 2>), grid, block, 0, st, tasks, n_tasks, segs,
                               bases, tb, tile_task);
            break;
        case GEMM_CFG_16x16:
            if constexpr (std::is_same<TC, SplitBf16>::value) return hipErrorInvalidValue;
            else
                hipLaunchKernelGGL((gemm_tiny_kernel<TA, TB, TC, MODE>), grid, block, 0, st, tasks, n_tasks, segs,
                                   bases, tb, tile_task);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_gemm(hipStream_t st, int cfg, int mode, const char* types, const GemmTask* tasks, uint32_t n_tasks,
                       const GemmSeg* segs, uint32_t n_tiles, Bases bases, GemmTables tb, const uint32_t* tile_task) {
    if (n_tiles == 0) return hipSuccess;
    if (mode == MODE_F64)
        return launch_typed<double, double, double, MODE_F64>(st, cfg, tasks, n_tasks, segs, n_tiles, bases, tb, tile_task);
    const bool da = types[0] == 'd', db = types[1] == 'd', dc = types[2] == 'd', fc = types[2] == 'f';
    if (mode == MODE_F32) {
        if (da && db && fc)
            return launch_typed<double, double, float, MODE_F32>(st, cfg, tasks, n_tasks, segs, n_tiles, bases, tb, tile_task);
        if (!da && !db && fc)
            return launch_typed<float, float, float, MODE_F32>(st, cfg, tasks, n_tasks, segs, n_tiles, bases, tb, tile_task);
        if (da && !db && dc)
            return launch_typed<double, float, double, MODE_F32>(st, cfg, tasks, n_tasks, segs, n_tiles, bases, tb, tile_task);
        if (da && db && types[2] == 's' && cfg != GEMM_CFG_16x16)
            return launch_typed<double, double, SplitBf16, MODE_F32>(st, cfg, tasks, n_tasks, segs, n_tiles, bases, tb, tile_task);
    }
    return hipErrorInvalidValue;
}

}  // namespace mq
