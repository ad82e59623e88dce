// Device task tables and kernel launchers of the sigma / boundary engine (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace mq {

// Operand bases: every task addresses memory as (base selector, element offset), so one
// task table serves any psi / sigma pair handed to an apply call.
enum Base : uint8_t {
    BASE_WS = 0,      // plan / step workspace (T and S buffers)
    BASE_IN = 1,      // psi (sigma) or ket (boundary steps)
    BASE_OUT = 2,     // sigma (sigma) or new boundary data (steps)
    BASE_LEFT = 3,    // left boundary data
    BASE_RIGHT = 4,   // right boundary data
    BASE_BRA = 5,     // bra data (boundary steps)
    N_BASES = 6
};

// Base pointers of one launch.  They are carried as double*; a GEMM reads each operand in its
// own storage type (double or float, a template argument of the kernel).
struct Bases {
    double* p[N_BASES];
};

// C (M x N, ldc) = beta*C + sum_seg alpha_seg op(A_seg) (M x K_seg) * op(B_seg) (K_seg x N)
//
// Optional remaps (MQ_NONE = off):
//  column-block store remap: column n of C is column n % cbw of block j = n / cbw, stored at
//      c_off + maps[cmap + j] * cldk + (n % cbw) * ldc
//  row-block load remap of B (non-transposed B only): row k of B is row k % bbh of block
//      j = k / bbh, read at b_off + maps[bmap + j] * bldk + (k % bbh) + n * ldb
//  column units of B: output column n of a task with colunit != MAP_NONE belongs to unit
//      u = maps[colunit + n] (u < nunits); a segment with units != MAP_NONE reads, for unit u,
//      the descriptor U = unit table[units + u] and
//          B(k, n) = b_off + U.src + maps[U.map + k / bbh] * U.ld + (k % bbh) + bbh * (n - U.col0)
//      (zero if U.valid == 0).  Used where the output columns of one task gather their B blocks
//      from different workspace matrices.
struct GemmTask {
    uint64_t c_off;
    uint64_t cldk;
    uint32_t M, N, ldc;
    uint32_t seg0, seg1;
    uint32_t tiles_n;
    uint32_t tile_begin;
    uint32_t cmap, cbw;
    uint32_t colunit, nunits;
    uint32_t m_off;   // row offset of the task in op(A) and C (tasks sharing one segment range)
    uint32_t rowunit; // row units (MAP_NONE = off), below
    uint8_t c_base, beta, pad0, pad1;
};

// Row units of a stacked task (rowunit != MAP_NONE): several A blocks that multiply the same B are
// stacked into the rows of one task, so the B panel of a column tile is read once for all of them.
// Unit u occupies rows [R.row0, R.row0 + span) of the task, span = nrows rounded up to a multiple
// of 16 (row0 too is a multiple of 16, so every 16-row MFMA block lies in one unit); row m belongs
// to unit u = maps[rowunit + m / 16] (one entry per 16-row block), with descriptor R = row
// table[u] and local row i = m - R.row0:
//     A(m, k)  = R.a_scale * A[a_off + R.a_src + i * R.a_si + k * R.a_sk]   i < R.nrows, else 0
//     C(m, n)  = C[c_off + R.c_src + maps[R.cmap + n / cbw] * R.cldk + (n % cbw) * R.ldc + i * R.c_si]
// (R.cmap == MAP_NONE: C[c_off + R.c_src + n * R.ldc + i * R.c_si]); padding rows are not stored.
// A segment with amap != MAP_NONE addresses A per unit through its own block map instead: unit j
// = R.jblk reads block maps[amap + j] (stride bldk) of the segment's operand, op(A) with the
// segment's lda / trans_a,
//     A(m, k)  = R.a_scale * A[a_off + maps[amap + j] * bldk + (trans_a ? i * lda + k : i + k * lda)]
// so one task can sum segments whose blocks sit at per-unit, per-segment places.  The task's own
// cmap, ldc and m_off are unused.  Tiles of a stacked task are numbered row tile fastest: the row
// tiles that read one B panel run next to each other and share it in L2.
// Split output (C type SplitBf16, stacked tasks only): the element goes to three bf16 copies,
//     C[c_off + R.c_split + o] = hi,  C[... + o + R.split_stride] = lo,  C[... + o + 2 R.split_stride] = hi
// with o the offset above without c_src, hi = bf16(x), lo = bf16(x - hi).
struct GemmRow {
    uint64_t a_src, c_src, c_split, split_stride;
    double a_scale;
    uint32_t a_si, a_sk, row0, cldk, ldc, cmap, nrows, jblk, c_si, pad;
};

// C storage tag of the split output (three bf16 runs [hi | lo | hi]: the K-concatenated left
// operand of the one-GEMM BF16x3 product)
struct SplitBf16 {};

// x = hi + lo with hi = bf16(x), lo = bf16(x - hi), round to nearest even (finite inputs)
__device__ inline uint16_t bf16_rne(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

// Zeroing of scattered ranges: chunk c clears elements [off, off + len) of the base array
// (len <= ZERO_CHUNK).
struct ZeroChunk {
    uint64_t off;
    uint32_t len, pad;
};
constexpr uint32_t ZERO_CHUNK = 4096;

//  row-block remap of a transposed A (trans_a only): column k of op(A) lies in block j = k / bbh,
//      op(A)(m, k) = alpha * scales[ascale + j] * A[a_off + maps[amap + j] + (k % bbh) + m * lda]
//      (the transposed Hermitian partners of a cohort's blocks, each with its conjugation scale)
struct GemmSeg {
    uint64_t a_off, b_off;
    uint64_t bldk;
    double alpha;
    uint32_t K, lda, ldb;
    uint32_t bmap, bbh;
    uint32_t units;
    uint32_t amap, ascale;
    uint8_t a_base, b_base, trans_a, trans_b;
};

struct GemmUnit {
    uint64_t src;
    uint32_t ld, map, col0, valid;
};

// S-stage region: dst (R x C, column stride ld_dst) = sum_{t in [t0,t1)} alpha_t * src_t,
// where src_t is R*C contiguous elements of the workspace starting at src_off[t].
struct GatherRegion {
    uint64_t dst_off;
    uint32_t R, C, ld_dst;
    uint32_t t0, t1;
    uint32_t chunk_begin;
};

// Workgroups are dealt to the 8 XCDs round-robin (workgroup b runs on XCD b % 8), each XCD with its
// own L2.  Logical tiles are renumbered so that each XCD runs one contiguous range of them:
// neighbouring tiles (same A row panel, same task) then share an L2.  A bijection on [0, n).
__device__ inline uint32_t xcd_tile(uint32_t b, uint32_t n) {
    constexpr uint32_t X = 8;
    const uint32_t q = n / X, r = n % X, x = b % X, i = b / X;
    return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

constexpr uint32_t MAP_NONE = 0xffffffffu;
// Arithmetic of an application (mq_precision).  MODE_BF16X3: split products (hi*lo + lo*hi + hi*hi
// of bf16 hi / lo parts, FP32 accumulation) in the S-stage kernel and the lean 128 x 128 T and
// final kernels; the generic GEMM kernel runs MODE_F64 and MODE_F32 only and takes BF16X3's other
// tasks in MODE_F32.
enum GemmMode { MODE_F64 = 0, MODE_F32 = 1, MODE_BF16X3 = 2 };

// Tile configurations of the GEMM (gemm.hip); a task runs with the one that fits its (M, N).
enum GemmCfg { GEMM_CFG_128x128 = 0, GEMM_CFG_64x64 = 1, GEMM_CFG_32x128 = 2, GEMM_CFG_16x128 = 3,
               GEMM_CFG_16x16 = 4, GEMM_N_CFG = 5 };
constexpr int GEMM_CFG_BM[GEMM_N_CFG] = {128, 64, 32, 16, 16};
constexpr int GEMM_CFG_BN[GEMM_N_CFG] = {128, 64, 128, 128, 16};
inline int gemm_pick_cfg(uint32_t M, uint32_t N) {
    if (M <= 16) return N <= 16 ? GEMM_CFG_16x16 : GEMM_CFG_16x128;
    if (M <= 32) return GEMM_CFG_32x128;
    if (M >= 128 && N >= 128) return GEMM_CFG_128x128;
    return GEMM_CFG_64x64;
}
constexpr int GEMM_THREADS = 256;

// Block-sparse S stage (s_stage.hip).  A task is one 128 x 128 tile of S_D = T_D W_D of one group:
// rows [0, M) of the row panel starting at a_off / c_off (T_D and S_D column-major with leading
// dimension ld, element offsets in the workspace), columns [0, N) of the column panel; its K steps
// are the runs [run0, run1).  A run covers keys [k0, k0 + nk) (k0 a multiple of 16) whose 16-key
// tiles of this panel's W_D are stored compactly from tile index `tile` on, each tile 128 x 16 as
// [n][k] (zero padded).
constexpr int S_BSR_TILE_M = 128, S_BSR_TILE_N = 128, S_BSR_TILE_K = 16;
struct BsrTask {
    uint64_t a_off, c_off;
    uint32_t M, N, ld, run0, run1, pad;
};
struct BsrRun {
    uint32_t k0, nk, tile, pad;
};
// mode MODE_F64: ws, W tiles (bt0) double; MODE_F32: float; MODE_BF16X3: ws float, bt0 / bt1 the
// bf16 hi / lo tiles
// narrow: every task's T_D step window, (32 * ld + S_BSR_TILE_M) elements of 8 bytes, fits in 2^32
// bytes, so the kernel addresses it with 32-bit offsets from a uniform base
hipError_t launch_s_bsr(hipStream_t st, int mode, const BsrTask* tasks, uint32_t n_tasks, const BsrRun* runs,
                        const void* bt0, const void* bt1, void* ws, bool narrow);

// Side tables of a launch: remap entries, column units, row units, block scales.
struct GemmTables {
    const uint32_t* maps;
    const GemmUnit* units;
    const GemmRow* rows;
    const double* scales;
};

// One launch of all tasks of one tile configuration.  mode MODE_F64: all operands double.
// Otherwise `types` gives the storage of A, B and C as 'd' (double), 'f' (float) or 's' (split
// bf16 output of stacked tasks): "ddf", "fff", "dfd" or "dds".
// tile_task (may be null): task index of every tile of the launch.
hipError_t launch_gemm(hipStream_t st, int cfg, int mode, const char* types, const GemmTask* tasks, uint32_t n_tasks,
                       const GemmSeg* segs, uint32_t n_tiles, Bases bases, GemmTables tables,
                       const uint32_t* tile_task);
// Stacked T-stage tasks (row units, one segment, no A / B remaps, no beta) of tile configuration
// cfg (GEMM_CFG_128x128 or GEMM_CFG_64x64) on the lean kernel of t_stage.hip: mode MODE_F64
// (C double), MODE_F32 (C float if c_float, else double) or MODE_BF16X3 (C float), A and B double;
// same tables, tiles per task and results as launch_gemm with that cfg.
hipError_t launch_t_stack(hipStream_t st, int cfg, int mode, bool c_float, const GemmTask* tasks, uint32_t n_tasks,
                          const GemmSeg* segs, uint32_t n_tiles, Bases bases, GemmTables tb, const uint32_t* tile_task);
// Final-stage tasks of tile configuration cfg (GEMM_CFG_128x128 or GEMM_CFG_64x64) whose segments
// all read B through column units (no row units, no B block map, no beta, m_off 0) on the lean
// kernel of f_stage.hip: MODE_F64 (all double), MODE_F32 or MODE_BF16X3 (A double, B float,
// C double); same tables, tiles per task and results as launch_gemm with that cfg.
hipError_t launch_f_unit(hipStream_t st, int cfg, int mode, const GemmTask* tasks, uint32_t n_tasks, const GemmSeg* segs,
                         uint32_t n_tiles, Bases bases, GemmTables tb, const uint32_t* tile_task);
constexpr int GATHER_THREADS = 256;
constexpr int GATHER_CHUNK = 1024;   // elements per workgroup
constexpr int RED_BLOCKS = 512;      // fixed grid of the deterministic reductions
constexpr int GATHER_SMALL = 256;    // regions up to this many elements: one wavefront each

hipError_t launch_gather_small(hipStream_t st, const GatherRegion* regions, uint32_t n_regions,
                               const uint64_t* src_off, const double* alpha, double* ws, double* dst_base);

hipError_t launch_gather(hipStream_t st, const GatherRegion* regions, uint32_t n_regions,
                         const uint64_t* src_off, const double* alpha, uint32_t n_chunks, double* ws,
                         double* dst_base);
hipError_t launch_fill(hipStream_t st, double* x, uint64_t n, double v);
// grad[t] = sum_{e in [ptr[t], ptr[t+1])} coef[e] * g[idx[e]]   (one thread per term, fixed order)
hipError_t launch_term_grad(hipStream_t st, uint32_t n_terms, const uint32_t* ptr, const uint64_t* idx,
                            const double* coef, const double* g, double* grad);
// clear every chunk's range of base, elements of elem_bytes (2, 4 or 8) bytes
hipError_t launch_zero_chunks(hipStream_t st, const ZeroChunk* chunks, uint32_t n_chunks, void* base, int elem_bytes);
// dst[0, n) = hi, dst[n, 2n) = lo, dst[2n, 3n) = hi of x, with hi = bf16(x), lo = bf16(x - hi)
// (round to nearest even)
hipError_t launch_split3_bf16(hipStream_t st, const float* x, uint16_t* dst, uint64_t n);
// dst[i] = float(src[i])
hipError_t launch_to_float(hipStream_t st, const double* src, float* dst, uint64_t n);
// B' = [W_lo; W_hi; W_hi] (3K x N, column-major, ld 3K) of W (K x N, column stride ldw), with
// hi = bf16(float(w)), lo = bf16(float(w) - hi)
hipError_t launch_split_w_bf16(hipStream_t st, const double* W, uint64_t ldw, uint64_t K, uint64_t N, uint16_t* dst);
// n W-tile elements: FP32 value and bf16 split hi / lo
hipError_t launch_split_tiles(hipStream_t st, const double* w, uint64_t n, float* w32, uint16_t* hi, uint16_t* lo);
// out[pos[i]] = val[i], i < n (distinct positions)
hipError_t launch_scatter(hipStream_t st, uint64_t n, const uint64_t* pos, const double* val, double* out);
hipError_t launch_dot(hipStream_t st, uint64_t n, const double* x, const double* y, double* partial,
                      double* result);
hipError_t launch_axpy(hipStream_t st, uint64_t n, double a, const double* x, double* y);
// batched dot: result[j] = V_j . x for j < k (V_j = V + j * ldv; result on the device);
// partial holds k * RED_BLOCKS doubles.  Fixed grid and order: bitwise reproducible.
constexpr int MDOT_MAXK = 16;   // vectors per pass of the batched kernels
hipError_t launch_multi_dot(hipStream_t st, uint64_t n, int k, const double* V, uint64_t ldv, const double* x,
                            double* partial, double* result);
// batched axpy: y += alpha * sum_{j<k} c_j V_j, coefficients from device memory (c_dev) or the host
// (c_host, passed by value); exactly one of them is non-null
hipError_t launch_multi_axpy(hipStream_t st, uint64_t n, int k, double alpha, const double* c_dev, const double* c_host,
                             const double* V, uint64_t ldv, double* y);
hipError_t launch_scal(hipStream_t st, uint64_t n, double a, double* x);
hipError_t launch_permute(hipStream_t st, uint64_t n, const uint64_t* idx, const double* src, double* dst);
// per-b sums of boundary blocks: out[b] = sum of block b over all cohorts
hipError_t launch_block_sums(hipStream_t st, const uint64_t* blk_off, const uint32_t* blk_len,
                             const uint32_t* blk_b, uint32_t n_blk, const double* data, double* out_partial);

}  // namespace mq
