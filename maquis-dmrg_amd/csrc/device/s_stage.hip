// Block-sparse S stage of the grouped sigma / boundary steps on gfx950 MFMA:
//
//   S_D[:, panel] = T_D[:, K runs] * W_D[K runs, panel]      per charge-transfer group D
//
// W_D's keys are in pattern-class order (engine.hip, reorder_group_keys), so its nonzeros sit in
// dense rectangles; every 128-column panel of W_D keeps only its 16-key tiles that hold a nonzero,
// stored compactly ([n][16 k] per tile, zero padded) and grouped into runs of consecutive tiles.
// One workgroup computes one 128 x 128 tile of S_D: it walks the runs of its panel, staging the
// T_D rows (column-major, rows contiguous) and the compact W tiles through LDS, double buffered,
// and accumulates with MFMA.  Three arithmetics:
//
//   MODE_F64     v_mfma_f64_16x16x4f64 on FP64 T_D / W tiles (the plain path)
//   MODE_F32     v_mfma_f32_16x16x4f32 on FP32 T_D / W tiles
//   MODE_BF16X3  FP32 T_D split on the fly into bf16 hi + lo while it is staged, W tiles split once
//                per plan; three v_mfma_f32_16x16x32_bf16 products per block and 32-key step
//                (hi*lo + lo*hi + hi*hi) with FP32 accumulation, the split-product scheme of the
//                library path (gfx950's K = 32 bf16 MFMA: twice the rate per key of the K = 16 form)
//
// The reference's S stage is a gather over the nonzero MPO couplings alpha
// (solver/numeric/gpu_dgemm_ddot.cu:282-359, solver/tasks.cpp:466-504); here the couplings of one
// group form W_D and its empty tiles are never visited.  Each output element is written once, by
// one workgroup, with a fixed summation order: the stage is deterministic.
//
// MFMA lane layouts (lane l, lk = l / 16):
//   f64 16x16x4        A row l%16, k lk;  B col l%16, k lk;  D[r] row 4r + lk, col l%16
//   f32 16x16x4        A row l%16, k lk;  B col l%16, k lk;  D[r] row 4lk + r, col l%16
//   bf16 16x16x32      A row l%16, 8 k per lane;  B col l%16, the same 8 k;  D[r] row 4lk + r
// (the bf16 kernel gives lane group lk the keys 8 lk .. 8 lk + 7 of its row and of its column; the
// sum over k does not depend on which of the 32 k positions each key takes, as long as A and B
// agree, and they use one layout)
#include <hip/hip_runtime.h>

#include <type_traits>

#include "kernels.hpp"

namespace mq {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int BM = S_BSR_TILE_M, BN = S_BSR_TILE_N, BK = S_BSR_TILE_K, NT = 256;
static_assert(BM == 128 && BN == 128 && BK == 16, "tile shape");

// Two elements: hi = bf16(x), lo = bf16(x - hi), round to nearest even (the rule of bf16_rne, the T
// stage's split), each pair packed with the lower element in the low half.  The float -> bf16
// conversions are gfx950's v_cvt_pk_bf16_f32 (two values per instruction).
__device__ inline void split_bf16x2(float x0, float x1, uint32_t& hi, uint32_t& lo) {
    hi = __builtin_bit_cast(uint32_t, bf2{(__bf16)x0, (__bf16)x1});
    const float h0 = __uint_as_float(hi << 16), h1 = __uint_as_float(hi & 0xffff0000u);
    lo = __builtin_bit_cast(uint32_t, bf2{(__bf16)(x0 - h0), (__bf16)(x1 - h1)});
}

template <int MODE> struct BsrTraits;
template <> struct BsrTraits<MODE_F64> { typedef double T; typedef d4 Acc; };
template <> struct BsrTraits<MODE_F32> { typedef float T; typedef f4 Acc; };
template <> struct BsrTraits<MODE_BF16X3> { typedef float T; typedef f4 Acc; };

// LDS of one stage buffer.  FP64 / FP32: A as [k][m] (T elements), B as [n][k].  BF16X3: A hi and
// lo as [m][k] bf16 (the MFMA reads 8 consecutive k of one row), B hi and lo as [n][k] bf16.
template <int MODE> struct BsrSmem;
template <int MODE> struct BsrSmemPlain {
    typedef typename BsrTraits<MODE>::T T;
    T a[BK][BM + 2];
    T b[BN][BK + 1];
};
template <> struct BsrSmem<MODE_F64> : BsrSmemPlain<MODE_F64> {};
template <> struct BsrSmem<MODE_F32> : BsrSmemPlain<MODE_F32> {};
// BF16X3 steps over two 16-key tiles (32 keys): its MFMAs are 16x faster per key than FP32's, and
// one tile per step left too little work between two waits on global memory.  Rows of 40 bf16
// (80 bytes): 16-byte aligned runs of 8 keys.
template <int MODE> struct BsrStep { static constexpr int KS = BK; };
template <> struct BsrStep<MODE_BF16X3> { static constexpr int KS = 2 * BK; };
template <> struct BsrSmem<MODE_BF16X3> {
    static constexpr int P = 2 * BK + 8;
    uint16_t ahi[BM][P], alo[BM][P];
    uint16_t bhi[BN][P], blo[BN][P];
};

template <int MODE, bool NARROW>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2))) void s_bsr_kernel(const BsrTask* __restrict__ tasks, const BsrRun* __restrict__ runs,
                                                   const void* __restrict__ bt0, const void* __restrict__ bt1,
                                                   void* __restrict__ ws) {
    typedef typename BsrTraits<MODE>::T T;
    typedef typename BsrTraits<MODE>::Acc Acc;
    constexpr bool BF = MODE == MODE_BF16X3;
    __shared__ __align__(16) BsrSmem<MODE> sm[2];

    // consecutive tasks (blocks of 8 x 8 panels of one group) run next to each other on one XCD
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const BsrTask tk = tasks[tile];
    const T* __restrict__ A = reinterpret_cast<const T*>(ws) + tk.a_off;
    T* __restrict__ C = reinterpret_cast<T*>(ws) + tk.c_off;
    const uint32_t M = tk.M, N = tk.N;
    const uint64_t ld = tk.ld;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;   // 2 x 2 waves of 64 x 64
    const int l16 = lane & 15, lk = lane >> 4;

    Acc acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = Acc{0, 0, 0, 0};

    // ---- staging of one step of KS keys (run r, key offset kk within the run)
    // A: T_D rows m0 .. m0+127 of KS key columns.  Thread t reads row t % 128 of the KS / 2 keys
    // (t / 128) * KS / 2 .. : every load instruction covers 64 consecutive rows of one column.  Rows
    // beyond M read valid workspace memory (the plan checks 128 rows past every panel); they only
    // feed output rows that are not stored.  Keys beyond the run's end read its last key: their W
    // entries are zero (the run's last tile is zero padded, and a tile past the run's end is not
    // read), so they add nothing.  The addresses are a wave-uniform base (the step's first key
    // column) plus 32-bit byte offsets when the plan allows it (NARROW: (KS * ld + BM) elements of
    // every task fit in 2^32 bytes), else full 64-bit offsets.
    // B: the compact tile(s) (128 columns x 16 keys each, [n][k], 16-byte aligned runs of 8 keys),
    // thread t reads column t / 2: 8 keys of the one tile (KS = 16), or the 16 keys of tile t % 2
    // (KS = 32; a tile past the run's end is not read and counts as zero).
    constexpr int KS = BsrStep<MODE>::KS, AK = KS / 2;
    const int a_m = tid & (BM - 1), a_k = (tid >> 7) * AK;
    const int b_n = tid >> 1, b_k = (tid & 1) * (KS / 2);
    const uint32_t ld_b = uint32_t(ld) * uint32_t(sizeof(T));             // narrow plans only
    const uint32_t a_b = uint32_t(a_m) * uint32_t(sizeof(T)) + uint32_t(a_k) * ld_b;
    T va[AK];
    T vb[8];
    uint4 vbh[KS / 16], vbl[KS / 16];
    auto gload = [&](const BsrRun& R, uint32_t kk) {
        const T* col = A + (uint64_t(R.k0) + kk) * ld;   // the step's first key column (uniform)
        const uint32_t last = R.nk - 1 - kk;              // last key of the run, relative to the step
        if constexpr (NARROW) {
            const char* base = reinterpret_cast<const char*>(col);
            if (kk + uint32_t(a_k + AK) <= R.nk) {        // uniform per wave: every key in the run
#pragma unroll
                for (int e = 0; e < AK; ++e) va[e] = *reinterpret_cast<const T*>(base + (a_b + uint32_t(e) * ld_b));
            } else {
#pragma unroll
                for (int e = 0; e < AK; ++e) {
                    const uint32_t key = min(uint32_t(a_k + e), last);
                    va[e] = *reinterpret_cast<const T*>(base + (uint32_t(a_m) * uint32_t(sizeof(T)) + key * ld_b));
                }
            }
        } else {
#pragma unroll
            for (int e = 0; e < AK; ++e) va[e] = col[uint64_t(min(uint32_t(a_k + e), last)) * ld + a_m];
        }
        if constexpr (BF) {
            // KS = 32: tile kk / 16 + (t % 2), its 16 keys of column t / 2
            const bool tok = kk + uint32_t(b_k) < R.nk;
            const uint64_t toff = (uint64_t(R.tile) + (kk + b_k) / BK) * (BN * BK) + uint64_t(b_n) * BK;
            const uint4* hi = reinterpret_cast<const uint4*>(reinterpret_cast<const uint16_t*>(bt0) + toff);
            const uint4* lo = reinterpret_cast<const uint4*>(reinterpret_cast<const uint16_t*>(bt1) + toff);
#pragma unroll
            for (int v = 0; v < 2; ++v) {
                vbh[v] = tok ? hi[v] : uint4{0, 0, 0, 0};
                vbl[v] = tok ? lo[v] : uint4{0, 0, 0, 0};
            }
        } else {
            const uint64_t toff = (uint64_t(R.tile) + kk / BK) * (BN * BK) + uint64_t(b_n) * BK + b_k;
            typedef typename std::conditional<MODE == MODE_F64, double2, float4>::type V;
            constexpr int W = sizeof(V) / sizeof(T);
            const V* bsrc = reinterpret_cast<const V*>(reinterpret_cast<const T*>(bt0) + toff);
#pragma unroll
            for (int v = 0; v < 8 / W; ++v) {
                const V x = bsrc[v];
                const T* px = reinterpret_cast<const T*>(&x);
#pragma unroll
                for (int w = 0; w < W; ++w) vb[v * W + w] = px[w];
            }
        }
    };
    auto lstore = [&](int buf) {
        BsrSmem<MODE>& s = sm[buf];
        if constexpr (BF) {
            uint32_t h[AK / 2], l[AK / 2];   // pairs of bf16, the lower key in the low half
#pragma unroll
            for (int e = 0; e < AK / 2; ++e) split_bf16x2(va[2 * e], va[2 * e + 1], h[e], l[e]);
#pragma unroll
            for (int v = 0; v < AK / 8; ++v) {
                *reinterpret_cast<uint4*>(&s.ahi[a_m][a_k + 8 * v]) = uint4{h[4 * v], h[4 * v + 1], h[4 * v + 2], h[4 * v + 3]};
                *reinterpret_cast<uint4*>(&s.alo[a_m][a_k + 8 * v]) = uint4{l[4 * v], l[4 * v + 1], l[4 * v + 2], l[4 * v + 3]};
            }
#pragma unroll
            for (int v = 0; v < KS / 16; ++v) {
                *reinterpret_cast<uint4*>(&s.bhi[b_n][b_k + 8 * v]) = vbh[v];
                *reinterpret_cast<uint4*>(&s.blo[b_n][b_k + 8 * v]) = vbl[v];
            }
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) s.a[a_k + e][a_m] = va[e];
#pragma unroll
            for (int e = 0; e < 8; ++e) s.b[b_n][b_k + e] = vb[e];
        }
    };
    auto compute = [&](int buf) {
        const BsrSmem<MODE>& s = sm[buf];
        if constexpr (BF) {
            static_assert(KS == 32, "one K = 32 MFMA step");
            bf8 ah[4], al[4], bh[4], bl[4];   // 8 consecutive keys of one row / column: one 16-byte read each
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int r = wm + 16 * i + l16;
                ah[i] = *reinterpret_cast<const bf8*>(&s.ahi[r][8 * lk]);
                al[i] = *reinterpret_cast<const bf8*>(&s.alo[r][8 * lk]);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int c = wn + 16 * j + l16;
                bh[j] = *reinterpret_cast<const bf8*>(&s.bhi[c][8 * lk]);
                bl[j] = *reinterpret_cast<const bf8*>(&s.blo[c][8 * lk]);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
                }
        } else {
#pragma unroll
            for (int ks = 0; ks < BK / 4; ++ks) {
                T a[4], b[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) a[i] = s.a[4 * ks + lk][wm + 16 * i + l16];
#pragma unroll
                for (int j = 0; j < 4; ++j) b[j] = s.b[wn + 16 * j + l16][4 * ks + lk];
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        if constexpr (MODE == MODE_F64)
                            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
                        else
                            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
                    }
            }
        }
    };

    // ---- the steps of all runs of the panel: prologue, then compute step t while step t+1 loads
    uint32_t r = tk.run0, kk = 0;
    if (r < tk.run1) {
        BsrRun R = runs[r];
        gload(R, kk);
        lstore(0);
        __syncthreads();
        int buf = 0;
        while (true) {
            // advance to the next step
            kk += KS;
            bool more = true;
            if (kk >= R.nk) {
                ++r;
                kk = 0;
                more = r < tk.run1;
                if (more) R = runs[r];
            }
            if (more) gload(R, kk);
            compute(buf);
            if (!more) break;
            lstore(buf ^ 1);
            __syncthreads();
            buf ^= 1;
        }
    }

    // ---- epilogue: C = acc (rows < M, columns < N).  FP32 accumulators of a lane are 4 consecutive
    // rows of one column: one 16-byte store when all four are in range (element-aligned vector
    // stores are legal for global memory)
    typedef float __attribute__((ext_vector_type(4), aligned(4))) f4_a4;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t col = wn + 16 * j + l16;
        if (col >= N) continue;
        T* cc = C + uint64_t(col) * ld;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if constexpr (MODE != MODE_F64) {
                const uint32_t row0 = wm + 16 * i + 4 * lk;
                if (row0 + 3 < M) {
                    *reinterpret_cast<f4_a4*>(cc + row0) = f4_a4{acc[i][j][0], acc[i][j][1], acc[i][j][2], acc[i][j][3]};
                    continue;
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t row = MODE == MODE_F64 ? wm + 16 * i + 4 * q + lk : wm + 16 * i + 4 * lk + q;
                if (row < M) cc[row] = acc[i][j][q];
            }
        }
    }
}

}  // namespace

hipError_t launch_s_bsr(hipStream_t st, int mode, const BsrTask* tasks, uint32_t n_tasks, const BsrRun* runs,
                        const void* bt0, const void* bt1, void* ws, bool narrow) {
    if (n_tasks == 0) return hipSuccess;
    const dim3 grid(n_tasks), block(NT);
#define MQ_S_BSR(M_)                                                                                  \
    if (narrow)                                                                                       \
        hipLaunchKernelGGL((s_bsr_kernel<M_, true>), grid, block, 0, st, tasks, runs, bt0, bt1, ws);  \
    else                                                                                              \
        hipLaunchKernelGGL((s_bsr_kernel<M_, false>), grid, block, 0, st, tasks, runs, bt0, bt1, ws)
    switch (mode) {
        case MODE_F64:
            MQ_S_BSR(MODE_F64);
            break;
        case MODE_F32:
            MQ_S_BSR(MODE_F32);
            break;
        case MODE_BF16X3:
            MQ_S_BSR(MODE_BF16X3);
            break;
        default:
            return hipErrorInvalidValue;
    }
#undef MQ_S_BSR
    return hipGetLastError();
}

}  // namespace mq
