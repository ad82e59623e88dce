// Grouped, K-segmented GEMM on gfx950 MFMA, templated on operand storage types and compute mode.
//
//   MODE_F64    v_mfma_f64_16x16x4f64, FP64 operands and accumulation (the plain path)
//   MODE_F32    v_mfma_f32_16x16x4f32, FP32 operands and accumulation
//   MODE_BF16X3 v_mfma_f32_16x16x32_bf16 on a two-term split of each FP32 operand,
//               x = hi + lo (hi, lo bf16): a*b ~ a_hi*b_hi + a_hi*b_lo + a_lo*b_hi, FP32 accumulation
//
// Operands are read in their storage type (TA, TB: double or float) and converted when staged
// into LDS; C is written in TC.  One launch covers every tile of every task (the reference issues
// one cuBLAS call per block product, solver/tasks.cpp:226-229, 277-280, 335-336, 630-639, 738-748).
// Workgroup = 4 wavefronts = one 64x64 tile of C; wave w owns the 32x32 quadrant (w&1, w>>1) as
// 2x2 MFMA 16x16 sub-tiles.
//
// MFMA lane layouts on gfx950 (lane l = 0..63), measured with one-hot operands:
//   f64 16x16x4 : A row l%16, k l/16;  B col l%16, k l/16;  D[r] row 4*r + l/16, col l%16
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace mq {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef short s8 __attribute__((ext_vector_type(8)));

template <int MODE> struct ModeTraits;
template <> struct ModeTraits<MODE_F64> { typedef double T; typedef d4 Acc; static constexpr int BK = 16; };
template <> struct ModeTraits<MODE_F32> { typedef float T; typedef f4 Acc; static constexpr int BK = 16; };
template <> struct ModeTraits<MODE_BF16X3> { typedef float T; typedef f4 Acc; static constexpr int BK = 32; };

// output row of accumulator register r for lanes of k-group lk = l/16
template <int MODE> __device__ inline int d_row(int r, int lk);
template <> __device__ inline int d_row<MODE_F64>(int r, int lk) { return 4 * r + lk; }
template <> __device__ inline int d_row<MODE_F32>(int r, int lk) { return MQ_F32_DROW(r, lk); }
template <> __device__ inline int d_row<MODE_BF16X3>(int r, int lk) { return MQ_BF16_DROW(r, lk); }

// bf16 of x, round to nearest even (finite inputs)
__device__ inline short bf16_bits(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return short(u >> 16);
}
__device__ inline float bf16_value(short h) { return __uint_as_float(uint32_t(uint16_t(h)) << 16); }

template <int MODE, typename T, typename Acc>
__device__ inline void mma_tile(const T (*As)[GEMM_BM + 1], const T (*Bs)[GEMM_BN + 1], int wm, int wn, int l16,
                                int lk, Acc& acc00, Acc& acc01, Acc& acc10, Acc& acc11) {
    constexpr int BK = ModeTraits<MODE>::BK;
    if constexpr (MODE == MODE_F64) {
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            const int kr = kk + lk;
            const double a0 = As[kr][wm + l16], a1 = As[kr][wm + 16 + l16];
            const double b0 = Bs[kr][wn + l16], b1 = Bs[kr][wn + 16 + l16];
            acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc00, 0, 0, 0);
            acc01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc01, 0, 0, 0);
            acc10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc10, 0, 0, 0);
            acc11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc11, 0, 0, 0);
        }
    } else if constexpr (MODE == MODE_F32) {
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            const int kr = kk + lk;
            const float a0 = As[kr][wm + l16], a1 = As[kr][wm + 16 + l16];
            const float b0 = Bs[kr][wn + l16], b1 = Bs[kr][wn + 16 + l16];
            acc00 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc00, 0, 0, 0);
            acc01 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc01, 0, 0, 0);
            acc10 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc10, 0, 0, 0);
            acc11 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc11, 0, 0, 0);
        }
    } else {
        // one 16x16x32 step: lane holds k = 8*lk + j, j = 0..7
        s8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int kr = MQ_BF16_K(lk, j);
                const float a = As[kr][wm + 16 * q + l16], b = Bs[kr][wn + 16 * q + l16];
                const short ahi = bf16_bits(a), bhi = bf16_bits(b);
                ah[q][j] = ahi;
                al[q][j] = bf16_bits(a - bf16_value(ahi));
                bh[q][j] = bhi;
                bl[q][j] = bf16_bits(b - bf16_value(bhi));
            }
        Acc* accs[4] = {&acc00, &acc01, &acc10, &acc11};
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                Acc& c = *accs[2 * i + j];
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], c, 0, 0, 0);
            }
    }
}

template <typename TA, typename TB, typename TC, int MODE>
__global__ __launch_bounds__(GEMM_THREADS) void gemm_kernel(const GemmTask* __restrict__ tasks, uint32_t n_tasks,
                                                            const GemmSeg* __restrict__ segs, Bases bases,
                                                            const uint32_t* __restrict__ maps) {
    typedef typename ModeTraits<MODE>::T T;
    typedef typename ModeTraits<MODE>::Acc Acc;
    constexpr int BK = ModeTraits<MODE>::BK;
    constexpr int KSTEP = GEMM_THREADS / GEMM_BM;   // k rows per pass of the m-contiguous loaders
    __shared__ T As[BK][GEMM_BM + 1];
    __shared__ T Bs[BK][GEMM_BN + 1];

    const uint32_t tile = blockIdx.x;
    uint32_t lo = 0, hi = n_tasks;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (tasks[mid].tile_begin <= tile) lo = mid; else hi = mid;
    }
    const GemmTask tk = tasks[lo];
    const uint32_t local = tile - tk.tile_begin;
    const uint32_t m0 = (local / tk.tiles_n) * GEMM_BM;
    const uint32_t n0 = (local % tk.tiles_n) * GEMM_BN;
    const uint32_t M = tk.M, N = tk.N;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    const int l16 = lane & 15, lk = lane >> 4;

    Acc acc00 = {0, 0, 0, 0}, acc01 = {0, 0, 0, 0}, acc10 = {0, 0, 0, 0}, acc11 = {0, 0, 0, 0};

    for (uint32_t sgi = tk.seg0; sgi < tk.seg1; ++sgi) {
        const GemmSeg sg = segs[sgi];
        const TA* __restrict__ A = reinterpret_cast<const TA*>(bases.p[sg.a_base]) + sg.a_off;
        const TB* __restrict__ B = reinterpret_cast<const TB*>(bases.p[sg.b_base]) + sg.b_off;
        const uint32_t K = sg.K;
        const T alpha = T(sg.alpha);
        for (uint32_t k0 = 0; k0 < K; k0 += BK) {
            if (!sg.trans_a) {
                const int m = tid & 63;
#pragma unroll
                for (int i = 0; i < BK / KSTEP; ++i) {
                    const int k = (tid >> 6) + KSTEP * i;
                    const uint32_t gm = m0 + m, gk = k0 + k;
                    As[k][m] = (gm < M && gk < K) ? alpha * T(A[uint64_t(gk) * sg.lda + gm]) : T(0);
                }
            } else {
#pragma unroll
                for (int i = 0; i < (BK * GEMM_BM) / GEMM_THREADS; ++i) {
                    const int e = tid + GEMM_THREADS * i;
                    const int k = e % BK, m = e / BK;
                    const uint32_t gm = m0 + m, gk = k0 + k;
                    As[k][m] = (gm < M && gk < K) ? alpha * T(A[uint64_t(gm) * sg.lda + gk]) : T(0);
                }
            }
            if (sg.bmap != MAP_NONE) {
#pragma unroll
                for (int i = 0; i < (BK * GEMM_BN) / GEMM_THREADS; ++i) {
                    const int e = tid + GEMM_THREADS * i;
                    const int k = e % BK, n = e / BK;
                    const uint32_t gk = k0 + k, gn = n0 + n;
                    T v = T(0);
                    if (gk < K && gn < N) {
                        const uint32_t blk = gk / sg.bbh;
                        v = T(B[uint64_t(maps[sg.bmap + blk]) * sg.bldk + (gk - blk * sg.bbh) + uint64_t(gn) * sg.ldb]);
                    }
                    Bs[k][n] = v;
                }
            } else if (!sg.trans_b) {
#pragma unroll
                for (int i = 0; i < (BK * GEMM_BN) / GEMM_THREADS; ++i) {
                    const int e = tid + GEMM_THREADS * i;
                    const int k = e % BK, n = e / BK;
                    const uint32_t gn = n0 + n, gk = k0 + k;
                    Bs[k][n] = (gn < N && gk < K) ? T(B[uint64_t(gn) * sg.ldb + gk]) : T(0);
                }
            } else {
                const int n = tid & 63;
#pragma unroll
                for (int i = 0; i < BK / KSTEP; ++i) {
                    const int k = (tid >> 6) + KSTEP * i;
                    const uint32_t gn = n0 + n, gk = k0 + k;
                    Bs[k][n] = (gn < N && gk < K) ? T(B[uint64_t(gk) * sg.ldb + gn]) : T(0);
                }
            }
            __syncthreads();
            mma_tile<MODE, T, Acc>(As, Bs, wm, wn, l16, lk, acc00, acc01, acc10, acc11);
            __syncthreads();
        }
    }

    TC* __restrict__ C = reinterpret_cast<TC*>(bases.p[tk.c_base]) + tk.c_off;
    const uint64_t ldc = tk.ldc;
    const Acc* accs[4] = {&acc00, &acc01, &acc10, &acc11};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int i = q >> 1, j = q & 1;
        const uint32_t gn = n0 + wn + j * 16 + l16;
        if (gn >= N) continue;
        TC* __restrict__ Cn;
        if (tk.cmap != MAP_NONE) {
            const uint32_t blk = gn / tk.cbw;
            Cn = C + uint64_t(maps[tk.cmap + blk]) * tk.cldk + uint64_t(gn - blk * tk.cbw) * ldc;
        } else {
            Cn = C + uint64_t(gn) * ldc;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t gm = m0 + wm + i * 16 + d_row<MODE>(r, lk);
            if (gm < M) {
                TC v = TC((*accs[q])[r]);
                TC* p = Cn + gm;
                if (tk.beta) v += *p;
                *p = v;
            }
        }
    }
}

template <typename TA, typename TB, typename TC, int MODE>
static hipError_t launch(hipStream_t st, const GemmTask* tasks, uint32_t n_tasks, const GemmSeg* segs,
                         uint32_t n_tiles, Bases bases, const uint32_t* maps) {
    hipLaunchKernelGGL((gemm_kernel<TA, TB, TC, MODE>), dim3(n_tiles), dim3(GEMM_THREADS), 0, st, tasks, n_tasks,
                       segs, bases, maps);
    return hipGetLastError();
}

hipError_t launch_gemm(hipStream_t st, const GemmTask* tasks, uint32_t n_tasks, const GemmSeg* segs,
                       uint32_t n_tiles, Bases bases, const uint32_t* maps) {
    if (n_tiles == 0) return hipSuccess;
    return launch<double, double, double, MODE_F64>(st, tasks, n_tasks, segs, n_tiles, bases, maps);
}

// Mixed-precision stages.  types: operand storage of A, B, C as 'd' (double) or 'f' (float).
hipError_t launch_gemm_mixed(hipStream_t st, int mode, const char* types, const GemmTask* tasks, uint32_t n_tasks,
                             const GemmSeg* segs, uint32_t n_tiles, Bases bases, const uint32_t* maps) {
    if (n_tiles == 0) return hipSuccess;
    const bool da = types[0] == 'd', db = types[1] == 'd', dc = types[2] == 'd';
#define MQ_DISPATCH(MD)                                                                                       \
    if (da && db && !dc) return launch<double, double, float, MD>(st, tasks, n_tasks, segs, n_tiles, bases, maps); \
    if (!da && !db && !dc) return launch<float, float, float, MD>(st, tasks, n_tasks, segs, n_tiles, bases, maps); \
    if (da && !db && dc) return launch<double, float, double, MD>(st, tasks, n_tasks, segs, n_tiles, bases, maps); \
    if (da && db && dc) return launch<double, double, double, MD>(st, tasks, n_tasks, segs, n_tiles, bases, maps);
    if (mode == MODE_F32) { MQ_DISPATCH(MODE_F32) }
    if (mode == MODE_BF16X3) { MQ_DISPATCH(MODE_BF16X3) }
#undef MQ_DISPATCH
    return hipErrorInvalidValue;
}

}  // namespace mq
