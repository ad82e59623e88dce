// Stacked T-stage GEMM of the grouped sigma and right boundary steps on gfx950 MFMA:
//
//   T_D[pair rows, keys] = psi blocks (stacked row units) * R cohort      one task per right cohort
//
// The tasks are the stacked 128 x 128 and 64 x 64 tasks of the generic grouped GEMM (GemmTask / GemmRow /
// GemmSeg, one segment, no A / B remaps, no beta): the psi blocks that meet one right cohort are
// the row units of the task, the cohort (or a Hermitian partner block, transposed and scaled) is B,
// and every 16-row block is stored to its own pair slab through the unit's column map.  This
// kernel is the lean form of that case (the generic kernel keeps every other task shape):
//
//   * operands are staged k-contiguous, A as [m][k] and B as [n][k] (16 keys + a 16-byte pad per
//     row), and each lane reads the 4 consecutive keys it needs for the 4 MFMA k-steps of a 16-key
//     step with one 16-byte LDS read: lane group lk takes keys 4 lk .. 4 lk + 3 and k-step t uses
//     key 4 lk + t, for A and B alike (the sum over k does not depend on the order);
//   * with 128 x 128 tiles thread t stages row t % 128 of A (8 keys: one row unit, one address
//     stream) and column t / 2 of B (8 keys: one contiguous 64-byte run of a non-transposed
//     cohort, four 16-byte loads); with 64 x 64 tiles row t % 64 and column t / 4, 4 keys each,
//     so the per-element address work of the generic kernel is gone;
//   * keys beyond the segment read the last key and are zeroed on the A side only; rows and
//     columns beyond the tile's range read clamped, finite data into outputs that are not stored.
//
// MODE_F64: FP64 operands in LDS, v_mfma_f64_16x16x4f64; MODE_F32: FP32 in LDS (converted while
// staged, A scaled by its unit and segment factors), v_mfma_f32_16x16x4f32; MODE_BF16X3: each FP32
// value split while staged into hi = bf16(x), lo = bf16(x - hi) (gfx950's v_cvt_pk_bf16_f32, round to
// nearest even) and three v_mfma_f32_16x16x16bf16 products per block and 16-key step (hi*lo + lo*hi +
// hi*hi, FP32 accumulation: the S stage's split-product scheme; lane group lk supplies keys
// 4 lk .. 4 lk + 3 as one 8-byte LDS read).  C is FP64 or FP32.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "kernels.hpp"

namespace mq {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d2_a8 __attribute__((ext_vector_type(2), aligned(8)));   // 8-byte aligned global runs
typedef short s4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));

constexpr int TK = 16, TT = GEMM_THREADS;

template <int MODE> struct TsTraits;
template <> struct TsTraits<MODE_F64> { typedef double T; typedef d4 Acc; static constexpr int P = TK + 2; };
template <> struct TsTraits<MODE_F32> { typedef float T; typedef f4 Acc; static constexpr int P = TK + 4; };

template <> struct TsTraits<MODE_BF16X3> { typedef float T; typedef f4 Acc; static constexpr int P = TK + 8; };

template <int MODE, int BM> struct TsSmem {
    typedef typename TsTraits<MODE>::T T;
    T a[BM][TsTraits<MODE>::P];
    T b[BM][TsTraits<MODE>::P];
};
// BF16X3: hi and lo bf16 of A and B, rows of 16 keys + 8 pad (48 bytes, 16-byte aligned runs)
template <int BM> struct TsSmem<MODE_BF16X3, BM> {
    static constexpr int P = TsTraits<MODE_BF16X3>::P;
    uint16_t ahi[BM][P], alo[BM][P], bhi[BM][P], blo[BM][P];
};

// waves per SIMD the kernel is compiled for (register budget): 128 x 128 tiles hold 16 blocks of
// accumulators per wave, 64 x 64 tiles 4
constexpr int ts_waves(int mode, int bm) {
    return bm == 64 ? 4 : mode == MODE_F32 ? 3 : mode == MODE_F64 ? 1 : 2;
}

// hi = bf16(x), lo = bf16(x - hi) of two elements, each pair packed lower element first
__device__ inline void split2(float x0, float x1, uint32_t& hi, uint32_t& lo) {
    hi = __builtin_bit_cast(uint32_t, bf2{(__bf16)x0, (__bf16)x1});
    lo = __builtin_bit_cast(uint32_t, bf2{(__bf16)(x0 - __uint_as_float(hi << 16)),
                                          (__bf16)(x1 - __uint_as_float(hi & 0xffff0000u))});
}

__device__ inline uint32_t task_of(const GemmTask* tasks, uint32_t n_tasks, const uint32_t* tile_task, uint32_t tile) {
    if (tile_task) return tile_task[tile];
    if (n_tasks == gridDim.x) return tile;
    uint32_t lo = 0, hi = n_tasks;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (tasks[mid].tile_begin <= tile) lo = mid; else hi = mid;
    }
    return lo;
}

// BM x BM tiles (BM = 128 or 64), 2 x 2 waves of WB x WB 16 x 16 blocks; each thread stages KPT
// keys of one A row and of one B column per 16-key step
template <int MODE, typename TC, int BM>
__global__ __launch_bounds__(TT) __attribute__((amdgpu_waves_per_eu(ts_waves(MODE, BM))))
void t_stack_kernel(const GemmTask* __restrict__ tasks, uint32_t n_tasks, const GemmSeg* __restrict__ segs,
                    Bases bases, GemmTables tables, const uint32_t* __restrict__ tile_task) {
    typedef typename TsTraits<MODE>::T T;
    typedef typename TsTraits<MODE>::Acc Acc;
    constexpr int TM = BM, TN = BM, KPT = BM / 16, WB = BM / 32, BT = TK / KPT;
    static_assert(BM == 128 || BM == 64, "tile");
    __shared__ __align__(16) TsSmem<MODE, BM> sm[2];
    const uint32_t* __restrict__ maps = tables.maps;
    const GemmRow* __restrict__ rows = tables.rows;

    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const GemmTask tk = tasks[task_of(tasks, n_tasks, tile_task, tile)];
    const GemmSeg sg = segs[tk.seg0];
    const uint32_t M = tk.M, N = tk.N, K = sg.K;
    const uint32_t local = tile - tk.tile_begin, tiles_m = (M + TM - 1) / TM;
    const uint32_t m0 = (local % tiles_m) * TM, n0 = (local / tiles_m) * TN;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave & 1) * (BM / 2), wn = (wave >> 1) * (BM / 2);
    const int l16 = lane & 15, lk = lane >> 4;

    // A: row a_m of the tile, KPT keys from a_k; its row unit's psi block row (clamped into the block)
    const int a_m = tid % TM, a_k = (tid / TM) * KPT;
    const double* __restrict__ A = bases.p[sg.a_base] + sg.a_off;
    const double* a_row;
    uint64_t a_sk;
    T a_sc;
    {
        const uint32_t m = min(m0 + uint32_t(a_m), M - 1);
        const GemmRow R = rows[maps[tk.rowunit + m / 16]];
        const uint32_t i = min(m - R.row0, R.nrows - 1);
        a_row = A + R.a_src + uint64_t(i) * R.a_si;
        a_sk = R.a_sk;
        a_sc = T(R.a_scale * sg.alpha);
    }
    // B: column b_n of the tile (clamped into op(B)), KPT keys from b_k
    const int b_n = tid / BT, b_k = (tid % BT) * KPT;
    const double* __restrict__ B = bases.p[sg.b_base] + sg.b_off;
    const uint32_t bn = min(n0 + uint32_t(b_n), N - 1);
    const bool tb = sg.trans_b;
    const double* b_col = tb ? B + bn : B + uint64_t(bn) * sg.ldb;   // op(B)(k, n) = b_col[k * (tb ? ldb : 1)]
    const uint64_t b_sk = tb ? sg.ldb : 1;

    Acc acc[WB][WB];
#pragma unroll
    for (int i = 0; i < WB; ++i)
#pragma unroll
        for (int j = 0; j < WB; ++j) acc[i][j] = Acc{0, 0, 0, 0};

    double va[KPT], vb[KPT];
    bool part = false;   // the staged step holds keys beyond K
    auto gload = [&](uint32_t k0) {
        part = k0 + TK > K;
        if (!part) {
            const double* pa = a_row + uint64_t(k0 + a_k) * a_sk;
#pragma unroll
            for (int e = 0; e < KPT; ++e) va[e] = pa[uint64_t(e) * a_sk];
            if (!tb) {
                const d2_a8* pb = reinterpret_cast<const d2_a8*>(b_col + k0 + b_k);
#pragma unroll
                for (int v = 0; v < KPT / 2; ++v) {
                    const d2_a8 x = pb[v];
                    vb[2 * v] = x[0];
                    vb[2 * v + 1] = x[1];
                }
            } else {
                const double* pb = b_col + uint64_t(k0 + b_k) * b_sk;
#pragma unroll
                for (int e = 0; e < KPT; ++e) vb[e] = pb[uint64_t(e) * b_sk];
            }
        } else {
#pragma unroll
            for (int e = 0; e < KPT; ++e) {
                va[e] = a_row[uint64_t(min(k0 + uint32_t(a_k + e), K - 1)) * a_sk];
                vb[e] = b_col[uint64_t(min(k0 + uint32_t(b_k + e), K - 1)) * b_sk];
            }
        }
    };
    auto lstore = [&](int buf, uint32_t k0) {
        TsSmem<MODE, BM>& s = sm[buf];
        T xa[KPT], xb[KPT];
#pragma unroll
        for (int e = 0; e < KPT; ++e) {
            xa[e] = T(va[e]) * a_sc;
            if (part && k0 + uint32_t(a_k + e) >= K) xa[e] = T(0);
            xb[e] = T(vb[e]);
        }
        if constexpr (MODE == MODE_BF16X3) {
            uint32_t ah[KPT / 2], al[KPT / 2], bh[KPT / 2], bl[KPT / 2];
#pragma unroll
            for (int e = 0; e < KPT / 2; ++e) {
                split2(xa[2 * e], xa[2 * e + 1], ah[e], al[e]);
                split2(xb[2 * e], xb[2 * e + 1], bh[e], bl[e]);
            }
            if constexpr (KPT == 8) {
                *reinterpret_cast<uint4*>(&s.ahi[a_m][a_k]) = uint4{ah[0], ah[1], ah[2], ah[3]};
                *reinterpret_cast<uint4*>(&s.alo[a_m][a_k]) = uint4{al[0], al[1], al[2], al[3]};
                *reinterpret_cast<uint4*>(&s.bhi[b_n][b_k]) = uint4{bh[0], bh[1], bh[2], bh[3]};
                *reinterpret_cast<uint4*>(&s.blo[b_n][b_k]) = uint4{bl[0], bl[1], bl[2], bl[3]};
            } else {
                *reinterpret_cast<uint2*>(&s.ahi[a_m][a_k]) = uint2{ah[0], ah[1]};
                *reinterpret_cast<uint2*>(&s.alo[a_m][a_k]) = uint2{al[0], al[1]};
                *reinterpret_cast<uint2*>(&s.bhi[b_n][b_k]) = uint2{bh[0], bh[1]};
                *reinterpret_cast<uint2*>(&s.blo[b_n][b_k]) = uint2{bl[0], bl[1]};
            }
        } else if constexpr (MODE == MODE_F32) {
#pragma unroll
            for (int v = 0; v < KPT / 4; ++v) {
                *reinterpret_cast<float4*>(&s.a[a_m][a_k + 4 * v]) = float4{xa[4 * v], xa[4 * v + 1], xa[4 * v + 2], xa[4 * v + 3]};
                *reinterpret_cast<float4*>(&s.b[b_n][b_k + 4 * v]) = float4{xb[4 * v], xb[4 * v + 1], xb[4 * v + 2], xb[4 * v + 3]};
            }
        } else {
#pragma unroll
            for (int v = 0; v < KPT / 2; ++v) {
                *reinterpret_cast<double2*>(&s.a[a_m][a_k + 2 * v]) = double2{xa[2 * v], xa[2 * v + 1]};
                *reinterpret_cast<double2*>(&s.b[b_n][b_k + 2 * v]) = double2{xb[2 * v], xb[2 * v + 1]};
            }
        }
    };
    auto compute = [&](int buf) {
        const TsSmem<MODE, BM>& s = sm[buf];
        if constexpr (MODE == MODE_BF16X3) {
            s4 ah[WB], al[WB], bh[WB], bl[WB];   // keys 4 lk .. 4 lk + 3 of one row / column
#pragma unroll
            for (int i = 0; i < WB; ++i) {
                ah[i] = *reinterpret_cast<const s4*>(&s.ahi[wm + 16 * i + l16][4 * lk]);
                al[i] = *reinterpret_cast<const s4*>(&s.alo[wm + 16 * i + l16][4 * lk]);
                bh[i] = *reinterpret_cast<const s4*>(&s.bhi[wn + 16 * i + l16][4 * lk]);
                bl[i] = *reinterpret_cast<const s4*>(&s.blo[wn + 16 * i + l16][4 * lk]);
            }
#pragma unroll
            for (int i = 0; i < WB; ++i)
#pragma unroll
                for (int j = 0; j < WB; ++j) {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[i], bl[j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al[i], bh[j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[i], bh[j], acc[i][j], 0, 0, 0);
                }
        } else {
            T a[WB][4], b[WB][4];   // [block][k-step]: keys 4 lk + t
#pragma unroll
            for (int i = 0; i < WB; ++i) {
                const T* pa = &s.a[wm + 16 * i + l16][4 * lk];
                const T* pb = &s.b[wn + 16 * i + l16][4 * lk];
                if constexpr (MODE == MODE_F32) {
                    const float4 x = *reinterpret_cast<const float4*>(pa);
                    const float4 y = *reinterpret_cast<const float4*>(pb);
                    a[i][0] = x.x, a[i][1] = x.y, a[i][2] = x.z, a[i][3] = x.w;
                    b[i][0] = y.x, b[i][1] = y.y, b[i][2] = y.z, b[i][3] = y.w;
                } else {
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const double2 x = *reinterpret_cast<const double2*>(pa + 2 * h);
                        const double2 y = *reinterpret_cast<const double2*>(pb + 2 * h);
                        a[i][2 * h] = x.x, a[i][2 * h + 1] = x.y;
                        b[i][2 * h] = y.x, b[i][2 * h + 1] = y.y;
                    }
                }
            }
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int i = 0; i < WB; ++i)
#pragma unroll
                    for (int j = 0; j < WB; ++j) {
                        if constexpr (MODE == MODE_F64)
                            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][t], b[j][t], acc[i][j], 0, 0, 0);
                        else
                            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i][t], b[j][t], acc[i][j], 0, 0, 0);
                    }
        }
    };

    // ---- the K steps: compute step k0 while step k0 + 16 loads (one barrier per step)
    if (K > 0) {
        uint32_t k0 = 0;
        gload(k0);
        lstore(0, k0);
        __syncthreads();
        int buf = 0;
        while (true) {
            const uint32_t k1 = k0 + TK;
            const bool more = k1 < K;
            if (more) gload(k1);
            compute(buf);
            if (!more) break;
            lstore(buf ^ 1, k1);
            __syncthreads();
            buf ^= 1;
            k0 = k1;
        }
    }

    // ---- epilogue: each 16-row block of the wave lies in one row unit, stored to its slab through
    // the unit's column map; the 4 FP32 accumulator rows of a lane are consecutive (one 16-byte
    // store when all are inside the unit)
    TC* __restrict__ C = reinterpret_cast<TC*>(bases.p[tk.c_base]) + tk.c_off;
    // the lane's column of block j: its column block in the unit's column map and the column in it
    // (the same for every row block)
    uint32_t cblk[WB], cin[WB];
    {
        const uint32_t cbw = max(tk.cbw, 1u);
#pragma unroll
        for (int j = 0; j < WB; ++j) {
            const uint32_t gn = n0 + wn + 16 * j + l16;
            cblk[j] = gn / cbw;
            cin[j] = gn - cblk[j] * cbw;
        }
    }
#pragma unroll
    for (int i = 0; i < WB; ++i) {
        const uint32_t mb = m0 + wm + 16 * i;
        if (mb >= M) continue;
        const GemmRow R = rows[maps[tk.rowunit + mb / 16]];
#pragma unroll
        for (int j = 0; j < WB; ++j) {
            const uint32_t gn = n0 + wn + 16 * j + l16;
            if (gn >= N) continue;
            uint64_t coff;
            if (R.cmap != MAP_NONE) {
                coff = uint64_t(maps[R.cmap + cblk[j]]) * R.cldk + uint64_t(cin[j]) * R.ldc;
            } else {
                coff = uint64_t(gn) * R.ldc;
            }
            TC* __restrict__ Cc = C + R.c_src + coff;
            if constexpr (MODE != MODE_F64 && std::is_same<TC, float>::value) {
                const uint32_t li0 = mb + 4 * lk - R.row0;
                if (R.c_si == 1 && mb + 4 * lk + 3 < M && li0 + 3 < R.nrows) {
                    typedef float __attribute__((ext_vector_type(4), aligned(4))) f4_a4;
                    *reinterpret_cast<f4_a4*>(Cc + li0) = f4_a4{acc[i][j][0], acc[i][j][1], acc[i][j][2], acc[i][j][3]};
                    continue;
                }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t gm = mb + (MODE == MODE_F64 ? 4 * r + lk : 4 * lk + r);
                const uint32_t li = gm - R.row0;
                if (gm < M && li < R.nrows) Cc[uint64_t(li) * R.c_si] = TC(acc[i][j][r]);
            }
        }
    }
}

template <int BM>
hipError_t launch_t_tile(hipStream_t st, int mode, bool c_float, const GemmTask* tasks, uint32_t n_tasks,
                         const GemmSeg* segs, uint32_t n_tiles, Bases bases, GemmTables tb, const uint32_t* tile_task) {
    const dim3 grid(n_tiles), block(TT);
    if (mode == MODE_F64 && !c_float)
        hipLaunchKernelGGL((t_stack_kernel<MODE_F64, double, BM>), grid, block, 0, st, tasks, n_tasks, segs, bases, tb, tile_task);
    else if (mode == MODE_F32 && c_float)
        hipLaunchKernelGGL((t_stack_kernel<MODE_F32, float, BM>), grid, block, 0, st, tasks, n_tasks, segs, bases, tb, tile_task);
    else if (mode == MODE_BF16X3 && c_float)
        hipLaunchKernelGGL((t_stack_kernel<MODE_BF16X3, float, BM>), grid, block, 0, st, tasks, n_tasks, segs, bases, tb, tile_task);
    else if (mode == MODE_F32)
        hipLaunchKernelGGL((t_stack_kernel<MODE_F32, double, BM>), grid, block, 0, st, tasks, n_tasks, segs, bases, tb, tile_task);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace

hipError_t launch_t_stack(hipStream_t st, int cfg, int mode, bool c_float, const GemmTask* tasks, uint32_t n_tasks,
                          const GemmSeg* segs, uint32_t n_tiles, Bases bases, GemmTables tb, const uint32_t* tile_task) {
    if (n_tiles == 0) return hipSuccess;
    if (cfg == GEMM_CFG_128x128)
        return launch_t_tile<128>(st, mode, c_float, tasks, n_tasks, segs, n_tiles, bases, tb, tile_task);
    if (cfg == GEMM_CFG_64x64)
        return launch_t_tile<64>(st, mode, c_float, tasks, n_tasks, segs, n_tiles, bases, tb, tile_task);
    return hipErrorInvalidValue;
}

}  // namespace mq
