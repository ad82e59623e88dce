// Final stage of the grouped sigma on gfx950 MFMA: one output block of sigma (ls_out x its columns)
// is a sum over K segments, one per piece of a left cohort,
//
//   C[:, n] += op(L piece) * S_D(unit of column n)          C: sigma or an FP64 split-K partial
//
// with op(L piece) either the cohort's blocks side by side (A(m, k) = L[a_off + m + k lda]) or the
// Hermitian partner's blocks transposed and scaled (A(m, k) = alpha * scale_j * L[a_off +
// maps[amap + j] + k % bbh + m lda], j = k / bbh), and B read through the task's column units (the
// S_D slab of the pair of each unit, column maps[U.map + k / bbh] of it, rows k % bbh of the unit's
// column).  These are the 128 x 128 and 64 x 64 tasks of the generic grouped GEMM whose segments
// all use column units (GemmTask / GemmSeg / GemmUnit, no B block map, no beta); this kernel is
// their lean form:
//
//   * with BM x BM tiles (BM = 128 or 64) thread t stages row t % BM of A and column t % BM of B
//     for KPT = BM / 16 consecutive keys of the 16-key step (keys KPT (t / BM) ..): within a wave
//     the keys, and so the block j = k / bbh and the offset in it, are uniform, so block lookups
//     are scalar loads and the KPT keys of one block are one contiguous run (16-byte loads of FP32
//     S_D or FP64; a transposed A block's run likewise);
//   * operands are staged k-contiguous ([m][k], [n][k]) and read with one 16-byte LDS read per
//     4 keys, k-step t using key 4 lk + t (the t_stage.hip scheme);
//   * a column unit is resolved once per segment; a B row (maps[U.map + j]) is reloaded only when
//     the step's block changes; keys beyond the segment are zeroed on the A side, invalid units on
//     the B side.
//
// MODE_F64: FP64 in LDS, v_mfma_f64_16x16x4f64; MODE_F32: FP32 in LDS, v_mfma_f32_16x16x4f32 (TB
// float: the reduced-precision S_D); MODE_BF16X3: A and B split while staged into bf16 hi / lo and
// three v_mfma_f32_16x16x16bf16 products per block and step (the t_stage.hip scheme).  C is FP64.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "kernels.hpp"

namespace mq {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d2_a8 __attribute__((ext_vector_type(2), aligned(8)));
typedef float f4_a4 __attribute__((ext_vector_type(4), aligned(4)));
typedef short s4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));

constexpr int FK = 16, FT = GEMM_THREADS;

template <int MODE> struct FsTraits;
template <> struct FsTraits<MODE_F64> { typedef double T; typedef d4 Acc; static constexpr int P = FK + 2; };
template <> struct FsTraits<MODE_F32> { typedef float T; typedef f4 Acc; static constexpr int P = FK + 4; };

template <> struct FsTraits<MODE_BF16X3> { typedef float T; typedef f4 Acc; static constexpr int P = FK + 8; };

template <int MODE, int BM> struct FsSmem {
    typedef typename FsTraits<MODE>::T T;
    T a[BM][FsTraits<MODE>::P];
    T b[BM][FsTraits<MODE>::P];
};
template <int BM> struct FsSmem<MODE_BF16X3, BM> {
    static constexpr int P = FsTraits<MODE_BF16X3>::P;
    uint16_t ahi[BM][P], alo[BM][P], bhi[BM][P], blo[BM][P];
};

// hi = bf16(x), lo = bf16(x - hi) of two elements, each pair packed lower element first
__device__ inline void fsplit2(float x0, float x1, uint32_t& hi, uint32_t& lo) {
    hi = __builtin_bit_cast(uint32_t, bf2{(__bf16)x0, (__bf16)x1});
    lo = __builtin_bit_cast(uint32_t, bf2{(__bf16)(x0 - __uint_as_float(hi << 16)),
                                          (__bf16)(x1 - __uint_as_float(hi & 0xffff0000u))});
}

__device__ inline uint32_t f_task_of(const GemmTask* tasks, uint32_t n_tasks, const uint32_t* tile_task, uint32_t tile) {
    if (tile_task) return tile_task[tile];
    if (n_tasks == gridDim.x) return tile;
    uint32_t lo = 0, hi = n_tasks;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (tasks[mid].tile_begin <= tile) lo = mid; else hi = mid;
    }
    return lo;
}

// N consecutive elements from p (8-byte aligned doubles or 4-byte aligned floats; N = 4 or 8)
template <int N> __device__ inline void loadn(const double* p, double* v) {
#pragma unroll
    for (int q = 0; q < N / 2; ++q) {
        const d2_a8 x = reinterpret_cast<const d2_a8*>(p)[q];
        v[2 * q] = x[0];
        v[2 * q + 1] = x[1];
    }
}
template <int N> __device__ inline void loadn(const float* p, float* v) {
#pragma unroll
    for (int q = 0; q < N / 4; ++q) {
        const f4_a4 x = reinterpret_cast<const f4_a4*>(p)[q];
        v[4 * q] = x[0], v[4 * q + 1] = x[1], v[4 * q + 2] = x[2], v[4 * q + 3] = x[3];
    }
}

// BM x BM tiles, 2 x 2 waves of WB x WB 16 x 16 blocks
template <int MODE, typename TB, int BM>
__global__ __launch_bounds__(FT) __attribute__((amdgpu_waves_per_eu(BM == 64 ? 4 : 2)))
void f_unit_kernel(const GemmTask* __restrict__ tasks, uint32_t n_tasks, const GemmSeg* __restrict__ segs,
                   Bases bases, GemmTables tables, const uint32_t* __restrict__ tile_task) {
    typedef typename FsTraits<MODE>::T T;
    typedef typename FsTraits<MODE>::Acc Acc;
    constexpr int FM = BM, FN = BM, KPT = BM / 16, WB = BM / 32;
    static_assert(BM == 128 || BM == 64, "tile");
    __shared__ __align__(16) FsSmem<MODE, BM> sm[2];
    const uint32_t* __restrict__ maps = tables.maps;
    const GemmUnit* __restrict__ units = tables.units;
    const double* __restrict__ scales = tables.scales;

    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const GemmTask tk = tasks[f_task_of(tasks, n_tasks, tile_task, tile)];
    const uint32_t M = tk.M, N = tk.N;
    const uint32_t local = tile - tk.tile_begin;
    const uint32_t m0 = (local / tk.tiles_n) * FM, n0 = (local % tk.tiles_n) * FN;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave & 1) * (BM / 2), wn = (wave >> 1) * (BM / 2);
    const int l16 = lane & 15, lk = lane >> 4;
    const int r_i = tid % BM, r_k = (tid / BM) * KPT;   // staged row (of A) / column (of B) and first key
    const uint32_t am = min(m0 + uint32_t(r_i), M - 1);
    const uint32_t bn = min(n0 + uint32_t(r_i), N - 1);
    const uint32_t unit = maps[tk.colunit + bn];

    Acc acc[WB][WB];
#pragma unroll
    for (int i = 0; i < WB; ++i)
#pragma unroll
        for (int j = 0; j < WB; ++j) acc[i][j] = Acc{0, 0, 0, 0};

    // per-segment state
    GemmSeg sg;
    uint32_t sgi = tk.seg0, k0 = 0;
    const double* A = nullptr;   // segment's A operand base
    const TB* Bcol = nullptr;    // this thread's unit column of the segment (row 0 of block 0)
    uint32_t u_ld = 0, u_map = 0;
    bool u_ok = false;
    uint32_t blk_c = 0xffffffffu;   // block of the cached B row
    const TB* brow = nullptr;
    auto begin_seg = [&]() {
        sg = segs[sgi];
        A = bases.p[sg.a_base] + sg.a_off;
        const GemmUnit U = units[sg.units + unit];
        u_ok = U.valid != 0;
        u_ld = U.ld;
        u_map = U.map;
        Bcol = reinterpret_cast<const TB*>(bases.p[sg.b_base]) + sg.b_off + U.src + uint64_t(sg.bbh) * (bn - U.col0);
        blk_c = 0xffffffffu;
    };
    auto settle = [&]() {   // (sgi, k0) at a step inside a segment; false after the last
        while (sgi < tk.seg1) {
            if (k0 < segs[sgi].K) return true;
            ++sgi;
            k0 = 0;
            if (sgi < tk.seg1) begin_seg();
        }
        return false;
    };

    double va[KPT];
    TB vb[KPT];
    bool part = false;
    auto gload = [&]() {
        const uint32_t K = sg.K, bbh = sg.bbh;
        const uint32_t kb = k0 + uint32_t(r_k);   // first key of this thread (uniform per wave)
        part = kb + KPT > K;
        const uint32_t j0 = min(kb, K - 1) / bbh, kin0 = min(kb, K - 1) - j0 * bbh;
        const bool run = !part && kin0 + KPT <= bbh;   // the KPT keys lie in one block (uniform)
        // ---- A
        if (sg.amap == MAP_NONE) {
            const double* p = A + am;
#pragma unroll
            for (int e = 0; e < KPT; ++e) va[e] = sg.alpha * p[uint64_t(min(kb + e, K - 1)) * sg.lda];
        } else if (run) {
            const double sc = sg.alpha * scales[sg.ascale + j0];
            const double* p = A + maps[sg.amap + j0] + kin0 + uint64_t(am) * sg.lda;
            loadn<KPT>(p, va);
#pragma unroll
            for (int e = 0; e < KPT; ++e) va[e] *= sc;
        } else {
#pragma unroll
            for (int e = 0; e < KPT; ++e) {
                const uint32_t k = min(kb + e, K - 1), j = k / bbh;
                va[e] = sg.alpha * scales[sg.ascale + j] * A[maps[sg.amap + j] + (k - j * bbh) + uint64_t(am) * sg.lda];
            }
        }
        // ---- B (the cached block row when the step stays in its block)
        if (run) {
            if (j0 != blk_c) {
                blk_c = j0;
                brow = Bcol + uint64_t(maps[u_map + j0]) * u_ld;
            }
            loadn<KPT>(brow + kin0, vb);
        } else {
#pragma unroll
            for (int e = 0; e < KPT; ++e) {
                const uint32_t k = min(kb + e, K - 1), j = k / bbh;
                vb[e] = Bcol[uint64_t(maps[u_map + j]) * u_ld + (k - j * bbh)];
            }
        }
    };
    auto lstore = [&](int buf) {
        FsSmem<MODE, BM>& s = sm[buf];
        const uint32_t kb = k0 + uint32_t(r_k);
        T xa[KPT], xb[KPT];
#pragma unroll
        for (int e = 0; e < KPT; ++e) {
            xa[e] = T(va[e]);
            if (part && kb + e >= sg.K) xa[e] = T(0);
            xb[e] = u_ok ? T(vb[e]) : T(0);
        }
        if constexpr (MODE == MODE_BF16X3) {
            uint32_t ah[KPT / 2], al[KPT / 2], bh[KPT / 2], bl[KPT / 2];
#pragma unroll
            for (int e = 0; e < KPT / 2; ++e) {
                fsplit2(xa[2 * e], xa[2 * e + 1], ah[e], al[e]);
                fsplit2(xb[2 * e], xb[2 * e + 1], bh[e], bl[e]);
            }
            if constexpr (KPT == 8) {
                *reinterpret_cast<uint4*>(&s.ahi[r_i][r_k]) = uint4{ah[0], ah[1], ah[2], ah[3]};
                *reinterpret_cast<uint4*>(&s.alo[r_i][r_k]) = uint4{al[0], al[1], al[2], al[3]};
                *reinterpret_cast<uint4*>(&s.bhi[r_i][r_k]) = uint4{bh[0], bh[1], bh[2], bh[3]};
                *reinterpret_cast<uint4*>(&s.blo[r_i][r_k]) = uint4{bl[0], bl[1], bl[2], bl[3]};
            } else {
                *reinterpret_cast<uint2*>(&s.ahi[r_i][r_k]) = uint2{ah[0], ah[1]};
                *reinterpret_cast<uint2*>(&s.alo[r_i][r_k]) = uint2{al[0], al[1]};
                *reinterpret_cast<uint2*>(&s.bhi[r_i][r_k]) = uint2{bh[0], bh[1]};
                *reinterpret_cast<uint2*>(&s.blo[r_i][r_k]) = uint2{bl[0], bl[1]};
            }
        } else if constexpr (MODE == MODE_F32) {
#pragma unroll
            for (int v = 0; v < KPT / 4; ++v) {
                *reinterpret_cast<float4*>(&s.a[r_i][r_k + 4 * v]) = float4{xa[4 * v], xa[4 * v + 1], xa[4 * v + 2], xa[4 * v + 3]};
                *reinterpret_cast<float4*>(&s.b[r_i][r_k + 4 * v]) = float4{xb[4 * v], xb[4 * v + 1], xb[4 * v + 2], xb[4 * v + 3]};
            }
        } else {
#pragma unroll
            for (int v = 0; v < KPT / 2; ++v) {
                *reinterpret_cast<double2*>(&s.a[r_i][r_k + 2 * v]) = double2{xa[2 * v], xa[2 * v + 1]};
                *reinterpret_cast<double2*>(&s.b[r_i][r_k + 2 * v]) = double2{xb[2 * v], xb[2 * v + 1]};
            }
        }
    };
    auto compute = [&](int buf) {
        const FsSmem<MODE, BM>& s = sm[buf];
        if constexpr (MODE == MODE_BF16X3) {
            s4 ah[WB], al[WB], bh[WB], bl[WB];
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
            T a[WB][4], b[WB][4];
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

    // ---- the K steps of all segments: compute step t while step t + 1 loads (one barrier per step)
    if (sgi < tk.seg1) begin_seg();
    if (settle()) {
        gload();
        lstore(0);
        __syncthreads();
        int buf = 0;
        while (true) {
            k0 += FK;
            const bool more = settle();
            if (more) gload();
            compute(buf);
            if (!more) break;
            lstore(buf ^ 1);
            __syncthreads();
            buf ^= 1;
        }
    }

    // ---- epilogue: C (FP64) rows < M, columns < N; plain column-major block at c_off + m_off
    double* __restrict__ C = bases.p[tk.c_base] + tk.c_off + tk.m_off;
#pragma unroll
    for (int j = 0; j < WB; ++j) {
        const uint32_t gn = n0 + wn + 16 * j + l16;
        if (gn >= N) continue;
        double* __restrict__ Cn = C + uint64_t(gn) * tk.ldc;
#pragma unroll
        for (int i = 0; i < WB; ++i) {
            if constexpr (MODE != MODE_F64) {
                const uint32_t gm0 = m0 + wm + 16 * i + 4 * lk;
                if (gm0 + 3 < M) {
                    d2_a8* q = reinterpret_cast<d2_a8*>(Cn + gm0);
                    q[0] = d2_a8{double(acc[i][j][0]), double(acc[i][j][1])};
                    q[1] = d2_a8{double(acc[i][j][2]), double(acc[i][j][3])};
                    continue;
                }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t gm = m0 + wm + 16 * i + (MODE == MODE_F64 ? 4 * r + lk : 4 * lk + r);
                if (gm < M) Cn[gm] = double(acc[i][j][r]);
            }
        }
    }
}

template <int BM>
hipError_t launch_f_tile(hipStream_t st, int mode, const GemmTask* tasks, uint32_t n_tasks, const GemmSeg* segs,
                         uint32_t n_tiles, Bases bases, GemmTables tb, const uint32_t* tile_task) {
    const dim3 grid(n_tiles), block(FT);
    if (mode == MODE_F64)
        hipLaunchKernelGGL((f_unit_kernel<MODE_F64, double, BM>), grid, block, 0, st, tasks, n_tasks, segs, bases, tb, tile_task);
    else if (mode == MODE_F32)
        hipLaunchKernelGGL((f_unit_kernel<MODE_F32, float, BM>), grid, block, 0, st, tasks, n_tasks, segs, bases, tb, tile_task);
    else if (mode == MODE_BF16X3)
        hipLaunchKernelGGL((f_unit_kernel<MODE_BF16X3, float, BM>), grid, block, 0, st, tasks, n_tasks, segs, bases, tb, tile_task);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace

hipError_t launch_f_unit(hipStream_t st, int cfg, int mode, const GemmTask* tasks, uint32_t n_tasks, const GemmSeg* segs,
                         uint32_t n_tiles, Bases bases, GemmTables tb, const uint32_t* tile_task) {
    if (n_tiles == 0) return hipSuccess;
    if (cfg == GEMM_CFG_128x128)
        return launch_f_tile<128>(st, mode, tasks, n_tasks, segs, n_tiles, bases, tb, tile_task);
    if (cfg == GEMM_CFG_64x64)
        return launch_f_tile<64>(st, mode, tasks, n_tasks, segs, n_tiles, bases, tb, tile_task);
    return hipErrorInvalidValue;
}

}  // namespace mq
