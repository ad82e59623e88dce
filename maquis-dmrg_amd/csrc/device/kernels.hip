// gfx950 kernels of the sigma / boundary engine.
//
//  (the grouped MFMA GEMM lives in gemm.hip)
//  gather        the MPO-term gather/scale of T slabs into S (reference kernels
//                compute_s_stackedv / compute_s_left_stackedv, solver/numeric/gpu_dgemm_ddot.cu:215-359):
//                every region is owned by one workgroup chunk and written exactly once, so no
//                atomics and no memset of S are needed for covered regions.
//  dot/axpy/...  Davidson vector kernels with a fixed-shape (deterministic) reduction tree.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace mq {

// ------------------------------------------------------------------------------- gather
// One workgroup = GATHER_CHUNK consecutive elements of one region.  The (src, alpha) list of
// the region is uniform across the workgroup (scalar loads); every T slab read is a
// contiguous, coalesced stream.
__global__ __launch_bounds__(GATHER_THREADS) void gather_kernel(const GatherRegion* __restrict__ regions,
                                                                uint32_t n_regions,
                                                                const uint64_t* __restrict__ src_off,
                                                                const double* __restrict__ alpha,
                                                                const double* __restrict__ ws,
                                                                double* __restrict__ dst_base) {
    const uint32_t chunk = blockIdx.x;
    uint32_t lo = 0, hi = n_regions;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (regions[mid].chunk_begin <= chunk) lo = mid; else hi = mid;
    }
    const GatherRegion rg = regions[lo];
    const uint64_t total = uint64_t(rg.R) * rg.C;
    const uint64_t e0 = uint64_t(chunk - rg.chunk_begin) * GATHER_CHUNK + threadIdx.x;
    constexpr int PER = GATHER_CHUNK / GATHER_THREADS;
    double acc[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) acc[j] = 0.0;
    for (uint32_t t = rg.t0; t < rg.t1; ++t) {
        const double a = alpha[t];
        const double* __restrict__ src = ws + src_off[t];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const uint64_t e = e0 + uint64_t(j) * GATHER_THREADS;
            if (e < total) acc[j] += a * src[e];
        }
    }
    double* __restrict__ dst = dst_base + rg.dst_off;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const uint64_t e = e0 + uint64_t(j) * GATHER_THREADS;
        if (e < total) {
            const uint64_t c = e / rg.R, r = e - c * rg.R;
            dst[c * rg.ld_dst + r] = acc[j];
        }
    }
}

hipError_t launch_gather(hipStream_t st, const GatherRegion* regions, uint32_t n_regions, const uint64_t* src_off,
                         const double* alpha, uint32_t n_chunks, double* ws, double* dst_base) {
    if (n_chunks == 0) return hipSuccess;
    hipLaunchKernelGGL(gather_kernel, dim3(n_chunks), dim3(GATHER_THREADS), 0, st, regions, n_regions, src_off,
                       alpha, ws, dst_base);
    return hipGetLastError();
}

// Small regions (R*C <= GATHER_SMALL): one wavefront per region, four regions per workgroup.
// Regions are listed contiguously; chunk_begin is unused.
__global__ __launch_bounds__(256) void gather_small_kernel(const GatherRegion* __restrict__ regions,
                                                           uint32_t n_regions,
                                                           const uint64_t* __restrict__ src_off,
                                                           const double* __restrict__ alpha,
                                                           const double* __restrict__ ws,
                                                           double* __restrict__ dst_base) {
    const uint32_t rid = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (rid >= n_regions) return;
    const int lane = threadIdx.x & 63;
    const GatherRegion rg = regions[rid];
    const uint32_t total = rg.R * rg.C;
    constexpr int PER = GATHER_SMALL / 64;
    double acc[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) acc[j] = 0.0;
    for (uint32_t t = rg.t0; t < rg.t1; ++t) {
        const double a = alpha[t];
        const double* __restrict__ src = ws + src_off[t];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const uint32_t e = lane + 64 * j;
            if (e < total) acc[j] += a * src[e];
        }
    }
    double* __restrict__ dst = dst_base + rg.dst_off;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const uint32_t e = lane + 64 * j;
        if (e < total) {
            const uint32_t c = e / rg.R, r = e - c * rg.R;
            dst[uint64_t(c) * rg.ld_dst + r] = acc[j];
        }
    }
}

hipError_t launch_gather_small(hipStream_t st, const GatherRegion* regions, uint32_t n_regions,
                               const uint64_t* src_off, const double* alpha, double* ws, double* dst_base) {
    if (n_regions == 0) return hipSuccess;
    hipLaunchKernelGGL(gather_small_kernel, dim3((n_regions + 3) / 4), dim3(256), 0, st, regions, n_regions,
                       src_off, alpha, ws, dst_base);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------- elementwise
__global__ void fill_kernel(double* __restrict__ x, uint64_t n, double v) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        x[i] = v;
}

static inline unsigned grid_for(uint64_t n, unsigned threads) {
    uint64_t g = (n + threads - 1) / threads;
    if (g > 8192) g = 8192;
    if (g == 0) g = 1;
    return unsigned(g);
}

hipError_t launch_fill(hipStream_t st, double* x, uint64_t n, double v) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n, 256)), dim3(256), 0, st, x, n, v);
    return hipGetLastError();
}

// dst = [hi | lo | hi] of x (three consecutive n-element runs): the K-concatenated left operand of
// the one-GEMM form of the BF16x3 product [A_hi A_lo A_hi] [B_lo; B_hi; B_hi].
__global__ void split3_bf16_kernel(const float* __restrict__ x, uint16_t* __restrict__ dst, uint64_t n) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
        const float v = x[i];
        const uint16_t h = bf16_rne(v);
        dst[i] = h;
        dst[n + i] = bf16_rne(v - __uint_as_float(uint32_t(h) << 16));
        dst[2 * n + i] = h;
    }
}

hipError_t launch_split3_bf16(hipStream_t st, const float* x, uint16_t* dst, uint64_t n) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(split3_bf16_kernel, dim3(grid_for(n, 256)), dim3(256), 0, st, x, dst, n);
    return hipGetLastError();
}

template <typename E>
__global__ void zero_chunks_kernel(const ZeroChunk* __restrict__ chunks, E* __restrict__ base) {
    const ZeroChunk c = chunks[blockIdx.x];
    for (uint32_t i = threadIdx.x; i < c.len; i += blockDim.x) base[c.off + i] = E(0);
}
hipError_t launch_zero_chunks(hipStream_t st, const ZeroChunk* chunks, uint32_t n_chunks, void* base, int elem_bytes) {
    if (n_chunks == 0) return hipSuccess;
    if (elem_bytes == 8)
        hipLaunchKernelGGL(zero_chunks_kernel<uint64_t>, dim3(n_chunks), dim3(256), 0, st, chunks,
                           static_cast<uint64_t*>(base));
    else if (elem_bytes == 4)
        hipLaunchKernelGGL(zero_chunks_kernel<uint32_t>, dim3(n_chunks), dim3(256), 0, st, chunks,
                           static_cast<uint32_t*>(base));
    else if (elem_bytes == 2)
        hipLaunchKernelGGL(zero_chunks_kernel<uint16_t>, dim3(n_chunks), dim3(256), 0, st, chunks,
                           static_cast<uint16_t*>(base));
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

__global__ void term_grad_kernel(uint32_t n_terms, const uint32_t* __restrict__ ptr, const uint64_t* __restrict__ idx,
                                 const double* __restrict__ coef, const double* __restrict__ g, double* __restrict__ grad) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_terms) return;
    double acc = 0.0;
    for (uint32_t e = ptr[t]; e < ptr[t + 1]; ++e) acc += coef[e] * g[idx[e]];
    grad[t] = acc;
}
hipError_t launch_term_grad(hipStream_t st, uint32_t n_terms, const uint32_t* ptr, const uint64_t* idx,
                            const double* coef, const double* g, double* grad) {
    if (n_terms == 0) return hipSuccess;
    hipLaunchKernelGGL(term_grad_kernel, dim3((n_terms + 255) / 256), dim3(256), 0, st, n_terms, ptr, idx, coef, g,
                       grad);
    return hipGetLastError();
}

__global__ void to_float_kernel(const double* __restrict__ src, float* __restrict__ dst, uint64_t n) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        dst[i] = float(src[i]);
}
hipError_t launch_to_float(hipStream_t st, const double* src, float* dst, uint64_t n) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(to_float_kernel, dim3(grid_for(n, 256)), dim3(256), 0, st, src, dst, n);
    return hipGetLastError();
}
__global__ void split_w_bf16_kernel(const double* __restrict__ W, uint64_t ldw, uint64_t K, uint64_t N,
                                    uint16_t* __restrict__ dst) {
    const uint64_t total = K * N;
    for (uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total; e += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t n = e / K, k = e - n * K;
        const float v = float(W[n * ldw + k]);
        const uint16_t hi = bf16_rne(v);
        const uint16_t lo = bf16_rne(v - __uint_as_float(uint32_t(hi) << 16));
        uint16_t* col = dst + n * 3 * K;
        col[k] = lo;
        col[K + k] = hi;
        col[2 * K + k] = hi;
    }
}
hipError_t launch_split_w_bf16(hipStream_t st, const double* W, uint64_t ldw, uint64_t K, uint64_t N, uint16_t* dst) {
    if (K == 0 || N == 0) return hipSuccess;
    hipLaunchKernelGGL(split_w_bf16_kernel, dim3(grid_for(K * N, 256)), dim3(256), 0, st, W, ldw, K, N, dst);
    return hipGetLastError();
}

// W tiles of the block-sparse S stage: FP32 value and its bf16 split hi + lo (same rounding as
// the T stage's and the staging of T_D)
__global__ void split_tiles_kernel(const double* __restrict__ w, uint64_t n, float* __restrict__ w32,
                                   uint16_t* __restrict__ hi, uint16_t* __restrict__ lo) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
        const float v = float(w[i]);
        const uint16_t h = bf16_rne(v);
        w32[i] = v;
        hi[i] = h;
        lo[i] = bf16_rne(v - __uint_as_float(uint32_t(h) << 16));
    }
}
hipError_t launch_split_tiles(hipStream_t st, const double* w, uint64_t n, float* w32, uint16_t* hi, uint16_t* lo) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(split_tiles_kernel, dim3(grid_for(n, 256)), dim3(256), 0, st, w, n, w32, hi, lo);
    return hipGetLastError();
}

// out[pos[i]] = val[i] (positions distinct): the nonzeros of the block-sparse S stage's W tiles
// placed into their zero-filled tiles
__global__ void scatter_kernel(uint64_t n, const uint64_t* __restrict__ pos, const double* __restrict__ val,
                               double* __restrict__ out) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        out[pos[i]] = val[i];
}
hipError_t launch_scatter(hipStream_t st, uint64_t n, const uint64_t* pos, const double* val, double* out) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(scatter_kernel, dim3(grid_for(n, 256)), dim3(256), 0, st, n, pos, val, out);
    return hipGetLastError();
}

__global__ void axpy_kernel(uint64_t n, double a, const double* __restrict__ x, double* __restrict__ y) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        y[i] += a * x[i];
}
__global__ void scal_kernel(uint64_t n, double a, double* __restrict__ x) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        x[i] *= a;
}
hipError_t launch_axpy(hipStream_t st, uint64_t n, double a, const double* x, double* y) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(n, 256)), dim3(256), 0, st, n, a, x, y);
    return hipGetLastError();
}
hipError_t launch_scal(hipStream_t st, uint64_t n, double a, double* x) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(scal_kernel, dim3(grid_for(n, 256)), dim3(256), 0, st, n, a, x);
    return hipGetLastError();
}

__global__ void permute_kernel(uint64_t n, const uint64_t* __restrict__ idx, const double* __restrict__ src,
                               double* __restrict__ dst) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        dst[i] = src[idx[i]];
}
hipError_t launch_permute(hipStream_t st, uint64_t n, const uint64_t* idx, const double* src, double* dst) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(permute_kernel, dim3(grid_for(n, 256)), dim3(256), 0, st, n, idx, src, dst);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------- reductions
// Wavefront reduction (64 lanes) followed by an LDS reduction over the 4 wavefronts.
__device__ inline double block_sum_256(double v) {
    __shared__ double red[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) r = (red[0] + red[1]) + (red[2] + red[3]);
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(256) void dot_partial_kernel(uint64_t n, const double* __restrict__ x,
                                                          const double* __restrict__ y, double* __restrict__ partial) {
    double s = 0.0;
    for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += uint64_t(gridDim.x) * 256) s += x[i] * y[i];
    s = block_sum_256(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void reduce_final_kernel(const double* __restrict__ partial, int n,
                                                           double* __restrict__ out) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += partial[i];
    s = block_sum_256(s);
    if (threadIdx.x == 0) *out = s;
}

hipError_t launch_dot(hipStream_t st, uint64_t n, const double* x, const double* y, double* partial, double* result) {
    hipLaunchKernelGGL(dot_partial_kernel, dim3(RED_BLOCKS), dim3(256), 0, st, n, x, y, partial);
    hipLaunchKernelGGL(reduce_final_kernel, dim3(1), dim3(256), 0, st, partial, RED_BLOCKS, result);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------- batched dot / axpy
// partial[j * RED_BLOCKS + b] = sum over the b-th grid stride of V_j[i] x[i], j < k <= MDOT_MAXK: each
// workgroup reads its part of x once for all k vectors (V_j = V + j * ldv).
__global__ __launch_bounds__(256) void multi_dot_partial_kernel(uint64_t n, int k, const double* __restrict__ V,
                                                                uint64_t ldv, const double* __restrict__ x,
                                                                double* __restrict__ partial) {
    double s[MDOT_MAXK];
#pragma unroll
    for (int j = 0; j < MDOT_MAXK; ++j) s[j] = 0.0;
    for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += uint64_t(gridDim.x) * 256) {
        const double xi = x[i];
#pragma unroll
        for (int j = 0; j < MDOT_MAXK; ++j)
            if (j < k) s[j] += V[uint64_t(j) * ldv + i] * xi;
    }
#pragma unroll
    for (int j = 0; j < MDOT_MAXK; ++j) {
        if (j >= k) break;   // uniform across the workgroup
        const double r = block_sum_256(s[j]);
        if (threadIdx.x == 0) partial[size_t(j) * RED_BLOCKS + blockIdx.x] = r;
    }
}

// out[j] = sum_b partial[j * RED_BLOCKS + b], one workgroup per vector (fixed order)
__global__ __launch_bounds__(256) void multi_reduce_kernel(const double* __restrict__ partial, double* __restrict__ out) {
    double s = 0.0;
    const double* p = partial + size_t(blockIdx.x) * RED_BLOCKS;
    for (int i = threadIdx.x; i < RED_BLOCKS; i += 256) s += p[i];
    s = block_sum_256(s);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

hipError_t launch_multi_dot(hipStream_t st, uint64_t n, int k, const double* V, uint64_t ldv, const double* x,
                            double* partial, double* result) {
    for (int j0 = 0; j0 < k; j0 += MDOT_MAXK) {
        const int kk = k - j0 < MDOT_MAXK ? k - j0 : MDOT_MAXK;
        hipLaunchKernelGGL(multi_dot_partial_kernel, dim3(RED_BLOCKS), dim3(256), 0, st, n, kk, V + uint64_t(j0) * ldv,
                           ldv, x, partial + size_t(j0) * RED_BLOCKS);
    }
    hipLaunchKernelGGL(multi_reduce_kernel, dim3(k), dim3(256), 0, st, partial, result);
    return hipGetLastError();
}

// y[i] += alpha * sum_{j<k} c_j V_j[i], with c from device memory (c_dev) or by value (ch)
struct MdotCoef { double c[MDOT_MAXK]; };
__global__ void multi_axpy_kernel(uint64_t n, int k, double alpha, const double* __restrict__ c_dev, MdotCoef ch,
                                  const double* __restrict__ V, uint64_t ldv, double* __restrict__ y) {
    double c[MDOT_MAXK];
#pragma unroll
    for (int j = 0; j < MDOT_MAXK; ++j) c[j] = j < k ? alpha * (c_dev ? c_dev[j] : ch.c[j]) : 0.0;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
        double acc = y[i];
#pragma unroll
        for (int j = 0; j < MDOT_MAXK; ++j)
            if (j < k) acc += c[j] * V[uint64_t(j) * ldv + i];
        y[i] = acc;
    }
}

hipError_t launch_multi_axpy(hipStream_t st, uint64_t n, int k, double alpha, const double* c_dev, const double* c_host,
                             const double* V, uint64_t ldv, double* y) {
    if (n == 0 || k <= 0) return hipSuccess;
    for (int j0 = 0; j0 < k; j0 += MDOT_MAXK) {
        const int kk = k - j0 < MDOT_MAXK ? k - j0 : MDOT_MAXK;
        MdotCoef ch{};
        if (c_host)
            for (int j = 0; j < kk; ++j) ch.c[j] = c_host[j0 + j];
        hipLaunchKernelGGL(multi_axpy_kernel, dim3(grid_for(n, 256)), dim3(256), 0, st, n, kk, alpha,
                           c_dev ? c_dev + j0 : nullptr, ch, V + uint64_t(j0) * ldv, ldv, y);
    }
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void block_sums_kernel(const uint64_t* __restrict__ blk_off,
                                                         const uint32_t* __restrict__ blk_len,
                                                         const double* __restrict__ data, double* __restrict__ out) {
    const uint64_t off = blk_off[blockIdx.x];
    const uint32_t len = blk_len[blockIdx.x];
    double s = 0.0;
    for (uint32_t i = threadIdx.x; i < len; i += 256) s += data[off + i];
    s = block_sum_256(s);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

hipError_t launch_block_sums(hipStream_t st, const uint64_t* blk_off, const uint32_t* blk_len, const uint32_t*,
                             uint32_t n_blk, const double* data, double* out_partial) {
    if (n_blk == 0) return hipSuccess;
    hipLaunchKernelGGL(block_sums_kernel, dim3(n_blk), dim3(256), 0, st, blk_off, blk_len, data, out_partial);
    return hipGetLastError();
}

}  // namespace mq
