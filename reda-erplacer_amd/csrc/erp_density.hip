// Density scatter (gfx950): movable cells and fixed cells into 64-bit
// fixed-point bin accumulators.
//
// The reference accumulates float contributions per bin, sequentially for
// iteration < 200 (density.rs:174-237) and under per-bin Mutexes afterwards
// (density.rs:102-171), whose sum order depends on the thread schedule.  Here
// each contribution is computed with exactly the reference's f32 operations
// (this file is compiled with -ffp-contract=off), rounded once to a multiple of
// 1/S and added with an int64 atomic: integer addition is associative, so the
// map is bit-identical for any launch order and equal to a CPU restatement.
#include "erp_internal.h"
#include "erp_kernels.h"

namespace erp {
namespace {

// density.rs:7-13 triangle_density_function (clamped at 0), xl = 0
__device__ __forceinline__ float tri_overlap(float x, float s, int k, float bs) {
    const float bin_xl = 0.f + (float)k * bs;
    return fmaxf(fminf(x + s, bin_xl + bs) - fmaxf(x, bin_xl), 0.f);
}
// density.rs:16-23 compute_density_overflow against a bin centre
__device__ __forceinline__ float center_overlap(float x, float s, float c, float bs) {
    const float two = 1.f + 1.f;
    return fmaxf(fminf(x + s, c + bs / two) - fmaxf(x, c - bs / two), 0.f);
}
// placement.rs:81-94 bin centre i (last one clipped to the die)
__device__ __forceinline__ float bin_center(int i, float bs, float hi) {
    const float bl = 0.f + (float)i * bs;
    const float bh = fminf(bl + bs, hi);
    return (bl + bh) / 2.f;
}
// bin range [floor(x/bs), ceil((x+s)/bs)+1) capped to [0, nb-1)  (density.rs:202-218)
__device__ __forceinline__ void bin_range(float x, float s, float inv, int nb, int &lo, int &hi) {
    int l = (int)floorf((x - 0.f) * inv);
    int h = (int)ceilf((x - 0.f + s) * inv) + 1;
    lo = l < 0 ? 0 : l;
    hi = h > nb - 1 ? nb - 1 : h;
}
__device__ __forceinline__ int64_t quantize(float c, double scale) {
    return (int64_t)__double2ll_rn((double)c * scale);
}
__device__ __forceinline__ void add_q(int64_t *q, int64_t v) {
    atomicAdd(reinterpret_cast<unsigned long long *>(q), (unsigned long long)v);
}
__device__ __forceinline__ int64_t wave_sum(int64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Movable scatter: one thread per cell.  When every lane of a wavefront covers
// the same bin window (all cells start at the die centre, placement.rs:129-130,
// and stay clustered for many iterations) the wave sums its contributions per
// bin and issues one atomic per bin instead of 64.
__global__ void __launch_bounds__(kBlock) k_scatter(
    const float2 *__restrict__ pos, const float2 *__restrict__ csize,
    const float2 *__restrict__ coff, const float *__restrict__ ratio, int64_t num_cells,
    int m, int n, float bw, float bh, float inv_bw, float inv_bh, double scale,
    int64_t *__restrict__ q) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < num_cells;
    float xi = 0.f, yi = 0.f, sx = 0.f, sy = 0.f, r = 0.f;
    int bxl = 0, bxh = 0, byl = 0, byh = 0;
    if (valid) {
        const float2 p = pos[i], s = csize[i], o = coff[i];
        xi = p.x + o.x;
        yi = p.y + o.y;
        sx = s.x;
        sy = s.y;
        r = ratio[i];
        bin_range(xi, sx, inv_bw, m, bxl, bxh);
        bin_range(yi, sy, inv_bh, n, byl, byh);
    }
    const int l_bxl = __shfl(bxl, 0), l_bxh = __shfl(bxh, 0);
    const int l_byl = __shfl(byl, 0), l_byh = __shfl(byh, 0);
    const bool same = !valid || (bxl == l_bxl && bxh == l_bxh && byl == l_byl && byh == l_byh);
    if (__all(same)) {
        if (!__shfl(valid ? 1 : 0, 0)) return;   // whole wave past the end
        for (int bx = l_bxl; bx < l_bxh; ++bx) {
            const float px = r * tri_overlap(xi, sx, bx, bw);
            for (int by = l_byl; by < l_byh; ++by) {
                const float py = tri_overlap(yi, sy, by, bh);
                int64_t v = valid ? quantize(px * py, scale) : 0;
                v = wave_sum(v);
                if ((threadIdx.x & 63) == 0 && v != 0) add_q(q + (size_t)bx * n + by, v);
            }
        }
        return;
    }
    if (!valid) return;
    for (int bx = bxl; bx < bxh; ++bx) {
        const float px = r * tri_overlap(xi, sx, bx, bw);
        for (int by = byl; by < byh; ++by) {
            const float py = tri_overlap(yi, sy, by, bh);
            const int64_t v = quantize(px * py, scale);
            if (v != 0) add_q(q + (size_t)bx * n + by, v);
        }
    }
}

// Fixed cells (density.rs:26-99): one workgroup per cell over its bin window,
// overlap against bin centres, no clamp offset applied.
__global__ void __launch_bounds__(kBlock) k_fixed_map(const float2 *__restrict__ pos,
                                                      const float2 *__restrict__ csize,
                                                      int m, int n, float bw, float bh,
                                                      float die_w, float die_h, double scale,
                                                      int64_t *__restrict__ q) {
    const int i = blockIdx.x;
    const float xi = pos[i].x, yi = pos[i].y, sxi = csize[i].x, syi = csize[i].y;
    const float inv_x = 1.f / bw, inv_y = 1.f / bh;
    int bxl, bxh, byl, byh;
    bin_range(xi, sxi, inv_x, m, bxl, bxh);
    bin_range(yi, syi, inv_y, n, byl, byh);
    const int wx = bxh - bxl, wy = byh - byl;
    if (wx <= 0 || wy <= 0) return;
    for (int t = threadIdx.x; t < wx * wy; t += blockDim.x) {
        const int bx = bxl + t / wy, by = byl + t % wy;
        const float px = center_overlap(xi, sxi, bin_center(bx, bw, die_w), bw);
        const float py = center_overlap(yi, syi, bin_center(by, bh, die_h), bh);
        const int64_t v = quantize(px * py, scale);
        if (v != 0) add_q(q + (size_t)bx * n + by, v);
    }
}

// q_fix = round(target_density * q_init)  (epotential.rs:138-140 in fixed point)
__global__ void k_scale_fixed(int64_t *__restrict__ q, int64_t nb, double t) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) q[b] = (int64_t)__double2ll_rn(t * (double)q[b]);
}

// Leaf finalize: map = float(q / S) (+ initial), density.rs:164-170 / :231-236
__global__ void k_finalize_leaf(const int64_t *__restrict__ q, const float *__restrict__ initial,
                                float *__restrict__ map, int64_t nb, double inv_scale) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const float v = (float)((double)q[b] * inv_scale);
    map[b] = initial ? v + initial[b] : v;
}

__global__ void k_zero_i64(int64_t *__restrict__ q, int64_t count) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b < count) q[b] = 0;
}

inline unsigned grid_for(int64_t count) { return (unsigned)((count + kBlock - 1) / kBlock); }

}  // namespace

int launch_scatter(const float2 *pos, const float2 *csize, const float2 *coff, const float *ratio,
                   int64_t num_cells, int m, int n, float bw, float bh, double scale, int64_t *q,
                   hipStream_t s) {
    if (num_cells <= 0) return ERP_OK;
    const float inv_bw = 1.f / bw, inv_bh = 1.f / bh;
    hipLaunchKernelGGL(k_scatter, dim3(grid_for(num_cells)), dim3(kBlock), 0, s, pos, csize, coff,
                       ratio, num_cells, m, n, bw, bh, inv_bw, inv_bh, scale, q);
    ERP_HIP_TRY(hipGetLastError());
    return ERP_OK;
}

int launch_fixed_map(const float2 *pos, const float2 *csize, int64_t num_fixed, int m, int n,
                     float bw, float bh, float die_w, float die_h, double scale, int64_t *q,
                     hipStream_t s) {
    if (num_fixed <= 0) return ERP_OK;
    hipLaunchKernelGGL(k_fixed_map, dim3((unsigned)num_fixed), dim3(kBlock), 0, s, pos, csize, m,
                       n, bw, bh, die_w, die_h, scale, q);
    ERP_HIP_TRY(hipGetLastError());
    return ERP_OK;
}

int launch_scale_fixed(int64_t *q, int64_t nb, double t, hipStream_t s) {
    hipLaunchKernelGGL(k_scale_fixed, dim3(grid_for(nb)), dim3(kBlock), 0, s, q, nb, t);
    ERP_HIP_TRY(hipGetLastError());
    return ERP_OK;
}

int launch_finalize_leaf(const int64_t *q, const float *initial, float *map, int64_t nb,
                         double inv_scale, hipStream_t s) {
    hipLaunchKernelGGL(k_finalize_leaf, dim3(grid_for(nb)), dim3(kBlock), 0, s, q, initial, map,
                       nb, inv_scale);
    ERP_HIP_TRY(hipGetLastError());
    return ERP_OK;
}

int launch_zero_i64(int64_t *q, int64_t count, hipStream_t s) {
    if (count <= 0) return ERP_OK;
    hipLaunchKernelGGL(k_zero_i64, dim3(grid_for(count)), dim3(kBlock), 0, s, q, count);
    ERP_HIP_TRY(hipGetLastError());
    return ERP_OK;
}

}  // namespace erp
