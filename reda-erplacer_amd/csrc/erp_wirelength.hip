// Weighted-average wirelength gradient and HPWL (gfx950), hpwl.rs:139-343.
//
// Nets are net-parallel: two-pin nets one per thread (the reference's own fast
// path, hpwl.rs:173-246), multi-pin nets one per thread up to a degree bound,
// and the high-degree tail one per 64-wide wavefront with shuffle reductions.
// Pins are laid out in net order (slots), so a net's pin records are contiguous
// and a wavefront reads them coalesced; instance positions are gathered as one
// float2 per pin.
//
// The gradient is the reference's  (a + b x) e - (a' + b' x) e'  rewritten with
// coordinates relative to the net's max / min (identical in exact arithmetic):
//   g = e/S (1 + (dx - D/S)/gamma) - e'/S' (1 - (dn - D'/S')/gamma),
//   dx = x - max, e = exp(dx/gamma), S = sum e, D = sum dx e,
//   dn = x - min, e' = exp(-dn/gamma), S' = sum e', D' = sum dn e'.
// which avoids the cancellation of a + b x (both ~ x/gamma) in f32.
#include <climits>

#include "erp_internal.h"
#include "erp_kernels.h"
#include "erp_timer.h"

namespace erp {
namespace {

__device__ __forceinline__ double block_sum_f64(double v) {
    __shared__ double red[kBlock];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    return red[0];
}

// one dimension of a two-pin net
__device__ __forceinline__ float2 wa2(float a, float b, float ig) {
    const float mx = fmaxf(a, b), mn = fminf(a, b);
    const float da = a - mx, db = b - mx;
    const float ea = expf(da * ig), eb = expf(db * ig);
    const float S = ea + eb, off = (da * ea + db * eb) / S;
    const float na = a - mn, nb = b - mn;
    const float fa = expf(-na * ig), fb = expf(-nb * ig);
    const float S2 = fa + fb, off2 = (na * fa + nb * fb) / S2;
    const float ga = ea / S * (1.f + (da - off) * ig) - fa / S2 * (1.f - (na - off2) * ig);
    const float gb = eb / S * (1.f + (db - off) * ig) - fb / S2 * (1.f - (nb - off2) * ig);
    return make_float2(ga, gb);
}

__global__ void __launch_bounds__(kBlock) k_wa_2pin(const int32_t *__restrict__ slot_node,
                                                    const float2 *__restrict__ slot_off,
                                                    const float2 *__restrict__ pos, int64_t n2,
                                                    const double *__restrict__ gamma,
                                                    float2 *__restrict__ slot_grad,
                                                    double *__restrict__ part) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double h = 0.0;
    if (e < n2) {
        const float ig = 1.f / (float)(*gamma);
        const int2 nd = reinterpret_cast<const int2 *>(slot_node)[e];
        const float4 of = reinterpret_cast<const float4 *>(slot_off)[e];
        const float2 p0 = pos[nd.x], p1 = pos[nd.y];
        const float x0 = p0.x + of.x, y0 = p0.y + of.y;
        const float x1 = p1.x + of.z, y1 = p1.y + of.w;
        h = (double)((fmaxf(x0, x1) - fminf(x0, x1)) + (fmaxf(y0, y1) - fminf(y0, y1)));
        const float2 gx = wa2(x0, x1, ig), gy = wa2(y0, y1, ig);
        reinterpret_cast<float4 *>(slot_grad)[e] = make_float4(gx.x, gy.x, gx.y, gy.y);
    }
    const double s = block_sum_f64(h);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kBlock) k_wa_small(const int32_t *__restrict__ slot_node,
                                                     const float2 *__restrict__ slot_off,
                                                     const int32_t *__restrict__ net_start,
                                                     const float2 *__restrict__ pos,
                                                     int64_t e0, int64_t e1,
                                                     const double *__restrict__ gamma,
                                                     float2 *__restrict__ slot_grad,
                                                     double *__restrict__ part) {
    const int64_t e = e0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double h = 0.0;
    if (e < e1) {
        const float ig = 1.f / (float)(*gamma);
        const int s0 = net_start[e], s1 = net_start[e + 1];
        float mx = -INFINITY, mn = INFINITY, my = -INFINITY, ny = INFINITY;
        for (int k = s0; k < s1; ++k) {
            const float2 p = pos[slot_node[k]], o = slot_off[k];
            const float x = p.x + o.x, y = p.y + o.y;
            mx = fmaxf(mx, x); mn = fminf(mn, x);
            my = fmaxf(my, y); ny = fminf(ny, y);
        }
        float Sx = 0.f, Dx = 0.f, Sxn = 0.f, Dxn = 0.f, Sy = 0.f, Dy = 0.f, Syn = 0.f, Dyn = 0.f;
        for (int k = s0; k < s1; ++k) {
            const float2 p = pos[slot_node[k]], o = slot_off[k];
            const float x = p.x + o.x, y = p.y + o.y;
            const float dx = x - mx, nx = x - mn, dy = y - my, nyy = y - ny;
            const float ex = expf(dx * ig), fx = expf(-nx * ig);
            const float ey = expf(dy * ig), fy = expf(-nyy * ig);
            Sx += ex; Dx += dx * ex; Sxn += fx; Dxn += nx * fx;
            Sy += ey; Dy += dy * ey; Syn += fy; Dyn += nyy * fy;
        }
        h = (double)((mx - mn) + (my - ny));
        const float ox = Dx / Sx, oxn = Dxn / Sxn, oy = Dy / Sy, oyn = Dyn / Syn;
        const float iSx = 1.f / Sx, iSxn = 1.f / Sxn, iSy = 1.f / Sy, iSyn = 1.f / Syn;
        for (int k = s0; k < s1; ++k) {
            const float2 p = pos[slot_node[k]], o = slot_off[k];
            const float x = p.x + o.x, y = p.y + o.y;
            const float dx = x - mx, nx = x - mn, dy = y - my, nyy = y - ny;
            const float ex = expf(dx * ig), fx = expf(-nx * ig);
            const float ey = expf(dy * ig), fy = expf(-nyy * ig);
            const float gx = ex * iSx * (1.f + (dx - ox) * ig) - fx * iSxn * (1.f - (nx - oxn) * ig);
            const float gy = ey * iSy * (1.f + (dy - oy) * ig) - fy * iSyn * (1.f - (nyy - oyn) * ig);
            slot_grad[k] = make_float2(gx, gy);
        }
    }
    const double s = block_sum_f64(h);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
    return v;
}
__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off));
    return v;
}
__device__ __forceinline__ float wave_fsum(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// One net per wavefront (4 per workgroup).
__global__ void __launch_bounds__(kBlock) k_wa_large(const int32_t *__restrict__ slot_node,
                                                     const float2 *__restrict__ slot_off,
                                                     const int32_t *__restrict__ net_start,
                                                     const float2 *__restrict__ pos,
                                                     int64_t e0, int64_t e1,
                                                     const double *__restrict__ gamma,
                                                     float2 *__restrict__ slot_grad,
                                                     double *__restrict__ part) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t e = e0 + (int64_t)blockIdx.x * (kBlock / 64) + wave;
    double h = 0.0;
    if (e < e1) {
        const float ig = 1.f / (float)(*gamma);
        const int s0 = net_start[e], s1 = net_start[e + 1];
        float mx = -INFINITY, mn = INFINITY, my = -INFINITY, ny = INFINITY;
        for (int k = s0 + lane; k < s1; k += 64) {
            const float2 p = pos[slot_node[k]], o = slot_off[k];
            const float x = p.x + o.x, y = p.y + o.y;
            mx = fmaxf(mx, x); mn = fminf(mn, x);
            my = fmaxf(my, y); ny = fminf(ny, y);
        }
        mx = wave_max(mx); mn = wave_min(mn); my = wave_max(my); ny = wave_min(ny);
        float Sx = 0.f, Dx = 0.f, Sxn = 0.f, Dxn = 0.f, Sy = 0.f, Dy = 0.f, Syn = 0.f, Dyn = 0.f;
        for (int k = s0 + lane; k < s1; k += 64) {
            const float2 p = pos[slot_node[k]], o = slot_off[k];
            const float x = p.x + o.x, y = p.y + o.y;
            const float dx = x - mx, nx = x - mn, dy = y - my, nyy = y - ny;
            const float ex = expf(dx * ig), fx = expf(-nx * ig);
            const float ey = expf(dy * ig), fy = expf(-nyy * ig);
            Sx += ex; Dx += dx * ex; Sxn += fx; Dxn += nx * fx;
            Sy += ey; Dy += dy * ey; Syn += fy; Dyn += nyy * fy;
        }
        Sx = wave_fsum(Sx); Dx = wave_fsum(Dx); Sxn = wave_fsum(Sxn); Dxn = wave_fsum(Dxn);
        Sy = wave_fsum(Sy); Dy = wave_fsum(Dy); Syn = wave_fsum(Syn); Dyn = wave_fsum(Dyn);
        if (lane == 0) h = (double)((mx - mn) + (my - ny));
        const float ox = Dx / Sx, oxn = Dxn / Sxn, oy = Dy / Sy, oyn = Dyn / Syn;
        const float iSx = 1.f / Sx, iSxn = 1.f / Sxn, iSy = 1.f / Sy, iSyn = 1.f / Syn;
        for (int k = s0 + lane; k < s1; k += 64) {
            const float2 p = pos[slot_node[k]], o = slot_off[k];
            const float x = p.x + o.x, y = p.y + o.y;
            const float dx = x - mx, nx = x - mn, dy = y - my, nyy = y - ny;
            const float ex = expf(dx * ig), fx = expf(-nx * ig);
            const float ey = expf(dy * ig), fy = expf(-nyy * ig);
            const float gx = ex * iSx * (1.f + (dx - ox) * ig) - fx * iSxn * (1.f - (nx - oxn) * ig);
            const float gy = ey * iSy * (1.f + (dy - oy) * ig) - fy * iSyn * (1.f - (nyy - oyn) * ig);
            slot_grad[k] = make_float2(gx, gy);
        }
    }
    const double s = block_sum_f64(h);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Exact HPWL: pin coordinates rounded to integer DBU (nearest-even).
__global__ void __launch_bounds__(kBlock) k_hpwl_dbu(const int32_t *__restrict__ slot_node,
                                                     const float2 *__restrict__ slot_off,
                                                     const int32_t *__restrict__ net_start,
                                                     const float2 *__restrict__ pos,
                                                     int64_t num_nets,
                                                     unsigned long long *__restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    long long v = 0;
    if (e < num_nets) {
        long long mx = LLONG_MIN, mn = LLONG_MAX, my = LLONG_MIN, ny = LLONG_MAX;
        for (int k = net_start[e]; k < net_start[e + 1]; ++k) {
            const float2 p = pos[slot_node[k]], o = slot_off[k];
            const long long x = __float2ll_rn(p.x + o.x), y = __float2ll_rn(p.y + o.y);
            mx = x > mx ? x : mx; mn = x < mn ? x : mn;
            my = y > my ? y : my; ny = y < ny ? y : ny;
        }
        if (net_start[e + 1] > net_start[e]) v = (mx - mn) + (my - ny);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0 && v != 0) atomicAdd(out, (unsigned long long)v);
}

inline unsigned blocks(int64_t count, int per) { return (unsigned)((count + per - 1) / per); }

}  // namespace

int wa_partials(const WaNets &w) {
    return (int)(blocks(w.n2, kBlock) + blocks(w.small_end - w.small_begin, kBlock) +
                 blocks(w.large_end - w.large_begin, kBlock / 64));
}

int launch_wa(const WaNets &w, const float2 *pos, const double *gamma_dev, hipStream_t s,
              KernelTimer *t) {
    double *part = w.hpwl_part;
    const unsigned b2 = blocks(w.n2, kBlock);
    if (b2) {
        if (t) t->begin("wa_2pin");
        hipLaunchKernelGGL(k_wa_2pin, dim3(b2), dim3(kBlock), 0, s, w.slot_node, w.slot_off, pos,
                           w.n2, gamma_dev, w.slot_grad, part);
        if (t) t->end();
        ERP_HIP_TRY(hipGetLastError());
    }
    part += b2;
    const unsigned bs = blocks(w.small_end - w.small_begin, kBlock);
    if (bs) {
        if (t) t->begin("wa_small");
        hipLaunchKernelGGL(k_wa_small, dim3(bs), dim3(kBlock), 0, s, w.slot_node, w.slot_off,
                           w.net_start, pos, w.small_begin, w.small_end, gamma_dev, w.slot_grad,
                           part);
        if (t) t->end();
        ERP_HIP_TRY(hipGetLastError());
    }
    part += bs;
    const unsigned bl = blocks(w.large_end - w.large_begin, kBlock / 64);
    if (bl) {
        if (t) t->begin("wa_large");
        hipLaunchKernelGGL(k_wa_large, dim3(bl), dim3(kBlock), 0, s, w.slot_node, w.slot_off,
                           w.net_start, pos, w.large_begin, w.large_end, gamma_dev, w.slot_grad,
                           part);
        if (t) t->end();
        ERP_HIP_TRY(hipGetLastError());
    }
    return ERP_OK;
}

int launch_hpwl_dbu(const WaNets &w, const float2 *pos, int64_t num_nets, unsigned long long *out,
                    hipStream_t s) {
    if (num_nets <= 0) return ERP_OK;
    hipLaunchKernelGGL(k_hpwl_dbu, dim3(blocks(num_nets, kBlock)), dim3(kBlock), 0, s, w.slot_node,
                       w.slot_off, w.net_start, pos, num_nets, out);
    ERP_HIP_TRY(hipGetLastError());
    return ERP_OK;
}

}  // namespace erp
