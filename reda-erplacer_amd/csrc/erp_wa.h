// Device functions of the weighted-average wirelength kernels (hpwl.rs:139-343),
// shared by the WA launches (erp_wirelength.hip) and the launch that runs them
// beside the first spectral pass (erp_spectral.hip).  Header-only device code:
// every translation unit that includes it gets its own copy.
//
// Nets are net-parallel: two-pin nets one per thread (the reference's own fast
// path, hpwl.rs:173-246), multi-pin nets one per group of lanes up to a degree
// bound, and the high-degree tail one per 64-wide wavefront with shuffle
// reductions.  Pins are laid out in net order (slots), so a net's pin records
// are contiguous and a wavefront reads them coalesced; instance positions are
// gathered as one float2 per pin.
//
// The gradient is the reference's  (a + b x) e - (a' + b' x) e'  rewritten with
// coordinates relative to the net's max / min (identical in exact arithmetic):
//   g = e/S (1 + (dx - D/S)/gamma) - e'/S' (1 - (dn - D'/S')/gamma),
//   dx = x - max, e = exp(dx/gamma), S = sum e, D = sum dx e,
//   dn = x - min, e' = exp(-dn/gamma), S' = sum e', D' = sum dn e'.
// which avoids the cancellation of a + b x (both ~ x/gamma) in f32.
#pragma once
#include <cmath>

#include "erp_kernels.h"

namespace erp {
namespace {

// the workgroup's HPWL partial (valid in thread 0)
__device__ __forceinline__ double wa_block_sum(double v) {
    __shared__ double red[kBlock];
    return block_sum_tree(v, red);
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

// Workgroup `blk` of the two-pin nets.
__device__ __forceinline__ void wa_2pin(const WaNets &w, int64_t blk) {
    const int64_t n2 = w.n2;
    if (blk * kBlock >= n2) return;                    // this job needs fewer workgroups
    const int64_t e = blk * blockDim.x + threadIdx.x;
    double h = 0.0;
    if (e < n2) {
        const float ig = 1.f / (float)(*w.gamma);
        const int2 nd = reinterpret_cast<const int2 *>(w.slot_node)[e];
        const float4 of = reinterpret_cast<const float4 *>(w.slot_off)[e];
        const float2 p0 = w.pos[nd.x], p1 = w.pos[nd.y];
        const float x0 = p0.x + of.x, y0 = p0.y + of.y;
        const float x1 = p1.x + of.z, y1 = p1.y + of.w;
        h = (double)((fmaxf(x0, x1) - fminf(x0, x1)) + (fmaxf(y0, y1) - fminf(y0, y1)));
        const float2 gx = wa2(x0, x1, ig), gy = wa2(y0, y1, ig);
        reinterpret_cast<float4 *>(w.slot_grad)[e] = make_float4(gx.x, gy.x, gx.y, gy.y);
    }
    const double sum = wa_block_sum(h);
    if (threadIdx.x == 0) w.hpwl_part[w.part_offset[0] + blk] = sum;
}

// One net per group of G lanes (G divides 64), each lane holding up to K of the
// net's pins in registers: all pin loads of a net are issued together, the
// three passes of hpwl.rs:248-334 run on registers, and the max/min and the
// exponential sums are reduced with shuffles inside the group.
template <int G, int K>
__device__ __forceinline__ void wa_group(const WaNets &w, int bucket, int64_t blk) {
    const int64_t e0 = w.bucket_begin[bucket], e1 = w.bucket_begin[bucket + 1];
    if (blk * kBlock >= (e1 - e0) * G) return;   // uniform per workgroup
    const int32_t *__restrict__ slot_node = w.slot_node;
    const float2 *__restrict__ slot_off = w.slot_off;
    const int32_t *__restrict__ net_start = w.net_start;
    const float2 *__restrict__ pos = w.pos;
    float2 *__restrict__ slot_grad = w.slot_grad;
    const double *gamma = w.gamma;
    const int64_t t = blk * blockDim.x + threadIdx.x;
    const int64_t e = e0 + t / G;
    const int r = (int)(t & (G - 1));
    const bool valid = e < e1;
    const float ig = 1.f / (float)(*gamma);
    int s0 = 0, s1 = 0;
    if (valid) {
        s0 = net_start[e];
        s1 = net_start[e + 1];
    }
    int nd[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int k = s0 + r + G * j;
        nd[j] = k < s1 ? slot_node[k] : -1;
    }
    float x[K], y[K];
    float mx = -INFINITY, mn = INFINITY, my = -INFINITY, ny = INFINITY;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        x[j] = 0.f;
        y[j] = 0.f;
        if (nd[j] >= 0) {
            const int k = s0 + r + G * j;
            const float2 p = pos[nd[j]], o = slot_off[k];
            x[j] = p.x + o.x;
            y[j] = p.y + o.y;
            mx = fmaxf(mx, x[j]); mn = fminf(mn, x[j]);
            my = fmaxf(my, y[j]); ny = fminf(ny, y[j]);
        }
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        mx = fmaxf(mx, __shfl_xor(mx, off)); mn = fminf(mn, __shfl_xor(mn, off));
        my = fmaxf(my, __shfl_xor(my, off)); ny = fminf(ny, __shfl_xor(ny, off));
    }
    float ex[K], fx[K], ey[K], fy[K];
    float Sx = 0.f, Dx = 0.f, Sxn = 0.f, Dxn = 0.f, Sy = 0.f, Dy = 0.f, Syn = 0.f, Dyn = 0.f;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        ex[j] = fx[j] = ey[j] = fy[j] = 0.f;
        if (nd[j] >= 0) {
            const float dx = x[j] - mx, nx = x[j] - mn, dy = y[j] - my, nyy = y[j] - ny;
            ex[j] = expf(dx * ig); fx[j] = expf(-nx * ig);
            ey[j] = expf(dy * ig); fy[j] = expf(-nyy * ig);
            Sx += ex[j]; Dx += dx * ex[j]; Sxn += fx[j]; Dxn += nx * fx[j];
            Sy += ey[j]; Dy += dy * ey[j]; Syn += fy[j]; Dyn += nyy * fy[j];
        }
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        Sx += __shfl_xor(Sx, off); Dx += __shfl_xor(Dx, off);
        Sxn += __shfl_xor(Sxn, off); Dxn += __shfl_xor(Dxn, off);
        Sy += __shfl_xor(Sy, off); Dy += __shfl_xor(Dy, off);
        Syn += __shfl_xor(Syn, off); Dyn += __shfl_xor(Dyn, off);
    }
    double h = 0.0;
    if (valid) {
        if (r == 0) h = (double)((mx - mn) + (my - ny));
        const float ox = Dx / Sx, oxn = Dxn / Sxn, oy = Dy / Sy, oyn = Dyn / Syn;
        const float iSx = 1.f / Sx, iSxn = 1.f / Sxn, iSy = 1.f / Sy, iSyn = 1.f / Syn;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            if (nd[j] >= 0) {
                const float dx = x[j] - mx, nx = x[j] - mn, dy = y[j] - my, nyy = y[j] - ny;
                const float gx = ex[j] * iSx * (1.f + (dx - ox) * ig) - fx[j] * iSxn * (1.f - (nx - oxn) * ig);
                const float gy = ey[j] * iSy * (1.f + (dy - oy) * ig) - fy[j] * iSyn * (1.f - (nyy - oyn) * ig);
                slot_grad[s0 + r + G * j] = make_float2(gx, gy);
            }
        }
    }
    const double sum = wa_block_sum(h);
    if (threadIdx.x == 0) w.hpwl_part[w.part_offset[bucket + 1] + blk] = sum;
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
__device__ __forceinline__ void wa_large(const WaNets &w, int bucket, int64_t blk) {
    const int64_t e0 = w.bucket_begin[bucket], e1 = w.bucket_begin[bucket + 1];
    if (blk * (kBlock / 64) >= e1 - e0) return;  // uniform per workgroup
    const int32_t *__restrict__ slot_node = w.slot_node;
    const float2 *__restrict__ slot_off = w.slot_off;
    const int32_t *__restrict__ net_start = w.net_start;
    const float2 *__restrict__ pos = w.pos;
    float2 *__restrict__ slot_grad = w.slot_grad;
    const double *gamma = w.gamma;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t e = e0 + blk * (kBlock / 64) + wave;
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
    const double sum = wa_block_sum(h);
    if (threadIdx.x == 0) w.hpwl_part[w.part_offset[bucket + 1] + blk] = sum;
}



// A grid smaller than the plan loops over its workgroups (grid stride), as do
// the WA workgroups that share a spectral pass's launch.

// A range of the WA plan's workgroups run inside another launch.
struct WaShare {
    WaPlan plan;
    unsigned first, last;   // plan workgroups [first, last)
};

// The workgroups first, first + stride, ... below `last` of a WA plan: segment
// lookup, then the bucket's net group (the same work and partial as the
// workgroup of that index in a full launch).
__device__ __forceinline__ void wa_narrow_blocks(const WaNets &w, const WaPlan &plan, unsigned first,
                                                 unsigned stride, unsigned last) {
    for (unsigned vb = first; vb < last; vb += stride) {
        int sg = 0;
        while (sg + 1 < plan.nseg && vb >= plan.begin[sg + 1]) ++sg;
        const int64_t blk = (int64_t)vb - plan.begin[sg];
        switch (plan.bucket[sg]) {
        case -1: wa_2pin(w, blk); break;
        case 0: wa_group<1, 4>(w, 0, blk); break;
        case 1: wa_group<2, 4>(w, 1, blk); break;
        case 2: wa_group<4, 4>(w, 2, blk); break;
        case 3: wa_group<8, 4>(w, 3, blk); break;
        case 4: wa_group<16, 4>(w, 4, blk); break;
        case 5: wa_group<32, 4>(w, 5, blk); break;
        default: wa_group<64, 4>(w, 6, blk); break;
        }
    }
}

}  // namespace
}  // namespace erp
