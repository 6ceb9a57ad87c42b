// Per-cell kernels (field gather, WA pin->cell gather, preconditioned Nesterov
// step) and the device-side scalar state (Nesterov momentum/step size and the
// Scheduler), gfx950.
#include <climits>

#include "erp_internal.h"
#include "erp_kernels.h"

namespace erp {
namespace {

// Workgroup sum in a fixed order: butterfly inside each wavefront, then the
// wavefront results in order.  red: kBlock / 64 doubles of LDS.
__device__ __forceinline__ double block_sum_f64(double v, double *red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) r += red[w];
    __syncthreads();
    return r;
}

// Both sums of block_sum_f64 for two values at once (the same order, one
// barrier); valid in thread 0, and red is not free again until a barrier.
__device__ __forceinline__ void block_sum2_f64(double &a, double &b, double (*red)[kBlock / 64]) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = a;
        red[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    double ra = 0.0, rb = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        ra += red[0][w];
        rb += red[1][w];
    }
    a = ra;
    b = rb;
}

// eforce.rs:6-12: overlap of [x, x+s] with bin k, NOT clamped at 0
__device__ __forceinline__ float tri_unclamped(float x, float s, int k, float bs) {
    const float bin_xl = 0.f + (float)k * bs;
    return fminf(x + s, bin_xl + bs) - fmaxf(x, bin_xl);
}

// The bins [xl, xh) x [yl, yh) of a cell's field integral (eforce.rs:40-90):
// [max(trunc(x/bw), 0), min(trunc((x+s)/bw) + 1, M)) per axis, given the cell's
// clamped lower-left corner (nx, ny) = position + clamp offset and clamped size.
__device__ __forceinline__ void field_window(const CellBufs &c, float nx, float ny, float2 sz,
                                             float inv_bw, float inv_bh, int &xl, int &xh,
                                             int &yl, int &yh) {
    xl = (int)((nx - 0.f) * inv_bw);
    xl = xl < 0 ? 0 : xl;
    xh = (int)((nx + sz.x - 0.f) * inv_bw) + 1;
    xh = xh > c.m ? c.m : xh;
    yl = (int)((ny - 0.f) * inv_bh);
    yl = yl < 0 ? 0 : yl;
    yh = (int)((ny + sz.y - 0.f) * inv_bh) + 1;
    yh = yh > c.n ? c.n : yh;
}

// eforce.rs:40-90 for one cell: -ratio * sum_b overlap_x overlap_y field_b over
// its window, summed row by row (k), each row in increasing h.
__device__ __forceinline__ float2 efield_at(const CellBufs &c, float nx, float ny, float2 sz,
                                            float r, float inv_bw, float inv_bh) {
    int xl, xh, yl, yh;
    field_window(c, nx, ny, sz, inv_bw, inv_bh, xl, xh, yl, yh);
    float gx = 0.f, gy = 0.f;
    if (yh - yl <= 5) {
        // the window's columns lie in the 4-bin block at or below yl and the
        // next one: per row, two aligned 16-byte loads per block (the next
        // block only when the window reaches it)
        const int hb = yl & ~3;
        const bool two = yh > hb + 4;
        for (int k = xl; k < xh; ++k) {
            const float px = tri_unclamped(nx, sz.x, k, c.bw);
            const float4 *b0 = reinterpret_cast<const float4 *>(c.fxy + fxy_index(k, hb, c.n));
            const float4 v0 = gld(b0), v1 = gld(b0 + 1);
            float4 v2 = make_float4(0.f, 0.f, 0.f, 0.f), v3 = v2;
            if (two) {
                const float4 *b1 = reinterpret_cast<const float4 *>(c.fxy + fxy_index(k, hb + 4, c.n));
                v2 = gld(b1);
                v3 = gld(b1 + 1);
            }
            const float vx[8] = {v0.x, v0.z, v1.x, v1.z, v2.x, v2.z, v3.x, v3.z};
            const float vy[8] = {v0.y, v0.w, v1.y, v1.w, v2.y, v2.w, v3.y, v3.w};
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int h = hb + j;
                if (h >= yl && h < yh) {
                    const float area = px * tri_unclamped(ny, sz.y, h, c.bh);
                    gx += area * vx[j];
                    gy += area * vy[j];
                }
            }
        }
    } else {
        for (int k = xl; k < xh; ++k) {
            const float px = tri_unclamped(nx, sz.x, k, c.bw);
            for (int h = yl; h < yh; ++h) {
                const float2 v = gld(c.fxy + fxy_index(k, h, c.n));
                const float area = px * tri_unclamped(ny, sz.y, h, c.bh);
                gx += area * v.x;
                gy += area * v.y;
            }
        }
    }
    return make_float2(-gx * r, -gy * r);
}

// the field force of cell i in id order (its inputs from pos, cgeo, cprop)
__device__ __forceinline__ float2 efield_cell(const CellBufs &c, int64_t i, float inv_bw,
                                              float inv_bh) {
    const float2 p = gld(c.pos + i);
    const float4 gq = gld(c.cgeo + i);
    const float r = gld(c.cprop + i).x;
    return efield_at(c, p.x + gq.x, p.y + gq.y, make_float2(gq.z, gq.w), r, inv_bw, inv_bh);
}

// hpwl.rs:120-134: sum of the cell's pin gradients in macro_2_pins order
// (slot indices and gradients of four pins are loaded before they are summed).
__device__ __forceinline__ float2 pin2cell(const CellBufs &c, int64_t i) {
    const int k0 = gld(c.cstart + i), k1 = gld(c.cstart + i + 1);
    float gx = 0.f, gy = 0.f;
    for (int k = k0; k < k1; k += 4) {
        int sl[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) sl[j] = k + j < k1 ? gld(c.cslots + k + j) : -1;
        float2 g[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) g[j] = sl[j] >= 0 ? gld(c.slot_grad + sl[j]) : make_float2(0.f, 0.f);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (sl[j] >= 0) {
                gx += g[j].x;
                gy += g[j].y;
            }
        }
    }
    return make_float2(gx, gy);
}

// nesterov.rs:75-95: preconditioned combined gradient
__device__ __forceinline__ float2 combine(const CellBufs &c, int64_t i, float2 gwl, float2 gd,
                                          float dw) {
    const float4 pr = gld(c.cprop + i);
    const float inv_p = 1.f / fminf(pr.y + dw * pr.z, 1.f);
    return make_float2((gwl.x + dw * gd.x) * inv_p, (gwl.y + dw * gd.y) * inv_p);
}

#define ERP_CELL_JOB()                                                              \
    if (gridDim.y > 1 && !jd[blockIdx.y].active) return; /* empty batch slot */    \
    const CellBufs &c = jd[blockIdx.y].cb;                                         \
    const int64_t nblk = (c.C + kBlock - 1) / kBlock;                              \
    if ((int64_t)blockIdx.x >= nblk) return; /* uniform per workgroup */           \
    const int64_t blk = xcd_block(blockIdx.x, nblk);                               \
    const int64_t i = blk * kBlock + threadIdx.x;

__global__ void __launch_bounds__(kBlock) k_pin2cell(const JobDesc *__restrict__ jd) {
    ERP_CELL_JOB();
    if (i < c.C) c.gwl[i] = pin2cell(c, i);
}

__global__ void __launch_bounds__(kBlock) k_efield(const JobDesc *__restrict__ jd) {
    ERP_CELL_JOB();
    if (i < c.C) c.gd[i] = efield_cell(c, i, 1.f / c.bw, 1.f / c.bh);
}

// The field gather in the spatial order of this iteration's sort: sorted
// record i holds the cell's clamped corner and size (rec) and its ratio and id
// (aux), the values efield_cell derives from pos, cgeo and cprop (the same f32
// sums).  A workgroup's windows then lie in one region of the map: it copies
// the 4 x 4-bin blocks that its cells' windows touch into LDS once (whole
// 128-byte lines, coalesced) and every cell sums its window from there, in the
// order of efield_at.  Cells with windows above kFieldSmall bins a side, and
// every cell of a workgroup whose box exceeds kFieldBoxBlocks, read the map
// directly.  The force is written to the cell's id.
constexpr int kFieldBoxBlocks = 128;   // LDS box: 4 x 4-bin blocks (16 KB)
constexpr int kFieldSmall = 8;         // window side (bins) of the cells that read the box

__device__ __forceinline__ void block_minmax(int &lo, int &hi, int (*red)[kBlock / 64], int slot) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, __shfl_xor(lo, off));
        hi = max(hi, __shfl_xor(hi, off));
    }
    if ((threadIdx.x & 63) == 0) {
        red[slot][threadIdx.x >> 6] = lo;
        red[slot + 1][threadIdx.x >> 6] = hi;
    }
}

__global__ void __launch_bounds__(kBlock) k_l1(const JobDesc *__restrict__ jd) {
    __shared__ double red[kBlock / 64];
    ERP_CELL_JOB();
    double a = 0.0, b = 0.0;
    if (i < c.C) {
        const float2 w = c.gwl[i], d = c.gd[i];
        a = (double)fabsf(w.x) + (double)fabsf(w.y);
        b = (double)fabsf(d.x) + (double)fabsf(d.y);
    }
    a = block_sum_f64(a, red);
    b = block_sum_f64(b, red);
    if (threadIdx.x == 0) {
        c.part_a[blk] = a;
        c.part_b[blk] = b;
    }
}

// g and the BB dot-product partials sum |dx|^2, sum |dg|^2 (nesterov.rs:87-93)
__device__ __forceinline__ void bb_terms(const CellBufs &c, int64_t i, float2 g, double &a,
                                         double &b) {
    const float2 p = gld(c.pos + i), pp = gld(c.prev_pos + i), pg = gld(c.prev_g + i);
    const double dx = (double)(p.x - pp.x), dy = (double)(p.y - pp.y);
    const double gx = (double)(g.x - pg.x), gy = (double)(g.y - pg.y);
    a = dx * dx + dy * dy;
    b = gx * gx + gy * gy;
}

__global__ void __launch_bounds__(kBlock) k_combine(const JobDesc *__restrict__ jd) {
    __shared__ double red[kBlock / 64];
    ERP_CELL_JOB();
    const DevState *__restrict__ st = c.st;
    double a = 0.0, b = 0.0;
    if (i < c.C) {
        const float2 g = combine(c, i, c.gwl[i], c.gd[i], (float)st->density_weight);
        c.g[i] = g;
        bb_terms(c, i, g, a, b);
    }
    a = block_sum_f64(a, red);
    b = block_sum_f64(b, red);
    if (threadIdx.x == 0) {
        c.part_a[blk] = a;
        c.part_b[blk] = b;
    }
}

// READ_GD: the field force was gathered already (k_field_sorted) and is read
// from gd.
template <bool STORE_PARTS, bool READ_GD>
__global__ void __launch_bounds__(kBlock) k_cell_fused(const JobDesc *__restrict__ jd) {
    __shared__ double red[2][kBlock / 64];
    ERP_CELL_JOB();
    const DevState *__restrict__ st = c.st;
    double a = 0.0, b = 0.0;
    if (i < c.C) {
        const float2 gwl = pin2cell(c, i);
        const float2 gd = READ_GD ? gld(c.gd + i) : efield_cell(c, i, 1.f / c.bw, 1.f / c.bh);
        const float2 g = combine(c, i, gwl, gd, (float)st->density_weight);
        c.g[i] = g;
        if (STORE_PARTS) {
            c.gwl[i] = gwl;
            if (!READ_GD) c.gd[i] = gd;
        }
        bb_terms(c, i, g, a, b);
    }
    block_sum2_f64(a, b, red);
    if (threadIdx.x == 0) {
        c.part_a[blk] = a;
        c.part_b[blk] = b;
    }
}

// nesterov.rs:111-145
__global__ void __launch_bounds__(kBlock) k_update(const JobDesc *__restrict__ jd) {
    ERP_CELL_JOB();
    const DevState *__restrict__ st = c.st;
    if (i >= c.C) return;
    const float alpha = (float)st->alpha, beta = (float)st->beta;
    const float2 p = c.pos[i], pp = c.prev_pos[i], g = c.g[i], hi = c.hi[i];
    const float xh = p.x + beta * (p.x - pp.x);
    const float yh = p.y + beta * (p.y - pp.y);
    float nx = xh - alpha * g.x, ny = yh - alpha * g.y;
    if (nx < 0.f) nx = 0.f;      // f32::clamp(0, W - w - 1)
    if (nx > hi.x) nx = hi.x;
    if (ny < 0.f) ny = 0.f;
    if (ny > hi.y) ny = hi.y;
    c.prev_pos[i] = p;
    c.pos_out[i] = make_float2(nx, ny);
    c.prev_g[i] = g;
}

// Deterministic sums of the partial arrays, then the FP64 Nesterov and
// scheduler updates, by one workgroup per job.  The sums are those of
// kScalarThreads virtual threads: virtual thread v adds the entries
// k = v (mod kScalarThreads) in increasing k, the lanes of each virtual
// wavefront are summed by butterfly, then the wavefronts in order.  A workgroup
// of T threads carries kScalarThreads / T virtual threads per thread (thread t
// the virtual threads t + T j, which sit in lane t mod 64 of virtual wavefront
// t / 64 + (T / 64) j), so the sums have the same bits for every T.
constexpr int kScalarThreads = 1024;

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}
__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// LOADS: loads per array in flight per thread (the order of the sums does not
// depend on it).
template <int T, int LOADS>
__device__ __forceinline__ void scalar_body(const ScalarArgs &a, int ops) {
    constexpr int V = kScalarThreads / T;            // virtual threads per thread
    constexpr int U = LOADS / V;                     // strides per round
    static_assert(V * T == kScalarThreads && U * V == LOADS && T % 64 == 0 && T >= 256,
                  "scalar workgroup shape");
    constexpr int kWaves = kScalarThreads / 64;      // virtual wavefronts
    __shared__ double redd[4][kWaves];
    __shared__ int64_t redi[kWaves];
    __shared__ double tot[4];
    __shared__ int64_t toti;
    __shared__ double pw[3];
    // the device state is staged in LDS: its words are loaded together with the
    // partials, updated there by one thread and stored back together
    constexpr int kStateWords = (int)(sizeof(DevState) / 8);
    static_assert(sizeof(DevState) % 8 == 0 && kStateWords <= T, "DevState is copied as 8-byte words");
    __shared__ unsigned long long ssw[kStateWords];
    const unsigned long long *gsw = reinterpret_cast<const unsigned long long *>(a.st);
    const unsigned long long sword = threadIdx.x < kStateWords ? gld(gsw + threadIdx.x) : 0ull;
    double h[V], e[V], pa[V], pb[V];
    int64_t o[V];
#pragma unroll
    for (int j = 0; j < V; ++j) {
        h[j] = e[j] = pa[j] = pb[j] = 0.0;
        o[j] = 0;
    }
    const int nh = (ops & SC_HPWL) ? a.n_hpwl : 0;
    const int no = (ops & SC_POT) ? a.n_ovf : 0;
    const int ne = (ops & SC_POT) ? a.n_energy : 0;
    const int nc = (ops & (SC_L1 | SC_NEST)) ? a.n_cell : 0;
    const int most = max(max(nh, no), max(ne, nc));
    const double *hp = a.hpwl_part, *ep = a.energy_part, *ap = a.part_a, *bp = a.part_b;
    const int64_t *op = a.ovf_part;
    // U strides of every array (LOADS loads per thread) are issued before any
    // is added
    for (int k0 = 0; k0 < most; k0 += U * kScalarThreads) {
        double vh[U][V], ve[U][V], va[U][V], vb[U][V];
        int64_t vo[U][V];
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const int k = k0 + u * kScalarThreads + (int)threadIdx.x + T * j;
                vh[u][j] = k < nh ? gld(hp + k) : 0.0;
                vo[u][j] = k < no ? gld(op + k) : 0;
                ve[u][j] = k < ne ? gld(ep + k) : 0.0;
                va[u][j] = k < nc ? gld(ap + k) : 0.0;
                vb[u][j] = k < nc ? gld(bp + k) : 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int j = 0; j < V; ++j) {
                h[j] += vh[u][j]; o[j] += vo[u][j]; e[j] += ve[u][j]; pa[j] += va[u][j]; pb[j] += vb[u][j];
            }
        }
    }
    if (threadIdx.x < kStateWords) ssw[threadIdx.x] = sword;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < V; ++j) {
        h[j] = wave_sum_f64(h[j]); e[j] = wave_sum_f64(e[j]);
        pa[j] = wave_sum_f64(pa[j]); pb[j] = wave_sum_f64(pb[j]);
        o[j] = wave_sum_i64(o[j]);
        if (lane == 0) {
            const int w = wave + (T / 64) * j;
            redd[0][w] = h[j]; redd[1][w] = e[j]; redd[2][w] = pa[j]; redd[3][w] = pb[j];
            redi[w] = o[j];
        }
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double v = 0.0;
        for (int w = 0; w < kWaves; ++w) v += redd[threadIdx.x][w];
        tot[threadIdx.x] = v;
    } else if (threadIdx.x == 4) {
        int64_t v = 0;
        for (int w = 0; w < kWaves; ++w) v += redi[w];
        toti = v;
    }
    __syncthreads();
    const double hsum = tot[0], esum = tot[1], asum = tot[2], bsum = tot[3];
    const int64_t osum = toti;
    DevState *sp = reinterpret_cast<DevState *>(ssw);
    const bool sched = (ops & (SC_SCHED | SC_SCHED_STAGE)) != 0;
    // scheduler.rs:92-130 needs three powers; lane 0 of waves 1..3 evaluates them
    // concurrently from the values the update below will see.
    if (sched && lane == 0 && wave >= 1 && wave <= 3) {
        const DevState &s = *sp;
        const double hcur = (ops & SC_HPWL) ? hsum : s.hpwl;
        const double ocur = (ops & SC_POT) ? (double)osum * a.inv_scale_over_area : s.overflow;
        if (wave == 1) {
            pw[0] = pow(0.9999, (double)(int32_t)(s.iteration + 1));
        } else if (wave == 2) {
            const double init = s.hpwl_has_initial ? s.hpwl_initial : hcur;
            pw[1] = pow(1.05, -(hcur - s.hpwl_cur) / init);
        } else {
            pw[2] = pow(10.0, (ocur - 0.1) * 20.0 / 9.0 - 1.0);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        DevState &s = *sp;
        if (ops & SC_HPWL) s.hpwl = hsum;
        if (ops & SC_POT) {
            s.overflow = (double)osum * a.inv_scale_over_area;
            s.energy = esum;
        }
        if (ops & SC_L1) {                       // nesterov.rs:47-55
            s.l1_wl = asum;
            s.l1_d = bsum;
            s.density_weight = s.density_weight * (asum / bsum);
        }
        if (ops & SC_NEST) {
            // update_momentum (nesterov.rs:148-158), then the BB step (:98-100)
            const double a_prev = s.alpha;
            const double na = (1.0 + sqrt(4.0 * a_prev * a_prev + 1.0)) / 2.0;
            s.beta = (a_prev - 1.0) / na;
            s.alpha = na;
            s.num = asum;
            s.den = bsum;
            if (s.iteration > 0 && bsum > 0.0) s.alpha = sqrt(asum / bsum);
            s.dw_used = s.density_weight;
        }
        if (sched) {                             // scheduler.rs:68-130 but the iteration count
            if (!s.hpwl_has_initial) { s.hpwl_has_initial = 1; s.hpwl_initial = s.hpwl; }
            s.hpwl_prev = s.hpwl_cur;
            s.hpwl_cur = s.hpwl;
            if (!s.ovf_has_initial) { s.ovf_has_initial = 1; s.ovf_initial = s.overflow; }
            s.ovf_prev = s.ovf_cur;
            s.ovf_cur = s.overflow;
            const double delta = s.hpwl_cur - s.hpwl_prev;
            double power;
            if (delta < 0.0) {
                power = pw[0];
                power = power > 0.98 ? power : 0.98;
            } else {
                power = pw[1];
                power = power < 0.95 ? 0.95 : (power > 1.05 ? 1.05 : power);
            }
            s.dw_next = s.density_weight * (1.05 * power);
            double coef = pw[2];
            coef = coef < 0.05 ? 0.05 : (coef > 10.0 ? 10.0 : coef);
            s.gamma_next = s.base_gamma * coef;
            s.should_stop = s.hpwl_prev != 0.0 && s.hpwl_cur != 0.0 && s.hpwl_prev > s.hpwl_cur * 1.5;
        }
        if (ops & (SC_SCHED | SC_COMMIT)) {      // scheduler.rs:69, and the values staged above
            s.iteration += 1;
            s.density_weight = s.dw_next;
            s.gamma = s.gamma_next;
        }
        const double chk = s.hpwl + s.overflow + s.energy + s.alpha + s.beta + s.density_weight + s.gamma;
        s.nonfinite = isfinite(chk) ? 0 : 1;
    }
    __syncthreads();
    if (threadIdx.x < kStateWords) reinterpret_cast<unsigned long long *>(a.st)[threadIdx.x] = ssw[threadIdx.x];
}

__global__ void __launch_bounds__(kScalarThreads) k_scalar(const JobDesc *__restrict__ jd, int ops) {
    if (gridDim.x > 1 && !jd[blockIdx.x].active) return;   // an empty slot of a batch
    scalar_body<kScalarThreads, 8>(jd[blockIdx.x].sa, ops);
}

// WITH_SCHED: workgroup 0 runs the step's HPWL, overflow and energy sums and
// the staged half of the scheduler (SC_HPWL | SC_POT | SC_SCHED_STAGE) beside
// the gather, workgroups 1-7 are empty, and the gather's workgroups start at 8
// (so that launch index mod 8 still names the XCD).  Every partial those sums
// read is complete before this launch (WA and pass 1 to 3).
template <bool WITH_SCHED>
__global__ void __launch_bounds__(kBlock) k_field_sorted(const JobDesc *__restrict__ jd) {
    __shared__ float4 box[kFieldBoxBlocks * 8];
    __shared__ int red[4][kBlock / 64];
    if (gridDim.y > 1 && !jd[blockIdx.y].active) return;   // an empty slot of a batch
    int64_t bx = blockIdx.x;
    if (WITH_SCHED) {
        if (bx < 8) {                                      // uniform per workgroup
            if (bx == 0) scalar_body<kBlock, 4>(jd[blockIdx.y].sa, SC_HPWL | SC_POT | SC_SCHED_STAGE);
            return;
        }
        bx -= 8;
    }
    const CellBufs &c = jd[blockIdx.y].cb;
    const int64_t nblk = (c.C + kBlock - 1) / kBlock;
    if (bx >= nblk) return;                                 // uniform per workgroup
    const int64_t i = xcd_block(bx, nblk) * kBlock + threadIdx.x;
    const bool valid = i < c.C;
    const SortBufs &sb = jd[blockIdx.y].sb;
    float4 r4 = make_float4(0.f, 0.f, 0.f, 0.f), a4 = r4;
    if (valid) {
        r4 = gld(sb.rec + i);
        a4 = gld(sb.aux + i);
    }
    const float inv_bw = 1.f / c.bw, inv_bh = 1.f / c.bh;
    const float2 sz = make_float2(r4.z, r4.w);
    int xl = 0, xh = 0, yl = 0, yh = 0;
    if (valid) field_window(c, r4.x, r4.y, sz, inv_bw, inv_bh, xl, xh, yl, yh);
    const bool small = valid && xh > xl && yh > yl && xh - xl <= kFieldSmall && yh - yl <= kFieldSmall;
    // the box of the small windows, in 4 x 4-bin blocks [lx, hx] x [ly, hy]
    int lx = small ? xl >> 2 : INT_MAX, hx = small ? (xh - 1) >> 2 : INT_MIN;
    int ly = small ? yl >> 2 : INT_MAX, hy = small ? (yh - 1) >> 2 : INT_MIN;
    block_minmax(lx, hx, red, 0);
    block_minmax(ly, hy, red, 2);
    __syncthreads();
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        lx = min(lx, red[0][w]);
        hx = max(hx, red[1][w]);
        ly = min(ly, red[2][w]);
        hy = max(hy, red[3][w]);
    }
    const int wxb = hx >= lx ? hx - lx + 1 : 0, wyb = hy >= ly ? hy - ly + 1 : 0;
    const bool use_box = wxb > 0 && wyb > 0 && wxb * wyb <= kFieldBoxBlocks;   // uniform
    if (use_box) {
        const int nbr = (c.n + 3) >> 2;                    // blocks per block row of the map
        const float4 *src = reinterpret_cast<const float4 *>(c.fxy);
        for (int t = threadIdx.x; t < wxb * wyb * 8; t += kBlock) {
            const int b = t >> 3;
            const int kb = lx + b / wyb, hb = ly + b % wyb;
            box[t] = gld(src + ((size_t)kb * nbr + hb) * 8 + (t & 7));
        }
        __syncthreads();
    }
    if (!valid) return;
    float2 gd;
    if (use_box && small) {
        const float2 *bins = reinterpret_cast<const float2 *>(box);
        float gx = 0.f, gy = 0.f;
        for (int k = xl; k < xh; ++k) {
            const float px = tri_unclamped(r4.x, sz.x, k, c.bw);
            const int row = ((k >> 2) - lx) * wyb;
            for (int h = yl; h < yh; ++h) {
                const float2 v = bins[(row + (h >> 2) - ly) * 16 + (((k & 3) << 2) | (h & 3))];
                const float area = px * tri_unclamped(r4.y, sz.y, h, c.bh);
                gx += area * v.x;
                gy += area * v.y;
            }
        }
        gd = make_float2(-gx * a4.z, -gy * a4.z);
    } else {
        gd = efield_at(c, r4.x, r4.y, sz, a4.z, inv_bw, inv_bh);
    }
    c.gd[__float_as_int(a4.w)] = gd;
}

// (hpwl, overflow, nonfinite) of every job of a batch into one array.
__global__ void k_gather_state(const JobDesc *__restrict__ jd, int nj, double *__restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nj || !jd[j].active) return;
    const DevState *st = jd[j].sa.st;
    out[3 * j] = st->hpwl;
    out[3 * j + 1] = st->overflow;
    out[3 * j + 2] = (double)st->nonfinite;
}

__global__ void k_reset_state(DevState *st, double gamma0, double dw0) {
    DevState s = {};
    s.iteration = 0;
    s.gamma = gamma0;
    s.base_gamma = gamma0;
    s.density_weight = dw0;
    s.alpha = 1.0;
    s.beta = 0.0;
    *st = s;
}

// PlacementSolution::new (placement.rs:109-139) and Nesterov::new (nesterov.rs:15-24)
__global__ void k_init_positions(float2 *pos, const float *w, const float *h, const float *x,
                                 const float *y, int64_t N, int64_t C, float die_w, float die_h,
                                 float2 *prev_pos, float2 *prev_g) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float two = 2.f;
    if (i >= C) {
        pos[i] = make_float2(x[i], y[i]);
    } else {
        pos[i] = make_float2(die_w / two - w[i] / two, die_h / two - h[i] / two);
        prev_pos[i] = make_float2(x[i], y[i]);
        prev_g[i] = make_float2(0.f, 0.f);
    }
}

// epotential.rs:83-103 and the per-cell constants of nesterov.rs:75-79,131-132
__global__ void k_setup_cells(const float *w, const float *h, const float *npins, int64_t N,
                              int64_t C, float bw, float bh, float t, float die_w, float die_h,
                              float2 *csize, float2 *coff, float *ratio, float *area,
                              float *npins_out, float2 *hi) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float sqrt2 = 1.41421356237309504880f, half = 0.5f;
    const float wi = w[i], hi_ = h[i];
    const float cw = fmaxf(wi, bw * sqrt2), ch = fmaxf(hi_, bh * sqrt2);
    csize[i] = make_float2(cw, ch);
    coff[i] = make_float2((wi - cw) * half, (hi_ - ch) * half);
    ratio[i] = i < C ? (wi * hi_) / (cw * ch) : t;
    if (i < C) {
        area[i] = wi * hi_;
        npins_out[i] = npins[i];
        hi[i] = make_float2(die_w - wi - 1.f, die_h - hi_ - 1.f);
    }
}

inline unsigned grid_for(int64_t count) { return (unsigned)((count + kBlock - 1) / kBlock); }

}  // namespace

int cell_partials(int64_t C) { return (int)grid_for(C > 0 ? C : 1); }

inline dim3 cell_grid(const BatchDims &bd, int nj) { return dim3(grid_for(bd.max_cells), nj); }

#define ERP_LAUNCH_CELLS(kern, ...)                                                       \
    do {                                                                                  \
        if (bd.max_cells <= 0) return ERP_OK;                                             \
        hipLaunchKernelGGL(kern, cell_grid(bd, nj), dim3(kBlock), 0, s, __VA_ARGS__);     \
        ERP_HIP_TRY(hipGetLastError());                                                   \
        return ERP_OK;                                                                    \
    } while (0)

int launch_pin2cell(const JobDesc *d, const BatchDims &bd, int nj, hipStream_t s) {
    ERP_LAUNCH_CELLS(k_pin2cell, d);
}
int launch_efield(const JobDesc *d, const BatchDims &bd, int nj, hipStream_t s) {
    ERP_LAUNCH_CELLS(k_efield, d);
}
int launch_l1(const JobDesc *d, const BatchDims &bd, int nj, hipStream_t s) { ERP_LAUNCH_CELLS(k_l1, d); }
int launch_combine(const JobDesc *d, const BatchDims &bd, int nj, hipStream_t s) {
    ERP_LAUNCH_CELLS(k_combine, d);
}
int launch_field_sorted(const JobDesc *d, const BatchDims &bd, int nj, hipStream_t s, bool with_sched) {
    if (!with_sched) ERP_LAUNCH_CELLS(k_field_sorted<false>, d);
    const unsigned nb = bd.max_cells > 0 ? grid_for(bd.max_cells) : 0u;
    hipLaunchKernelGGL(k_field_sorted<true>, dim3(8u + nb, nj), dim3(kBlock), 0, s, d);
    ERP_HIP_TRY(hipGetLastError());
    return ERP_OK;
}
int launch_cell_fused(const JobDesc *d, const BatchDims &bd, int nj, bool store_parts,
                      hipStream_t s, bool read_gd) {
    if (read_gd) {
        if (store_parts) ERP_LAUNCH_CELLS((k_cell_fused<true, true>), d);
        ERP_LAUNCH_CELLS((k_cell_fused<false, true>), d);
    }
    if (store_parts) ERP_LAUNCH_CELLS((k_cell_fused<true, false>), d);
    ERP_LAUNCH_CELLS((k_cell_fused<false, false>), d);
}
int launch_update(const JobDesc *d, const BatchDims &bd, int nj, hipStream_t s) {
    ERP_LAUNCH_CELLS(k_update, d);
}

int launch_scalar(const JobDesc *d, int nj, int ops, hipStream_t s) {
    hipLaunchKernelGGL(k_scalar, dim3(nj), dim3(kScalarThreads), 0, s, d, ops);
    ERP_HIP_TRY(hipGetLastError());
    return ERP_OK;
}

int launch_gather_state(const JobDesc *d, int nj, double *out, hipStream_t s) {
    if (nj <= 0) return ERP_OK;
    hipLaunchKernelGGL(k_gather_state, dim3((unsigned)((nj + 63) / 64)), dim3(64), 0, s, d, nj, out);
    ERP_HIP_TRY(hipGetLastError());
    return ERP_OK;
}

int launch_reset_state(DevState *st, double gamma0, double dw0, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_state, dim3(1), dim3(1), 0, s, st, gamma0, dw0);
    ERP_HIP_TRY(hipGetLastError());
    return ERP_OK;
}

int launch_init_positions(float2 *pos, const float *w, const float *h, const float *x,
                          const float *y, int64_t num_nodes, int64_t num_movable, float die_w,
                          float die_h, float2 *prev_pos, float2 *prev_g, hipStream_t s) {
    hipLaunchKernelGGL(k_init_positions, dim3(grid_for(num_nodes)), dim3(kBlock), 0, s, pos, w, h,
                       x, y, num_nodes, num_movable, die_w, die_h, prev_pos, prev_g);
    ERP_HIP_TRY(hipGetLastError());
    return ERP_OK;
}

namespace {
__global__ void __launch_bounds__(kBlock) k_pack_cell_consts(
    const float2 *__restrict__ csize, const float2 *__restrict__ coff, const float *__restrict__ ratio,
    const float *__restrict__ npins, const float *__restrict__ area, int64_t num_movable,
    float4 *__restrict__ cgeo, float4 *__restrict__ cprop) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num_movable) return;
    const float2 o = coff[i], sz = csize[i];
    cgeo[i] = make_float4(o.x, o.y, sz.x, sz.y);
    cprop[i] = make_float4(ratio[i], npins[i], area[i], 0.f);
}
}  // namespace

int launch_pack_cell_consts(const float2 *csize, const float2 *coff, const float *ratio,
                            const float *npins, const float *area, int64_t num_movable,
                            float4 *cgeo, float4 *cprop, hipStream_t s) {
    if (num_movable <= 0) return ERP_OK;
    hipLaunchKernelGGL(k_pack_cell_consts, dim3(grid_for(num_movable)), dim3(kBlock), 0, s, csize,
                       coff, ratio, npins, area, num_movable, cgeo, cprop);
    ERP_HIP_TRY(hipGetLastError());
    return ERP_OK;
}

int launch_setup_cells(const float *w, const float *h, const float *npins, int64_t num_nodes,
                       int64_t num_movable, float bw, float bh, float target_density,
                       float die_w, float die_h, float2 *csize, float2 *coff, float *ratio,
                       float *area, float *npins_out, float2 *hi, hipStream_t s) {
    hipLaunchKernelGGL(k_setup_cells, dim3(grid_for(num_nodes)), dim3(kBlock), 0, s, w, h, npins,
                       num_nodes, num_movable, bw, bh, target_density, die_w, die_h, csize, coff,
                       ratio, area, npins_out, hi);
    ERP_HIP_TRY(hipGetLastError());
    return ERP_OK;
}

}  // namespace erp
