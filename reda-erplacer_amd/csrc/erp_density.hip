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
#include <algorithm>
#include <climits>

#include "erp_internal.h"
#include "erp_kernels.h"
#include "erp_timer.h"

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

// ------------------------------------------------------------------ movable scatter
// Cells are processed in a spatially sorted order, 1024 per workgroup.  rec
// holds each cell's clamped lower-left corner (x + ox, y + oy) and clamped size,
// aux its clamp offset, ratio and id, both in that order (written by the sort
// below).  A
// workgroup accumulates into an LDS int64 tile of kScatterTile bins placed over
// its cells; each touched tile bin then costs one global atomic per workgroup
// instead of one per contribution.  A cell whose window does not fit the tile
// adds its contributions with global atomics directly.  Global 64-bit atomics
// are the expensive resource here: they are coherent device-wide and far
// slower than LDS atomics.
__global__ void __launch_bounds__(kScatterThreads) k_scatter(const JobDesc *__restrict__ jd, int copy_perm) {
    if (gridDim.y > 1 && !jd[blockIdx.y].active) return;   // an empty slot of a batch
    __shared__ unsigned long long tile[kScatterTile];
    __shared__ int wbox[4][kScatterThreads / 64];
    const SortBufs &sb = jd[blockIdx.y].sb;
    if (copy_perm && blockIdx.x == 0)   // after the sort: its tile counts start at 0 next time
        for (int t = threadIdx.x; t < kSortTiles; t += kScatterThreads) sb.hist[t] = 0;
    const int64_t num_cells = sb.num_cells;
    const int64_t base = (int64_t)blockIdx.x * kScatterCells;
    if (base >= num_cells) return;                      // this job needs fewer workgroups
    const float4 *__restrict__ rec = sb.rec;
    const float4 *__restrict__ aux = sb.aux;
    int32_t *__restrict__ perm_dst = copy_perm ? sb.perm : nullptr;
    float4 *__restrict__ sstat = sb.sstat;
    float *__restrict__ sratio = sb.sratio;
    int64_t *__restrict__ q = jd[blockIdx.y].pb.q_mov;
    const int m = sb.m, n = sb.n;
    const float bw = sb.bw, bh = sb.bh;
    const float inv_bw = 1.f / bw, inv_bh = 1.f / bh;
    const double scale = jd[blockIdx.y].pb.scale;
    float xi[kScatterPer], yi[kScatterPer], sx[kScatterPer], sy[kScatterPer], r[kScatterPer];
    int bxl[kScatterPer], bxh[kScatterPer], byl[kScatterPer], byh[kScatterPer];
    int lx = INT_MAX, hx = INT_MIN, ly = INT_MAX, hy = INT_MIN;
#pragma unroll
    for (int j = 0; j < kScatterPer; ++j) {
        const int64_t i = base + threadIdx.x + (int64_t)j * kScatterThreads;
        bxl[j] = bxh[j] = byl[j] = byh[j] = 0;
        xi[j] = yi[j] = sx[j] = sy[j] = r[j] = 0.f;
        if (i < num_cells) {
            const float4 v = rec[i];
            xi[j] = v.x;
            yi[j] = v.y;
            sx[j] = v.z;
            sy[j] = v.w;
            const float4 a = aux[i];
            r[j] = a.z;
            if (perm_dst) {          // the sorted order becomes the next sort's input order
                perm_dst[i] = __float_as_int(a.w);
                sstat[i] = make_float4(v.z, v.w, a.x, a.y);
                sratio[i] = a.z;
            }
            bin_range(xi[j], sx[j], inv_bw, m, bxl[j], bxh[j]);
            bin_range(yi[j], sy[j], inv_bh, n, byl[j], byh[j]);
            if (bxh[j] > bxl[j] && byh[j] > byl[j]) {
                lx = min(lx, bxl[j]); hx = max(hx, bxh[j]);
                ly = min(ly, byl[j]); hy = max(hy, byh[j]);
            }
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lx = min(lx, __shfl_xor(lx, off)); hx = max(hx, __shfl_xor(hx, off));
        ly = min(ly, __shfl_xor(ly, off)); hy = max(hy, __shfl_xor(hy, off));
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        wbox[0][wave] = lx; wbox[1][wave] = hx; wbox[2][wave] = ly; wbox[3][wave] = hy;
    }
    __syncthreads();
    lx = wbox[0][0]; hx = wbox[1][0]; ly = wbox[2][0]; hy = wbox[3][0];
#pragma unroll
    for (int w = 1; w < kScatterThreads / 64; ++w) {
        lx = min(lx, wbox[0][w]); hx = max(hx, wbox[1][w]);
        ly = min(ly, wbox[2][w]); hy = max(hy, wbox[3][w]);
    }
    if (hx <= lx || hy <= ly) return;      // no contribution in this workgroup (uniform)
    // LDS tile: the bounding box when it fits, else a kScatterSide^2 window
    // centred on the box (cells outside it go to global memory directly)
    int tx0 = lx, ty0 = ly, wx = hx - lx, wy = hy - ly;
    if ((int64_t)wx * wy > kScatterTile) {
        wx = min(wx, kScatterSide);
        wy = min(wy, kScatterTile / wx);
        tx0 = (lx + hx - wx) / 2;
        ty0 = (ly + hy - wy) / 2;
    }
    const int area = wx * wy;
    for (int t = threadIdx.x; t < area; t += kScatterThreads) tile[t] = 0ull;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kScatterPer; ++j) {
        const bool inside = bxl[j] >= tx0 && bxh[j] <= tx0 + wx && byl[j] >= ty0 && byh[j] <= ty0 + wy;
        for (int bx = bxl[j]; bx < bxh[j]; ++bx) {
            const float px = r[j] * tri_overlap(xi[j], sx[j], bx, bw);
            for (int by = byl[j]; by < byh[j]; ++by) {
                const float py = tri_overlap(yi[j], sy[j], by, bh);
                const int64_t v = quantize(px * py, scale);
                if (v == 0) continue;
                if (inside)
                    atomicAdd(&tile[(bx - tx0) * wy + (by - ty0)], (unsigned long long)v);
                else
                    add_q(q + (size_t)bx * n + by, v);
            }
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < area; t += kScatterThreads) {
        const unsigned long long v = tile[t];
        if (v != 0ull) add_q(q + (size_t)(tx0 + t / wy) * n + (ty0 + t % wy), (int64_t)v);
    }
}

// ------------------------------------------------------- spatial counting sort
// Every iteration the cells are re-bucketed by the sort tile (tiles of 16x16
// bins, coarser on grids above 1024, at most kSortTiles tiles, in Morton order
// so that consecutive tiles are close) of their clamped lower-left corner.  The
// input order is the previous iteration's sorted order, so each workgroup sees
// cells of only a few tiles and the global histogram / reservation atomics stay
// few.  Pass 1: keys + histogram.  Pass 2: per-workgroup ranks in LDS, one
// reservation atomic per tile and workgroup, and the sorted scatter records;
// every workgroup scans the tile counts itself (4096 at most, from L2) and
// reserves relative to its tile's start; pass 1 clears the cursors and the
// scatter clears the counts for the next sort.  The owner
// scatter also needs each tile's start and its work items: for it a separate
// scan (k_sort_scan) runs between the two passes and the cursors start at the
// tile starts.  Order inside a tile depends on the schedule; the fixed-point
// scatter is order independent, so the density map is not.
__device__ __forceinline__ unsigned spread_bits(unsigned v) {   // 0b abc -> 0b a0b0c
    v &= 0xffu;
    v = (v | (v << 4)) & 0x0f0fu;
    v = (v | (v << 2)) & 0x3333u;
    v = (v | (v << 1)) & 0x5555u;
    return v;
}
__device__ __forceinline__ int sort_key(float lx, float ly, float inv_bw, float inv_bh, int m,
                                        int n, int shift) {
    int bx = (int)floorf(lx * inv_bw), by = (int)floorf(ly * inv_bh);
    bx = bx < 0 ? 0 : (bx > m - 1 ? m - 1 : bx);
    by = by < 0 ? 0 : (by > n - 1 ? n - 1 : by);
    return (int)((spread_bits((unsigned)(bx >> shift)) << 1) | spread_bits((unsigned)(by >> shift)));
}

__device__ __forceinline__ unsigned compact_bits(unsigned v) {  // 0b a0b0c -> 0b abc
    v &= 0x5555u;
    v = (v | (v >> 1)) & 0x3333u;
    v = (v | (v >> 2)) & 0x0f0fu;
    v = (v | (v >> 4)) & 0x00ffu;
    return v;
}
__device__ __forceinline__ int tile_key(int tx, int ty) {
    return (int)((spread_bits((unsigned)tx) << 1) | spread_bits((unsigned)ty));
}

__global__ void __launch_bounds__(kScatterThreads) k_sort_keys(const JobDesc *__restrict__ jd, int shift,
                                                      int ntiles) {
    if (gridDim.y > 1 && !jd[blockIdx.y].active) return;   // an empty slot of a batch
    __shared__ int lh[kSortTiles];
    const SortBufs &sb = jd[blockIdx.y].sb;
    const int64_t num_cells = sb.num_cells;
    const int64_t nblk = (num_cells + kScatterCells - 1) / kScatterCells;
    if ((int64_t)blockIdx.x >= nblk) return;
    const int32_t *__restrict__ perm = sb.perm;
    const float2 *__restrict__ pos = sb.pos;
    const float4 *__restrict__ sstat = sb.sstat;
    int32_t *__restrict__ keys = sb.keys;
    float2 *__restrict__ ll = sb.ll;
    int32_t *__restrict__ hist = sb.hist;
    const int m = sb.m, n = sb.n;
    const float inv_bw = 1.f / sb.bw, inv_bh = 1.f / sb.bh;
    for (int t = threadIdx.x; t < ntiles; t += kScatterThreads) {
        lh[t] = 0;
        if (blockIdx.x == 0) sb.cursor[t] = 0;   // the placement pass's cursors count from 0
    }
    __syncthreads();
    const int64_t base = xcd_block(blockIdx.x, nblk) * kScatterCells;
    // all loads of the thread's cells first (they are then in flight together)
    int32_t cid[kScatterPer];
    float4 stv[kScatterPer];
    float2 pv[kScatterPer];
#pragma unroll
    for (int j = 0; j < kScatterPer; ++j) {
        const int64_t i = base + threadIdx.x + (int64_t)j * kScatterThreads;
        cid[j] = i < num_cells ? perm[i] : -1;
        stv[j] = i < num_cells ? sstat[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int j = 0; j < kScatterPer; ++j) pv[j] = cid[j] >= 0 ? pos[cid[j]] : make_float2(0.f, 0.f);
#pragma unroll
    for (int j = 0; j < kScatterPer; ++j) {
        const int64_t i = base + threadIdx.x + (int64_t)j * kScatterThreads;
        const bool valid = i < num_cells;
        int k = -1;
        if (valid) {
            const float2 l = make_float2(pv[j].x + stv[j].z, pv[j].y + stv[j].w);
            k = sort_key(l.x, l.y, inv_bw, inv_bh, m, n, shift);
            keys[i] = k;
            ll[i] = l;
        }
        // a wavefront whose cells share one tile adds its count once
        const int k0 = __shfl(k, 0);
        if (__all(k == k0 || !valid)) {
            const int cnt = __popcll(__ballot(valid));
            if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(&lh[k0], cnt);
        } else if (valid) {
            atomicAdd(&lh[k], 1);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < ntiles; t += kScatterThreads)
        if (lh[t]) atomicAdd(&hist[t], lh[t]);
}

// Owner scatter only: exclusive scan of the tile counts into the placement
// cursors and the tile starts; then the number of workgroups each tile needs
// (its own cells and those of the `reach` tiles to its left and below, in chunks
// of kOwnerVisits; 0 outside the grid) and their exclusive scan.
__device__ __forceinline__ int block_exclusive_scan(int sum, int *wsum, int &total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int inc = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(inc, off);
        if (lane >= off) inc += y;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    int before = 0;
    total = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        if (w < wave) before += wsum[w];
        total += wsum[w];
    }
    __syncthreads();
    return before + inc - sum;
}

__global__ void __launch_bounds__(1024) k_sort_scan(const JobDesc *__restrict__ jd, int ntiles, int shift) {
    if (gridDim.x > 1 && !jd[blockIdx.x].active) return;   // an empty slot of a batch
    __shared__ int wsum[16];
    __shared__ int cnt[kSortTiles];
    const SortBufs &sb = jd[blockIdx.x].sb;
    int32_t *__restrict__ hist = sb.hist;
    int32_t *__restrict__ cursor = sb.cursor;
    const int per = (ntiles + 1023) / 1024;          // ntiles <= kSortTiles = 4 * 1024
    const int t0 = threadIdx.x * per;
    int v[4] = {0, 0, 0, 0};
    int sum = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (k < per && t0 + k < ntiles) { v[k] = hist[t0 + k]; sum += v[k]; cnt[t0 + k] = v[k]; }
    int total;
    int run = block_exclusive_scan(sum, wsum, total);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k < per && t0 + k < ntiles) {
            cursor[t0 + k] = run;
            sb.tstart[t0 + k] = run;
            run += v[k];
            hist[t0 + k] = 0;
        }
    }
    if (threadIdx.x == 0) sb.tstart[ntiles] = total;
    // work items (cnt is complete: the scan above synchronised)
    const int T = 1 << shift, R = sb.reach;
    const int tiles_x = (sb.m + T - 1) >> shift, tiles_y = (sb.n + T - 1) >> shift;
    int nch[4] = {0, 0, 0, 0};
    sum = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int t = t0 + k;
        if (k >= per || t >= ntiles) continue;
        const int tx = (int)compact_bits((unsigned)t >> 1), ty = (int)compact_bits((unsigned)t);
        if (tx >= tiles_x || ty >= tiles_y) continue;
        int vis = 0;
        for (int a = 0; a <= R && a <= tx; ++a)
            for (int b = 0; b <= R && b <= ty; ++b) vis += cnt[tile_key(tx - a, ty - b)];
        nch[k] = vis ? (vis + kOwnerVisits - 1) / kOwnerVisits : 1;
        sum += nch[k];
    }
    run = block_exclusive_scan(sum, wsum, total);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k < per && t0 + k < ntiles) {
            sb.items[t0 + k] = run;
            run += nch[k];
        }
    }
    if (threadIdx.x == 0) sb.items[ntiles] = total;
}

// OWN: the workgroup scans the tile counts itself and the cursors count from 0
// (k_sort_keys cleared them; the scatter that follows clears the counts), else
// k_sort_scan has set each cursor to its tile's start and cleared the counts.
template <bool OWN>
__global__ void __launch_bounds__(kScatterThreads) k_sort_place(const JobDesc *__restrict__ jd, int ntiles) {
    if (gridDim.y > 1 && !jd[blockIdx.y].active) return;   // an empty slot of a batch
    constexpr int kPer = kSortTiles / kScatterThreads;      // tile counts per thread
    __shared__ int lh[kSortTiles];
    __shared__ int pre[OWN ? kSortTiles : 1];
    __shared__ int wsum[kScatterThreads / 64];
    const SortBufs &sb = jd[blockIdx.y].sb;
    const int64_t num_cells = sb.num_cells;
    const int64_t nblk = (num_cells + kScatterCells - 1) / kScatterCells;
    if ((int64_t)blockIdx.x >= nblk) return;
    const int32_t *__restrict__ perm = sb.perm;
    const int32_t *__restrict__ keys = sb.keys;
    const float2 *__restrict__ ll = sb.ll;
    const float4 *__restrict__ sstat = sb.sstat;
    const float *__restrict__ sratio = sb.sratio;
    const int32_t *__restrict__ hist = sb.hist;
    int32_t *__restrict__ cursor = sb.cursor;

    float4 *__restrict__ rec = sb.rec;
    float4 *__restrict__ aux = sb.aux;
    const int t0 = threadIdx.x * kPer;
    int hv[kPer];
    if (OWN) {
#pragma unroll
        for (int k = 0; k < kPer; ++k) hv[k] = t0 + k < ntiles ? hist[t0 + k] : 0;
    }
    for (int t = threadIdx.x; t < ntiles; t += kScatterThreads) lh[t] = 0;
    __syncthreads();
    const int64_t base = xcd_block(blockIdx.x, nblk) * kScatterCells;
    int key[kScatterPer], rank[kScatterPer];
    int32_t cperm[kScatterPer];
    float2 cll[kScatterPer];
    float4 cst[kScatterPer];
    float crat[kScatterPer];
    // the thread's records are loaded up front, all in flight together
#pragma unroll
    for (int j = 0; j < kScatterPer; ++j) {
        const int64_t i = base + threadIdx.x + (int64_t)j * kScatterThreads;
        const bool valid = i < num_cells;
        key[j] = valid ? keys[i] : -1;
        cperm[j] = valid ? perm[i] : 0;
        cll[j] = valid ? ll[i] : make_float2(0.f, 0.f);
        cst[j] = valid ? sstat[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        crat[j] = valid ? sratio[i] : 0.f;
    }
    if (OWN) {   // exclusive prefix of the tile counts: each tile's first sorted index
        int sum = 0;
#pragma unroll
        for (int k = 0; k < kPer; ++k) sum += hv[k];
        int total;
        int run = block_exclusive_scan(sum, wsum, total);
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            if (t0 + k < ntiles) pre[t0 + k] = run;
            run += hv[k];
        }
    }
#pragma unroll
    for (int j = 0; j < kScatterPer; ++j) {
        const int64_t i = base + threadIdx.x + (int64_t)j * kScatterThreads;
        const bool valid = i < num_cells;
        rank[j] = 0;
        // valid lanes are a prefix of the wavefront; if they share one tile the
        // wavefront reserves its ranks with one LDS atomic
        const int k0 = __shfl(key[j], 0);
        if (__all(key[j] == k0 || !valid)) {
            const int cnt = __popcll(__ballot(valid));
            int b0 = 0;
            if ((threadIdx.x & 63) == 0 && cnt) b0 = atomicAdd(&lh[k0], cnt);
            rank[j] = __shfl(b0, 0) + (int)(threadIdx.x & 63);
        } else if (valid) {
            rank[j] = atomicAdd(&lh[key[j]], 1);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < ntiles; t += kScatterThreads)
        if (lh[t]) lh[t] = (OWN ? pre[t] : 0) + atomicAdd(&cursor[t], lh[t]);   // count -> start
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kScatterPer; ++j) {
        if (key[j] >= 0) {
            const int dst = lh[key[j]] + rank[j];
            rec[dst] = make_float4(cll[j].x, cll[j].y, cst[j].x, cst[j].y);
            aux[dst] = make_float4(cst[j].z, cst[j].w, crat[j], __int_as_float(cperm[j]));
        }
    }
}

// Owner scatter: workgroup b takes item b of the list built by k_sort_scan --
// a bin tile (T x T bins, Morton index t) and a chunk of the cells that can touch
// it: the sorted cells of tile t and of the `reach` tiles to its left and below
// (a cell's lower-left bin lies in its own tile and it spans at most `reach`
// tiles further).  Contributions are clipped to the tile and summed in an LDS
// int64 tile; a tile served by one workgroup stores its bins (every bin, no
// atomics, the map needs no clearing there); a tile split over several
// workgroups adds its nonzero bins atomically into the cleared map.  The
// workgroup that visits a cell in its own tile also copies the cell's sorted
// record back as the next sort's input.
__global__ void __launch_bounds__(kBlock) k_scatter_owner(const JobDesc *__restrict__ jd, int shift,
                                                          int ntiles) {
    if (gridDim.y > 1 && !jd[blockIdx.y].active) return;   // an empty slot of a batch
    __shared__ unsigned long long tile[32 * 32];
    __shared__ int src_start[9], src_cnt[9];
    const SortBufs &sb = jd[blockIdx.y].sb;
    const int32_t *__restrict__ items = sb.items;
    const int b = (int)blockIdx.x;
    if (b >= items[ntiles]) return;                 // uniform per workgroup
    int lo = 0, hi = ntiles;                        // last t with items[t] <= b
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (items[mid] <= b) lo = mid;
        else hi = mid;
    }
    const int t = lo;
    const int chunk = b - items[t];
    const bool exclusive = items[t + 1] - items[t] == 1;
    const int T = 1 << shift, R = sb.reach, W = R + 1;
    const int tx = (int)compact_bits((unsigned)t >> 1), ty = (int)compact_bits((unsigned)t);
    const int x0 = tx << shift, y0 = ty << shift;
    const int m = sb.m, n = sb.n;
    if ((int)threadIdx.x < W * W) {                 // own tile first
        const int a = threadIdx.x % W, c = threadIdx.x / W;
        int st = 0, ct = 0;
        if (a <= tx && c <= ty) {
            const int s = tile_key(tx - a, ty - c);
            st = sb.tstart[s];
            ct = sb.tstart[s + 1] - st;
        }
        src_start[threadIdx.x] = st;
        src_cnt[threadIdx.x] = ct;
    }
    for (int e = threadIdx.x; e < T * T; e += kBlock) tile[e] = 0ull;
    __syncthreads();
    int total = 0;
    for (int k = 0; k < W * W; ++k) total += src_cnt[k];
    const int v0 = chunk * kOwnerVisits, v1 = min(v0 + kOwnerVisits, total);
    const float4 *__restrict__ rec = sb.rec;
    const float4 *__restrict__ aux = sb.aux;
    const float bw = sb.bw, bh = sb.bh;
    const float inv_bw = 1.f / bw, inv_bh = 1.f / bh;
    const double scale = jd[blockIdx.y].pb.scale;
    for (int v = v0 + (int)threadIdx.x; v < v1; v += kBlock) {
        int k = 0, off = v;
        while (off >= src_cnt[k]) { off -= src_cnt[k]; ++k; }
        const int i = src_start[k] + off;
        const float4 r4 = rec[i];
        const float4 a4 = aux[i];
        const float r = a4.z;
        if (k == 0) {                               // the cell's own tile
            sb.perm[i] = __float_as_int(a4.w);
            sb.sstat[i] = make_float4(r4.z, r4.w, a4.x, a4.y);
            sb.sratio[i] = r;
        }
        int bxl, bxh, byl, byh;
        bin_range(r4.x, r4.z, inv_bw, m, bxl, bxh);
        bin_range(r4.y, r4.w, inv_bh, n, byl, byh);
        bxl = max(bxl, x0); bxh = min(bxh, x0 + T);
        byl = max(byl, y0); byh = min(byh, y0 + T);
        for (int bx = bxl; bx < bxh; ++bx) {
            const float px = r * tri_overlap(r4.x, r4.z, bx, bw);
            for (int by = byl; by < byh; ++by) {
                const float py = tri_overlap(r4.y, r4.w, by, bh);
                const int64_t q = quantize(px * py, scale);
                if (q != 0) atomicAdd(&tile[(bx - x0) * T + (by - y0)], (unsigned long long)q);
            }
        }
    }
    __syncthreads();
    int64_t *__restrict__ qm = jd[blockIdx.y].pb.q_mov;
    for (int e = threadIdx.x; e < T * T; e += kBlock) {
        const int bx = x0 + e / T, by = y0 + e % T;
        if (bx >= m || by >= n) continue;
        const unsigned long long val = tile[e];
        if (exclusive) qm[(size_t)bx * n + by] = (int64_t)val;
        else if (val != 0ull) add_q(qm + (size_t)bx * n + by, (int64_t)val);
    }
}

// Scatter records in identity order (leaf entry point, no sort).
__global__ void k_pack_records(const float2 *__restrict__ pos, const float2 *__restrict__ csize,
                               const float2 *__restrict__ coff, int64_t num_cells,
                               float4 *__restrict__ rec) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < num_cells) {
        const float2 p = pos[i], sz = csize[i], o = coff[i];
        rec[i] = make_float4(p.x + o.x, p.y + o.y, sz.x, sz.y);
    }
}

// Sort constants in identity order (the order of a fresh perm).
__global__ void k_init_sort_static(const float2 *__restrict__ csize, const float2 *__restrict__ coff,
                                   const float *__restrict__ ratio, int64_t num_cells,
                                   float4 *__restrict__ sstat, float *__restrict__ sratio) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < num_cells) {
        const float2 sz = csize[i], o = coff[i];
        sstat[i] = make_float4(sz.x, sz.y, o.x, o.y);
        sratio[i] = ratio[i];
    }
}

__global__ void k_iota(int32_t *__restrict__ v, int64_t count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) v[i] = (int32_t)i;
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

int launch_scatter(const JobDesc *d, const BatchDims &bd, int nj, bool copy_perm, hipStream_t s) {
    if (bd.max_cells <= 0) return ERP_OK;
    const unsigned nb = (unsigned)((bd.max_cells + kScatterCells - 1) / kScatterCells);
    hipLaunchKernelGGL(k_scatter, dim3(nb, nj), dim3(kScatterThreads), 0, s, d, copy_perm ? 1 : 0);
    ERP_HIP_TRY(hipGetLastError());
    return ERP_OK;
}

int launch_scatter_owner(const JobDesc *d, const BatchDims &bd, int nj, hipStream_t s) {
    if (bd.max_cells <= 0) return ERP_OK;
    const SortGeom g = sort_geometry(bd.m, bd.n);
    const int64_t w = bd.max_reach + 1;
    // upper bound of the items of any job: one per tile plus the chunks
    const int64_t nb = g.ntiles + (w * w * bd.max_cells + kOwnerVisits - 1) / kOwnerVisits;
    hipLaunchKernelGGL(k_scatter_owner, dim3((unsigned)nb, nj), dim3(kBlock), 0, s, d, g.shift, g.ntiles);
    ERP_HIP_TRY(hipGetLastError());
    return ERP_OK;
}

SortGeom sort_geometry(int m, int n) {
    SortGeom g;
    // Morton keys over a square of side^2 tiles, side = next power of two of the
    // larger tile count; coarsen the tiles until that square fits kSortTiles.
    for (g.shift = 4;; ++g.shift) {
        const int tx = std::max(m >> g.shift, 1), ty = std::max(n >> g.shift, 1);
        int side = 1;
        while (side < tx || side < ty) side <<= 1;
        g.ntiles = side * side;
        if (g.ntiles <= kSortTiles) break;
    }
    return g;
}

int launch_sort(const JobDesc *d, const BatchDims &bd, int nj, bool owner_items, hipStream_t s,
                KernelTimer *t) {
    if (bd.max_cells <= 0) return ERP_OK;
    const SortGeom g = sort_geometry(bd.m, bd.n);
    const unsigned nb = (unsigned)((bd.max_cells + kScatterCells - 1) / kScatterCells);
    if (t) t->begin("sort_keys");
    hipLaunchKernelGGL(k_sort_keys, dim3(nb, nj), dim3(kScatterThreads), 0, s, d, g.shift, g.ntiles);
    if (t) t->end();
    ERP_HIP_TRY(hipGetLastError());
    if (owner_items) {
        if (t) t->begin("sort_scan");
        hipLaunchKernelGGL(k_sort_scan, dim3(nj), dim3(1024), 0, s, d, g.ntiles, g.shift);
        if (t) t->end();
        ERP_HIP_TRY(hipGetLastError());
    }
    if (t) t->begin("sort_place");
    if (owner_items)
        hipLaunchKernelGGL(k_sort_place<false>, dim3(nb, nj), dim3(kScatterThreads), 0, s, d, g.ntiles);
    else
        hipLaunchKernelGGL(k_sort_place<true>, dim3(nb, nj), dim3(kScatterThreads), 0, s, d, g.ntiles);
    if (t) t->end();
    ERP_HIP_TRY(hipGetLastError());
    return ERP_OK;
}

int launch_init_sort_static(const float2 *csize, const float2 *coff, const float *ratio,
                            int64_t num_cells, float4 *sstat, float *sratio, hipStream_t s) {
    if (num_cells <= 0) return ERP_OK;
    hipLaunchKernelGGL(k_init_sort_static, dim3((unsigned)((num_cells + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, s, csize, coff, ratio, num_cells, sstat, sratio);
    ERP_HIP_TRY(hipGetLastError());
    return ERP_OK;
}

int launch_pack_records(const float2 *pos, const float2 *csize, const float2 *coff,
                        int64_t num_cells, float4 *rec, hipStream_t s) {
    if (num_cells <= 0) return ERP_OK;
    hipLaunchKernelGGL(k_pack_records, dim3((unsigned)((num_cells + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, s, pos, csize, coff, num_cells, rec);
    ERP_HIP_TRY(hipGetLastError());
    return ERP_OK;
}

int launch_iota(int32_t *v, int64_t count, hipStream_t s) {
    if (count <= 0) return ERP_OK;
    hipLaunchKernelGGL(k_iota, dim3((unsigned)((count + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                       v, count);
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
