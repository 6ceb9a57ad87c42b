// Weighted-average wirelength gradient and HPWL launches (gfx950), hpwl.rs:139-343;
// the device functions are in erp_wa.h.
#include <algorithm>
#include <climits>

#include "erp_internal.h"
#include "erp_kernels.h"
#include "erp_timer.h"
#include "erp_wa.h"

namespace erp {
namespace {

__global__ void __launch_bounds__(kBlock) k_wa_narrow(const JobDesc *__restrict__ jd, WaPlan plan) {
    if (gridDim.y > 1 && !jd[blockIdx.y].active) return;   // an empty slot of a batch
    wa_narrow_blocks(jd[blockIdx.y].wa, plan, blockIdx.x, gridDim.x, plan.begin[plan.nseg]);
}

__global__ void __launch_bounds__(kBlock) k_wa_wide(const JobDesc *__restrict__ jd, WaPlan plan) {
    if (gridDim.y > 1 && !jd[blockIdx.y].active) return;   // an empty slot of a batch
    const WaNets &w = jd[blockIdx.y].wa;
    for (unsigned vb = blockIdx.x; vb < plan.begin[plan.nseg]; vb += gridDim.x) {
        int sg = 0;
        while (sg + 1 < plan.nseg && vb >= plan.begin[sg + 1]) ++sg;
        const int64_t blk = (int64_t)vb - plan.begin[sg];
        if (plan.bucket[sg] == 7) wa_group<64, 8>(w, 7, blk);
        else wa_large(w, 8, blk);
    }
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

// Degree buckets of the multi-pin nets: bucket b holds degrees in
// (kWaDegree[b-1], kWaDegree[b]] and runs wa_group<G, K> with G*K >= degree.
const int kWaDegree[kWaBuckets] = {4, 8, 16, 32, 64, 128, 256, 512, 1 << 30};

namespace {
// lanes per net of bucket b (the last bucket loops with one wavefront per net)
constexpr int kWaLanes[kWaBuckets] = {1, 2, 4, 8, 16, 32, 64, 64, 64};
unsigned bucket_blocks(int64_t nets, int b) {
    return b == kWaBuckets - 1 ? blocks(nets, kBlock / 64) : blocks(nets * kWaLanes[b], kBlock);
}
}  // namespace

void BatchDims::add(const JobDesc &j) {
    m = j.pb.m;
    n = j.pb.n;
    max_cells = std::max<int64_t>(max_cells, j.cb.C);
    max_n2 = std::max<int64_t>(max_n2, j.wa.n2);
    for (int b = 0; b < kWaBuckets; ++b)
        max_bucket[b] = std::max<int64_t>(max_bucket[b], j.wa.bucket_begin[b + 1] - j.wa.bucket_begin[b]);
    max_reach = std::max(max_reach, j.sb.reach);
}

int wa_partials(WaNets &w) {
    int total = (int)blocks(w.n2, kBlock);
    w.part_offset[0] = 0;
    for (int b = 0; b < kWaBuckets; ++b) {
        w.part_offset[b + 1] = total;
        total += (int)bucket_blocks(w.bucket_begin[b + 1] - w.bucket_begin[b], b);
    }
    return total;
}

WaPlan wa_narrow_plan(const BatchDims &bd) {
    // heavier workgroups first: high-degree buckets, then the small ones
    WaPlan p{};
    unsigned total = 0;
    for (int b = 6; b >= -1; --b) {
        const unsigned nb = b < 0 ? blocks(bd.max_n2, kBlock) : bucket_blocks(bd.max_bucket[b], b);
        if (!nb) continue;
        p.bucket[p.nseg] = b;
        p.begin[p.nseg++] = total;
        total += nb;
    }
    p.begin[p.nseg] = total;
    return p;
}

WaPlan wa_wide_plan(const BatchDims &bd) {
    WaPlan p{};
    unsigned total = 0;
    for (int b = kWaBuckets - 1; b >= 7; --b) {
        const unsigned nb = bucket_blocks(bd.max_bucket[b], b);
        if (!nb) continue;
        p.bucket[p.nseg] = b;
        p.begin[p.nseg++] = total;
        total += nb;
    }
    p.begin[p.nseg] = total;
    return p;
}

int launch_wa_wide(const JobDesc *d, const BatchDims &bd, int nj, hipStream_t s, KernelTimer *t) {
    const WaPlan wide = wa_wide_plan(bd);
    const unsigned total = wide.begin[wide.nseg];
    if (total) {
        if (t) t->begin("wa_wide");
        const unsigned grid = total;
        hipLaunchKernelGGL(k_wa_wide, dim3(grid, nj), dim3(kBlock), 0, s, d, wide);
        if (t) t->end();
        ERP_HIP_TRY(hipGetLastError());
    }
    return ERP_OK;
}

int launch_wa(const JobDesc *d, const BatchDims &bd, int nj, hipStream_t s, KernelTimer *t) {
    const WaPlan narrow = wa_narrow_plan(bd);
    const unsigned total = narrow.begin[narrow.nseg];
    if (total) {
        if (t) t->begin("wa_narrow");
        const unsigned grid = total;
        hipLaunchKernelGGL(k_wa_narrow, dim3(grid, nj), dim3(kBlock), 0, s, d, narrow);
        if (t) t->end();
        ERP_HIP_TRY(hipGetLastError());
    }
    return launch_wa_wide(d, bd, nj, s, t);
}

int launch_hpwl_dbu(const WaNets &w, int64_t num_nets, unsigned long long *out, hipStream_t s) {
    if (num_nets <= 0) return ERP_OK;
    hipLaunchKernelGGL(k_hpwl_dbu, dim3(blocks(num_nets, kBlock)), dim3(kBlock), 0, s, w.slot_node,
                       w.slot_off, w.net_start, w.pos, num_nets, out);
    ERP_HIP_TRY(hipGetLastError());
    return ERP_OK;
}

}  // namespace erp
