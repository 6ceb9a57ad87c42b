// Launchers of the per-iteration kernels (definitions in erp_density.hip,
// erp_wirelength.hip, erp_optimizer.hip).
#pragma once
#include "erp_internal.h"

namespace erp {

// Loads through the global address space.  A job's buffers are reached through
// pointers read from its descriptor in memory, which the compiler cannot place
// in an address space: it emits flat loads, and a flat load counts against the
// same wait counter as the scalar loads of those pointers, so every pointer load
// then waits for all vector loads in flight.
// gld(p) loads *p (a type of 4, 8 or 16 bytes) through a global pointer to an
// integer vector of the same size and reinterprets the bits.
#define ERP_GLOBAL __attribute__((address_space(1)))
template <int B> struct GldWord;
template <> struct GldWord<4> { typedef unsigned type; };
template <> struct GldWord<8> { typedef unsigned type __attribute__((ext_vector_type(2))); };
template <> struct GldWord<16> { typedef unsigned type __attribute__((ext_vector_type(4))); };
template <class T>
__device__ __forceinline__ T gld(const T *p) {
    typedef typename GldWord<sizeof(T)>::type W;
    return __builtin_bit_cast(T, *(const ERP_GLOBAL W *)p);
}

// Workgroups are handed to the eight XCDs round robin (launch index w runs on
// XCD w mod 8), and each XCD has its own L2.  xcd_block(w, nb) renumbers the nb
// workgroups of one job so that every XCD processes one contiguous range of
// logical blocks: neighbouring cells, nets and bins then share one L2 instead
// of being fetched into all eight.  A bijection on [0, nb); the identity above.
__device__ __forceinline__ int64_t xcd_block(int64_t w, int64_t nb) {
    if (w >= nb) return w;
    const int64_t q = nb >> 3, r = nb & 7, x = w & 7;
    return x * q + (x < r ? x : r) + (w >> 3);
}

// Sum of v over a kBlock-thread workgroup in the order of the LDS tree
// red[t] += red[t + s] for s = kBlock / 2, ..., 1, with the same bits: the two
// levels across wavefronts through LDS, the six inside one wavefront by
// shuffles (two barriers instead of nine).  The result is valid in thread 0.
// red: kBlock elements of LDS, free again when the call returns.
template <class T>
__device__ __forceinline__ T block_sum_tree(T v, T *red) {
    static_assert(kBlock == 256, "the tree's first two levels cross the four wavefronts");
    red[threadIdx.x] = v;
    __syncthreads();
    T x = 0;
    if (threadIdx.x < 64) {
        const T a = red[threadIdx.x] + red[threadIdx.x + 128];        // level s = 128, t < 64
        const T b = red[threadIdx.x + 64] + red[threadIdx.x + 192];   // level s = 128, t in [64, 128)
        x = a + b;                                                     // level s = 64
#pragma unroll
        for (int s = 32; s > 0; s >>= 1) x += __shfl_down(x, s);     // levels 32 .. 1
    }
    __syncthreads();
    return x;
}

// ------------------------------------------------------------ density
// The sort and cell-major scatter kernels: 1024 cells per workgroup as 512
// threads x 2 cells (at 1M cells 4141-4148 it/s as 256 x 4, 4272-4279 as
// 512 x 2, 4280-4283 as 1024 x 1: more waves per CU hide the gathers).
constexpr int kScatterThreads = 512;
constexpr int kScatterPer = 2;                                 // cells per thread
constexpr int kScatterCells = kScatterPer * kScatterThreads;   // cells per workgroup
constexpr int kScatterTile = 8192;                    // LDS int64 bins per workgroup
constexpr int kScatterSide = 128;                     // tile side when the box is too big
constexpr int kSortTiles = 4096;                      // max sort tiles (64 x 64 Morton)
constexpr int kOwnerVisits = 4096;                    // cell visits per owner-scatter workgroup
struct SortGeom {
    int shift, ntiles;
};
SortGeom sort_geometry(int m, int n);
struct SortBufs {
    int64_t num_cells;
    const float2 *pos, *csize, *coff;
    const float *ratio;
    int m, n;
    float bw, bh;
    int32_t *perm;       // [C] current order (input of the sort)
    int32_t *keys, *hist, *cursor;   // keys [C], hist/cursor [kSortTiles]
    float4 *rec;         // [C] sorted scatter records {clamped lower-left, clamped size}
    // [C] sorted side records {clamp offset x, y, ratio, cell id (int bits)}: one
    // 16-byte write per cell in the sort (the cell id is copied back into perm
    // by the scatter)
    float4 *aux;
    // per-cell constants carried along the order so that the sort reads them
    // sequentially: {clamped w, h, offset x, y} and ratio in the order of perm
    float4 *sstat;
    float *sratio;
    float2 *ll;          // [C] clamped lower-left corners in the order of perm (sort temp)
    // owner scatter: first sorted index of every tile ([ntiles + 1], Morton
    // order) and the workgroup items per tile (exclusive prefix, [ntiles + 1])
    int32_t *tstart, *items;
    int reach;           // tiles a cell can reach beyond its own, per axis (>= 1)
};
int launch_init_sort_static(const float2 *csize, const float2 *coff, const float *ratio,
                            int64_t num_cells, float4 *sstat, float *sratio, hipStream_t s);
int launch_pack_records(const float2 *pos, const float2 *csize, const float2 *coff,
                        int64_t num_cells, float4 *rec, hipStream_t s);
int launch_iota(int32_t *v, int64_t count, hipStream_t s);
int launch_fixed_map(const float2 *pos, const float2 *csize, int64_t num_fixed, int m, int n, float bw, float bh, float die_w,
                     float die_h, double scale, int64_t *q, hipStream_t s);
int launch_scale_fixed(int64_t *q, int64_t nb, double t, hipStream_t s);
int launch_finalize_leaf(const int64_t *q, const float *initial, float *map, int64_t nb,
                         double inv_scale, hipStream_t s);
int launch_zero_i64(int64_t *q, int64_t count, hipStream_t s);

// ------------------------------------------------------------ wirelength
// Net-ordered pin slots: the pins of net e are slots [net_start[e], net_start[e+1]).
constexpr int kWaBuckets = 9;
extern const int kWaDegree[kWaBuckets];   // upper degree of each bucket
struct WaNets {
    const int32_t *slot_node;   // [P]
    const float2 *slot_off;     // [P] pin offset from the instance origin
    const int32_t *net_start;   // [E+1]
    int64_t n2;                 // nets [0, n2) have two pins, slots 2e, 2e+1
    int64_t bucket_begin[kWaBuckets + 1];  // multi-pin nets, sorted by degree, per bucket
    int32_t part_offset[kWaBuckets + 2];   // first partial of the 2-pin launch / bucket b (b+1)
    const float2 *pos;          // [N] instance positions
    const double *gamma;        // WA gamma (the scheduler's, or an explicit one)
    float2 *slot_grad;          // [P] output
    double *hpwl_part;          // partial sums, one per workgroup (see wa_partials)
};
// One launch for several net groups: the grid is cut into segments of
// workgroups, each segment running one group (-1: two-pin nets, b >= 0: degree
// bucket b), so small buckets overlap instead of each draining the GPU alone.
struct WaPlan {
    int nseg;
    int bucket[kWaBuckets + 1];
    unsigned begin[kWaBuckets + 2];
};
// The plans of a batch: net groups up to degree 256 (k_wa_narrow) and above.
struct BatchDims;
WaPlan wa_narrow_plan(const BatchDims &bd);
WaPlan wa_wide_plan(const BatchDims &bd);
// sets part_offset and returns the number of partials
int wa_partials(WaNets &w);
// Exact HPWL in int64 DBU over all nets.
int launch_hpwl_dbu(const WaNets &w, int64_t num_nets, unsigned long long *out, hipStream_t s);

// ------------------------------------------------------------ optimizer
struct CellBufs {
    int64_t C;
    const float2 *pos;        // [N] (movable first)
    const float2 *csize, *coff;
    const float *ratio;
    const float *area, *npins;  // [C] w*h and num_pins
    // [C] the same constants packed for the cell kernels (one 16-byte load each):
    // {clamp offset x, y, clamped w, h} and {ratio, num_pins, w*h, 0}
    const float4 *cgeo, *cprop;
    const float2 *hi;           // [C] (W - w - 1, H - h - 1)
    const int32_t *cstart, *cslots;   // cell -> pin slots CSR (macro_2_pins order)
    const float2 *slot_grad;
    float2 *gwl, *gd, *g;       // [C]
    float2 *prev_pos, *prev_g;  // [C]
    float2 *pos_out;            // == pos (movable part written)
    double *part_a, *part_b;    // per-workgroup FP64 partials
    const float2 *fxy;          // field maps, (x, y) per bin at fxy_index (4 x 4-bin blocks)
    int m, n;
    float bw, bh;
    DevState *st;               // scalar state (density weight, alpha, beta)
};
int cell_partials(int64_t C);

// Scalar work on the device state (one workgroup).
enum ScalarOps {
    SC_HPWL = 1,    // hpwl = sum(hpwl partials)
    SC_POT = 2,     // overflow, energy from their partials
    SC_L1 = 4,      // density_weight *= l1_wl / l1_d (iteration 0)
    SC_NEST = 8,    // update_momentum, then BB step size
    SC_SCHED = 16,  // Scheduler::update (SC_SCHED_STAGE, then SC_COMMIT)
    // Scheduler::update in two halves around the Nesterov update of the same
    // step.  STAGE: the history, should_stop, and the next density weight and
    // gamma into dw_next / gamma_next (it reads the density weight, the step's
    // HPWL and overflow and the iteration count, none of which SC_NEST writes,
    // and writes nothing SC_NEST reads).  COMMIT: iteration + 1 and the staged
    // density weight and gamma.
    SC_SCHED_STAGE = 32,
    SC_COMMIT = 64,
};
struct ScalarArgs {
    DevState *st;
    const double *hpwl_part; int n_hpwl;
    const int64_t *ovf_part; int n_ovf;
    const double *energy_part; int n_energy;
    const double *part_a, *part_b; int n_cell;
    double inv_scale_over_area;   // 1 / (S * movable_area)
};
int launch_reset_state(DevState *st, double gamma0, double dw0, hipStream_t s);
int launch_init_positions(float2 *pos, const float *w, const float *h, const float *x,
                          const float *y, int64_t num_nodes, int64_t num_movable, float die_w,
                          float die_h, float2 *prev_pos, float2 *prev_g, hipStream_t s);
int launch_pack_cell_consts(const float2 *csize, const float2 *coff, const float *ratio,
                            const float *npins, const float *area, int64_t num_movable,
                            float4 *cgeo, float4 *cprop, hipStream_t s);
int launch_setup_cells(const float *w, const float *h, const float *npins, int64_t num_nodes,
                       int64_t num_movable, float bw, float bh, float target_density,
                       float die_w, float die_h, float2 *csize, float2 *coff, float *ratio,
                       float *area, float *npins_out, float2 *hi, hipStream_t s);

// ------------------------------------------------------------ batched execution
// Everything one placement's per-iteration kernels read, resident in device
// memory.  Every per-iteration kernel takes an array of these and serves job
// blockIdx.y (or blockIdx.x for one-workgroup kernels), so jobs that share a bin
// grid run as one launch per kernel; a single placement is a batch of one.
struct JobDesc {
    // 0: an empty slot of a serving batch (every batched kernel returns at once).
    // Launches over one job slot skip the check (gridDim 1 in the job
    // dimension): a single context, or a batch of one, which is never empty.
    int32_t active = 1;
    WaNets wa;
    SortBufs sb;
    PotentialBufs pb;
    CellBufs cb;
    ScalarArgs sa;
};

// Launch geometry of a batch: the largest per-job work of each kernel.
struct BatchDims {
    int m = 0, n = 0;                   // common bin grid
    int64_t max_cells = 0;              // movable cells
    int64_t max_n2 = 0;                 // two-pin nets
    int64_t max_bucket[kWaBuckets] = {};
    int max_reach = 1;                  // owner scatter neighbourhood
    void add(const JobDesc &j);         // widen to include job j
};

int launch_wa(const JobDesc *d, const BatchDims &bd, int nj, hipStream_t s,
              KernelTimer *t = nullptr);
// the degree buckets above 256 only (k_wa_wide); launch_wa runs both
int launch_wa_wide(const JobDesc *d, const BatchDims &bd, int nj, hipStream_t s,
                   KernelTimer *t = nullptr);
int launch_sort(const JobDesc *d, const BatchDims &bd, int nj, bool owner_items, hipStream_t s,
                KernelTimer *t = nullptr);   // owner_items: also build the owner-scatter work list
// copy_perm: the sorted cell ids (aux.w) are copied back into perm, and the
// per-cell constants into sstat / sratio, as the next sort's input order
int launch_scatter(const JobDesc *d, const BatchDims &bd, int nj, bool copy_perm, hipStream_t s);
// Density scatter of the sorted cells, one workgroup per bin tile (plain stores
// of the tile's bins; atomics only for tiles split over several workgroups), and
// the sorted order copied back as the next sort's input.
int launch_scatter_owner(const JobDesc *d, const BatchDims &bd, int nj, hipStream_t s);
int launch_pin2cell(const JobDesc *d, const BatchDims &bd, int nj, hipStream_t s);
int launch_efield(const JobDesc *d, const BatchDims &bd, int nj, hipStream_t s);
int launch_l1(const JobDesc *d, const BatchDims &bd, int nj, hipStream_t s);
int launch_combine(const JobDesc *d, const BatchDims &bd, int nj, hipStream_t s);
// Field force per cell gathered in the sorted order of this iteration's sort
// (rec / aux must hold it), written to gd in id order.  with_sched: one more
// workgroup per job runs launch_scalar's SC_HPWL | SC_POT | SC_SCHED_STAGE
// beside it (the WA, overflow and energy partials must be complete).
int launch_field_sorted(const JobDesc *d, const BatchDims &bd, int nj, hipStream_t s,
                        bool with_sched = false);
// read_gd: the field force was gathered by launch_field_sorted already
int launch_cell_fused(const JobDesc *d, const BatchDims &bd, int nj, bool store_parts,
                      hipStream_t s, bool read_gd = false);
int launch_update(const JobDesc *d, const BatchDims &bd, int nj, hipStream_t s);
// out[3j..3j+2] = hpwl, overflow, nonfinite of job j
int launch_gather_state(const JobDesc *d, int nj, double *out, hipStream_t s);
int launch_scalar(const JobDesc *d, int nj, int ops, hipStream_t s);
// with_wa: pass 1 shares its launch with the WA net groups up to degree 256
// (the groups above run just before)
int launch_potential_fft(const JobDesc *d, const BatchDims &bd, int nj, hipStream_t s,
                         KernelTimer *t = nullptr, bool with_wa = false);

}  // namespace erp
