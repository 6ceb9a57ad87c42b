// Launchers of the per-iteration kernels (definitions in erp_density.hip,
// erp_wirelength.hip, erp_optimizer.hip).
#pragma once
#include "erp_internal.h"

namespace erp {

// ------------------------------------------------------------ density
constexpr int kScatterPer = 4;                       // cells per thread
constexpr int kScatterCells = kScatterPer * kBlock;   // cells per workgroup
constexpr int kScatterTile = 4096;                    // LDS int64 bins per workgroup
constexpr int kScatterSide = 64;                      // tile side when the box is too big
constexpr int kSortTiles = 4096;                      // max sort tiles (64 x 64 Morton)
// rec = (x + ox, y + oy, cw, ch) and rratio per cell, in processing order;
// perm_src -> perm_dst (both nullable) is copied on the way.
int launch_scatter(const float4 *rec, const float *rratio, int64_t num_cells, int m, int n,
                   float bw, float bh, double scale, const int32_t *perm_src, int32_t *perm_dst,
                   int64_t *q, hipStream_t s);
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
    int32_t *perm_out;   // [C] sorted order (copied back into perm by the scatter)
    int32_t *keys, *hist, *cursor;   // keys [C], hist/cursor [kSortTiles]
    float4 *rec;         // [C] sorted scatter records
    float *rratio;       // [C]
};
int launch_sort(const SortBufs &b, hipStream_t s, KernelTimer *t = nullptr);
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
    float2 *slot_grad;          // [P] output
    double *hpwl_part;          // partial sums, one per workgroup (see wa_partials)
};
int wa_partials(const WaNets &w);
int launch_wa(const WaNets &w, const float2 *pos, const double *gamma_dev, hipStream_t s,
              KernelTimer *t = nullptr);
// Exact HPWL in int64 DBU over all nets.
int launch_hpwl_dbu(const WaNets &w, const float2 *pos, int64_t num_nets, unsigned long long *out,
                    hipStream_t s);

// ------------------------------------------------------------ optimizer
struct CellBufs {
    int64_t C;
    const float2 *pos;        // [N] (movable first)
    const float2 *csize, *coff;
    const float *ratio;
    const float *area, *npins;  // [C] w*h and num_pins
    const float2 *hi;           // [C] (W - w - 1, H - h - 1)
    const int32_t *cstart, *cslots;   // cell -> pin slots CSR (macro_2_pins order)
    const float2 *slot_grad;
    float2 *gwl, *gd, *g;       // [C]
    float2 *prev_pos, *prev_g;  // [C]
    float2 *pos_out;            // == pos (movable part written)
    double *part_a, *part_b;    // per-workgroup FP64 partials
    const float *fx, *fy;       // field maps
    int m, n;
    float bw, bh;
};
int cell_partials(int64_t C);
// pin -> cell gather of the WA gradient (hpwl.rs:113-135)
int launch_pin2cell(const CellBufs &c, hipStream_t s);
// field gather (eforce.rs:15-91)
int launch_efield(const CellBufs &c, hipStream_t s);
// L1 norms for the first iteration's density-weight scaling (nesterov.rs:47-55)
int launch_l1(const CellBufs &c, hipStream_t s);
// combined preconditioned gradient + BB dot-product partials (nesterov.rs:74-95)
int launch_combine(const CellBufs &c, const DevState *st, hipStream_t s);
// fused per-cell pass: pin->cell gather, field gather, combine (iterations > 0)
int launch_cell_fused(const CellBufs &c, const DevState *st, bool store_parts, hipStream_t s);
// look-ahead step with die clamp and history (nesterov.rs:111-145)
int launch_update(const CellBufs &c, const DevState *st, hipStream_t s);

// Scalar work on the device state (one workgroup).
enum ScalarOps {
    SC_HPWL = 1,    // hpwl = sum(hpwl partials)
    SC_POT = 2,     // overflow, energy from their partials
    SC_L1 = 4,      // density_weight *= l1_wl / l1_d (iteration 0)
    SC_NEST = 8,    // update_momentum, then BB step size
    SC_SCHED = 16,  // Scheduler::update
};
struct ScalarArgs {
    DevState *st;
    const double *hpwl_part; int n_hpwl;
    const int64_t *ovf_part; int n_ovf;
    const double *energy_part; int n_energy;
    const double *part_a, *part_b; int n_cell;
    double inv_scale_over_area;   // 1 / (S * movable_area)
};
int launch_scalar(const ScalarArgs &a, int ops, hipStream_t s);
int launch_reset_state(DevState *st, double gamma0, double dw0, hipStream_t s);
int launch_init_positions(float2 *pos, const float *w, const float *h, const float *x,
                          const float *y, int64_t num_nodes, int64_t num_movable, float die_w,
                          float die_h, float2 *prev_pos, float2 *prev_g, hipStream_t s);
int launch_setup_cells(const float *w, const float *h, const float *npins, int64_t num_nodes,
                       int64_t num_movable, float bw, float bh, float target_density,
                       float die_w, float die_h, float2 *csize, float2 *coff, float *ratio,
                       float *area, float *npins_out, float2 *hi, hipStream_t s);

}  // namespace erp
