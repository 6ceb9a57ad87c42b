// The placement context (definition shared by the C ABI and the serving loop).
#pragma once
#include "erp_internal.h"
#include "erp_kernels.h"

using namespace erp;

struct erp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t N = 0, C = 0, F = 0, P = 0, E = 0, n2 = 0, Pm = 0;
    int m = 0, n = 0;
    int64_t B = 0;
    float bw = 0, bh = 0, die_w = 0, die_h = 0, t = 0;
    double movable_area = 0, fx_scale = 1, coef = 0, gamma0 = 0, dw0 = 0;
    int64_t coef_q = 0;
    int spectral_mode = ERP_SPECTRAL_FFT;
    int64_t iter_host = 0;
    bool debug_parts = true;   // keep per-cell WA / density gradients in the fused loop
    bool keep_q = false;       // keep the movable accumulators after pass 1
    bool q_dirty = false;
    // device buffers: carved from one slab sized by the admission estimate
    // (declared first, released last)
    DevBuf slab_mem;
    Slab slab;
    DevBuf d_x, d_y, d_w, d_h, d_npins_all;
    DevBuf pos, prev_pos, prev_g, g, gwl, gd;
    DevBuf csize, coff, ratio, area, npins, hi, cgeo, cprop;
    DevBuf slot_node, slot_off, net_start, slot_grad, cstart, cslots;
    DevBuf q_mov, q_fix, rho, rowt, p1, p2, p3, fxy, pot, auv;
    DevBuf hpwl_part, ovf_part, energy_part, part_a, part_b;
    DevBuf state, gamma_override, hpwl_dbu;
    DevBuf keys, perm, hist, cursor, tstart, items, rec, aux, sstat, sratio, sll;
    bool owner_scatter = true;     // every movable cell reaches at most 2 tiles past its own
    bool owner_auto = true;        // owner_scatter as chosen at creation (ERP_SCATTER_AUTO)
    int reach = 1;                 // tiles a movable cell reaches past its own
    SpecTables tm, tn;
    MfmaPlan *mfma = nullptr;
    JobDesc job{};                 // host copy of the job descriptor
    DevBuf djob, djob_gamma;       // device copies (the second reads an explicit gamma)
    BatchDims dims;                // launch geometry of this placement alone
    int n_hpwl_part = 0, n_pot_part = 0, n_ovf_part = 0, n_energy_part = 0, n_cell_part = 0;
    hipGraphExec_t graph_many = nullptr, graph_one = nullptr;
    KernelTimer *timer = nullptr;
    bool timing = false;
};


namespace erp {
// Context helpers defined in erp_api.hip.
int upload_job(erp_ctx *ctx);
int read_info(erp_ctx *ctx, erp_step_info *info);
// `iterations` steps with eager launches (erp_run without graph capture).
int run_eager(erp_ctx *ctx, int64_t iterations, erp_step_info *info);
// Launch one fused iteration of a batch of jobs that share a bin grid.
int launch_batched_iteration(const JobDesc *d, const BatchDims &bd, int nj, hipStream_t s);
// Device memory a context for this design/config would hold.
size_t estimate_ctx_bytes(const erp_design *d, const erp_config *cfg);
size_t ctx_device_bytes(const erp_ctx *ctx);
}  // namespace erp
