// C ABI of liberp_gpu.so: the placement context (Placement::new and
// Placement::step of placement.rs:159-228 on the device), the per-phase entry
// points, the hipGraph-replayed iteration loop and the stand-alone leaf kernels.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>

#include "erp_ctx.h"
#include "erp_timer.h"

namespace erp {

static thread_local std::string g_error;
void set_error(const std::string &msg) { g_error = msg; }
int fail(int code, const std::string &msg) {
    set_error(msg);
    return code;
}

constexpr int kGraphIters = 8;       // iterations per captured graph

double fx_scale(double mass_bound) {
    if (!(mass_bound > 0)) return 1.0;
    int e;
    std::frexp(4611686018427387904.0 / mass_bound, &e);
    return std::ldexp(1.0, e - 1);
}

}  // namespace erp

using namespace erp;

namespace {

int check_grid(int v) {
    return v >= kMinBins && v <= kMaxBins && is_pow2(v);
}

struct DeviceGuard {
    int rc = ERP_OK;
    explicit DeviceGuard(int dev) {
        hipError_t e = hipSetDevice(dev);
        if (e != hipSuccess) rc = fail(ERP_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    }
};

template <class T> int upload(DevBuf &b, const T *src, size_t count, hipStream_t s) {
    int rc = b.alloc(sizeof(T) * count);
    if (rc) return rc;
    if (count) ERP_HIP_TRY(hipMemcpyAsync(b.p, src, sizeof(T) * count, hipMemcpyHostToDevice, s));
    return ERP_OK;
}

int validate(const erp_design *d, const erp_config *c) {
    if (!d || !c) return fail(ERP_ERR_INVALID, "null design or config");
    if (!check_grid(c->num_bins_x) || !check_grid(c->num_bins_y))
        return fail(ERP_ERR_UNSUPPORTED, "bin grid must be a power of two in [64, 2048] per axis");
    if (d->num_nodes <= 0 || d->num_movable < 0 || d->num_movable > d->num_nodes)
        return fail(ERP_ERR_INVALID, "bad num_nodes / num_movable");
    if (d->num_nodes > INT32_MAX || d->num_pins_total > INT32_MAX)
        return fail(ERP_ERR_UNSUPPORTED, "more than 2^31 instances or pins");
    if (!d->x || !d->y || !d->w || !d->h || !d->num_pins)
        return fail(ERP_ERR_INVALID, "missing instance arrays");
    if (d->num_pins_total > 0 && (!d->pin_offset_x || !d->pin_offset_y || !d->pin_node))
        return fail(ERP_ERR_INVALID, "missing pin arrays");
    if (d->num_nets < 0 || d->n_2nets < 0 || d->n_2nets > d->num_nets)
        return fail(ERP_ERR_INVALID, "bad num_nets / n_2nets");
    if (d->num_nets > 0 && (!d->net_start || !d->net_pins))
        return fail(ERP_ERR_INVALID, "missing net arrays");
    if (!d->node_pin_start || !d->node_pins) return fail(ERP_ERR_INVALID, "missing macro_2_pins");
    if (!(d->die_w > 0) || !(d->die_h > 0)) return fail(ERP_ERR_INVALID, "empty die");
    if (!(d->movable_area >= 0)) return fail(ERP_ERR_INVALID, "bad movable_area");
    return ERP_OK;
}

int timer_kernel_names(erp_ctx *ctx, double *ms, int32_t cap, int32_t *count, const char **names) {
    if (!ctx->timer) { *count = 0; return ERP_OK; }
    return ctx->timer->report(ms, cap, count, names);
}

// ---------------------------------------------------------------- phases
const JobDesc *djob(erp_ctx *ctx) { return ctx->djob.as<JobDesc>(); }

}  // namespace

int erp::upload_job(erp_ctx *ctx) {
    ERP_HIP_TRY(hipMemcpyAsync(ctx->djob.p, &ctx->job, sizeof(JobDesc), hipMemcpyHostToDevice, ctx->stream));
    JobDesc g = ctx->job;
    g.wa.gamma = ctx->gamma_override.as<double>();
    ERP_HIP_TRY(hipMemcpyAsync(ctx->djob_gamma.p, &g, sizeof(JobDesc), hipMemcpyHostToDevice, ctx->stream));
    ERP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return ERP_OK;
}

namespace {

int phase_hpwl(erp_ctx *ctx, const JobDesc *d) {
    KernelTimer *t = ctx->timing ? ctx->timer : nullptr;
    int rc = launch_wa(d, ctx->dims, 1, ctx->stream, t);
    if (rc) return rc;
    if ((rc = launch_pin2cell(d, ctx->dims, 1, ctx->stream))) return rc;
    return launch_scalar(d, 1, SC_HPWL, ctx->stream);
}

// Clears the movable accumulators if a step in keep mode left them filled.
int clear_kept_density(erp_ctx *ctx) {
    if (!ctx->q_dirty) return ERP_OK;
    int rc = launch_zero_i64(ctx->q_mov.as<int64_t>(), ctx->B, ctx->stream);
    if (rc == ERP_OK) ctx->q_dirty = false;
    return rc;
}

// compute_potential (epotential.rs:225-374) in two halves: the movable density
// (sort + scatter), then the spectral solve (density map -> field maps).
int run_scatter(erp_ctx *ctx) {
    KernelTimer *t = ctx->timing ? ctx->timer : nullptr;
    int rc;
    if ((rc = clear_kept_density(ctx))) return rc;
    if ((rc = launch_sort(djob(ctx), ctx->dims, 1, ctx->owner_scatter, ctx->stream, t))) return rc;
    if (t) t->begin("scatter");
    rc = ctx->owner_scatter ? launch_scatter_owner(djob(ctx), ctx->dims, 1, ctx->stream)
                            : launch_scatter(djob(ctx), ctx->dims, 1, true, ctx->stream);
    if (t) t->end();
    return rc;
}

int run_spectral(erp_ctx *ctx) {
    KernelTimer *t = ctx->timing ? ctx->timer : nullptr;
    int rc;
    if (ctx->spectral_mode == ERP_SPECTRAL_MFMA_BF16X3)
        rc = launch_potential_mfma(ctx->job.pb, ctx->mfma, ctx->stream, t);
    else
        rc = launch_potential_fft(djob(ctx), ctx->dims, 1, ctx->stream, t);
    if (ctx->keep_q) ctx->q_dirty = true;
    return rc;
}

int run_potential(erp_ctx *ctx) {
    int rc = run_scatter(ctx);
    return rc ? rc : run_spectral(ctx);
}

int phase_potential(erp_ctx *ctx) {
    int rc = run_potential(ctx);
    if (rc) return rc;
    if ((rc = launch_efield(djob(ctx), ctx->dims, 1, ctx->stream))) return rc;
    return launch_scalar(djob(ctx), 1, SC_POT, ctx->stream);
}

int phase_nesterov(erp_ctx *ctx) {
    int rc;
    if (ctx->iter_host == 0) {
        if ((rc = launch_l1(djob(ctx), ctx->dims, 1, ctx->stream))) return rc;
        if ((rc = launch_scalar(djob(ctx), 1, SC_L1, ctx->stream))) return rc;
    }
    if ((rc = launch_combine(djob(ctx), ctx->dims, 1, ctx->stream))) return rc;
    if ((rc = launch_scalar(djob(ctx), 1, SC_NEST, ctx->stream))) return rc;
    return launch_update(djob(ctx), ctx->dims, 1, ctx->stream);
}

int phase_scheduler(erp_ctx *ctx) {
    int rc = launch_scalar(djob(ctx), 1, SC_SCHED, ctx->stream);
    if (rc == ERP_OK) ctx->iter_host += 1;
    return rc;
}

// One iteration (iteration >= 1) as a fused kernel sequence: what the graphs
// hold.  The WA kernel shares the first spectral launch on the one stream (no
// fork and join): pass 1 in the FFT form (k_pass1_wa), the density finalize in
// the MFMA form (k_finalize_wa); each leaves room on every CU for the WA
// workgroups, and the cell kernel then sums the pin gradients itself.  WA only
// reads the positions; the results are bit-identical to the phase sequence.
int launch_fused_iteration(erp_ctx *ctx) {
    KernelTimer *t = ctx->timing ? ctx->timer : nullptr;
    int rc;
    if ((rc = run_scatter(ctx))) return rc;
    if (ctx->spectral_mode == ERP_SPECTRAL_MFMA_BF16X3)
        rc = launch_potential_mfma(ctx->job.pb, ctx->mfma, ctx->stream, t, djob(ctx), &ctx->dims);
    else
        rc = launch_potential_fft(djob(ctx), ctx->dims, 1, ctx->stream, t, /*with_wa=*/true);
    if (ctx->keep_q) ctx->q_dirty = true;
    if (rc) return rc;
    // every form above ran this iteration's sort: the field gather follows its
    // order; the step's HPWL / overflow / energy and the scheduler's staged half
    // run beside it, the Nesterov update and the commit after the cell kernel
    if (t) t->begin("field_sorted");
    rc = launch_field_sorted(djob(ctx), ctx->dims, 1, ctx->stream, /*with_sched=*/true);
    if (t) t->end();
    if (rc) return rc;
    if (t) t->begin("cell_fused");
    rc = launch_cell_fused(djob(ctx), ctx->dims, 1, ctx->debug_parts, ctx->stream, /*read_gd=*/true);
    if (t) t->end();
    if (rc) return rc;
    if (t) t->begin("scalar");
    rc = launch_scalar(djob(ctx), 1, SC_NEST | SC_COMMIT, ctx->stream);
    if (t) t->end();
    if (rc) return rc;
    if (t) t->begin("update");
    rc = launch_update(djob(ctx), ctx->dims, 1, ctx->stream);
    if (t) t->end();
    return rc;
}

void drop_graphs(erp_ctx *ctx) {
    if (ctx->graph_many) (void)hipGraphExecDestroy(ctx->graph_many);
    if (ctx->graph_one) (void)hipGraphExecDestroy(ctx->graph_one);
    ctx->graph_many = ctx->graph_one = nullptr;
}

int capture(erp_ctx *ctx, int iters, hipGraphExec_t *out) {
    ERP_HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = ERP_OK;
    for (int k = 0; k < iters && rc == ERP_OK; ++k) rc = launch_fused_iteration(ctx);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return fail(ERP_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(ERP_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    return ERP_OK;
}

}  // namespace

int erp::read_info(erp_ctx *ctx, erp_step_info *info) {
    if (!info) return ERP_OK;
    DevState s;
    ERP_HIP_TRY(hipMemcpyAsync(&s, ctx->state.p, sizeof(DevState), hipMemcpyDeviceToHost, ctx->stream));
    ERP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    info->hpwl = s.hpwl;
    info->overflow = s.overflow;
    info->energy = s.energy;
    info->density_weight = s.density_weight;
    info->gamma = s.gamma;
    info->alpha = s.alpha;
    info->beta = s.beta;
    info->iteration = s.iteration;
    info->should_stop = s.should_stop;
    info->nonfinite = s.nonfinite;
    return ERP_OK;
}

namespace {

int reset_ctx(erp_ctx *ctx) {
    int rc = launch_init_positions(ctx->pos.as<float2>(), ctx->d_w.as<float>(), ctx->d_h.as<float>(),
                                   ctx->d_x.as<float>(), ctx->d_y.as<float>(), ctx->N, ctx->C,
                                   ctx->die_w, ctx->die_h, ctx->prev_pos.as<float2>(),
                                   ctx->prev_g.as<float2>(), ctx->stream);
    if (rc) return rc;
    if ((rc = launch_reset_state(ctx->state.as<DevState>(), ctx->gamma0, ctx->dw0, ctx->stream))) return rc;
    if ((rc = launch_zero_i64(ctx->q_mov.as<int64_t>(), ctx->B, ctx->stream))) return rc;
    ctx->q_dirty = false;
    ctx->iter_host = 0;
    ERP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return ERP_OK;
}

}  // namespace

int erp::launch_batched_iteration(const JobDesc *d, const BatchDims &bd, int nj, hipStream_t s) {
    const bool owner = bd.max_reach <= 2 && (double)bd.max_cells < 0.75 * bd.m * bd.n;
    int rc = launch_sort(d, bd, nj, owner, s);
    if (!rc) rc = owner ? launch_scatter_owner(d, bd, nj, s) : launch_scatter(d, bd, nj, true, s);
    if (!rc) rc = launch_potential_fft(d, bd, nj, s, nullptr, /*with_wa=*/true);   // WA inside pass 1
    if (!rc) rc = launch_field_sorted(d, bd, nj, s, /*with_sched=*/true);
    if (!rc) rc = launch_cell_fused(d, bd, nj, false, s, /*read_gd=*/true);
    if (!rc) rc = launch_scalar(d, nj, SC_NEST | SC_COMMIT, s);
    if (!rc) rc = launch_update(d, bd, nj, s);
    return rc;
}

// erp_run without graphs (no capture cost; bit-identical results): the serving
// loop's first chunk of each job.
int erp::run_eager(erp_ctx *ctx, int64_t iterations, erp_step_info *info) {
    DeviceGuard g(ctx->device);
    if (g.rc) return g.rc;
    int rc = ERP_OK;
    while (iterations > 0 && ctx->iter_host == 0) {
        rc = phase_hpwl(ctx, djob(ctx));
        if (!rc) rc = phase_potential(ctx);
        if (!rc) rc = phase_nesterov(ctx);
        if (!rc) rc = phase_scheduler(ctx);
        if (rc) return rc;
        --iterations;
    }
    for (int64_t k = 0; k < iterations && !rc; ++k) {
        rc = launch_fused_iteration(ctx);
        ctx->iter_host += 1;
    }
    if (rc) return rc;
    if (info) return read_info(ctx, info);
    ERP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return ERP_OK;
}

size_t erp::estimate_ctx_bytes(const erp_design *d, const erp_config *cfg) {
    const size_t N = (size_t)d->num_nodes, C = (size_t)d->num_movable, P = (size_t)d->num_pins_total;
    const size_t E = (size_t)d->num_nets, B = (size_t)cfg->num_bins_x * (size_t)cfg->num_bins_y;
    const size_t per_node = 5 * 4 + 3 * 8 + 4;                       // DB copies, pos/csize/coff, ratio
    const size_t per_cell = 6 * 8 + 2 * 4 + 8 + 4 * 4 + 16 + 4 + 4 + 36 + 32;  // state, area/pins, hi, sort, packed consts
    const size_t per_pin = 4 + 8 + 8 + 4;                             // slots, grads, cell->slot
    const size_t per_bin = 2 * 8 + 9 * 4;                             // q_mov/q_fix + float maps
    const size_t partials = 16 * (P / 64 + C / 128 + (size_t)cfg->num_bins_x) + 64 * 1024;
    size_t mfma = 0;
    if (cfg->spectral_mode == ERP_SPECTRAL_MFMA_BF16X3)     // operator planes and split maps
        mfma = mfma_plan_estimate((int)cfg->num_bins_x, (int)cfg->num_bins_y) + 8 * B / 4096 * 4;
    return N * per_node + C * per_cell + P * per_pin + E * 4 + B * per_bin + partials + mfma +
           2 * 8 * (size_t)(cfg->num_bins_x + cfg->num_bins_y) + (1u << 20);
}

size_t erp::ctx_device_bytes(const erp_ctx *c) {
    const DevBuf *bufs[] = {&c->d_x, &c->d_y, &c->d_w, &c->d_h, &c->d_npins_all, &c->pos,
                            &c->prev_pos, &c->prev_g, &c->g, &c->gwl, &c->gd, &c->csize, &c->coff,
                            &c->ratio, &c->area, &c->npins, &c->hi, &c->cgeo, &c->cprop, &c->slot_node, &c->slot_off,
                            &c->net_start, &c->slot_grad, &c->cstart, &c->cslots, &c->q_mov,
                            &c->q_fix, &c->rho, &c->rowt, &c->p1, &c->p2, &c->p3, &c->fxy,
                            &c->pot, &c->auv, &c->hpwl_part, &c->ovf_part, &c->energy_part,
                            &c->part_a, &c->part_b, &c->state, &c->gamma_override, &c->hpwl_dbu,
                            &c->keys, &c->perm, &c->hist, &c->cursor, &c->tstart, &c->items, &c->rec,
                            &c->aux, &c->sstat, &c->sratio, &c->sll, &c->djob, &c->djob_gamma, &c->tm.fft_tw, &c->tm.dct_tw,
                            &c->tn.fft_tw, &c->tn.dct_tw};
    size_t total = mfma_plan_bytes(c->mfma);
    for (const DevBuf *b : bufs) total += b->bytes;
    return total;
}

extern "C" {

const char *erp_last_error(void) { return g_error.c_str(); }

int erp_device_count(int *count) {
    if (!count) return fail(ERP_ERR_INVALID, "null count");
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) { *count = 0; return fail(ERP_ERR_NO_DEVICE, hipGetErrorString(e)); }
    return ERP_OK;
}

void erp_config_default(erp_config *cfg) {
    if (!cfg) return;
    cfg->num_bins_x = 512;            // placement.rs:50
    cfg->num_bins_y = 512;
    cfg->target_density = NAN;        // db.cell_utilization (placement.rs:134)
    cfg->gamma_scale = 4.0;           // scheduler.rs:56
    cfg->density_weight0 = 8e-5;      // scheduler.rs:57
    cfg->spectral_mode = ERP_SPECTRAL_FFT;
    cfg->device = 0;
}

int erp_ctx_create(const erp_design *d, const erp_config *cfg, erp_ctx **out) {
    if (!out) return fail(ERP_ERR_INVALID, "null output pointer");
    *out = nullptr;
    int rc = validate(d, cfg);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(ERP_ERR_NO_DEVICE, "no HIP device");
    if (cfg->device < 0 || cfg->device >= ndev) return fail(ERP_ERR_INVALID, "bad device ordinal");
    if (cfg->spectral_mode != ERP_SPECTRAL_FFT && cfg->spectral_mode != ERP_SPECTRAL_MFMA_BF16X3)
        return fail(ERP_ERR_UNSUPPORTED, "unknown spectral mode");
    if (cfg->spectral_mode == ERP_SPECTRAL_MFMA_BF16X3 && !mfma_supported(cfg->num_bins_x, cfg->num_bins_y))
        return fail(ERP_ERR_UNSUPPORTED, "the MFMA spectral form needs at least 128 bins per axis");

    // ----------------------------------------------------- host-side index setup
    const int64_t N = d->num_nodes, C = d->num_movable, P = d->num_pins_total, E = d->num_nets;
    const float die_w = (float)d->die_w, die_h = (float)d->die_h;
    for (int64_t i = 0; i < C; ++i)
        if (!(die_w - d->w[i] - 1.f >= 0.f) || !(die_h - d->h[i] - 1.f >= 0.f))
            return fail(ERP_ERR_INVALID, "movable instance " + std::to_string(i) +
                                             " does not fit the die (W - w - 1 < 0)");
    for (int64_t i = C; i < N; ++i)
        if (!(d->x[i] >= 0.f) || !(d->y[i] >= 0.f))
            return fail(ERP_ERR_INVALID, "fixed instance " + std::to_string(i) + " has a negative coordinate");
    if (E > 0 && (d->net_start[0] != 0)) return fail(ERP_ERR_INVALID, "net_start[0] != 0");
    std::vector<int32_t> slot_of_pin((size_t)P, -1);
    std::vector<int64_t> order;                 // permuted net order
    order.reserve((size_t)E);
    for (int64_t e = 0; e < E; ++e) {
        const int64_t a = d->net_start[e], b = d->net_start[e + 1];
        if (b < a) return fail(ERP_ERR_INVALID, "net_start not monotonic");
        if (b == a) return fail(ERP_ERR_INVALID, "net " + std::to_string(e) + " has no pins");
        if (e < d->n_2nets && b - a != 2)
            return fail(ERP_ERR_INVALID, "net " + std::to_string(e) + " below n_2nets is not two-pin");
    }
    // Net order on the device: two-pin nets first, then multi-pin nets by
    // degree; inside each group by the lowest instance id they touch, so that
    // consecutive nets (one wavefront's work) gather nearby instances and the
    // per-instance gradient gather reads nearby pin slots.  The order only
    // changes the rounding of the FP64 HPWL sum.
    std::vector<int32_t> low_node((size_t)E, INT32_MAX);
    for (int64_t e = 0; e < E; ++e)
        for (int64_t k = d->net_start[e]; k < d->net_start[e + 1]; ++k) {
            const int32_t pin = d->net_pins[k];
            if (pin < 0 || pin >= P) return fail(ERP_ERR_INVALID, "pin id out of range");
            low_node[e] = std::min(low_node[e], d->pin_node[pin]);
        }
    for (int64_t e = 0; e < d->n_2nets; ++e) order.push_back(e);
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return low_node[a] < low_node[b]; });
    std::vector<int64_t> multi;
    for (int64_t e = d->n_2nets; e < E; ++e) multi.push_back(e);
    std::stable_sort(multi.begin(), multi.end(), [&](int64_t a, int64_t b) {
        const int64_t da = d->net_start[a + 1] - d->net_start[a];
        const int64_t db = d->net_start[b + 1] - d->net_start[b];
        return da != db ? da < db : low_node[a] < low_node[b];
    });
    int64_t bucket_count[kWaBuckets] = {0};
    for (int64_t e : multi) {
        order.push_back(e);
        const int64_t deg = d->net_start[e + 1] - d->net_start[e];
        int b = 0;
        while (deg > kWaDegree[b]) ++b;
        ++bucket_count[b];
    }
    std::vector<int32_t> h_slot_node, h_net_start;
    std::vector<float2> h_slot_off;
    h_slot_node.reserve((size_t)P);
    h_slot_off.reserve((size_t)P);
    h_net_start.reserve((size_t)E + 1);
    for (int64_t e : order) {
        h_net_start.push_back((int32_t)h_slot_node.size());
        for (int64_t k = d->net_start[e]; k < d->net_start[e + 1]; ++k) {
            const int32_t pin = d->net_pins[k];
            if (pin < 0 || pin >= P) return fail(ERP_ERR_INVALID, "pin id out of range");
            if (slot_of_pin[pin] >= 0) return fail(ERP_ERR_INVALID, "pin " + std::to_string(pin) + " is in two nets");
            const int32_t node = d->pin_node[pin];
            if (node < 0 || node >= N) return fail(ERP_ERR_INVALID, "pin_node out of range");
            slot_of_pin[pin] = (int32_t)h_slot_node.size();
            h_slot_node.push_back(node);
            h_slot_off.push_back(make_float2(d->pin_offset_x[pin], d->pin_offset_y[pin]));
        }
    }
    h_net_start.push_back((int32_t)h_slot_node.size());
    std::vector<int32_t> h_cstart((size_t)C + 1, 0), h_cslots;
    for (int64_t c = 0; c < C; ++c) {
        h_cstart[c] = (int32_t)h_cslots.size();
        for (int64_t k = d->node_pin_start[c]; k < d->node_pin_start[c + 1]; ++k) {
            const int32_t pin = d->node_pins[k];
            if (pin < 0 || pin >= P || d->pin_node[pin] != c)
                return fail(ERP_ERR_INVALID, "macro_2_pins inconsistent with pin_node");
            if (slot_of_pin[pin] >= 0) h_cslots.push_back(slot_of_pin[pin]);
        }
    }
    h_cstart[C] = (int32_t)h_cslots.size();

    // ------------------------------------------------------------- scalars
    erp_ctx *ctx = new erp_ctx();
    ctx->device = cfg->device;
    DeviceGuard g(ctx->device);
    if (g.rc) { delete ctx; return g.rc; }
    if (hipStreamCreate(&ctx->stream) != hipSuccess) {
        delete ctx;
        return fail(ERP_ERR_HIP, "hipStreamCreate failed");
    }
    hipStream_t s = ctx->stream;
    ctx->N = N; ctx->C = C; ctx->F = N - C; ctx->P = P; ctx->E = E; ctx->n2 = d->n_2nets;
    ctx->Pm = (int64_t)h_cslots.size();
    ctx->m = cfg->num_bins_x; ctx->n = cfg->num_bins_y; ctx->B = (int64_t)ctx->m * ctx->n;
    ctx->die_w = die_w; ctx->die_h = die_h;
    ctx->bw = die_w / (float)ctx->m;                       // placement.rs:52-53
    ctx->bh = die_h / (float)ctx->n;
    ctx->t = std::isnan(cfg->target_density) ? (float)d->cell_utilization : (float)cfg->target_density;
    ctx->movable_area = d->movable_area;
    ctx->gamma0 = cfg->gamma_scale * ((double)ctx->bw + (double)ctx->bh);
    ctx->dw0 = cfg->density_weight0;
    ctx->spectral_mode = cfg->spectral_mode;
    double mass = d->movable_area;
    const double sq2bw = (double)(ctx->bw * 1.41421356237309504880f);
    const double sq2bh = (double)(ctx->bh * 1.41421356237309504880f);
    for (int64_t i = C; i < N; ++i)
        mass += (double)ctx->t * std::max((double)d->w[i], sq2bw) * std::max((double)d->h[i], sq2bh);
    ctx->fx_scale = fx_scale(mass * 1.01 + 1.0);
    ctx->coef = (double)ctx->t * (double)ctx->bw * (double)ctx->bh;
    ctx->coef_q = (int64_t)std::llrint(ctx->coef * ctx->fx_scale);

    // ------------------------------------------------------------- device buffers
    // one slab for every buffer below (one hipMalloc / hipFree per context); a
    // buffer that does not fit gets its own allocation
    {
        const size_t est = estimate_ctx_bytes(d, cfg);
        if (ctx->slab_mem.alloc(est) == ERP_OK) {
            ctx->slab.base = ctx->slab_mem.as<char>();
            ctx->slab.cap = est;
        }
    }
    SlabScope slab_scope(&ctx->slab);
#define TRY(x) do { rc = (x); if (rc) { erp_ctx_destroy(ctx); return rc; } } while (0)
    TRY(upload(ctx->d_x, d->x, (size_t)N, s));
    TRY(upload(ctx->d_y, d->y, (size_t)N, s));
    TRY(upload(ctx->d_w, d->w, (size_t)N, s));
    TRY(upload(ctx->d_h, d->h, (size_t)N, s));
    TRY(upload(ctx->d_npins_all, d->num_pins, (size_t)N, s));
    TRY(upload(ctx->slot_node, h_slot_node.data(), h_slot_node.size(), s));
    TRY(upload(ctx->slot_off, h_slot_off.data(), h_slot_off.size(), s));
    TRY(upload(ctx->net_start, h_net_start.data(), h_net_start.size(), s));
    TRY(upload(ctx->cstart, h_cstart.data(), h_cstart.size(), s));
    TRY(upload(ctx->cslots, h_cslots.data(), h_cslots.size(), s));
    const size_t f2 = sizeof(float2);
    TRY(ctx->pos.alloc(f2 * (size_t)N));
    TRY(ctx->csize.alloc(f2 * (size_t)N));
    TRY(ctx->coff.alloc(f2 * (size_t)N));
    TRY(ctx->ratio.alloc(sizeof(float) * (size_t)N));
    for (DevBuf *b : {&ctx->prev_pos, &ctx->prev_g, &ctx->g, &ctx->gwl, &ctx->gd, &ctx->hi})
        TRY(b->alloc(f2 * (size_t)std::max<int64_t>(C, 1)));
    TRY(ctx->area.alloc(sizeof(float) * (size_t)std::max<int64_t>(C, 1)));
    TRY(ctx->cgeo.alloc(sizeof(float4) * (size_t)std::max<int64_t>(C, 1)));
    TRY(ctx->cprop.alloc(sizeof(float4) * (size_t)std::max<int64_t>(C, 1)));
    TRY(ctx->npins.alloc(sizeof(float) * (size_t)std::max<int64_t>(C, 1)));
    TRY(ctx->slot_grad.alloc(f2 * (size_t)std::max<int64_t>(P, 1)));
    TRY(ctx->q_mov.alloc(sizeof(int64_t) * (size_t)ctx->B));
    TRY(ctx->q_fix.alloc(sizeof(int64_t) * (size_t)ctx->B));
    for (DevBuf *b : {&ctx->rho, &ctx->rowt, &ctx->p1, &ctx->p2, &ctx->p3, &ctx->pot, &ctx->auv})
        TRY(b->alloc(sizeof(float) * (size_t)ctx->B));
    TRY(ctx->fxy.alloc(sizeof(float2) * fxy_bins(ctx->m, ctx->n)));
    TRY(ctx->state.alloc(sizeof(DevState)));
    TRY(ctx->gamma_override.alloc(sizeof(double)));
    TRY(ctx->hpwl_dbu.alloc(sizeof(unsigned long long)));

    // ------------------------------------------------------------- kernel argument blocks
    WaNets &w = ctx->job.wa;
    w.slot_node = ctx->slot_node.as<int32_t>();
    w.slot_off = ctx->slot_off.as<float2>();
    w.net_start = ctx->net_start.as<int32_t>();
    w.n2 = d->n_2nets;
    w.bucket_begin[0] = d->n_2nets;
    for (int b = 0; b < kWaBuckets; ++b) w.bucket_begin[b + 1] = w.bucket_begin[b] + bucket_count[b];
    w.slot_grad = ctx->slot_grad.as<float2>();
    w.pos = ctx->pos.as<float2>();
    w.gamma = &ctx->state.as<DevState>()->gamma;
    ctx->n_hpwl_part = wa_partials(w);
    TRY(ctx->hpwl_part.alloc(sizeof(double) * (size_t)std::max(ctx->n_hpwl_part, 1)));
    w.hpwl_part = ctx->hpwl_part.as<double>();

    ctx->n_pot_part = potential_partials(ctx->m);
    const bool mfma = ctx->spectral_mode == ERP_SPECTRAL_MFMA_BF16X3;
    ctx->n_energy_part = mfma ? mfma_energy_partials(ctx->m, ctx->n) : ctx->n_pot_part;
    // the scalar kernel sums exactly the overflow partials the form's first pass writes
    ctx->n_ovf_part = mfma ? mfma_overflow_partials(ctx->m) : ctx->n_pot_part;
    TRY(ctx->ovf_part.alloc(sizeof(int64_t) * (size_t)std::max(ctx->n_pot_part, ctx->n_ovf_part)));
    TRY(ctx->energy_part.alloc(sizeof(double) * (size_t)ctx->n_energy_part));
    PotentialBufs &pb = ctx->job.pb;
    pb.q_fix = ctx->q_fix.as<int64_t>();
    pb.q_mov = ctx->q_mov.as<int64_t>();
    pb.rho = ctx->rho.as<float>();
    pb.rowt = ctx->rowt.as<float>();
    pb.p1 = ctx->p1.as<float>();
    pb.p2 = ctx->p2.as<float>();
    pb.p3 = ctx->p3.as<float>();
    pb.fxy = ctx->fxy.as<float2>();
    pb.pot = ctx->pot.as<float>();
    pb.auv = ctx->auv.as<float>();
    pb.ovf_part = ctx->ovf_part.as<int64_t>();
    pb.energy_part = ctx->energy_part.as<double>();
    pb.m = ctx->m;
    pb.n = ctx->n;
    pb.inv_scale = 1.0 / ctx->fx_scale;
    pb.coef_q = ctx->coef_q;
    {
        const float bin_area = ctx->bw * ctx->bh;           // epotential.rs:272
        pb.inv_bin_area = 1.f / bin_area;                   // epotential.rs:284
    }
    pb.bw = ctx->bw;
    pb.bh = ctx->bh;
    pb.keep_q = false;
    pb.potential = 1;
    pb.scale = ctx->fx_scale;

    ctx->n_cell_part = cell_partials(C);
    TRY(ctx->part_a.alloc(sizeof(double) * (size_t)ctx->n_cell_part));
    TRY(ctx->part_b.alloc(sizeof(double) * (size_t)ctx->n_cell_part));
    CellBufs &cb = ctx->job.cb;
    cb.C = C;
    cb.pos = ctx->pos.as<float2>();
    cb.csize = ctx->csize.as<float2>();
    cb.coff = ctx->coff.as<float2>();
    cb.ratio = ctx->ratio.as<float>();
    cb.area = ctx->area.as<float>();
    cb.npins = ctx->npins.as<float>();
    cb.cgeo = ctx->cgeo.as<float4>();
    cb.cprop = ctx->cprop.as<float4>();
    cb.hi = ctx->hi.as<float2>();
    cb.cstart = ctx->cstart.as<int32_t>();
    cb.cslots = ctx->cslots.as<int32_t>();
    cb.slot_grad = ctx->slot_grad.as<float2>();
    cb.gwl = ctx->gwl.as<float2>();
    cb.gd = ctx->gd.as<float2>();
    cb.g = ctx->g.as<float2>();
    cb.prev_pos = ctx->prev_pos.as<float2>();
    cb.prev_g = ctx->prev_g.as<float2>();
    cb.pos_out = ctx->pos.as<float2>();
    cb.part_a = ctx->part_a.as<double>();
    cb.part_b = ctx->part_b.as<double>();
    cb.fxy = ctx->fxy.as<float2>();
    cb.m = ctx->m;
    cb.n = ctx->n;
    cb.bw = ctx->bw;
    cb.bh = ctx->bh;
    cb.st = ctx->state.as<DevState>();

    ScalarArgs &sa = ctx->job.sa;
    sa.st = ctx->state.as<DevState>();
    sa.hpwl_part = ctx->hpwl_part.as<double>();
    sa.n_hpwl = ctx->n_hpwl_part;
    sa.ovf_part = ctx->ovf_part.as<int64_t>();
    sa.n_ovf = ctx->n_ovf_part;
    sa.energy_part = ctx->energy_part.as<double>();
    sa.n_energy = ctx->n_energy_part;
    sa.part_a = ctx->part_a.as<double>();
    sa.part_b = ctx->part_b.as<double>();
    sa.n_cell = ctx->n_cell_part;
    sa.inv_scale_over_area = 1.0 / (ctx->fx_scale * ctx->movable_area);

    // ------------------------------------------------------------- device setup kernels
    TRY(ctx->keys.alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(C, 1)));
    TRY(ctx->perm.alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(C, 1)));
    TRY(ctx->rec.alloc(sizeof(float4) * (size_t)std::max<int64_t>(C, 1)));
    TRY(ctx->aux.alloc(sizeof(float4) * (size_t)std::max<int64_t>(C, 1)));
    TRY(ctx->sstat.alloc(sizeof(float4) * (size_t)std::max<int64_t>(C, 1)));
    TRY(ctx->sratio.alloc(sizeof(float) * (size_t)std::max<int64_t>(C, 1)));
    TRY(ctx->sll.alloc(sizeof(float2) * (size_t)std::max<int64_t>(C, 1)));
    TRY(launch_iota(ctx->perm.as<int32_t>(), C, s));
    TRY(ctx->hist.alloc(sizeof(int32_t) * (size_t)kSortTiles));
    TRY(ctx->cursor.alloc(sizeof(int32_t) * (size_t)kSortTiles));
    TRY(ctx->tstart.alloc(sizeof(int32_t) * (size_t)(kSortTiles + 1)));
    TRY(ctx->items.alloc(sizeof(int32_t) * (size_t)(kSortTiles + 1)));
    {
        SortBufs &sb = ctx->job.sb;
        sb.num_cells = C;
        sb.pos = ctx->pos.as<float2>();
        sb.csize = ctx->csize.as<float2>();
        sb.coff = ctx->coff.as<float2>();
        sb.ratio = ctx->ratio.as<float>();
        sb.m = ctx->m;
        sb.n = ctx->n;
        sb.bw = ctx->bw;
        sb.bh = ctx->bh;
        sb.keys = ctx->keys.as<int32_t>();
        sb.hist = ctx->hist.as<int32_t>();
        sb.cursor = ctx->cursor.as<int32_t>();
        sb.perm = ctx->perm.as<int32_t>();

        sb.rec = ctx->rec.as<float4>();
        sb.aux = ctx->aux.as<float4>();
        sb.sstat = ctx->sstat.as<float4>();
        sb.sratio = ctx->sratio.as<float>();
        sb.ll = ctx->sll.as<float2>();

        sb.tstart = ctx->tstart.as<int32_t>();
        sb.items = ctx->items.as<int32_t>();
        // tiles a movable cell can reach past its own: its clamped size in bins
        // (+1 for the bin range's extra bin and rounding) over the tile side
        float wmax = 0.f, hmax = 0.f;
        for (int64_t i = 0; i < C; ++i) {
            wmax = std::max(wmax, d->w[i]);
            hmax = std::max(hmax, d->h[i]);
        }
        const float sq2 = 1.41421356237309504880f;
        const double cw = std::max((double)wmax, (double)(ctx->bw * sq2));
        const double ch = std::max((double)hmax, (double)(ctx->bh * sq2));
        const int T = 1 << sort_geometry(ctx->m, ctx->n).shift;
        const int rx = 1 + (int)((std::ceil(cw / ctx->bw) + 1.0) / T);
        const int ry = 1 + (int)((std::ceil(ch / ctx->bh) + 1.0) / T);
        sb.reach = std::max(rx, ry);
        // The owner scatter reads every cell (reach + 1)^2 times but writes each
        // bin once without atomics; the cell-major scatter reads each cell once
        // but issues about one global atomic per touched bin per workgroup.
        // Measured on MI355X: owner wins with fewer cells than ~0.75 per bin
        // (2M cells / 2048^2: 157 -> 88 us), cell-major above (1M / 1024^2:
        // 40 vs 64 us).
        ctx->owner_scatter = sb.reach <= 2 && (double)C < 0.75 * (double)ctx->B;
        ctx->owner_auto = ctx->owner_scatter;
        ctx->reach = sb.reach;
    }
    for (DevBuf *b : {&ctx->hpwl_part, &ctx->ovf_part, &ctx->energy_part, &ctx->part_a, &ctx->part_b, &ctx->hist, &ctx->cursor})
        if (hipMemsetAsync(b->p, 0, b->bytes, s) != hipSuccess) { erp_ctx_destroy(ctx); return fail(ERP_ERR_HIP, "hipMemsetAsync failed"); }
    TRY(launch_setup_cells(ctx->d_w.as<float>(), ctx->d_h.as<float>(), ctx->d_npins_all.as<float>(),
                           N, C, ctx->bw, ctx->bh, ctx->t, die_w, die_h, ctx->csize.as<float2>(),
                           ctx->coff.as<float2>(), ctx->ratio.as<float>(), ctx->area.as<float>(),
                           ctx->npins.as<float>(), ctx->hi.as<float2>(), s));
    TRY(launch_pack_cell_consts(ctx->csize.as<float2>(), ctx->coff.as<float2>(), ctx->ratio.as<float>(),
                                ctx->npins.as<float>(), ctx->area.as<float>(), C, ctx->cgeo.as<float4>(),
                                ctx->cprop.as<float4>(), s));
    TRY(launch_init_sort_static(ctx->csize.as<float2>(), ctx->coff.as<float2>(), ctx->ratio.as<float>(),
                                C, ctx->sstat.as<float4>(), ctx->sratio.as<float>(), s));
    TRY(reset_ctx(ctx));
    TRY(launch_zero_i64(ctx->q_fix.as<int64_t>(), ctx->B, s));
    TRY(launch_fixed_map(ctx->pos.as<float2>() + C, ctx->csize.as<float2>() + C, N - C, ctx->m,
                         ctx->n, ctx->bw, ctx->bh, die_w, die_h, ctx->fx_scale,
                         ctx->q_fix.as<int64_t>(), s));
    TRY(launch_scale_fixed(ctx->q_fix.as<int64_t>(), ctx->B, (double)ctx->t, s));
    TRY(make_spec_tables(ctx->tm, ctx->m, s));
    TRY(make_spec_tables(ctx->tn, ctx->n, s));
    pb.tm_fft = ctx->tm.fft_tw.as<float2>();
    pb.tm_dct = ctx->tm.dct_tw.as<float2>();
    pb.tn_fft = ctx->tn.fft_tw.as<float2>();
    pb.tn_dct = ctx->tn.dct_tw.as<float2>();
    TRY(ctx->djob.alloc(sizeof(JobDesc)));
    TRY(ctx->djob_gamma.alloc(sizeof(JobDesc)));
    ctx->dims = BatchDims();
    ctx->dims.add(ctx->job);
    TRY(upload_job(ctx));
    if (ctx->spectral_mode == ERP_SPECTRAL_MFMA_BF16X3)
        TRY(mfma_plan_create(&ctx->mfma, ctx->m, ctx->n, ctx->job.pb.bw / ctx->job.pb.bh, s));
    ctx->timer = new KernelTimer(s);
    if (hipStreamSynchronize(s) != hipSuccess) { erp_ctx_destroy(ctx); return fail(ERP_ERR_HIP, "setup failed"); }
#undef TRY
    *out = ctx;
    return ERP_OK;
}

int erp_ctx_destroy(erp_ctx *ctx) {
    if (!ctx) return ERP_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    drop_graphs(ctx);
    if (ctx->mfma) mfma_plan_destroy(ctx->mfma);
    delete ctx->timer;
    hipStream_t s = ctx->stream;
    delete ctx;   // releases every DevBuf
    if (s) (void)hipStreamDestroy(s);
    return ERP_OK;
}

int erp_ctx_reset(erp_ctx *ctx) {
    if (!ctx) return fail(ERP_ERR_INVALID, "null context");
    DeviceGuard g(ctx->device);
    if (g.rc) return g.rc;
    return reset_ctx(ctx);
}

int erp_set_positions(erp_ctx *ctx, const float *x, const float *y) {
    if (!ctx || !x || !y) return fail(ERP_ERR_INVALID, "null argument");
    DeviceGuard g(ctx->device);
    if (g.rc) return g.rc;
    std::vector<float2> p((size_t)ctx->N);
    for (int64_t i = 0; i < ctx->N; ++i) p[i] = make_float2(x[i], y[i]);
    ERP_HIP_TRY(hipMemcpyAsync(ctx->pos.p, p.data(), sizeof(float2) * p.size(), hipMemcpyHostToDevice, ctx->stream));
    ERP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return ERP_OK;
}

int erp_get_positions(erp_ctx *ctx, float *x, float *y) {
    if (!ctx || !x || !y) return fail(ERP_ERR_INVALID, "null argument");
    DeviceGuard g(ctx->device);
    if (g.rc) return g.rc;
    std::vector<float2> p((size_t)ctx->N);
    ERP_HIP_TRY(hipMemcpyAsync(p.data(), ctx->pos.p, sizeof(float2) * p.size(), hipMemcpyDeviceToHost, ctx->stream));
    ERP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int64_t i = 0; i < ctx->N; ++i) { x[i] = p[i].x; y[i] = p[i].y; }
    return ERP_OK;
}

static int read_scalar(erp_ctx *ctx, size_t offset, double *out) {
    ERP_HIP_TRY(hipMemcpyAsync(out, (char *)ctx->state.p + offset, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    ERP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return ERP_OK;
}

int erp_compute_hpwl(erp_ctx *ctx, double *hpwl) {
    if (!ctx) return fail(ERP_ERR_INVALID, "null context");
    DeviceGuard g(ctx->device);
    if (g.rc) return g.rc;
    int rc = phase_hpwl(ctx, djob(ctx));
    if (rc || !hpwl) return rc;
    return read_scalar(ctx, offsetof(DevState, hpwl), hpwl);
}

int erp_compute_hpwl_gamma(erp_ctx *ctx, double gamma, double *hpwl) {
    if (!ctx) return fail(ERP_ERR_INVALID, "null context");
    if (!(gamma > 0)) return fail(ERP_ERR_INVALID, "gamma must be positive");
    DeviceGuard g(ctx->device);
    if (g.rc) return g.rc;
    // the copy reads `gamma` from this frame: it completes before the call returns
    ERP_HIP_TRY(hipMemcpyAsync(ctx->gamma_override.p, &gamma, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    int rc = phase_hpwl(ctx, ctx->djob_gamma.as<JobDesc>());
    if (!rc && hpwl) return read_scalar(ctx, offsetof(DevState, hpwl), hpwl);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess && !rc) rc = fail(ERP_ERR_HIP, "hipStreamSynchronize failed");
    return rc;
}

int erp_compute_potential(erp_ctx *ctx, double *energy, double *overflow) {
    if (!ctx) return fail(ERP_ERR_INVALID, "null context");
    DeviceGuard g(ctx->device);
    if (g.rc) return g.rc;
    int rc = phase_potential(ctx);
    if (rc) return rc;
    if (energy && (rc = read_scalar(ctx, offsetof(DevState, energy), energy))) return rc;
    if (overflow && (rc = read_scalar(ctx, offsetof(DevState, overflow), overflow))) return rc;
    return ERP_OK;
}

int erp_nesterov_step(erp_ctx *ctx) {
    if (!ctx) return fail(ERP_ERR_INVALID, "null context");
    DeviceGuard g(ctx->device);
    if (g.rc) return g.rc;
    return phase_nesterov(ctx);
}

int erp_scheduler_update(erp_ctx *ctx) {
    if (!ctx) return fail(ERP_ERR_INVALID, "null context");
    DeviceGuard g(ctx->device);
    if (g.rc) return g.rc;
    return phase_scheduler(ctx);
}

int erp_step(erp_ctx *ctx, erp_step_info *info) {
    if (!ctx) return fail(ERP_ERR_INVALID, "null context");
    DeviceGuard g(ctx->device);
    if (g.rc) return g.rc;
    if (ctx->timing) ctx->timer->clear();
    int rc = phase_hpwl(ctx, djob(ctx));
    if (!rc) rc = phase_potential(ctx);
    if (!rc) rc = phase_nesterov(ctx);
    if (!rc) rc = phase_scheduler(ctx);
    if (rc) return rc;
    if (ctx->timing) { if ((rc = ctx->timer->collect())) return rc; }
    return read_info(ctx, info);
}

int erp_run(erp_ctx *ctx, int64_t iterations, erp_step_info *info) {
    if (!ctx) return fail(ERP_ERR_INVALID, "null context");
    if (iterations < 0) return fail(ERP_ERR_INVALID, "negative iteration count");
    DeviceGuard g(ctx->device);
    if (g.rc) return g.rc;
    int rc = ERP_OK;
    if (ctx->timing) ctx->timer->clear();
    while (iterations > 0 && ctx->iter_host == 0) {   // iteration 0: density-weight scaling
        rc = phase_hpwl(ctx, djob(ctx));
        if (!rc) rc = phase_potential(ctx);
        if (!rc) rc = phase_nesterov(ctx);
        if (!rc) rc = phase_scheduler(ctx);
        if (rc) return rc;
        --iterations;
    }
    if (ctx->timing || ctx->keep_q) {
        for (int64_t k = 0; k < iterations && !rc; ++k) {
            rc = launch_fused_iteration(ctx);
            ctx->iter_host += 1;
            if (!rc && ctx->timing && (k % 16 == 15)) rc = ctx->timer->collect();
        }
        if (!rc && ctx->timing) rc = ctx->timer->collect();
    } else if (iterations > 0) {
        // graphs hold no clearing node: accumulators left by a kept step are cleared first
        if ((rc = clear_kept_density(ctx))) return rc;
        if (!ctx->graph_many && (rc = capture(ctx, kGraphIters, &ctx->graph_many))) return rc;
        if (!ctx->graph_one && (rc = capture(ctx, 1, &ctx->graph_one))) return rc;
        int64_t k = 0;
        for (; k + kGraphIters <= iterations; k += kGraphIters) {
            ERP_HIP_TRY(hipGraphLaunch(ctx->graph_many, ctx->stream));
            ctx->iter_host += kGraphIters;
        }
        for (; k < iterations; ++k) {
            ERP_HIP_TRY(hipGraphLaunch(ctx->graph_one, ctx->stream));
            ctx->iter_host += 1;
        }
    }
    if (rc) return rc;
    if (info) return read_info(ctx, info);
    ERP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return ERP_OK;
}

int erp_run_until(erp_ctx *ctx, int64_t max_iterations, double stop_overflow, int32_t check_every,
                  erp_step_info *info) {
    if (!ctx) return fail(ERP_ERR_INVALID, "null context");
    if (check_every <= 0) check_every = 1;
    erp_step_info tmp{};
    int64_t done = 0;
    while (done < max_iterations) {
        const int64_t chunk = std::min<int64_t>(check_every, max_iterations - done);
        int rc = erp_run(ctx, chunk, &tmp);
        if (rc) return rc;
        done += chunk;
        if (tmp.overflow < stop_overflow || tmp.nonfinite) break;
    }
    if (info) *info = tmp;
    return ERP_OK;
}

int erp_get_step_info(erp_ctx *ctx, erp_step_info *info) {
    if (!ctx || !info) return fail(ERP_ERR_INVALID, "null argument");
    DeviceGuard g(ctx->device);
    if (g.rc) return g.rc;
    return read_info(ctx, info);
}

int erp_get_map(erp_ctx *ctx, int which, float *out) {
    if (!ctx || !out) return fail(ERP_ERR_INVALID, "null argument");
    DeviceGuard g(ctx->device);
    if (g.rc) return g.rc;
    ERP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (which == ERP_MAP_INITIAL) {
        std::vector<int64_t> q((size_t)ctx->B);
        ERP_HIP_TRY(hipMemcpy(q.data(), ctx->q_fix.p, sizeof(int64_t) * q.size(), hipMemcpyDeviceToHost));
        for (int64_t b = 0; b < ctx->B; ++b) out[b] = (float)((double)q[b] / ctx->fx_scale);
        return ERP_OK;
    }
    if (which == ERP_MAP_FIELD_X || which == ERP_MAP_FIELD_Y) {   // one float of each pair
        std::vector<float2> f(fxy_bins(ctx->m, ctx->n));
        ERP_HIP_TRY(hipMemcpy(f.data(), ctx->fxy.p, sizeof(float2) * f.size(), hipMemcpyDeviceToHost));
        for (int k = 0; k < ctx->m; ++k)
            for (int h = 0; h < ctx->n; ++h) {
                const float2 v = f[fxy_index(k, h, ctx->n)];
                out[(size_t)k * ctx->n + h] = which == ERP_MAP_FIELD_Y ? v.y : v.x;
            }
        return ERP_OK;
    }
    const DevBuf *src = which == ERP_MAP_DENSITY ? &ctx->rho : which == ERP_MAP_AUV ? &ctx->auv
                      : which == ERP_MAP_POTENTIAL ? &ctx->pot : nullptr;
    if (!src) return fail(ERP_ERR_INVALID, "unknown map");
    ERP_HIP_TRY(hipMemcpy(out, src->p, sizeof(float) * (size_t)ctx->B, hipMemcpyDeviceToHost));
    return ERP_OK;
}

int erp_get_cell_grads(erp_ctx *ctx, int which, float *gx, float *gy) {
    if (!ctx || !gx || !gy) return fail(ERP_ERR_INVALID, "null argument");
    DeviceGuard g(ctx->device);
    if (g.rc) return g.rc;
    const DevBuf *src = which == ERP_GRAD_WIRELENGTH ? &ctx->gwl : which == ERP_GRAD_DENSITY ? &ctx->gd
                      : which == ERP_GRAD_COMBINED ? &ctx->g : nullptr;
    if (!src) return fail(ERP_ERR_INVALID, "unknown gradient");
    ERP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::vector<float2> v((size_t)ctx->C);
    if (ctx->C) ERP_HIP_TRY(hipMemcpy(v.data(), src->p, sizeof(float2) * v.size(), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < ctx->C; ++i) { gx[i] = v[i].x; gy[i] = v[i].y; }
    return ERP_OK;
}

int erp_get_density_fixed(erp_ctx *ctx, int64_t *q_mov, int64_t *q_fix, double *scale) {
    if (!ctx) return fail(ERP_ERR_INVALID, "null context");
    DeviceGuard g(ctx->device);
    if (g.rc) return g.rc;
    ERP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (q_mov) {
        if (!ctx->q_dirty) return fail(ERP_ERR_INVALID, "movable accumulators were cleared; enable keep mode with erp_set_keep_density");
        ERP_HIP_TRY(hipMemcpy(q_mov, ctx->q_mov.p, sizeof(int64_t) * (size_t)ctx->B, hipMemcpyDeviceToHost));
    }
    if (q_fix) ERP_HIP_TRY(hipMemcpy(q_fix, ctx->q_fix.p, sizeof(int64_t) * (size_t)ctx->B, hipMemcpyDeviceToHost));
    if (scale) *scale = ctx->fx_scale;
    return ERP_OK;
}

int erp_set_keep_density(erp_ctx *ctx, int keep) {
    if (!ctx) return fail(ERP_ERR_INVALID, "null context");
    DeviceGuard g(ctx->device);
    if (g.rc) return g.rc;
    ctx->keep_q = keep != 0;
    ctx->job.pb.keep_q = ctx->keep_q;   // read by pass 1 from the device descriptor
    int rc = upload_job(ctx);
    if (rc) return rc;
    // the accumulators kept by the last step must not seed the next scatter
    return clear_kept_density(ctx);
}

int erp_set_store_gradients(erp_ctx *ctx, int store) {
    if (!ctx) return fail(ERP_ERR_INVALID, "null context");
    DeviceGuard g(ctx->device);
    if (g.rc) return g.rc;
    if (ctx->debug_parts == (store != 0)) return ERP_OK;
    drop_graphs(ctx);
    ctx->debug_parts = store != 0;
    // the spectral pass writes the a_uv map only while the parts are kept
    ctx->job.pb.auv = ctx->debug_parts ? ctx->auv.as<float>() : nullptr;
    return upload_job(ctx);
}

int erp_set_compute_potential(erp_ctx *ctx, int on) {
    if (!ctx) return fail(ERP_ERR_INVALID, "null context");
    DeviceGuard g(ctx->device);
    if (g.rc) return g.rc;
    if ((ctx->job.pb.potential != 0) == (on != 0)) return ERP_OK;
    drop_graphs(ctx);
    ctx->job.pb.potential = on ? 1 : 0;
    ctx->job.sa.n_energy = on ? ctx->n_energy_part : 0;   // energy reads 0 while skipped
    return upload_job(ctx);
}

int erp_set_scatter_kind(erp_ctx *ctx, int kind) {
    if (!ctx) return fail(ERP_ERR_INVALID, "null context");
    bool owner;
    if (kind == ERP_SCATTER_AUTO) owner = ctx->owner_auto;
    else if (kind == ERP_SCATTER_CELL_MAJOR) owner = false;
    else if (kind == ERP_SCATTER_OWNER) owner = true;
    else return fail(ERP_ERR_INVALID, "unknown scatter kind");
    if (owner && ctx->reach > 2)
        return fail(ERP_ERR_UNSUPPORTED, "the owner scatter needs cells that reach at most two tiles past their own");
    if (owner == ctx->owner_scatter) return ERP_OK;
    DeviceGuard g(ctx->device);
    if (g.rc) return g.rc;
    ERP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    drop_graphs(ctx);          // the graphs hold the scatter kernel they were captured with
    ctx->owner_scatter = owner;
    return ERP_OK;
}

int erp_get_scatter_kind(erp_ctx *ctx, int32_t *kind) {
    if (!ctx || !kind) return fail(ERP_ERR_INVALID, "null argument");
    *kind = ctx->owner_scatter ? ERP_SCATTER_OWNER : ERP_SCATTER_CELL_MAJOR;
    return ERP_OK;
}

int erp_get_setup(erp_ctx *ctx, float *cw, float *ch, float *ox, float *oy, float *ratio,
                  double *scalars) {
    if (!ctx) return fail(ERP_ERR_INVALID, "null context");
    DeviceGuard g(ctx->device);
    if (g.rc) return g.rc;
    ERP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::vector<float2> a((size_t)ctx->N), b((size_t)ctx->N);
    ERP_HIP_TRY(hipMemcpy(a.data(), ctx->csize.p, sizeof(float2) * a.size(), hipMemcpyDeviceToHost));
    ERP_HIP_TRY(hipMemcpy(b.data(), ctx->coff.p, sizeof(float2) * b.size(), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < ctx->N; ++i) {
        if (cw) cw[i] = a[i].x;
        if (ch) ch[i] = a[i].y;
        if (ox) ox[i] = b[i].x;
        if (oy) oy[i] = b[i].y;
    }
    if (ratio) ERP_HIP_TRY(hipMemcpy(ratio, ctx->ratio.p, sizeof(float) * (size_t)ctx->N, hipMemcpyDeviceToHost));
    if (scalars) {
        scalars[0] = ctx->bw;
        scalars[1] = ctx->bh;
        scalars[2] = ctx->t;
        scalars[3] = ctx->fx_scale;
    }
    return ERP_OK;
}

int erp_final_hpwl_dbu(erp_ctx *ctx, int64_t *hpwl_dbu) {
    if (!ctx || !hpwl_dbu) return fail(ERP_ERR_INVALID, "null argument");
    DeviceGuard g(ctx->device);
    if (g.rc) return g.rc;
    unsigned long long zero = 0, v = 0;
    ERP_HIP_TRY(hipMemcpyAsync(ctx->hpwl_dbu.p, &zero, sizeof(zero), hipMemcpyHostToDevice, ctx->stream));
    int rc = launch_hpwl_dbu(ctx->job.wa, ctx->E, ctx->hpwl_dbu.as<unsigned long long>(), ctx->stream);
    if (rc) return rc;
    ERP_HIP_TRY(hipMemcpyAsync(&v, ctx->hpwl_dbu.p, sizeof(v), hipMemcpyDeviceToHost, ctx->stream));
    ERP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    *hpwl_dbu = (int64_t)v;
    return ERP_OK;
}

int erp_set_timing(erp_ctx *ctx, int enabled) {
    if (!ctx) return fail(ERP_ERR_INVALID, "null context");
    if (enabled && !ctx->timing && ctx->timer) ctx->timer->reset_totals();
    ctx->timing = enabled != 0;
    return ERP_OK;
}

int erp_get_timing(erp_ctx *ctx, double *ms, int32_t capacity, int32_t *count, const char **names) {
    if (!ctx || !count) return fail(ERP_ERR_INVALID, "null argument");
    return timer_kernel_names(ctx, ms, capacity, count, names);
}

int erp_ctx_device_bytes(erp_ctx *ctx, int64_t *bytes) {
    if (!ctx || !bytes) return fail(ERP_ERR_INVALID, "null argument");
    *bytes = (int64_t)ctx_device_bytes(ctx);
    return ERP_OK;
}

int erp_estimate_device_bytes(const erp_design *d, const erp_config *cfg, int64_t *bytes) {
    if (!d || !cfg || !bytes) return fail(ERP_ERR_INVALID, "null argument");
    int rc = validate(d, cfg);
    if (rc) return rc;
    *bytes = (int64_t)estimate_ctx_bytes(d, cfg);
    return ERP_OK;
}

// ------------------------------------------------------------------ leaf kernels
namespace {
struct LeafScope {
    hipStream_t s = nullptr;
    int rc = ERP_OK;
    LeafScope() {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) { rc = fail(ERP_ERR_NO_DEVICE, "no HIP device"); return; }
        if (hipStreamCreate(&s) != hipSuccess) rc = fail(ERP_ERR_HIP, "hipStreamCreate failed");
    }
    ~LeafScope() { if (s) { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); } }
};

int leaf_spectral(const float *x, float *out, int32_t m, int32_t n, int32_t mode, int which) {
    if (!x || !out) return fail(ERP_ERR_INVALID, "null argument");
    if (!check_grid(m) || !check_grid(n))
        return fail(ERP_ERR_UNSUPPORTED, "sizes must be powers of two in [64, 2048]");
    if (mode != ERP_SPECTRAL_FFT && mode != ERP_SPECTRAL_MFMA_BF16X3)
        return fail(ERP_ERR_UNSUPPORTED, "unknown spectral mode");
    if (mode == ERP_SPECTRAL_MFMA_BF16X3 && !mfma_supported(m, n))
        return fail(ERP_ERR_UNSUPPORTED, "the MFMA spectral form needs at least 128 bins per axis");
    LeafScope L;
    if (L.rc) return L.rc;
    const size_t B = (size_t)m * n;
    if (mode == ERP_SPECTRAL_MFMA_BF16X3) {
        DevBuf dx, dout;
        MfmaPlan *plan = nullptr;
        int rc;
        if ((rc = upload(dx, x, B, L.s)) || (rc = dout.alloc(sizeof(float) * B)) ||
            (rc = mfma_plan_create(&plan, m, n, 1.0f, L.s)))
            return rc;
        // which: 0 dct2, 1 idct2, 2 idct_idxst, 3 idxst_idct
        rc = which == 0 ? launch_dct2_mfma(dx.as<float>(), dout.as<float>(), plan, L.s)
                        : launch_inverse_mfma(dx.as<float>(), dout.as<float>(), which - 1, plan, L.s);
        if (!rc && hipMemcpyAsync(out, dout.p, sizeof(float) * B, hipMemcpyDeviceToHost, L.s) != hipSuccess)
            rc = fail(ERP_ERR_HIP, "hipMemcpyAsync failed");
        if (!rc && hipStreamSynchronize(L.s) != hipSuccess) rc = fail(ERP_ERR_HIP, "hipStreamSynchronize failed");
        (void)hipStreamSynchronize(L.s);
        mfma_plan_destroy(plan);
        return rc;
    }
    DevBuf dx, dout, dscr;
    SpecTables tm, tn;
    int rc;
    if ((rc = upload(dx, x, B, L.s)) || (rc = dout.alloc(sizeof(float) * B)) ||
        (rc = dscr.alloc(sizeof(float) * B)) || (rc = make_spec_tables(tm, m, L.s)) ||
        (rc = make_spec_tables(tn, n, L.s)))
        return rc;
    const float *px = dx.as<float>();
    float *po = dout.as<float>(), *ps = dscr.as<float>();
    if ((rc = launch_spectral_leaf(px, po, ps, m, n, which, spec_view(tm), spec_view(tn), L.s))) return rc;
    ERP_HIP_TRY(hipMemcpyAsync(out, po, sizeof(float) * B, hipMemcpyDeviceToHost, L.s));
    ERP_HIP_TRY(hipStreamSynchronize(L.s));
    return ERP_OK;
}

void pack2(std::vector<float2> &v, const float *a, const float *b, int64_t n) {
    v.resize((size_t)n);
    for (int64_t i = 0; i < n; ++i) v[i] = make_float2(a[i], b[i]);
}
}  // namespace

int erp_dct2_fft2_forward(const float *x, float *out, int32_t m, int32_t n, int32_t mode) {
    return leaf_spectral(x, out, m, n, mode, 0);
}
int erp_idct2_fft2_forward(const float *x, float *out, int32_t m, int32_t n, int32_t mode) {
    return leaf_spectral(x, out, m, n, mode, 1);
}
int erp_idct_idxst_forward(const float *x, float *out, int32_t m, int32_t n, int32_t mode) {
    return leaf_spectral(x, out, m, n, mode, 2);
}
int erp_idxst_idct_forward(const float *x, float *out, int32_t m, int32_t n, int32_t mode) {
    return leaf_spectral(x, out, m, n, mode, 3);
}

int erp_compute_density_map(int64_t nc, const float *x, const float *y, const float *cw,
                            const float *ch, const float *ox, const float *oy, const float *ratio,
                            int32_t m, int32_t n, float bw, float bh, double scale,
                            const float *initial, float *map, int64_t *q_out) {
    if (nc < 0 || (nc > 0 && (!x || !y || !cw || !ch || !ox || !oy || !ratio)) || !map)
        return fail(ERP_ERR_INVALID, "null argument");
    if (m < 2 || n < 2 || !(bw > 0) || !(bh > 0) || !(scale > 0))
        return fail(ERP_ERR_INVALID, "bad grid, bin size or scale");
    LeafScope L;
    if (L.rc) return L.rc;
    const size_t B = (size_t)m * n;
    std::vector<float2> p, sz, of;
    pack2(p, x, y, nc);
    pack2(sz, cw, ch, nc);
    pack2(of, ox, oy, nc);
    std::vector<float4> aux((size_t)nc);            // only the ratio is read (no order copy)
    for (int64_t i = 0; i < nc; ++i) aux[(size_t)i] = make_float4(0.f, 0.f, ratio[i], 0.f);
    DevBuf dp, ds, dof, dr, dq, dinit, dmap, drec;
    int rc;
    if ((rc = upload(dp, p.data(), p.size(), L.s)) || (rc = upload(ds, sz.data(), sz.size(), L.s)) ||
        (rc = upload(dof, of.data(), of.size(), L.s)) || (rc = upload(dr, aux.data(), aux.size(), L.s)) ||
        (rc = dq.alloc(sizeof(int64_t) * B)) || (rc = dmap.alloc(sizeof(float) * B)) ||
        (rc = drec.alloc(sizeof(float4) * (size_t)std::max<int64_t>(nc, 1))))
        return rc;
    if (initial && (rc = upload(dinit, initial, B, L.s))) return rc;
    if ((rc = launch_zero_i64(dq.as<int64_t>(), (int64_t)B, L.s))) return rc;
    if ((rc = launch_pack_records(dp.as<float2>(), ds.as<float2>(), dof.as<float2>(), nc,
                                  drec.as<float4>(), L.s)))
        return rc;
    JobDesc j{};
    j.sb.num_cells = nc;
    j.sb.rec = drec.as<float4>();
    j.sb.aux = dr.as<float4>();
    j.sb.m = m;
    j.sb.n = n;
    j.sb.bw = bw;
    j.sb.bh = bh;
    j.pb.q_mov = dq.as<int64_t>();
    j.pb.scale = scale;
    j.pb.m = m;
    j.pb.n = n;
    j.cb.C = nc;
    DevBuf dj;
    if ((rc = upload(dj, &j, 1, L.s))) return rc;
    BatchDims bd;
    bd.add(j);
    if ((rc = launch_scatter(dj.as<JobDesc>(), bd, 1, false, L.s))) return rc;
    if ((rc = launch_finalize_leaf(dq.as<int64_t>(), initial ? dinit.as<float>() : nullptr,
                                   dmap.as<float>(), (int64_t)B, 1.0 / scale, L.s)))
        return rc;
    ERP_HIP_TRY(hipMemcpyAsync(map, dmap.p, sizeof(float) * B, hipMemcpyDeviceToHost, L.s));
    if (q_out) ERP_HIP_TRY(hipMemcpyAsync(q_out, dq.p, sizeof(int64_t) * B, hipMemcpyDeviceToHost, L.s));
    ERP_HIP_TRY(hipStreamSynchronize(L.s));
    return ERP_OK;
}

int erp_compute_initial_density_map(int64_t nf, const float *x, const float *y, const float *cw,
                                    const float *ch, int32_t m, int32_t n, float bw, float bh,
                                    float die_w, float die_h, double scale, float *map,
                                    int64_t *q_out) {
    if (nf < 0 || (nf > 0 && (!x || !y || !cw || !ch)) || !map)
        return fail(ERP_ERR_INVALID, "null argument");
    if (m < 2 || n < 2 || !(bw > 0) || !(bh > 0) || !(scale > 0))
        return fail(ERP_ERR_INVALID, "bad grid, bin size or scale");
    for (int64_t i = 0; i < nf; ++i)
        if (!(x[i] >= 0.f) || !(y[i] >= 0.f)) return fail(ERP_ERR_INVALID, "negative fixed-cell coordinate");
    LeafScope L;
    if (L.rc) return L.rc;
    const size_t B = (size_t)m * n;
    std::vector<float2> p, sz;
    pack2(p, x, y, nf);
    pack2(sz, cw, ch, nf);
    DevBuf dp, ds, dq, dmap;
    int rc;
    if ((rc = upload(dp, p.data(), p.size(), L.s)) || (rc = upload(ds, sz.data(), sz.size(), L.s)) ||
        (rc = dq.alloc(sizeof(int64_t) * B)) || (rc = dmap.alloc(sizeof(float) * B)))
        return rc;
    if ((rc = launch_zero_i64(dq.as<int64_t>(), (int64_t)B, L.s))) return rc;
    if ((rc = launch_fixed_map(dp.as<float2>(), ds.as<float2>(), nf, m, n, bw, bh, die_w, die_h,
                               scale, dq.as<int64_t>(), L.s)))
        return rc;
    if ((rc = launch_finalize_leaf(dq.as<int64_t>(), nullptr, dmap.as<float>(), (int64_t)B,
                                   1.0 / scale, L.s)))
        return rc;
    ERP_HIP_TRY(hipMemcpyAsync(map, dmap.p, sizeof(float) * B, hipMemcpyDeviceToHost, L.s));
    if (q_out) ERP_HIP_TRY(hipMemcpyAsync(q_out, dq.p, sizeof(int64_t) * B, hipMemcpyDeviceToHost, L.s));
    ERP_HIP_TRY(hipStreamSynchronize(L.s));
    return ERP_OK;
}

int erp_compute_electric_force(int64_t nc, int32_t m, int32_t n, const float *field_x,
                               const float *field_y, const float *x, const float *y,
                               const float *cw, const float *ch, const float *ox, const float *oy,
                               const float *ratio, float bw, float bh, float *grad_x,
                               float *grad_y) {
    if (nc < 0 || !field_x || !field_y || (nc > 0 && (!x || !y || !cw || !ch || !ox || !oy ||
                                                      !ratio || !grad_x || !grad_y)))
        return fail(ERP_ERR_INVALID, "null argument");
    if (m < 1 || n < 1 || !(bw > 0) || !(bh > 0)) return fail(ERP_ERR_INVALID, "bad grid");
    if (nc == 0) return ERP_OK;
    LeafScope L;
    if (L.rc) return L.rc;
    std::vector<float2> p, sz, of;
    pack2(p, x, y, nc);
    pack2(sz, cw, ch, nc);
    pack2(of, ox, oy, nc);
    std::vector<float2> fxy(fxy_bins(m, n), make_float2(0.f, 0.f));   // 4 x 4-bin blocks
    for (int k = 0; k < m; ++k)
        for (int h = 0; h < n; ++h)
            fxy[fxy_index(k, h, n)] = make_float2(field_x[(size_t)k * n + h], field_y[(size_t)k * n + h]);
    std::vector<float4> geo((size_t)nc), prop((size_t)nc);   // the cell kernels' packed constants
    for (int64_t i = 0; i < nc; ++i) {
        geo[(size_t)i] = make_float4(ox[i], oy[i], cw[i], ch[i]);
        prop[(size_t)i] = make_float4(ratio[i], 0.f, 0.f, 0.f);
    }
    DevBuf dp, ds, dof, dr, dfxy, dg, dgeo, dprop;
    int rc;
    if ((rc = upload(dp, p.data(), p.size(), L.s)) || (rc = upload(ds, sz.data(), sz.size(), L.s)) ||
        (rc = upload(dof, of.data(), of.size(), L.s)) || (rc = upload(dr, ratio, (size_t)nc, L.s)) ||
        (rc = upload(dgeo, geo.data(), geo.size(), L.s)) || (rc = upload(dprop, prop.data(), prop.size(), L.s)) ||
        (rc = upload(dfxy, fxy.data(), fxy.size(), L.s)) || (rc = dg.alloc(sizeof(float2) * (size_t)nc)))
        return rc;
    CellBufs c{};
    c.C = nc;
    c.pos = dp.as<float2>();
    c.csize = ds.as<float2>();
    c.coff = dof.as<float2>();
    c.ratio = dr.as<float>();
    c.cgeo = dgeo.as<float4>();
    c.cprop = dprop.as<float4>();
    c.gd = dg.as<float2>();
    c.fxy = dfxy.as<float2>();
    c.m = m;
    c.n = n;
    c.bw = bw;
    c.bh = bh;
    JobDesc j{};
    j.cb = c;
    j.pb.m = m;
    j.pb.n = n;
    DevBuf dj;
    if ((rc = upload(dj, &j, 1, L.s))) return rc;
    BatchDims bd;
    bd.add(j);
    if ((rc = launch_efield(dj.as<JobDesc>(), bd, 1, L.s))) return rc;
    std::vector<float2> gv((size_t)nc);
    ERP_HIP_TRY(hipMemcpyAsync(gv.data(), dg.p, sizeof(float2) * gv.size(), hipMemcpyDeviceToHost, L.s));
    ERP_HIP_TRY(hipStreamSynchronize(L.s));
    for (int64_t i = 0; i < nc; ++i) { grad_x[i] = gv[i].x; grad_y[i] = gv[i].y; }
    return ERP_OK;
}

}  // extern "C"
