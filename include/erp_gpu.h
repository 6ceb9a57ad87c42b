/*
 * erp_gpu.h -- C ABI of the MI355X (gfx950) electrostatic global-placement
 * iteration: density scatter, spectral Poisson solve, field gather, weighted-
 * average wirelength gradient, preconditioned Nesterov step and scheduler.
 *
 * The boundary replaces the crate-internal Rust surface that ERPlacer's
 * Placement::step calls (giammirove/reda-erplacer, src/placement.rs:180-228).
 * Each entry point names the reference item it stands in for.
 *
 * Conventions
 *   - plain C: pointers, sizes, scalars; no exceptions cross the ABI;
 *   - every function returns ERP_OK (0) or a negative ERP_ERR_* code, and
 *     erp_last_error() returns a thread-local message for the last failure
 *     (the reference panics instead; a Rust shim turns non-zero into panic!);
 *   - arrays passed in are host memory and are copied; a context owns all of its
 *     device memory and one HIP stream; calls on one context are serialised by
 *     the caller, distinct contexts are independent;
 *   - the iteration computes in FP32 like the reference binary (main.rs:69 reads
 *     f32); the density map is accumulated in 64-bit fixed point (deterministic),
 *     step-size dot products, overflow, energy and the per-iteration HPWL sum are
 *     reduced in FP64, and erp_final_hpwl_dbu is exact in int64.
 */
#ifndef ERP_GPU_H
#define ERP_GPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ERP_OK = 0,
    ERP_ERR_INVALID = -1,     /* bad argument or inconsistent design */
    ERP_ERR_HIP = -2,         /* HIP runtime failure */
    ERP_ERR_NO_DEVICE = -3,   /* no usable GPU */
    ERP_ERR_UNSUPPORTED = -4  /* size or mode not supported */
};

/* Spectral-solve implementations. */
enum {
    ERP_SPECTRAL_FFT = 0,        /* FP32 FFT-based DCT/IDCT/IDXST in LDS */
    ERP_SPECTRAL_MFMA_BF16X3 = 1 /* dense transform products on MFMA, split-BF16 */
};

/* Maps for erp_get_map (row-major [bx * num_bins_y + by], x-bin major). */
enum {
    ERP_MAP_DENSITY = 0,   /* density_map after the 1/bin_area scaling (epotential.rs:284-285) */
    ERP_MAP_AUV = 1,       /* DCT coefficients a_uv (epotential.rs:287-297); kept only
                              while gradients are stored (erp_set_store_gradients) */
    ERP_MAP_FIELD_X = 2,   /* field_map_x (epotential.rs:310-319) */
    ERP_MAP_FIELD_Y = 3,   /* field_map_y (epotential.rs:320-329) */
    ERP_MAP_POTENTIAL = 4, /* potential_map (epotential.rs:333-343) */
    ERP_MAP_INITIAL = 5    /* fixed-cell map x target density (epotential.rs:127-140) */
};

/* Per-cell gradients for erp_get_cell_grads. */
enum { ERP_GRAD_WIRELENGTH = 0, ERP_GRAD_DENSITY = 1, ERP_GRAD_COMBINED = 2 };

/* The reda-db data contract the loop reads (main.rs:68-70, placement.rs:52-134,
 * hpwl.rs:25-45): instances with movable ones first, nets with the n_2nets
 * two-pin nets first, CSR instead of Vec<Vec<usize>>. */
typedef struct erp_design {
    int64_t num_nodes, num_movable;        /* instances.len(), num_movable */
    const float *x, *y;                    /* instances.coords (DB, lower-left) */
    const float *w, *h;                    /* instances.sizes */
    const float *num_pins;                 /* instances.num_pins (as T) */
    int64_t num_pins_total;                /* netlist.pins.len() */
    const float *pin_offset_x, *pin_offset_y; /* netlist.pins.x/y */
    const int32_t *pin_node;               /* owning instance of each pin */
    int64_t num_nets, n_2nets;             /* netlist.nets.len(), n_2nets */
    const int32_t *net_start, *net_pins;   /* nets[i].pin_ids as CSR */
    const int32_t *node_pin_start, *node_pins; /* macro_2_pins as CSR */
    double die_w, die_h;                   /* diearea.width()/height(), origin 0 */
    double movable_area, cell_utilization; /* db.movable_area, db.cell_utilization */
} erp_design;

typedef struct erp_config {
    int32_t num_bins_x, num_bins_y; /* placement.rs:50 hard-codes 512x512; powers of two 64..2048 */
    double target_density;          /* NaN: db.cell_utilization (placement.rs:134) */
    double gamma_scale;             /* gamma0 = gamma_scale * (bw + bh), scheduler.rs:56 (4) */
    double density_weight0;         /* scheduler.rs:57 (8e-5) */
    int32_t spectral_mode;          /* ERP_SPECTRAL_* */
    int32_t device;                 /* HIP device ordinal */
} erp_config;

/* Scalars of one Placement::step: what the step returned and the scheduler /
 * Nesterov state after it. */
typedef struct erp_step_info {
    double hpwl, overflow, energy;
    double density_weight, gamma;
    double alpha, beta;
    int64_t iteration;
    int32_t should_stop;            /* Scheduler::should_stop (scheduler.rs:75-81) */
    int32_t nonfinite;              /* 1 if any reduced scalar was NaN/Inf */
} erp_step_info;

typedef struct erp_ctx erp_ctx;

const char *erp_last_error(void);
int erp_device_count(int *count);
void erp_config_default(erp_config *cfg);

/* ---------------------------------------------------------------- context
 * Placement::new (placement.rs:159-178): Bins::new, PlacementSolution::new
 * (movable cells at the die centre), PotentialComputation::new (clamped sizes,
 * fixed-cell map, spectral tables), HpwlComputation::new, Scheduler::new,
 * Nesterov::new; everything is uploaded and precomputed on the device. */
int erp_ctx_create(const erp_design *design, const erp_config *cfg, erp_ctx **out);
int erp_ctx_destroy(erp_ctx *ctx);
/* Back to the state right after erp_ctx_create (positions, optimizer, scheduler). */
int erp_ctx_reset(erp_ctx *ctx);
/* PlacementSolution.instances (all nodes; fixed nodes keep their DB coordinates). */
int erp_set_positions(erp_ctx *ctx, const float *x, const float *y);
int erp_get_positions(erp_ctx *ctx, float *x, float *y);

/* compute_hpwl (hpwl.rs:139-343) with the scheduler's gamma: WA gradient into
 * the context's wirelength gradient, returns sum of per-net HPWL. */
int erp_compute_hpwl(erp_ctx *ctx, double *hpwl);
/* Same with an explicit gamma (the reference's `gamma` argument). */
int erp_compute_hpwl_gamma(erp_ctx *ctx, double gamma, double *hpwl);
/* compute_potential (epotential.rs:225-374): density map, overflow, DCT,
 * field and potential maps, energy, density gradient. */
int erp_compute_potential(erp_ctx *ctx, double *energy, double *overflow);
/* Nesterov::step_adaptive (nesterov.rs:27-146) incl. the preconditioner
 * (nesterov.rs:75-82; preconditioner.rs is empty). */
int erp_nesterov_step(erp_ctx *ctx);
/* Scheduler::update (scheduler.rs:68-73) with the last hpwl/overflow. */
int erp_scheduler_update(erp_ctx *ctx);
/* Placement::step (placement.rs:180-228): the four calls above in order. */
int erp_step(erp_ctx *ctx, erp_step_info *info);
/* `iterations` calls of Placement::step (main.rs:78-92), replayed from one
 * captured hipGraph; info (nullable) receives the last step's scalars. */
int erp_run(erp_ctx *ctx, int64_t iterations, erp_step_info *info);
/* The same loop that additionally stops once overflow < stop_overflow
 * (checked every check_every iterations). */
int erp_run_until(erp_ctx *ctx, int64_t max_iterations, double stop_overflow,
                  int32_t check_every, erp_step_info *info);
/* Scalars of the last completed step. */
int erp_get_step_info(erp_ctx *ctx, erp_step_info *info);

int erp_get_map(erp_ctx *ctx, int which, float *out);
/* Movable-cell gradients (length num_movable). */
int erp_get_cell_grads(erp_ctx *ctx, int which, float *gx, float *gy);
/* The fixed-point density accumulators of the last compute_potential:
 * movable part, fixed part (x target density), and the scale S (value = q / S). */
int erp_get_density_fixed(erp_ctx *ctx, int64_t *q_mov, int64_t *q_fix, double *scale);
/* Keep the movable accumulators after each compute_potential so that
 * erp_get_density_fixed can return them (off by default: the loop clears them
 * in the same pass that reads them). */
int erp_set_keep_density(erp_ctx *ctx, int keep);
/* Keep the per-cell wirelength and density gradients and the a_uv map in the
 * fused loop (on by default; off saves 16 bytes per cell and a strided write of
 * the whole a_uv map per iteration). */
int erp_set_store_gradients(erp_ctx *ctx, int store);
/* Compute the potential map and the energy every step (on by default, as
 * compute_potential does, epotential.rs:333-346).  Placement::step discards
 * both (placement.rs:185), so off ("trajectory only") skips one of the three
 * inverse transforms with an unchanged trajectory; energy then reads 0 and the
 * potential map is not refreshed. */
int erp_set_compute_potential(erp_ctx *ctx, int on);
/* Movable-scatter kernel of the loop (compute_density_map / _fallback,
 * density.rs:240-299; both give the same int64 map bit for bit).  AUTO picks the
 * owner scatter (one workgroup per tile neighbourhood, no atomics on exclusive
 * tiles) for sparse grids whose cells reach at most two tiles past their own,
 * the cell-major scatter (LDS tiles over sorted cells, global atomics) otherwise.
 * Forcing OWNER where a cell reaches further is ERP_ERR_UNSUPPORTED. */
enum { ERP_SCATTER_AUTO = -1, ERP_SCATTER_CELL_MAJOR = 0, ERP_SCATTER_OWNER = 1 };
int erp_set_scatter_kind(erp_ctx *ctx, int kind);
int erp_get_scatter_kind(erp_ctx *ctx, int32_t *kind);
/* Clamped sizes/offsets/ratios of PotentialComputation::new (epotential.rs:83-103),
 * length num_nodes; scalars = {bin_w, bin_h, target_density, fx_scale}. */
int erp_get_setup(erp_ctx *ctx, float *cw, float *ch, float *ox, float *oy, float *ratio,
                  double *scalars);
/* Exact HPWL of the current positions in int64 database units: pin coordinates
 * rounded to nearest-even integers, sum over nets of (max-min)x + (max-min)y. */
int erp_final_hpwl_dbu(erp_ctx *ctx, int64_t *hpwl_dbu);
/* Per-launch device time (ms, HIP events on the context's stream) of the
 * kernels of the last erp_step/erp_run call when timing is enabled. */
int erp_set_timing(erp_ctx *ctx, int enabled);
int erp_get_timing(erp_ctx *ctx, double *ms_per_kernel, int32_t capacity, int32_t *count,
                   const char **names);
/* Device memory a context holds (erp_ctx_device_bytes) and the upper estimate
 * used for admission before one exists (erp_estimate_device_bytes). */
int erp_ctx_device_bytes(erp_ctx *ctx, int64_t *bytes);
int erp_estimate_device_bytes(const erp_design *design, const erp_config *config, int64_t *bytes);

/* ----------------------------------------------------------------- serving
 * No counterpart in the reference (main.rs places one design per process).
 * A persistent loop that runs many independent placements on one GPU with
 * continuous batching: submitted jobs are admitted in arrival order while
 * job slots and the device-memory budget allow, keep their state resident,
 * and jobs with the same bin grid advance together in one captured hipGraph
 * of `check_every` batched iterations.  After each graph every job's overflow
 * is read; a job stops like erp_run_until(max_iterations, stop_overflow,
 * check_every) and its result is bit-identical to that call on the design
 * alone (FFT spectral form). */
typedef struct erp_server erp_server;

typedef struct {
    int64_t hbm_budget_bytes;   /* sum of admitted jobs' estimated device memory */
    int32_t max_active;         /* jobs resident at once */
    int32_t check_every;        /* iterations between convergence checks */
    int32_t device;
    /* 0 (default): each batch's check_every iterations are one captured hipGraph;
     * 1: the same batched kernels are launched directly each round (no capture;
     * identical results, for kernel tracers) */
    int32_t eager;
} erp_server_config;

typedef struct {
    int64_t job_id;
    int32_t status;             /* ERP_OK or the error that ended the job */
    int32_t converged;          /* overflow < stop_overflow */
    int64_t iterations;
    double hpwl, overflow;      /* scalars of the last step */
    int64_t hpwl_dbu;           /* exact final HPWL (erp_final_hpwl_dbu) */
    double t_submit, t_admit, t_done;  /* seconds on the server's clock */
} erp_job_result;

typedef struct {
    int64_t pending, active, finished, bytes_in_use;
    int64_t rounds, batched_launches, max_batch;
    double clock;               /* seconds since erp_server_create */
    /* worker time split: admitting (context creation + first chunk), batch
     * rounds (graph launches, capture, state reads), retiring (copy-out, free) */
    double admit_seconds, round_seconds, retire_seconds;
} erp_server_stat;

void erp_server_config_default(erp_server_config *config);
int erp_server_create(const erp_server_config *config, erp_server **out);
/* Stops the worker after the running round; queued jobs are dropped. */
int erp_server_destroy(erp_server *server);
/* Copies the design; returns at once with the job id. */
int erp_server_submit(erp_server *server, const erp_design *design, const erp_config *config,
                      int64_t max_iterations, double stop_overflow, int64_t *job_id);
/* Results of jobs finished since the last poll (up to capacity). */
int erp_server_poll(erp_server *server, erp_job_result *out, int32_t capacity, int32_t *count);
/* Blocks until the job has finished. */
int erp_server_wait(erp_server *server, int64_t job_id, erp_job_result *out);
/* Blocks until no job is queued or running. */
int erp_server_drain(erp_server *server);
/* Final positions of a finished job (num_nodes each). */
int erp_server_get_positions(erp_server *server, int64_t job_id, float *x, float *y);
/* Forget a finished job's result and positions. */
int erp_server_release(erp_server *server, int64_t job_id);
int erp_server_stats(erp_server *server, erp_server_stat *stats);

/* --------------------------------------------------- leaf kernels (host buffers)
 * Stand-alone forms of the reference's leaf functions, for per-kernel parity.
 * m = rows (x bins), n = columns (y bins), powers of two in [64, 2048]. */
/* dct2_fft2_forward (dct2d.rs:256-279): out = dctn(x, type 2) / (m n) */
int erp_dct2_fft2_forward(const float *x, float *out, int32_t m, int32_t n, int32_t mode);
/* idct2_fft2_forward (idct2.rs:144-166): out = dctn(x, type 3) */
int erp_idct2_fft2_forward(const float *x, float *out, int32_t m, int32_t n, int32_t mode);
/* idct_idxst_forward (idct_idxst.rs:137-161): cosine along x, sine along y */
int erp_idct_idxst_forward(const float *x, float *out, int32_t m, int32_t n, int32_t mode);
/* idxst_idct_forward (idxst_idct.rs:156-180): sine along x, cosine along y */
int erp_idxst_idct_forward(const float *x, float *out, int32_t m, int32_t n, int32_t mode);
/* Allocation-free form of the four transforms above, on device buffers.  It has
 * no counterpart in the reference; it is for callers that cannot allocate
 * device memory: code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  x and out are device float[m*n] of the current
 * device (not aliased), which: 0 dct2_fft2_forward, 1 idct2_fft2_forward,
 * 2 idct_idxst_forward, 3 idxst_idct_forward (FFT form, the same kernels and
 * results as the host-buffer entry points), stream a hipStream_t of the current
 * device (NULL: the default stream).  The call only enqueues kernels on that
 * stream (out is valid once the stream reaches them) and makes no allocation,
 * stream or event: its scratch is one fixed array per GPU of 16 MiB + 48 KiB
 * (one 2048 x 2048 map and the tables of two lengths), so m and n are powers
 * of two in [64, 2048] and larger sizes are ERP_ERR_UNSUPPORTED.  Calls that
 * overlap on one GPU share that array and race; serialise them (one stream, or
 * events).  The first call on a device looks the array's address up; make it
 * outside graph capture. */
int erp_spectral_transform_noalloc(const float *x, float *out, int32_t m, int32_t n,
                                   int32_t which, void *stream);
/* compute_density_map / _fallback (density.rs:240-299): movable scatter of
 * clamped cells (x, y are cell origins; ox, oy the clamp offsets) accumulated in
 * fixed point at `scale`, then + initial.  q_out (nullable) receives the
 * movable int64 accumulators. */
int erp_compute_density_map(int64_t num_cells, const float *x, const float *y,
                            const float *cw, const float *ch, const float *ox,
                            const float *oy, const float *ratio, int32_t m, int32_t n,
                            float bin_w, float bin_h, double scale, const float *initial,
                            float *map, int64_t *q_out);
/* compute_initial_density_map (density.rs:301-325): fixed cells against bin
 * centres, accumulated in fixed point at `scale` (q_out: int64 accumulators). */
int erp_compute_initial_density_map(int64_t num_fixed, const float *x, const float *y,
                                    const float *cw, const float *ch, int32_t m, int32_t n,
                                    float bin_w, float bin_h, float die_w, float die_h,
                                    double scale, float *map, int64_t *q_out);
/* compute_electric_force (eforce.rs:15-91) */
int erp_compute_electric_force(int64_t num_cells, int32_t m, int32_t n, const float *field_x,
                               const float *field_y, const float *x, const float *y,
                               const float *cw, const float *ch, const float *ox,
                               const float *oy, const float *ratio, float bin_w, float bin_h,
                               float *grad_x, float *grad_y);

#ifdef __cplusplus
}
#endif
#endif
