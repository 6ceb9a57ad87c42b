/*
 * erp_oracle.h -- CPU oracle for ERPlacer's global-placement iteration.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/, the smoke
 * check and the CPU-baseline leg of bench.py load it.  The product (the HIP
 * library under reda-erplacer_amd/) never links, loads or calls it.
 *
 * It restates, function by function, the reference's Rust path
 * (giammirove/reda-erplacer, src/ directory); each definition in erp_oracle_impl.inc
 * cites the file:line it follows.  Every function exists three times:
 *   *_f64  -- the FP64 parity oracle (run with nthreads=1: deterministic);
 *   *_f32  -- the reference binary's own precision (main.rs:69 reads f32), used
 *             as the timed CPU baseline with nthreads = host cores, parallelised
 *             with OpenMP the way the reference parallelises with Rayon;
 *   *_f32d -- f32 arrays and per-element arithmetic with FP64 scalars and sums
 *             (step size, momentum, weight, gamma; HPWL, overflow, energy and
 *             dot-product reductions): the precision split the GPU path uses.
 *
 * The *_fx64 functions are not in the reference: they restate, on the CPU, the
 * deterministic 64-bit fixed-point density accumulation the GPU path uses in
 * place of the reference's per-bin Mutex float sums (density.rs:158), so the GPU
 * map can be checked bit for bit.
 *
 * Pinning (see tests/test_oracle_*.py): make_expk / frequency matrices by closed
 * form; the four spectral transforms against scipy.fft.dctn type 2/3 and dense
 * sin/cos products; irfft_2d against numpy.fft.irfft2; the FFT against
 * numpy.fft; density scatter/gather by brute force and the adjoint identity;
 * WA gradient by finite differences of the WA value and by translation
 * invariance; Nesterov and the scheduler by hand-computed recurrences.
 * Parity unpinned: bit-level agreement with rustfft's f32 rounding and with the
 * reference's thread-order-dependent float sums (no fixtures ship with the
 * reference and it cannot be built here).
 */
#ifndef ERP_ORACLE_H
#define ERP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ERO_PI_F32 = 0, ERO_PI_EXACT = 1 };

/* reda-db data contract, CSR form (see the generator for field meanings). */
typedef struct {
    int64_t num_nodes, num_movable;
    const float *x, *y, *w, *h, *num_pins;
    int64_t num_pins_total;
    const float *pin_offset_x, *pin_offset_y;
    const int32_t *pin_node;
    int64_t num_nets, n_2nets;
    const int32_t *net_start, *net_pins;
    const int32_t *node_pin_start, *node_pins;
    double die_w, die_h;
    double movable_area, cell_utilization;
} ero_design;

typedef struct {
    int64_t num_bins_x, num_bins_y;
    double target_density;   /* NaN: use cell_utilization (placement.rs:134) */
    int pi_mode;             /* ERO_PI_F32 reproduces utils.rs:3 */
    int nthreads;
    int fixed_point;         /* 1: density map accumulated as in the GPU path */
} ero_config;

typedef struct {
    double hpwl, overflow, energy;
    double density_weight, gamma;     /* scheduler state after the step */
    double alpha, beta;               /* Nesterov state after the step */
    int64_t iteration;                /* scheduler iteration after the step */
    int should_stop;
} ero_step_out;

#define ERO_DECL(SFX, REAL)                                                        \
    typedef struct ero_state##SFX ero_state##SFX;                                 \
    void ero_make_expk##SFX(int64_t n, int pi_mode, REAL *re, REAL *im);          \
    void ero_frequency_matrices##SFX(int64_t m, int64_t n, REAL bw, REAL bh,      \
                                     REAL *wu_half, REAL *wv_half, REAL *inv);   \
    void ero_fft##SFX(REAL *re, REAL *im, int64_t n, int inverse);                \
    void ero_initial_density_map##SFX(int64_t nf, const REAL *x, const REAL *y,   \
                                      const REAL *sx, const REAL *sy, int64_t m,  \
                                      int64_t n, REAL bw, REAL bh, REAL die_w,    \
                                      REAL die_h, int nthreads, REAL *map);       \
    void ero_density_map_sequential##SFX(int64_t nc, const REAL *x, const REAL *y,\
        const REAL *sx, const REAL *sy, const REAL *ox, const REAL *oy,           \
        const REAL *ratio, int64_t m, int64_t n, REAL bw, REAL bh,                \
        const REAL *initial, REAL *map);                                          \
    void ero_density_map_parallel##SFX(int64_t nc, const REAL *x, const REAL *y,  \
        const REAL *sx, const REAL *sy, const REAL *ox, const REAL *oy,           \
        const REAL *ratio, int64_t m, int64_t n, REAL bw, REAL bh,                \
        const REAL *initial, REAL *map, int nthreads);                            \
    void ero_density_map_fx64##SFX(int64_t nc, const REAL *x, const REAL *y,      \
        const REAL *sx, const REAL *sy, const REAL *ox, const REAL *oy,           \
        const REAL *ratio, int64_t m, int64_t n, REAL bw, REAL bh, double scale,  \
        int64_t *q);                                                              \
    void ero_initial_density_fx64##SFX(int64_t nf, const REAL *x, const REAL *y,  \
        const REAL *sx, const REAL *sy, int64_t m, int64_t n, REAL bw, REAL bh,   \
        REAL die_w, REAL die_h, double scale, int64_t *q);                        \
    void ero_dct2_fft2_forward##SFX(const REAL *x, REAL *out, int64_t m,          \
                                    int64_t n, int pi_mode, int nthreads);        \
    void ero_idct2_fft2_forward##SFX(const REAL *x, REAL *out, int64_t m,         \
                                     int64_t n, int pi_mode, int nthreads);       \
    void ero_idct_idxst_forward##SFX(const REAL *x, REAL *out, int64_t m,         \
                                     int64_t n, int pi_mode, int nthreads);       \
    void ero_idxst_idct_forward##SFX(const REAL *x, REAL *out, int64_t m,         \
                                     int64_t n, int pi_mode, int nthreads);       \
    void ero_irfft_2d##SFX(REAL *in_re, REAL *in_im, REAL *out, int64_t m,        \
                           int64_t n, int nthreads);                              \
    void ero_dct2d_preprocess##SFX(const REAL *x, REAL *y_re, int64_t m, int64_t n);\
    void ero_compute_electric_force##SFX(int64_t nc, int64_t m, int64_t n,        \
        const REAL *fx, const REAL *fy, const REAL *x, const REAL *y,             \
        const REAL *sx, const REAL *sy, const REAL *ox, const REAL *oy,           \
        const REAL *ratio, REAL bw, REAL bh, REAL *gx, REAL *gy, int nthreads);   \
    REAL ero_compute_hpwl##SFX(const ero_design *d, const REAL *x, const REAL *y, \
        REAL gamma, REAL *grad_x, REAL *grad_y, REAL *pin_grad_x,                 \
        REAL *pin_grad_y, int nthreads);                                          \
    REAL ero_wa_value##SFX(const ero_design *d, const REAL *x, const REAL *y,     \
                           REAL gamma);                                           \
    ero_state##SFX *ero_create##SFX(const ero_design *d, const ero_config *c);    \
    void ero_destroy##SFX(ero_state##SFX *s);                                     \
    void ero_step##SFX(ero_state##SFX *s, ero_step_out *out);                     \
    void ero_set_positions##SFX(ero_state##SFX *s, const REAL *x, const REAL *y); \
    void ero_get_positions##SFX(const ero_state##SFX *s, REAL *x, REAL *y);       \
    void ero_get_map##SFX(const ero_state##SFX *s, int which, REAL *out);         \
    void ero_get_cell_grads##SFX(const ero_state##SFX *s, int which, REAL *gx,    \
                                 REAL *gy);                                       \
    void ero_get_setup##SFX(const ero_state##SFX *s, REAL *cw, REAL *ch,          \
                            REAL *ox, REAL *oy, REAL *ratio, double *scalars);    \
    void ero_scheduler_sequence##SFX(int64_t n, const REAL *hpwl, const REAL *ovf,\
        REAL base_gamma, REAL dw0, REAL *out_dw, REAL *out_gamma, int32_t *out_stop);

ERO_DECL(_f64, double)
ERO_DECL(_f32, float)
ERO_DECL(_f32d, float)   /* f32 arrays, FP64 scalars and sums (the GPU's split) */

/* maps for ero_get_map */
enum { ERO_MAP_DENSITY = 0, ERO_MAP_AUV = 1, ERO_MAP_FIELD_X = 2, ERO_MAP_FIELD_Y = 3,
       ERO_MAP_POTENTIAL = 4, ERO_MAP_INITIAL = 5 };
/* grads for ero_get_cell_grads */
enum { ERO_GRAD_WL = 0, ERO_GRAD_DENSITY = 1 };

/* fixed-point helpers shared with the GPU semantics */
double ero_fx_scale(double mass_bound);
void ero_fx_finalize_f32(int64_t nbins, const int64_t *q_mov, const int64_t *q_fix,
                         double scale, double coef, float *rho, int64_t *ovf_q);

#ifdef __cplusplus
}
#endif
#endif
