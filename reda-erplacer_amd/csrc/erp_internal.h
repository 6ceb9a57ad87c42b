// Internal declarations shared by the HIP translation units of liberp_gpu.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/erp_gpu.h"

namespace erp {

struct KernelTimer;

// ------------------------------------------------------------------ errors
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define ERP_HIP_TRY(expr)                                                              \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess)                                                           \
            return ::erp::fail(ERP_ERR_HIP, std::string(#expr " failed: ") +           \
                                                hipGetErrorString(_e));                 \
    } while (0)

// ---------------------------------------------------------- device buffers
// A slab of device memory that buffers are carved from (256-byte aligned,
// never freed one by one).  While a slab is active on a thread (SlabScope),
// DevBuf::alloc takes memory from it and falls back to hipMalloc when it is
// full: a context is created with one hipMalloc and destroyed with one hipFree.
struct Slab {
    char *base = nullptr;
    size_t cap = 0, used = 0;
    void *take(size_t nbytes) {
        const size_t off = (used + 255) & ~size_t(255);
        if (!base || off > cap || nbytes > cap - off) return nullptr;
        used = off + nbytes;
        return base + off;
    }
};
inline thread_local Slab *tl_slab = nullptr;
struct SlabScope {
    Slab *prev;
    explicit SlabScope(Slab *s) : prev(tl_slab) { tl_slab = s; }
    ~SlabScope() { tl_slab = prev; }
    SlabScope(const SlabScope &) = delete;
    SlabScope &operator=(const SlabScope &) = delete;
};

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    bool owned = false;          // from hipMalloc (else carved from a slab)
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    int alloc(size_t nbytes) {
        release();
        if (nbytes == 0) nbytes = 16;
        if (tl_slab) {
            if (void *q = tl_slab->take(nbytes)) {
                p = q;
                bytes = nbytes;
                return ERP_OK;
            }
        }
        hipError_t e = hipMalloc(&p, nbytes);
        if (e != hipSuccess) {
            p = nullptr;
            return fail(ERP_ERR_HIP, std::string("hipMalloc(") + std::to_string(nbytes) +
                                         ") failed: " + hipGetErrorString(e));
        }
        bytes = nbytes;
        owned = true;
        return ERP_OK;
    }
    void release() {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        owned = false;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// --------------------------------------------------------------- constants
constexpr int kBlock = 256;          // threads per workgroup for streaming kernels
constexpr int kMinBins = 64, kMaxBins = 2048;

// Field map layout: blocks of 4 x 4 bins, one 128-byte line each, blocks
// row-major over the (padded) grid, both field components of a bin in one
// float2.  A cell's window of up to 4 x 4 bins touches 1-4 lines (3-6 in a
// row-major map).
__host__ __device__ inline size_t fxy_index(int k, int h, int n) {
    return ((size_t)(k >> 2) * (size_t)((n + 3) >> 2) + (size_t)(h >> 2)) * 16 +
           (size_t)(((k & 3) << 2) | (h & 3));
}
__host__ __device__ inline size_t fxy_bins(int m, int n) {
    return (size_t)((m + 3) >> 2) * (size_t)((n + 3) >> 2) * 16;
}

inline bool is_pow2(int64_t v) { return v > 0 && (v & (v - 1)) == 0; }
inline int ilog2(int64_t v) { int r = 0; while ((int64_t(1) << r) < v) ++r; return r; }

// Device-resident scalar state of one placement (Scheduler + Nesterov + the
// reductions of the current step).  FP64 throughout.
struct DevState {
    // Scheduler (scheduler.rs:36-46)
    int64_t iteration;
    double gamma, base_gamma, density_weight;
    int32_t hpwl_has_initial, ovf_has_initial;
    double hpwl_initial, hpwl_cur, hpwl_prev;
    double ovf_initial, ovf_cur, ovf_prev;
    // Nesterov (nesterov.rs:6-13)
    double alpha, beta;
    // density weight the current step's Nesterov update used
    double dw_used;
    // reductions of the current step
    double hpwl, overflow, energy, l1_wl, l1_d, num, den;
    int32_t nonfinite, should_stop;
    // the scheduler's next density weight and gamma, computed ahead of the
    // Nesterov update of the same step and applied after it
    double dw_next, gamma_next;
};

// Spectral transform tables for one length L (power of two).
struct SpecTables {
    int L = 0;
    DevBuf fft_tw;   // float2[L/2]: e^{-2 pi i k / L}
    DevBuf dct_tw;   // float2[L]:   e^{-i pi k / (2L)}
};

// Device pointers to the tables of one length: a SpecTables or memory the
// caller provides (the allocation-free entry point's scratch).
struct SpecView {
    const float2 *fft_tw, *dct_tw;
};
inline SpecView spec_view(const SpecTables &t) { return {t.fft_tw.as<float2>(), t.dct_tw.as<float2>()}; }

// ----------------------------------------------------------- launch helpers
int make_spec_tables(SpecTables &t, int L, hipStream_t s);
// fills fft_tw[L/2] and dct_tw[L] (no allocation)
int launch_make_tables(float2 *fft_tw, float2 *dct_tw, int L, hipStream_t s);

// Spectral operators (erp_spectral.hip).  All buffers are device float[m*n].
// Kinds of one-dimensional inverse transforms:
enum InvKind { INV_IDCT = 0, INV_IDXST = 1 };

// a = dctn(x, 2) / (m n)     (scratch: float[m*n])
int launch_dct2(const float *x, float *out, float *scratch, int m, int n, SpecView tm,
                SpecView tn, hipStream_t s);
// out = inverse along x (kind_x) then along y (kind_y); scale 4
int launch_inverse2(const float *x, float *out, float *scratch, int m, int n, InvKind kind_x,
                    InvKind kind_y, SpecView tm, SpecView tn, hipStream_t s);
// The four leaf transforms of the reference (which: 0 dct2_fft2_forward,
// 1 idct2_fft2_forward, 2 idct_idxst_forward, 3 idxst_idct_forward).
int launch_spectral_leaf(const float *x, float *out, float *scratch, int m, int n, int which,
                         SpecView tm, SpecView tn, hipStream_t s);

// Fused potential pipeline (FFT form).
struct PotentialBufs {
    const int64_t *q_fix;   // [B]
    int64_t *q_mov;         // [B]  read and cleared by pass 1
    float *rho;             // [B]  density map after 1/bin_area scaling
    float *rowt;            // [B]  row-DCT intermediate
    float *p1, *p2, *p3;    // [B]  column-inverse intermediates
    float2 *fxy;            // field (x, y) per bin, at fxy_index
    float *pot;             // [B]
    float *auv;             // [B] or nullptr (store DCT coefficients)
    int64_t *ovf_part;      // [m / rows-per-block] int64 overflow partials
    double *energy_part;    // [m / rows-per-block] FP64 energy partials
    int m, n;
    double inv_scale;       // 1/S of the fixed-point accumulators
    int64_t coef_q;         // round(target_density * bin_area * S)
    float inv_bin_area;     // T::one() / bin_area in f32
    float bw, bh;
    bool keep_q;            // leave q_mov intact (parity inspection)
    // 1: the third inverse (potential map) and the energy are computed
    // (epotential.rs:333-346); 0: skipped -- Placement::step discards them
    // (placement.rs:185), the trajectory is unchanged
    int potential;
    double scale;           // S of the fixed-point accumulators
    const float2 *tm_fft, *tm_dct, *tn_fft, *tn_dct;   // spectral tables (length m / n)
};
int potential_partials(int m);   // number of partials pass 1 / pass 3 write
// Potential with MFMA dense transforms (erp_mfma.hip).
struct MfmaPlan;
struct JobDesc;
struct BatchDims;
// wv_div: bw / bh, the aspect factor of the y frequency operator
int mfma_plan_create(MfmaPlan **plan, int m, int n, float wv_div, hipStream_t s);
void mfma_plan_destroy(MfmaPlan *plan);
// jd / bd (the placement's device descriptor and its dimensions): run the WA
// net groups inside the finalize launch
int launch_potential_mfma(const PotentialBufs &b, MfmaPlan *plan, hipStream_t s,
                          KernelTimer *t = nullptr, const JobDesc *jd = nullptr,
                          const BatchDims *bd = nullptr);
bool mfma_supported(int m, int n);             // at least 128 bins per axis
int mfma_energy_partials(int m, int n);        // energy partials the MFMA pipeline writes
int mfma_overflow_partials(int m);             // overflow partials the MFMA pipeline writes
size_t mfma_plan_bytes(const MfmaPlan *plan);
size_t mfma_plan_estimate(int m, int n);       // device bytes a plan for an m x n grid takes
// leaf transforms on device buffers (which: 0 idct2, 1 idct_idxst, 2 idxst_idct)
int launch_dct2_mfma(const float *x, float *out, MfmaPlan *plan, hipStream_t s);
int launch_inverse_mfma(const float *x, float *out, int which, MfmaPlan *plan, hipStream_t s);

}  // namespace erp
