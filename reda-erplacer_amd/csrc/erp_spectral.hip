// Spectral Poisson solve, FFT form (FP32), for gfx950.
//
// The reference computes a 2-D DCT-II (dct2d.rs:256-279) and three inverse
// transforms (idct2.rs:144-166, idct_idxst.rs:137-161, idxst_idct.rs:156-180),
// each as a whole-map pre-process, a 2-D real FFT and a post-process.  Here every
// 2-D transform is separable: a row pass and a column pass, each a batch of 1-D
// transforms done in LDS by a workgroup.  A 1-D DCT-II of length L is Makhoul's
// reordering v[p]=x[2p], v[L-1-p]=x[2p+1], one complex FFT, and a twiddle
// e^{-i pi k/2L}; two real sequences share one complex FFT (real and imaginary
// parts).  The inverse (2 sum_k eps_k Y_k cos(pi(2i+1)k/2L), eps_0=1/2) runs the
// same steps backwards; the sine variant (2 sum_k Y_k sin(pi(2i+1)k/2L)) is the
// cosine one on the reversed spectrum times (-1)^i.
//
// The fused potential pipeline is three passes over the map:
//   pass 1 (rows):    fixed-point accumulators -> density map, overflow partials,
//                     row DCT-II                                  (epotential.rs:271-297)
//   pass 2 (columns): column DCT-II, the three spectral products, three column
//                     inverses                                    (epotential.rs:299-331)
//   pass 3 (rows):    three row inverses -> field_x, field_y, potential; energy
//                     partials                                    (epotential.rs:310-346)
#include <map>
#include <mutex>

#include "erp_internal.h"
#include "erp_kernels.h"
#include "erp_timer.h"
#include "erp_wa.h"

namespace erp {
namespace {

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ int brev(int v, int logL) {
    return (int)(__brev((unsigned)v) >> (32 - logL));
}
// Makhoul order: element c of x sits at v[perm(c)]; x[i] = v[perm(i)] inverts it.
__device__ __forceinline__ int perm(int c, int L) { return (c & 1) ? (L - 1 - (c >> 1)) : (c >> 1); }

// In-place decimation-in-time FFT of F sequences of length 2^logL held in LDS
// at buf[f << logL], input already in bit-reversed order.  tw holds
// e^{-2 pi i k / L}, k < L/2 (conjugated for the inverse).  Unnormalised.
// Two radix-2 stages (half-size h and 2h) are done per LDS round trip: a thread
// loads x[i0 + {0,h,2h,3h}], applies both stages in registers and stores them
// back, so a length-1024 transform takes five barriers instead of ten.
template <bool INV>
__device__ __forceinline__ float2 twid(const float2 *tw, int idx) {
    float2 w = tw[idx];
    if (INV) w.y = -w.y;
    return w;
}

template <bool INV>
__device__ void fft_stages(float2 *buf, int F, int logL, const float2 *tw) {
    const int L = 1 << logL;
    int s = 0;
    for (; s + 1 < logL; s += 2) {
        __syncthreads();
        const int h = 1 << s;
        const int nq = F << (logL - 2);          // radix-4 butterflies over all F sequences
        for (int b = threadIdx.x; b < nq; b += blockDim.x) {
            const int f = b >> (logL - 2);
            const int r = b & ((L >> 2) - 1);
            const int k = r & (h - 1);
            const int i0 = (f << logL) + ((r >> s) << (s + 2)) + k;
            const float2 w1 = twid<INV>(tw, k << (logL - 1 - s));        // W_{2h}^k
            const float2 w2 = twid<INV>(tw, k << (logL - 2 - s));        // W_{4h}^k
            const float2 w3 = twid<INV>(tw, (k + h) << (logL - 2 - s));  // W_{4h}^{k+h}
            const float2 a0 = buf[i0], a1 = buf[i0 + h], a2 = buf[i0 + 2 * h], a3 = buf[i0 + 3 * h];
            const float2 t1 = cmulf(a1, w1), t3 = cmulf(a3, w1);
            const float2 c0 = make_float2(a0.x + t1.x, a0.y + t1.y), c1 = make_float2(a0.x - t1.x, a0.y - t1.y);
            const float2 c2 = make_float2(a2.x + t3.x, a2.y + t3.y), c3 = make_float2(a2.x - t3.x, a2.y - t3.y);
            const float2 u2 = cmulf(c2, w2), u3 = cmulf(c3, w3);
            buf[i0] = make_float2(c0.x + u2.x, c0.y + u2.y);
            buf[i0 + 2 * h] = make_float2(c0.x - u2.x, c0.y - u2.y);
            buf[i0 + h] = make_float2(c1.x + u3.x, c1.y + u3.y);
            buf[i0 + 3 * h] = make_float2(c1.x - u3.x, c1.y - u3.y);
        }
    }
    if (s < logL) {                               // odd log2 length: one radix-2 stage
        __syncthreads();
        const int half = 1 << s;
        const int nb = F << (logL - 1);
        for (int b = threadIdx.x; b < nb; b += blockDim.x) {
            const int f = b >> (logL - 1);
            const int r = b & ((L >> 1) - 1);
            const int k = r & (half - 1);
            const int i0 = (f << logL) + ((r >> s) << (s + 1)) + k;
            const float2 w = twid<INV>(tw, k << (logL - 1 - s));
            const float2 u = buf[i0];
            const float2 v = cmulf(buf[i0 + half], w);
            buf[i0] = make_float2(u.x + v.x, u.y + v.y);
            buf[i0 + half] = make_float2(u.x - v.x, u.y - v.y);
        }
    }
    __syncthreads();
}

__device__ __forceinline__ void load_tw(float2 *dst, const float2 *src, int count) {
    for (int k = threadIdx.x; k < count; k += blockDim.x) dst[k] = src[k];
}

// DCT-II of the two real sequences packed in Z (FFT output), at frequency k:
// returns (y0[k], y1[k]) = (Re(w_k V0[k]), Re(w_k V1[k])) with the 1/2 of the
// unpacking left to the caller's scale.
__device__ __forceinline__ float2 dct_unpack(const float2 *Z, int k, int L, float2 w) {
    const float2 zk = Z[k];
    const float2 zr = Z[(L - k) & (L - 1)];
    const float v0x = zk.x + zr.x, v0y = zk.y - zr.y;    // 2 V0
    const float v1x = zk.y + zr.y, v1y = zr.x - zk.x;    // 2 V1
    return make_float2(w.x * v0x - w.y * v0y, w.x * v1x - w.y * v1y);
}

// Inverse pre-twiddle of one real spectrum at k: e^{+i pi k/2L}(Yk - i Yr).
__device__ __forceinline__ float2 inv_twiddle(float yk, float yr, float2 w) {
    return make_float2(w.x * yk - w.y * yr, -w.x * yr - w.y * yk);
}
// Two real spectra packed for one complex inverse FFT: V0 + i V1.
__device__ __forceinline__ float2 pack2(float2 v0, float2 v1) {
    return make_float2(v0.x - v1.y, v0.y + v1.x);
}
// (Yk, Yr) for frequency k of an IDCT or (reversed spectrum) IDXST input.
template <int KIND>
__device__ __forceinline__ float2 spec_pair(const float *Y, int k, int L, int stride) {
    if (KIND == INV_IDCT) {
        return make_float2(Y[(size_t)k * stride], k ? Y[(size_t)(L - k) * stride] : 0.f);
    } else {
        return make_float2(k ? Y[(size_t)(L - k) * stride] : 0.f, k ? Y[(size_t)k * stride] : 0.f);
    }
}

// ---------------------------------------------------------------- row DCT-II
// R = 2F rows per block.  SRC_FIXED: rows come from the fixed-point accumulators
// (pass 1); otherwise from a float map.  Output T = unscaled row DCT sums.
template <bool SRC_FIXED>
__device__ __forceinline__ void row_dct_body(
    const float *__restrict__ x, int64_t *__restrict__ q_mov, const int64_t *__restrict__ q_fix,
    float *__restrict__ rho_out, int64_t *__restrict__ ovf_part, float *__restrict__ T, int n,
    int logN, int F, const float2 *__restrict__ fft_tw, const float2 *__restrict__ dct_tw,
    double inv_scale, int64_t coef_q, float inv_bin_area, bool clear_q, int bx) {
    extern __shared__ float2 lds[];
    float2 *tw = lds;
    float2 *dtw_s = lds + (n >> 1);   // e^{-i pi k / 2L} staged in LDS
    float2 *buf = dtw_s + n;
    load_tw(tw, fft_tw, n >> 1);
    load_tw(dtw_s, dct_tw, n);
    __syncthreads();
    const int row_base = bx * 2 * F;
    int64_t ovf = 0;
    // chunks of kCh elements per thread: every load of a chunk is issued
    // before any is used
    constexpr int kCh = 4;
    const int total = F << logN;
    for (int idx0 = threadIdx.x; idx0 < total; idx0 += kCh * blockDim.x) {
        int64_t t0[kCh], t1[kCh];
        float xa[kCh], xb[kCh];
#pragma unroll
        for (int r = 0; r < kCh; ++r) {
            const int idx = idx0 + r * blockDim.x;
            if (idx < total) {
                const int f = idx >> logN, c = idx & (n - 1);
                const size_t o0 = (size_t)(row_base + 2 * f) * n + c, o1 = o0 + n;
                if (SRC_FIXED) {
                    t0[r] = q_mov[o0] + q_fix[o0];
                    t1[r] = q_mov[o1] + q_fix[o1];
                } else {
                    xa[r] = x[o0];
                    xb[r] = x[o1];
                }
            }
        }
#pragma unroll
        for (int r = 0; r < kCh; ++r) {
            const int idx = idx0 + r * blockDim.x;
            if (idx >= total) continue;
            const int f = idx >> logN, c = idx & (n - 1);
            const size_t o0 = (size_t)(row_base + 2 * f) * n + c, o1 = o0 + n;
            float a, b;
            if (SRC_FIXED) {
                if (clear_q) {
                    q_mov[o0] = 0;
                    q_mov[o1] = 0;
                }
                ovf += (t0[r] > coef_q ? t0[r] - coef_q : 0) + (t1[r] > coef_q ? t1[r] - coef_q : 0);
                a = (float)((double)t0[r] * inv_scale) * inv_bin_area;
                b = (float)((double)t1[r] * inv_scale) * inv_bin_area;
                rho_out[o0] = a;
                rho_out[o1] = b;
            } else {
                a = xa[r];
                b = xb[r];
            }
            buf[(f << logN) + brev(perm(c, n), logN)] = make_float2(a, b);
        }
    }
    if (SRC_FIXED) {
        __shared__ int64_t red[kBlock];
        const int64_t sum = block_sum_tree(ovf, red);
        if (threadIdx.x == 0) ovf_part[bx] = sum;
    }
    fft_stages<false>(buf, F, logN, tw);
    for (int idx = threadIdx.x; idx < (F << logN); idx += blockDim.x) {
        const int f = idx >> logN, k = idx & (n - 1);
        const float2 y = dct_unpack(buf + (f << logN), k, n, dtw_s[k]);
        const size_t o0 = (size_t)(row_base + 2 * f) * n + k;
        T[o0] = 0.5f * y.x;
        T[o0 + n] = 0.5f * y.y;
    }
}

// leaf: rows of a float map
__global__ void __launch_bounds__(kBlock) k_row_dct(const float *__restrict__ x,
                                                    float *__restrict__ T, int n, int logN, int F,
                                                    const float2 *__restrict__ fft_tw,
                                                    const float2 *__restrict__ dct_tw) {
    row_dct_body<false>(x, nullptr, nullptr, nullptr, nullptr, T, n, logN, F, fft_tw, dct_tw, 0.0,
                        0, 0.f, false, (int)blockIdx.x);
}

// pass 1 of the potential pipeline for job blockIdx.y
__global__ void __launch_bounds__(kBlock) k_pass1(const JobDesc *__restrict__ jd, int logN, int F) {
    if (gridDim.y > 1 && !jd[blockIdx.y].active) return;   // an empty slot of a batch
    const PotentialBufs &b = jd[blockIdx.y].pb;
    row_dct_body<true>(nullptr, b.q_mov, b.q_fix, b.rho, b.ovf_part, b.rowt, b.n, logN, F,
                       b.tn_fft, b.tn_dct, b.inv_scale, b.coef_q, b.inv_bin_area, !b.keep_q,
                       (int)blockIdx.x);
}

// All of the WA plan's workgroups run inside pass 1's launch: 4377 it/s at 1M
// cells against 4367 / 4345 / 4299 with 70 / 50 / 30 % there and the rest inside
// pass 3's, whose 36 KB of LDS per workgroup leave the WA workgroups little room.

// Pass 1 and part of the WA net groups up to degree 256 in one launch:
// workgroups [0, npass) are pass-1 workgroups, the rest loop over their share
// of the WA plan with their own stride.  The two are independent (WA reads the
// positions, pass 1 the accumulators), each workgroup computes exactly what it
// does in its own launch, and the launch needs no second stream and no events
// to overlap them (pass 1's two workgroups per CU leave room).
__global__ void __launch_bounds__(kBlock) k_pass1_wa(const JobDesc *__restrict__ jd, int logN, int F,
                                                     int npass, WaShare wa) {
    if (gridDim.y > 1 && !jd[blockIdx.y].active) return;   // an empty slot of a batch
    if ((int)blockIdx.x < npass) {
        const PotentialBufs &b = jd[blockIdx.y].pb;
        row_dct_body<true>(nullptr, b.q_mov, b.q_fix, b.rho, b.ovf_part, b.rowt, b.n, logN, F,
                           b.tn_fft, b.tn_dct, b.inv_scale, b.coef_q, b.inv_bin_area, !b.keep_q,
                           (int)blockIdx.x);
    } else {
        wa_narrow_blocks(jd[blockIdx.y].wa, wa.plan, wa.first + (blockIdx.x - npass), gridDim.x - npass,
                         wa.last);
    }
}

// -------------------------------------------------------------- column DCT-II
// S = 2G columns per block starting at column blockIdx -> strip (XCD-aware).
// out[u][v] = scale * sum_i T[i][v] cos(pi(2i+1)u/2M)
__device__ __forceinline__ int strip_of_block(int b, int nblocks) {
    // Consecutive workgroups go round-robin over the 8 XCDs; give each XCD a
    // contiguous range of strips so neighbouring strips share its L2.
    const int per = nblocks >> 3;
    return (nblocks & 7) ? b : (b & 7) * per + (b >> 3);
}

__global__ void __launch_bounds__(kBlock) k_col_dct(const float *__restrict__ T,
                                                    float *__restrict__ out, int m, int logM,
                                                    int n, int G,
                                                    const float2 *__restrict__ fft_tw,
                                                    const float2 *__restrict__ dct_tw,
                                                    float scale) {
    extern __shared__ float2 lds[];
    float2 *tw = lds;
    float2 *dtw_s = lds + (m >> 1);   // e^{-i pi k / 2L} staged in LDS
    float2 *buf = dtw_s + m;
    load_tw(tw, fft_tw, m >> 1);
    load_tw(dtw_s, dct_tw, m);
    __syncthreads();
    const int S = 2 * G;
    const int v0 = strip_of_block(blockIdx.x, gridDim.x) * S;
    for (int idx = threadIdx.x; idx < m * S; idx += blockDim.x) {
        const int i = idx / S, s = idx - i * S;
        const float val = T[(size_t)i * n + v0 + s];
        float *dst = reinterpret_cast<float *>(&buf[((s >> 1) << logM) + brev(perm(i, m), logM)]);
        dst[s & 1] = val;
    }
    fft_stages<false>(buf, G, logM, tw);
    for (int idx = threadIdx.x; idx < m * G; idx += blockDim.x) {
        const int u = idx / G, g = idx - u * G;
        const float2 y = dct_unpack(buf + (g << logM), u, m, dtw_s[u]);
        out[(size_t)u * n + v0 + 2 * g] = 0.5f * scale * y.x;
        out[(size_t)u * n + v0 + 2 * g + 1] = 0.5f * scale * y.y;
    }
}

// ---------------------------------------------------------- column inverse
// out[i][v] = 2 sum_u eps_u X[u][v] cos(...)   (KIND = INV_IDCT)
//           = 2 sum_u X[u][v] sin(...)         (KIND = INV_IDXST)
template <int KIND>
__global__ void __launch_bounds__(kBlock) k_col_inv(const float *__restrict__ X,
                                                    float *__restrict__ out, int m, int logM,
                                                    int n, int G,
                                                    const float2 *__restrict__ fft_tw,
                                                    const float2 *__restrict__ dct_tw) {
    extern __shared__ float2 lds[];
    float2 *tw = lds;
    float2 *dtw_s = lds + (m >> 1);   // e^{-i pi k / 2L} staged in LDS
    float2 *buf = dtw_s + m;
    load_tw(tw, fft_tw, m >> 1);
    load_tw(dtw_s, dct_tw, m);
    __syncthreads();
    const int S = 2 * G;
    const int v0 = strip_of_block(blockIdx.x, gridDim.x) * S;
    for (int idx = threadIdx.x; idx < m * G; idx += blockDim.x) {
        const int k = idx / G, g = idx - k * G;
        const float2 w = dtw_s[k];
        const float2 p0 = spec_pair<KIND>(X + v0 + 2 * g, k, m, n);
        const float2 p1 = spec_pair<KIND>(X + v0 + 2 * g + 1, k, m, n);
        buf[(g << logM) + brev(k, logM)] = pack2(inv_twiddle(p0.x, p0.y, w), inv_twiddle(p1.x, p1.y, w));
    }
    fft_stages<true>(buf, G, logM, tw);
    for (int idx = threadIdx.x; idx < m * S; idx += blockDim.x) {
        const int i = idx / S, s = idx - i * S;
        const float2 v = buf[((s >> 1) << logM) + perm(i, m)];
        float r = (s & 1) ? v.y : v.x;
        if (KIND == INV_IDXST && (i & 1)) r = -r;
        out[(size_t)i * n + v0 + s] = r;
    }
}

// ------------------------------------------------------------- row inverse
// R = 2F rows per block; along y.
template <int KIND>
__global__ void __launch_bounds__(kBlock) k_row_inv(const float *__restrict__ X,
                                                    float *__restrict__ out, int n, int logN,
                                                    int F, const float2 *__restrict__ fft_tw,
                                                    const float2 *__restrict__ dct_tw) {
    extern __shared__ float2 lds[];
    float2 *tw = lds;
    float2 *dtw_s = lds + (n >> 1);   // e^{-i pi k / 2L} staged in LDS
    float2 *buf = dtw_s + n;
    load_tw(tw, fft_tw, n >> 1);
    load_tw(dtw_s, dct_tw, n);
    __syncthreads();
    const int row_base = blockIdx.x * 2 * F;
    for (int idx = threadIdx.x; idx < (F << logN); idx += blockDim.x) {
        const int f = idx >> logN, k = idx & (n - 1);
        const float *r0 = X + (size_t)(row_base + 2 * f) * n;
        const float2 w = dtw_s[k];
        const float2 p0 = spec_pair<KIND>(r0, k, n, 1);
        const float2 p1 = spec_pair<KIND>(r0 + n, k, n, 1);
        buf[(f << logN) + brev(k, logN)] = pack2(inv_twiddle(p0.x, p0.y, w), inv_twiddle(p1.x, p1.y, w));
    }
    fft_stages<true>(buf, F, logN, tw);
    for (int idx = threadIdx.x; idx < (F << logN); idx += blockDim.x) {
        const int f = idx >> logN, j = idx & (n - 1);
        const float2 v = buf[(f << logN) + perm(j, n)];
        const float sg = (KIND == INV_IDXST && (j & 1)) ? -1.f : 1.f;
        const size_t o0 = (size_t)(row_base + 2 * f) * n + j;
        out[o0] = sg * v.x;
        out[o0 + n] = sg * v.y;
    }
}

// ------------------------------------------------- fused column pass (pass 2)
// Threads per workgroup of the column pass on grids of 1024 rows and more (at
// 1024^2 the loop runs 4106-4116 it/s with 256, 4135-4149 with 512 and
// 3928-3949 with 768); smaller grids keep kBlock.
constexpr int kColThreadsLarge = 512;

// Column DCT-II of the row-transformed map, a_uv = 4/(MN) * sums; the three
// spectral products (frequencies.rs:4-72 evaluated in f32 on the fly, in the
// reference's operation order); three column inverses:
//   P1 = IDXST_x(a Wx), P2 = IDCT_x(a Wy), P3 = IDCT_x(a inv).
__global__ void __launch_bounds__(kColThreadsLarge) k_col_spectral(const JobDesc *__restrict__ jd, int logM,
                                                         int G, float scale, float two_pi_over_m,
                                                         float two_pi_over_n) {
    if (gridDim.y > 1 && !jd[blockIdx.y].active) return;   // an empty slot of a batch
    extern __shared__ float2 lds[];
    const PotentialBufs &pb = jd[blockIdx.y].pb;
    const float *__restrict__ T = pb.rowt;
    float *__restrict__ P1 = pb.p1;
    float *__restrict__ P2 = pb.p2;
    float *__restrict__ P3 = pb.p3;
    float *__restrict__ auv_out = pb.auv;
    const int ninv = pb.potential ? 3 : 2;         // inverses: field x, field y (, potential)
    const int m = pb.m, n = pb.n;
    const float2 *__restrict__ fft_tw = pb.tm_fft;
    const float2 *__restrict__ dct_tw = pb.tm_dct;
    const float wv_div = pb.bw / pb.bh;
    const int S = 2 * G;
    float2 *tw = lds;
    float2 *dtw_s = lds + (m >> 1);   // e^{-i pi k / 2L} staged in LDS
    float2 *buf = dtw_s + m;                          // 3G complex arrays of m
    // S real arrays of m in the second inverse array's space (unused until the
    // products below write the inverse inputs): 72 KB of LDS at m = 2048, two
    // workgroups per CU
    float *acol = reinterpret_cast<float *>(buf + G * m);
    const int v0 = strip_of_block(blockIdx.x, gridDim.x) * S;
    // the strip's column segments are loaded into registers first (all loads
    // in flight together: one workgroup per CU cannot hide them one by one)
    constexpr int kMaxPer = 16;                      // m * S / blockDim <= 16 for m <= 2048
    const int total = m * S;
    const int per = (total + (int)blockDim.x - 1) / (int)blockDim.x;
    float vals[kMaxPer];
#pragma unroll
    for (int r = 0; r < kMaxPer; ++r) {
        const int idx = threadIdx.x + r * blockDim.x;
        if (r < per && idx < total) {
            const int i = idx / S, s = idx - i * S;
            vals[r] = T[(size_t)i * n + v0 + s];
        }
    }
    load_tw(tw, fft_tw, m >> 1);
    load_tw(dtw_s, dct_tw, m);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kMaxPer; ++r) {
        const int idx = threadIdx.x + r * blockDim.x;
        if (r < per && idx < total) {
            const int i = idx / S, s = idx - i * S;
            float *dst = reinterpret_cast<float *>(&buf[((s >> 1) << logM) + brev(perm(i, m), logM)]);
            dst[s & 1] = vals[r];
        }
    }
    fft_stages<false>(buf, G, logM, tw);
    for (int idx = threadIdx.x; idx < m * G; idx += blockDim.x) {
        const int u = idx / G, g = idx - u * G;
        const float2 y = dct_unpack(buf + (g << logM), u, m, dtw_s[u]);
        const float a0 = 0.5f * scale * y.x, a1 = 0.5f * scale * y.y;
        acol[(2 * g) * m + u] = a0;
        acol[(2 * g + 1) * m + u] = a1;
        if (auv_out) {
            auv_out[(size_t)u * n + v0 + 2 * g] = a0;
            auv_out[(size_t)u * n + v0 + 2 * g + 1] = a1;
        }
    }
    __syncthreads();
    // spectral products and inverse pre-twiddles, all three kinds.  The
    // coefficients share LDS with the inverse inputs, so every thread reads all
    // of its coefficients before any inverse input is written.
    constexpr int kMaxK = 8;                   // m * G / blockDim <= 8 (m <= 2048, G = 1)
    const int nk = (m * G + (int)blockDim.x - 1) / (int)blockDim.x;
    float ca[kMaxK][2], cr[kMaxK][2];
#pragma unroll
    for (int r = 0; r < kMaxK; ++r) {
        const int idx = threadIdx.x + r * blockDim.x;
        if (r < nk && idx < m * G) {
            const int k = idx / G, g = idx - k * G;
            const int kr = (m - k) & (m - 1);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                ca[r][h] = acol[(2 * g + h) * m + k];
                cr[r][h] = k ? acol[(2 * g + h) * m + kr] : 0.f;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kMaxK; ++r) {
        const int idx = threadIdx.x + r * blockDim.x;
        if (r >= nk || idx >= m * G) continue;
        const int k = idx / G, g = idx - k * G;
        const int kr = (m - k) & (m - 1);
        const float2 w = dtw_s[k];
        float2 vx[2], vy[2], vp[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int s = 2 * g + h;
            const int v = v0 + s;
            const float wv = (float)v * two_pi_over_n * wv_div;
            // operator values at rows k and m-k (frequencies.rs:22-57)
            float wu_k = (float)k * two_pi_over_m, wu_r = (float)kr * two_pi_over_m;
            float d_k = wu_k * wu_k + wv * wv, d_r = wu_r * wu_r + wv * wv;
            float inv_k = (k == 0 && v == 0) ? 0.f : 1.f / d_k;
            float inv_r = (kr == 0 && v == 0) ? 0.f : 1.f / d_r;
            const float ak = ca[r][h];
            const float ar = cr[r][h];
            const float wxk = wu_k * inv_k * 0.5f, wxr = wu_r * inv_r * 0.5f;
            const float wyk = wv * inv_k * 0.5f, wyr = wv * inv_r * 0.5f;
            // IDXST along x of a*Wx: reversed spectrum (k -> m-k)
            const float xk = k ? ar * wxr : 0.f, xr = k ? ak * wxk : 0.f;
            vx[h] = inv_twiddle(xk, xr, w);
            vy[h] = inv_twiddle(ak * wyk, ar * wyr, w);
            vp[h] = inv_twiddle(ak * inv_k, ar * inv_r, w);
        }
        const int pos = brev(k, logM);
        buf[((0 * G + g) << logM) + pos] = pack2(vx[0], vx[1]);
        buf[((1 * G + g) << logM) + pos] = pack2(vy[0], vy[1]);
        if (ninv == 3) buf[((2 * G + g) << logM) + pos] = pack2(vp[0], vp[1]);
    }
    fft_stages<true>(buf, ninv * G, logM, tw);
    for (int idx = threadIdx.x; idx < ninv * m * S; idx += blockDim.x) {
        const int kind = idx / (m * S);
        const int rem = idx - kind * m * S;
        const int i = rem / S, s = rem - i * S;
        const float2 v = buf[((kind * G + (s >> 1)) << logM) + perm(i, m)];
        float r = (s & 1) ? v.y : v.x;
        if (kind == 0 && (i & 1)) r = -r;
        float *dst = kind == 0 ? P1 : (kind == 1 ? P2 : P3);
        dst[(size_t)i * n + v0 + s] = r;
    }
}

// ----------------------------------------------------- fused row pass (pass 3)
// field_x = IDCT_y(P1), field_y = IDXST_y(P2), potential = IDCT_y(P3);
// energy partial = sum potential * density (epotential.rs:344-346).
__global__ void __launch_bounds__(kBlock) k_row_final(const JobDesc *__restrict__ jd, int logN, int F) {
    if (gridDim.y > 1 && !jd[blockIdx.y].active) return;   // an empty slot of a batch
    extern __shared__ float2 lds[];
    const PotentialBufs &pb = jd[blockIdx.y].pb;
    const float *__restrict__ P1 = pb.p1;
    const float *__restrict__ P2 = pb.p2;
    const float *__restrict__ P3 = pb.p3;
    const float *__restrict__ rho = pb.rho;
    float2 *__restrict__ fxy = pb.fxy;
    float *__restrict__ pot = pb.pot;
    double *__restrict__ energy_part = pb.energy_part;
    const bool pot_on = pb.potential != 0;   // else: field maps only (energy partials 0)
    const int n = pb.n;
    const float2 *__restrict__ fft_tw = pb.tn_fft;
    const float2 *__restrict__ dct_tw = pb.tn_dct;
    float2 *tw = lds;
    float2 *dtw_s = lds + (n >> 1);   // e^{-i pi k / 2L} staged in LDS
    float2 *buf = dtw_s + n;   // 3F arrays
    load_tw(tw, fft_tw, n >> 1);
    load_tw(dtw_s, dct_tw, n);
    __syncthreads();
    const int row_base = blockIdx.x * 2 * F;
    constexpr int kCh = 4;      // elements per thread whose loads are issued together
    const int total = F << logN;
    for (int idx0 = threadIdx.x; idx0 < total; idx0 += kCh * blockDim.x) {
        float2 q[kCh][6];
#pragma unroll
        for (int r = 0; r < kCh; ++r) {
            const int idx = idx0 + r * blockDim.x;
            if (idx < total) {
                const int f = idx >> logN, k = idx & (n - 1);
                const size_t r0 = (size_t)(row_base + 2 * f) * n;
                q[r][0] = spec_pair<INV_IDCT>(P1 + r0, k, n, 1);
                q[r][1] = spec_pair<INV_IDCT>(P1 + r0 + n, k, n, 1);
                q[r][2] = spec_pair<INV_IDXST>(P2 + r0, k, n, 1);
                q[r][3] = spec_pair<INV_IDXST>(P2 + r0 + n, k, n, 1);
                if (pot_on) {
                    q[r][4] = spec_pair<INV_IDCT>(P3 + r0, k, n, 1);
                    q[r][5] = spec_pair<INV_IDCT>(P3 + r0 + n, k, n, 1);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < kCh; ++r) {
            const int idx = idx0 + r * blockDim.x;
            if (idx >= total) continue;
            const int f = idx >> logN, k = idx & (n - 1);
            const float2 w = dtw_s[k];
            const int pos = (f << logN) + brev(k, logN);
            buf[pos] = pack2(inv_twiddle(q[r][0].x, q[r][0].y, w), inv_twiddle(q[r][1].x, q[r][1].y, w));
            buf[pos + (F << logN)] = pack2(inv_twiddle(q[r][2].x, q[r][2].y, w),
                                           inv_twiddle(q[r][3].x, q[r][3].y, w));
            if (pot_on)
                buf[pos + (2 * F << logN)] = pack2(inv_twiddle(q[r][4].x, q[r][4].y, w),
                                                   inv_twiddle(q[r][5].x, q[r][5].y, w));
        }
    }
    fft_stages<true>(buf, (pot_on ? 3 : 2) * F, logN, tw);
    double e = 0.0;
    for (int idx = threadIdx.x; idx < (F << logN); idx += blockDim.x) {
        const int f = idx >> logN, j = idx & (n - 1);
        const int pos = (f << logN) + perm(j, n);
        const float2 vx = buf[pos], vy = buf[pos + (F << logN)];
        const float sg = (j & 1) ? -1.f : 1.f;
        const int row = row_base + 2 * f;
        fxy[fxy_index(row, j, n)] = make_float2(vx.x, sg * vy.x);
        fxy[fxy_index(row + 1, j, n)] = make_float2(vx.y, sg * vy.y);
        if (pot_on) {
            const float2 vp = buf[pos + (2 * F << logN)];
            const size_t o0 = (size_t)row * n + j;
            pot[o0] = vp.x;
            pot[o0 + n] = vp.y;
            e += (double)(vp.x * rho[o0]) + (double)(vp.y * rho[o0 + n]);
        }
    }
    __shared__ double red[kBlock];
    const double sum = block_sum_tree(e, red);
    if (threadIdx.x == 0) energy_part[blockIdx.x] = sum;
}


__global__ void k_make_tables(float2 *fft_tw, float2 *dct_tw, int L) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < L / 2) {
        double s, c;
        sincospi(-2.0 * (double)k / (double)L, &s, &c);
        fft_tw[k] = make_float2((float)c, (float)s);
    }
    if (k < L) {
        double s, c;
        sincospi(-(double)k / (2.0 * (double)L), &s, &c);
        dct_tw[k] = make_float2((float)c, (float)s);
    }
}

// rows per block of the row passes, columns per block of the column pass
constexpr int kRowPairs = 1;   // F: packed pairs of rows per workgroup (2 rows)
inline int col_pairs(int m) { return 1; }   // one packed column pair per workgroup (more workgroups in flight)

size_t row_lds(int n, int arrays) { return sizeof(float2) * ((size_t)n / 2 + (size_t)n + (size_t)arrays * n); }
constexpr size_t kWaShareLds = 24 * 1024;   // largest pass-1 LDS that WA shares its launch with

// Raise a kernel's dynamic-LDS limit once per (device, kernel) to the largest size
// requested so far, so that captured launches never need the attribute call.
int set_lds_limit(const void *fn, size_t bytes) {
    if (bytes <= 48 * 1024) return ERP_OK;
    static std::mutex mu;
    static std::map<std::pair<int, const void *>, size_t> done;
    int dev = 0;
    ERP_HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    size_t &cur = done[{dev, fn}];
    if (bytes <= cur) return ERP_OK;
    ERP_HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    cur = bytes;
    return ERP_OK;
}

}  // namespace

int launch_make_tables(float2 *fft_tw, float2 *dct_tw, int L, hipStream_t s) {
    hipLaunchKernelGGL(k_make_tables, dim3((L + 255) / 256), dim3(256), 0, s, fft_tw, dct_tw, L);
    ERP_HIP_TRY(hipGetLastError());
    return ERP_OK;
}

int make_spec_tables(SpecTables &t, int L, hipStream_t s) {
    t.L = L;
    int rc = t.fft_tw.alloc(sizeof(float2) * (size_t)(L / 2));
    if (rc) return rc;
    if ((rc = t.dct_tw.alloc(sizeof(float2) * (size_t)L))) return rc;
    return launch_make_tables(t.fft_tw.as<float2>(), t.dct_tw.as<float2>(), L, s);
}

int potential_partials(int m) { return m / (2 * kRowPairs); }

int launch_dct2(const float *x, float *out, float *scratch, int m, int n, SpecView tm,
                SpecView tn, hipStream_t s) {
    const int F = kRowPairs, logN = ilog2(n), logM = ilog2(m), G = col_pairs(m);
    size_t lr = row_lds(n, F);
    int rc = set_lds_limit((const void *)k_row_dct, lr);
    if (rc) return rc;
    hipLaunchKernelGGL(k_row_dct, dim3(m / (2 * F)), dim3(kBlock), lr, s, x, scratch, n, logN, F,
                       tn.fft_tw, tn.dct_tw);
    ERP_HIP_TRY(hipGetLastError());
    size_t lc = row_lds(m, G);
    if ((rc = set_lds_limit((const void *)k_col_dct, lc))) return rc;
    hipLaunchKernelGGL(k_col_dct, dim3(n / (2 * G)), dim3(kBlock), lc, s, scratch, out, m, logM,
                       n, G, tm.fft_tw, tm.dct_tw,
                       (float)(4.0 / ((double)m * (double)n)));
    ERP_HIP_TRY(hipGetLastError());
    return ERP_OK;
}

int launch_inverse2(const float *x, float *out, float *scratch, int m, int n, InvKind kind_x,
                    InvKind kind_y, SpecView tm, SpecView tn, hipStream_t s) {
    const int F = kRowPairs, logN = ilog2(n), logM = ilog2(m), G = col_pairs(m);
    size_t lc = row_lds(m, G);
    const void *colk = kind_x == INV_IDCT ? (const void *)k_col_inv<INV_IDCT>
                                          : (const void *)k_col_inv<INV_IDXST>;
    int rc = set_lds_limit(colk, lc);
    if (rc) return rc;
    if (kind_x == INV_IDCT)
        hipLaunchKernelGGL(k_col_inv<INV_IDCT>, dim3(n / (2 * G)), dim3(kBlock), lc, s, x, scratch,
                           m, logM, n, G, tm.fft_tw, tm.dct_tw);
    else
        hipLaunchKernelGGL(k_col_inv<INV_IDXST>, dim3(n / (2 * G)), dim3(kBlock), lc, s, x,
                           scratch, m, logM, n, G, tm.fft_tw, tm.dct_tw);
    ERP_HIP_TRY(hipGetLastError());
    size_t lr = row_lds(n, F);
    const void *rowk = kind_y == INV_IDCT ? (const void *)k_row_inv<INV_IDCT>
                                          : (const void *)k_row_inv<INV_IDXST>;
    if ((rc = set_lds_limit(rowk, lr))) return rc;
    if (kind_y == INV_IDCT)
        hipLaunchKernelGGL(k_row_inv<INV_IDCT>, dim3(m / (2 * F)), dim3(kBlock), lr, s, scratch,
                           out, n, logN, F, tn.fft_tw, tn.dct_tw);
    else
        hipLaunchKernelGGL(k_row_inv<INV_IDXST>, dim3(m / (2 * F)), dim3(kBlock), lr, s, scratch,
                           out, n, logN, F, tn.fft_tw, tn.dct_tw);
    ERP_HIP_TRY(hipGetLastError());
    return ERP_OK;
}

int launch_spectral_leaf(const float *x, float *out, float *scratch, int m, int n, int which,
                         SpecView tm, SpecView tn, hipStream_t s) {
    switch (which) {
    case 0: return launch_dct2(x, out, scratch, m, n, tm, tn, s);
    case 1: return launch_inverse2(x, out, scratch, m, n, INV_IDCT, INV_IDCT, tm, tn, s);
    case 2: return launch_inverse2(x, out, scratch, m, n, INV_IDCT, INV_IDXST, tm, tn, s);
    case 3: return launch_inverse2(x, out, scratch, m, n, INV_IDXST, INV_IDCT, tm, tn, s);
    default: return fail(ERP_ERR_INVALID, "unknown transform");
    }
}

int launch_potential_fft(const JobDesc *d, const BatchDims &bd, int nj, hipStream_t s,
                         KernelTimer *t, bool with_wa) {
    const int m = bd.m, n = bd.n, F = kRowPairs, logN = ilog2(n), logM = ilog2(m), G = col_pairs(m);
    // pass 1 (with the WA nets: the wide buckets first, on their own)
    size_t l1 = row_lds(n, F);
    int rc;
    const int npass = m / (2 * F);
    // A launch's dynamic LDS is reserved by each of its workgroups, the WA ones
    // included: above kWaShareLds per workgroup (rows of 2048 bins: 40 KB) too
    // few WA workgroups fit a CU, and WA runs in a launch of its own before
    // pass 1 (2M cells at 2048^2: 147 us shared against 128 us apart).
    if (with_wa && l1 > kWaShareLds) {
        if ((rc = launch_wa(d, bd, nj, s, t))) return rc;
        with_wa = false;
    }
    if (with_wa) {
        if ((rc = launch_wa_wide(d, bd, nj, s))) return rc;
        const WaPlan plan = wa_narrow_plan(bd);
        const unsigned total = plan.begin[plan.nseg];
        const WaShare wa{plan, 0u, total};
        const unsigned wa_grid = total;
        if ((rc = set_lds_limit((const void *)k_pass1_wa, l1))) return rc;
        if (t) t->begin("pass1_wa");
        hipLaunchKernelGGL(k_pass1_wa, dim3((unsigned)npass + wa_grid, nj), dim3(kBlock), l1, s, d, logN, F,
                           npass, wa);
        if (t) t->end();
    } else {
        if ((rc = set_lds_limit((const void *)k_pass1, l1))) return rc;
        if (t) t->begin("pass1_row_dct");
        hipLaunchKernelGGL(k_pass1, dim3(npass, nj), dim3(kBlock), l1, s, d, logN, F);
        if (t) t->end();
    }
    ERP_HIP_TRY(hipGetLastError());
    // pass 2 (operator constants exactly as frequencies.rs computes them in T=f32)
    const float two_pi_over_m = (float)(2.0 * M_PI) / (float)m;
    const float two_pi_over_n = (float)(2.0 * M_PI) / (float)n;
    if ((size_t)m * G > 8 * (size_t)kBlock) return fail(ERP_ERR_UNSUPPORTED, "column pass: m * G > 2048");
    size_t l2 = sizeof(float2) * ((size_t)m / 2 + (size_t)m + (size_t)3 * G * m);
    if ((rc = set_lds_limit((const void *)k_col_spectral, l2))) return rc;
    if (t) t->begin("pass2_col_spectral");
    const unsigned col_threads = m >= 1024 ? kColThreadsLarge : kBlock;
    hipLaunchKernelGGL(k_col_spectral, dim3(n / (2 * G), nj), dim3(col_threads), l2, s, d, logM, G,
                       (float)(4.0 / ((double)m * (double)n)), two_pi_over_m, two_pi_over_n);
    if (t) t->end();
    ERP_HIP_TRY(hipGetLastError());
    // pass 3
    size_t l3 = row_lds(n, 3 * F);
    if ((rc = set_lds_limit((const void *)k_row_final, l3))) return rc;
    if (t) t->begin("pass3_row_final");
    hipLaunchKernelGGL(k_row_final, dim3(npass, nj), dim3(kBlock), l3, s, d, logN, F);
    if (t) t->end();
    ERP_HIP_TRY(hipGetLastError());
    return ERP_OK;
}

}  // namespace erp
