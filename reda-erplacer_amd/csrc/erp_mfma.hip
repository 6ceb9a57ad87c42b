// Spectral Poisson solve as dense transform products on MFMA (gfx950).
//
// The 2-D transforms of the potential pipeline are the matrix products
//   a   = 4/(MN) C_M^T rho C_N                         (dct2_fft2_forward)
//   xi_x = 4 S_M (a.Wx diag eps) C_N^T                  (idxst_idct_forward)
//   xi_y = 4 C_M (diag eps a.Wy) S_N^T                  (idct_idxst_forward)
//   phi  = 4 C_M (eps eps a.inv) C_N^T                  (idct2_fft2_forward)
// with C_L[i][k] = cos(pi (2i+1) k / 2L), S_L[i][k] = sin(pi (2i+1) k / 2L).
// Each 1-D transform of length L is halved with the symmetry
// C_L[L-1-i][k] = (-1)^k C_L[i][k] (S_L: (-1)^(k+1)): a forward transform
// folds its input (x_i +- x_{L-1-i}) and multiplies by the even / odd columns of
// C; an inverse multiplies even / odd frequencies separately and unfolds
// (E + O at i, E - O at L-1-i).  Every product is then a pair of
// (rows x L/2) x (L/2 x cols) GEMMs.
//
// Precision: operands are split into bf16 pairs (x = hi + lo, |lo| <= 2^-9 |x|)
// and each product is hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_bf16 with
// FP32 accumulation (relative error ~2^-17 per term).  The operator matrices
// are made once in FP64 and stored split.
//
// Layout between the passes (all row-major m x n floats): the row-transformed
// map and the spectra keep the frequency index de-interleaved (even
// frequencies in the first half, odd in the second) along every transformed
// axis, so each half-GEMM reads contiguous blocks.
#include <cmath>

#include "erp_internal.h"
#include "erp_timer.h"

namespace erp {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

enum { FWD_ROW = 0, FWD_COL = 1, INV_COL = 2, INV_ROW = 3 };
// operator planes of one length (each (L/2)^2, hi then lo):
//   forward  [k'][i] = C[i][2k' + p]         (FE: p = 0, FO: p = 1)
//   inverse  [i][k'] = C[i][2k' + p], S[...]  (CE, CO, SE, SO)
enum { T_FE = 0, T_FO = 1, T_CE = 2, T_CO = 3, T_SE = 4, T_SO = 5, T_COUNT = 6 };

constexpr int kBM = 64;              // output rows per workgroup (each half product)
constexpr int kBN = 32;              // output columns per workgroup (each half product)
constexpr int kBK = 64;              // K per stage (two MFMA steps of 32)
constexpr int kLd = kBK + 8;         // LDS row length in bf16 (16-byte pad)
constexpr int kAPlane = kBM * kLd;
constexpr int kBPlane = kBN * kLd;
constexpr int kStage = 4 * kAPlane + 4 * kBPlane;   // planes: [product e/o][hi/lo]

struct MfmaArgs {
    const float *src[3];
    float *dst[3];
    int dst_stride[3];        // INV_ROW: element stride of each output (2 for the field pair)
    const __bf16 *tab;        // operator planes of the transformed length
    int H;                    // half length (the K of every product)
    int m, n;
    int sine_mask;            // bit z: launch z uses the sine operator
    int spectral;             // FWD_COL: write the three spectra (else plain a)
    float scale, two_pi_over_m, two_pi_over_n, wv_div;
    float *auv;               // FWD_COL: a in natural order (nullable)
    const float *rho;         // INV_ROW: density for the energy of launch z = 2
    double *energy_part;
};

__device__ __forceinline__ void split8(const float *v, bf16x8 &hi, bf16x8 &lo) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const __bf16 h = (__bf16)v[j];
        hi[j] = h;
        lo[j] = (__bf16)(v[j] - (float)h);
    }
}

__device__ __forceinline__ const __bf16 *plane(const MfmaArgs &a, int t, int hl) {
    return a.tab + (size_t)(2 * t + hl) * a.H * a.H;
}

// Global -> register staging of one K stage (two stages are in flight).
struct Stage {
    float4 d[8];          // data operand
    bf16x8 o[8];          // operator operand
};

// Data as the A operand (rows x K, K contiguous): FWD_ROW folds column k with
// n-1-k, INV_ROW takes the two de-interleaved halves.  Thread: one row, 16 K.
template <int MODE>
__device__ __forceinline__ void load_a_data(const MfmaArgs &a, const float *src, int r0, int k0,
                                            Stage &s) {
    const int row = threadIdx.x >> 2, kk = (threadIdx.x & 3) * 16;
    const float *p = src + (size_t)(r0 + row) * a.n;
    const float4 *x0 = reinterpret_cast<const float4 *>(p + k0 + kk);
    const float4 *x1 = MODE == FWD_ROW ? reinterpret_cast<const float4 *>(p + a.n - 16 - k0 - kk)
                                       : reinterpret_cast<const float4 *>(p + a.H + k0 + kk);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        s.d[q] = x0[q];
        s.d[4 + q] = x1[q];
    }
}
template <int MODE>
__device__ __forceinline__ void store_a_data(const Stage &s, __bf16 *lds) {
    const int row = threadIdx.x >> 2, kk = (threadIdx.x & 3) * 16;
    const float *lo = reinterpret_cast<const float *>(&s.d[0]);
    const float *hi = reinterpret_cast<const float *>(&s.d[4]);
#pragma unroll
    for (int h8 = 0; h8 < 2; ++h8) {
        float e[8], o[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int jj = 8 * h8 + j;
            if (MODE == FWD_ROW) { e[j] = lo[jj] + hi[15 - jj]; o[j] = lo[jj] - hi[15 - jj]; }
            else { e[j] = lo[jj]; o[j] = hi[jj]; }
        }
        const int at = row * kLd + kk + 8 * h8;
        bf16x8 vh, vl;
        split8(e, vh, vl);
        *reinterpret_cast<bf16x8 *>(lds + 0 * kAPlane + at) = vh;
        *reinterpret_cast<bf16x8 *>(lds + 1 * kAPlane + at) = vl;
        split8(o, vh, vl);
        *reinterpret_cast<bf16x8 *>(lds + 2 * kAPlane + at) = vh;
        *reinterpret_cast<bf16x8 *>(lds + 3 * kAPlane + at) = vl;
    }
}

// Data as the B operand (K x cols, K along rows of the map): FWD_COL folds row
// k with 2H-1-k, INV_COL takes the two de-interleaved halves.  Stored
// transposed (K contiguous per column).  Thread: two adjacent K rows, 4 columns.
template <int MODE>
__device__ __forceinline__ void load_b_data(const MfmaArgs &a, const float *src, int c0, int k0,
                                            Stage &s) {
    const int k2 = (threadIdx.x >> 3) * 2, cl = (threadIdx.x & 7) * 4;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int r = k0 + k2 + q;
        const int r1 = MODE == FWD_COL ? 2 * a.H - 1 - r : a.H + r;
        s.d[q] = *reinterpret_cast<const float4 *>(src + (size_t)r * a.n + c0 + cl);
        s.d[2 + q] = *reinterpret_cast<const float4 *>(src + (size_t)r1 * a.n + c0 + cl);
    }
}
template <int MODE>
__device__ __forceinline__ void store_b_data(const Stage &s, __bf16 *lds) {
    const int k2 = (threadIdx.x >> 3) * 2, cl = (threadIdx.x & 7) * 4;
    const float *x0 = reinterpret_cast<const float *>(&s.d[0]);
    const float *x1 = reinterpret_cast<const float *>(&s.d[1]);
    const float *y0 = reinterpret_cast<const float *>(&s.d[2]);
    const float *y1 = reinterpret_cast<const float *>(&s.d[3]);
    __bf16 *B = lds + 4 * kAPlane;
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float e0 = MODE == FWD_COL ? x0[q] + y0[q] : x0[q];
        const float e1 = MODE == FWD_COL ? x1[q] + y1[q] : x1[q];
        const float o0 = MODE == FWD_COL ? x0[q] - y0[q] : y0[q];
        const float o1 = MODE == FWD_COL ? x1[q] - y1[q] : y1[q];
        bf16x2 eh, el, oh, ol;
        eh[0] = (__bf16)e0; eh[1] = (__bf16)e1;
        el[0] = (__bf16)(e0 - (float)eh[0]); el[1] = (__bf16)(e1 - (float)eh[1]);
        oh[0] = (__bf16)o0; oh[1] = (__bf16)o1;
        ol[0] = (__bf16)(o0 - (float)oh[0]); ol[1] = (__bf16)(o1 - (float)oh[1]);
        const int at = (cl + q) * kLd + k2;
        *reinterpret_cast<bf16x2 *>(B + 0 * kBPlane + at) = eh;
        *reinterpret_cast<bf16x2 *>(B + 1 * kBPlane + at) = el;
        *reinterpret_cast<bf16x2 *>(B + 2 * kBPlane + at) = oh;
        *reinterpret_cast<bf16x2 *>(B + 3 * kBPlane + at) = ol;
    }
}

// Operator as the A operand: 64 rows x 64 K of four planes (thread: one row of
// one plane).
__device__ __forceinline__ void load_a_op(const MfmaArgs &a, int te, int r0, int k0, Stage &s) {
    const int pl = threadIdx.x >> 6, row = threadIdx.x & 63;
    const int t = te + (pl >> 1);   // even / odd product
    const bf16x8 *p = reinterpret_cast<const bf16x8 *>(plane(a, t, pl & 1) + (size_t)(r0 + row) * a.H + k0);
#pragma unroll
    for (int q = 0; q < 8; ++q) s.o[q] = p[q];
}
__device__ __forceinline__ void store_a_op(const Stage &s, __bf16 *lds) {
    const int pl = threadIdx.x >> 6, row = threadIdx.x & 63;
    bf16x8 *d = reinterpret_cast<bf16x8 *>(lds + pl * kAPlane + row * kLd);
#pragma unroll
    for (int q = 0; q < 8; ++q) d[q] = s.o[q];
}
// Operator as the B operand: 32 columns x 64 K of four planes (thread: half a
// column row of one plane).
__device__ __forceinline__ void load_b_op(const MfmaArgs &a, int te, int c0, int k0, Stage &s) {
    const int pl = threadIdx.x >> 6, col = (threadIdx.x >> 1) & 31, half = threadIdx.x & 1;
    const int t = te + (pl >> 1);
    const bf16x8 *p = reinterpret_cast<const bf16x8 *>(plane(a, t, pl & 1) + (size_t)(c0 + col) * a.H + k0 + 32 * half);
#pragma unroll
    for (int q = 0; q < 4; ++q) s.o[q] = p[q];
}
__device__ __forceinline__ void store_b_op(const Stage &s, __bf16 *lds) {
    const int pl = threadIdx.x >> 6, col = (threadIdx.x >> 1) & 31, half = threadIdx.x & 1;
    bf16x8 *d = reinterpret_cast<bf16x8 *>(lds + 4 * kAPlane + pl * kBPlane + col * kLd + 32 * half);
#pragma unroll
    for (int q = 0; q < 4; ++q) d[q] = s.o[q];
}

// XCD-aware tile order: consecutive workgroups go round-robin over the 8 XCDs;
// give each XCD a contiguous run of tiles (tiles of one row band share data).
__device__ __forceinline__ int tile_of_block(int b, int nb) {
    return (nb & 7) ? b : (b & 7) * (nb >> 3) + (b >> 3);
}

template <int MODE>
__global__ void __launch_bounds__(256) k_mfma(MfmaArgs a) {
    __shared__ __attribute__((aligned(16))) __bf16 lds[kStage];
    const int z = blockIdx.y;
    const bool sine = (a.sine_mask >> z) & 1;
    const int te = (MODE == FWD_ROW || MODE == FWD_COL) ? T_FE : (sine ? T_SE : T_CE);
    // output grid of the half products: rows x cols
    const int cols = (MODE == FWD_ROW || MODE == INV_ROW) ? a.n / 2 : a.n;
    const int ncolt = cols / kBN;
    const int tile = tile_of_block(blockIdx.x, gridDim.x);
    const int r0 = (tile / ncolt) * kBM, c0 = (tile % ncolt) * kBN;
    const float *src = a.src[z];
    const int nk = a.H / kBK;

    auto load = [&](int ks, Stage &st) {
        const int k0 = ks * kBK;
        if (MODE == FWD_ROW || MODE == INV_ROW) {
            load_a_data<MODE>(a, src, r0, k0, st);
            load_b_op(a, te, c0, k0, st);
        } else {
            load_a_op(a, te, r0, k0, st);
            load_b_data<MODE>(a, src, c0, k0, st);
        }
    };
    auto store = [&](const Stage &st) {
        if (MODE == FWD_ROW || MODE == INV_ROW) {
            store_a_data<MODE>(st, lds);
            store_b_op(st, lds);
        } else {
            store_a_op(st, lds);
            store_b_data<MODE>(st, lds);
        }
    };

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int lr = lane & 15, lk = 8 * (lane >> 4);
    f32x4 acc[2][2];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int rt = 0; rt < 2; ++rt) acc[p][rt] = f32x4{0.f, 0.f, 0.f, 0.f};

    auto compute = [&]() {
#pragma unroll
        for (int kq = 0; kq < kBK; kq += 32) {
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const __bf16 *B = lds + 4 * kAPlane + (2 * p) * kBPlane + (wc * 16 + lr) * kLd + lk + kq;
                const bf16x8 bh = *reinterpret_cast<const bf16x8 *>(B);
                const bf16x8 bl = *reinterpret_cast<const bf16x8 *>(B + kBPlane);
#pragma unroll
                for (int rt = 0; rt < 2; ++rt) {
                    const __bf16 *A = lds + (2 * p) * kAPlane + (wr * 32 + rt * 16 + lr) * kLd + lk + kq;
                    const bf16x8 ah = *reinterpret_cast<const bf16x8 *>(A);
                    const bf16x8 al = *reinterpret_cast<const bf16x8 *>(A + kAPlane);
                    acc[p][rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc[p][rt], 0, 0, 0);
                    acc[p][rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc[p][rt], 0, 0, 0);
                    acc[p][rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc[p][rt], 0, 0, 0);
                }
            }
        }
    };

    // two register stages in flight; one LDS stage
    Stage s0, s1;
    load(0, s0);
    if (nk > 1) load(1, s1);
    for (int ks = 0; ks < nk; ks += 2) {
        __syncthreads();                 // the previous stage has been consumed
        store(s0);
        __syncthreads();
        if (ks + 2 < nk) load(ks + 2, s0);
        compute();
        if (ks + 1 < nk) {
            __syncthreads();
            store(s1);
            __syncthreads();
            if (ks + 3 < nk) load(ks + 3, s1);
            compute();
        }
    }

    // ------------------------------------------------------------- epilogue
    // lane holds P[row = wr*32 + rt*16 + 4*(lane/16) + j][col = wc*16 + lane%16]
    const int n = a.n, m = a.m;
    double energy = 0.0;
    const int col = c0 + wc * 16 + lr;
#pragma unroll
    for (int rt = 0; rt < 2; ++rt) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int row = r0 + wr * 32 + rt * 16 + 4 * (lane >> 4) + j;
            const float pe = acc[0][rt][j], po = acc[1][rt][j];
            if (MODE == FWD_ROW) {
                float *d = a.dst[z] + (size_t)row * n;
                d[col] = pe;
                d[n / 2 + col] = po;
            } else if (MODE == FWD_COL) {
                // row = k' (frequency pair), col = de-interleaved column position
                const int v = col < n / 2 ? 2 * col : 2 * (col - n / 2) + 1;
#pragma unroll
                for (int p = 0; p < 2; ++p) {
                    const int u = 2 * row + p;
                    const float av = a.scale * (p ? po : pe);
                    const size_t at = (size_t)(p * (m / 2) + row) * n + col;
                    if (a.auv) a.auv[(size_t)u * n + v] = av;
                    if (!a.spectral) {
                        a.dst[0][(size_t)u * n + v] = av;
                        continue;
                    }
                    // frequencies.rs:22-57 in f32
                    const float wu = (float)u * a.two_pi_over_m;
                    const float wv = (float)v * a.two_pi_over_n * a.wv_div;
                    const float den = wu * wu + wv * wv;
                    const float inv = (u == 0 && v == 0) ? 0.f : 1.f / den;
                    const float wx = wu * inv * 0.5f, wy = wv * inv * 0.5f;
                    const float eu = u ? 1.f : 0.5f, ev = v ? 1.f : 0.5f;
                    a.dst[0][at] = av * wx * ev;
                    a.dst[1][at] = eu * (av * wy);
                    a.dst[2][at] = eu * ev * (av * inv);
                }
            } else if (MODE == INV_COL) {
                float *d = a.dst[z];
                d[(size_t)row * n + col] = pe + po;
                d[(size_t)(m - 1 - row) * n + col] = sine ? po - pe : pe - po;
            } else {
                const int st = a.dst_stride[z];   // 2: a component of the field map (fxy_index)
                const float x0 = 4.f * (pe + po);
                const float x1 = 4.f * (sine ? po - pe : pe - po);
                if (st == 2) {
                    a.dst[z][2 * fxy_index(row, col, n)] = x0;
                    a.dst[z][2 * fxy_index(row, n - 1 - col, n)] = x1;
                } else {
                    float *d = a.dst[z] + (size_t)row * n;
                    d[col] = x0;
                    d[n - 1 - col] = x1;
                }
                if (a.energy_part && z == 2) {
                    const float *rr = a.rho + (size_t)row * n;
                    energy += (double)(x0 * rr[col]) + (double)(x1 * rr[n - 1 - col]);
                }
            }
        }
    }
    if (MODE == INV_ROW) {
        if (a.energy_part && z == 2) {
            __shared__ double red[256];
            red[threadIdx.x] = energy;
            __syncthreads();
            for (int s = 128; s > 0; s >>= 1) {
                if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
                __syncthreads();
            }
            if (threadIdx.x == 0) a.energy_part[blockIdx.x] = red[0];
        }
    }
}

// Density from the fixed-point accumulators (epotential.rs:271-285): rho,
// overflow partials per 4 rows, movable accumulators cleared.
__global__ void __launch_bounds__(1024) k_finalize(PotentialBufs b) {
    const int n = b.n;
    const size_t base = (size_t)blockIdx.x * 4 * n;
    const longlong2 *__restrict__ qf = reinterpret_cast<const longlong2 *>(b.q_fix + base);
    longlong2 *__restrict__ qm = reinterpret_cast<longlong2 *>(b.q_mov + base);
    float2 *__restrict__ rho = reinterpret_cast<float2 *>(b.rho + base);
    const int pairs = 2 * n;                  // 4 rows as int64 pairs
    int64_t ovf = 0;
#pragma unroll 2
    for (int k = threadIdx.x; k < pairs; k += 1024) {
        const longlong2 a = qm[k], f = qf[k];
        if (!b.keep_q) qm[k] = make_longlong2(0, 0);
        const int64_t t0 = a.x + f.x, t1 = a.y + f.y;
        ovf += (t0 > b.coef_q ? t0 - b.coef_q : 0) + (t1 > b.coef_q ? t1 - b.coef_q : 0);
        rho[k] = make_float2((float)((double)t0 * b.inv_scale) * b.inv_bin_area,
                             (float)((double)t1 * b.inv_scale) * b.inv_bin_area);
    }
    __shared__ int64_t red[16];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) ovf += __shfl_xor(ovf, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ovf;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t v = 0;
        for (int w = 0; w < 16; ++w) v += red[w];
        b.ovf_part[blockIdx.x] = v;
    }
}

// Natural-order spectrum -> de-interleaved layout with the eps factors of the
// inverse (leaf entry points): out[p(u)][p(v)] = eps_u^eu eps_v^ev X[u][v].
__global__ void k_deinterleave(const float *__restrict__ x, float *__restrict__ out, int m, int n,
                               int eps_u, int eps_v) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)m * n) return;
    const int u = (int)(idx / n), v = (int)(idx % n);
    float f = x[idx];
    if (eps_u && u == 0) f *= 0.5f;
    if (eps_v && v == 0) f *= 0.5f;
    const int pu = (u & 1) ? m / 2 + u / 2 : u / 2;
    const int pv = (v & 1) ? n / 2 + v / 2 : v / 2;
    out[(size_t)pu * n + pv] = f;
}

// Operator planes of length L in FP64, split to bf16 pairs.
__global__ void k_make_planes(__bf16 *tab, int L) {
    const int H = L / 2;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)T_COUNT * H * H) return;
    const int t = (int)(idx / ((int64_t)H * H));
    const int rem = (int)(idx % ((int64_t)H * H));
    const int r = rem / H, c = rem % H;
    int i, k;
    if (t == T_FE || t == T_FO) { k = 2 * r + (t == T_FO); i = c; }
    else { i = r; k = 2 * c + (t & 1); }
    // angle pi (2i+1) k / 2L, reduced exactly modulo 2 pi
    const int64_t num = ((int64_t)(2 * i + 1) * k) % (4 * (int64_t)L);
    double s, co;
    sincospi((double)num / (double)(2 * L), &s, &co);
    const double v = (t == T_SE || t == T_SO) ? s : co;
    const float f = (float)v;
    const __bf16 hi = (__bf16)f;
    const __bf16 lo = (__bf16)(float)(v - (double)(float)hi);
    tab[(size_t)(2 * t) * H * H + rem] = hi;
    tab[(size_t)(2 * t + 1) * H * H + rem] = lo;
}

inline unsigned cdiv(int64_t a, int b) { return (unsigned)((a + b - 1) / b); }

}  // namespace

struct MfmaPlan {
    int m = 0, n = 0;
    DevBuf tab_m, tab_n;          // operator planes (tab_n unused when m == n)
    DevBuf t0, y0, y1, y2;        // scratch maps for the leaf entry points / inverses
    const __bf16 *tm() const { return tab_m.as<__bf16>(); }
    const __bf16 *tn() const { return m == n ? tab_m.as<__bf16>() : tab_n.as<__bf16>(); }
};

bool mfma_supported(int m, int n) {
    return m >= 2 * kBM && n >= 2 * kBN * 2 && m % (2 * kBM) == 0 && n % (2 * kBN) == 0;
}

int mfma_energy_partials(int m, int n) { return (m / kBM) * ((n / 2) / kBN); }

// k_finalize runs one workgroup per four rows and writes one overflow partial each.
int mfma_overflow_partials(int m) { return m / 4; }

int mfma_plan_create(MfmaPlan **out, int m, int n, hipStream_t s) {
    if (!mfma_supported(m, n))
        return fail(ERP_ERR_UNSUPPORTED, "the MFMA spectral form needs at least 128 bins per axis");
    MfmaPlan *p = new MfmaPlan();
    p->m = m;
    p->n = n;
    const size_t B = (size_t)m * n;
    int rc = ERP_OK;
    auto make = [&](DevBuf &b, int L) {
        const int64_t cnt = (int64_t)T_COUNT * (L / 2) * (L / 2);
        if ((rc = b.alloc(sizeof(__bf16) * 2 * (size_t)cnt))) return;
        hipLaunchKernelGGL(k_make_planes, dim3(cdiv(cnt, 256)), dim3(256), 0, s, b.as<__bf16>(), L);
        if (hipGetLastError() != hipSuccess) rc = fail(ERP_ERR_HIP, "k_make_planes launch failed");
    };
    make(p->tab_m, m);
    if (!rc && n != m) make(p->tab_n, n);
    for (DevBuf *b : {&p->t0, &p->y0, &p->y1, &p->y2})
        if (!rc) rc = b->alloc(sizeof(float) * B);
    if (rc) {
        delete p;
        return rc;
    }
    *out = p;
    return ERP_OK;
}

void mfma_plan_destroy(MfmaPlan *plan) { delete plan; }

size_t mfma_plan_bytes(const MfmaPlan *p) {
    return p ? p->tab_m.bytes + p->tab_n.bytes + p->t0.bytes + p->y0.bytes + p->y1.bytes + p->y2.bytes : 0;
}

namespace {

template <int MODE>
int launch_mode(const MfmaArgs &a, int nz, hipStream_t s) {
    const int rows = (MODE == FWD_ROW || MODE == INV_ROW) ? a.m : a.m / 2;
    const int cols = (MODE == FWD_ROW || MODE == INV_ROW) ? a.n / 2 : a.n;
    const unsigned nb = (unsigned)((rows / kBM) * (cols / kBN));
    hipLaunchKernelGGL(k_mfma<MODE>, dim3(nb, nz), dim3(256), 0, s, a);
    ERP_HIP_TRY(hipGetLastError());
    return ERP_OK;
}

MfmaArgs base_args(const MfmaPlan *p, int m, int n) {
    MfmaArgs a{};
    a.m = m;
    a.n = n;
    (void)p;
    return a;
}

// rows of x (m x n, natural) -> de-interleaved row transform T; then columns.
int forward(const MfmaPlan *p, const float *x, float *T, const MfmaArgs &colargs, hipStream_t s,
            KernelTimer *t) {
    MfmaArgs a = base_args(p, p->m, p->n);
    a.src[0] = x;
    a.dst[0] = T;
    a.tab = p->tn();
    a.H = p->n / 2;
    if (t) t->begin("mfma_fwd_row");
    int rc = launch_mode<FWD_ROW>(a, 1, s);
    if (t) t->end();
    if (rc) return rc;
    MfmaArgs c = colargs;
    c.src[0] = T;
    c.tab = p->tm();
    c.H = p->m / 2;
    if (t) t->begin("mfma_fwd_col");
    rc = launch_mode<FWD_COL>(c, 1, s);
    if (t) t->end();
    return rc;
}

// de-interleaved spectra X[z] -> column inverses Y[z] -> row inverses out[z].
int inverse(const MfmaPlan *p, const float *const *X, float *const *Y, float *const *out,
            const int *out_stride, int nz, int col_sine, int row_sine, const float *rho,
            double *energy_part, hipStream_t s, KernelTimer *t) {
    MfmaArgs a = base_args(p, p->m, p->n);
    for (int z = 0; z < nz; ++z) { a.src[z] = X[z]; a.dst[z] = Y[z]; }
    a.tab = p->tm();
    a.H = p->m / 2;
    a.sine_mask = col_sine;
    if (t) t->begin("mfma_inv_col");
    int rc = launch_mode<INV_COL>(a, nz, s);
    if (t) t->end();
    if (rc) return rc;
    MfmaArgs b = base_args(p, p->m, p->n);
    for (int z = 0; z < nz; ++z) { b.src[z] = Y[z]; b.dst[z] = out[z]; b.dst_stride[z] = out_stride ? out_stride[z] : 1; }
    b.tab = p->tn();
    b.H = p->n / 2;
    b.sine_mask = row_sine;
    b.rho = rho;
    b.energy_part = energy_part;
    if (t) t->begin("mfma_inv_row");
    rc = launch_mode<INV_ROW>(b, nz, s);
    if (t) t->end();
    return rc;
}

}  // namespace

// The potential pipeline of one placement (epotential.rs:225-374) in MFMA form.
int launch_potential_mfma(const PotentialBufs &b, MfmaPlan *p, hipStream_t s, KernelTimer *t) {
    if (!p || p->m != b.m || p->n != b.n) return fail(ERP_ERR_INVALID, "MFMA plan does not match the grid");
    const int m = b.m, n = b.n;
    if (t) t->begin("mfma_finalize");
    hipLaunchKernelGGL(k_finalize, dim3((unsigned)mfma_overflow_partials(m)), dim3(1024), 0, s, b);
    if (t) t->end();
    ERP_HIP_TRY(hipGetLastError());
    MfmaArgs c = base_args(p, m, n);
    c.dst[0] = b.p1;
    c.dst[1] = b.p2;
    c.dst[2] = b.p3;
    c.spectral = 1;
    c.scale = (float)(4.0 / ((double)m * (double)n));
    c.two_pi_over_m = (float)(2.0 * M_PI) / (float)m;
    c.two_pi_over_n = (float)(2.0 * M_PI) / (float)n;
    c.wv_div = b.bw / b.bh;
    c.auv = b.auv;
    int rc = forward(p, b.rho, b.rowt, c, s, t);
    if (rc) return rc;
    const float *X[3] = {b.p1, b.p2, b.p3};
    float *Y[3] = {p->y0.as<float>(), p->y1.as<float>(), p->y2.as<float>()};
    float *fxy = reinterpret_cast<float *>(b.fxy);
    float *out[3] = {fxy, fxy + 1, b.pot};
    const int stride[3] = {2, 2, 1};
    // xi_x: sine along x, cosine along y; xi_y: cosine along x, sine along y; phi: cosines
    // the potential (third inverse) and its energy only when requested
    return inverse(p, X, Y, out, stride, b.potential ? 3 : 2, /*col_sine=*/1, /*row_sine=*/2, b.rho,
                   b.potential ? b.energy_part : nullptr, s, t);
}

// Leaf transforms (host-facing entry points run these on device buffers).
int launch_dct2_mfma(const float *x, float *out, MfmaPlan *p, hipStream_t s) {
    MfmaArgs c = base_args(p, p->m, p->n);
    c.dst[0] = out;
    c.spectral = 0;
    c.scale = (float)(4.0 / ((double)p->m * (double)p->n));
    return forward(p, x, p->t0.as<float>(), c, s, nullptr);
}

// which: 0 idct2 (cos, cos), 1 idct_idxst (cos along x, sin along y),
//        2 idxst_idct (sin along x, cos along y)
int launch_inverse_mfma(const float *x, float *out, int which, MfmaPlan *p, hipStream_t s) {
    const int m = p->m, n = p->n;
    const int eps_u = which != 2, eps_v = which != 1;
    hipLaunchKernelGGL(k_deinterleave, dim3(cdiv((int64_t)m * n, 256)), dim3(256), 0, s, x,
                       p->t0.as<float>(), m, n, eps_u, eps_v);
    ERP_HIP_TRY(hipGetLastError());
    const float *X[1] = {p->t0.as<float>()};
    float *Y[1] = {p->y0.as<float>()};
    float *O[1] = {out};
    return inverse(p, X, Y, O, nullptr, 1, which == 2 ? 1 : 0, which == 1 ? 1 : 0, nullptr, nullptr, s,
                   nullptr);
}

}  // namespace erp
