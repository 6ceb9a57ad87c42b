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
// (rows x L/2) x (L/2 x cols) GEMMs, K = L/2.
//
// The three inverses share one spectrum.  With b = a.inv (frequencies.rs:22-57:
// Wx = wu inv / 2, Wy = wv inv / 2) the diagonal factors move into the
// operators:
//   xi_x = 4 [S_M diag(wu/2)] b [diag(eps) C_N^T] = 4 V [diag(eps) C_N^T]
//   xi_y = 4 [C_M diag(eps)]  b [diag(wv/2) S_N^T] = 4 U [diag(wv/2) S_N^T]
//   phi  = 4 [C_M diag(eps)]  b [diag(eps) C_N^T]  = 4 U [diag(eps) C_N^T]
// so the inverse along x is two products of the one map b (U and V) and the
// inverse along y three (xi_x from V, xi_y and phi from U): five products
// where the three separate 2-D inverses take six, and the forward pass writes
// one spectrum instead of three.  The scaled operators are made in FP64.
//
// Precision: operands are split into bf16 pairs (x = hi + lo, |lo| <= 2^-9 |x|)
// and each product is hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_bf16 with
// FP32 accumulation (relative error ~2^-17 per term).  The operator matrices
// are made once in FP64 and stored split.
//
// Data flow: every pass writes its result already folded, split into bf16
// hi / lo and laid out the way the next pass reads it -- four "planes" per map
// ([product half e / o][hi / lo], K contiguous) -- so the GEMM kernel only
// copies bf16 tiles into LDS:
//   k_finalize : q -> rho (f32) and the row pass operand, rows folded along y;
//                in the iteration the WA net groups share its launch
//   FWD_ROW    : -> the column pass operand, folded along x (each workgroup owns
//                a band of rows and its mirror image, so it can fold)
//   FWD_COL    : -> the spectrum b = a.inv, split by the parity of the x
//                frequency
//   INV_COL    : -> U and V, unfolded along x, split by the parity of the y
//                frequency
//   INV_ROW    : -> xi_x, xi_y (the interleaved field map), phi, energy
#include <cmath>

#include "erp_internal.h"
#include "erp_kernels.h"
#include "erp_timer.h"
#include "erp_wa.h"

namespace erp {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));   // 16 bytes of operand tile

enum { FWD_ROW = 0, FWD_COL = 1, INV_COL = 2, INV_ROW = 3 };
// operator planes of one length (each (L/2)^2, hi then lo):
//   forward  [k'][i] = C[i][2k' + p]         (FE: p = 0, FO: p = 1)
//   inverse  [i][k'] = C[i][2k' + p], S[...]  (CE, CO, SE, SO)
enum { T_FE = 0, T_FO = 1, T_CE = 2, T_CO = 3, T_SE = 4, T_SO = 5, T_COUNT = 6 };
// the scaled inverse operators of the potential pipeline along one axis, same
// layout: [i][k'] = C[i][k] eps_k (k = 2k' + p) and S[i][k] w_k / 2
enum { W_CE = 0, W_CO = 1, W_SE = 2, W_SO = 3, W_COUNT = 4 };

constexpr int kKb = 32;                 // K per stage: one 16x16x32 MFMA step
// Output tile of a workgroup (rows and columns per product half): TT = 64
// (four waves of 32 x 32) or, on grids with many tiles, TT = 128 (eight waves
// of 32 x 64): each workgroup reads its whole A and B panels from L2, so the
// larger tile halves the L2 -> CU traffic per output, and the wider wave tile
// reads fewer LDS fragments per MFMA.
template <int TT>
struct Tile {
    static constexpr int WR = TT / 32;              // waves along the rows (32 rows each)
    static constexpr int WC = 2;                    // waves along the columns
    static constexpr int WJ = TT / (16 * WC);       // 16-column MFMA tiles per wave
    static constexpr int THREADS = WR * WC * 64;
    static constexpr int PLANE = TT * kKb;          // bf16 per LDS operand tile
    static constexpr int TP = TT + 1;               // padded row of the f32 epilogue tiles
};
constexpr int kBigTileMin = 2048;       // both axes at least this long: TT = 128
// Row pitch of every bf16 plane: K plus 64 (128 bytes), so that the 16 rows one
// fragment load touches do not all fall into the same few L2 channels (a
// power-of-two pitch maps them onto a quarter of the channels or fewer).
constexpr int kPad = 64;
__host__ __device__ constexpr int pitch_of(int k) { return k + kPad; }

// Four planes of one split map: [half][hi/lo], each `rows` x `ld` bf16, K
// contiguous (K < ld: the last kPad of each row are padding).
struct Split {
    __bf16 *base;
    int rows, ld;
};
__host__ __device__ inline __bf16 *plane(const Split &s, int half, int hl) {
    return s.base + (size_t)(2 * half + hl) * s.rows * s.ld;
}

struct GemmArgs {
    // operator of transform z: its even-half hi plane; the even lo, odd hi and
    // odd lo planes follow at H * ldo apart (the layout of both tables)
    const __bf16 *op[3];
    int H;                    // half length: the K of every product
    int ldo;                  // row pitch of the operator planes
    Split in[3];              // data operand per transform z
    Split out[3];             // FWD_ROW / FWD_COL / INV_COL: split result per z
    float *dst[3];            // INV_ROW result per z; FWD_COL with spectral == 0: natural a
    int dst_stride[3];        // INV_ROW: element stride (2 for a component of the field map)
    int m, n;
    int sine_mask;            // bit z: transform z's operator is a sine (its unfold sign)
    int spectral;             // FWD_COL: write the spectrum b = a.inv (else natural a into dst[0])
    float scale, two_pi_over_m, two_pi_over_n, wv_div;
    float *auv;               // FWD_COL: a in natural order (nullable)
    const float *rho;         // INV_ROW: density for the energy of z = 2
    double *energy_part;
};

// A kernel-argument value made opaque to the optimizer (kept in scalar
// registers): a select between such values cannot be folded back into one load
// at a computed offset, which would make the argument block addressable and
// copy it to scratch memory in every lane.
template <class T>
__device__ __forceinline__ T opaque(T v) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "one or two scalar registers");
    if constexpr (sizeof(T) == 8) {
        unsigned long long u;
        __builtin_memcpy(&u, &v, 8);
        asm volatile("" : "+s"(u));
        __builtin_memcpy(&v, &u, 8);
    } else {
        unsigned u;
        __builtin_memcpy(&u, &v, 4);
        asm volatile("" : "+s"(u));
        __builtin_memcpy(&v, &u, 4);
    }
    return v;
}
__device__ __forceinline__ Split opaque(const Split &v) {
    return Split{opaque(v.base), opaque(v.rows), opaque(v.ld)};
}

// element z (wave-uniform) of a kernel-argument array
template <class T>
__device__ __forceinline__ T pick3(const T (&v)[3], int z) {
    const T v0 = opaque(v[0]), v1 = opaque(v[1]), v2 = opaque(v[2]);
    return z == 0 ? v0 : (z == 1 ? v1 : v2);
}

__device__ __forceinline__ const __bf16 *op_plane(const GemmArgs &a, int z, int half, int hl) {
    return pick3(a.op, z) + (size_t)(2 * half + hl) * a.H * a.ldo;
}

__device__ __forceinline__ void split8(const float *v, bf16x8 &hi, bf16x8 &lo) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const __bf16 h = (__bf16)v[j];
        hi[j] = h;
        lo[j] = (__bf16)(v[j] - (float)h);
    }
}

// LDS tile of 64 rows x 32 K bf16: 16-byte chunk c of row r at chunk
// c ^ ((r >> 1) & 3), so the eight lanes of a ds_read_b128 phase (eight rows,
// one chunk) hit distinct banks.
__device__ __forceinline__ int swz(int row, int chunk) {
    return row * kKb + 8 * (chunk ^ ((row >> 1) & 3));
}

// XCD-aware tile order: consecutive workgroups go round-robin over the 8 XCDs;
// give each XCD a contiguous run of tiles (tiles of one row band share data).
__device__ __forceinline__ int tile_of_block(int b, int nb, int nct) {
    if (nb & 7) return b;
    const int per = nb >> 3, nrt = nb / nct;
    int bc = 1;
    while (bc * bc < per) bc <<= 1;           // 2^ceil(p/2) for per = 2^p
    bc = bc > nct ? nct : bc;
    int br = per / bc;
    if (br > nrt) { br = nrt; bc = per / nrt; }
    const int x = b & 7, l = b >> 3;
    const int bpr = nct / bc;                 // XCD blocks per block row
    return ((x / bpr) * br + l / bc) * nct + (x % bpr) * bc + l % bc;
}



template <int MODE, int TT>
struct GemmShape {
    static constexpr bool BOTH = MODE == INV_COL || MODE == INV_ROW;   // both product halves per workgroup
    static constexpr int NH = BOTH ? 2 : 1;
    static constexpr int NPL = 4 * NH;      // tiles per stage: (A, B) x half x (hi, lo)
    // a thread stages one 16-byte chunk of row THREADS/4 of each plane
    static_assert(Tile<TT>::THREADS == 4 * TT, "one thread per row chunk of a tile");
    static constexpr int NPLT = NPL;            // planes a thread stages
    // stages of global loads in flight (TT = 128 with both halves: two, so the
    // staging registers fit beside the 32 x 64 accumulators of a wave)
    static constexpr int DEPTH = BOTH ? (TT == 64 ? 3 : 2) : 4;
    static constexpr int KLOOP_BYTES = 2 * NPL * Tile<TT>::PLANE * 2;
    static constexpr int EPI_BYTES = (MODE == INV_COL ? 2 : 1) * TT * Tile<TT>::TP * 4;
    static constexpr int LDS_BYTES = KLOOP_BYTES > EPI_BYTES ? KLOOP_BYTES : EPI_BYTES;
};

// One stage's tiles: global -> registers, registers -> LDS (16 bytes per tile per thread).
template <int NPLT>
__device__ __forceinline__ void load_set(u32x4 (&st)[NPLT], const __bf16 *const (&src)[NPLT], int koff) {
#pragma unroll
    for (int q = 0; q < NPLT; ++q) st[q] = *reinterpret_cast<const u32x4 *>(src[q] + koff);
}
template <int NPLT, int PLANE>
__device__ __forceinline__ void store_set(const u32x4 (&st)[NPLT], __bf16 *buf, int soff) {
#pragma unroll
    for (int q = 0; q < NPLT; ++q) *reinterpret_cast<u32x4 *>(buf + q * PLANE + soff) = st[q];
}

// The products of one LDS stage: per half, 2 x WJ MFMA tiles of 16 x 16 per wave,
// each hi*hi + hi*lo + lo*hi.
template <int MODE, int TT>
__device__ __forceinline__ void compute_stage(const __bf16 *buf,
                                              f32x4 (&acc)[GemmShape<MODE, TT>::NH][2][Tile<TT>::WJ],
                                              int wr, int wc, int lane) {
    constexpr int NH = GemmShape<MODE, TT>::NH, PL = Tile<TT>::PLANE, WJ = Tile<TT>::WJ;
    const int fr = lane & 15, fc = lane >> 4;
#pragma unroll
    for (int h = 0; h < NH; ++h) {
        const __bf16 *A = buf + (2 * h) * PL;
        const __bf16 *B = buf + (2 * NH + 2 * h) * PL;
        bf16x8 ah[2], al[2], bh[WJ], bl[WJ];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int o = swz(wr * 32 + i * 16 + fr, fc);
            ah[i] = *reinterpret_cast<const bf16x8 *>(A + o);
            al[i] = *reinterpret_cast<const bf16x8 *>(A + PL + o);
        }
#pragma unroll
        for (int j = 0; j < WJ; ++j) {
            const int o = swz(wc * 16 * WJ + j * 16 + fr, fc);
            bh[j] = *reinterpret_cast<const bf16x8 *>(B + o);
            bl[j] = *reinterpret_cast<const bf16x8 *>(B + PL + o);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < WJ; ++j) {
                acc[h][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[h][i][j], 0, 0, 0);
                acc[h][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[h][i][j], 0, 0, 0);
                acc[h][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[h][i][j], 0, 0, 0);
            }
    }
}

// The register sets of the K loop, each a separately named array (a set index
// is resolved at compile time, so no set needs addressable storage).
template <int NPL>
struct Sets {
    u32x4 s0[NPL], s1[NPL], s2[NPL], s3[NPL];
    template <int I>
    __device__ __forceinline__ u32x4 (&get())[NPL] {
        if constexpr (I == 0) return s0;
        else if constexpr (I == 1) return s1;
        else if constexpr (I == 2) return s2;
        else return s3;
    }
};

// Stage T of the K loop (a compile-time recursion): multiply the LDS stage T,
// store stage T+1 from its register set into the other LDS stage, refill that
// set with stage T+1+DEPTH.
template <int MODE, int NK, int TT, int T>
__device__ __forceinline__ void k_stages(Sets<GemmShape<MODE, TT>::NPLT> &st,
                                         const __bf16 *const (&src)[GemmShape<MODE, TT>::NPLT], __bf16 *lds,
                                         int soff, f32x4 (&acc)[GemmShape<MODE, TT>::NH][2][Tile<TT>::WJ], int wr,
                                         int wc, int lane) {
    if constexpr (T < NK) {
        using G = GemmShape<MODE, TT>;
        constexpr int NPL = G::NPL, NPLT = G::NPLT, DEPTH = G::DEPTH, PL = Tile<TT>::PLANE;
        compute_stage<MODE, TT>(lds + (T & 1) * NPL * PL, acc, wr, wc, lane);
        if constexpr (T + 1 < NK) {
            store_set<NPLT, PL>(st.template get<(T + 1) % DEPTH>(), lds + ((T + 1) & 1) * NPL * PL, soff);
            if constexpr (T + 1 + DEPTH < NK)
                load_set<NPLT>(st.template get<(T + 1) % DEPTH>(), src, (T + 1 + DEPTH) * kKb);
        }
        __syncthreads();
        k_stages<MODE, NK, TT, T + 1>(st, src, lds, soff, acc, wr, wc, lane);
    }
}

template <int MODE, int NK, int TT, int D>
__device__ __forceinline__ void k_prologue(Sets<GemmShape<MODE, TT>::NPLT> &st,
                                           const __bf16 *const (&src)[GemmShape<MODE, TT>::NPLT]) {
    if constexpr (D < GemmShape<MODE, TT>::DEPTH && D < NK) {
        load_set<GemmShape<MODE, TT>::NPLT>(st.template get<D>(), src, D * kKb);
        k_prologue<MODE, NK, TT, D + 1>(st, src);
    }
}

// Workgroup sum of one double per thread in a fixed order (256 or 512 threads).
template <int NT>
__device__ __forceinline__ double wg_sum(double v, double *red) {
    if constexpr (NT == kBlock) {
        return block_sum_tree(v, red);
    } else {
        static_assert(NT == 2 * kBlock, "256 or 512 threads");
        red[threadIdx.x] = v;
        __syncthreads();
        const double x = threadIdx.x < kBlock ? red[threadIdx.x] + red[threadIdx.x + kBlock] : 0.0;
        __syncthreads();
        return block_sum_tree(x, red);   // threads >= kBlock contribute 0
    }
}

// NK: K stages (H / 32); TT: output tile (64 or 128).
template <int MODE, int NK, int TT>
__global__ void __launch_bounds__(Tile<TT>::THREADS) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_gemm(GemmArgs a) {
    using G = GemmShape<MODE, TT>;
    constexpr bool BOTH = G::BOTH;
    constexpr int NH = G::NH, NPLT = G::NPLT, DEPTH = G::DEPTH;
    constexpr int WC = Tile<TT>::WC, WJ = Tile<TT>::WJ, TP = Tile<TT>::TP, H2 = TT / 2;
    constexpr int NT = Tile<TT>::THREADS;
    __shared__ __attribute__((aligned(16))) unsigned char lds_raw[G::LDS_BYTES];
    __bf16 *lds = reinterpret_cast<__bf16 *>(lds_raw);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave / WC, wc = wave % WC;                    // 32 x 16 WJ outputs per wave
    const int z = BOTH ? (int)blockIdx.y : 0;
    const int hsel = BOTH ? 0 : (int)blockIdx.y;                 // single-half passes: which half
    const int m = a.m, n = a.n;
    const bool sine = (a.sine_mask >> z) & 1;

    const int out_cols = (MODE == FWD_ROW || MODE == INV_ROW) ? n / 2 : n;
    const int nct = out_cols / TT;
    const int tile = tile_of_block(blockIdx.x, gridDim.x, nct);
    const int rt = tile / nct, ct = tile % nct;
    const int r0 = rt * TT, c0 = ct * TT;

    // ---- operand rows this thread stages: row `lr` of each TT-row tile, chunk `lc`
    const int lr = tid >> 2, lc = tid & 3;
    // FWD_ROW: the workgroup's rows are a band and its mirror image
    const int arow = MODE == FWD_ROW ? (lr < H2 ? rt * H2 + lr : m - rt * H2 - TT + lr) : r0 + lr;
    const int brow = c0 + lr;
    const __bf16 *src[NPLT];
    const Split din = pick3(a.in, z);
#pragma unroll
    for (int h = 0; h < NH; ++h) {
        const int half = BOTH ? h : hsel;
#pragma unroll
        for (int hl = 0; hl < 2; ++hl) {
            const Split &d = din;
            const __bf16 *pa, *pb;
            if (MODE == FWD_ROW || MODE == INV_ROW) {            // A = data rows, B = operator
                pa = plane(d, half, hl) + (size_t)arow * d.ld;
                pb = op_plane(a, z, half, hl) + (size_t)brow * a.ldo;
            } else {                                             // A = operator, B = data columns
                pa = op_plane(a, z, half, hl) + (size_t)arow * a.ldo;
                pb = plane(d, half, hl) + (size_t)brow * d.ld;
            }
            src[2 * h + hl] = pa + 8 * lc;
            src[2 * NH + 2 * h + hl] = pb + 8 * lc;
        }
    }
    const int soff = swz(lr, lc);

    f32x4 acc[NH][2][WJ];
#pragma unroll
    for (int h = 0; h < NH; ++h)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < WJ; ++j) acc[h][i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    // K loop over a double-buffered LDS stage, one barrier per stage; the global
    // loads of stage t go to register set t % DEPTH and stay in flight for DEPTH
    // stages before they are stored to LDS
    Sets<NPLT> st;
    k_prologue<MODE, NK, TT, 0>(st, src);
    store_set<NPLT, Tile<TT>::PLANE>(st.template get<0>(), lds, soff);
    if constexpr (DEPTH < NK) load_set<NPLT>(st.template get<0>(), src, DEPTH * kKb);
    __syncthreads();
    k_stages<MODE, NK, TT, 0>(st, src, lds, soff, acc, wr, wc, lane);

    // ------------------------------------------------------------- epilogues
    // lane holds C[row = wr*32 + i*16 + 4*(lane/16) + j][col = wc*16*WJ + jj*16 + lane%16]
    float *T = reinterpret_cast<float *>(lds);
    auto to_lds = [&](float *dstT, int h) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int jj = 0; jj < WJ; ++jj)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    dstT[(wr * 32 + i * 16 + 4 * (lane >> 4) + j) * TP + wc * 16 * WJ + jj * 16 + (lane & 15)] =
                        acc[h][i][jj][j];
    };

    if (MODE == FWD_ROW) {
        // fold the band with its mirror (rows l and TT-1-l) for the column pass:
        // its operand row = the row-transform column hsel * n/2 + c, K = the row
        to_lds(T, 0);
        __syncthreads();
        const int c = tid % TT;
        const Split &O = a.out[0];
        for (int lb = 8 * (tid / TT); lb < H2; lb += 8 * (NT / TT)) {
            float e[8], o[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const float v1 = T[(lb + q) * TP + c], v2 = T[(TT - 1 - lb - q) * TP + c];
                e[q] = v1 + v2;
                o[q] = v1 - v2;
            }
            const size_t at = (size_t)(hsel * (n / 2) + c0 + c) * O.ld + rt * H2 + lb;
            bf16x8 hi, lo;
            split8(e, hi, lo);
            *reinterpret_cast<bf16x8 *>(plane(O, 0, 0) + at) = hi;
            *reinterpret_cast<bf16x8 *>(plane(O, 0, 1) + at) = lo;
            split8(o, hi, lo);
            *reinterpret_cast<bf16x8 *>(plane(O, 1, 0) + at) = hi;
            *reinterpret_cast<bf16x8 *>(plane(O, 1, 1) + at) = lo;
        }
    } else if (MODE == FWD_COL) {
        // rows k' (x frequency u = 2k' + half), columns c (de-interleaved y
        // frequency): the spectrum b = a.inv, split by the parity of u, K = k'
        to_lds(T, 0);
        __syncthreads();
        const int c = tid % TT;
        const int cg = c0 + c;
        const int v = cg < n / 2 ? 2 * cg : 2 * (cg - n / 2) + 1;
        for (int kb = 16 * (tid / TT); kb < TT; kb += 16 * (NT / TT)) {
        float sp[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int u = 2 * (r0 + kb + q) + hsel;
            const float av = a.scale * T[(kb + q) * TP + c];
            if (a.auv) a.auv[(size_t)u * n + v] = av;
            if (!a.spectral) {
                a.dst[0][(size_t)u * n + v] = av;
                continue;
            }
            // frequencies.rs:22-57 in f32 (the inverse operators carry wu / 2,
            // wv / 2 and eps)
            const float wu = (float)u * a.two_pi_over_m;
            const float wv = (float)v * a.two_pi_over_n * a.wv_div;
            const float den = wu * wu + wv * wv;
            const float inv = (u == 0 && v == 0) ? 0.f : 1.f / den;
            sp[q] = av * inv;
        }
        if (a.spectral) {
            const Split &O = a.out[0];
            const size_t at = (size_t)cg * O.ld + r0 + kb;
#pragma unroll
            for (int g = 0; g < 2; ++g) {
                bf16x8 hi, lo;
                split8(&sp[8 * g], hi, lo);
                *reinterpret_cast<bf16x8 *>(plane(O, hsel, 0) + at + 8 * g) = hi;
                *reinterpret_cast<bf16x8 *>(plane(O, hsel, 1) + at + 8 * g) = lo;
            }
        }
        }
    } else if (MODE == INV_COL) {
        // unfold along x (rows i and m-1-i); the row pass reads rows, split by
        // the parity of the y frequency (column half), K = column within it
        float *TE = T, *TO = T + TT * TP;
        to_lds(TE, 0);
        to_lds(TO, NH - 1);
        __syncthreads();
        const int cb = 16 * (tid % (TT / 16));
        const int cg = c0 + cb;
        const int q = cg < n / 2 ? 0 : 1, k = cg - q * (n / 2);
        const Split O = pick3(a.out, z);
        for (int l = tid / (TT / 16); l < TT; l += NT / (TT / 16)) {
        const int i = r0 + l, mi = m - 1 - i;
        float y0[16], y1[16];
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const float e = TE[l * TP + cb + t], o = TO[l * TP + cb + t];
            y0[t] = e + o;
            y1[t] = sine ? o - e : e - o;
        }
#pragma unroll
        for (int g = 0; g < 2; ++g) {
            bf16x8 hi, lo;
            split8(&y0[8 * g], hi, lo);
            const size_t a0 = (size_t)i * O.ld + k + 8 * g, a1 = (size_t)mi * O.ld + k + 8 * g;
            *reinterpret_cast<bf16x8 *>(plane(O, q, 0) + a0) = hi;
            *reinterpret_cast<bf16x8 *>(plane(O, q, 1) + a0) = lo;
            split8(&y1[8 * g], hi, lo);
            *reinterpret_cast<bf16x8 *>(plane(O, q, 0) + a1) = hi;
            *reinterpret_cast<bf16x8 *>(plane(O, q, 1) + a1) = lo;
        }
        }
    } else {
        // INV_ROW: unfold along y (columns i and n-1-i), x4; energy of phi
        double energy = 0.0;
        const int st = pick3(a.dst_stride, z);
        float *d = pick3(a.dst, z);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int jj = 0; jj < WJ; ++jj)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int row = r0 + wr * 32 + i * 16 + 4 * (lane >> 4) + j;
                    const int col = c0 + wc * 16 * WJ + jj * 16 + (lane & 15);
                    const float pe = acc[0][i][jj][j], po = acc[NH - 1][i][jj][j];
                    const float x0 = 4.f * (pe + po);
                    const float x1 = 4.f * (sine ? po - pe : pe - po);
                    if (st == 2) {
                        d[2 * fxy_index(row, col, n)] = x0;
                        d[2 * fxy_index(row, n - 1 - col, n)] = x1;
                    } else {
                        d[(size_t)row * n + col] = x0;
                        d[(size_t)row * n + n - 1 - col] = x1;
                    }
                    if (a.energy_part && z == 2) {
                        const float *rr = a.rho + (size_t)row * n;
                        energy += (double)(x0 * rr[col]) + (double)(x1 * rr[n - 1 - col]);
                    }
                }
        if (a.energy_part && z == 2) {
            const double sum = wg_sum<NT>(energy, reinterpret_cast<double *>(lds));
            if (tid == 0) a.energy_part[blockIdx.x] = sum;
        }
    }
}

// Density from the fixed-point accumulators (epotential.rs:271-285) of row r:
// rho, the row's overflow partial, movable accumulators cleared; then the row
// pass operand: rho folded along y (k with n-1-k), split.  One 256-thread
// workgroup per row; R holds the row (n floats of LDS).
__device__ __forceinline__ void finalize_row(const PotentialBufs &b, const Split &rows, int r, float *R) {
    const int n = b.n;
    const size_t base = (size_t)r * n;
    const longlong2 *__restrict__ qf = reinterpret_cast<const longlong2 *>(b.q_fix + base);
    longlong2 *__restrict__ qm = reinterpret_cast<longlong2 *>(b.q_mov + base);
    float2 *__restrict__ rho = reinterpret_cast<float2 *>(b.rho + base);
    int64_t ovf = 0;
    for (int k = threadIdx.x; k < n / 2; k += kBlock) {
        const longlong2 a = qm[k], f = qf[k];
        if (!b.keep_q) qm[k] = make_longlong2(0, 0);
        const int64_t t0 = a.x + f.x, t1 = a.y + f.y;
        ovf += (t0 > b.coef_q ? t0 - b.coef_q : 0) + (t1 > b.coef_q ? t1 - b.coef_q : 0);
        const float2 v = make_float2((float)((double)t0 * b.inv_scale) * b.inv_bin_area,
                                     (float)((double)t1 * b.inv_scale) * b.inv_bin_area);
        rho[k] = v;
        R[2 * k] = v.x;
        R[2 * k + 1] = v.y;
    }
    __shared__ int64_t red[kBlock / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) ovf += __shfl_xor(ovf, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ovf;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t v = 0;
        for (int w = 0; w < kBlock / 64; ++w) v += red[w];
        b.ovf_part[r] = v;
    }
    // fold: n/2 columns in runs of 8
    for (int t = threadIdx.x; t < n / 16; t += kBlock) {
        const int k = 8 * t;
        float e[8], o[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const float x0 = R[k + q], x1 = R[n - 1 - k - q];
            e[q] = x0 + x1;
            o[q] = x0 - x1;
        }
        const size_t at = (size_t)r * rows.ld + k;
        bf16x8 hi, lo;
        split8(e, hi, lo);
        *reinterpret_cast<bf16x8 *>(plane(rows, 0, 0) + at) = hi;
        *reinterpret_cast<bf16x8 *>(plane(rows, 0, 1) + at) = lo;
        split8(o, hi, lo);
        *reinterpret_cast<bf16x8 *>(plane(rows, 1, 0) + at) = hi;
        *reinterpret_cast<bf16x8 *>(plane(rows, 1, 1) + at) = lo;
    }
}

__global__ void __launch_bounds__(kBlock) k_finalize(PotentialBufs b, Split rows) {
    extern __shared__ __attribute__((aligned(16))) float R[];
    finalize_row(b, rows, (int)blockIdx.x, R);
}

// The finalize and the WA net groups up to degree 256 in one launch:
// workgroups [0, m) finalize a row each, the rest loop over the WA plan (the
// same work and partial per plan workgroup as in its own launch).  The two are
// independent (WA reads the positions, the finalize the accumulators), so they
// overlap with no second stream and no events.
__global__ void __launch_bounds__(kBlock) k_finalize_wa(const JobDesc *__restrict__ jd, Split rows, WaShare wa) {
    extern __shared__ __attribute__((aligned(16))) float R[];
    const PotentialBufs &b = jd->pb;
    if ((int)blockIdx.x < b.m)
        finalize_row(b, rows, (int)blockIdx.x, R);
    else
        wa_narrow_blocks(jd->wa, wa.plan, wa.first + (blockIdx.x - b.m), gridDim.x - b.m, wa.last);
}

// Leaf forward transform: a natural m x n map -> the row pass operand.
__global__ void k_prep_rows(const float *__restrict__ x, Split rows, int m, int n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int runs = n / 16;
    if (t >= (int64_t)m * runs) return;
    const int r = (int)(t / runs), k = 8 * (int)(t % runs);
    const float *xr = x + (size_t)r * n;
    float e[8], o[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const float x0 = xr[k + q], x1 = xr[n - 1 - k - q];
        e[q] = x0 + x1;
        o[q] = x0 - x1;
    }
    const size_t at = (size_t)r * rows.ld + k;
    bf16x8 hi, lo;
    split8(e, hi, lo);
    *reinterpret_cast<bf16x8 *>(plane(rows, 0, 0) + at) = hi;
    *reinterpret_cast<bf16x8 *>(plane(rows, 0, 1) + at) = lo;
    split8(o, hi, lo);
    *reinterpret_cast<bf16x8 *>(plane(rows, 1, 0) + at) = hi;
    *reinterpret_cast<bf16x8 *>(plane(rows, 1, 1) + at) = lo;
}

// Leaf inverse transforms: a natural-order spectrum X (with the eps factors of
// the inverse) -> the column pass operand: column c' = de-interleaved y
// frequency, split by the parity of the x frequency u, K = u / 2.
__global__ void k_prep_spectrum(const float *__restrict__ x, Split cols, int m, int n, int eps_u,
                                int eps_v) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int runs = m / 16;                   // runs of 8 k per (column, half)
    if (t >= (int64_t)n * 2 * runs) return;
    const int c = (int)(t / (2 * runs)), rem = (int)(t % (2 * runs));
    const int q = rem / runs, k0 = 8 * (rem % runs);
    const int v = c < n / 2 ? 2 * c : 2 * (c - n / 2) + 1;
    float f[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int u = 2 * (k0 + j) + q;
        float val = x[(size_t)u * n + v];
        if (eps_u && u == 0) val *= 0.5f;
        if (eps_v && v == 0) val *= 0.5f;
        f[j] = val;
    }
    const size_t at = (size_t)c * cols.ld + k0;
    bf16x8 hi, lo;
    split8(f, hi, lo);
    *reinterpret_cast<bf16x8 *>(plane(cols, q, 0) + at) = hi;
    *reinterpret_cast<bf16x8 *>(plane(cols, q, 1) + at) = lo;
}

// Operator planes of length L in FP64, split to bf16 pairs (row pitch pitch_of(H)).
__global__ void k_make_planes(__bf16 *tab, int L) {
    const int H = L / 2, ld = pitch_of(H);
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
    tab[(size_t)(2 * t) * H * ld + (size_t)r * ld + c] = hi;
    tab[(size_t)(2 * t + 1) * H * ld + (size_t)r * ld + c] = lo;
}

// The scaled inverse operators of length L (W_CE, W_CO, W_SE, W_SO) in FP64,
// split to bf16 pairs: [i][k'] = C[i][k] eps_k (eps_0 = 1/2, else 1) and
// S[i][k] k w1 / 2, k = 2k' + p, where k w1 is the axis' frequency operator
// w_k (frequencies.rs:22-57: 2 pi k / L, times bw / bh along y).
__global__ void k_make_inv_planes(__bf16 *tab, int L, double w1) {
    const int H = L / 2, ld = pitch_of(H);
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)W_COUNT * H * H) return;
    const int t = (int)(idx / ((int64_t)H * H));
    const int rem = (int)(idx % ((int64_t)H * H));
    const int i = rem / H, k = 2 * (rem % H) + (t & 1);
    const int64_t num = ((int64_t)(2 * i + 1) * k) % (4 * (int64_t)L);
    double s, co;
    sincospi((double)num / (double)(2 * L), &s, &co);
    const double v = t < W_SE ? (k ? co : 0.5 * co) : s * (0.5 * w1 * (double)k);
    const __bf16 hi = (__bf16)(float)v;
    const __bf16 lo = (__bf16)(float)(v - (double)(float)hi);
    tab[(size_t)(2 * t) * H * ld + (size_t)rem / H * ld + rem % H] = hi;
    tab[(size_t)(2 * t + 1) * H * ld + (size_t)rem / H * ld + rem % H] = lo;
}

inline unsigned cdiv(int64_t a, int b) { return (unsigned)((a + b - 1) / b); }

}  // namespace

// Split maps between the passes: the row pass operand (m rows, K = n/2), the
// column pass operand and the spectrum (n rows, K = m/2), U and V (m rows,
// K = n/2).
struct MfmaPlan {
    int m = 0, n = 0;
    DevBuf tab_m, tab_n;          // operator planes (tab_n unused when m == n)
    DevBuf inv_m, inv_n;          // scaled inverse operators along x and y (inv_n unused
                                  // when m == n and bw == bh: the same operators)
    DevBuf row, col, x, y[2];
    const __bf16 *tm() const { return tab_m.as<__bf16>(); }
    const __bf16 *tn() const { return m == n ? tab_m.as<__bf16>() : tab_n.as<__bf16>(); }
    const __bf16 *im() const { return inv_m.as<__bf16>(); }
    const __bf16 *in() const { return inv_n.p ? inv_n.as<__bf16>() : inv_m.as<__bf16>(); }
};

namespace {
size_t table_bytes(int L) { return sizeof(__bf16) * 2 * T_COUNT * (size_t)(L / 2) * pitch_of(L / 2); }
size_t inv_table_bytes(int L) { return sizeof(__bf16) * 2 * W_COUNT * (size_t)(L / 2) * pitch_of(L / 2); }
size_t split_bytes(int rows, int k) { return sizeof(__bf16) * 4 * (size_t)rows * pitch_of(k); }
// the hi plane of the even half of operator t of a table of length L
const __bf16 *op_of(const __bf16 *tab, int L, int t) {
    return tab + (size_t)(2 * t) * (L / 2) * pitch_of(L / 2);
}
}  // namespace

size_t mfma_plan_estimate(int m, int n) {
    return table_bytes(m) + (m == n ? 0 : table_bytes(n)) + inv_table_bytes(m) + inv_table_bytes(n) +
           3 * split_bytes(m, n / 2) + 2 * split_bytes(n, m / 2);
}

bool mfma_supported(int m, int n) {
    return m >= 128 && n >= 128 && m % 128 == 0 && n % 128 == 0;
}

namespace {
// the output tile of the GEMM passes of an m x n grid (powers of two)
int mfma_tile(int m, int n) { return m >= kBigTileMin && n >= kBigTileMin ? 128 : 64; }
}  // namespace

// the last inverse pass writes one energy partial per tile of the potential
int mfma_energy_partials(int m, int n) {
    const int t = mfma_tile(m, n);
    return (m / t) * ((n / 2) / t);
}

// k_finalize runs one workgroup per row and writes one overflow partial each.
int mfma_overflow_partials(int m) { return m; }

int mfma_plan_create(MfmaPlan **out, int m, int n, float wv_div, hipStream_t s) {
    if (!mfma_supported(m, n))
        return fail(ERP_ERR_UNSUPPORTED, "the MFMA spectral form needs at least 128 bins per axis");
    MfmaPlan *p = new MfmaPlan();
    p->m = m;
    p->n = n;
    int rc = ERP_OK;
    auto make = [&](DevBuf &b, int L) {
        const int64_t cnt = (int64_t)T_COUNT * (L / 2) * (L / 2);
        if ((rc = b.alloc(table_bytes(L)))) return;
        hipLaunchKernelGGL(k_make_planes, dim3(cdiv(cnt, 256)), dim3(256), 0, s, b.as<__bf16>(), L);
        if (hipGetLastError() != hipSuccess) rc = fail(ERP_ERR_HIP, "k_make_planes launch failed");
    };
    auto make_inv = [&](DevBuf &b, int L, double w1) {
        const int64_t cnt = (int64_t)W_COUNT * (L / 2) * (L / 2);
        if ((rc = b.alloc(inv_table_bytes(L)))) return;
        hipLaunchKernelGGL(k_make_inv_planes, dim3(cdiv(cnt, 256)), dim3(256), 0, s, b.as<__bf16>(), L, w1);
        if (hipGetLastError() != hipSuccess) rc = fail(ERP_ERR_HIP, "k_make_inv_planes launch failed");
    };
    make(p->tab_m, m);
    if (!rc && n != m) make(p->tab_n, n);
    if (!rc) make_inv(p->inv_m, m, 2.0 * M_PI / (double)m);
    if (!rc && (n != m || wv_div != 1.0f)) make_inv(p->inv_n, n, 2.0 * M_PI / (double)n * (double)wv_div);
    for (DevBuf *b : {&p->row, &p->y[0], &p->y[1]})
        if (!rc) rc = b->alloc(split_bytes(m, n / 2));
    for (DevBuf *b : {&p->col, &p->x})
        if (!rc) rc = b->alloc(split_bytes(n, m / 2));
    if (rc) {
        delete p;
        return rc;
    }
    *out = p;
    return ERP_OK;
}

void mfma_plan_destroy(MfmaPlan *plan) { delete plan; }

size_t mfma_plan_bytes(const MfmaPlan *p) {
    if (!p) return 0;
    return p->tab_m.bytes + p->tab_n.bytes + p->inv_m.bytes + p->inv_n.bytes + p->row.bytes + p->col.bytes +
           p->x.bytes + p->y[0].bytes + p->y[1].bytes;
}

namespace {

// Split planes of `rows` rows and K = k over a plan buffer (row pitch pitch_of(k)).
Split split_of(const DevBuf &buf, int rows, int k) {
    return Split{reinterpret_cast<__bf16 *>(buf.p), rows, pitch_of(k)};
}

template <int MODE>
int launch_mode(const GemmArgs &a, int nz, hipStream_t s) {
    const int rows = (MODE == FWD_ROW || MODE == INV_ROW) ? a.m : a.m / 2;
    const int cols = (MODE == FWD_ROW || MODE == INV_ROW) ? a.n / 2 : a.n;
    const int tt = mfma_tile(a.m, a.n);
    const dim3 grid((unsigned)((rows / tt) * (cols / tt)),
                    (MODE == FWD_ROW || MODE == FWD_COL) ? 2u : (unsigned)nz);
    if (tt == 128) {       // both lengths 2048 (K = 1024)
        if (a.H / kKb != 32) return fail(ERP_ERR_UNSUPPORTED, "MFMA 128-bin tiles need lengths of 2048");
        hipLaunchKernelGGL((k_gemm<MODE, 32, 128>), grid, dim3(Tile<128>::THREADS), 0, s, a);
        ERP_HIP_TRY(hipGetLastError());
        return ERP_OK;
    }
    switch (a.H / kKb) {   // H = L/2 with L a power of two in [128, 2048]
    case 2: hipLaunchKernelGGL((k_gemm<MODE, 2, 64>), grid, dim3(Tile<64>::THREADS), 0, s, a); break;
    case 4: hipLaunchKernelGGL((k_gemm<MODE, 4, 64>), grid, dim3(Tile<64>::THREADS), 0, s, a); break;
    case 8: hipLaunchKernelGGL((k_gemm<MODE, 8, 64>), grid, dim3(Tile<64>::THREADS), 0, s, a); break;
    case 16: hipLaunchKernelGGL((k_gemm<MODE, 16, 64>), grid, dim3(Tile<64>::THREADS), 0, s, a); break;
    case 32: hipLaunchKernelGGL((k_gemm<MODE, 32, 64>), grid, dim3(Tile<64>::THREADS), 0, s, a); break;
    default: return fail(ERP_ERR_UNSUPPORTED, "MFMA transform length must be a power of two in [128, 2048]");
    }
    ERP_HIP_TRY(hipGetLastError());
    return ERP_OK;
}

// row-pass operand -> column-pass operand -> (spectrum | natural a)
int forward(const MfmaPlan *p, Split rowop, Split colop, GemmArgs c, hipStream_t s, KernelTimer *t) {
    GemmArgs a{};
    a.m = p->m;
    a.n = p->n;
    a.in[0] = rowop;
    a.out[0] = colop;
    a.op[0] = op_of(p->tn(), p->n, T_FE);
    a.H = p->n / 2;
    a.ldo = pitch_of(a.H);
    if (t) t->begin("mfma_fwd_row");
    int rc = launch_mode<FWD_ROW>(a, 1, s);
    if (t) t->end();
    if (rc) return rc;
    c.m = p->m;
    c.n = p->n;
    c.in[0] = colop;
    c.op[0] = op_of(p->tm(), p->m, T_FE);
    c.H = p->m / 2;
    c.ldo = pitch_of(c.H);
    if (t) t->begin("mfma_fwd_col");
    rc = launch_mode<FWD_COL>(c, 1, s);
    if (t) t->end();
    return rc;
}

// One inverse along x (INV_COL: the split spectrum -> split U and V, one
// product of the spectrum each) then along y (INV_ROW: U, V -> natural maps); per transform z its input, operator and
// unfold sign (bit z of the sine masks).
struct InverseSpec {
    int ncol, nrow;                      // transforms of each pass
    Split col_in[3], col_out[3];
    const __bf16 *col_op[3];
    int col_sine;
    Split row_in[3];
    const __bf16 *row_op[3];
    int row_sine;
    float *out[3];
    int out_stride[3];
    const float *rho;                    // energy of row transform 2 (nullable)
    double *energy_part;
};

int inverse(const MfmaPlan *p, const InverseSpec &v, hipStream_t s, KernelTimer *t) {
    GemmArgs a{};
    a.m = p->m;
    a.n = p->n;
    for (int z = 0; z < v.ncol; ++z) {
        a.in[z] = v.col_in[z];
        a.out[z] = v.col_out[z];
        a.op[z] = v.col_op[z];
    }
    a.H = p->m / 2;
    a.ldo = pitch_of(a.H);
    a.sine_mask = v.col_sine;
    if (t) t->begin("mfma_inv_col");
    int rc = launch_mode<INV_COL>(a, v.ncol, s);
    if (t) t->end();
    if (rc) return rc;
    GemmArgs b{};
    b.m = p->m;
    b.n = p->n;
    for (int z = 0; z < v.nrow; ++z) {
        b.in[z] = v.row_in[z];
        b.op[z] = v.row_op[z];
        b.dst[z] = v.out[z];
        b.dst_stride[z] = v.out_stride[z];
    }
    b.H = p->n / 2;
    b.ldo = pitch_of(b.H);
    b.sine_mask = v.row_sine;
    b.rho = v.rho;
    b.energy_part = v.energy_part;
    if (t) t->begin("mfma_inv_row");
    rc = launch_mode<INV_ROW>(b, v.nrow, s);
    if (t) t->end();
    return rc;
}

}  // namespace

// The potential pipeline of one placement (epotential.rs:225-374) in MFMA form.
// With `jd` (the placement's device descriptor) the WA net groups run inside
// the finalize launch (the wide buckets first, on their own).
int launch_potential_mfma(const PotentialBufs &b, MfmaPlan *p, hipStream_t s, KernelTimer *t,
                          const JobDesc *jd, const BatchDims *bd) {
    if (!p || p->m != b.m || p->n != b.n) return fail(ERP_ERR_INVALID, "MFMA plan does not match the grid");
    const int m = b.m, n = b.n;
    const Split rowop = split_of(p->row, m, n / 2);         // [half][hl][row][k < n/2]
    const Split colop = split_of(p->col, n, m / 2);         // [half][hl][column][k < m/2]
    const size_t lds = sizeof(float) * (size_t)n;
    int rc;
    if (jd) {
        if ((rc = launch_wa_wide(jd, *bd, 1, s))) return rc;
        const WaPlan plan = wa_narrow_plan(*bd);
        const WaShare wa{plan, 0u, plan.begin[plan.nseg]};
        if (t) t->begin("mfma_finalize_wa");
        hipLaunchKernelGGL(k_finalize_wa, dim3((unsigned)m + wa.last), dim3(kBlock), lds, s, jd, rowop, wa);
        if (t) t->end();
    } else {
        if (t) t->begin("mfma_finalize");
        hipLaunchKernelGGL(k_finalize, dim3((unsigned)m), dim3(kBlock), lds, s, b, rowop);
        if (t) t->end();
    }
    ERP_HIP_TRY(hipGetLastError());
    GemmArgs c{};
    const Split X = split_of(p->x, n, m / 2);
    c.out[0] = X;
    c.spectral = 1;
    c.scale = (float)(4.0 / ((double)m * (double)n));
    c.two_pi_over_m = (float)(2.0 * M_PI) / (float)m;
    c.two_pi_over_n = (float)(2.0 * M_PI) / (float)n;
    c.wv_div = b.bw / b.bh;
    c.auv = b.auv;
    if ((rc = forward(p, rowop, colop, c, s, t))) return rc;
    // U = [C_M diag eps] b (z = 0), V = [S_M diag wu/2] b (z = 1); then
    // xi_x = 4 V [diag eps C_N^T], xi_y = 4 U [diag wv/2 S_N^T], phi = 4 U [diag eps C_N^T]
    const Split U = split_of(p->y[0], m, n / 2), V = split_of(p->y[1], m, n / 2);
    const __bf16 *cm = op_of(p->im(), m, W_CE), *sm = op_of(p->im(), m, W_SE);
    const __bf16 *cn = op_of(p->in(), n, W_CE), *sn = op_of(p->in(), n, W_SE);
    float *fxy = reinterpret_cast<float *>(b.fxy);
    InverseSpec v{};
    v.ncol = 2;
    v.col_in[0] = v.col_in[1] = X;
    v.col_out[0] = U;
    v.col_out[1] = V;
    v.col_op[0] = cm;
    v.col_op[1] = sm;
    v.col_sine = 2;
    // the potential (third row transform) and its energy only when requested
    v.nrow = b.potential ? 3 : 2;
    v.row_in[0] = V;
    v.row_in[1] = v.row_in[2] = U;
    v.row_op[0] = v.row_op[2] = cn;
    v.row_op[1] = sn;
    v.row_sine = 2;
    v.out[0] = fxy;
    v.out[1] = fxy + 1;
    v.out[2] = b.pot;
    v.out_stride[0] = v.out_stride[1] = 2;
    v.out_stride[2] = 1;
    v.rho = b.rho;
    v.energy_part = b.potential ? b.energy_part : nullptr;
    return inverse(p, v, s, t);
}

// Leaf transforms (host-facing entry points run these on device buffers).
int launch_dct2_mfma(const float *x, float *out, MfmaPlan *p, hipStream_t s) {
    const int m = p->m, n = p->n;
    const Split rowop = split_of(p->row, m, n / 2);
    const Split colop = split_of(p->col, n, m / 2);
    hipLaunchKernelGGL(k_prep_rows, dim3(cdiv((int64_t)m * (n / 16), 256)), dim3(256), 0, s, x, rowop, m, n);
    ERP_HIP_TRY(hipGetLastError());
    GemmArgs c{};
    c.dst[0] = out;
    c.spectral = 0;
    c.scale = (float)(4.0 / ((double)m * (double)n));
    return forward(p, rowop, colop, c, s, nullptr);
}

// which: 0 idct2 (cos, cos), 1 idct_idxst (cos along x, sin along y),
//        2 idxst_idct (sin along x, cos along y)
int launch_inverse_mfma(const float *x, float *out, int which, MfmaPlan *p, hipStream_t s) {
    const int m = p->m, n = p->n;
    const int eps_u = which != 2, eps_v = which != 1;
    const Split X = split_of(p->x, n, m / 2);
    hipLaunchKernelGGL(k_prep_spectrum, dim3(cdiv((int64_t)n * 2 * (m / 16), 256)), dim3(256), 0, s, x,
                       X, m, n, eps_u, eps_v);
    ERP_HIP_TRY(hipGetLastError());
    const bool sine_x = which == 2, sine_y = which == 1;
    InverseSpec v{};
    v.ncol = v.nrow = 1;
    v.col_in[0] = X;
    v.col_out[0] = split_of(p->y[0], m, n / 2);
    v.col_op[0] = op_of(p->tm(), m, sine_x ? T_SE : T_CE);
    v.col_sine = sine_x;
    v.row_in[0] = v.col_out[0];
    v.row_op[0] = op_of(p->tn(), n, sine_y ? T_SE : T_CE);
    v.row_sine = sine_y;
    v.out[0] = out;
    v.out_stride[0] = 1;
    return inverse(p, v, s, nullptr);
}

}  // namespace erp
