// Allocation-free form of the four leaf transforms (erp_spectral_transform_noalloc).
//
// Its scratch -- the spectral tables of both lengths and the intermediate map
// between the column and row passes -- lives in one file-scope device array of
// fixed size, so a call makes no hipMalloc, hipFree, stream or event: it only
// enqueues kernels on the caller's stream.  Each GPU holds its own copy of the
// array until the process ends; its device address is looked up once per
// device (hipGetSymbolAddress) and kept.  Calls that overlap on one GPU share
// the array and race.
#include <mutex>

#include "erp_ctx.h"

// Tables of one length L <= 2048: fft_tw float2[L/2], dct_tw float2[L].
constexpr size_t kTabBytes = sizeof(float2) * (erp::kMaxBins / 2 + erp::kMaxBins);        // 24 KiB
constexpr size_t kMapBytes = sizeof(float) * (size_t)erp::kMaxBins * erp::kMaxBins;       // 16 MiB
constexpr size_t kScratchBytes = 2 * kTabBytes + kMapBytes;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace erp {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char *g_ptr[kMaxDevices] = {};   // device address of `scratch` per device

// The device address of this device's copy of `scratch` (looked up on the
// device's first call, with that device current).
int scratch_ptr(unsigned char **out) {
    int dev = 0;
    ERP_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(ERP_ERR_UNSUPPORTED, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(ERP_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_ptr[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_ptr[dev];
    return ERP_OK;
}

}  // namespace
}  // namespace erp

using namespace erp;

extern "C" {

int erp_spectral_transform_noalloc(const float *x, float *out, int32_t m, int32_t n,
                                   int32_t which, void *stream) {
    if (!x || !out) return fail(ERP_ERR_INVALID, "null argument");
    if (which < 0 || which > 3) return fail(ERP_ERR_INVALID, "unknown transform (0..3)");
    if (m < kMinBins || n < kMinBins || m > kMaxBins || n > kMaxBins || !is_pow2(m) || !is_pow2(n))
        return fail(ERP_ERR_UNSUPPORTED,
                    "sizes must be powers of two in [64, 2048]: the scratch holds one "
                    "2048 x 2048 map and the tables of two lengths up to 2048");
    if (x == out) return fail(ERP_ERR_INVALID, "input and output must not alias");
    unsigned char *base = nullptr;
    int rc = scratch_ptr(&base);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    // layout: [tables of m][tables of n][map]; every part is fully written
    // before it is read, so nothing is cleared
    float2 *tm_fft = reinterpret_cast<float2 *>(base);
    float2 *tm_dct = tm_fft + kMaxBins / 2;
    float2 *tn_fft = reinterpret_cast<float2 *>(base + kTabBytes);
    float2 *tn_dct = tn_fft + kMaxBins / 2;
    float *map = reinterpret_cast<float *>(base + 2 * kTabBytes);
    if ((rc = launch_make_tables(tm_fft, tm_dct, m, s))) return rc;
    if ((rc = launch_make_tables(tn_fft, tn_dct, n, s))) return rc;
    return launch_spectral_leaf(x, out, map, m, n, which, SpecView{tm_fft, tm_dct},
                                SpecView{tn_fft, tn_dct}, s);
}

}  // extern "C"
