// Spectral Poisson solve as dense transform products on MFMA (gfx950).
#include "erp_internal.h"
#include "erp_timer.h"

namespace erp {

struct MfmaPlan {
    int m = 0, n = 0;
};

int mfma_plan_create(MfmaPlan **plan, int m, int n, hipStream_t) {
    (void)plan; (void)m; (void)n;
    return fail(ERP_ERR_UNSUPPORTED, "MFMA spectral form not built yet");
}

void mfma_plan_destroy(MfmaPlan *plan) { delete plan; }

int launch_potential_mfma(const PotentialBufs &, MfmaPlan *, hipStream_t, KernelTimer *) {
    return fail(ERP_ERR_UNSUPPORTED, "MFMA spectral form not built yet");
}

}  // namespace erp
