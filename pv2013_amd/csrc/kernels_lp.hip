// FP32 / FP16 shifted-storage kernels (populations stored as f - w, collision
// in FP32).  Built with FMA contraction allowed.
#include "step_kernels.h"

using namespace lbm;

hipError_t launch_step_lp(int storage, int mode, const StepArgs& a, int block, hipStream_t s) {
    if (storage == LBM_STORE_F32) {
        if (a.q == 19) return launch_step_t<19, PrecF32S>(mode, a, block, s);
        return launch_step_t<27, PrecF32S>(mode, a, block, s);
    }
    if (a.q == 19) return launch_step_t<19, PrecF16S>(mode, a, block, s);
    return launch_step_t<27, PrecF16S>(mode, a, block, s);
}

hipError_t launch_init_lp(int storage, int q, const InitArgs& a, int block, hipStream_t s) {
    (void)block;
    if (storage == LBM_STORE_F32) {
        if (q == 19) return launch_init_t<19, PrecF32S>(a, s);
        return launch_init_t<27, PrecF32S>(a, s);
    }
    if (q == 19) return launch_init_t<19, PrecF16S>(a, s);
    return launch_init_t<27, PrecF16S>(a, s);
}
