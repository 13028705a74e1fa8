// FP32 / FP16 shifted-storage kernels (populations stored as f - w, collision
// in FP32).
#include "step_kernels.h"

using namespace lbm;

hipError_t launch_step_lp(int storage, int mode, int sch, int layout, const StepArgs& a, hipStream_t s) {
    if (storage == LBM_STORE_F32) {
        if (a.q == 19) return launch_step_t<19, PrecF32S>(mode, sch, layout, a, s);
        return launch_step_t<27, PrecF32S>(mode, sch, layout, a, s);
    }
    if (a.q == 19) return launch_step_t<19, PrecF16S>(mode, sch, layout, a, s);
    return launch_step_t<27, PrecF16S>(mode, sch, layout, a, s);
}

hipError_t launch_init_lp(int storage, int q, bool swap, const InitArgs& a, hipStream_t s) {
    if (storage == LBM_STORE_F32) {
        if (q == 19) return launch_init_t<19, PrecF32S>(a, swap, s);
        return launch_init_t<27, PrecF32S>(a, swap, s);
    }
    if (q == 19) return launch_init_t<19, PrecF16S>(a, swap, s);
    return launch_init_t<27, PrecF16S>(a, swap, s);
}
