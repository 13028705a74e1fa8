// FP16 shifted-storage step and initialisation kernels (collision in FP32).
#include "step_kernels.h"

using namespace lbm;

hipError_t launch_step_f16s(int mode, int sch, int layout, const StepArgs& a, hipStream_t s) {
    if (a.q == 19) return launch_step_t<19, PrecF16S>(mode, sch, layout, a, s);
    return launch_step_t<27, PrecF16S>(mode, sch, layout, a, s);
}

hipError_t launch_init_f16s(int q, bool swap, const InitArgs& a, hipStream_t s) {
    if (q == 19) return launch_init_t<19, PrecF16S>(a, swap, s);
    return launch_init_t<27, PrecF16S>(a, swap, s);
}
