// Dispatch of the FP32 / FP16 shifted-storage kernels (populations stored as
// f - w, collision in FP32).  The instantiations live in kernels_f32.hip and
// kernels_f16.hip so the two compile in parallel.
#include "lbm_internal.h"

hipError_t launch_step_f32s(int mode, int sch, int layout, const StepArgs& a, hipStream_t s);
hipError_t launch_step_f16s(int mode, int sch, int layout, const StepArgs& a, hipStream_t s);
hipError_t launch_init_f32s(int q, bool swap, const InitArgs& a, hipStream_t s);
hipError_t launch_init_f16s(int q, bool swap, const InitArgs& a, hipStream_t s);

hipError_t launch_step_lp(int storage, int mode, int sch, int layout, const StepArgs& a, hipStream_t s) {
    return storage == LBM_STORE_F32 ? launch_step_f32s(mode, sch, layout, a, s) : launch_step_f16s(mode, sch, layout, a, s);
}

hipError_t launch_init_lp(int storage, int q, bool swap, const InitArgs& a, hipStream_t s) {
    return storage == LBM_STORE_F32 ? launch_init_f32s(q, swap, a, s) : launch_init_f16s(q, swap, a, s);
}
