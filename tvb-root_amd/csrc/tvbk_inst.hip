// Explicit instantiations of the step, init and dense kernels.  The build
// compiles this file once per (precision, model, coupling) combination,
// selected by TVBK_INST_R (0: float history and state, 1: double history
// and state, 2: float history, double state) and TVBK_INST_COMBO, and once
// with TVBK_INST_DENSE for the dense zero-delay kernels, so that the heavy
// template instantiations compile in parallel.
#include "tvbk_step.hpp"

namespace tvbk {

#if defined(TVBK_INST_DENSE)
template hipError_t launch_dense_t<float, G2D>(const StepArgs&, int, int, size_t, hipStream_t);
template hipError_t launch_dense_t<float, MPR>(const StepArgs&, int, int, size_t, hipStream_t);
template hipError_t launch_dense_t<double, G2D>(const StepArgs&, int, int, size_t, hipStream_t);
template hipError_t launch_dense_t<double, MPR>(const StepArgs&, int, int, size_t, hipStream_t);
#elif defined(TVBK_INST_R) && defined(TVBK_INST_COMBO)
#if TVBK_INST_R == 1
using InstRH = double;
using InstR = double;
#elif TVBK_INST_R == 2
using InstRH = float;
using InstR = double;
#else
using InstRH = float;
using InstR = float;
#endif
#if TVBK_INST_COMBO == 0
#define TVBK_INST_M G2D
#define TVBK_INST_CF CF_LINEAR
#elif TVBK_INST_COMBO == 1
#define TVBK_INST_M MPR
#define TVBK_INST_CF CF_LINEAR
#elif TVBK_INST_COMBO == 2
#define TVBK_INST_M JR
#define TVBK_INST_CF CF_LINEAR
#elif TVBK_INST_COMBO == 3
#define TVBK_INST_M JR
#define TVBK_INST_CF CF_SIGMOID_JR
#elif TVBK_INST_COMBO == 4
#define TVBK_INST_M SUM
#define TVBK_INST_CF CF_IDENTITY
#elif TVBK_INST_COMBO == 5
#define TVBK_INST_M SUM
#define TVBK_INST_CF CF_LINEAR
#elif TVBK_INST_COMBO == 6
#define TVBK_INST_M MPR
#define TVBK_INST_CF CF_LINEAR1
#else
#error "unknown TVBK_INST_COMBO"
#endif
template hipError_t launch_step_t<InstRH, InstR, TVBK_INST_M, TVBK_INST_CF>(const StepArgs&, int, int, size_t,
                                                                          hipStream_t);
template hipError_t launch_init_t<InstRH, InstR, TVBK_INST_M, TVBK_INST_CF>(const InitArgs&, int, hipStream_t);
#else
#error "tvbk_inst.hip needs TVBK_INST_DENSE or TVBK_INST_R and TVBK_INST_COMBO"
#endif

}  // namespace tvbk
