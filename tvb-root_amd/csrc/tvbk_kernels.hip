// Small kernels (history scatter/gather, slot arithmetic check) and the
// dispatch from the plan's runtime (precision, model, coupling) to the
// instantiated step, init and dense kernels of tvbk_inst.hip.
#include "tvbk_kernels.hpp"
#include "tvbk_device.hpp"

namespace tvbk {

template <class RH, class R, class M, int CF>
hipError_t launch_step_t(const StepArgs& a, int ntiles, int nthreads, size_t lds, hipStream_t st);
template <class RH, class R, class M, int CF>
hipError_t launch_init_t(const InitArgs& a, int ntiles, hipStream_t st);
template <class R, class M>
hipError_t launch_dense_t(const StepArgs& a, int ntiles, int nthreads, size_t lds, hipStream_t st);

template <class RH, class R>
__global__ void k_scatter(bool gather, RH* ring, R* st, RH* hist, R* state, int H, int N, int nch, int nsv, int T,
                          int t) {
  const int64_t nh = (int64_t)H * N * nch;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nh + (int64_t)nsv * N;
       q += (int64_t)gridDim.x * blockDim.x) {
    if (q < nh) {
      const int i = (int)(q % N), ch = (int)((q / N) % nch), slot = (int)(q / ((int64_t)N * nch));
      RH* r = ring + (((int64_t)slot * N + i) * nch + ch) * T + t;
      if (gather) hist[q] = *r; else *r = hist[q];
    } else {
      const int64_t u = q - nh;
      const int i = (int)(u % N), v = (int)(u / N);
      R* r = st + ((int64_t)i * nsv + v) * T + t;
      if (gather) state[u] = *r; else *r = state[u];
    }
  }
}

// Slot read by every edge of a connectome at `step`, with the kernels' slot
// arithmetic.
__global__ void k_debug_slots(const uint32_t* edges, int nnz, int H, int64_t step, int32_t* out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  const int smod = (int)((step - 1) % H);
  const int d = (int)(edges[e] >> 16);
  int sl = smod - d;
  sl += (sl < 0) ? H : 0;
  out[e] = sl;
}

// ---------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------
template <class RH, class R>
static hipError_t dispatch(int model, int cf, bool init, const StepArgs* a, const InitArgs* ia, int ntiles,
                           int nthreads, size_t lds, hipStream_t st) {
#define TVBK_CASE(MODEL, CFV)                                                       \
  if (init) return launch_init_t<RH, R, MODEL, CFV>(*ia, ntiles, st);               \
  return launch_step_t<RH, R, MODEL, CFV>(*a, ntiles, nthreads, lds, st);
  if (model == 0 && cf == CF_LINEAR) { TVBK_CASE(G2D, CF_LINEAR) }
  if (model == 1 && cf == CF_LINEAR) { TVBK_CASE(MPR, CF_LINEAR) }
  if (model == 2 && cf == CF_LINEAR) { TVBK_CASE(JR, CF_LINEAR) }
  if (model == 2 && cf == CF_SIGMOID_JR) { TVBK_CASE(JR, CF_SIGMOID_JR) }
  if (model == 3 && cf == CF_IDENTITY) { TVBK_CASE(SUM, CF_IDENTITY) }
  if (model == 3 && cf == CF_LINEAR) { TVBK_CASE(SUM, CF_LINEAR) }
#undef TVBK_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_dense(int precision, int model, const StepArgs& a, int ntiles, int nthreads, size_t lds,
                        hipStream_t st) {
  if (precision == 2) {
    if (model == 0) return launch_dense_t<double, G2D>(a, ntiles, nthreads, lds, st);
    if (model == 1) return launch_dense_t<double, MPR>(a, ntiles, nthreads, lds, st);
    return hipErrorInvalidValue;
  }
  if (precision != 0) return hipErrorInvalidValue;
  if (model == 0) return launch_dense_t<float, G2D>(a, ntiles, nthreads, lds, st);
  if (model == 1) return launch_dense_t<float, MPR>(a, ntiles, nthreads, lds, st);
  return hipErrorInvalidValue;
}

// precision: 0 float history and state, 1 double both, 2 float history with
// double state
hipError_t launch_step(int precision, int model, int cf, const StepArgs& a, int ntiles, int nthreads, size_t lds,
                       hipStream_t st) {
  if (precision == 1) return dispatch<double, double>(model, cf, false, &a, nullptr, ntiles, nthreads, lds, st);
  if (precision == 2) return dispatch<float, double>(model, cf, false, &a, nullptr, ntiles, nthreads, lds, st);
  return dispatch<float, float>(model, cf, false, &a, nullptr, ntiles, nthreads, lds, st);
}

hipError_t launch_init(int precision, int model, int cf, const InitArgs& a, int ntiles, hipStream_t st) {
  if (precision == 1) return dispatch<double, double>(model, cf, true, nullptr, &a, ntiles, 0, 0, st);
  if (precision == 2) return dispatch<float, double>(model, cf, true, nullptr, &a, ntiles, 0, 0, st);
  return dispatch<float, float>(model, cf, true, nullptr, &a, ntiles, 0, 0, st);
}

hipError_t launch_scatter(int precision, bool gather, void* ring, unsigned char* img_state, void* hist,
                          void* state, int H, int N, int nch, int nsv, int T, int t, hipStream_t st) {
  if (precision == 1)
    hipLaunchKernelGGL((k_scatter<double, double>), dim3(64), dim3(256), 0, st, gather, (double*)ring,
                       (double*)img_state, (double*)hist, (double*)state, H, N, nch, nsv, T, t);
  else if (precision == 2)
    hipLaunchKernelGGL((k_scatter<float, double>), dim3(64), dim3(256), 0, st, gather, (float*)ring,
                       (double*)img_state, (float*)hist, (double*)state, H, N, nch, nsv, T, t);
  else
    hipLaunchKernelGGL((k_scatter<float, float>), dim3(64), dim3(256), 0, st, gather, (float*)ring,
                       (float*)img_state, (float*)hist, (float*)state, H, N, nch, nsv, T, t);
  return hipGetLastError();
}

hipError_t launch_debug_slots(const uint32_t* edges, int nnz, int H, int64_t step, int32_t* out, hipStream_t st) {
  if (nnz <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_debug_slots, dim3((nnz + 255) / 256), dim3(256), 0, st, edges, nnz, H, step, out);
  return hipGetLastError();
}

}  // namespace tvbk
