// Balloon-Windkessel BOLD of neural-input series (BalloonModel.evaluate,
// tvb_library/tvb/analyzers/fmri_balloon.py:198-275, neural input
// transformation "none", HeunDeterministic scheme, integrators.py:309-318).
//
// One thread per series: the model state (s, f, v, q) stays in registers
// while the thread walks the samples in order; input loads and output
// stores are coalesced across series (series innermost).  Float64
// throughout, in the reference's operation order (the library is built
// without contraction, so no multiply-add is fused).
#include "tvbk_kernels.hpp"

namespace tvbk {

struct BalloonPar {
  double inv_tau_s, inv_tau_f, inv_tau_o, inv_alpha, one_m_e0, e0, v0, k1, k2, k3, dt;
  int32_t linear, pad;
};

// balloon_dfun (fmri_balloon.py:330-357)
__device__ inline void balloon_dfun(const BalloonPar& p, double s, double f, double v, double q, double x,
                                    double* d) {
  const double vpow = pow(v, p.inv_alpha);
  d[0] = x - p.inv_tau_s * s - p.inv_tau_f * (f - 1.0);
  d[1] = s;
  d[2] = p.inv_tau_o * (f - vpow);
  d[3] = p.inv_tau_o * ((f * (1.0 - pow(p.one_m_e0, 1.0 / f)) / p.e0) - vpow * (q / v));
}

// BOLD equations (fmri_balloon.py:253-268)
__device__ inline double balloon_bold(const BalloonPar& p, double v, double q) {
  if (p.linear) return p.v0 * ((p.k1 + p.k2) * (1.0 - q) + (p.k3 - p.k2) * (1.0 - v));
  return p.v0 * (p.k1 * (1.0 - q) + p.k2 * (1.0 - q / v) + p.k3 * (1.0 - v));
}

template <class X>
__global__ void __launch_bounds__(256) k_balloon(const X* __restrict__ x, int64_t x_stride, int64_t n_samples,
                                                 int64_t n_series, BalloonPar p, int start, int normalize,
                                                 double* __restrict__ state, double* __restrict__ y) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_series) return;
  double mean = 0.0;
  if (normalize) {
    // numpy's mean over the time axis: rows added in order, then divided
    for (int64_t t = 0; t < n_samples; ++t) mean += (double)x[t * x_stride + j];
    mean = mean / (double)n_samples;
  }
  double s, f, v, q;
  int64_t t0 = 0;
  if (start) {
    s = 0.0, f = 1.0, v = 1.0, q = 1.0;       // fmri_balloon.py:208-213
    if (n_samples > 0) y[j] = balloon_bold(p, v, q);
    t0 = 1;
  } else {
    s = state[j], f = state[n_series + j], v = state[2 * n_series + j], q = state[3 * n_series + j];
  }
  for (int64_t t = t0; t < n_samples; ++t) {
    double xv = (double)x[t * x_stride + j];
    if (normalize) xv = xv - mean;
    // HeunDeterministic.scheme with stimulus 0, no state bounds
    double m[4], m2[4];
    balloon_dfun(p, s, f, v, q, xv, m);
    const double is = s + p.dt * (m[0] + 0.0), if_ = f + p.dt * (m[1] + 0.0), iv = v + p.dt * (m[2] + 0.0),
                 iq = q + p.dt * (m[3] + 0.0);
    balloon_dfun(p, is, if_, iv, iq, xv, m2);
    s = s + (m[0] + m2[0]) * p.dt / 2.0 + p.dt * 0.0;
    f = f + (m[1] + m2[1]) * p.dt / 2.0 + p.dt * 0.0;
    v = v + (m[2] + m2[2]) * p.dt / 2.0 + p.dt * 0.0;
    q = q + (m[3] + m2[3]) * p.dt / 2.0 + p.dt * 0.0;
    y[t * n_series + j] = balloon_bold(p, v, q);
  }
  state[j] = s;
  state[n_series + j] = f;
  state[2 * n_series + j] = v;
  state[3 * n_series + j] = q;
}

hipError_t launch_balloon(const void* x, int x_elem_bytes, int64_t x_stride, int64_t n_samples, int64_t n_series,
                          const double* par, int linear, int classical, int start, int normalize,
                          double sample_period_ms, double* state, double* y, hipStream_t st) {
  // par: tau_s, tau_f, tau_o, alpha, TE, V0, E0, epsilon, nu_0, r_0
  BalloonPar p{};
  p.inv_tau_s = 1.0 / par[0];
  p.inv_tau_f = 1.0 / par[1];
  p.inv_tau_o = 1.0 / par[2];
  p.inv_alpha = 1.0 / par[3];
  p.one_m_e0 = 1.0 - par[6];
  p.e0 = par[6];
  p.v0 = par[5];
  if (classical) {   // fmri_balloon.py:283-291
    p.k1 = 7.0 * par[6];
    p.k2 = 2.0 * par[6];
    p.k3 = 1.0 - par[7];
  } else {           // :292-301
    p.k1 = 4.3 * par[8] * par[6] * par[4];
    p.k2 = par[7] * par[9] * par[6] * par[4];
    p.k3 = 1 - par[7];
  }
  p.dt = 1.0 / (1000.0 / sample_period_ms);   // 1 / sample_rate, seconds
  p.linear = linear;
  const int nb = (int)((n_series + 255) / 256);
  if (x_elem_bytes == 8)
    hipLaunchKernelGGL(k_balloon<double>, dim3(nb), dim3(256), 0, st, (const double*)x, x_stride, n_samples,
                       n_series, p, start, normalize, state, y);
  else
    hipLaunchKernelGGL(k_balloon<float>, dim3(nb), dim3(256), 0, st, (const float*)x, x_stride, n_samples,
                       n_series, p, start, normalize, state, y);
  return hipGetLastError();
}

}  // namespace tvbk
