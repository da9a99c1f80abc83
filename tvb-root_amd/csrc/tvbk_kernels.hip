// Small kernels (history scatter/gather, slot arithmetic check) and the
// dispatch from the plan's runtime (precision, model, coupling) to the
// instantiated step, init and dense kernels of tvbk_inst.hip.
#include <algorithm>
#include <mutex>

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
      RH* r = ring + ring_index(T, H, N, nch, slot, i, ch, t);
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
  if (model == 1 && cf == CF_LINEAR1) { TVBK_CASE(MPR, CF_LINEAR1) }
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

// numpy's pairwise summation (the add.reduce inner loop over a contiguous
// axis): under 8 values a plain sum from 0, up to 128 eight interleaved
// partial sums combined as ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)) and the rest
// added in order, beyond that the two halves (the first a multiple of 8) summed
// separately.  Templated on the remaining split depth (2^12 * 128 values).
template <class R, int D>
struct PairwiseSum {
  __device__ static R run(const R* x, int64_t st, int64_t n) {
    if (n <= 128) return PairwiseSum<R, 0>::run(x, st, n);
    int64_t n2 = n / 2;
    n2 -= n2 % 8;
    return PairwiseSum<R, D - 1>::run(x, st, n2) + PairwiseSum<R, D - 1>::run(x + n2 * st, st, n - n2);
  }
};
template <class R>
struct PairwiseSum<R, 0> {
  __device__ static R run(const R* x, int64_t st, int64_t n) {
    if (n < 8) {
      R res = R(0);
      for (int64_t i = 0; i < n; ++i) res = res + x[i * st];
      return res;
    }
    R r[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = x[j * st];
    int64_t i = 8;
    for (; i < n - (n % 8); i += 8)
#pragma unroll
      for (int j = 0; j < 8; ++j) r[j] = r[j] + x[(i + j) * st];
    R res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; ++i) res = res + x[i * st];
    return res;
  }
};

// GlobalAverage (monitors.py:352-357: numpy.mean over the node axis) and
// SpatialAverage (:328-333: spatial_mean . state, each area's members in node
// order) of subsampled monitor data, one thread per (sample, voi, area) of an
// instance (blockIdx.y strides over instances).
template <class R>
__global__ void k_areas(int mode, const R* __restrict__ raw, R* __restrict__ out, const AreaJob* __restrict__ jobs,
                        int n_jobs, int max_count, int nvoi, int n_area, const int32_t* __restrict__ area_of,
                        const double* __restrict__ weight) {
  const int64_t per = (int64_t)max_count * nvoi * n_area;
  for (int b = blockIdx.y; b < n_jobs; b += gridDim.y) {
    const AreaJob jb = jobs[b];
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < per; q += (int64_t)gridDim.x * blockDim.x) {
      const int a = (int)(q % n_area);
      const int v = (int)((q / n_area) % nvoi);
      const int j = (int)(q / ((int64_t)n_area * nvoi));
      if (j >= jb.count) continue;
      const R* xs = raw + jb.raw_off + (int64_t)j * jb.rsj + (int64_t)v * jb.rsv;
      R res;
      if (mode == 0) {
        res = PairwiseSum<R, 12>::run(xs, jb.rsn, jb.n_node) / (R)jb.n_node;
      } else {
        res = R(0);
        for (int n = 0; n < jb.n_node; ++n)
          if (area_of[n] == a) res = res + (R)weight[n] * xs[(int64_t)n * jb.rsn];
      }
      out[jb.out_off + (int64_t)j * jb.osj + (int64_t)v * jb.osv + (int64_t)a * jb.osa] = res;
    }
  }
}

hipError_t launch_areas(int esize, int mode, const void* raw, void* out, const AreaJob* jobs, int n_jobs,
                        int max_count, int nvoi, int n_area, const int32_t* area_of, const double* weight,
                        hipStream_t st) {
  if (n_jobs <= 0 || max_count <= 0) return hipSuccess;
  const int64_t per = (int64_t)max_count * nvoi * n_area;
  const int bx = (int)std::min<int64_t>((per + 255) / 256, 4096);
  const int by = std::min(n_jobs, 65535);
  if (esize == 8)
    hipLaunchKernelGGL(k_areas<double>, dim3(bx, by), dim3(256), 0, st, mode, (const double*)raw, (double*)out, jobs,
                       n_jobs, max_count, nvoi, n_area, area_of, weight);
  else
    hipLaunchKernelGGL(k_areas<float>, dim3(bx, by), dim3(256), 0, st, mode, (const float*)raw, (float*)out, jobs,
                       n_jobs, max_count, nvoi, n_area, area_of, weight);
  return hipGetLastError();
}

// Commit of a block of a split plan: every split tile copies the history its
// nodes produced at steps k0 .. k0 + nb - 1 of the call (staged as
// [j][node][channel][T]) into ring slots (step0 + k0 + j) % H.
template <class RH>
__global__ void k_commit(const TileDesc* __restrict__ tiles, const ConnDesc* __restrict__ conns, RH* ring,
                         const RH* __restrict__ stage, int T, int nch, int k0, int kn, int ntiles) {
  // grid.y strides over the tiles (a plan may have more than 65535)
  for (int ty = blockIdx.y; ty < ntiles; ty += gridDim.y) {
  const TileDesc td = tiles[ty];
  if (td.parts <= 1) continue;
  const int nb = (int)std::min<int64_t>(std::max<int64_t>(td.n_steps - k0, 0), kn);
  const ConnDesc cd = conns[td.conn];
  const int N = cd.n_node, H = cd.horizon;
  const int rs = nch * T, nn = td.node1 - td.node0;
  const int64_t per = (int64_t)nn * rs;
  const int64_t total = per * nb;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)(q / per);
    const int64_t r = q - (int64_t)j * per;
    const int i = td.node0 + (int)(r / rs), e = (int)(r % rs);
    const int ch = e / T, t = e - (e / T) * T;
    const int slot = (int)((td.step0 + k0 + j) % H);
    ring[td.ring_off + ring_index(T, H, N, nch, slot, i, ch, t)] = stage[td.stage_off + ((int64_t)j * N + i) * rs + e];
  }
  }
}

hipError_t launch_commit(int hsize, const TileDesc* tiles, const ConnDesc* conns, void* ring, const void* stage,
                         int T, int nch, int k0, int kn, int ntiles, hipStream_t st) {
  if (ntiles <= 0) return hipSuccess;
  const dim3 grid(16, std::min(ntiles, 65535));
  if (hsize == 8)
    hipLaunchKernelGGL(k_commit<double>, grid, dim3(256), 0, st, tiles, conns, (double*)ring, (const double*)stage, T,
                       nch, k0, kn, ntiles);
  else
    hipLaunchKernelGGL(k_commit<float>, grid, dim3(256), 0, st, tiles, conns, (float*)ring, (const float*)stage, T,
                       nch, k0, kn, ntiles);
  return hipGetLastError();
}

// Standard normals of the Philox noise source for steps step0 .. step0+n-1
// of one instance stream, z[k][sv][node], with the step kernels' generator
// (philox_normals in the history precision NR, widened exactly to double).
template <class NR>
__global__ void k_debug_normals(int nsv, int N, uint64_t stream, int64_t step0, int64_t n, uint32_t key0,
                                uint32_t key1, double* z) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n * N; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = q / N;
    const int i = (int)(q - k * N);
    NR zz[8];
    philox_normals<NR>(nsv, (uint32_t)i, stream, step0 + k, key0, key1, zz);
    for (int v = 0; v < nsv; ++v) z[(k * nsv + v) * N + i] = (double)zz[v];
  }
}

hipError_t launch_debug_normals(bool f64, int nsv, int N, uint64_t stream, int64_t step0, int64_t n, uint32_t key0,
                                uint32_t key1, double* z, hipStream_t st) {
  if (n <= 0 || N <= 0) return hipSuccess;
  const int64_t work = n * N;
  const int grid = (int)std::min<int64_t>((work + 255) / 256, 4096);
  if (f64)
    hipLaunchKernelGGL(k_debug_normals<double>, dim3(grid), dim3(256), 0, st, nsv, N, stream, step0, n, key0, key1, z);
  else
    hipLaunchKernelGGL(k_debug_normals<float>, dim3(grid), dim3(256), 0, st, nsv, N, stream, step0, n, key0, key1, z);
  return hipGetLastError();
}

// Device memory of the allocation-free entry points: one array per device
// (each GPU the library runs on holds its own copy until the process ends).
// Host code never uses the array's name as an address: scratch_device_ptr
// returns the device address of the current device's copy, looked up once
// per device with hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[TVBK_SCRATCH_BYTES];

hipError_t scratch_device_ptr(void** ptr) {
  static std::mutex mu;
  static void* addr[TVBK_SCRATCH_MAX_DEVICES] = {nullptr};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= TVBK_SCRATCH_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[dev] = p;
  }
  *ptr = addr[dev];
  return hipSuccess;
}

hipError_t launch_debug_slots(const uint32_t* edges, int nnz, int H, int64_t step, int32_t* out, hipStream_t st) {
  if (nnz <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_debug_slots, dim3((nnz + 255) / 256), dim3(256), 0, st, edges, nnz, H, step, out);
  return hipGetLastError();
}

}  // namespace tvbk
