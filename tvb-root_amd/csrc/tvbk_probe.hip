// Calibration of the delayed gather's memory pattern on this device.
//
// The step kernel reads the history ring as 256-byte pieces (one wave load:
// two channels x 32 instances x 4 bytes of one source node and slot) at
// offsets set by the connectome's delays, i.e. scattered over a ring of
// gigabytes, and writes one slot per step (a few percent of the bytes).
// k_probe_gather reproduces that pattern without the model (and
// k_probe_gather_half the 128-byte pieces of a one-channel tile of 32,
// whose wave loads read two of them): every wave
// reads 16 pieces per batch (K consecutive pieces per segment, at random
// segment-aligned offsets of a large buffer) and, with WR, stores one
// segment every K batches (1/17 of the bytes).  Its rate is the ceiling the
// step kernel's HBM traffic can be held against.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "tvbk_kernels.hpp"

namespace tvbk {

namespace {

__device__ inline uint32_t probe_hash(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

// buf holds nseg segments of K pieces (64 floats each); every address below
// is seg * 64K + 64k + lane with seg < nseg, k < K, lane < 64: inside buf
template <int K, bool WR>
__global__ void __launch_bounds__(1024) k_probe_gather(float* __restrict__ buf, uint32_t nseg, int iters,
                                                       float* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  constexpr int S = 16 / K;
  float acc = 0.f;
  for (int it = 0; it < iters; ++it) {
    float v[16];
#pragma unroll
    for (int q = 0; q < S; ++q) {
      const uint32_t seg = probe_hash(gw * 7919u + (uint32_t)(it * S + q) * 104729u + 17u) % nseg;
      const float* p = buf + (size_t)seg * (64 * K) + lane;
#pragma unroll
      for (int k = 0; k < K; ++k) v[q * K + k] = p[64 * k];
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) acc += v[q];
    if (WR && (it % K) == K - 1) {
      const uint32_t seg = probe_hash(gw * 31u + (uint32_t)it * 15485863u + 5u) % nseg;
      float* p = buf + (size_t)seg * (64 * K) + lane;
#pragma unroll
      for (int k = 0; k < K; ++k) p[64 * k] = acc * 1e-30f;
    }
  }
  out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

// 128-byte pieces: the two halves of a wave read different random
// segments of 32 floats (the gather of a one-channel tile of 32 instances);
// every address is seg * 32 + (lane & 31) with seg < nseg: inside buf
template <bool WR>
__global__ void __launch_bounds__(1024) k_probe_gather_half(float* __restrict__ buf, uint32_t nseg, int iters,
                                                            float* __restrict__ out) {
  const int lane = threadIdx.x & 63, half = lane >> 5;
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  float acc = 0.f;
  for (int it = 0; it < iters; ++it) {
    float v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const uint32_t seg = probe_hash(gw * 7919u + (uint32_t)(it * 16 + q) * 104729u + 17u + (uint32_t)half * 31337u) % nseg;
      v[q] = buf[(size_t)seg * 32 + (lane & 31)];
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) acc += v[q];
    if (WR) {
      const uint32_t seg = probe_hash(gw * 31u + (uint32_t)it * 15485863u + 5u + (uint32_t)half * 7717u) % nseg;
      buf[(size_t)seg * 32 + (lane & 31)] = acc * 1e-30f;
    }
  }
  out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

template <int K>
hipError_t launch_k(float* buf, uint32_t nseg, bool wr, int grid, int iters, float* out, hipStream_t st) {
  if (wr) k_probe_gather<K, true><<<grid, 1024, 0, st>>>(buf, nseg, iters, out);
  else k_probe_gather<K, false><<<grid, 1024, 0, st>>>(buf, nseg, iters, out);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_probe_gather(float* buf, size_t n_float, int k, bool wr, int grid, int iters, float* out,
                               hipStream_t st) {
  if (k == 0) {   // 128-byte pieces, one write per 16 reads
    const uint32_t nseg = (uint32_t)(n_float / 32);
    if (wr) k_probe_gather_half<true><<<grid, 1024, 0, st>>>(buf, nseg, iters, out);
    else k_probe_gather_half<false><<<grid, 1024, 0, st>>>(buf, nseg, iters, out);
    return hipGetLastError();
  }
  const uint32_t nseg = (uint32_t)(n_float / (64 * (size_t)k));
  switch (k) {
    case 1: return launch_k<1>(buf, nseg, wr, grid, iters, out, st);
    case 2: return launch_k<2>(buf, nseg, wr, grid, iters, out, st);
    case 4: return launch_k<4>(buf, nseg, wr, grid, iters, out, st);
    case 8: return launch_k<8>(buf, nseg, wr, grid, iters, out, st);
    case 16: return launch_k<16>(buf, nseg, wr, grid, iters, out, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace tvbk
