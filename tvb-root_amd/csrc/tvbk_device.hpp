// Device-side building blocks of the step kernels: neural-mass model dfuns,
// integrator schemes, the history-channel (pre-synaptic) transform, the
// numpy-order row reduction and the Philox4x32-10 noise source.
//
// Arithmetic follows the reference's operation order term by term (the host
// derives the parameter combinations the reference evaluates first, e.g.
// d*tau or 1/tau, in float64 exactly as numpy does).  The library is built
// with -ffp-contract=off so no multiply-add is fused where the reference
// rounds twice.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>

namespace tvbk {

// CF_LINEAR1: Linear coupling of the first coupling variable only -- the
// kernel form of a plan whose model multiplies every other coupling term by
// a zero coefficient (MontbrioPazoRoxin with cv = 0, infinite_theta.py:205)
enum { CF_LINEAR = 0, CF_SIGMOID_JR = 1, CF_IDENTITY = 2, CF_LINEAR1 = 3 };
enum { SCH_EULER_DET = 0, SCH_EULER_STO = 1, SCH_HEUN_DET = 2, SCH_HEUN_STO = 3, SCH_IDENTITY = 4 };
enum { NOISE_NONE = 0, NOISE_INJECTED = 1, NOISE_PHILOX = 2 };
enum { MON_RAW = 0, MON_TAVG = 1, MON_BOLD = 2 };

// ---------------------------------------------------------------------------
// Models.  NDP = number of derived parameters the host prepares per instance.
// ---------------------------------------------------------------------------

// Generic2dOscillator, _numba_dfun_g2d (models/oscillator.py:385-395).
// derived: d*tau, alpha, f, e, g, gamma*I, gamma, a, b, c, d, beta, tau
struct G2D {
  static constexpr int NSV = 2, NCV = 1, NDP = 13;
  static constexpr int cvar(int k) { return 0; }
  template <class R>
  __device__ static inline void dfun(const R* x, const R* c, const R* p, R* dx) {
    const R V = x[0], W = x[1];
    const R V2 = V * V;
    dx[0] = p[0] * (((((p[1] * W - p[2] * V2 * V) + p[3] * V2) + p[4] * V) + p[5]) + p[6] * c[0]);
    dx[1] = (p[10] * (((p[7] + p[8] * V) + p[9] * V2) - p[11] * W)) / p[12];
  }
  template <class R> __device__ static inline void bound(R* x) {}
};

// MontbrioPazoRoxin.dfun (models/infinite_theta.py:171-207), bound r >= 0
// (infinite_theta.py:129-134, applied by integrators.py:181-219).
// derived: 1/tau, Delta/(pi*tau), pi^2*tau^2, eta, J*tau, I, cr, cv
struct MPR {
  static constexpr int NSV = 2, NCV = 2, NDP = 8;
  static constexpr int cvar(int k) { return k; }
  template <class R>
  __device__ static inline void dfun(const R* x, const R* c, const R* p, R* dx) {
    const R r = x[0], V = x[1];
    dx[0] = p[0] * (p[1] + (R(2) * V) * r);
    dx[1] = p[0] * ((((((V * V - p[2] * (r * r)) + p[3]) + p[4] * r) + p[5]) + p[6] * c[0]) + p[7] * c[1]);
  }
  template <class R> __device__ static inline void bound(R* x) {
    if (x[0] < R(0)) x[0] = R(0);
    if (x[0] > R(DBL_MAX)) x[0] = R(DBL_MAX);
  }
};

// JansenRit, _numba_dfun_jr (models/jansen_rit.py:276-289).  Coupling enters
// through row 0 only (jansen_rit.py:268, :288).
// derived: 2*nu_max, r, v0, a_1*J, a_3*J, A*a, 2*a, a^2, mu, a_2*J, B*b,
//          a_4*J, 2*b, b^2
struct JR {
  static constexpr int NSV = 6, NCV = 2, NDP = 14;
  static constexpr int cvar(int k) { return 1 + k; }
  template <class R>
  __device__ static inline void dfun(const R* y, const R* c, const R* p, R* dx) {
    const R one = R(1);
    const R s12 = p[0] / (one + exp(p[1] * (p[2] - (y[1] - y[2]))));
    const R s01 = p[0] / (one + exp(p[1] * (p[2] - p[3] * y[0])));
    const R s03 = p[0] / (one + exp(p[1] * (p[2] - p[4] * y[0])));
    dx[0] = y[3];
    dx[1] = y[4];
    dx[2] = y[5];
    dx[3] = (p[5] * s12 - p[6] * y[3]) - p[7] * y[0];
    dx[4] = (p[5] * ((p[8] + p[9] * s01) + c[0]) - p[6] * y[4]) - p[7] * y[1];
    dx[5] = (p[10] * (p[11] * s03) - p[12] * y[5]) - p[13] * y[2];
  }
  template <class R> __device__ static inline void bound(R* x) {}
};

// Sum: dx = x + c, the integer test model of history_test.py:55-65.
struct SUM {
  static constexpr int NSV = 1, NCV = 1, NDP = 0;
  static constexpr int cvar(int k) { return 0; }
  template <class R>
  __device__ static inline void dfun(const R* x, const R* c, const R* p, R* dx) { dx[0] = x[0] + c[0]; }
  template <class R> __device__ static inline void bound(R* x) {}
};

// ---------------------------------------------------------------------------
// Coupling.  History channels hold what the coupling sums: the cvars for
// Linear/IdCoupling, and the pre-synaptic sigmoid of (y1 - y2) for
// SigmoidalJansenRit, whose pre depends on the source node only
// (coupling.py:370-373), so it is applied once when the state is written.
// NCH = history channels, NCT = coupling terms summed from them, NCD =
// coupling terms the dfun reads (terms NCT .. NCD-1 are the post of a zero
// sum).
// derived coupling params: Linear a, b; SigmoidalJR cmin, cmax-cmin, r,
// midpoint, a.
// ---------------------------------------------------------------------------
template <class M, int CF> struct Cfun;
template <class M> struct Cfun<M, CF_LINEAR> {
  static constexpr int NCH = M::NCV, NCT = M::NCV, NCD = M::NCV, NCP = 2;
  template <class R> __device__ static inline R channel(const R* x, int ch, const R* cp) { return x[M::cvar(ch)]; }
  template <class R> __device__ static inline R post(R gx, const R* cp) { return cp[0] * gx + cp[1]; }
  // parameters strided in LDS
  template <class R> __device__ static inline R post(R gx, const R* cp, int st) { return cp[0] * gx + cp[st]; }
};
template <class M> struct Cfun<M, CF_IDENTITY> {
  static constexpr int NCH = M::NCV, NCT = M::NCV, NCD = M::NCV, NCP = 0;
  template <class R> __device__ static inline R channel(const R* x, int ch, const R* cp) { return x[M::cvar(ch)]; }
  template <class R> __device__ static inline R post(R gx, const R* cp) { return gx; }
  template <class R> __device__ static inline R post(R gx, const R* cp, int st) { return gx; }
};
template <class M> struct Cfun<M, CF_SIGMOID_JR> {
  static constexpr int NCH = 1, NCT = 1, NCD = 1, NCP = 5;
  // SigmoidalJansenRit.pre (coupling.py:370-373) on x_j[:,0] - x_j[:,1]
  template <class R> __device__ static inline R channel(const R* x, int ch, const R* cp) {
    return cp[0] + cp[1] / (R(1) + exp(cp[2] * (cp[3] - (x[M::cvar(0)] - x[M::cvar(1)]))));
  }
  template <class R> __device__ static inline R post(R gx, const R* cp) { return cp[4] * gx; }
  template <class R> __device__ static inline R post(R gx, const R* cp, int st) { return cp[4 * st] * gx; }
};
// Linear on the first coupling variable: one history channel, one summed
// term; the dfun's other terms are a * 0 + b (their coefficient is zero)
template <class M> struct Cfun<M, CF_LINEAR1> {
  static constexpr int NCH = 1, NCT = 1, NCD = M::NCV, NCP = 2;
  template <class R> __device__ static inline R channel(const R* x, int ch, const R* cp) { return x[M::cvar(0)]; }
  template <class R> __device__ static inline R post(R gx, const R* cp) { return cp[0] * gx + cp[1]; }
  template <class R> __device__ static inline R post(R gx, const R* cp, int st) { return cp[0] * gx + cp[st]; }
};

// ---------------------------------------------------------------------------
// numpy add.reduceat order for one segment: x[0] + pairwise_sum(x[1:]),
// pairwise_sum as in numpy's loops: n < 8 sequential from 0, n <= 128 eight
// accumulators combined ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)) then the tail,
// larger n split at a multiple of 8 near n/2.  Handles n <= 1025.
// ---------------------------------------------------------------------------
template <class R, class F>
__device__ inline R pw_leaf(int s, int n, F& f) {
  if (n < 8) {
    R res = R(0);
    for (int i = 0; i < n; ++i) res += f(s + i);
    return res;
  }
  R r0 = f(s), r1 = f(s + 1), r2 = f(s + 2), r3 = f(s + 3);
  R r4 = f(s + 4), r5 = f(s + 5), r6 = f(s + 6), r7 = f(s + 7);
  const int nfull = n - (n & 7);
  int i = 8;
  for (; i < nfull; i += 8) {
    r0 += f(s + i);     r1 += f(s + i + 1);
    r2 += f(s + i + 2); r3 += f(s + i + 3);
    r4 += f(s + i + 4); r5 += f(s + i + 5);
    r6 += f(s + i + 6); r7 += f(s + i + 7);
  }
  R res = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
  for (; i < n; ++i) res += f(s + i);
  return res;
}

template <int DEPTH, class R, class F>
__device__ inline R pw_sum(int s, int n, F& f) {
  if constexpr (DEPTH == 0) {
    return pw_leaf<R>(s, n, f);
  } else {
    if (n <= 128) return pw_leaf<R>(s, n, f);
    int n2 = n / 2;
    n2 -= n2 & 7;
    R a = pw_sum<DEPTH - 1, R>(s, n2, f);
    R b = pw_sum<DEPTH - 1, R>(s + n2, n - n2, f);
    return a + b;
  }
}

// Same order for rows of at most RMAX entries whose values are already
// staged in registers v[0..n-1]: every load of the row is issued before the
// first add, one memory round trip per row.
template <int RMAX, class R>
__device__ inline R reduceat_regs(int n, const R* v) {
  if (n <= 1) return n == 1 ? v[0] : R(0);
  const int m = n - 1;                       // pairwise part v[1..n-1]
  R res;
  if (m < 8) {
    res = R(0);
#pragma unroll
    for (int i = 0; i < 7; ++i)
      if (i < m) res += v[1 + i];
  } else {
    const int nfull = m - (m & 7);
    R r[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) r[k] = v[1 + k];
#pragma unroll
    for (int b = 8; b + 8 <= RMAX - 1; b += 8)
      if (b < nfull) {
#pragma unroll
        for (int k = 0; k < 8; ++k) r[k] += v[1 + b + k];
      }
    res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
#pragma unroll
    for (int i = 8; i < RMAX - 1; ++i)
      if (i >= nfull && i < m) res += v[1 + i];
  }
  return v[0] + res;
}

// The same order with the products formed on use: x_e = w_e * v_e (rounded,
// as numpy multiplies before it sums); only entries e < n are touched.  The
// tail after the blocks of eight is selected, not branched, per entry.
template <int RMAX, class R, class W>
__device__ inline R reduceat_regs_w(int n, const R* v, const W* w) {
  auto x = [&](int e) -> R { return (R)w[e] * v[e]; };
  if (n <= 1) return n == 1 ? x(0) : R(0);
  const int m = n - 1;                       // pairwise part x[1..n-1]
  R res;
  if (m < 8) {
    res = R(0);
#pragma unroll
    for (int i = 0; i < 7; ++i)
      if (i < m) res += x(1 + i);
  } else {
    const int nfull = m - (m & 7), rem = m & 7;
    R r[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) r[k] = x(1 + k);
#pragma unroll
    for (int b = 8; b + 8 <= RMAX - 1; b += 8)
      if (b < nfull) {
#pragma unroll
        for (int k = 0; k < 8; ++k) r[k] += x(1 + b + k);
      }
    res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    // tail x[1 + nfull .. m - 1]: one uniform branch per possible nfull
#pragma unroll
    for (int b = 8; b < RMAX - 1; b += 8)
      if (nfull == b) {
#pragma unroll
        for (int j = 0; j < 7 && b + j < RMAX - 1; ++j)
          if (j < rem) res += x(1 + b + j);
      }
  }
  return x(0) + res;
}

// numpy add.reduceat order for rows longer than the register path: the
// pairwise blocks of eight are loaded as a batch before they are used.
// ld(e) returns the raw value of entry e (zero past the row end), wt(e) its
// weight; products are rounded before they are summed, as numpy does.
// PL: the next block's loads are issued before the current block is summed
// (two blocks in flight, the sums in the same order); every entry read lies
// below s + n + 8, as without it.
template <class R, bool PL, class LD, class WT>
__device__ inline R pw_leaf_blocked(int s, int n, LD& ld, WT& wt) {
  R x[8];
  if (n < 8) {
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = ld(s + k);
    R res = R(0);
#pragma unroll
    for (int k = 0; k < 7; ++k)
      if (k < n) res += wt(s + k) * x[k];
    return res;
  }
  if constexpr (PL) {
    R r[8], y[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = ld(s + k);
#pragma unroll
    for (int k = 0; k < 8; ++k) y[k] = ld(s + 8 + k);
#pragma unroll
    for (int k = 0; k < 8; ++k) r[k] = wt(s + k) * x[k];
    const int nfull = n - (n & 7);
    int i = 8;
    for (; i < nfull; i += 8) {
#pragma unroll
      for (int k = 0; k < 8; ++k) x[k] = ld(s + i + 8 + k);
#pragma unroll
      for (int k = 0; k < 8; ++k) r[k] += wt(s + i + k) * y[k];
#pragma unroll
      for (int k = 0; k < 8; ++k) y[k] = x[k];
    }
    R res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    const int rem = n - nfull;
#pragma unroll
    for (int k = 0; k < 7; ++k)
      if (k < rem) res += wt(s + i + k) * y[k];
    return res;
  }
  R r[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) x[k] = ld(s + k);
#pragma unroll
  for (int k = 0; k < 8; ++k) r[k] = wt(s + k) * x[k];
  const int nfull = n - (n & 7);
  int i = 8;
  for (; i < nfull; i += 8) {
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = ld(s + i + k);
#pragma unroll
    for (int k = 0; k < 8; ++k) r[k] += wt(s + i + k) * x[k];
  }
  R res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
  const int rem = n - nfull;
#pragma unroll
  for (int k = 0; k < 8; ++k) x[k] = ld(s + i + k);
#pragma unroll
  for (int k = 0; k < 7; ++k)
    if (k < rem) res += wt(s + i + k) * x[k];
  return res;
}

template <int DEPTH, class R, bool PL, class LD, class WT>
__device__ inline R pw_sum_blocked(int s, int n, LD& ld, WT& wt) {
  if constexpr (DEPTH == 0) {
    return pw_leaf_blocked<R, PL>(s, n, ld, wt);
  } else {
    if (n <= 128) return pw_leaf_blocked<R, PL>(s, n, ld, wt);
    int n2 = n / 2;
    n2 -= n2 & 7;
    R a = pw_sum_blocked<DEPTH - 1, R, PL>(s, n2, ld, wt);
    R b = pw_sum_blocked<DEPTH - 1, R, PL>(s + n2, n - n2, ld, wt);
    return a + b;
  }
}

template <class R, bool PL = false, class LD, class WT>
__device__ inline R reduceat_row_blocked(int n, LD& ld, WT& wt) {
  if (n <= 0) return R(0);
  const R first = wt(0) * ld(0);
  if (n == 1) return first;
  auto ld1 = [&](int e) -> R { return ld(e + 1); };
  auto wt1 = [&](int e) -> R { return wt(e + 1); };
  return first + pw_sum_blocked<3, R, PL>(0, n - 1, ld1, wt1);
}

// Raw buffer loads: 32-bit byte offsets from a per-tile base; offsets past
// num_records read as zero without a memory access.
__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
template <class R> __device__ inline R buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off);
template <> __device__ inline float buf_load<float>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <> __device__ inline double buf_load<double>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

template <class R, class F>
__device__ inline R reduceat_row(int n, F& f) {
  if (n <= 0) return R(0);
  R first = f(0);
  if (n == 1) return first;
  if (n - 1 <= 128) return first + pw_leaf<R>(1, n - 1, f);
  return first + pw_sum<3, R>(1, n - 1, f);
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11) and Box-Muller normals.
// ---------------------------------------------------------------------------
struct U4 { uint32_t x, y, z, w; };

__device__ inline U4 philox10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    // one 32 x 32 -> 64-bit product gives both halves (v_mad_u64_u32)
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
  }
  return c;
}

__device__ inline void box_muller(uint32_t a, uint32_t b, double* z0, double* z1) {
  const double u1 = ((double)a + 1.0) * 2.3283064365386963e-10;   // 2^-32
  const double u2 = (double)b * 2.3283064365386963e-10;
  const double rad = sqrt(-2.0 * log(u1));
  double s, c;
  sincos(2.0 * 3.141592653589793 * u2, &s, &c);
  *z0 = rad * c;
  *z1 = rad * s;
}

// float32: the hardware transcendentals -- v_log_f32 (log2), v_sqrt_f32 and
// v_sin_f32 / v_cos_f32, which take the angle in revolutions, so
// sin(2 pi u2) is one instruction; each within a few ulp.
__device__ inline void box_muller(uint32_t a, uint32_t b, float* z0, float* z1) {
  const float u1 = (float)((a >> 8) + 1u) * 5.9604644775390625e-8f;  // 2^-24, in (0, 1]
  const float u2 = (float)(b >> 8) * 5.9604644775390625e-8f;         // [0, 1)
  const float rad = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1));   // -2 ln u1
  *z0 = rad * __builtin_amdgcn_cosf(u2);
  *z1 = rad * __builtin_amdgcn_sinf(u2);
}

// Standard normals for state variables 0..nsv-1 of (node, stream, step).
template <class R>
__device__ inline void philox_normals(int nsv, uint32_t node, uint64_t stream, int64_t step,
                                      uint32_t k0, uint32_t k1, R* z) {
  for (int call = 0; call * 4 < nsv; ++call) {
    U4 c{node | ((uint32_t)call << 20), (uint32_t)stream, (uint32_t)(uint64_t)step,
         (uint32_t)((uint64_t)step >> 32)};
    U4 w = philox10(c, k0, k1);
    R a, b, cc, d;
    box_muller(w.x, w.y, &a, &b);
    box_muller(w.z, w.w, &cc, &d);
    const int o = call * 4;
    if (o + 0 < nsv) z[o + 0] = a;
    if (o + 1 < nsv) z[o + 1] = b;
    if (o + 2 < nsv) z[o + 2] = cc;
    if (o + 3 < nsv) z[o + 3] = d;
  }
}

// ---------------------------------------------------------------------------
// Integrator schemes (integrators.py:286-426, :496-508).  xi[sv] is the noise
// increment sqrt(dt)*z*gfun of this step (zero for deterministic schemes);
// with STIM, st[sv] is the region stimulus of this step, entering each scheme
// where the reference adds `stimulus` (:310, :315, :354, :359, :388, :423,
// :506), in the same operation order.
// ---------------------------------------------------------------------------
template <class R, class M, bool STIM = false>
__device__ inline void scheme_step(int sch, R* x, const R* c, const R* p, R dt, const R* xi,
                                   const R* st = nullptr) {
  constexpr int S = M::NSV;
  R m[S], y[S], m2[S];
  M::dfun(x, c, p, m);
  switch (sch) {
    case SCH_HEUN_DET:
    case SCH_HEUN_STO: {
      const bool sto = (sch == SCH_HEUN_STO);
#pragma unroll
      for (int k = 0; k < S; ++k) {
        if constexpr (STIM)
          y[k] = sto ? ((x[k] + dt * m[k]) + xi[k]) + dt * st[k] : x[k] + dt * (m[k] + st[k]);
        else
          y[k] = sto ? (x[k] + dt * m[k]) + xi[k] : x[k] + dt * m[k];
      }
      M::template bound<R>(y);
      M::dfun(y, c, p, m2);
#pragma unroll
      for (int k = 0; k < S; ++k) {
        const R dX = ((m[k] + m2[k]) * dt) * R(0.5);
        x[k] = sto ? (x[k] + dX) + xi[k] : x[k] + dX;
        if constexpr (STIM) x[k] = x[k] + dt * st[k];
      }
      break;
    }
    case SCH_EULER_DET:
#pragma unroll
      for (int k = 0; k < S; ++k) {
        if constexpr (STIM) x[k] = x[k] + dt * (m[k] + st[k]);
        else x[k] = x[k] + dt * m[k];
      }
      break;
    case SCH_EULER_STO:
#pragma unroll
      for (int k = 0; k < S; ++k) {
        x[k] = (x[k] + m[k] * dt) + xi[k];
        if constexpr (STIM) x[k] = x[k] + dt * st[k];
      }
      break;
    default:  // identity
#pragma unroll
      for (int k = 0; k < S; ++k) {
        if constexpr (STIM) x[k] = m[k] + st[k];
        else x[k] = m[k];
      }
      break;
  }
  M::template bound<R>(x);
}

}  // namespace tvbk
