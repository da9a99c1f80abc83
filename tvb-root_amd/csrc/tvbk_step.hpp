// Persistent fused step kernel for TVB's simulator loop on gfx950.
//
// One workgroup owns a "tile": T simulation instances that share one
// connectome (T = 16/32/64 for parameter sweeps over a shared connectome;
// T = 1 for instances with their own connectome, one per workgroup).  The
// workgroup advances its tile through all steps of a call without any
// inter-workgroup synchronisation: instances are independent.
//
// Per step s (Simulator.__call__, simulator.py:356-363, with the coupling of
// step s evaluated from the history as it stood after step s-1):
//   Phase A  coupling c(s): every (target node, history channel) pair is
//            reduced by a group of T lanes (one lane per instance) over the
//            node's CSR row in numpy add.reduceat order, gathering
//            history[(s-1-d) % H][source][channel][lane] -- T consecutive
//            values per edge, so each gather is one coalesced segment.
//   barrier
//   Phase B  integrator step of every (node, instance) with c(s), write of
//            history slot s % H, on-device monitors.
// With every delay >= 1, phase A of step s+1 only reads slots written before
// the barrier of step s, so one barrier per step suffices; coupling terms are
// double-buffered by step parity.  With zero delays a second barrier orders
// the history write before the next gather.
//
// Layouts (elements of the plan precision R unless stated):
//   history  ring[tile][slot][node][channel][T]; single-instance tiles
//            ring[tile][node][channel][slot] (ring_index)
//   LDS      bold interim (f64) | state [node][svar][T] | tavg sum, comp
//            [node][voi][T] | c[2][node][term][T] | CSR (packed col|delay,
//            weight, row_ptr) when it fits
//   the persistent part of LDS (bold interim, state, tavg) is copied from and
//   back to a per-tile image in HBM around each call.
#pragma once
#include <type_traits>

#include "tvbk_kernels.hpp"
#include "tvbk_device.hpp"

namespace tvbk {

static_assert(CF_LINEAR1 == KCF_LINEAR1, "the host's kernel coupling id");

__device__ inline void copy_bytes(unsigned char* dst, const unsigned char* src, int64_t n) {
  // 16-byte granularity copy; all images are padded to 16 bytes.
  const int64_t n16 = n >> 4;
  uint4* d = reinterpret_cast<uint4*>(dst);
  const uint4* s = reinterpret_cast<const uint4*>(src);
  for (int64_t k = threadIdx.x; k < n16; k += blockDim.x) d[k] = s[k];
}

// Tile, monitor and observable configuration staged in LDS once per launch,
// plus the per-step monitor bookkeeping (double-buffered by step parity,
// written by thread 0 for the next step before it reaches that step's
// barrier, read by everyone in phase B).
struct StepInfo {
  int32_t emit, push, msample, pad;
  int64_t mp, kk;
  // Bold push with istep = rat * ii: emissions klo .. klo + fir_n - 1 take the
  // interim sample, at accumulator slots fir_slot0 + j (mod Q) with ages
  // fir_age0 + j * rat
  int32_t fir_n, fir_slot0, fir_age0, pad2;
};
static_assert(sizeof(StepInfo) == TVBK_STEPINFO_BYTES, "StepInfo size is part of the LDS layout");

struct KCfg {
  TileDesc td;
  MonArgs mon[TVBK_KMAX_MON];
  StepInfo si[2][TVBK_KMAX_MON];
  int32_t obs_sv[TVBK_KMAX_OBS];
  double obs_coef[TVBK_KMAX_OBS][8];
};

// One CSR entry as staged in LDS: byte offset of the source node inside a
// history slot, delay, and the weight's bits.
struct EdgeRec {
  uint32_t coff, delay, w0, w1;
};

template <class R> __device__ inline R rec_weight(const EdgeRec& e);
template <> __device__ inline float rec_weight<float>(const EdgeRec& e) { return __uint_as_float(e.w0); }
template <> __device__ inline double rec_weight<double>(const EdgeRec& e) {
  return __hiloint2double((int)e.w1, (int)e.w0);
}

// Edge arrays of the wave-uniform row path, read through the scalar cache
// in 32-byte chunks: offsets {delay * slot bytes, source node * node bytes}
// (four edges a chunk) and weights (eight float32 or four float64 a chunk).
typedef unsigned int u32x8a __attribute__((ext_vector_type(8), aligned(4)));
typedef const __attribute__((address_space(4))) u32x8a ScalarChunk;
typedef const __attribute__((address_space(4))) uint32_t ScalarU32;
typedef const __attribute__((address_space(4))) int32_t ScalarInt;

template <class R> __device__ inline R chunk_weight(const u32x8a& c, int j);
template <> __device__ inline float chunk_weight<float>(const u32x8a& c, int j) { return __uint_as_float(c[j]); }
template <> __device__ inline double chunk_weight<double>(const u32x8a& c, int j) {
  return __hiloint2double((int)c[2 * j + 1], (int)c[2 * j]);
}

// history value at vector offset voff (per lane) + scalar offset soff
template <class R> __device__ inline R buf_load2(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff);
template <> __device__ inline float buf_load2<float>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <> __device__ inline double buf_load2<double>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// Per-step monitor bookkeeping for step s = s0 + k (uniform): emission when
// s % istep == 0 (sample index within the call), Bold push when s % ii == 0
// (interim sample index mp = s / ii) and emission index kk = s / istep.
__device__ inline void step_info(const MonArgs& mo, int k, StepInfo& si) {
  const int ek = mo.r0 + k;
  si.emit = (ek % mo.istep) == 0;
  // sample index among the call's emissions; afferent monitors keep the
  // phase of step s0 - 1 (their samples of step s are taken at step s + 1)
  si.msample = ek / mo.istep - (mo.src ? 1 : (mo.r0 > 0 ? 1 : 0));
  si.kk = mo.qe0 + ek / mo.istep;
  const int pk = mo.rb0 + k;
  si.push = (pk % mo.ii) == 0;
  si.mp = mo.qb0 + pk / mo.ii;
  if (mo.kind == MON_BOLD && mo.rat > 0 && si.push) {
    const int64_t klo = (si.mp + mo.rat - 1) / mo.rat;
    const int64_t khi = (si.mp + mo.S - 1) / mo.rat;
    si.fir_n = (int32_t)(khi - klo + 1);
    si.fir_slot0 = (int32_t)(klo % mo.Q);
    si.fir_age0 = (int32_t)(klo * mo.rat - si.mp);
  }
}

// Parameter k of this lane from `par` (stride pst: registers or LDS).
template <class R>
__device__ __forceinline__ R lane_param(const StepArgs& a, const R* par, int pst, int k) {
  return par[k * pst];
}

// Per-lane parameter rows (model, coupling, noise gfun; node 0's row for
// spatial parameters) staged in LDS as [param][lane] with stride T, read by
// phase B instead of being held in registers across the step loop.
template <class R>
__device__ inline void stage_params(const StepArgs& a, unsigned char* smem, int inst0, int n_inst, int N) {
  R* s_par = reinterpret_cast<R*>(smem + a.lds_par);
  const int T = a.T;
  const int64_t row = a.spatial ? (int64_t)N * a.prow : (int64_t)a.prow;
  for (int q = threadIdx.x; q < a.prow * T; q += blockDim.x) {
    const int k = q / T, tt = q - (q / T) * T;
    const int inst = inst0 + (tt < n_inst ? tt : 0);
    s_par[q] = reinterpret_cast<const R*>(a.params)[(int64_t)inst * row + k];
  }
}

// Monitors of one (node i, lane t) pair after step k (monitors.py:160-190,
// :205-214, :363-431, :834-962): samples of the observed state x, or of the
// node coupling c (afferent monitors, which sample the coupling of step
// s + 1 at step s, simulator.py:311-316).  EXT: afferent monitors may be
// present; aff_only: the pass after the last step, afferent monitors only.
template <class R, class M, int CF, bool EXT>
__device__ __forceinline__ void monitors_node(const StepArgs& a, const KCfg& kc, unsigned char* smem, int i, int li,
                                              int t, int T, int N, bool live, int k, const StepInfo* sir, const R* x,
                                              const R* c, bool aff_only) {
  using C = Cfun<M, CF>;
  constexpr int NSV = M::NSV, NCT = C::NCT;
  // this tile's descriptor: read-only during the kernel, scalar loads
  const auto& td = *(const __attribute__((address_space(4))) TileDesc*)(a.tiles + blockIdx.x);
  auto sample_val = [&](const MonArgs& mo, int v) -> R {
    const int o = mo.voi[v];
    R val;
    if (EXT && mo.src) {       // coupling term o of this node
      val = c[0];
#pragma unroll
      for (int u = 1; u < NCT; ++u) val = (u == o) ? c[u] : val;
      return val;
    }
    const int osv = a.obs_sv[o];
    if (osv >= 0) {
      val = x[0];
#pragma unroll
      for (int u = 1; u < NSV; ++u) val = (u == osv) ? x[u] : val;
    } else {
      val = R(0);
#pragma unroll
      for (int u = 0; u < NSV; ++u) val = val + (R)a.obs_coef[o][u] * x[u];
    }
    return val;
  };
  for (int m = 0; m < a.n_mon; ++m) {
    const MonArgs& mo = a.mon[m];   // plan constants: kernel-argument (scalar) loads
    if constexpr (EXT) {
      // afferent samples of step s - 1 are taken at step s (k >= 1) and by
      // the tail pass after the loop; the others at their own step
      if (mo.src ? (k == a.k0) : aff_only) continue;
    }
    const StepInfo& si = sir[m];
    const int nv = mo.nvoi;
    const bool emit = si.emit != 0;
    if (mo.kind == MON_RAW || mo.kind == MON_TAVG) {
      auto outp = [&]() -> R* {
        return reinterpret_cast<R*>(a.out[m]) + td.out_off[m] + (int64_t)si.msample * td.out_sj[m] +
               (int64_t)i * td.out_si[m] + t;
      };
      R* ts = reinterpret_cast<R*>(smem + mo.lds_off);
      // running sum over the averaging window: Kahan-compensated in float32
      // (sum su, compensation cm), plain sequential (numpy's mean order) in
      // float64 (sum only); su and cm are the LDS values of value v
      constexpr int NSL = sizeof(R) == 8 ? 1 : 2;
      auto tavg_value = [&](int v, R val, R& su, R& cm) {
        R tv;
        if constexpr (sizeof(R) == 8) {
          tv = su + val;
        } else {
          const R yv = val - cm;
          tv = su + yv;
          cm = (tv - su) - yv;
        }
        su = tv;
        if (emit) {
          if (live) outp()[v * td.out_sv[m]] = tv / (R)mo.istep;
          su = R(0);
          cm = R(0);
        }
      };
      auto slot_of = [&](int v, int q) -> R* { return ts + ((li * nv + v) * NSL + q) * T + t; };
      int v0 = 0;
      if (mo.kind == MON_RAW) {
        for (int v = 0; v < nv; ++v)
          if (live && emit) outp()[v * td.out_sv[m]] = sample_val(mo, v);
        v0 = nv;
      }
      for (int v = v0; v < nv; ++v) {
        R su = *slot_of(v, 0), cm = NSL == 2 ? *slot_of(v, NSL - 1) : R(0);
        tavg_value(v, sample_val(mo, v), su, cm);
        *slot_of(v, 0) = su;
        if constexpr (NSL == 2) *slot_of(v, NSL - 1) = cm;
      }
    } else {  // Bold: interim mean every ii steps pushed into FIR accumulators
      double* bi = reinterpret_cast<double*>(smem + mo.lds_off);
      const bool push = si.push != 0;
      const int64_t mp = si.mp;                      // interim sample index s / ii
      double* acc = a.bacc + td.bacc_off[m];
      // interim sum bs (the LDS value of value v) and this step's sample
      auto bold_value = [&](int v, R val, double& bs) {
        bs = bs + (double)val;
        if (push) {
          const double mean = bs / (double)mo.ii;
          bs = 0.0;
          if (mo.stock) {   // the reference's stock slot (mp - 1) % S (monitors.py:948), mp >= 1
            const int64_t sl = (mp - 1) % mo.S;
            mo.stock[td.bst_off[m] + ((sl * nv + v) * (int64_t)N + i) * T + t] = mean;
          }
          // sample mp contributes to emissions kk with kk*istep >= mp*ii
          // and kk*istep < (mp+S)*ii; weight by age a = floor(kk*istep/ii) - mp:
          // a == 0 -> G[S-1], else G[a-1]
          if (mo.rat > 0) {
            // istep = rat * ii: the schedule of this push, from the step's
            // bookkeeping (consecutive slots, ages rat apart)
            // in pairs: both accumulators are loaded before either is
            // stored, two read-modify-writes in flight (the slots of one
            // push are distinct; larger groups cost registers elsewhere)
            constexpr int GRP = 2;
            int slot = si.fir_slot0, age = si.fir_age0;
            const int64_t sstr = (int64_t)nv * N * T;           // doubles between slots
            double* ap0 = acc + ((int64_t)v * N + i) * T + t;
            for (int j0 = 0; j0 < si.fir_n; j0 += GRP) {
              double* ap[GRP];
              double av[GRP], wv[GRP];
#pragma unroll
              for (int q = 0; q < GRP; ++q) {
                if (j0 + q < si.fir_n) {
                  wv[q] = mo.taps[age == 0 ? mo.S - 1 : age - 1];
                  ap[q] = ap0 + slot * sstr;
                  av[q] = *ap[q];
                  age += mo.rat;
                  slot = (slot + 1 == mo.Q) ? 0 : slot + 1;
                }
              }
#pragma unroll
              for (int q = 0; q < GRP; ++q)
                if (j0 + q < si.fir_n) *ap[q] = av[q] + wv[q] * mean;
            }
          } else {
            const int64_t klo = (mp * mo.ii + mo.istep - 1) / mo.istep;
            const int64_t khi = ((mp + mo.S) * mo.ii - 1) / mo.istep;
            for (int64_t kk = klo; kk <= khi; ++kk) {
              const int64_t age = (kk * mo.istep) / mo.ii - mp;
              const double w = mo.taps[age == 0 ? mo.S - 1 : age - 1];
              double* ap = acc + (((kk % mo.Q) * nv + v) * (int64_t)N + i) * T + t;
              *ap = *ap + w * mean;
            }
          }
        }
        if (emit) {
          const int64_t kk = si.kk;                  // s / istep
          double* ap = acc + (((kk % mo.Q) * nv + v) * (int64_t)N + i) * T + t;
          double* out = reinterpret_cast<double*>(a.out[m]) + td.out_off[m] +
                        (int64_t)si.msample * td.out_sj[m] + (int64_t)v * td.out_sv[m] +
                        (int64_t)i * td.out_si[m] + t;
          const double bold = mo.sub1 ? (*ap - 1.0) * mo.k1v0 : *ap;
          if (live) *out = bold;
          *ap = 0.0;
        }
      };
      auto bslot = [&](int v) -> double* { return bi + (li * nv + v) * T + t; };
      for (int v = 0; v < nv; ++v) {
        double bs = *bslot(v);
        bold_value(v, sample_val(mo, v), bs);
        *bslot(v) = bs;
      }
    }
  }
}

// Integrator step, history write and monitors of one (node i, lane t) pair
// with its coupling terms c (phase B of the step kernels).  `sink(ch, v)`
// stores history channel ch of the new state.  EXT: region stimulus and
// afferent monitors may be present (the sequential loops of k_step).
// zpre: this step's injected normals when the caller has loaded them
// already (else they are loaded here).
template <class R, class M, int CF, class NR, bool UM = false, bool EXT = false, class SINK>
__device__ __forceinline__ void update_node(const StepArgs& a, const KCfg& kc, unsigned char* smem, R* s_x, int i,
                                            int li, int t, int T, int N, bool live, int inst, uint64_t stream, int k,
                                            int64_t s, const StepInfo* sir, const R* par, int pst, const R* c,
                                            const double* zpre, int& bad, SINK sink) {
  using C = Cfun<M, CF>;
  constexpr int NSV = M::NSV, NDP = M::NDP, NCP = C::NCP, NCH = C::NCH;
  // this tile's descriptor: read-only during the kernel, scalar loads
  const auto& td = *(const __attribute__((address_space(4))) TileDesc*)(a.tiles + blockIdx.x);
  const R dt = (R)a.dt;
  const int64_t spatial_row = (int64_t)N * a.prow;
  R x[NSV], xi[NSV];
#pragma unroll
  for (int v = 0; v < NSV; ++v) x[v] = s_x[(li * NSV + v) * T + t];
  // this lane's parameters (registers, stride 1, or LDS, stride T): model,
  // coupling, noise gfun
  R pp[NDP > 0 ? NDP : 1], cp[NCP > 0 ? NCP : 1], gf[NSV];
  // model parameters: plan constants from the kernel arguments (UM), else
  // this lane's
#pragma unroll
  for (int v = 0; v < NDP; ++v) pp[v] = UM ? (R)a.par_uval[v] : lane_param<R>(a, par, pst, v);
#pragma unroll
  for (int v = 0; v < NCP; ++v) cp[v] = lane_param<R>(a, par, pst, NDP + v);
#pragma unroll
  for (int v = 0; v < NSV; ++v) gf[v] = (a.noise != NOISE_NONE) ? lane_param<R>(a, par, pst, NDP + NCP + v) : R(0);
  if (a.spatial) {
    const R* pr = reinterpret_cast<const R*>(a.params) + (int64_t)inst * spatial_row + (int64_t)i * a.prow;
#pragma unroll
    for (int v = 0; v < NDP; ++v) pp[v] = pr[v];
#pragma unroll
    for (int v = 0; v < NSV; ++v) gf[v] = (a.noise != NOISE_NONE) ? pr[NDP + NCP + v] : R(0);
  }
  if (a.noise == NOISE_INJECTED) {
    // padded lanes (t >= n_inst) read lane 0's noise: never past the array
    const double* z = a.z + td.z_off + k * td.z_sk + (int64_t)i * td.z_si + (live ? t : 0);
#pragma unroll
    for (int v = 0; v < NSV; ++v) xi[v] = (R)((a.sqrt_dt * (zpre ? zpre[v] : z[v * td.z_ssv])) * (double)gf[v]);
  } else if (a.noise == NOISE_PHILOX) {
    // normals in the history precision NR (float for the float32-history
    // modes), the increment in the state precision
    NR zz[NSV];
    philox_normals<NR>(NSV, (uint32_t)i, stream, s, a.key0, a.key1, zz);
#pragma unroll
    for (int v = 0; v < NSV; ++v) xi[v] = ((R)a.sqrt_dt * (R)zz[v]) * gf[v];
  } else {
#pragma unroll
    for (int v = 0; v < NSV; ++v) xi[v] = R(0);
  }
  if constexpr (EXT) {
    if (a.stim_tp) {   // simulator.py:300-306: spatial pattern x temporal value of this step
      const R sv = (R)(a.stim_sp[i] * a.stim_tp[k]);
      R st[NSV];
#pragma unroll
      for (int v = 0; v < NSV; ++v) st[v] = ((a.stim_mask >> v) & 1u) ? sv : R(0);
      scheme_step<R, M, true>(a.scheme, x, c, pp, dt, xi, st);
    } else {
      scheme_step<R, M>(a.scheme, x, c, pp, dt, xi);
    }
  } else {
    scheme_step<R, M>(a.scheme, x, c, pp, dt, xi);
  }
#pragma unroll
  for (int v = 0; v < NSV; ++v) {
    bad |= !isfinite(x[v]);
    s_x[(li * NSV + v) * T + t] = x[v];
  }
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) sink(ch, C::template channel<R>(x, ch, cp));

  monitors_node<R, M, CF, EXT>(a, kc, smem, i, li, t, T, N, live, k, sir, x, c, false);
}

// RH: history / weight / coupling-sum type, R: state, integration and
// monitor type (float/float, double/double, or float/double -- the
// reference's own mix).  Float64 histories get up to 8 waves per workgroup
// and twice the registers.
// PIPE = 1: the pipelined loop only (wave-uniform rows, every delay >= 1,
// n_channel * T <= 64); PIPE = 0: the sequential loop with every gather
// variant; PIPE = 2: the time-blocked sequential loop (blocks of up to
// d_min steps, no afferent monitor); PIPE = 3: the pipelined loop with two
// steps per pass (rows of 32 values, every delay >= 3).  Separate
// instantiations keep each
// loop's register allocation to itself.  UM: every model parameter is a plan constant (a coupling or noise
// sweep) and is read from the kernel arguments.
template <class RH, class R, class M, int CF, int PIPE, bool UM>
__global__ void __launch_bounds__(sizeof(RH) == 8 ? 512 : 1024) k_step(StepArgs a) {
  using C = Cfun<M, CF>;
  constexpr int NSV = M::NSV, NCH = C::NCH, NCT = C::NCT, NDP = M::NDP, NCP = C::NCP;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ KCfg kc;

  if (a.tiles[blockIdx.x].n_steps <= a.k0) return;   // uniform: whole block leaves
  // scalars of this tile and its connectome: loaded straight from global
  // memory (uniform addresses -> scalar loads, values stay in SGPRs)
  const TileDesc* tdg = a.tiles + blockIdx.x;
  const int conn = tdg->conn;
  const ConnDesc* cdg = a.conns + conn;
  const int N = cdg->n_node, H = cdg->horizon;
  const int cd_max_row = cdg->max_row;
  const int64_t cd_e_off = cdg->e_off, cd_rp_off = cdg->rp_off;
  const int td_inst0 = tdg->inst0, td_n_inst = tdg->n_inst;
  const int64_t td_step0 = tdg->step0, td_n_steps = tdg->n_steps, td_ring_off = tdg->ring_off,
                td_img_off = tdg->img_off;
  const int T = a.T;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int lane = tid & 63, wave = tid >> 6, nwave = nthr >> 6;
  const int nsteps = (int)td_n_steps;                            // steps of the whole call
  // this launch: steps [kbeg, kfin) of the call, target nodes [node0, node1)
  const int kbeg = a.k0, kfin = nsteps < a.k1 ? nsteps : a.k1;
  const int node0 = tdg->node0, node1 = tdg->node1;
  const bool staged = tdg->parts > 1;
  const int64_t stage_off = tdg->stage_off;
  const int rstride = NCH * T;                                   // elements per (slot, node)
  const uint32_t cstride = (uint32_t)rstride * (uint32_t)sizeof(RH);  // bytes per node in a slot
  const uint32_t sstride = (uint32_t)N * cstride;                // bytes per slot (< 2^24)
  // byte strides of the gather (ring_index): slot, source node, channel;
  // single-instance tiles keep a node's slots together
  const bool tmaj = T == 1;
  const uint32_t g_sl = tmaj ? (uint32_t)sizeof(RH) : sstride;
  const uint32_t g_nd = tmaj ? (uint32_t)(NCH * H) * (uint32_t)sizeof(RH) : cstride;
  const uint32_t g_ch = tmaj ? (uint32_t)H * (uint32_t)sizeof(RH) : (uint32_t)T * (uint32_t)sizeof(RH);
  const int64_t w_ch = tmaj ? H : T;                             // elements between channels

  if (tid == 0) {
    kc.td = *tdg;
    for (int m = 0; m < TVBK_KMAX_MON; ++m) {
      MonArgs mo = a.mon[m];
      if (m < a.n_mon) {
        const int64_t p0 = td_step0 - (mo.src ? 1 : 0);   // step of the call's first sample
        mo.r0 = (int)(p0 % mo.istep);
        mo.qe0 = p0 / mo.istep;
        mo.rb0 = (int)(p0 % mo.ii);
        mo.qb0 = p0 / mo.ii;
        step_info(mo, kbeg, kc.si[kbeg & 1][m]);
      }
      kc.mon[m] = mo;
    }
    for (int o = 0; o < TVBK_KMAX_OBS; ++o) {
      kc.obs_sv[o] = a.obs_sv[o];
      for (int u = 0; u < 8; ++u) kc.obs_coef[o][u] = a.obs_coef[o][u];
    }
  }
  const TileDesc& td = kc.td;

  // ---- LDS carve ---------------------------------------------------------
  unsigned char* img = smem;                              // persistent image
  R* s_x = reinterpret_cast<R*>(smem + a.lds_state);
  RH* s_c = reinterpret_cast<RH*>(smem + a.lds_c);     // raw coupling sums
  const uint32_t* g_edge = a.edges + cd_e_off;
  const RH* g_w = reinterpret_cast<const RH*>(a.weights) + cd_e_off;
  const int32_t* g_rp = a.row_ptr + cd_rp_off;
  EdgeRec* l_rec = reinterpret_cast<EdgeRec*>(smem + a.lds_edge);
  int32_t* l_rp = reinterpret_cast<int32_t*>(smem + a.lds_rp);
  // the tile's rows only (a split tile holds a node range): records of
  // edges ebase .. of its rows plus the padding
  const int ebase = a.csr_in_lds ? g_rp[node0] : 0;
  if (a.csr_in_lds) {
    const int ecnt = g_rp[node1] - ebase + TVBK_ROW_PAD;
    for (int k = tid; k < ecnt; k += nthr) {
      const uint32_t ed = g_edge[ebase + k];
      EdgeRec r;
      r.coff = (ed & 0xffffu) * g_nd;
      r.delay = ed >> 16;
      if constexpr (sizeof(RH) == 4) {
        r.w0 = __float_as_uint((float)g_w[ebase + k]);
        r.w1 = 0;
      } else {
        const double wd = (double)g_w[ebase + k];
        r.w0 = (uint32_t)__double2loint(wd);
        r.w1 = (uint32_t)__double2hiint(wd);
      }
      l_rec[k] = r;
    }
    for (int k = tid; k <= N; k += nthr) l_rp[k] = g_rp[k];
  }
  stage_params<R>(a, smem, td_inst0, td_n_inst, N);
  // the tile image's regions (node-major), this tile's node range: from
  // the HBM image (every node) to the LDS image (the range from position 0)
  for (int r = 0; r < a.n_img_reg; ++r) {
    const int64_t nw = (int64_t)(node1 - node0) * a.img_reg_bpn[r] / 4;
    const uint32_t* sr = reinterpret_cast<const uint32_t*>(a.img + td_img_off + a.img_reg_off[r] +
                                                           (int64_t)node0 * a.img_reg_bpn[r]);
    uint32_t* d = reinterpret_cast<uint32_t*>(img + a.img_reg_lds[r]);
    for (int64_t q = tid; q < nw; q += nthr) d[q] = sr[q];
  }
  __syncthreads();

  // ---- per-lane instance parameters (phase B lane and phase A lane agree:
  //      blockDim and 64 are multiples of T) -------------------------------
  const int t = tid % T;
  const bool live = t < td_n_inst;
  const int inst = td_inst0 + (live ? t : 0);
  // this lane's parameters: float64 states read them from LDS (registers are
  // short there), float32 states keep them in registers
  constexpr bool PAR_LDS = sizeof(R) == 8;
  constexpr int NPAR = NDP + NCP + NSV;
  R preg[PAR_LDS ? 1 : NPAR];
  if constexpr (!PAR_LDS) {
    const R* prow = reinterpret_cast<const R*>(a.params) + (int64_t)inst * (a.spatial ? (int64_t)N * a.prow : a.prow);
#pragma unroll
    for (int k = 0; k < NPAR; ++k) preg[k] = prow[k];
  }
  const R* par = PAR_LDS ? reinterpret_cast<const R*>(smem + a.lds_par) + t : preg;
  const int pst = PAR_LDS ? T : 1;
  const R dt = (R)a.dt;
  const uint64_t stream = a.streams ? a.streams[inst] : (uint64_t)inst;
  RH* ring = reinterpret_cast<RH*>(a.ring) + td_ring_off;

  // ---- step counters (wave-uniform) --------------------------------------
  int64_t s = td_step0 + kbeg;
  int smod = (int)((s - 1) % H);          // (s-1) % H ; s >= 1
  int bad = 0;
  const int P = N * NCH;
  const int NG = 64 / T;        // lane groups per wave
  const int g = lane / T;

  // ---- phase A: per (node, channel) pair, a group of T lanes reduces the
  // node's CSR row; the history is gathered with buffer loads at byte offset
  // slot*sstride + col*cstride + lane_off.  Instantiated once for the LDS and
  // once for the global CSR so every access has a known address space.
  constexpr int RMAX = sizeof(RH) == 4 ? 32 : 16;
  constexpr int WPE = sizeof(RH) / 4;                   // 32-bit words per weight
  using RMc = std::integral_constant<int, RMAX>;
  // the pipelined loop keeps two rows of RP registers in flight
  constexpr int RP = sizeof(RH) == 4 ? 24 : 12;
  using RPc = std::integral_constant<int, RP>;
  // (the two-step pipelined loop's rings hold a mirror of slot 0 at slot H)
  const __amdgpu_buffer_rsrc_t rsrc = make_rsrc(ring, (uint32_t)(PIPE == 3 ? H + 1 : H) * sstride);
  const int rmax = cd_max_row < RMAX ? cd_max_row : RMAX;
  // byte offset of the slot `delay` steps before the step whose (s - 1) % H
  // is sm
  auto slot_off = [&](int sm, uint32_t delay) -> uint32_t {
    int sl = sm - (int)delay;
    sl += (sl < 0) ? H : 0;
    return __umul24((uint32_t)sl, g_sl);
  };
  // phase A of nk consecutive steps, the first with (s - 1) % H = smod
  // (nk <= d_min: every slot read was written before the first of them):
  // items (node-channel pair q, step j), row-major so that the lanes of a
  // row read consecutive slots of each edge; step j's sums into
  // cbuf0 + j * cstep
  // LDS arrays indexed by node hold the tile's node range [node0, node1):
  // node i sits at position i - node0 of a.lds_nodes
  const int64_t cstep = (int64_t)a.lds_nodes * NCT * T;
  auto phase_a = [&](auto getrec, const int32_t* RP, RH* cbuf0, int nk) {
    const int q0 = node0 * NCH, P = (node1 - node0) * NCH, tot = nk * P;
    for (int it = wave * NG + g; it - g < tot; it += nwave * NG) {
      if (it < tot) {
        const int qq = it / nk, j = it - qq * nk, q = q0 + qq;
        const int sm = smod + j < H ? smod + j : smod + j - H;
        RH* cbuf = cbuf0 + j * cstep;
        const int i = q / NCH, ch = q - (q / NCH) * NCH;
        const int e0 = RP[i];
        const int n = RP[i + 1] - e0;
        const uint32_t lane_off = (uint32_t)ch * g_ch + (uint32_t)t * (uint32_t)sizeof(RH);
        RH acc;
        // one path per wave (both give the same sums): a wave whose rows
        // are not all short takes the blocked loads for every row
        if (__all(n <= RMAX)) {
          // three passes: offsets, all loads back to back, then the weights
          // -- no load result is used before the row's last load issues
          // (float32 keeps the weights in registers; float64 re-reads them)
          constexpr bool KEEPW = sizeof(RH) == 4;
          uint32_t off[RMAX];
          RH wv[KEEPW ? RMAX : 1];
#pragma unroll
          for (int e = 0; e < RMAX; ++e) {
            if (e < rmax) {
              const EdgeRec r = getrec(e0 + e);
              off[e] = (e < n) ? slot_off(sm, r.delay) + r.coff + lane_off : 0xFFFFFFF0u;
              if constexpr (KEEPW) wv[e] = rec_weight<RH>(r);
            } else {
              off[e] = 0xFFFFFFF0u;
              if constexpr (KEEPW) wv[e] = RH(0);
            }
          }
          RH v[RMAX];
#pragma unroll
          for (int e = 0; e < RMAX; ++e) v[e] = (e < rmax) ? buf_load<RH>(rsrc, off[e]) : RH(0);
#pragma unroll
          for (int e = 0; e < RMAX; ++e) {
            if constexpr (KEEPW) v[e] = wv[e] * v[e];
            else v[e] = (e < rmax ? rec_weight<RH>(getrec(e0 + e)) : RH(0)) * v[e];
          }
          acc = reduceat_regs<RMAX, RH>(n, v);
        } else {
          // long rows: the same order in blocks of eight loads
          auto ld = [&](int e) -> RH {
            const EdgeRec r = getrec(e0 + e);
            return buf_load<RH>(rsrc, e < n ? slot_off(sm, r.delay) + r.coff + lane_off : 0xFFFFFFF0u);
          };
          auto wt = [&](int e) -> RH { return rec_weight<RH>(getrec(e0 + e)); };
          acc = reduceat_row_blocked<RH, true>(n, ld, wt);
        }
        // the raw sum; the coupling's post runs in the state precision in
        // phase B (coupling.py:166-167: post of the float32 sum in float64)
        if (CF == CF_SIGMOID_JR) {
          cbuf[((i - node0) * NCT) * T + t] = acc;
        } else {
          cbuf[((i - node0) * NCT + ch) * T + t] = acc;
        }
      }
    }
  };
  // wave-uniform rows (NCH * T a multiple of 32, one shared connectome): a
  // wave reduces all channels and instances of one target node; the row's
  // edge records come through the scalar cache and each edge's history
  // offset is a scalar buffer offset, so an edge costs each lane one load,
  // one multiply and one add
  // padded row layout of the one shared connectome
  ScalarU32* c_ofs = (ScalarU32*)a.sofs;                                  // {ds, coff} per entry
  ScalarU32* c_w = (ScalarU32*)a.swts;                                    // weight words per entry
  ScalarInt* c_prp = (ScalarInt*)a.srp;                                   // padded row starts
  ScalarInt* c_rp = (ScalarInt*)(a.row_ptr + cd_rp_off);
  const uint32_t Hs = (uint32_t)H * sstride;
  const uint32_t hs = __builtin_amdgcn_readfirstlane(Hs);
  const int QN = NCH * T;                               // lanes per target node
  // history byte offset of edge {ds, coff} for slot base `base` ((s-1) % H
  // times the slot bytes) and baseh = base + H * slot bytes: the slot
  // difference wraps at H, min over unsigned picks the in-range one
  auto soff = [&](uint32_t base, uint32_t baseh, uint32_t ds, uint32_t coff) -> uint32_t {
    const uint32_t o1 = base - ds, o2 = baseh - ds;
    return (o1 < o2 ? o1 : o2) + coff;
  };
  // issue the loads of row i (n <= RMAX) in blocks of four edges.  The
  // uniform-path arrays pad every row to a multiple of four with copies of
  // the row's first edge (re-read, a cache hit, never summed), from row
  // start c_prp[i]; registers of blocks past the row stay unset and unread.
  // Returns n; rows longer than RMAX load nothing here.
  // rm: std::integral_constant with the register row length RM
  auto row_issue = [&](auto rm, int i, uint32_t base, uint32_t lane_off, RH* v, int& e0) -> int {
    constexpr int RM = decltype(rm)::value;
    e0 = __builtin_amdgcn_readfirstlane(c_prp[i]);
    const int n = __builtin_amdgcn_readfirstlane(c_rp[i + 1]) - __builtin_amdgcn_readfirstlane(c_rp[i]);
    if (n <= RM) {
      const uint32_t baseh = base + hs;
      ScalarChunk* q = (ScalarChunk*)(c_ofs + 2 * e0);
#pragma unroll
      for (int b = 0; b < RM / 4; ++b) {
        if (4 * b < n) {
          const u32x8a r = q[b];
#pragma unroll
          for (int k = 0; k < 4; ++k)
            v[4 * b + k] = buf_load2<RH>(rsrc, lane_off,
                                         __builtin_amdgcn_readfirstlane(soff(base, baseh, r[2 * k], r[2 * k + 1])));
        }
      }
    }
    return n;
  };
  // weights and the numpy reduceat order; long rows are loaded here
  auto row_finish = [&](auto rm, int e0, int n, uint32_t base, uint32_t lane_off, RH* v) -> RH {
    constexpr int RM = decltype(rm)::value;
    constexpr int EPC = 8 / WPE;                          // weights per 32-byte chunk
    if (n <= RM) {
      // weights in 32-byte scalar chunks; the products are formed inside the
      // reduction, on entries < n only
      ScalarChunk* wq = (ScalarChunk*)(c_w + WPE * e0);
      RH w[RM];
#pragma unroll
      for (int c = 0; c < RM / EPC; ++c) {
        if (c * EPC < n) {
          const u32x8a wc = wq[c];
#pragma unroll
          for (int j = 0; j < EPC; ++j) w[c * EPC + j] = chunk_weight<RH>(wc, j);
        }
      }
      return reduceat_regs_w<RM, RH>(n, v, w);
    }
    const uint32_t baseh = base + hs;
    auto ld = [&](int e) -> RH {
      if (e >= n) return RH(0);
      const uint32_t so = soff(base, baseh, c_ofs[2 * (e0 + e)], c_ofs[2 * (e0 + e) + 1]);
      return buf_load2<RH>(rsrc, lane_off, __builtin_amdgcn_readfirstlane(so));
    };
    auto wt = [&](int e) -> RH {
      if constexpr (WPE == 1) return __uint_as_float(c_w[e0 + e]);
      else return __hiloint2double((int)c_w[2 * (e0 + e) + 1], (int)c_w[2 * (e0 + e)]);
    };
    return reduceat_row_blocked<RH>(n, ld, wt);
  };
  auto row_store = [&](RH* cbuf, int i, int q, RH acc) {
    if (q < QN) {
      const int ch = q / T, tq = q - (q / T) * T;
      cbuf[((i - node0) * NCT + (CF == CF_SIGMOID_JR ? 0 : ch)) * T + tq] = acc;
    }
  };
  // the same for wave-uniform rows: items (step kk, node i), one per wave
  auto phase_a_uni = [&](RH* cbuf0, int nk) {
    const int nn = node1 - node0;
    for (int j = 0; j * 64 < QN; ++j) {                 // uniform trip count
      const int q = j * 64 + lane;
      // lanes past the node's values read lane 0's (discarded)
      const uint32_t lane_off = q < QN ? (uint32_t)q * (uint32_t)sizeof(RH) : 0u;
      for (int it = __builtin_amdgcn_readfirstlane(wave); it < nk * nn; it += nwave) {   // uniform: scalar loads
        const int kk = it / nn, i = node0 + (it - kk * nn);
        const int sm = smod + kk < H ? smod + kk : smod + kk - H;
        const uint32_t base = __builtin_amdgcn_readfirstlane((uint32_t)sm * sstride);   // slot (s - 1 + kk) % H
        RH v[RMAX];
        int e0;
        const int n = row_issue(RMc{}, i, base, lane_off, v, e0);
        row_store(cbuf0 + kk * cstep, i, q, row_finish(RMc{}, e0, n, base, lane_off, v));
      }
    }
  };
  auto rec_lds = [&](int e) -> EdgeRec { return l_rec[e - ebase]; };
  auto rec_glb = [&](int e) -> EdgeRec {
    const uint32_t ed = g_edge[e];
    EdgeRec r;
    r.coff = (ed & 0xffffu) * g_nd;
    r.delay = ed >> 16;
    if constexpr (sizeof(RH) == 4) {
      r.w0 = __float_as_uint((float)g_w[e]);
      r.w1 = 0;
    } else {
      const double wd = (double)g_w[e];
      r.w0 = (uint32_t)__double2loint(wd);
      r.w1 = (uint32_t)__double2hiint(wd);
    }
    return r;
  };

  // one (node i, lane t) update of phase B for step s = s0 + k
  // With `stage`, the new history channels go to this item's entries of the
  // coupling buffer it has just read (NCT == NCH) instead of the ring, and
  // ring_out() stores them later.
  auto b_item = [&](int i, RH* cbuf, int k, int64_t s_, const StepInfo* sir, int wslot, bool stage,
                    const double* zpre) {
    R c[C::NCD];
    R cpv[NCP > 0 ? NCP : 1];
#pragma unroll
    for (int v = 0; v < NCP; ++v) cpv[v] = lane_param<R>(a, par, pst, NDP + v);
    const int li = i - node0;
#pragma unroll
    for (int v = 0; v < NCT; ++v) c[v] = C::post((R)cbuf[(li * NCT + v) * T + t], cpv);
#pragma unroll
    for (int v = NCT; v < C::NCD; ++v) c[v] = C::post(R(0), cpv);
    RH* wr = staged ? reinterpret_cast<RH*>(a.stage) + stage_off + ((int64_t)(k - kbeg) * N + i) * rstride + t
                    : ring + ring_index(T, H, N, NCH, wslot, i, 0, t);
    const int64_t wch = staged ? T : w_ch;
    RH* st = cbuf + (li * NCT) * T + t;
    update_node<R, M, CF, RH, UM, PIPE != 1 && PIPE != 3>(a, kc, smem, s_x, i, li, t, T, N, live, inst, stream, k, s_, sir, par,
                                             pst, c, zpre, bad,
                                  [&](int ch, R v) {
                                    if (stage) st[ch * T] = (RH)v;
                                    else wr[ch * wch] = (RH)v;
                                  });
  };
  // store this thread's staged history channels of the step into slot wslot
  auto ring_out = [&](const RH* cbuf, int wslot) {
    for (int i = tid / T; i < N; i += nthr / T) {
      RH* wr = ring + (wslot * N + i) * rstride + t;
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch) wr[ch * T] = cbuf[(i * NCT + ch) * T + t];
    }
  };

  // ---- pipelined loop (uniform rows, one wave pass per node, every delay
  // >= 1): the coupling of step s+1 reads slots <= s-1, written before the
  // previous barrier, so each wave interleaves its rows of phase A(s+1) with
  // its updates of phase B(s) -- a row's loads are in flight while the wave
  // integrates -- and one barrier per step orders both against the next.
  // rows of this wave in the pipelined loops: a contiguous range (the
  // scalar records of a wave's rows share cache lines).  The first E waves
  // carry one more integration pass than the others (when the passes do
  // not divide the nodes); they take a.row_shift rows fewer than
  // N / nwave, the others share the rest
  auto row_range = [&](int npass, int& r0, int& nrow) {
    const int npw = 64 / T;                              // nodes per wave and pass
    const int E = (N - (npass - 1) * (nthr / T) + npw - 1) / npw;
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    if (E > 0 && E < nwave) {
      const int re0 = N / nwave - a.row_shift, re = re0 > 0 ? re0 : 0;
      const int rest = N - E * re, q = nwave - E, bq = rest / q, xq = rest - (rest / q) * q;
      if (wv < E) {
        r0 = wv * re;
        nrow = re;
      } else {
        const int ix = wv - E;
        nrow = bq + (ix < xq ? 1 : 0);
        r0 = E * re + ix * bq + (ix < xq ? ix : xq);
      }
    } else {
      const int bq = N / nwave, xq = N - (N / nwave) * nwave;
      nrow = bq + (wv < xq ? 1 : 0);
      r0 = wv * bq + (wv < xq ? wv : xq);
    }
  };
  if constexpr (PIPE == 1) {
    phase_a_uni(s_c, 1);                                // coupling of the first step
    __syncthreads();
    const int npass = (N + nthr / T - 1) / (nthr / T);  // phase B passes
    int r0, nrow;
    row_range(npass, r0, nrow);
    const int nu = nrow > npass ? nrow : npass;
    const uint32_t lane_off = lane < QN ? (uint32_t)lane * (uint32_t)sizeof(RH) : 0u;
    for (int k = 0; k < nsteps; ++k, ++s) {
      const int buf = k & 1;
      RH* cbuf = s_c + (int64_t)buf * a.lds_nodes * NCT * T;
      RH* cnext = s_c + (int64_t)(buf ^ 1) * a.lds_nodes * NCT * T;
      const int wslot = (smod + 1 == H) ? 0 : smod + 1;   // s % H
      const uint32_t base_n = __builtin_amdgcn_readfirstlane((uint32_t)wslot * sstride);   // (s+1-1) % H
      // history stores wait until the step's gathers are reduced: a wave's
      // memory counter retires in issue order, so a store issued between a
      // row's loads and their use would delay the use until it completes
      // rows u and u+1 in flight: row u+1 is issued, then this wave's
      // update u runs, then row u is reduced; va / vb alternate
      RH va[RP], vb[RP];
      int ea = 0, na = 0, eb = 0, nb = 0;
      if (nrow > 0) na = row_issue(RPc{}, r0, base_n, lane_off, va, ea);
      auto body = [&](int u, RH* vf, int ef, int nf, RH* vi, int& ei, int& ni) {
        if (u + 1 < nrow) ni = row_issue(RPc{}, r0 + u + 1, base_n, lane_off, vi, ei);
        const int ib = tid / T + u * (nthr / T);
        if (ib < N) b_item(ib, cbuf, k, s, kc.si[buf], wslot, NCT == NCH, nullptr);
        if (u < nrow) row_store(cnext, r0 + u, lane, row_finish(RPc{}, ef, nf, base_n, lane_off, vf));
      };
      for (int u = 0; u < nu; u += 2) {
        body(u, va, ea, na, vb, eb, nb);
        if (u + 1 < nu) body(u + 1, vb, eb, nb, va, ea, na);
      }
      if (NCT == NCH) ring_out(cbuf, wslot);
      if (tid == 0)
        for (int m = 0; m < a.n_mon; ++m) step_info(kc.mon[m], k + 1, kc.si[buf ^ 1][m]);
      smod = wslot;
      __syncthreads();
    }
  } else if constexpr (PIPE == 3) {
    // ---- two-step pipelined loop (uniform rows of QN = 32 values, every
    // delay >= 3): a pass integrates steps s and s+1 while it gathers the
    // coupling of steps s+2 and s+3, which read slots <= s-1 only.  One wave
    // load serves both gathered steps of an edge: lanes 0..31 read the
    // edge's slot for step s+2, lanes 32..63 the next slot (step s+3), so
    // every lane carries a value, each edge's scalar offset serves two steps
    // and the bytes in flight per row double.  The slot after H-1 is slot 0:
    // the ring holds a mirror of slot 0 at slot H, refreshed here and
    // whenever slot 0 is written.  Coupling buffers: two sets of two steps,
    // alternating by pass; monitor step records [4][TVBK_KMAX_MON] in LDS.
    StepInfo* si4 = reinterpret_cast<StepInfo*>(smem + a.lds_si);
    RH* ring_h = ring + (int64_t)H * N * rstride;        // the mirror slot
    for (int q = tid; q < N * rstride; q += nthr) ring_h[q] = ring[q];
    if (tid == 0)
      for (int m = 0; m < a.n_mon; ++m) {
        si4[m] = kc.si[0][m];                           // step 0 (kbeg = 0)
        step_info(kc.mon[m], 1, si4[TVBK_KMAX_MON + m]);
      }
    phase_a_uni(s_c, 2);                                // coupling of steps 0 and 1
    __syncthreads();
    const int npass = (N + nthr / T - 1) / (nthr / T);  // integration passes
    int r0, nrow;
    row_range(npass, r0, nrow);
    const int nu = nrow > npass ? nrow : npass;
    const int hl = lane >> 5, ql = lane & 31;           // step half, value of the row
    const uint32_t lane_off = (uint32_t)ql * (uint32_t)sizeof(RH) + (uint32_t)hl * sstride;
    // staged history of one step into slot wslot (and the mirror)
    auto ring_out2 = [&](const RH* cbuf, int wslot) {
      for (int i = tid / T; i < N; i += nthr / T) {
        RH* wr = ring + (wslot * N + i) * rstride + t;
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
          const RH v = cbuf[(i * NCT + ch) * T + t];
          wr[ch * T] = v;
          if (wslot == 0) wr[(int64_t)H * N * rstride + ch * T] = v;
        }
      }
    };
    for (int k = 0; k < nsteps; k += 2, s += 2) {
      const int set = (k >> 1) & 1;
      RH* cb = s_c + (int64_t)set * 2 * cstep;           // steps s, s+1
      RH* cn = s_c + (int64_t)(set ^ 1) * 2 * cstep;     // steps s+2, s+3
      const StepInfo* sir = si4 + set * 2 * TVBK_KMAX_MON;
      const bool two = k + 1 < nsteps;                   // uniform
      const int ws0 = (smod + 1 == H) ? 0 : smod + 1;     // s % H
      const int ws1 = (ws0 + 1 == H) ? 0 : ws0 + 1;       // (s+1) % H
      const uint32_t base_n = __builtin_amdgcn_readfirstlane((uint32_t)ws1 * sstride);   // (s+2-1) % H
      RH va[RP], vb[RP];
      int ea = 0, na = 0, eb = 0, nb = 0;
      if (nrow > 0) na = row_issue(RPc{}, r0, base_n, lane_off, va, ea);
      auto body = [&](int u, RH* vf, int ef, int nf, RH* vi, int& ei, int& ni) {
        if (u + 1 < nrow) ni = row_issue(RPc{}, r0 + u + 1, base_n, lane_off, vi, ei);
        const int ib = tid / T + u * (nthr / T);
        if (ib < N) {
          b_item(ib, cb, k, s, sir, ws0, true, nullptr);
          if (two) b_item(ib, cb + cstep, k + 1, s + 1, sir + TVBK_KMAX_MON, ws1, true, nullptr);
        }
        if (u < nrow) {
          const RH acc = row_finish(RPc{}, ef, nf, base_n, lane_off, vf);
          const int i = r0 + u, ch = ql / T, tq = ql - (ql / T) * T;
          cn[hl * cstep + ((i - node0) * NCT + ch) * T + tq] = acc;
        }
      };
      for (int u = 0; u < nu; u += 2) {
        body(u, va, ea, na, vb, eb, nb);
        if (u + 1 < nu) body(u + 1, vb, eb, nb, va, ea, na);
      }
      ring_out2(cb, ws0);
      if (two) ring_out2(cb + cstep, ws1);
      if (tid == 0)
        for (int m = 0; m < a.n_mon; ++m) {
          step_info(kc.mon[m], k + 2, si4[(set ^ 1) * 2 * TVBK_KMAX_MON + m]);
          step_info(kc.mon[m], k + 3, si4[((set ^ 1) * 2 + 1) * TVBK_KMAX_MON + m]);
        }
      smod = two ? ws1 : ws0;
      __syncthreads();
    }
  } else if constexpr (PIPE == 2) {
    // ---- time-blocked sequential loop: a block of nk <= d_min steps reads
    // only history written before the block, so phase A gathers the
    // coupling of all nk steps at once; then each thread integrates its
    // nodes through the nk steps (a node's update needs only its own state
    // and coupling).  Two barriers per block; the monitors' step records of
    // the block are kept in LDS.
    StepInfo* si_blk = reinterpret_cast<StepInfo*>(smem + a.lds_si);
    // blocks never exceed H steps (the plan keeps them <= d_min < H)
    const int KT = a.tb_k < H ? a.tb_k : H;
    for (int kb = kbeg; kb < kfin; kb += KT) {
      const int nk = kfin - kb < KT ? kfin - kb : KT;
      // one thread per (step, monitor) record
      for (int q = tid; q < nk * a.n_mon; q += nthr) {
        const int j = q / a.n_mon, m = q - j * a.n_mon;
        step_info(kc.mon[m], kb + j, si_blk[j * TVBK_KMAX_MON + m]);
      }
      if (a.uniform_rows) phase_a_uni(s_c, nk);
      else if (a.csr_in_lds) phase_a(rec_lds, l_rp, s_c, nk);
      else phase_a(rec_glb, g_rp, s_c, nk);
      __syncthreads();
      // injected normals (models of up to four state variables): the next
      // step's are loaded before the current step is integrated, so their
      // latency hides behind the update
      constexpr bool ZPF = NSV <= 4;
      const bool zpf = ZPF && a.noise == NOISE_INJECTED;
      for (int i = node0 + tid / T; i < node1; i += nthr / T) {
        double zn[ZPF ? NSV : 1] = {}, zc[ZPF ? NSV : 1];
        auto zload = [&](int k) {
          if constexpr (ZPF) {
            const double* z = a.z + td.z_off + (int64_t)k * td.z_sk + (int64_t)i * td.z_si + (live ? t : 0);
#pragma unroll
            for (int v = 0; v < NSV; ++v) zn[v] = z[v * td.z_ssv];
          }
        };
        if (zpf) zload(kb);
        int ws = smod;
        for (int j = 0; j < nk; ++j) {
#pragma unroll
          for (int v = 0; v < (ZPF ? NSV : 1); ++v) zc[v] = zn[v];
          if (zpf && j + 1 < nk) zload(kb + j + 1);
          ws = (ws + 1 == H) ? 0 : ws + 1;                // (s + j) % H
          b_item(i, s_c + j * cstep, kb + j, s + j, si_blk + j * TVBK_KMAX_MON, ws, false, zpf ? zc : nullptr);
        }
      }
      s += nk;
      smod = smod + nk < H ? smod + nk : smod + nk - H;
      __syncthreads();
    }
  } else {
  // afferent monitors sample, at step s, the coupling step s + 1 integrates
  // with (simulator.py:311-316): the call's last sample needs the coupling of
  // the step after it, gathered by one more phase A (the next call gathers it
  // again) followed by those monitors only
  const int kend = kfin + (a.has_aff ? 1 : 0);
  for (int k = kbeg; k < kend; ++k, ++s) {
    const int buf = k & 1;
    RH* cbuf = s_c + (int64_t)buf * a.lds_nodes * NCT * T;
    // ================= phase A: coupling of step s ========================
    if (a.uniform_rows) phase_a_uni(cbuf, 1);
    else if (a.csr_in_lds) phase_a(rec_lds, l_rp, cbuf, 1);
    else phase_a(rec_glb, g_rp, cbuf, 1);
    __syncthreads();
    if (k == kfin) {
      for (int i = node0 + tid / T; i < node1; i += nthr / T) {
        R c[C::NCD], x[NSV];
        R cpv[NCP > 0 ? NCP : 1];
#pragma unroll
        for (int v = 0; v < NCP; ++v) cpv[v] = lane_param<R>(a, par, pst, NDP + v);
        const int li = i - node0;
#pragma unroll
        for (int v = 0; v < NCT; ++v) c[v] = C::post((R)cbuf[(li * NCT + v) * T + t], cpv);
#pragma unroll
        for (int v = NCT; v < C::NCD; ++v) c[v] = C::post(R(0), cpv);
#pragma unroll
        for (int v = 0; v < NSV; ++v) x[v] = s_x[(li * NSV + v) * T + t];
        monitors_node<R, M, CF, true>(a, kc, smem, i, li, t, T, N, live, kfin, kc.si[buf], x, c, true);
      }
      break;
    }
    // ================= phase B: integrate, write history, monitors =========
    const int wslot = (smod + 1 == H) ? 0 : smod + 1;   // s % H
    for (int i = node0 + tid / T; i < node1; i += nthr / T) b_item(i, cbuf, k, s, kc.si[buf], wslot, false, nullptr);
    // bookkeeping of step k+1 into the other buffer (read after the next
    // phase A barrier; its previous readers finished before this step's)
    if (tid == 0)
      for (int m = 0; m < a.n_mon; ++m) step_info(kc.mon[m], k + 1, kc.si[buf ^ 1][m]);
    smod = wslot;
    if (a.zero_delay) __syncthreads();
  }
  }
  __syncthreads();
  // write back this tile's nodes of every region
  for (int r = 0; r < a.n_img_reg; ++r) {
    const int64_t nw = (int64_t)(node1 - node0) * a.img_reg_bpn[r] / 4;
    uint32_t* d = reinterpret_cast<uint32_t*>(a.img + td_img_off + a.img_reg_off[r] + (int64_t)node0 * a.img_reg_bpn[r]);
    const uint32_t* sr = reinterpret_cast<const uint32_t*>(img + a.img_reg_lds[r]);
    for (int64_t q = tid; q < nw; q += nthr) d[q] = sr[q];
  }
  if (bad && live) atomicOr(a.flags + inst, 1);
}

// Zero-delay coupling (all idelays 0, horizon 1) as a dense contraction on
// the matrix cores: per history channel ch, C = W X with W the N x N weights
// (float32, as the reference stores them) and X the channel values of the
// tile's 32 instances after the previous step.  v_mfma_f32_16x16x4f32:
// A (16 x 4): lane l holds A[l % 16][l / 16]; B (4 x 16): B[l / 16][l % 16];
// D (16 x 16): lane l holds D[4 (l / 16) + v][l % 16], v = 0..3.
// Wave w owns target rows 16 (w / 2) .. +15 and instances 16 (w % 2) .. +15
// of the tile; its W fragments stay in registers for the whole call.  X is
// double-buffered in LDS by step parity: one barrier per step.
typedef float f4_t __attribute__((ext_vector_type(4)));
#define TVBK_DENSE_KBMAX 32

template <class R, class M, int CF>
__global__ void __launch_bounds__(1024) k_dense(StepArgs a) {
  using RH = float;   // history channel values and weights; R: state
  using C = Cfun<M, CF>;
  constexpr int NSV = M::NSV, NCH = C::NCH, NDP = M::NDP, NCP = C::NCP;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ KCfg kc;

  if (a.tiles[blockIdx.x].n_steps <= 0) return;
  const TileDesc* tdg = a.tiles + blockIdx.x;
  const int conn = tdg->conn;
  const ConnDesc* cdg = a.conns + conn;
  const int N = cdg->n_node;
  const int td_inst0 = tdg->inst0, td_n_inst = tdg->n_inst;
  const int64_t td_step0 = tdg->step0, td_n_steps = tdg->n_steps, td_ring_off = tdg->ring_off,
                td_img_off = tdg->img_off;
  constexpr int T = 32;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int nsteps = (int)td_n_steps;
  const int KB = a.dense_kb, KP = 4 * a.dense_kb;
  const int rb = wave >> 1, cb = wave & 1;
  const int t = 16 * cb + (lane & 15);           // instance lane of the tile
  const int r0 = 16 * rb + 4 * (lane >> 4);      // first of the lane's 4 target rows

  if (tid == 0) {
    kc.td = *tdg;
    for (int m = 0; m < TVBK_KMAX_MON; ++m) {
      MonArgs mo = a.mon[m];
      if (m < a.n_mon) {
        mo.r0 = (int)(td_step0 % mo.istep);
        mo.qe0 = td_step0 / mo.istep;
        mo.rb0 = (int)(td_step0 % mo.ii);
        mo.qb0 = td_step0 / mo.ii;
        step_info(mo, 0, kc.si[0][m]);
      }
      kc.mon[m] = mo;
    }
    for (int o = 0; o < TVBK_KMAX_OBS; ++o) {
      kc.obs_sv[o] = a.obs_sv[o];
      for (int u = 0; u < 8; ++u) kc.obs_coef[o][u] = a.obs_coef[o][u];
    }
  }
  unsigned char* img = smem;
  R* s_x = reinterpret_cast<R*>(smem + a.lds_state);
  RH* X = reinterpret_cast<RH*>(smem + a.lds_x);   // [2][NCH][KP][T]
  RH* ring = reinterpret_cast<RH*>(a.ring) + td_ring_off;   // one slot: [node][ch][T]
  stage_params<R>(a, smem, td_inst0, td_n_inst, N);
  copy_bytes(img, a.img + td_img_off, a.img_bytes);
  for (int q = tid; q < 2 * NCH * KP * T; q += nthr) {
    const int tt = q % T, kk = (q / T) % KP, ch = (q / (T * KP)) % NCH, b = q / (T * KP * NCH);
    X[q] = (b == 0 && kk < N) ? ring[(kk * NCH + ch) * T + tt] : RH(0);
  }
  // W fragments of this wave's 16 rows
  float afr[TVBK_DENSE_KBMAX];
  const int arow = 16 * rb + (lane & 15);
#pragma unroll
  for (int kb = 0; kb < TVBK_DENSE_KBMAX; ++kb)
    afr[kb] = (kb < KB) ? a.dense_w[(int64_t)arow * KP + 4 * kb + (lane >> 4)] : 0.0f;
  __syncthreads();

  const bool live = t < td_n_inst;
  const int inst = td_inst0 + (live ? t : 0);
  constexpr bool PAR_LDS = sizeof(R) == 8;
  constexpr int NPAR = NDP + NCP + NSV;
  R preg[PAR_LDS ? 1 : NPAR];
  if constexpr (!PAR_LDS) {
    const R* prow = reinterpret_cast<const R*>(a.params) + (int64_t)inst * a.prow;
#pragma unroll
    for (int k = 0; k < NPAR; ++k) preg[k] = prow[k];
  }
  const R* par = PAR_LDS ? reinterpret_cast<const R*>(smem + a.lds_par) + t : preg;
  const int pst = PAR_LDS ? T : 1;
  const uint64_t stream = a.streams ? a.streams[inst] : (uint64_t)inst;
  int bad = 0;
  int64_t s = td_step0;
  const int krow = lane >> 4;                    // B operand row within a k block

  for (int k = 0; k < nsteps; ++k, ++s) {
    const int buf = k & 1;
    const RH* Xc = X + (int64_t)buf * NCH * KP * T;
    RH* Xn = X + (int64_t)(buf ^ 1) * NCH * KP * T;
    f4_t d[NCH];
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
      f4_t acc = {0.f, 0.f, 0.f, 0.f};
      const RH* xb = Xc + (int64_t)ch * KP * T + krow * T + t;
#pragma unroll
      for (int kb = 0; kb < TVBK_DENSE_KBMAX; ++kb)
        if (kb < KB) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(afr[kb], xb[4 * kb * T], acc, 0, 0, 0);
      d[ch] = acc;
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = r0 + v;
      if (i < N) {
        R c[NCH];
        R cpv[NCP > 0 ? NCP : 1];
#pragma unroll
        for (int u = 0; u < NCP; ++u) cpv[u] = lane_param<R>(a, par, pst, NDP + u);
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) c[ch] = C::post((R)d[ch][v], cpv);
        update_node<R, M, CF, RH>(a, kc, smem, s_x, i, i, t, T, N, live, inst, stream, k, s, kc.si[buf], par, pst, c,
                                  nullptr, bad,
                                  [&](int ch, R val) { Xn[((int64_t)ch * KP + i) * T + t] = (RH)val; });
      }
    }
    if (tid == 0)
      for (int m = 0; m < a.n_mon; ++m) step_info(kc.mon[m], k + 1, kc.si[buf ^ 1][m]);
    __syncthreads();
  }
  // the final channel values become the history's single slot
  const RH* Xf = X + (int64_t)(nsteps & 1) * NCH * KP * T;
  for (int q = tid; q < N * NCH * T; q += nthr) {
    const int tt = q % T, ch = (q / T) % NCH, i = q / (T * NCH);
    ring[(i * NCH + ch) * T + tt] = Xf[((int64_t)ch * KP + i) * T + tt];
  }
  copy_bytes(a.img + td_img_off, img, a.img_bytes);
  if (bad && live) atomicOr(a.flags + inst, 1);
}

// Fills the ring and state of every tile from Philox uniforms in [lo, hi),
// then bounds them (history.py:150-152).  The state is the value of slot
// current_step % H = 0.
template <class RH, class R, class M, int CF>
__global__ void k_init_uniform(InitArgs a) {
  using C = Cfun<M, CF>;
  constexpr int NSV = M::NSV, NCH = C::NCH, NDP = M::NDP;
  const TileDesc td = a.tiles[blockIdx.y];
  if (td.node0 != 0) return;                 // one tile of a split instance fills it
  const ConnDesc cd = a.conns[td.conn];
  const int T = a.T, N = cd.n_node, H = cd.horizon;
  const int64_t per_tile = (int64_t)H * N * T;
  RH* ring = reinterpret_cast<RH*>(a.ring) + td.ring_off;
  R* st = reinterpret_cast<R*>(a.img + td.img_off + a.lds_state);
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < per_tile; q += (int64_t)gridDim.x * blockDim.x) {
    const int t = (int)(q % T);
    const int i = (int)((q / T) % N);
    const int slot = (int)(q / ((int64_t)T * N));
    const int inst = td.inst0 + (t < td.n_inst ? t : 0);
    const uint64_t stream = a.streams ? a.streams[inst] : (uint64_t)inst;
    R x[NSV];
    for (int call = 0; call * 4 < NSV; ++call) {
      U4 c{(uint32_t)i | ((uint32_t)call << 20), (uint32_t)stream, (uint32_t)slot, 0x80000000u};
      U4 w = philox10(c, a.key0, a.key1);
      const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
      for (int u = 0; u < 4 && call * 4 + u < NSV; ++u) {
        const int v = call * 4 + u;
        const double uu = (double)ws[u] * 2.3283064365386963e-10;
        x[v] = (R)(a.lo[v] + (a.hi[v] - a.lo[v]) * uu);
      }
    }
    M::template bound<R>(x);
    const R* prow = reinterpret_cast<const R*>(a.params) + (int64_t)inst * a.prow;
    R cp[C::NCP > 0 ? C::NCP : 1];
    for (int k = 0; k < C::NCP; ++k) cp[k] = prow[NDP + k];
    for (int ch = 0; ch < NCH; ++ch) ring[ring_index(T, H, N, NCH, slot, i, ch, t)] = (RH)C::template channel<R>(x, ch, cp);
    if (slot == 0)
      for (int v = 0; v < NSV; ++v) st[(i * NSV + v) * T + t] = x[v];
  }
}

// ---------------------------------------------------------------------------
// launchers, explicitly instantiated per (precision, model, coupling) in
// tvbk_inst.hip so that the instantiations compile in parallel
// ---------------------------------------------------------------------------
template <class RH, class R, class M, int CF>
hipError_t launch_step_t(const StepArgs& a, int ntiles, int nthreads, size_t lds, hipStream_t st) {
  // the loop (step_loop); the pipelined one reads every model parameter from
  // the kernel arguments when all are plan constants (UM)
  const int loop = step_loop(a, Cfun<M, CF>::NCH);
  const bool pipe = loop == TVBK_LOOP_PIPELINED || loop == TVBK_LOOP_PIPELINED2;
  const uint64_t mmask = M::NDP >= 64 ? ~0ull : ((1ull << M::NDP) - 1);
  const bool um = pipe && (a.par_umask & mmask) == mmask && !a.spatial;
  // the two-step loop is built for models of up to two state variables
  // (the host plans it only for those)
  auto kern = pipe ? (um ? k_step<RH, R, M, CF, 1, true> : k_step<RH, R, M, CF, 1, false>)
                   : loop == TVBK_LOOP_TIME_BLOCKED ? k_step<RH, R, M, CF, 2, false> : k_step<RH, R, M, CF, 0, false>;
  if constexpr (M::NSV <= 2) {
    if (loop == TVBK_LOOP_PIPELINED2) kern = um ? k_step<RH, R, M, CF, 3, true> : k_step<RH, R, M, CF, 3, false>;
  }
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(ntiles), dim3(nthreads), lds, st, a);
  return hipGetLastError();
}

template <class RH, class R, class M, int CF>
hipError_t launch_init_t(const InitArgs& a, int ntiles, hipStream_t st) {
  hipLaunchKernelGGL((k_init_uniform<RH, R, M, CF>), dim3(256, ntiles), dim3(256), 0, st, a);
  return hipGetLastError();
}

template <class R, class M>
hipError_t launch_dense_t(const StepArgs& a, int ntiles, int nthreads, size_t lds, hipStream_t st) {
  auto kern = k_dense<R, M, CF_LINEAR>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(ntiles), dim3(nthreads), lds, st, a);
  return hipGetLastError();
}

}  // namespace tvbk
