// Host side of the C-ABI declared in include/tvbk.h: plan construction
// (validation, derived constants, connectome upload, buffer layout, kernel
// shape), state upload/export, step launches and monitor read-back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "tvbk.h"
#include "tvbk_kernels.hpp"

using namespace tvbk;

namespace {

thread_local std::string g_err;
// serialises the users of the static device scratch within this process
std::mutex g_scratch_mu;

struct ModelInfo {
  int nmp, nsv, ncv, ndp;
};

bool model_info(int model, ModelInfo* mi) {
  switch (model) {
    case TVBK_MODEL_G2D: *mi = {12, 2, 1, 13}; return true;
    case TVBK_MODEL_MPR: *mi = {8, 2, 2, 8}; return true;
    case TVBK_MODEL_JR: *mi = {15, 6, 2, 14}; return true;
    case TVBK_MODEL_SUM: *mi = {0, 1, 1, 0}; return true;
  }
  return false;
}

int cfun_nparam(int cf) {
  switch (cf) {
    case TVBK_CFUN_LINEAR: return 2;
    case TVBK_CFUN_SIGMOID_JR: return 5;
    case TVBK_CFUN_IDENTITY: return 0;
  }
  return -1;
}

bool combo_supported(int model, int cf) {
  if (cf == TVBK_CFUN_LINEAR) return true;
  if (cf == TVBK_CFUN_SIGMOID_JR) return model == TVBK_MODEL_JR;   // pre needs two cvars
  if (cf == TVBK_CFUN_IDENTITY) return model == TVBK_MODEL_SUM;
  return false;
}

int n_channel(int model, int cf) {
  ModelInfo mi;
  model_info(model, &mi);
  return cf == TVBK_CFUN_SIGMOID_JR ? 1 : mi.ncv;
}
int n_term(int model, int cf) { return n_channel(model, cf); }

// Derived constants the kernels use, evaluated in float64 in the order the
// reference evaluates them (see the model structs in tvbk_device.hpp).
void derive_model(int model, const double* p, double* out) {
  switch (model) {
    case TVBK_MODEL_G2D: {  // tau I a b c d e f g alpha beta gamma
      const double tau = p[0], I = p[1], a = p[2], b = p[3], c = p[4], d = p[5], e = p[6], f = p[7], g = p[8],
                   alpha = p[9], beta = p[10], gamma = p[11];
      const double o[13] = {d * tau, alpha, f, e, g, gamma * I, gamma, a, b, c, d, beta, tau};
      std::memcpy(out, o, sizeof(o));
      break;
    }
    case TVBK_MODEL_MPR: {  // tau I Delta J eta Gamma cr cv
      const double tau = p[0], I = p[1], Delta = p[2], J = p[3], eta = p[4], cr = p[6], cv = p[7];
      const double pi = 3.141592653589793;
      const double o[8] = {1 / tau, Delta / (pi * tau), (pi * pi) * (tau * tau), eta, J * tau, I, cr, cv};
      std::memcpy(out, o, sizeof(o));
      break;
    }
    case TVBK_MODEL_JR: {  // A B a b v0 nu_max r J a_1 a_2 a_3 a_4 p_min p_max mu
      const double A = p[0], B = p[1], a = p[2], b = p[3], v0 = p[4], nu = p[5], r = p[6], J = p[7], a1 = p[8],
                   a2 = p[9], a3 = p[10], a4 = p[11], mu = p[14];
      const double o[14] = {2.0 * nu, r,       v0,     a1 * J, a3 * J, A * a,   2.0 * a,
                            a * a,    mu,      a2 * J, B * b,  a4 * J, 2.0 * b, b * b};
      std::memcpy(out, o, sizeof(o));
      break;
    }
    default: break;
  }
}

void derive_cfun(int cf, const double* p, double* out) {
  if (cf == TVBK_CFUN_LINEAR) {
    out[0] = p[0];
    out[1] = p[1];
  } else if (cf == TVBK_CFUN_SIGMOID_JR) {  // cmin cmax midpoint r a
    out[0] = p[0];
    out[1] = p[1] - p[0];
    out[2] = p[3];
    out[3] = p[2];
    out[4] = p[4];
  }
}

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

}  // namespace

struct tvbk_plan {
  tvbk_plan_desc d;
  ModelInfo mi;
  int nch = 0, nct = 0, ncp = 0, prow = 0;
  // coupling function of the kernels: d.cfun, or CF_LINEAR1 when the plan
  // skips the coupling variables whose coefficient is zero
  int kcf = 0;
  int kernel = 0, T = 1, waves = 1, nthreads = 64;
  int esize = 4;    // state, parameters, Raw/TAvg outputs
  int hsize = 4;    // history ring, weights, coupling sums
  bool stochastic = false;
  hipStream_t stream = nullptr;
  std::string err;
  int max_n = 0;
  int64_t max_nnz = 0;
  int64_t csr_nnz = 0;   // edges a tile stages in LDS: the largest tile's rows
  int spatial = 0;
  // tiles
  std::vector<TileDesc> tiles;
  std::vector<int> tile_of, lane_of;  // per instance
  std::vector<int64_t> cur_step;      // per instance
  std::vector<int64_t> remaining;     // per instance (-1 = unlimited)
  std::vector<ConnDesc> conns;
  // LDS layout
  int lds_bold = 0, lds_state = 0, lds_c = 0, lds_edge = 0, lds_w = 0, lds_rp = 0, lds_par = 0, lds_total = 0;
  // time-blocked sequential loop: steps per block (0: one step at a time)
  // and the LDS offset of the block's monitor step records
  int tb_k = 0, lds_si = 0;
  // pipelined loop's steps per pass (2: TVBK_LOOP_PIPELINED2, whose rings
  // carry one mirror slot past the horizon)
  int pipe_k = 1;
  int row_shift = 0;   // two-step loop: rows fewer for the extra-pass waves
  int64_t img_bytes = 0;
  bool csr_in_lds = false;
  bool uniform_rows = false;   // batch-lane, n_channel * T a multiple of 32
  int last_loop = -1;          // TVBK_LOOP_* of the last run
  size_t srec_w_off = 0, srec_rp_off = 0;   // weights' and row starts' word offsets in d_srec
  uint64_t par_umask = 0;                     // derived parameters shared by every instance
  double par_uval[TVBK_KMAX_PAR] = {0};
  int mon_lds_off[TVBK_MAX_MON] = {0};       // image offsets (HBM, every node)
  // the LDS image: nodes it holds, its offsets and bytes
  int n_lds = 0, k_state = 0, k_mon_off[TVBK_MAX_MON] = {0};
  int64_t k_img_bytes = 0;
  bool batch_split = false;                  // batch-lane tiles split by node range
  int batch_parts = 0;
  int mon_Q[TVBK_MAX_MON] = {0};
  // device buffers
  DevBuf d_tiles, d_conns, d_edges, d_w, d_rp, d_ring, d_img, d_bacc, d_params, d_flags, d_streams, d_z, d_srec;
  DevBuf d_taps[TVBK_MAX_MON], d_out[TVBK_MAX_MON];
  DevBuf d_bstock[TVBK_MAX_MON];   // Bold monitors with keep_stock: interim-mean rings
  // monitor kinds as given (d.mon[m].kind holds the kernel's base kind:
  // RAW, TAVG or BOLD), sample source, areas of GAVG / SAVG
  int ukind[TVBK_MAX_MON] = {0}, src[TVBK_MAX_MON] = {0}, n_area[TVBK_MAX_MON] = {0};
  DevBuf d_raw[TVBK_MAX_MON], d_area_of[TVBK_MAX_MON], d_area_w[TVBK_MAX_MON], d_area_jobs[TVBK_MAX_MON];
  std::vector<AreaJob> jobs_host[TVBK_MAX_MON];
  // edge-lane instances split over several tiles: block length (steps,
  // 0 = no split) and the staged history of a block
  int split_k = 0;
  DevBuf d_stage;
  // region stimulus of the next run
  DevBuf d_stim_sp, d_stim_tp;
  const double* stim_sp = nullptr;
  const double* stim_tp = nullptr;
  int64_t stim_steps = 0;
  uint32_t stim_mask = 0;
  DevBuf d_dense;             // dense zero-delay kernel: W [16*RB][KP] float
  int dense_kb = 0, dense_rb = 0, lds_x = 0;
  int64_t z_steps = 0;
  const double* z_ptr = nullptr;
  // last run bookkeeping
  int64_t last_n = 0;
  std::vector<int64_t> mon_count[TVBK_MAX_MON], mon_first[TVBK_MAX_MON], mon_offset[TVBK_MAX_MON];
  int64_t mon_bytes[TVBK_MAX_MON] = {0};
  int64_t out_cap_samples[TVBK_MAX_MON] = {0};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool have_timing = false;
  // hipGraph of a split call's block launches (step kernel and commit per
  // block), captured on a private stream when a call repeats the length and
  // kernel arguments of the call before it, and replayed while calls keep
  // repeating them (all state the kernels read per call lives in device
  // buffers that are updated before the replay)
  hipStream_t cap_stream = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  StepArgs graph_args, prev_args;
  int64_t graph_n = -1, prev_n = -1;
  bool graph_off = false;                 // TVBK_GRAPH=0, or a capture failed
  int64_t graph_captures = 0, graph_replays = 0;
  std::vector<uint64_t> stream_host;   // Philox stream id per instance
  std::vector<double> taps_host[TVBK_MAX_MON];
};

namespace {

int fail(tvbk_plan* p, int code, const std::string& msg) {
  if (p) p->err = msg; else g_err = msg;
  return code;
}

#define HIPCHK(plan, expr)                                                                  \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess)                                                                   \
      return fail(plan, TVBK_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e));     \
  } while (0)

int dev_alloc(tvbk_plan* p, DevBuf* b, size_t bytes) {
  b->bytes = bytes;
  if (bytes == 0) { b->p = nullptr; return TVBK_OK; }
  hipError_t e = hipMalloc(&b->p, bytes);
  if (e != hipSuccess) {
    b->p = nullptr;
    return fail(p, TVBK_ENOMEM, "hipMalloc(" + std::to_string(bytes) + " bytes) failed: " + hipGetErrorString(e));
  }
  e = hipMemsetAsync(b->p, 0, bytes, p->stream);
  if (e != hipSuccess) return fail(p, TVBK_EHIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
  return TVBK_OK;
}

void dev_free(DevBuf* b) {
  if (b->p) hipFree(b->p);
  b->p = nullptr;
  b->bytes = 0;
}

int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

// Emission steps of a monitor with period istep within [s0, s0+n).
void emissions(int64_t s0, int64_t n, int istep, int64_t* count, int64_t* first) {
  if (n <= 0) { *count = 0; *first = s0; return; }
  const int64_t f = ((s0 + istep - 1) / istep) * istep;
  const int64_t last = s0 + n - 1;
  *first = f;
  *count = (f > last) ? 0 : (last - f) / istep + 1;
}

}  // namespace

extern "C" {

int tvbk_abi_version(void) { return TVBK_ABI_VERSION; }

int tvbk_model_info(int32_t model, int32_t cfun, int32_t* n_model_param, int32_t* n_cfun_param, int32_t* n_svar,
                    int32_t* n_ch) {
  ModelInfo mi;
  if (!model_info(model, &mi)) return fail(nullptr, TVBK_EINVAL, "unknown model");
  const int ncp = cfun_nparam(cfun);
  if (ncp < 0) return fail(nullptr, TVBK_EINVAL, "unknown coupling function");
  if (!combo_supported(model, cfun)) return fail(nullptr, TVBK_EUNSUPPORTED, "model/coupling combination not supported");
  if (n_model_param) *n_model_param = mi.nmp;
  if (n_cfun_param) *n_cfun_param = ncp;
  if (n_svar) *n_svar = mi.nsv;
  if (n_ch) *n_ch = n_channel(model, cfun);
  return TVBK_OK;
}

const char* tvbk_last_error(tvbk_plan* plan) { return plan ? plan->err.c_str() : g_err.c_str(); }

static void destroy(tvbk_plan* p) {
  (void)hipStreamSynchronize(p->stream);   // no pending copy or kernel touches the buffers
  DevBuf* bufs[] = {&p->d_tiles, &p->d_conns,  &p->d_edges, &p->d_w,       &p->d_rp, &p->d_ring, &p->d_img,
                    &p->d_bacc,  &p->d_params, &p->d_flags, &p->d_streams, &p->d_z,  &p->d_srec};
  for (DevBuf* b : bufs) dev_free(b);
  for (int m = 0; m < TVBK_MAX_MON; ++m) {
    dev_free(&p->d_taps[m]);
    dev_free(&p->d_out[m]);
    dev_free(&p->d_bstock[m]);
    dev_free(&p->d_raw[m]);
    dev_free(&p->d_area_of[m]);
    dev_free(&p->d_area_w[m]);
    dev_free(&p->d_area_jobs[m]);
  }
  dev_free(&p->d_stim_sp);
  dev_free(&p->d_stim_tp);
  dev_free(&p->d_stage);
  dev_free(&p->d_dense);
  if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
  if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
  if (p->ev0) hipEventDestroy(p->ev0);
  if (p->ev1) hipEventDestroy(p->ev1);
  delete p;
}

int tvbk_plan_destroy(tvbk_plan* plan) {
  if (!plan) return TVBK_OK;
  hipStreamSynchronize(plan->stream);
  destroy(plan);
  return TVBK_OK;
}

static int layout_lds(tvbk_plan* p, int T, int64_t max_lds, bool allow_csr) {
  const tvbk_plan_desc& d = p->d;
  const int N = p->max_n;
  // tile image of n nodes, node-major regions: Bold interim sums, state,
  // TemporalAverage sums; returns its bytes
  auto image = [&](int n, int* mon_off, int* state_off) -> int64_t {
    int64_t o = 0;
    for (int m = 0; m < d.n_mon; ++m)
      if (d.mon[m].kind == TVBK_MON_BOLD) {
        mon_off[m] = (int)o;
        o += (int64_t)n * d.mon[m].n_voi * T * 8;
      }
    *state_off = (int)o;
    o += (int64_t)n * p->mi.nsv * T * p->esize;
    for (int m = 0; m < d.n_mon; ++m)
      if (d.mon[m].kind == TVBK_MON_TAVG) {
        o = align_up(o, 8);
        mon_off[m] = (int)o;
        // float32: sum and Kahan compensation; float64: sum
        o += (int64_t)n * d.mon[m].n_voi * (p->esize == 8 ? 1 : 2) * T * p->esize;
      }
    return align_up(o, 16);
  };
  p->lds_bold = 0;
  // the HBM image holds every node; the LDS image the n_lds nodes of a
  // tile's node range (all of them unless tiles are split by node range)
  p->img_bytes = image(N, p->mon_lds_off, &p->lds_state);
  int64_t off = image(p->n_lds, p->k_mon_off, &p->k_state);
  p->k_img_bytes = off;
  p->lds_par = (int)off;                       // per-lane parameters [prow][T]
  off = align_up(off + (int64_t)p->prow * T * p->esize, 16);
  p->lds_c = (int)off;
  off += (int64_t)std::max(2 * p->pipe_k, p->tb_k) * p->n_lds * p->nct * T * p->hsize;
  off = align_up(off, 16);
  p->lds_si = (int)off;
  // monitor step records: one per step of a time block, or four (two passes
  // of two steps) for the two-step pipelined loop
  const int nsi = p->tb_k > 1 ? p->tb_k : (p->pipe_k == 2 ? 4 : 0);
  if (nsi) off = align_up(off + (int64_t)nsi * TVBK_KMAX_MON * TVBK_STEPINFO_BYTES, 16);
  // CSR staged as 16-byte records {col offset, delay, weight bits}
  const int64_t ne = p->csr_nnz + TVBK_ROW_PAD;
  const int64_t csr = ne * 16 + align_up((N + 1) * 4, 16);
  p->csr_in_lds = allow_csr && (off + csr <= max_lds);
  if (p->csr_in_lds) {
    p->lds_edge = (int)off;
    off += ne * 16;
    p->lds_w = 0;
    p->lds_rp = (int)off;
    off += align_up((N + 1) * 4, 16);
  }
  p->lds_total = (int)off;
  return off <= max_lds ? TVBK_OK : TVBK_EUNSUPPORTED;
}

static int plan_build(const tvbk_plan_desc* desc, void* stream, tvbk_plan** out, int64_t* footprint) {
  if (!desc || (!out && !footprint)) return fail(nullptr, TVBK_EINVAL, "null argument");
  if (out) *out = nullptr;
  if (desc->abi_version != TVBK_ABI_VERSION) return fail(nullptr, TVBK_EINVAL, "ABI version mismatch");
  tvbk_plan* p = new tvbk_plan();
  p->d = *desc;
  p->stream = (hipStream_t)stream;
  const tvbk_plan_desc& d = p->d;
  auto bail = [&](int code) {
    g_err = p->err;
    destroy(p);
    return code;
  };
  if (!model_info(d.model, &p->mi)) { p->err = "unknown model"; return bail(TVBK_EINVAL); }
  p->ncp = cfun_nparam(d.cfun);
  if (p->ncp < 0) { p->err = "unknown coupling function"; return bail(TVBK_EINVAL); }
  if (!combo_supported(d.model, d.cfun)) {
    p->err = "unsupported simulator component combination (model/coupling)";
    return bail(TVBK_EUNSUPPORTED);
  }
  if (d.scheme < TVBK_EULER_DET || d.scheme > TVBK_IDENTITY) { p->err = "unknown integrator"; return bail(TVBK_EINVAL); }
  p->stochastic = d.scheme == TVBK_EULER_STO || d.scheme == TVBK_HEUN_STO;
  if (p->stochastic && d.noise == TVBK_NOISE_NONE) { p->err = "stochastic scheme needs a noise source"; return bail(TVBK_EINVAL); }
  if (!p->stochastic && d.noise != TVBK_NOISE_NONE) { p->err = "deterministic scheme takes no noise"; return bail(TVBK_EINVAL); }
  if (d.precision != TVBK_F32 && d.precision != TVBK_F64 && d.precision != TVBK_MIXED) {
    p->err = "bad precision";
    return bail(TVBK_EINVAL);
  }
  if (d.precision == TVBK_MIXED && d.cfun == TVBK_CFUN_SIGMOID_JR) {
    p->err = "mixed precision is not offered with SigmoidalJansenRit (its reference pre and sum run in float64)";
    return bail(TVBK_EUNSUPPORTED);
  }
  p->esize = d.precision == TVBK_F32 ? 4 : 8;
  p->hsize = d.precision == TVBK_F64 ? 8 : 4;
  if (d.n_inst <= 0 || d.n_conn <= 0 || !d.conns || !d.params) { p->err = "empty plan"; return bail(TVBK_EINVAL); }
  const int nparam_expect = p->mi.nmp + p->ncp + (p->stochastic ? p->mi.nsv : 0);
  if (d.n_param != nparam_expect) {
    p->err = "n_param " + std::to_string(d.n_param) + " != " + std::to_string(nparam_expect);
    return bail(TVBK_EINVAL);
  }
  if (d.flags & ~(uint32_t)TVBK_PLAN_SKIP_ZERO_COUPLING) { p->err = "unknown plan flags"; return bail(TVBK_EINVAL); }
  p->nch = n_channel(d.model, d.cfun);
  p->nct = n_term(d.model, d.cfun);
  p->kcf = d.cfun;
  p->prow = p->mi.ndp + p->ncp + p->mi.nsv;
  if (d.n_obs <= 0 || d.n_obs > TVBK_MAX_OBS || !d.obs_coef) { p->err = "bad observables"; return bail(TVBK_EINVAL); }
  if (d.n_mon < 0 || d.n_mon > TVBK_MAX_MON) { p->err = "bad monitor count"; return bail(TVBK_EINVAL); }
  if (d.dt <= 0) { p->err = "dt must be positive"; return bail(TVBK_EINVAL); }
  if (d.max_steps <= 0) { p->err = "max_steps must be positive"; return bail(TVBK_EINVAL); }

  // ---- connectomes
  std::vector<uint32_t> edges;
  std::vector<double> w64;
  std::vector<int32_t> rp;
  for (int c = 0; c < d.n_conn; ++c) {
    const tvbk_connectome& cn = d.conns[c];
    if (cn.n_node <= 0 || cn.n_node > 65535 || cn.horizon <= 0 || cn.horizon > 65536 || cn.nnz < 0) {
      p->err = "connectome " + std::to_string(c) + ": n_node must be in [1, 65535] and horizon in [1, 65536]";
      return bail(TVBK_EINVAL);
    }
    ConnDesc cd{};
    cd.n_node = cn.n_node;
    cd.horizon = cn.horizon;
    cd.nnz = (int32_t)cn.nnz;
    cd.rp_off = (int64_t)rp.size();
    cd.e_off = (int64_t)edges.size();
    cd.d_min = cn.horizon;
    if (cn.row_ptr[0] != 0 || cn.row_ptr[cn.n_node] != cn.nnz) { p->err = "row_ptr inconsistent with nnz"; return bail(TVBK_EINVAL); }
    cd.max_row = 0;
    for (int i = 0; i < cn.n_node; ++i) {
      const int n = cn.row_ptr[i + 1] - cn.row_ptr[i];
      if (n < 0 || n > 1025) { p->err = "CSR rows must have 0..1025 entries"; return bail(TVBK_EINVAL); }
      cd.max_row = std::max(cd.max_row, n);
    }
    for (int i = 0; i <= cn.n_node; ++i) rp.push_back(cn.row_ptr[i]);
    for (int64_t e = 0; e < cn.nnz; ++e) {
      const int j = cn.col[e], dl = cn.idelay[e];
      if (j < 0 || j >= cn.n_node || dl < 0 || dl >= cn.horizon || dl > 65535) {
        p->err = "edge " + std::to_string(e) + " of connectome " + std::to_string(c) + " out of range";
        return bail(TVBK_EINVAL);
      }
      cd.d_min = std::min(cd.d_min, dl);
      edges.push_back((uint32_t)j | ((uint32_t)dl << 16));
      w64.push_back(cn.weight[e]);
    }
    p->max_n = std::max(p->max_n, cn.n_node);
    p->max_nnz = std::max<int64_t>(p->max_nnz, cn.nnz);
    p->csr_nnz = p->max_nnz;
    p->conns.push_back(cd);
  }
  // padding past the last edge: rows are read in fixed-width blocks
  for (int k = 0; k < TVBK_ROW_PAD; ++k) {
    edges.push_back(0);
    w64.push_back(0.0);
  }
  std::vector<int> inst_conn(d.n_inst, 0);
  bool shared = true;
  for (int b = 0; b < d.n_inst; ++b) {
    inst_conn[b] = d.inst_conn ? d.inst_conn[b] : 0;
    if (inst_conn[b] < 0 || inst_conn[b] >= d.n_conn) { p->err = "inst_conn out of range"; return bail(TVBK_EINVAL); }
    if (inst_conn[b] != inst_conn[0]) shared = false;
  }
  std::vector<int32_t> area_of[TVBK_MAX_MON];
  std::vector<double> area_w[TVBK_MAX_MON];
  for (int m = 0; m < d.n_mon; ++m) {
    tvbk_monitor& mo = p->d.mon[m];
    const int uk = mo.kind;
    if (uk < TVBK_MON_RAW || uk > TVBK_MON_AFF_TAVG || mo.istep <= 0 || mo.n_voi <= 0 || mo.n_voi > TVBK_MAX_OBS) {
      p->err = "bad monitor " + std::to_string(m);
      return bail(TVBK_EINVAL);
    }
    p->ukind[m] = uk;
    p->src[m] = (uk == TVBK_MON_AFF || uk == TVBK_MON_AFF_TAVG) ? 1 : 0;
    mo.kind = (uk == TVBK_MON_TAVG || uk == TVBK_MON_AFF_TAVG) ? TVBK_MON_TAVG
              : uk == TVBK_MON_BOLD                          ? TVBK_MON_BOLD
                                                              : TVBK_MON_RAW;
    if (uk == TVBK_MON_RAW && mo.istep != 1) { p->err = "Raw samples every step (istep 1)"; return bail(TVBK_EINVAL); }
    // afferent monitors index the model's coupling variables; the kernel's
    // coupling terms (one for SigmoidalJansenRit, whose two reference rows
    // are equal, coupling.py:358-376) hold them
    const int nvsrc = p->src[m] ? p->mi.ncv : d.n_obs;
    for (int v = 0; v < mo.n_voi; ++v) {
      if (mo.voi[v] < 0 || mo.voi[v] >= nvsrc) { p->err = "monitor voi out of range"; return bail(TVBK_EINVAL); }
      if (p->src[m]) mo.voi[v] = std::min(mo.voi[v], p->nct - 1);
    }
    if (uk == TVBK_MON_GAVG) {
      p->n_area[m] = 1;
    } else if (uk == TVBK_MON_SAVG) {
      const int A = mo.n_area;
      if (d.n_conn != 1 || !mo.spatial_mask || A <= 0) {
        p->err = "SpatialAverage needs one connectome, a spatial mask and n_area > 0";
        return bail(TVBK_EINVAL);
      }
      const int N0 = d.conns[0].n_node;
      std::vector<int> cnt(A, 0);
      area_of[m].assign(mo.spatial_mask, mo.spatial_mask + N0);
      for (int n = 0; n < N0; ++n) {
        if (area_of[m][n] < 0 || area_of[m][n] >= A) { p->err = "spatial mask value out of range"; return bail(TVBK_EINVAL); }
        ++cnt[area_of[m][n]];
      }
      for (int q = 0; q < A; ++q)   // monitors.py:315-320: contiguous area indices from zero
        if (cnt[q] == 0) { p->err = "every area of the spatial mask needs a node"; return bail(TVBK_EINVAL); }
      // spatial_mean = spatial_sum / nodes_per_area (monitors.py:325-331)
      area_w[m].resize(N0);
      for (int n = 0; n < N0; ++n) area_w[m][n] = 1.0 / (double)cnt[area_of[m][n]];
      p->n_area[m] = A;
    }
    if (mo.kind == TVBK_MON_BOLD && uk != TVBK_MON_BOLD) { p->err = "bad monitor kind"; return bail(TVBK_EINVAL); }
    if (mo.kind == TVBK_MON_BOLD) {
      if (mo.interim_istep <= 0 || mo.n_taps <= 0 || !mo.taps) { p->err = "bad Bold monitor"; return bail(TVBK_EINVAL); }
      // live accumulators: emissions within S interim samples of a sample
      p->mon_Q[m] = (int)(((int64_t)mo.n_taps * mo.interim_istep - 1) / mo.istep + 2);
      p->taps_host[m].assign(mo.taps, mo.taps + mo.n_taps);
    }
  }

  // ---- kernel shape
  p->n_lds = p->max_n;
  int dev = 0;
  HIPCHK(p, hipGetDevice(&dev));
  int max_lds = 65536;
  hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  max_lds -= 2048;   // the kernels' static LDS (tile and monitor configuration)
  int kernel = d.kernel;
  // dense zero-delay kernel: every delay zero (horizon 1), one shared
  // connectome of at most 128 nodes, Linear coupling, float32, G2D / MPR,
  // no afferent monitor
  bool any_aff = false;
  for (int m = 0; m < d.n_mon; ++m) any_aff |= p->src[m] != 0;
  const bool dense_ok = shared && d.precision != TVBK_F64 && d.cfun == TVBK_CFUN_LINEAR &&
                        (d.model == TVBK_MODEL_G2D || d.model == TVBK_MODEL_MPR) && p->conns.size() == 1 &&
                        p->conns[0].horizon == 1 && p->conns[0].n_node <= 128 && !any_aff;
  if (kernel == TVBK_KERNEL_AUTO) {
    if (dense_ok && d.n_inst >= 32) kernel = TVBK_KERNEL_DENSE_MFMA;
    else kernel = (shared && d.n_inst >= 32) ? TVBK_KERNEL_BATCH_LANE : TVBK_KERNEL_EDGE_LANE;
  }
  if (kernel == TVBK_KERNEL_DENSE_MFMA && !dense_ok) {
    p->err = "dense MFMA kernel needs a float32 history, Linear coupling, G2D or MPR, one connectome of <= 128 nodes "
             "with all delays zero, and no afferent monitor";
    return bail(TVBK_EUNSUPPORTED);
  }
  if (kernel == TVBK_KERNEL_BATCH_LANE && !shared) { p->err = "batch-lane kernel needs one shared connectome"; return bail(TVBK_EINVAL); }
  p->kernel = kernel;
  // MontbrioPazoRoxin multiplies its V-coupling term by cv
  // (infinite_theta.py:205): with cv == 0 in every instance (and node) the
  // V channel is neither stored nor gathered, when the caller allows it
  if ((d.flags & TVBK_PLAN_SKIP_ZERO_COUPLING) && d.model == TVBK_MODEL_MPR && d.cfun == TVBK_CFUN_LINEAR &&
      kernel != TVBK_KERNEL_DENSE_MFMA && !any_aff) {
    const int kcv = 7;   // raw MPR parameter cv
    bool zero = true;
    for (int b = 0; b < d.n_inst && zero; ++b) zero = d.params[(size_t)b * d.n_param + kcv] == 0.0;
    if (zero && ((d.node_param_mask >> kcv) & 1ull) && d.node_params) {
      const int N0 = d.conns[0].n_node;
      for (int i = 0; i < N0 && zero; ++i) zero = d.node_params[(size_t)kcv * N0 + i] == 0.0;
    }
    if (zero) {
      p->nch = p->nct = 1;
      p->kcf = KCF_LINEAR1;
    }
  }
  int T = 1;
  if (kernel == TVBK_KERNEL_DENSE_MFMA) {
    if (d.tile && d.tile != 32) { p->err = "dense kernel uses tile 32"; return bail(TVBK_EINVAL); }
    T = 32;
    layout_lds(p, T, max_lds, false);
    const int N0 = p->conns[0].n_node;
    p->dense_rb = (N0 + 15) / 16;
    p->dense_kb = (N0 + 3) / 4;
    p->lds_x = p->lds_c;
    p->lds_total = p->lds_x + 2 * p->nch * (4 * p->dense_kb) * T * 4;
    if (p->lds_total > max_lds) { p->err = "dense tile does not fit in LDS"; return bail(TVBK_EUNSUPPORTED); }
  } else if (kernel == TVBK_KERNEL_BATCH_LANE) {
    if (d.tile) {
      T = d.tile;
    } else {
      T = 16;
      for (int cand : {64, 32}) {
        if ((d.n_inst + cand - 1) / cand >= 256) { T = cand; break; }
      }
    }
    if (T != 16 && T != 32 && T != 64) { p->err = "tile must be 16, 32 or 64"; return bail(TVBK_EINVAL); }
    bool fit = false;
    while (true) {
      if (layout_lds(p, T, max_lds, true) == TVBK_OK) { fit = true; break; }
      if (d.tile || T == 16) break;
      T /= 2;
    }
    // a tile whose nodes do not fit LDS, or a batch of so few tiles that
    // they fill at most half of the device's CUs, is split by node range
    // over several workgroups: each holds its range's state and coupling
    // buffers, they step in blocks of K <= d_min steps (time-blocked loop),
    // stage their history writes and a commit stores them between blocks.
    // TVBK_BATCH_SPLIT: unset = when needed, 0 = never, n >= 2 = n parts
    const char* envb = std::getenv("TVBK_BATCH_SPLIT");
    const int forced_b = envb && *envb ? std::atoi(envb) : -1;
    int dmin_all = p->conns[0].horizon;
    for (const ConnDesc& cd : p->conns) dmin_all = std::min(dmin_all, cd.nnz > 0 ? cd.d_min : cd.horizon - 1);
    const int N0 = p->conns[0].n_node;
    int n_cu = 256;
    hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
    n_cu = std::max(n_cu, 1);
    const bool few = fit && !d.tile && 2 * ((d.n_inst + T - 1) / T) <= n_cu && N0 >= 32;
    const int T_fit = T;
    if ((!fit || few || forced_b >= 2) && forced_b != 0 && !any_aff && dmin_all >= 2 && N0 >= 16) {
      p->tb_k = std::min(std::min(dmin_all, (int)TVBK_TB_MAX), 16);
      // candidate tiles: rows filling a whole wave (n_channel * T = 64), then
      // half a wave; a given tile only
      int cand[2], nc = 0;
      if (d.tile) {
        cand[nc++] = d.tile;
      } else {
        for (int q : {64, 32})
          if (q % p->nch == 0 && (q / p->nch == 16 || q / p->nch == 32 || q / p->nch == 64)) cand[nc++] = q / p->nch;
        if (nc == 0) cand[nc++] = 16;
      }
      const int pmax = std::min(64, N0 / 8);
      int best_T = 0, best_P = 0;
      double best_cost = 0.0;
      for (int c = 0; c < nc; ++c) {
        const int Tc = cand[c];
        const bool uni = (p->nch * Tc) % 32 == 0;
        // fewest parts whose node range fits LDS
        int pmin = 0;
        for (int P = forced_b >= 2 ? forced_b : 2; P <= pmax || (forced_b >= 2 && P == forced_b); ++P) {
          p->n_lds = (N0 + P - 1) / P;
          if (layout_lds(p, Tc, max_lds, !uni) == TVBK_OK) { pmin = P; break; }
          if (forced_b >= 2) break;
        }
        if (pmin == 0) continue;
        // workgroups run one per CU: take as many parts as fill the rounds
        // that the fewest parts need, and rank tiles by rounds x rows per
        // workgroup x waves per row
        const int ntile0 = (d.n_inst + Tc - 1) / Tc;
        const int64_t rounds = ((int64_t)ntile0 * pmin + n_cu - 1) / n_cu;
        int P = pmin;
        if (forced_b < 2) P = (int)std::max<int64_t>(pmin, std::min<int64_t>(pmax, rounds * n_cu / ntile0));
        const double cost = (double)rounds * ((N0 + P - 1) / P) * (64.0 / std::min(64, p->nch * Tc));
        if (best_T == 0 || cost < best_cost) { best_T = Tc; best_P = P; best_cost = cost; }
      }
      bool ok = false;
      if (best_T) {
        p->n_lds = (N0 + best_P - 1) / best_P;
        ok = layout_lds(p, best_T, max_lds, (p->nch * best_T) % 32 != 0) == TVBK_OK;
      }
      if (ok) {
        p->batch_split = true;
        p->batch_parts = best_P;
        T = best_T;
        fit = true;
      } else {
        p->n_lds = p->max_n;
        p->tb_k = 0;
        if (fit) layout_lds(p, T_fit, max_lds, true);   // the unsplit tile that fits
      }
    }
    if (!fit) {
      if (d.kernel == TVBK_KERNEL_AUTO && !d.tile) {   // large nodes: one instance per workgroup
        kernel = p->kernel = TVBK_KERNEL_EDGE_LANE;
        T = 1;
      } else {
        p->err = "tile does not fit in LDS";
        return bail(TVBK_EUNSUPPORTED);
      }
    }
  }
  if (kernel == TVBK_KERNEL_EDGE_LANE) {
    if (d.tile && d.tile != 1) { p->err = "edge-lane kernel uses tile 1"; return bail(TVBK_EINVAL); }
    if (layout_lds(p, 1, max_lds, true) != TVBK_OK) { p->err = "node state does not fit in LDS"; return bail(TVBK_EUNSUPPORTED); }
  }
  p->T = T;
  // wave-uniform rows: one target node per wave, edge records through the
  // scalar cache (no CSR in LDS)
  p->uniform_rows = kernel == TVBK_KERNEL_BATCH_LANE && (p->nch * T) % 32 == 0;
  if (p->uniform_rows) layout_lds(p, T, max_lds, false);
  // two-step pipelined loop: rows of 32 values (half a wave, so a wave load
  // gathers two steps of an edge), every delay >= 3, no afferent monitor,
  // history channels staged in the coupling buffers (as many as coupling
  // terms), models of up to two state variables (the loop holds two rows of
  // loads beside the update: JansenRit's six variables spill); four
  // coupling buffers and step records in LDS.  TVBK_PIPE2=0 keeps the
  // one-step loop.
  p->pipe_k = 1;
  if (p->uniform_rows && !p->batch_split && p->nch * T == 32 && p->nct == p->nch && !any_aff && p->mi.nsv <= 2) {
    int dmin2 = 1 << 30;
    for (const ConnDesc& cd : p->conns) dmin2 = std::min(dmin2, cd.nnz > 0 ? cd.d_min : 0);
    const char* envp = std::getenv("TVBK_PIPE2");
    if (dmin2 >= 3 && !(envp && *envp && std::atoi(envp) == 0)) {
      p->pipe_k = 2;
      if (layout_lds(p, T, max_lds, false) != TVBK_OK) {
        p->pipe_k = 1;
        layout_lds(p, T, max_lds, false);
      }
    }
  }
  if (p->batch_split && p->lds_total > max_lds) { p->err = "split tile does not fit in LDS"; return bail(TVBK_EUNSUPPORTED); }
  {
    const int P = p->max_n * std::max(p->nch, 1);
    int nw;
    if (kernel == TVBK_KERNEL_DENSE_MFMA) nw = 2 * p->dense_rb;
    else if (d.waves) nw = d.waves;
    else if (p->uniform_rows) nw = std::min(16, p->max_n);
    else if (T > 1) nw = std::min(16, std::max(1, (P + 64 / T - 1) / (64 / T)));
    else nw = std::min(16, std::max(1, (std::max(P, p->max_n) + 63) / 64));
    const int max_waves = p->hsize == 8 ? 8 : 16;   // the float64-history kernels are built for 512 threads
    if (!d.waves) nw = std::min(nw, max_waves);
    if (nw < 1 || nw > max_waves) {
      p->err = "waves must be in [1, " + std::to_string(max_waves) + "]";
      return bail(TVBK_EINVAL);
    }
    p->waves = nw;
    p->nthreads = 64 * nw;
  }

  // ---- tiles
  const int B = d.n_inst;
  p->tile_of.assign(B, 0);
  p->lane_of.assign(B, 0);
  p->cur_step.assign(B, 0);
  p->remaining.assign(B, -1);
  std::vector<int> order(B);
  std::iota(order.begin(), order.end(), 0);
  if (T == 1) {
    // largest connectomes first: in-order dispatch then fills CUs
    // longest-processing-time first
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
      const ConnDesc& cx = p->conns[inst_conn[x]];
      const ConnDesc& cy = p->conns[inst_conn[y]];
      return (int64_t)cx.nnz + cx.n_node > (int64_t)cy.nnz + cy.n_node;
    });
  }
  int64_t ring_off = 0, img_off = 0, bacc_off = 0;
  int64_t bst_len[TVBK_MAX_MON] = {0};   // Bold stock rings (doubles)
  const int ntiles = (B + T - 1) / T;
  p->tiles.resize(ntiles);
  for (int k = 0; k < ntiles; ++k) {
    TileDesc& td = p->tiles[k];
    std::memset(&td, 0, sizeof(td));
    const int b0 = (T == 1) ? order[k] : k * T;
    td.inst0 = b0;
    td.n_inst = (T == 1) ? 1 : std::min(T, B - b0);
    td.conn = inst_conn[b0];
    const ConnDesc& cd = p->conns[td.conn];
    td.node0 = 0;
    td.node1 = cd.n_node;
    td.parts = 1;
    for (int l = 0; l < td.n_inst; ++l) {
      p->tile_of[b0 + l] = k;
      p->lane_of[b0 + l] = l;
    }
    td.ring_off = ring_off;
    const int64_t hslots = cd.horizon + (p->pipe_k == 2 ? 1 : 0);   // + the mirror slot
    ring_off += hslots * cd.n_node * p->nch * T;
    if (hslots * cd.n_node * p->nch * T * p->hsize >= (1LL << 31)) {
      p->err = "tile history exceeds 2 GiB";
      return bail(TVBK_EUNSUPPORTED);
    }
    if ((int64_t)cd.n_node * p->nch * T * p->hsize >= (1LL << 24)) {
      p->err = "one history slot of a tile exceeds 16 MiB";
      return bail(TVBK_EUNSUPPORTED);
    }
    td.img_off = img_off;
    img_off += align_up(p->img_bytes, 256);
    for (int m = 0; m < d.n_mon; ++m)
      if (d.mon[m].kind == TVBK_MON_BOLD) {
        td.bacc_off[m] = bacc_off;
        bacc_off += (int64_t)p->mon_Q[m] * d.mon[m].n_voi * cd.n_node * T;
        if (d.mon[m].keep_stock) {
          td.bst_off[m] = bst_len[m];
          bst_len[m] += (int64_t)d.mon[m].n_taps * d.mon[m].n_voi * cd.n_node * T;
        }
      }
  }

  // ---- batch-lane tiles split by node range: parts of equal node counts
  // (the LDS image holds n_lds nodes), sharing the tile's ring, image,
  // accumulators and outputs; history writes staged per block of split_k
  // steps
  int64_t stage_len = 0;
  if (p->batch_split) {
    const int N0 = p->conns[0].n_node, P = p->batch_parts, nn = p->n_lds;
    p->split_k = p->tb_k;
    std::vector<TileDesc> extra;
    for (int k = 0; k < ntiles; ++k) {
      TileDesc& td = p->tiles[k];
      td.parts = P;
      td.node0 = 0;
      td.node1 = std::min(nn, N0);
      td.stage_off = stage_len;
      stage_len += (int64_t)p->split_k * N0 * p->nch * T;
      for (int q = 1; q < P; ++q) {
        TileDesc t2 = td;
        t2.node0 = q * nn;
        t2.node1 = std::min((q + 1) * nn, N0);
        if (t2.node1 > t2.node0) extra.push_back(t2);
      }
    }
    p->tiles.insert(p->tiles.end(), extra.begin(), extra.end());
  }
  // ---- edge-lane kernel: split heavy instances over several tiles (node
  // ranges with balanced edge counts), so that one large connectome does not
  // hold one CU for the whole call while the others idle.  Split tiles step
  // in blocks of K <= d_min steps: a block reads only history written before
  // it, each tile stages its nodes' history writes, and a commit between
  // blocks stores them in the ring.  TVBK_SPLIT: unset = automatic, 0 = off,
  // n > 1 = n parts for every instance that can be split.
  if (T == 1 && kernel == TVBK_KERNEL_EDGE_LANE && B > 0) {
    const char* env = std::getenv("TVBK_SPLIT");
    const int forced = env ? std::atoi(env) : -1;
    int n_cu = 256;
    hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
    std::vector<int64_t> cost(B);
    int64_t total = 0;
    for (int b = 0; b < B; ++b) {
      const ConnDesc& cd = p->conns[inst_conn[b]];
      cost[b] = cd.nnz + cd.n_node;
      total += cost[b];
    }
    // aim at `div` tiles of equal per-step cost per CU (TVBK_SPLIT_DIV,
    // default 3), parts of at least 512 edges and 8 nodes: with the
    // time-blocked loop a part gathers a whole block at once, so finer parts
    // keep more CUs busy (measured on the uneven JansenRit batch: best at
    // 2-3 per CU, slower at 8)
    const char* envd = std::getenv("TVBK_SPLIT_DIV");
    const double div = std::max(0.05, envd ? std::atof(envd) : 3.0);
    const int64_t target = std::max<int64_t>((int64_t)((double)total / (div * std::max(n_cu, 1))), 512);
    int kmin = 1 << 30;
    std::vector<TileDesc> extra;
    for (int k = 0; k < ntiles && forced != 0; ++k) {
      TileDesc& td = p->tiles[k];
      const ConnDesc& cd = p->conns[td.conn];
      const tvbk_connectome& cn = d.conns[td.conn];
      if (cd.d_min < 4) continue;                      // blocks of at least four steps
      int parts = forced > 1 ? forced : (int)((cost[td.inst0] + target - 1) / target);
      parts = std::min(parts, std::min(64, cd.n_node / 8));
      if (parts < 2) continue;
      // node boundaries at equal shares of sum(row length + 1)
      std::vector<int> bound(parts + 1, cd.n_node);
      bound[0] = 0;
      int64_t acc = 0;
      const int64_t tot = cd.nnz + cd.n_node;
      int j = 1;
      for (int i = 0; i < cd.n_node && j < parts; ++i) {
        acc += (cn.row_ptr[i + 1] - cn.row_ptr[i]) + 1;
        if (acc * parts >= tot * j) bound[j++] = i + 1;
      }
      for (; j < parts; ++j) bound[j] = cd.n_node;
      td.parts = parts;
      td.node1 = bound[1];
      for (int q = 1; q < parts; ++q) {
        TileDesc t2 = td;
        t2.node0 = bound[q];
        t2.node1 = bound[q + 1];
        if (t2.node1 > t2.node0) extra.push_back(t2);
      }
      kmin = std::min(kmin, cd.d_min);
    }
    if (!extra.empty()) {
      p->split_k = std::min(kmin, 4096);
      // staging sized for K steps of each split instance (K <= its d_min)
      stage_len = 0;
      for (TileDesc& td : p->tiles)
        if (td.parts > 1) {
          td.stage_off = stage_len;
          stage_len += (int64_t)p->split_k * p->conns[td.conn].n_node * p->nch;
        }
      for (TileDesc& t2 : extra) t2.stage_off = p->tiles[p->tile_of[t2.inst0]].stage_off;
      p->tiles.insert(p->tiles.end(), extra.begin(), extra.end());
      // largest tiles first (edges of the node range): in-order dispatch
      // then fills the CUs longest-processing-time first
      auto tcost = [&](const TileDesc& t) -> int64_t {
        const tvbk_connectome& cn = d.conns[t.conn];
        return (int64_t)(cn.row_ptr[t.node1] - cn.row_ptr[t.node0]) + (t.node1 - t.node0);
      };
      std::stable_sort(p->tiles.begin(), p->tiles.end(),
                       [&](const TileDesc& x, const TileDesc& y) { return tcost(x) > tcost(y); });
      for (size_t k = 0; k < p->tiles.size(); ++k)
        if (p->tiles[k].node0 == 0) p->tile_of[p->tiles[k].inst0] = (int)k;
    } else {
      stage_len = 0;
      for (TileDesc& td : p->tiles) { td.parts = 1; td.node0 = 0; td.node1 = p->conns[td.conn].n_node; }
    }
  }
  // a tile stages only its own rows of the CSR and holds only its own node
  // range in the LDS image and coupling buffers: with split instances the
  // largest such range may fit where whole connectomes did not, and leaves
  // room for longer time blocks
  if (kernel == TVBK_KERNEL_EDGE_LANE) {
    int64_t rn = 0;
    int nr = 0;
    for (const TileDesc& td : p->tiles) {
      const tvbk_connectome& cn = d.conns[td.conn];
      rn = std::max<int64_t>(rn, (int64_t)cn.row_ptr[td.node1] - cn.row_ptr[td.node0]);
      nr = std::max(nr, td.node1 - td.node0);
    }
    if (rn < p->csr_nnz || nr < p->n_lds) {
      p->csr_nnz = std::min<int64_t>(p->csr_nnz, rn);
      p->n_lds = std::min(p->n_lds, nr);
      layout_lds(p, 1, max_lds, true);
    }
  }
  // time-blocked sequential loop: plans that cannot take the pipelined loop,
  // with every delay >= 2 and no afferent monitor, gather the coupling of
  // up to TVBK_TB_MAX (<= d_min) steps at once -- as many coupling buffers
  // as LDS holds without moving the CSR out of it.  TVBK_TIME_BLOCK: unset =
  // automatic, 0 or 1 = off, n = blocks of at most n steps
  p->tb_k = 0;
  int tb_cap = TVBK_TB_MAX;
  if (const char* env = std::getenv("TVBK_TIME_BLOCK"))
    if (*env) tb_cap = std::max(0, std::min(TVBK_TB_MAX, std::atoi(env)));
  if (kernel != TVBK_KERNEL_DENSE_MFMA && !(p->uniform_rows && p->nch * T <= 64) && !any_aff && tb_cap >= 2) {
    // (a connectome without edges reads nothing; its blocks stay below H
    // so that the slot counters wrap once per block)
    int dmin = tb_cap;
    for (const ConnDesc& cd : p->conns) dmin = std::min(dmin, cd.nnz > 0 ? cd.d_min : cd.horizon - 1);
    if (dmin >= 2) {
      const bool csr0 = p->csr_in_lds, allow = !p->uniform_rows;
      for (int K = dmin; K >= 2; --K) {
        p->tb_k = K;
        if (layout_lds(p, T, max_lds, allow) == TVBK_OK && p->csr_in_lds == csr0) break;
        p->tb_k = 0;
      }
      // split plans launch blocks of split_k steps: time blocks of equal
      // length inside them (33 = one block of 33, not 32 + 1)
      if (p->tb_k > 1 && p->split_k > p->tb_k) {
        const int nb = (p->split_k + p->tb_k - 1) / p->tb_k;
        p->tb_k = (p->split_k + nb - 1) / nb;
      }
      layout_lds(p, T, max_lds, allow);
    }
  }
  // a time block's gathers are (steps x rows) items: as many waves as the
  // kernel allows unless the caller chose (the 192-node instance: 5 % faster
  // with 12-16 waves than with one per 64 rows)
  if (p->tb_k > 1 && !d.waves) {
    p->waves = p->hsize == 8 ? 8 : 16;
    p->nthreads = 64 * p->waves;
  }

  // ---- derived parameters
  const bool spatial = d.node_param_mask != 0;
  if (spatial && (d.n_conn != 1 || !d.node_params)) { p->err = "node parameters need one connectome and node_params"; return bail(TVBK_EINVAL); }
  p->spatial = spatial ? 1 : 0;
  const int Nsp = spatial ? p->conns[0].n_node : 1;
  std::vector<double> prm((size_t)B * Nsp * p->prow, 0.0);
  std::vector<double> raw(d.n_param);
  for (int b = 0; b < B; ++b)
    for (int i = 0; i < Nsp; ++i) {
      for (int k = 0; k < d.n_param; ++k) {
        raw[k] = d.params[(size_t)b * d.n_param + k];
        if (spatial && ((d.node_param_mask >> k) & 1ull)) raw[k] = d.node_params[(size_t)k * Nsp + i];
      }
      double* o = &prm[((size_t)b * Nsp + i) * p->prow];
      derive_model(d.model, raw.data(), o);
      derive_cfun(d.cfun, raw.data() + p->mi.nmp, o + p->mi.ndp);
      for (int v = 0; v < p->mi.nsv; ++v)
        o[p->mi.ndp + p->ncp + v] = p->stochastic ? std::sqrt(2.0 * raw[p->mi.nmp + p->ncp + v]) : 0.0;
    }
  // derived parameters equal for every instance (and node): plan constants
  p->par_umask = 0;
  for (int k = 0; k < p->prow && k < TVBK_KMAX_PAR; ++k) {
    bool same = true;
    for (size_t r = 1; r < (size_t)B * Nsp && same; ++r) same = prm[r * p->prow + k] == prm[k];
    if (same) {
      p->par_umask |= 1ull << k;
      p->par_uval[k] = prm[k];
    }
  }

  // ---- device footprint (dry run stops here)
  {
    size_t tot = p->conns.size() * sizeof(ConnDesc) + edges.size() * 4 + 16 + rp.size() * 4 +
                 (w64.size() * p->hsize + 16) + prm.size() * p->esize + (d.inst_stream ? (size_t)B * 8 : 0) +
                 (size_t)(ring_off + stage_len) * p->hsize + (size_t)img_off + (size_t)bacc_off * 8 + (size_t)B * 4 +
                 p->tiles.size() * sizeof(TileDesc) +
                 (p->uniform_rows ? (size_t)(p->conns[0].nnz + 4 * p->conns[0].n_node + TVBK_ROW_PAD) * (8 + p->hsize) +
                                        (size_t)(p->conns[0].n_node + 1) * 4 + 64 : 0);
    for (int m = 0; m < d.n_mon; ++m) {
      if (d.mon[m].kind == TVBK_MON_BOLD) tot += p->taps_host[m].size() * 8 + (size_t)bst_len[m] * 8;
      const int64_t cap = d.max_steps / d.mon[m].istep + 1;
      int64_t elems = 0;
      if (T > 1) elems = cap * d.mon[m].n_voi * (int64_t)p->conns[0].n_node * B;
      else
        for (int b = 0; b < B; ++b) elems += cap * d.mon[m].n_voi * (int64_t)p->conns[inst_conn[b]].n_node;
      tot += (size_t)elems * (d.mon[m].kind == TVBK_MON_BOLD ? 8 : p->esize);
      if (p->n_area[m] > 0)   // node-level samples, then the per-area output
        tot += (size_t)(cap * d.mon[m].n_voi * p->n_area[m] * B) * p->esize +
               (size_t)p->conns[0].n_node * 12 + (size_t)B * sizeof(AreaJob);
    }
    if (d.noise == TVBK_NOISE_INJECTED) tot += (size_t)d.max_steps * p->mi.nsv * p->max_n * B * 8;
    if (footprint) *footprint = (int64_t)tot;
    if (!out) {
      destroy(p);
      return TVBK_OK;
    }
  }

  // ---- allocate and upload
  int rc;
  // zeroing and uploads are ordered on the plan's stream (which may be a
  // non-blocking stream, unordered with the null stream); the host arrays
  // stay alive until the stream synchronisation at the end of creation
  auto up = [&](DevBuf* b, const void* src, size_t bytes) -> int {
    int r = dev_alloc(p, b, bytes);
    if (r) return r;
    if (bytes) {
      hipError_t e = hipMemcpyAsync(b->p, src, bytes, hipMemcpyHostToDevice, p->stream);
      if (e != hipSuccess) return fail(p, TVBK_EHIP, std::string("upload: ") + hipGetErrorString(e));
    }
    return TVBK_OK;
  };
  if ((rc = up(&p->d_conns, p->conns.data(), p->conns.size() * sizeof(ConnDesc)))) return bail(rc);
  if ((rc = up(&p->d_edges, edges.data(), edges.size() * 4))) return bail(rc);
  if ((rc = up(&p->d_rp, rp.data(), rp.size() * 4))) return bail(rc);
  std::vector<float> w32, p32, wd;   // host sources of async uploads: alive until the final sync
  if (p->hsize == 8) {
    if ((rc = up(&p->d_w, w64.data(), w64.size() * 8))) return bail(rc);
  } else {
    w32.assign(w64.begin(), w64.end());
    if ((rc = up(&p->d_w, w32.data(), w32.size() * 4))) return bail(rc);
  }
  if (p->esize == 8) {
    if ((rc = up(&p->d_params, prm.data(), prm.size() * 8))) return bail(rc);
  } else {
    p32.assign(prm.begin(), prm.end());
    if ((rc = up(&p->d_params, p32.data(), p32.size() * 4))) return bail(rc);
  }
  if (d.inst_stream) {
    if ((rc = up(&p->d_streams, d.inst_stream, (size_t)B * 8))) return bail(rc);
  }
  p->stream_host.resize(B);
  for (int b = 0; b < B; ++b) p->stream_host[b] = d.inst_stream ? d.inst_stream[b] : (uint64_t)b;
  std::vector<uint32_t> srec;   // host source of an async upload: alive until the final sync
  size_t srec_w_off = 0, srec_rp_off = 0;   // word offsets of the weights and row starts
  if (p->uniform_rows) {
    // offsets {delay * slot bytes, source * node bytes} and weights of the
    // shared connectome's CSR, every row padded to a multiple of four
    // entries with copies of its first entry (weight 0), then padding
    const ConnDesc& cd = p->conns[0];
    const tvbk_connectome& cn = d.conns[0];
    const uint32_t cstride = (uint32_t)(p->nch * T * p->hsize), sstride = (uint32_t)cd.n_node * cstride;
    std::vector<int32_t> prp(cd.n_node + 1, 0);
    for (int i = 0; i < cd.n_node; ++i) prp[i + 1] = prp[i] + (cn.row_ptr[i + 1] - cn.row_ptr[i] + 3) / 4 * 4;
    const int64_t ne = prp[cd.n_node] + TVBK_ROW_PAD;
    const size_t wpe = (size_t)p->hsize / 4;
    srec_w_off = align_up((int64_t)ne * 2, 4);
    srec_rp_off = align_up((int64_t)(srec_w_off + (size_t)ne * wpe), 4);
    srec.assign(srec_rp_off + cd.n_node + 1, 0u);
    for (int i = 0; i < cd.n_node; ++i) {
      const int r0 = cn.row_ptr[i], n = cn.row_ptr[i + 1] - r0;
      for (int k = 0; k < prp[i + 1] - prp[i]; ++k) {
        const int64_t src = cd.e_off + r0 + (k < n ? k : 0);
        const uint32_t ed = edges[src];
        const size_t e = (size_t)prp[i] + k;
        srec[e * 2] = (ed >> 16) * sstride;
        srec[e * 2 + 1] = (ed & 0xffffu) * cstride;
        const double wv = k < n ? w64[src] : 0.0;
        if (p->hsize == 8) {
          std::memcpy(&srec[srec_w_off + e * 2], &wv, 8);
        } else {
          const float wf = (float)wv;
          std::memcpy(&srec[srec_w_off + e], &wf, 4);
        }
      }
    }
    std::memcpy(&srec[srec_rp_off], prp.data(), prp.size() * 4);
    if ((rc = up(&p->d_srec, srec.data(), srec.size() * 4))) return bail(rc);
  }
  p->srec_rp_off = srec_rp_off;
  p->srec_w_off = srec_w_off;
  if (kernel == TVBK_KERNEL_DENSE_MFMA) {
    const tvbk_connectome& cn = d.conns[0];
    const int KP = 4 * p->dense_kb, NR = 16 * p->dense_rb;
    wd.assign((size_t)NR * KP, 0.0f);
    for (int i = 0; i < cn.n_node; ++i)
      for (int e = cn.row_ptr[i]; e < cn.row_ptr[i + 1]; ++e) wd[(size_t)i * KP + cn.col[e]] = (float)cn.weight[e];
    if ((rc = up(&p->d_dense, wd.data(), wd.size() * 4))) return bail(rc);
  }
  for (int m = 0; m < d.n_mon; ++m)
    if (d.mon[m].kind == TVBK_MON_BOLD) {
      if ((rc = up(&p->d_taps[m], p->taps_host[m].data(), p->taps_host[m].size() * 8))) return bail(rc);
      if ((rc = dev_alloc(p, &p->d_bstock[m], (size_t)bst_len[m] * 8))) return bail(rc);
    }
  if ((rc = dev_alloc(p, &p->d_ring, (size_t)ring_off * p->hsize))) return bail(rc);
  if ((rc = dev_alloc(p, &p->d_stage, (size_t)stage_len * p->hsize))) return bail(rc);
  if ((rc = dev_alloc(p, &p->d_img, (size_t)img_off))) return bail(rc);
  if ((rc = dev_alloc(p, &p->d_bacc, (size_t)bacc_off * 8))) return bail(rc);
  if ((rc = dev_alloc(p, &p->d_flags, (size_t)B * 4))) return bail(rc);
  if ((rc = dev_alloc(p, &p->d_tiles, p->tiles.size() * sizeof(TileDesc)))) return bail(rc);
  // monitor outputs sized for max_steps per call
  for (int m = 0; m < d.n_mon; ++m) {
    const int64_t cap = d.max_steps / d.mon[m].istep + 1;
    p->out_cap_samples[m] = cap;
    const int es = d.mon[m].kind == TVBK_MON_BOLD ? 8 : p->esize;
    int64_t elems = 0;
    if (T > 1) {
      elems = cap * d.mon[m].n_voi * (int64_t)p->conns[0].n_node * B;
    } else {
      for (int b = 0; b < B; ++b) elems += cap * d.mon[m].n_voi * (int64_t)p->conns[inst_conn[b]].n_node;
    }
    if (p->n_area[m] > 0) {
      // GlobalAverage / SpatialAverage: the step kernel subsamples the nodes
      // into d_raw, k_areas reduces them into d_out
      if ((rc = dev_alloc(p, &p->d_raw[m], (size_t)elems * es))) return bail(rc);
      if ((rc = dev_alloc(p, &p->d_out[m], (size_t)(cap * d.mon[m].n_voi * p->n_area[m] * B) * es))) return bail(rc);
      if ((rc = dev_alloc(p, &p->d_area_jobs[m], (size_t)B * sizeof(AreaJob)))) return bail(rc);
      if (p->ukind[m] == TVBK_MON_SAVG) {
        if ((rc = up(&p->d_area_of[m], area_of[m].data(), area_of[m].size() * 4))) return bail(rc);
        if ((rc = up(&p->d_area_w[m], area_w[m].data(), area_w[m].size() * 8))) return bail(rc);
      }
    } else if ((rc = dev_alloc(p, &p->d_out[m], (size_t)elems * es))) {
      return bail(rc);
    }
    p->mon_count[m].assign(B, 0);
    p->mon_first[m].assign(B, 0);
    p->mon_offset[m].assign(B, 0);
  }
  HIPCHK(p, hipEventCreate(&p->ev0));
  HIPCHK(p, hipEventCreate(&p->ev1));
  if (const char* env = std::getenv("TVBK_GRAPH"))
    if (*env && std::atoi(env) == 0) p->graph_off = true;
  // pipelined loops: the first E waves carry one more integration pass (one
  // step of an item costs about two rows' time on the sweep; the two-step
  // loop's pass, two steps: rows shifted 0..4 -> 10.74, 10.65, 10.53, 10.39,
  // 10.99 us/step); they take that many rows fewer.  TVBK_ROW_SHIFT=n sets
  // the shift.
  if (p->uniform_rows && !p->batch_split) {
    const int N = p->conns[0].n_node, nw = p->waves, nthr = 64 * nw, T = p->T;
    const int npass = (N + nthr / T - 1) / (nthr / T), npw = 64 / T;
    const int E = (N - (npass - 1) * (nthr / T) + npw - 1) / npw;
    if (E > 0 && E < nw) {
      const double re = (double)N / nw - 2.0 * p->pipe_k * (nw - E) / nw;
      p->row_shift = std::max(0, N / nw - std::max(0, (int)std::lround(re)));
    }
  }
  if (const char* env = std::getenv("TVBK_ROW_SHIFT"))
    if (*env) p->row_shift = std::max(0, std::atoi(env));
  HIPCHK(p, hipStreamSynchronize(p->stream));
  *out = p;
  return TVBK_OK;
}

int tvbk_plan_create(const tvbk_plan_desc* desc, void* stream, tvbk_plan** out) {
  if (!out) return fail(nullptr, TVBK_EINVAL, "null argument");
  return plan_build(desc, stream, out, nullptr);
}

int tvbk_plan_footprint(const tvbk_plan_desc* desc, int64_t* bytes) {
  if (!bytes) return fail(nullptr, TVBK_EINVAL, "null argument");
  return plan_build(desc, nullptr, nullptr, bytes);
}

int tvbk_stream_create(void** stream) {
  if (!stream) return fail(nullptr, TVBK_EINVAL, "null argument");
  hipStream_t st = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e != hipSuccess) return fail(nullptr, TVBK_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  *stream = (void*)st;
  return TVBK_OK;
}

int tvbk_stream_destroy(void* stream) {
  if (!stream) return TVBK_OK;
  hipError_t e = hipStreamDestroy((hipStream_t)stream);
  if (e != hipSuccess) return fail(nullptr, TVBK_EHIP, std::string("hipStreamDestroy: ") + hipGetErrorString(e));
  return TVBK_OK;
}

int tvbk_poll(tvbk_plan* p, int32_t* done) {
  if (!p || !done) return fail(p, TVBK_EINVAL, "null argument");
  if (!p->have_timing) { *done = 1; return TVBK_OK; }
  hipError_t e = hipEventQuery(p->ev1);
  if (e == hipSuccess) { *done = 1; return TVBK_OK; }
  if (e == hipErrorNotReady) { *done = 0; return TVBK_OK; }
  return fail(p, TVBK_EHIP, std::string("hipEventQuery: ") + hipGetErrorString(e));
}

int tvbk_plan_shape(tvbk_plan* p, int32_t* kernel, int32_t* tile, int32_t* waves, int32_t* lds_bytes) {
  if (!p) return fail(nullptr, TVBK_EINVAL, "null plan");
  if (kernel) *kernel = p->kernel;
  if (tile) *tile = p->T;
  if (waves) *waves = p->waves;
  if (lds_bytes) *lds_bytes = p->lds_total;
  return TVBK_OK;
}

int tvbk_plan_graphs(tvbk_plan* p, int64_t* captured, int64_t* launched) {
  if (!p) return fail(nullptr, TVBK_EINVAL, "null plan");
  if (captured) *captured = p->graph_captures;
  if (launched) *launched = p->graph_replays;
  return TVBK_OK;
}

int tvbk_plan_channels(tvbk_plan* p, int32_t* n_channel) {
  if (!p || !n_channel) return fail(p, TVBK_EINVAL, "null argument");
  *n_channel = p->nch;
  return TVBK_OK;
}

int tvbk_plan_time_block(tvbk_plan* p, int32_t* steps) {
  if (!p) return fail(nullptr, TVBK_EINVAL, "null plan");
  if (steps) *steps = p->tb_k;
  return TVBK_OK;
}

int tvbk_plan_loop(tvbk_plan* p, int32_t* loop) {
  if (!p) return fail(nullptr, TVBK_EINVAL, "null plan");
  if (loop) *loop = p->last_loop;
  return TVBK_OK;
}

int tvbk_plan_tiles(tvbk_plan* p, int32_t* n_tiles, int32_t* split_k) {
  if (!p) return fail(nullptr, TVBK_EINVAL, "null plan");
  if (n_tiles) *n_tiles = (int32_t)p->tiles.size();
  if (split_k) *split_k = p->split_k;
  return TVBK_OK;
}

static int conn_of_inst(tvbk_plan* p, int b) { return p->tiles[p->tile_of[b]].conn; }

// History and state of instance(s) `inst` into the ring and tile images:
// converted on the host to the plan's precisions, staged on the device, and
// scattered into each instance's lane.  The staging is allocated per call,
// or (noalloc) the static device scratch array of this process and device.
static int set_history_impl(tvbk_plan* p, int32_t inst, const void* history, int32_t elem_bytes, const double* state,
                            int64_t current_step, bool noalloc) {
  if (!p) return fail(nullptr, TVBK_EINVAL, "null plan");
  if (elem_bytes != 4 && elem_bytes != 8) return fail(p, TVBK_EINVAL, "elem_bytes must be 4 or 8");
  if (current_step < 0) return fail(p, TVBK_EINVAL, "current_step must be >= 0");
  const int B = p->d.n_inst;
  if (inst < -1 || inst >= B) return fail(p, TVBK_EINVAL, "instance out of range");
  const int b0 = inst < 0 ? 0 : inst, b1 = inst < 0 ? B : inst + 1;
  const int conn = conn_of_inst(p, b0);
  if (inst < 0)
    for (int b = 0; b < B; ++b)
      if (conn_of_inst(p, b) != conn) return fail(p, TVBK_EINVAL, "inst=-1 needs one shared connectome");
  const ConnDesc& cd = p->conns[conn];
  const int H = cd.horizon, N = cd.n_node, nch = p->nch, nsv = p->mi.nsv;
  const size_t nh = (size_t)H * nch * N;
  // stage compact arrays on the device: history in the ring's precision,
  // state in the state precision
  std::vector<unsigned char> hbuf(nh * p->hsize), sbuf((size_t)nsv * N * p->esize);
  std::vector<unsigned char> hin(nh * elem_bytes);
  HIPCHK(p, hipMemcpy(hin.data(), history, hin.size(), hipMemcpyDefault));
  std::vector<double> sin((size_t)nsv * N);
  HIPCHK(p, hipMemcpy(sin.data(), state, sin.size() * 8, hipMemcpyDefault));
  for (size_t k = 0; k < nh; ++k) {
    const double v = elem_bytes == 8 ? reinterpret_cast<double*>(hin.data())[k] : reinterpret_cast<float*>(hin.data())[k];
    if (p->hsize == 8) reinterpret_cast<double*>(hbuf.data())[k] = v;
    else reinterpret_cast<float*>(hbuf.data())[k] = (float)v;
  }
  for (size_t k = 0; k < sin.size(); ++k) {
    if (p->esize == 8) reinterpret_cast<double*>(sbuf.data())[k] = sin[k];
    else reinterpret_cast<float*>(sbuf.data())[k] = (float)sin[k];
  }
  void *dh = nullptr, *ds = nullptr;
  std::unique_lock<std::mutex> scratch_lock;
  if (noalloc) {
    const size_t hb = (size_t)align_up((int64_t)hbuf.size(), 256);
    if (hb + sbuf.size() > (size_t)TVBK_NOALLOC_MAX_BYTES)
      return fail(p, TVBK_EINVAL, "history (" + std::to_string(hbuf.size()) + " bytes) and state (" +
                                      std::to_string(sbuf.size()) + " bytes) exceed the static scratch of " +
                                      std::to_string((long long)TVBK_NOALLOC_MAX_BYTES) + " bytes");
    scratch_lock = std::unique_lock<std::mutex>(g_scratch_mu);
    void* base = nullptr;
    HIPCHK(p, scratch_device_ptr(&base));
    dh = base;
    ds = (unsigned char*)base + hb;
  } else {
    hipError_t ea = hipMalloc(&dh, hbuf.size());
    if (ea == hipSuccess) ea = hipMalloc(&ds, sbuf.size());
    if (ea != hipSuccess) {
      if (dh) hipFree(dh);
      return fail(p, TVBK_ENOMEM, std::string("hipMalloc (history staging): ") + hipGetErrorString(ea));
    }
  }
  hipError_t e = hipMemcpy(dh, hbuf.data(), hbuf.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(ds, sbuf.data(), sbuf.size(), hipMemcpyHostToDevice);
  for (int b = b0; b < b1 && e == hipSuccess; ++b) {
    const TileDesc& td = p->tiles[p->tile_of[b]];
    e = launch_scatter(p->d.precision, false, (char*)p->d_ring.p + td.ring_off * p->hsize,
                       (unsigned char*)p->d_img.p + td.img_off + p->lds_state, dh, ds, H, N, nch, nsv, p->T,
                       p->lane_of[b], p->stream);
    p->cur_step[b] = current_step;
  }
  // the staging is consumed before it is freed or handed to the next call
  const hipError_t es = hipStreamSynchronize(p->stream);
  if (e == hipSuccess) e = es;
  if (!noalloc) {
    hipFree(dh);
    hipFree(ds);
  }
  if (e != hipSuccess) return fail(p, TVBK_EHIP, std::string("scatter: ") + hipGetErrorString(e));
  return TVBK_OK;
}

int tvbk_set_history(tvbk_plan* p, int32_t inst, const void* history, int32_t elem_bytes, const double* state,
                     int64_t current_step) {
  return set_history_impl(p, inst, history, elem_bytes, state, current_step, false);
}

int tvbk_set_history_noalloc(tvbk_plan* p, int32_t inst, const void* history, int32_t elem_bytes,
                             const double* state, int64_t current_step) {
  return set_history_impl(p, inst, history, elem_bytes, state, current_step, true);
}

int tvbk_export_state(tvbk_plan* p, int32_t inst, void* history, int32_t elem_bytes, double* state,
                      int64_t* current_step) {
  if (!p) return fail(nullptr, TVBK_EINVAL, "null plan");
  if (inst < 0 || inst >= p->d.n_inst) return fail(p, TVBK_EINVAL, "instance out of range");
  if (elem_bytes != 4 && elem_bytes != 8) return fail(p, TVBK_EINVAL, "elem_bytes must be 4 or 8");
  const TileDesc& td = p->tiles[p->tile_of[inst]];
  const ConnDesc& cd = p->conns[td.conn];
  const int H = cd.horizon, N = cd.n_node, nch = p->nch, nsv = p->mi.nsv;
  const size_t nh = (size_t)H * nch * N;
  void *dh = nullptr, *ds = nullptr;
  HIPCHK(p, hipMalloc(&dh, nh * p->hsize));
  HIPCHK(p, hipMalloc(&ds, (size_t)nsv * N * p->esize));
  hipError_t e = launch_scatter(p->d.precision, true, (char*)p->d_ring.p + td.ring_off * p->hsize,
                                (unsigned char*)p->d_img.p + td.img_off + p->lds_state, dh, ds, H, N, nch, nsv, p->T,
                                p->lane_of[inst], p->stream);
  std::vector<unsigned char> hb(nh * p->hsize), sb((size_t)nsv * N * p->esize);
  if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
  if (e == hipSuccess) e = hipMemcpy(hb.data(), dh, hb.size(), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(sb.data(), ds, sb.size(), hipMemcpyDeviceToHost);
  hipFree(dh);
  hipFree(ds);
  if (e != hipSuccess) return fail(p, TVBK_EHIP, std::string("export: ") + hipGetErrorString(e));
  if (history) {
    std::vector<unsigned char> ho(nh * elem_bytes);
    for (size_t k = 0; k < nh; ++k) {
      const double v = p->hsize == 8 ? reinterpret_cast<double*>(hb.data())[k] : reinterpret_cast<float*>(hb.data())[k];
      if (elem_bytes == 8) reinterpret_cast<double*>(ho.data())[k] = v;
      else reinterpret_cast<float*>(ho.data())[k] = (float)v;
    }
    HIPCHK(p, hipMemcpy(history, ho.data(), ho.size(), hipMemcpyDefault));
  }
  if (state) {
    std::vector<double> so((size_t)nsv * N);
    for (size_t k = 0; k < so.size(); ++k)
      so[k] = p->esize == 8 ? reinterpret_cast<double*>(sb.data())[k] : reinterpret_cast<float*>(sb.data())[k];
    HIPCHK(p, hipMemcpy(state, so.data(), so.size() * 8, hipMemcpyDefault));
  }
  if (current_step) *current_step = p->cur_step[inst];
  return TVBK_OK;
}

int tvbk_init_uniform(tvbk_plan* p, uint64_t seed, const double* lo, const double* hi) {
  if (!p) return fail(nullptr, TVBK_EINVAL, "null plan");
  if (!lo || !hi) return fail(p, TVBK_EINVAL, "null bounds");
  InitArgs a{};
  a.tiles = (const TileDesc*)p->d_tiles.p;
  a.conns = (const ConnDesc*)p->d_conns.p;
  a.ring = p->d_ring.p;
  a.img = (unsigned char*)p->d_img.p;
  a.params = p->d_params.p;
  a.streams = (const uint64_t*)p->d_streams.p;
  a.prow = p->prow * (p->spatial ? p->conns[0].n_node : 1);
  a.T = p->T;
  a.lds_state = p->lds_state;
  a.key0 = (uint32_t)seed;
  a.key1 = (uint32_t)(seed >> 32);
  for (int v = 0; v < p->mi.nsv; ++v) { a.lo[v] = lo[v]; a.hi[v] = hi[v]; }
  HIPCHK(p, hipMemcpyAsync(p->d_tiles.p, p->tiles.data(), p->tiles.size() * sizeof(TileDesc), hipMemcpyHostToDevice,
                           p->stream));
  HIPCHK(p, launch_init(p->d.precision, p->d.model, p->kcf, a, (int)p->tiles.size(), p->stream));
  HIPCHK(p, hipStreamSynchronize(p->stream));
  std::fill(p->cur_step.begin(), p->cur_step.end(), 0);
  return TVBK_OK;
}

int tvbk_set_noise(tvbk_plan* p, const double* z, int64_t n_steps) {
  if (!p) return fail(nullptr, TVBK_EINVAL, "null plan");
  if (p->d.noise != TVBK_NOISE_INJECTED) return fail(p, TVBK_ESTATE, "plan noise source is not INJECTED");
  if (!z || n_steps <= 0) return fail(p, TVBK_EINVAL, "empty noise");
  hipPointerAttribute_t attr;
  bool on_device = false;
  if (hipPointerGetAttributes(&attr, z) == hipSuccess) on_device = attr.type == hipMemoryTypeDevice;
  (void)hipGetLastError();
  const size_t bytes = (size_t)n_steps * p->mi.nsv * p->max_n * p->d.n_inst * 8;
  if (on_device) {
    p->z_ptr = z;
  } else {
    if (p->d_z.bytes < bytes) {
      dev_free(&p->d_z);
      int rc = dev_alloc(p, &p->d_z, bytes);
      if (rc) return rc;
    }
    HIPCHK(p, hipMemcpyAsync(p->d_z.p, z, bytes, hipMemcpyHostToDevice, p->stream));
    p->z_ptr = (const double*)p->d_z.p;
  }
  p->z_steps = n_steps;
  return TVBK_OK;
}

int tvbk_set_stimulus(tvbk_plan* p, const double* spatial, const double* temporal, int64_t n_steps,
                      uint32_t svar_mask) {
  if (!p) return fail(nullptr, TVBK_EINVAL, "null plan");
  HIPCHK(p, hipStreamSynchronize(p->stream));   // the previous run may still read the buffers
  if (!temporal) {
    p->stim_tp = nullptr;
    p->stim_steps = 0;
    return TVBK_OK;
  }
  if (!spatial || n_steps <= 0) return fail(p, TVBK_EINVAL, "empty stimulus");
  if (p->kernel == TVBK_KERNEL_DENSE_MFMA) return fail(p, TVBK_EUNSUPPORTED, "the dense kernel takes no stimulus");
  if (p->conns.size() != 1) return fail(p, TVBK_EUNSUPPORTED, "a region stimulus needs one connectome");
  if (svar_mask == 0 || (svar_mask >> p->mi.nsv) != 0) return fail(p, TVBK_EINVAL, "bad stimulus state-variable mask");
  const size_t nsp = (size_t)p->conns[0].n_node * 8, ntp = (size_t)n_steps * 8;
  int rc;
  if (p->d_stim_sp.bytes < nsp) {
    dev_free(&p->d_stim_sp);
    if ((rc = dev_alloc(p, &p->d_stim_sp, nsp))) return rc;
  }
  if (p->d_stim_tp.bytes < ntp) {
    dev_free(&p->d_stim_tp);
    if ((rc = dev_alloc(p, &p->d_stim_tp, ntp))) return rc;
  }
  HIPCHK(p, hipStreamSynchronize(p->stream));
  HIPCHK(p, hipMemcpy(p->d_stim_sp.p, spatial, nsp, hipMemcpyDefault));
  HIPCHK(p, hipMemcpy(p->d_stim_tp.p, temporal, ntp, hipMemcpyDefault));
  p->stim_sp = (const double*)p->d_stim_sp.p;
  p->stim_tp = (const double*)p->d_stim_tp.p;
  p->stim_steps = n_steps;
  p->stim_mask = svar_mask;
  return TVBK_OK;
}

int tvbk_set_steps(tvbk_plan* p, const int64_t* remaining) {
  if (!p) return fail(nullptr, TVBK_EINVAL, "null plan");
  for (int b = 0; b < p->d.n_inst; ++b) p->remaining[b] = remaining ? remaining[b] : -1;
  return TVBK_OK;
}

int tvbk_run_steps(tvbk_plan* p, int64_t n) {
  if (!p) return fail(nullptr, TVBK_EINVAL, "null plan");
  if (n < 0) return fail(p, TVBK_EINVAL, "n_steps must be >= 0");
  if (n > p->d.max_steps) return fail(p, TVBK_EINVAL, "n_steps exceeds the plan's max_steps");
  const tvbk_plan_desc& d = p->d;
  const int B = d.n_inst, T = p->T;
  if (p->stochastic && d.noise == TVBK_NOISE_INJECTED && (p->z_ptr == nullptr || p->z_steps < n))
    return fail(p, TVBK_ESTATE, "injected noise for this run not set (tvbk_set_noise)");
  // the tile table is rewritten below: the previous run must have consumed it
  HIPCHK(p, hipStreamSynchronize(p->stream));
  // per-tile steps and monitor layout
  int64_t out_off[TVBK_MAX_MON] = {0};
  for (size_t k = 0; k < p->tiles.size(); ++k) {
    TileDesc& td = p->tiles[k];
    if (td.node0 != 0) continue;            // further tiles of a split instance: copied below
    const int b0 = td.inst0;
    const int64_t cs = p->cur_step[b0];
    int64_t steps = n;
    for (int l = 0; l < td.n_inst; ++l) {
      const int b = b0 + l;
      if (p->cur_step[b] != cs) return fail(p, TVBK_ESTATE, "instances of one tile must share current_step");
      if (p->remaining[b] >= 0) steps = std::min(steps, p->remaining[b]);
    }
    td.step0 = cs + 1;
    td.n_steps = steps;
    const ConnDesc& cd = p->conns[td.conn];
    for (int m = 0; m < d.n_mon; ++m) {
      int64_t cnt, first;
      emissions(td.step0, steps, d.mon[m].istep, &cnt, &first);
      const int nv = d.mon[m].n_voi;
      if (T > 1) {
        td.out_off[m] = (int64_t)b0;
        td.out_si[m] = B;
        td.out_sv[m] = (int64_t)cd.n_node * B;
        td.out_sj[m] = (int64_t)nv * cd.n_node * B;
      } else {
        td.out_off[m] = out_off[m];
        td.out_si[m] = 1;
        td.out_sv[m] = cd.n_node;
        td.out_sj[m] = (int64_t)nv * cd.n_node;
        out_off[m] += cnt * nv * cd.n_node;
      }
      for (int l = 0; l < td.n_inst; ++l) {
        p->mon_count[m][b0 + l] = cnt;
        p->mon_first[m][b0 + l] = first;
        p->mon_offset[m][b0 + l] = (T > 1) ? 0 : td.out_off[m];
      }
    }
    // injected noise [k][sv][node][inst] over max_n nodes
    td.z_si = B;
    td.z_ssv = (int64_t)p->max_n * B;
    td.z_sk = (int64_t)p->mi.nsv * p->max_n * B;
    td.z_off = b0;
  }
  for (TileDesc& td : p->tiles)
    if (td.node0 != 0) {
      const TileDesc& pt = p->tiles[p->tile_of[td.inst0]];
      const int n0 = td.node0, n1 = td.node1;
      td = pt;                             // the instance's call layout, this tile's nodes
      td.node0 = n0;
      td.node1 = n1;
    }
  for (int m = 0; m < d.n_mon; ++m) {
    const int es = d.mon[m].kind == TVBK_MON_BOLD ? 8 : p->esize;
    int64_t cnt = 0;
    if (p->n_area[m] > 0) {
      // reduced output: [sample][voi][area][inst] (T > 1) or per instance
      // [sample][voi][area] one after the other
      const int A = p->n_area[m], nv = d.mon[m].n_voi;
      if (T > 1) {
        cnt = p->mon_count[m][0] * nv * (int64_t)A * B;
      } else {
        for (int b = 0; b < B; ++b) {
          p->mon_offset[m][b] = cnt;
          cnt += p->mon_count[m][b] * nv * (int64_t)A;
        }
      }
    } else if (T > 1) {
      cnt = p->mon_count[m][0] * d.mon[m].n_voi * (int64_t)p->conns[0].n_node * B;
    } else {
      cnt = out_off[m];
    }
    p->mon_bytes[m] = cnt * es;
  }
  if (p->stim_tp && p->stim_steps < n) return fail(p, TVBK_ESTATE, "the stimulus set for this run is shorter than it");
  HIPCHK(p, hipMemcpyAsync(p->d_tiles.p, p->tiles.data(), p->tiles.size() * sizeof(TileDesc), hipMemcpyHostToDevice,
                           p->stream));
  HIPCHK(p, hipMemsetAsync(p->d_flags.p, 0, (size_t)B * 4, p->stream));
  StepArgs a;
  std::memset(&a, 0, sizeof(a));
  a.tiles = (const TileDesc*)p->d_tiles.p;
  a.conns = (const ConnDesc*)p->d_conns.p;
  a.edges = (const uint32_t*)p->d_edges.p;
  a.weights = p->d_w.p;
  a.row_ptr = (const int32_t*)p->d_rp.p;
  a.ring = p->d_ring.p;
  a.img = (unsigned char*)p->d_img.p;
  a.bacc = (double*)p->d_bacc.p;
  for (int m = 0; m < d.n_mon; ++m) a.out[m] = p->n_area[m] > 0 ? p->d_raw[m].p : p->d_out[m].p;
  a.stim_sp = p->stim_tp ? p->stim_sp : nullptr;
  a.stim_tp = p->stim_tp;
  a.stim_mask = p->stim_mask;
  for (int m = 0; m < d.n_mon; ++m) a.has_aff |= p->src[m];
  a.z = p->z_ptr;
  a.params = p->d_params.p;
  a.flags = (int32_t*)p->d_flags.p;
  a.streams = (const uint64_t*)p->d_streams.p;
  a.prow = p->prow;
  a.spatial = p->spatial;
  a.T = T;
  a.n_mon = d.n_mon;
  a.scheme = d.scheme;
  a.noise = d.noise;
  a.csr_in_lds = p->csr_in_lds ? 1 : 0;
  a.sofs = (const uint32_t*)p->d_srec.p;
  a.swts = p->d_srec.p ? (const void*)((const uint32_t*)p->d_srec.p + p->srec_w_off) : nullptr;
  a.srp = p->d_srec.p ? (const int32_t*)((const uint32_t*)p->d_srec.p + p->srec_rp_off) : nullptr;
  a.uniform_rows = p->uniform_rows ? 1 : 0;

  int zero_delay = 0;
  for (const ConnDesc& cd : p->conns) zero_delay |= (cd.d_min == 0 && cd.nnz > 0);
  a.zero_delay = zero_delay;
  a.key0 = (uint32_t)d.seed;
  a.key1 = (uint32_t)(d.seed >> 32);
  a.dt = d.dt;
  a.sqrt_dt = std::sqrt(d.dt);
  a.lds_bold = p->lds_bold;
  a.lds_state = p->k_state;
  a.lds_nodes = p->n_lds;
  a.no_pipe = p->batch_split ? 1 : 0;
  a.pipe_k = p->pipe_k;
  a.row_shift = p->row_shift;
  a.lds_c = p->lds_c;
  a.lds_si = p->lds_si;
  a.tb_k = p->tb_k;
  a.lds_par = p->lds_par;
  a.par_umask = p->par_umask;
  for (int k = 0; k < TVBK_KMAX_PAR; ++k) a.par_uval[k] = p->par_uval[k];
  a.lds_edge = p->lds_edge;
  a.lds_w = p->lds_w;
  a.lds_rp = p->lds_rp;
  a.img_bytes = p->img_bytes;
  for (int o = 0; o < d.n_obs; ++o) {
    int single = -1, nz = 0;
    for (int v = 0; v < p->mi.nsv; ++v) {
      const double c = d.obs_coef[o * p->mi.nsv + v];
      a.obs_coef[o][v] = c;
      if (c != 0.0) { ++nz; if (c == 1.0) single = v; }
    }
    a.obs_sv[o] = (nz == 1 && single >= 0) ? single : -1;
  }
  for (int m = 0; m < d.n_mon; ++m) {
    const tvbk_monitor& mo = d.mon[m];
    MonArgs& ma = a.mon[m];
    ma.kind = mo.kind;
    ma.istep = mo.istep;
    ma.nvoi = mo.n_voi;
    for (int v = 0; v < mo.n_voi; ++v) ma.voi[v] = mo.voi[v];
    ma.ii = mo.kind == TVBK_MON_BOLD ? mo.interim_istep : 1;
    ma.rat = (mo.kind == TVBK_MON_BOLD && mo.istep % ma.ii == 0) ? mo.istep / ma.ii : 0;
    ma.S = mo.n_taps;
    ma.Q = p->mon_Q[m] > 0 ? p->mon_Q[m] : 1;
    ma.sub1 = mo.subtract_one;
    ma.k1v0 = mo.k1_v0;
    ma.taps = (const double*)p->d_taps[m].p;
    ma.stock = (double*)p->d_bstock[m].p;
    ma.src = p->src[m];
    ma.lds_off = p->k_mon_off[m];
  }
  a.dense_w = (const float*)p->d_dense.p;
  a.dense_kb = p->dense_kb;
  a.dense_rb = p->dense_rb;
  a.lds_x = p->lds_x;
  // image regions (node-major): HBM and LDS offsets, bytes per node
  a.n_img_reg = 0;
  for (int m = 0; m < d.n_mon; ++m)
    if (d.mon[m].kind == TVBK_MON_BOLD) {
      a.img_reg_off[a.n_img_reg] = p->mon_lds_off[m];
      a.img_reg_lds[a.n_img_reg] = p->k_mon_off[m];
      a.img_reg_bpn[a.n_img_reg++] = d.mon[m].n_voi * T * 8;
    }
  a.img_reg_off[a.n_img_reg] = p->lds_state;
  a.img_reg_lds[a.n_img_reg] = p->k_state;
  a.img_reg_bpn[a.n_img_reg++] = p->mi.nsv * T * p->esize;
  for (int m = 0; m < d.n_mon; ++m)
    if (d.mon[m].kind == TVBK_MON_TAVG) {
      a.img_reg_off[a.n_img_reg] = p->mon_lds_off[m];
      a.img_reg_lds[a.n_img_reg] = p->k_mon_off[m];
      a.img_reg_bpn[a.n_img_reg++] = d.mon[m].n_voi * (p->esize == 8 ? 1 : 2) * T * p->esize;
    }
  a.stage = p->d_stage.p;
  p->last_loop = p->kernel == TVBK_KERNEL_DENSE_MFMA ? TVBK_LOOP_DENSE : step_loop(a, p->nch);
  HIPCHK(p, hipEventRecord(p->ev0, p->stream));
  if (p->kernel == TVBK_KERNEL_DENSE_MFMA) {
    a.k0 = 0;
    a.k1 = (int32_t)n;
    HIPCHK(p, launch_dense(d.precision, d.model, a, (int)p->tiles.size(), p->nthreads, (size_t)p->lds_total,
                           p->stream));
  } else {
    // split plans run blocks of split_k steps, each followed by the commit
    // of the split tiles' staged history; others run the call in one launch
    const int64_t K = p->split_k > 0 ? p->split_k : std::max<int64_t>(n, 1);
    auto blocks = [&](hipStream_t st) -> hipError_t {
      for (int64_t k0 = 0; k0 < n; k0 += K) {
        StepArgs ab = a;
        ab.k0 = (int32_t)k0;
        ab.k1 = (int32_t)std::min<int64_t>(k0 + K, n);
        hipError_t e = launch_step(d.precision, d.model, p->kcf, ab, (int)p->tiles.size(), p->nthreads,
                                   (size_t)p->lds_total, st);
        if (e == hipSuccess && p->split_k > 0)
          e = launch_commit(p->hsize, (const TileDesc*)p->d_tiles.p, (const ConnDesc*)p->d_conns.p, p->d_ring.p,
                            p->d_stage.p, T, p->nch, (int)k0, (int)K, (int)p->tiles.size(), st);
        if (e != hipSuccess) return e;
      }
      return hipSuccess;
    };
    // several blocks: a call that repeats the previous call's length and
    // arguments is captured into a graph, replayed while calls repeat it
    bool graph = p->split_k > 0 && n > K && !p->graph_off;
    const bool cached = graph && p->graph_exec && p->graph_n == n && std::memcmp(&a, &p->graph_args, sizeof(a)) == 0;
    const bool repeat = graph && p->prev_n == n && std::memcmp(&a, &p->prev_args, sizeof(a)) == 0;
    p->prev_args = a;
    p->prev_n = n;
    if (graph && !cached && !repeat) graph = false;
    if (graph && !cached) {
      if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
      p->graph_exec = nullptr;
      hipError_t e = p->cap_stream ? hipSuccess : hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking);
      hipGraph_t g = nullptr;
      if (e == hipSuccess) e = hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal);
      if (e == hipSuccess) {
        const hipError_t eb = blocks(p->cap_stream);
        e = hipStreamEndCapture(p->cap_stream, &g);
        if (eb != hipSuccess) e = eb;
      }
      if (e == hipSuccess) e = hipGraphInstantiate(&p->graph_exec, g, nullptr, nullptr, 0);
      if (g) (void)hipGraphDestroy(g);
      if (e != hipSuccess) {            // run the blocks directly from now on
        p->graph_exec = nullptr;
        p->graph_off = true;
        graph = false;
        (void)hipGetLastError();
      } else {
        p->graph_args = a;
        p->graph_n = n;
        ++p->graph_captures;
      }
    }
    if (graph) {
      HIPCHK(p, hipGraphLaunch(p->graph_exec, p->stream));
      ++p->graph_replays;
    } else {
      HIPCHK(p, blocks(p->stream));
    }
  }
  // GlobalAverage / SpatialAverage of the call's node samples
  for (int m = 0; m < d.n_mon; ++m) {
    if (p->n_area[m] <= 0) continue;
    const int nv = d.mon[m].n_voi;
    std::vector<AreaJob>& jobs = p->jobs_host[m];
    jobs.assign(B, AreaJob{});
    int64_t maxc = 0;
    for (size_t k = 0; k < p->tiles.size(); ++k) {
      const TileDesc& td = p->tiles[k];
      const int Nt = p->conns[td.conn].n_node;
      for (int l = 0; l < td.n_inst; ++l) {
        const int b = td.inst0 + l;
        AreaJob& jb = jobs[b];
        jb.n_node = Nt;
        jb.count = (int32_t)p->mon_count[m][b];
        maxc = std::max<int64_t>(maxc, jb.count);
        jb.raw_off = td.out_off[m] + (T > 1 ? l : 0);
        jb.rsj = td.out_sj[m];
        jb.rsv = td.out_sv[m];
        jb.rsn = td.out_si[m];
        const int A = p->n_area[m];
        if (T > 1) {
          jb.out_off = b;
          jb.osa = B;
          jb.osv = (int64_t)A * B;
          jb.osj = (int64_t)nv * A * B;
        } else {
          jb.out_off = p->mon_offset[m][b];
          jb.osa = 1;
          jb.osv = A;
          jb.osj = (int64_t)nv * A;
        }
      }
    }
    if (maxc == 0) continue;
    HIPCHK(p, hipMemcpyAsync(p->d_area_jobs[m].p, jobs.data(), jobs.size() * sizeof(AreaJob), hipMemcpyHostToDevice,
                             p->stream));
    HIPCHK(p, launch_areas(p->esize, p->ukind[m] == TVBK_MON_SAVG ? 1 : 0, p->d_raw[m].p, p->d_out[m].p,
                           (const AreaJob*)p->d_area_jobs[m].p, B, (int)maxc, nv, p->n_area[m],
                           (const int32_t*)p->d_area_of[m].p, (const double*)p->d_area_w[m].p, p->stream));
  }
  HIPCHK(p, hipEventRecord(p->ev1, p->stream));
  p->have_timing = true;
  for (size_t k = 0; k < p->tiles.size(); ++k) {
    const TileDesc& td = p->tiles[k];
    if (td.node0 != 0) continue;            // one tile per instance counts its steps
    for (int l = 0; l < td.n_inst; ++l) {
      const int b = td.inst0 + l;
      p->cur_step[b] += td.n_steps;
      if (p->remaining[b] >= 0) p->remaining[b] -= td.n_steps;
    }
  }
  p->last_n = n;
  if (d.noise == TVBK_NOISE_INJECTED) p->z_steps = 0;
  p->stim_tp = nullptr;   // a stimulus is indexed by the call's steps: set it again for the next call
  p->stim_steps = 0;
  return TVBK_OK;
}

int tvbk_sync(tvbk_plan* p) {
  if (!p) return fail(nullptr, TVBK_EINVAL, "null plan");
  HIPCHK(p, hipStreamSynchronize(p->stream));
  return TVBK_OK;
}

int tvbk_last_run_ms(tvbk_plan* p, double* ms) {
  if (!p || !ms) return fail(p, TVBK_EINVAL, "null argument");
  *ms = 0;
  if (!p->have_timing) return TVBK_OK;
  HIPCHK(p, hipEventSynchronize(p->ev1));
  float f = 0;
  HIPCHK(p, hipEventElapsedTime(&f, p->ev0, p->ev1));
  *ms = f;
  return TVBK_OK;
}

int tvbk_balloon(const void* x, int32_t x_elem_bytes, int64_t x_stride, int64_t n_samples, int64_t n_series,
                 double sample_period_ms, const double* params, int32_t flags, int32_t start, double* state,
                 double* y, void* stream) {
  if (!x || !params || !state || !y) return fail(nullptr, TVBK_EINVAL, "null argument");
  if (x_elem_bytes != 4 && x_elem_bytes != 8) return fail(nullptr, TVBK_EINVAL, "x_elem_bytes must be 4 or 8");
  if (n_samples < 0 || n_series <= 0 || x_stride < n_series)
    return fail(nullptr, TVBK_EINVAL, "bad sizes (n_series > 0, n_samples >= 0, x_stride >= n_series)");
  if (!(sample_period_ms > 0)) return fail(nullptr, TVBK_EINVAL, "sample period must be positive");
  if ((flags & 4) && !start) return fail(nullptr, TVBK_EINVAL, "mean subtraction needs a run from the start");
  if (flags & ~7) return fail(nullptr, TVBK_EINVAL, "unknown flags");
  double par[10];
  hipError_t e = hipMemcpy(par, params, sizeof(par), hipMemcpyDefault);
  if (e != hipSuccess) return fail(nullptr, TVBK_EHIP, std::string("params: ") + hipGetErrorString(e));
  e = launch_balloon(x, x_elem_bytes, x_stride, n_samples, n_series, par, flags & 1, (flags >> 1) & 1, start ? 1 : 0,
                     (flags >> 2) & 1, sample_period_ms, state, y, (hipStream_t)stream);
  if (e != hipSuccess) return fail(nullptr, TVBK_EHIP, std::string("balloon: ") + hipGetErrorString(e));
  return TVBK_OK;
}

int tvbk_probe_gather(int64_t bytes, int32_t piece_bytes, int32_t write, double* bytes_per_s) {
  if (!bytes_per_s) return fail(nullptr, TVBK_EINVAL, "null argument");
  if (bytes < ((int64_t)64 << 20) || bytes > ((int64_t)64 << 30))
    return fail(nullptr, TVBK_EINVAL, "probe buffer must hold 64 MiB to 64 GiB");
  // k = 0: 128-byte pieces (two per wave load)
  const int k = piece_bytes == 128 ? 0 : piece_bytes / 256;
  if (piece_bytes != 128 && (piece_bytes % 256 || (k != 1 && k != 2 && k != 4 && k != 8 && k != 16)))
    return fail(nullptr, TVBK_EINVAL, "piece_bytes must be 128, 256, 512, 1024, 2048 or 4096");
  if (write != 0 && write != 1) return fail(nullptr, TVBK_EINVAL, "write must be 0 or 1");
  int dev = 0, n_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return fail(nullptr, TVBK_EHIP, std::string("device: ") + hipGetErrorString(e));
  const int grid = n_cu > 0 ? n_cu : 1;          // one 16-wave workgroup per CU, as the step kernel
  // about 2 GiB read per launch; a multiple of 16 batches (every k divides it)
  const double per_batch = (double)grid * 16 * 16 * 256;
  int iters = (int)((double)(1u << 31) / per_batch);
  iters = iters < 16 ? 16 : iters - iters % 16;
  float *buf = nullptr, *out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float best = 0.f;
  e = hipMalloc(&buf, (size_t)bytes);
  if (e == hipSuccess) e = hipMalloc(&out, (size_t)grid * 1024 * sizeof(float));
  if (e == hipSuccess) e = hipMemset(buf, 0, (size_t)bytes);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = launch_probe_gather(buf, (size_t)bytes / 4, k, write != 0, grid, 16, out, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  for (int r = 0; r < 3 && e == hipSuccess; ++r) {
    float ms = 0.f;
    e = hipEventRecord(e0, nullptr);
    if (e == hipSuccess) e = launch_probe_gather(buf, (size_t)bytes / 4, k, write != 0, grid, iters, out, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess && (r == 0 || ms < best)) best = ms;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (buf) (void)hipFree(buf);
  if (out) (void)hipFree(out);
  if (e != hipSuccess) return fail(nullptr, TVBK_EHIP, std::string("probe: ") + hipGetErrorString(e));
  const double rd = (double)grid * 16 * iters * 16 * 256;
  const double wr = !write ? 0.0 : k == 0 ? (double)grid * 16 * iters * 256 : (double)grid * 16 * (iters / k) * k * 256;
  *bytes_per_s = (rd + wr) / (best * 1e-3);
  return TVBK_OK;
}

int tvbk_monitor_layout(tvbk_plan* p, int32_t mon, int32_t* layout, int64_t* total_bytes, int64_t* offset,
                        int64_t* count, int64_t* first_step) {
  if (!p) return fail(nullptr, TVBK_EINVAL, "null plan");
  if (mon < 0 || mon >= p->d.n_mon) return fail(p, TVBK_EINVAL, "monitor out of range");
  if (layout) *layout = p->T > 1 ? 0 : 1;
  if (total_bytes) *total_bytes = p->mon_bytes[mon];
  for (int b = 0; b < p->d.n_inst; ++b) {
    if (offset) offset[b] = p->mon_offset[mon][b];
    if (count) count[b] = p->mon_count[mon][b];
    if (first_step) first_step[b] = p->mon_first[mon][b];
  }
  return TVBK_OK;
}

int tvbk_read_monitor(tvbk_plan* p, int32_t mon, void* out, int64_t cap_bytes) {
  if (!p) return fail(nullptr, TVBK_EINVAL, "null plan");
  if (mon < 0 || mon >= p->d.n_mon) return fail(p, TVBK_EINVAL, "monitor out of range");
  if (cap_bytes < p->mon_bytes[mon]) return fail(p, TVBK_EINVAL, "output buffer too small");
  HIPCHK(p, hipStreamSynchronize(p->stream));
  if (p->mon_bytes[mon]) HIPCHK(p, hipMemcpy(out, p->d_out[mon].p, p->mon_bytes[mon], hipMemcpyDefault));
  return TVBK_OK;
}

int tvbk_monitor_device_ptr(tvbk_plan* p, int32_t mon, void** ptr, int64_t* bytes) {
  if (!p) return fail(nullptr, TVBK_EINVAL, "null plan");
  if (mon < 0 || mon >= p->d.n_mon) return fail(p, TVBK_EINVAL, "monitor out of range");
  if (ptr) *ptr = p->d_out[mon].p;
  if (bytes) *bytes = p->mon_bytes[mon];
  return TVBK_OK;
}

// ---- monitor state in the reference's stock layout (checkpoints)

namespace {

// Host copy of one tile's state for monitor `m`: its image (TemporalAverage
// sums, Bold interim sums), Bold HRF accumulators and stock ring.
struct MonTile {
  std::vector<unsigned char> img;
  std::vector<double> acc, stk;
};

int mon_tile_io(tvbk_plan* p, int m, const TileDesc& td, MonTile& mt, bool get, bool with_stock) {
  const tvbk_monitor& mo = p->d.mon[m];
  const int N = p->conns[td.conn].n_node;
  unsigned char* dimg = (unsigned char*)p->d_img.p + td.img_off;
  const hipMemcpyKind k = get ? hipMemcpyDeviceToHost : hipMemcpyHostToDevice;
  mt.img.resize(p->img_bytes);
  HIPCHK(p, get ? hipMemcpy(mt.img.data(), dimg, mt.img.size(), k) : hipMemcpy(dimg, mt.img.data(), mt.img.size(), k));
  if (mo.kind == TVBK_MON_BOLD) {
    double* dacc = (double*)p->d_bacc.p + td.bacc_off[m];
    mt.acc.resize((size_t)p->mon_Q[m] * mo.n_voi * N * p->T);
    const size_t ab = mt.acc.size() * 8;
    HIPCHK(p, get ? hipMemcpy(mt.acc.data(), dacc, ab, k) : hipMemcpy(dacc, mt.acc.data(), ab, k));
    if (with_stock && mo.keep_stock) {
      double* dst = (double*)p->d_bstock[m].p + td.bst_off[m];
      mt.stk.resize((size_t)mo.n_taps * mo.n_voi * N * p->T);
      const size_t sb = mt.stk.size() * 8;
      HIPCHK(p, get ? hipMemcpy(mt.stk.data(), dst, sb, k) : hipMemcpy(dst, mt.stk.data(), sb, k));
    }
  }
  return TVBK_OK;
}

int mon_args_ok(tvbk_plan* p, int32_t inst, int32_t mon, bool allow_all) {
  if (!p) return fail(nullptr, TVBK_EINVAL, "null plan");
  if (inst < (allow_all ? -1 : 0) || inst >= p->d.n_inst) return fail(p, TVBK_EINVAL, "instance out of range");
  if (mon < 0 || mon >= p->d.n_mon) return fail(p, TVBK_EINVAL, "monitor out of range");
  return TVBK_OK;
}

}  // namespace

int tvbk_export_monitor(tvbk_plan* p, int32_t inst, int32_t mon, double* stock, double* interim) {
  int rc = mon_args_ok(p, inst, mon, false);
  if (rc) return rc;
  const tvbk_monitor& mo = p->d.mon[mon];
  if (mo.kind == TVBK_MON_RAW) return TVBK_OK;
  if (mo.kind == TVBK_MON_BOLD && stock && !mo.keep_stock)
    return fail(p, TVBK_ESTATE, "the Bold stock is only kept by a monitor created with keep_stock");
  HIPCHK(p, hipStreamSynchronize(p->stream));
  const TileDesc& td = p->tiles[p->tile_of[inst]];
  const int N = p->conns[td.conn].n_node, nv = mo.n_voi, T = p->T, t = p->lane_of[inst];
  MonTile mt;
  if ((rc = mon_tile_io(p, mon, td, mt, true, stock != nullptr))) return rc;
  const unsigned char* reg = mt.img.data() + p->mon_lds_off[mon];
  if (mo.kind == TVBK_MON_TAVG) {
    if (!stock) return TVBK_OK;
    std::vector<double> out((size_t)mo.istep * nv * N, 0.0);
    for (int i = 0; i < N; ++i)
      for (int v = 0; v < nv; ++v) {
        double sum;
        if (p->esize == 8) {
          sum = reinterpret_cast<const double*>(reg)[(i * nv + v) * T + t];
        } else {   // Kahan pair (sum, compensation): the true sum is sum - compensation
          const float* f = reinterpret_cast<const float*>(reg);
          sum = (double)f[((i * nv + v) * 2 + 0) * T + t] - (double)f[((i * nv + v) * 2 + 1) * T + t];
        }
        out[(size_t)v * N + i] = sum;
      }
    HIPCHK(p, hipMemcpy(stock, out.data(), out.size() * 8, hipMemcpyDefault));
    return TVBK_OK;
  }
  // Bold
  if (interim) {
    std::vector<double> out((size_t)mo.interim_istep * nv * N, 0.0);
    const double* bi = reinterpret_cast<const double*>(reg);
    for (int i = 0; i < N; ++i)
      for (int v = 0; v < nv; ++v) out[(size_t)v * N + i] = bi[(i * nv + v) * T + t];
    HIPCHK(p, hipMemcpy(interim, out.data(), out.size() * 8, hipMemcpyDefault));
  }
  if (stock) {
    const int S = mo.n_taps;
    std::vector<double> out((size_t)S * nv * N);
    for (int sl = 0; sl < S; ++sl)
      for (int v = 0; v < nv; ++v)
        for (int i = 0; i < N; ++i)
          out[((size_t)sl * nv + v) * N + i] = mt.stk[(((size_t)sl * nv + v) * N + i) * T + t];
    HIPCHK(p, hipMemcpy(stock, out.data(), out.size() * 8, hipMemcpyDefault));
  }
  return TVBK_OK;
}

int tvbk_import_monitor(tvbk_plan* p, int32_t inst, int32_t mon, const double* stock, const double* interim) {
  int rc = mon_args_ok(p, inst, mon, true);
  if (rc) return rc;
  const tvbk_monitor& mo = p->d.mon[mon];
  if (mo.kind == TVBK_MON_RAW) return TVBK_OK;
  if (!stock) return fail(p, TVBK_EINVAL, "null stock");
  const int B = p->d.n_inst, T = p->T;
  const int b0 = inst < 0 ? 0 : inst, b1 = inst < 0 ? B : inst + 1;
  const int conn = conn_of_inst(p, b0);
  for (int b = b0; b < b1; ++b)
    if (conn_of_inst(p, b) != conn) return fail(p, TVBK_EINVAL, "inst=-1 needs one shared connectome");
  const int N = p->conns[conn].n_node, nv = mo.n_voi;
  const bool bold = mo.kind == TVBK_MON_BOLD;
  const int S = mo.n_taps, ii = bold ? mo.interim_istep : 1;
  std::vector<double> hs((size_t)(bold ? S : mo.istep) * nv * N), hi;
  HIPCHK(p, hipMemcpy(hs.data(), stock, hs.size() * 8, hipMemcpyDefault));
  if (bold && interim) {
    hi.resize((size_t)ii * nv * N);
    HIPCHK(p, hipMemcpy(hi.data(), interim, hi.size() * 8, hipMemcpyDefault));
  }
  HIPCHK(p, hipStreamSynchronize(p->stream));
  const std::vector<double>& taps = p->taps_host[mon];
  std::vector<double> pend;       // Bold: pending HRF sums [Q][n_voi][n_node] at step pend_step
  int64_t pend_step = -1;
  // tiles holding the instances, each read, changed lane by lane and written back once
  std::vector<int> tiles;
  for (int b = b0; b < b1; ++b)
    if (tiles.empty() || tiles.back() != p->tile_of[b]) tiles.push_back(p->tile_of[b]);
  std::sort(tiles.begin(), tiles.end());
  tiles.erase(std::unique(tiles.begin(), tiles.end()), tiles.end());
  for (int k : tiles) {
    const TileDesc& td = p->tiles[k];
    MonTile mt;
    if ((rc = mon_tile_io(p, mon, td, mt, true, true))) return rc;
    unsigned char* reg = mt.img.data() + p->mon_lds_off[mon];
    for (int l = 0; l < td.n_inst; ++l) {
      const int b = td.inst0 + l;
      if (b < b0 || b >= b1) continue;
      const int t = p->lane_of[b];
      const int64_t s = p->cur_step[b];
      if (!bold) {
        // TemporalAverage: samples of steps s - kf + 1 .. s sit in slots 0 .. kf - 1 (monitors.py:386)
        const int kf = (int)(s % mo.istep);
        for (int i = 0; i < N; ++i)
          for (int v = 0; v < nv; ++v) {
            double sum = 0.0;
            for (int j = 0; j < kf; ++j) sum = sum + hs[((size_t)j * nv + v) * N + i];
            if (p->esize == 8) {
              reinterpret_cast<double*>(reg)[(i * nv + v) * T + t] = sum;
            } else {
              float* f = reinterpret_cast<float*>(reg);
              const float sf = (float)sum;
              f[((i * nv + v) * 2 + 0) * T + t] = sf;
              f[((i * nv + v) * 2 + 1) * T + t] = (float)((double)sf - sum);
            }
          }
        continue;
      }
      // Bold interim window (monitors.py:944)
      const int kf = (int)(s % ii);
      double* bi = reinterpret_cast<double*>(reg);
      for (int i = 0; i < N; ++i)
        for (int v = 0; v < nv; ++v) {
          double sum = 0.0;
          if (!hi.empty())
            for (int j = 0; j < kf; ++j) sum = sum + hi[((size_t)j * nv + v) * N + i];
          bi[(i * nv + v) * T + t] = sum;
        }
      // pending emissions kk (kk * istep > s): the HRF sum over the interim
      // samples mp <= s / ii already taken, in the order the step kernel adds
      // them (ascending mp); sample mp sits in stock slot (mp - 1) mod S and
      // enters emission kk with age floor(kk * istep / ii) - mp in [0, S)
      // (monitors.py:948-960)
      const int Q = p->mon_Q[mon];
      if (pend_step != s) {   // every instance reads the same stock: one evaluation per distinct step
        pend_step = s;
        pend.assign((size_t)Q * nv * N, 0.0);
        const int64_t mpc = s / ii, kk0 = s / mo.istep + 1;
        for (int64_t kk = kk0; kk < kk0 + Q; ++kk) {
          const int64_t me = kk * mo.istep / ii;
          const int64_t lo = me - S + 1, hi_mp = std::min(me, mpc);
          double* dst = &pend[(size_t)(kk % Q) * nv * N];
          for (int64_t mp = lo; mp <= hi_mp; ++mp) {
            const int64_t age = me - mp;
            const double w = taps[age == 0 ? S - 1 : age - 1];
            const double* src = &hs[(size_t)(((mp - 1) % S + S) % S) * nv * N];
            for (int q = 0; q < nv * N; ++q) dst[q] = dst[q] + w * src[q];
          }
        }
      }
      for (int q = 0; q < Q; ++q)
        for (int v = 0; v < nv; ++v)
          for (int i = 0; i < N; ++i)
            mt.acc[(((size_t)q * nv + v) * N + i) * T + t] = pend[((size_t)q * nv + v) * N + i];
      if (mo.keep_stock)
        for (int sl = 0; sl < S; ++sl)
          for (int v = 0; v < nv; ++v)
            for (int i = 0; i < N; ++i)
              mt.stk[(((size_t)sl * nv + v) * N + i) * T + t] = hs[((size_t)sl * nv + v) * N + i];
    }
    if ((rc = mon_tile_io(p, mon, td, mt, false, true))) return rc;
  }
  return TVBK_OK;
}

int tvbk_read_flags(tvbk_plan* p, int32_t* flags) {
  if (!p || !flags) return fail(p, TVBK_EINVAL, "null argument");
  HIPCHK(p, hipStreamSynchronize(p->stream));
  HIPCHK(p, hipMemcpy(flags, p->d_flags.p, (size_t)p->d.n_inst * 4, hipMemcpyDefault));
  return TVBK_OK;
}

int tvbk_debug_normals(tvbk_plan* p, int32_t inst, int64_t step0, int64_t n_steps, double* z) {
  if (!p || !z) return fail(p, TVBK_EINVAL, "null argument");
  if (p->d.noise != TVBK_NOISE_PHILOX) return fail(p, TVBK_ESTATE, "plan noise source is not PHILOX");
  if (inst < 0 || inst >= p->d.n_inst) return fail(p, TVBK_EINVAL, "instance out of range");
  if (n_steps <= 0 || step0 < 0) return fail(p, TVBK_EINVAL, "need n_steps > 0 and step0 >= 0");
  const int N = p->conns[conn_of_inst(p, inst)].n_node, nsv = p->mi.nsv;
  const size_t bytes = (size_t)n_steps * nsv * N * 8;
  void* dz = nullptr;
  HIPCHK(p, hipMalloc(&dz, bytes));
  hipError_t e = launch_debug_normals(p->hsize == 8, nsv, N, p->stream_host[inst], step0, n_steps,
                                      (uint32_t)p->d.seed, (uint32_t)(p->d.seed >> 32), (double*)dz, p->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
  if (e == hipSuccess) e = hipMemcpy(z, dz, bytes, hipMemcpyDefault);
  hipFree(dz);
  if (e != hipSuccess) return fail(p, TVBK_EHIP, std::string("debug normals: ") + hipGetErrorString(e));
  return TVBK_OK;
}

int tvbk_debug_slots(tvbk_plan* p, int32_t conn, int64_t step, int32_t* slots) {
  if (!p || !slots) return fail(p, TVBK_EINVAL, "null argument");
  if (conn < 0 || conn >= (int)p->conns.size()) return fail(p, TVBK_EINVAL, "connectome out of range");
  if (step < 1) return fail(p, TVBK_EINVAL, "step must be >= 1");
  const ConnDesc& cd = p->conns[conn];
  void* dout = nullptr;
  HIPCHK(p, hipMalloc(&dout, (size_t)std::max(cd.nnz, 1) * 4));
  hipError_t e = launch_debug_slots((const uint32_t*)p->d_edges.p + cd.e_off, cd.nnz, cd.horizon, step,
                                    (int32_t*)dout, p->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
  if (e == hipSuccess && cd.nnz) e = hipMemcpy(slots, dout, (size_t)cd.nnz * 4, hipMemcpyDefault);
  hipFree(dout);
  if (e != hipSuccess) return fail(p, TVBK_EHIP, std::string("debug slots: ") + hipGetErrorString(e));
  return TVBK_OK;
}

}  // extern "C"
