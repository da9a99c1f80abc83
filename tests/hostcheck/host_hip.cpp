// Host-only stand-in for the HIP runtime and the library's kernel launchers,
// linked with tvbk_api.cpp (built with -DTVBK_BOUNDS_CHECK and
// -fsanitize=address,undefined) so that the plan construction, tile layout,
// split and offset arithmetic of the C-ABI run on the CPU under the
// sanitizers.  Device allocations are host allocations; every launcher
// checks, before doing anything, that each address the kernel would form
// from the plan's tile table lies inside the allocation it targets:
//   step kernel  ring (every slot, node of the tile's range, channel, lane),
//                tile image, Bold accumulators and stock, monitor outputs
//                for the call's samples, injected noise, split staging,
//                CSR / padded row arrays, the LDS carve against the launch;
//   commit       staging and ring of every split tile's block;
//   scatter      the history / state of one lane (performed, so that a
//                history upload can be exported back and compared).
// Violations are counted and printed; the driver fails on any.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>

#include "tvbk.h"
#include "tvbk_kernels.hpp"

namespace {

std::map<const char*, size_t> g_allocs;   // live allocations: base -> bytes
int g_violations = 0;
hipError_t g_last = hipSuccess;

void violation(const std::string& what) {
  if (g_violations < 20) std::fprintf(stderr, "bounds: %s\n", what.c_str());
  ++g_violations;
}

// bytes from p to the end of the allocation holding it (0: none)
size_t room(const void* p) {
  const char* c = static_cast<const char*>(p);
  auto it = g_allocs.upper_bound(c);
  if (it == g_allocs.begin()) return 0;
  --it;
  if (c < it->first || c >= it->first + it->second) return 0;
  return (size_t)(it->first + it->second - c);
}

// elements [lo, hi] of size es from base must lie in one allocation.
// TVBK_HOSTCHECK_PLANT: the check itself is tested -- every allocation is
// taken one element shorter than it is, which must produce violations.
void need(const void* base, int64_t lo, int64_t hi, int es, const char* what) {
  if (hi < lo) return;
  if (!base) { violation(std::string(what) + ": null buffer"); return; }
  if (lo < 0) { violation(std::string(what) + ": negative index " + std::to_string(lo)); return; }
#ifdef TVBK_HOSTCHECK_PLANT
  const size_t r0 = room(base), r = r0 >= (size_t)es ? r0 - es : 0;
#else
  const size_t r = room(base);
#endif
  if ((size_t)(hi + 1) * (size_t)es > r)
    violation(std::string(what) + ": element " + std::to_string(hi) + " of " + std::to_string(es) +
              " bytes past the allocation (" + std::to_string(r) + " bytes)");
}

int64_t emissions(int64_t s0, int64_t n, int istep) {
  if (n <= 0) return 0;
  const int64_t f = ((s0 + istep - 1) / istep) * istep, last = s0 + n - 1;
  return f > last ? 0 : (last - f) / istep + 1;
}

}  // namespace

extern "C" int tvbk_hostcheck_violations() { return g_violations; }

namespace tvbk {
void bounds_failed(const char* cond, const char* file, int line) {
  violation(std::string("ring_index: ") + cond + " (" + file + ":" + std::to_string(line) + ")");
}
}  // namespace tvbk

// ---- runtime -----------------------------------------------------------------
hipError_t hipMalloc(void** p, size_t n) {
  char* q = static_cast<char*>(std::malloc(n ? n : 1));
  if (!q) return hipErrorOutOfMemory;
  g_allocs[q] = n;
  *p = q;
  return hipSuccess;
}
hipError_t hipFree(void* p) {
  if (!p) return hipSuccess;
  auto it = g_allocs.find(static_cast<char*>(p));
  if (it == g_allocs.end()) { violation("hipFree of an unknown pointer"); return hipErrorInvalidValue; }
  g_allocs.erase(it);
  std::free(p);
  return hipSuccess;
}
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
  if (n) std::memmove(d, s, n);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t) { return hipMemcpy(d, s, n, k); }
hipError_t hipMemset(void* d, int v, size_t n) {
  if (n) std::memset(d, v, n);
  return hipSuccess;
}
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) { return hipMemset(d, v, n); }
hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int) {
  if (a == hipDeviceAttributeMaxSharedMemoryPerBlock) *v = 160 * 1024;
  else if (a == hipDeviceAttributeMultiprocessorCount) *v = 256;
  else *v = 0;
  return hipSuccess;
}
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) {
  static int tag;
  *s = reinterpret_cast<hipStream_t>(&tag);
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t* e) {
  static int tag;
  *e = reinterpret_cast<hipEvent_t>(&tag);
  return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) { *ms = 0.f; return hipSuccess; }
hipError_t hipPointerGetAttributes(hipPointerAttribute_t*, const void*) { return hipErrorInvalidValue; }
hipError_t hipGetLastError() { hipError_t e = g_last; g_last = hipSuccess; return e; }
const char* hipGetErrorString(hipError_t e) { return e == hipSuccess ? "no error" : "stand-in runtime error"; }
// no graph capture here: the library falls back to direct launches
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorNotSupported; }
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t*) { return hipErrorNotSupported; }
hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, hipGraphNode_t*, char*, size_t) { return hipErrorNotSupported; }
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorNotSupported; }
hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }

// ---- launchers ---------------------------------------------------------------
namespace tvbk {

static const int NCH_OF_MODEL_CF[4][4] = {{1, 0, 1, 1}, {2, 0, 2, 1}, {2, 1, 2, 1}, {1, 0, 1, 1}};

static void check_tile_common(const StepArgs& a, const TileDesc& td, const ConnDesc& cd, int nch, int hsize,
                              int esize, int nsv, int64_t n_call, int nmon_kind_bold[TVBK_KMAX_MON], const char* who) {
  const int T = a.T, N = cd.n_node, H = cd.horizon;
  // ring: the corners of every coordinate range through ring_index; the
  // two-step pipelined loop also reads and writes slot H, the mirror of
  // slot 0
  const int hslots = a.pipe_k == 2 ? H + 1 : H;
  for (int slot : {0, hslots - 1})
    for (int i : {td.node0, td.node1 - 1})
      for (int ch : {0, nch - 1})
        for (int t : {0, T - 1}) {
          const int64_t e = td.ring_off + ring_index(T, hslots, N, nch, slot, i, ch, t);
          need(a.ring, e, e, hsize, who);
        }
  need(a.img, td.img_off, td.img_off + a.img_bytes - 1, 1, "tile image");
  for (int m = 0; m < a.n_mon; ++m) {
    const MonArgs& mo = a.mon[m];
    if (nmon_kind_bold[m]) {
      need(a.bacc, td.bacc_off[m], td.bacc_off[m] + (int64_t)mo.Q * mo.nvoi * N * T - 1, 8, "Bold accumulators");
      if (mo.stock)
        need(mo.stock, td.bst_off[m], td.bst_off[m] + (int64_t)mo.S * mo.nvoi * N * T - 1, 8, "Bold stock");
    }
    // samples of this call (afferent monitors keep the phase of step s0 - 1)
    const int64_t cnt = emissions(td.step0 - (mo.src ? 1 : 0), n_call, mo.istep);
    if (cnt > 0) {
      const int es = mo.kind == TVBK_MON_BOLD ? 8 : esize;
      const int64_t hi = td.out_off[m] + (cnt - 1) * td.out_sj[m] + (int64_t)(mo.nvoi - 1) * td.out_sv[m] +
                         (int64_t)(N - 1) * td.out_si[m] + (T > 1 ? td.n_inst - 1 : 0);
      need(a.out[m], td.out_off[m], hi, es, "monitor output");
    }
  }
  if (a.noise == TVBK_NOISE_INJECTED && n_call > 0) {
    const int64_t hi = td.z_off + (n_call - 1) * td.z_sk + (int64_t)(nsv - 1) * td.z_ssv + (int64_t)(N - 1) * td.z_si +
                       (T > 1 ? td.n_inst - 1 : 0);
    need(a.z, td.z_off, hi, 8, "injected noise");
  }
}

hipError_t launch_step(int precision, int model, int cf, const StepArgs& a, int ntiles, int nthreads, size_t lds,
                       hipStream_t) {
  const int hsize = precision == 1 ? 8 : 4, esize = precision == 0 ? 4 : 8;
  const int nch = (cf == 1) ? 1 : (cf == 3 ? 1 : NCH_OF_MODEL_CF[model][0]);
  const int nsv = model == 2 ? 6 : model == 3 ? 1 : 2;
  if (nthreads % 64 || nthreads > 1024 || nthreads <= 0) violation("block size " + std::to_string(nthreads));
  if (lds > 160 * 1024) violation("dynamic LDS " + std::to_string(lds));
  const int offs[] = {a.lds_state, a.lds_c, a.lds_par};
  for (int o : offs)
    if (o < 0 || (size_t)o > lds) violation("LDS offset " + std::to_string(o) + " past " + std::to_string(lds));
  if (a.pipe_k == 2) {
    // two-step pipelined loop: four coupling buffers of the tile's nodes and
    // four monitor step records per monitor slot, inside the dynamic LDS
    if (!a.uniform_rows || nch * a.T != 32) violation("two-step loop without rows of 32 values");
    const int64_t cb = (int64_t)a.lds_c + 4LL * a.lds_nodes * nch * a.T * hsize;
    if ((size_t)cb > lds) violation("two-step coupling buffers past the LDS");
    if ((size_t)a.lds_si + 4u * TVBK_KMAX_MON * TVBK_STEPINFO_BYTES > lds) violation("two-step step records past the LDS");
  }
  int is_bold[TVBK_KMAX_MON] = {0};
  for (int m = 0; m < a.n_mon; ++m) {
    is_bold[m] = a.mon[m].kind == TVBK_MON_BOLD;
    if (a.mon[m].lds_off < 0 || (size_t)a.mon[m].lds_off > lds) violation("monitor LDS offset");
  }
  need(a.tiles, 0, ntiles - 1, sizeof(TileDesc), "tile table");
  for (int k = 0; k < ntiles; ++k) {
    const TileDesc& td = a.tiles[k];
    const ConnDesc& cd = a.conns[td.conn];
    if (td.n_steps <= a.k0) continue;
    const int64_t kfin = td.n_steps < a.k1 ? td.n_steps : a.k1;
    if (td.node0 < 0 || td.node1 > cd.n_node || td.node0 >= td.node1) violation("tile node range");
    if (td.n_inst < 1 || td.n_inst > a.T) violation("tile instances");
    check_tile_common(a, td, cd, nch, hsize, esize, nsv, td.n_steps, is_bold, "ring");
    need(a.edges, cd.e_off, cd.e_off + cd.nnz + TVBK_ROW_PAD - 1, 4, "CSR edges");
    need(a.weights, cd.e_off, cd.e_off + cd.nnz + TVBK_ROW_PAD - 1, hsize, "CSR weights");
    need(a.row_ptr, cd.rp_off, cd.rp_off + cd.n_node, 4, "CSR row starts");
    if (a.uniform_rows) {
      // padded rows are read in chunks of up to 32 entries past their start
      const int64_t last = (int64_t)a.srp[cd.n_node - 1];
      need(a.sofs, 0, 2 * (last + 32) - 1, 4, "padded edge offsets");
      need(a.swts, 0, (last + 32) * (hsize / 4) - 1, 4, "padded edge weights");
      need(a.srp, 0, cd.n_node, 4, "padded row starts");
    }
    if (td.parts > 1) {
      const int64_t rs = (int64_t)nch * a.T;
      const int64_t hi = td.stage_off + ((kfin - 1 - a.k0) * cd.n_node + (td.node1 - 1)) * rs + a.T - 1;
      need(a.stage, td.stage_off, hi, hsize, "split staging");
    }
  }
  return hipSuccess;
}

hipError_t launch_init(int precision, int model, int cf, const InitArgs& a, int ntiles, hipStream_t) {
  const int hsize = precision == 1 ? 8 : 4;
  const int nch = (cf == 1 || cf == 3) ? 1 : NCH_OF_MODEL_CF[model][0];
  for (int k = 0; k < ntiles; ++k) {
    const TileDesc& td = a.tiles[k];
    const ConnDesc& cd = a.conns[td.conn];
    for (int slot : {0, cd.horizon - 1})
      for (int i : {0, cd.n_node - 1})
        for (int ch : {0, nch - 1}) {
          const int64_t e = td.ring_off + ring_index(a.T, cd.horizon, cd.n_node, nch, slot, i, ch, a.T - 1);
          need(a.ring, e, e, hsize, "init ring");
        }
  }
  return hipSuccess;
}

hipError_t launch_dense(int, int, const StepArgs& a, int ntiles, int nthreads, size_t lds, hipStream_t) {
  if (lds > 160 * 1024) violation("dense LDS");
  need(a.tiles, 0, ntiles - 1, sizeof(TileDesc), "tile table");
  need(a.dense_w, 0, (int64_t)16 * a.dense_rb * 4 * a.dense_kb - 1, 4, "dense weights");
  return hipSuccess;
}

hipError_t launch_areas(int esize, int, const void* raw, void* out, const AreaJob* jobs, int n_jobs, int max_count,
                        int nvoi, int n_area, const int32_t*, const double*, hipStream_t) {
  for (int b = 0; b < n_jobs; ++b) {
    const AreaJob& j = jobs[b];
    if (j.count <= 0) continue;
    need(raw, j.raw_off, j.raw_off + (j.count - 1) * j.rsj + (nvoi - 1) * j.rsv + (j.n_node - 1) * j.rsn, esize,
         "area input");
    need(out, j.out_off, j.out_off + (j.count - 1) * j.osj + (nvoi - 1) * j.osv + (n_area - 1) * j.osa, esize,
         "area output");
  }
  return hipSuccess;
}

hipError_t launch_commit(int hsize, const TileDesc* tiles, const ConnDesc* conns, void* ring, const void* stage, int T,
                         int nch, int k0, int kn, int ntiles, hipStream_t) {
  for (int k = 0; k < ntiles; ++k) {
    const TileDesc& td = tiles[k];
    if (td.parts <= 1) continue;
    const ConnDesc& cd = conns[td.conn];
    int64_t nb = td.n_steps - k0;
    nb = nb < 0 ? 0 : (nb > kn ? kn : nb);
    const int64_t rs = (int64_t)nch * T;
    for (int64_t j = 0; j < nb; ++j) {
      const int slot = (int)((td.step0 + k0 + j) % cd.horizon);
      for (int i : {td.node0, td.node1 - 1}) {
        const int64_t re = td.ring_off + ring_index(T, cd.horizon, cd.n_node, nch, slot, i, nch - 1, T - 1);
        need(ring, re, re, hsize, "commit ring");
        need(stage, td.stage_off + (j * cd.n_node + i) * rs, td.stage_off + (j * cd.n_node + i) * rs + rs - 1, hsize,
             "commit staging");
      }
    }
  }
  return hipSuccess;
}

hipError_t launch_scatter(int precision, bool gather, void* ring, unsigned char* img_state, void* hist, void* state,
                          int H, int N, int nch, int nsv, int T, int t, hipStream_t) {
  const int hs = precision == 1 ? 8 : 4, es = precision == 0 ? 4 : 8;
  for (int slot = 0; slot < H; ++slot)
    for (int ch = 0; ch < nch; ++ch)
      for (int i = 0; i < N; ++i) {
        const int64_t q = ((int64_t)slot * nch + ch) * N + i;
        const int64_t r = ring_index(T, H, N, nch, slot, i, ch, t);
        need(ring, r, r, hs, "scatter ring");
        need(hist, q, q, hs, "scatter history");
        char* rp = static_cast<char*>(ring) + r * hs;
        char* hp = static_cast<char*>(hist) + q * hs;
        if (gather) std::memcpy(hp, rp, hs); else std::memcpy(rp, hp, hs);
      }
  for (int v = 0; v < nsv; ++v)
    for (int i = 0; i < N; ++i) {
      const int64_t u = (int64_t)v * N + i, r = ((int64_t)i * nsv + v) * T + t;
      need(img_state, r * es, r * es + es - 1, 1, "scatter state");
      need(state, u, u, es, "scatter state input");
      char* sp = reinterpret_cast<char*>(img_state) + r * es;
      char* up = static_cast<char*>(state) + u * es;
      if (gather) std::memcpy(up, sp, es); else std::memcpy(sp, up, es);
    }
  return hipSuccess;
}

hipError_t launch_debug_slots(const uint32_t*, int, int, int64_t, int32_t*, hipStream_t) { return hipSuccess; }
hipError_t launch_debug_normals(bool, int, int, uint64_t, int64_t, int64_t, uint32_t, uint32_t, double*, hipStream_t) {
  return hipSuccess;
}
hipError_t launch_balloon(const void*, int, int64_t, int64_t, int64_t, const double*, int, int, int, int, double,
                          double*, double*, hipStream_t) {
  return hipSuccess;
}
hipError_t launch_probe_gather(float*, size_t, int, bool, int, int, float*, hipStream_t) { return hipSuccess; }

hipError_t scratch_device_ptr(void** ptr) {
  static unsigned char* buf = static_cast<unsigned char*>(std::malloc(TVBK_SCRATCH_BYTES));
  g_allocs[reinterpret_cast<const char*>(buf)] = TVBK_SCRATCH_BYTES;
  *ptr = buf;
  return hipSuccess;
}

}  // namespace tvbk
