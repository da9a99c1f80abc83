// Kernel argument structures shared by the step kernels and the host API.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tvbk.h"

#define TVBK_KMAX_MON 4
#define TVBK_KMAX_OBS 8
#define TVBK_KMAX_PAR 32
// static device scratch of the allocation-free entry points (bytes), and the
// devices per process it is looked up for
#define TVBK_SCRATCH_BYTES TVBK_NOALLOC_MAX_BYTES
#define TVBK_SCRATCH_MAX_DEVICES 64

namespace tvbk {

// kernel coupling id of Linear on the first coupling variable only (plans
// that skip zero-coefficient coupling channels; CF_LINEAR1 of the kernels)
constexpr int KCF_LINEAR1 = 3;

// One connectome as uploaded: CSR rows of packed (col | delay << 16) edges.
struct ConnDesc {
  int32_t n_node, horizon, nnz, d_min, max_row, pad;
  int64_t rp_off, e_off;   // element offsets into the concatenated arrays
};

// CSR arrays (global and LDS copies) carry this many padding entries past
// the last edge so that rows can be read in fixed-width blocks.
#define TVBK_ROW_PAD 32
// longest block of the time-blocked sequential loop, and the bytes of one
// monitor step record (StepInfo) it keeps per step and monitor in LDS
#define TVBK_TB_MAX 64
#define TVBK_STEPINFO_BYTES 48

// Element index of history value (slot, node i, channel ch, lane t) in a
// tile's ring.  Tiles of T > 1 instances keep a slot's values together
// ([slot][node][channel][T]: a wave's gather of one edge reads one run of
// channels x instances); single-instance tiles (T = 1) keep a node's slots
// together ([node][channel][slot]: consecutive steps of one edge read
// consecutive values, so a time block's gathers of an edge share lines).
// Host builds with TVBK_BOUNDS_CHECK assert that every coordinate lies in
// its range (the CPU check of the plans' addressing).
#if defined(TVBK_BOUNDS_CHECK) && !defined(__HIP_DEVICE_COMPILE__)
#define TVBK_CHECK(cond) ((cond) ? (void)0 : bounds_failed(#cond, __FILE__, __LINE__))
void bounds_failed(const char* cond, const char* file, int line);
#else
#define TVBK_CHECK(cond) ((void)0)
#endif
__host__ __device__ inline int64_t ring_index(int T, int H, int N, int nch, int slot, int i, int ch, int t) {
  TVBK_CHECK(slot >= 0 && slot < H && i >= 0 && i < N && ch >= 0 && ch < nch && t >= 0 && t < T);
  return T == 1 ? ((int64_t)i * nch + ch) * H + slot : (((int64_t)slot * N + i) * nch + ch) * T + t;
}

// One workgroup's tile of instances and where its data lives.
struct TileDesc {
  int32_t conn, n_inst, inst0;
  // target nodes [node0, node1) integrated by this tile; parts > 1: the
  // instance is split over `parts` tiles, which stage their history writes
  // at stage_off (elements) until the block's commit
  int32_t node0, node1, parts;
  int64_t stage_off;
  int64_t step0, n_steps;          // first step of this call, steps to run
  int64_t ring_off, img_off;       // ring elements, image bytes
  int64_t bacc_off[TVBK_KMAX_MON]; // Bold accumulators (doubles)
  int64_t bst_off[TVBK_KMAX_MON];  // Bold stock ring (doubles), monitors keeping one
  int64_t out_off[TVBK_KMAX_MON], out_sj[TVBK_KMAX_MON], out_sv[TVBK_KMAX_MON], out_si[TVBK_KMAX_MON];
  int64_t z_off, z_sk, z_ssv, z_si;  // injected noise addressing
};

struct MonArgs {
  int32_t kind, istep, nvoi, ii, S, Q, sub1, lds_off;
  int32_t rat;        // Bold: istep / ii when ii divides istep, else 0
  int32_t src;        // sample source: 0 observed state, 1 node coupling (afferent)
  int32_t voi[TVBK_KMAX_OBS];
  double k1v0;
  const double* taps;
  double* stock;      // Bold: ring of interim means [S][nvoi][N][T] per tile, or null
  // per-tile phase of the first step s0 of a call, filled on the device:
  // r0 = s0 % istep, qe0 = s0 / istep, rb0 = s0 % ii, qb0 = s0 / ii
  int32_t r0, rb0;
  int64_t qe0, qb0;
};

struct StepArgs {
  const TileDesc* tiles;
  const ConnDesc* conns;
  const uint32_t* edges;
  const void* weights;
  const int32_t* row_ptr;
  // per-plan edge arrays of the wave-uniform row path (one shared
  // connectome): offsets {delay * slot bytes, source * node bytes} and the
  // weights in the history precision
  const uint32_t* sofs;
  const void* swts;
  const int32_t* srp;   // row starts in those arrays (rows padded to multiples of 4)
  void* ring;
  unsigned char* img;
  double* bacc;
  void* out[TVBK_KMAX_MON];
  const double* z;
  const void* params;
  int32_t* flags;
  const uint64_t* streams;
  int32_t prow, spatial, T, n_mon, scheme, noise, csr_in_lds, zero_delay;
  int32_t uniform_rows;   // one target node per wave, edge records in scalar registers
  uint32_t key0, key1;
  double dt, sqrt_dt;
  int32_t lds_bold, lds_state, lds_c, lds_edge, lds_w, lds_rp;
  int32_t lds_par;   // per-lane parameters [prow][T] (state precision)
  // parameters equal for every instance and node: bit k set -> parameter k
  // is par_uval[k] (derived values, float64)
  uint64_t par_umask;
  double par_uval[TVBK_KMAX_PAR];
  // dense zero-delay kernel: W padded to [16*RB][KP] (float), KB = KP/4,
  // LDS offset of the double-buffered coupling-channel array X
  const float* dense_w;
  int32_t dense_kb, dense_rb, lds_x;
  int64_t img_bytes;
  int32_t obs_sv[TVBK_KMAX_OBS];
  double obs_coef[TVBK_KMAX_OBS][8];
  // region stimulus of this call: spatial[node] * temporal[k] for the state
  // variables in stim_mask (null temporal: none); afferent monitors present
  const double* stim_sp;
  const double* stim_tp;
  uint32_t stim_mask;
  int32_t has_aff;
  // this launch runs steps [k0, k1) of the call (a block of a split plan;
  // 0 and the call's length otherwise); history staging of split tiles
  int32_t k0, k1;
  void* stage;
  // regions of the tile image, node-major: byte offset in the HBM image
  // (every node), byte offset in the LDS image (the tile's node range from
  // position 0) and bytes per node; nodes held by the LDS arrays
  int32_t n_img_reg, lds_nodes;
  int32_t img_reg_off[1 + TVBK_KMAX_MON], img_reg_lds[1 + TVBK_KMAX_MON], img_reg_bpn[1 + TVBK_KMAX_MON];
  MonArgs mon[TVBK_KMAX_MON];
  // time-blocked sequential loop: steps per block (<= 1: off) and the LDS
  // offset of its monitor step records [tb_k][TVBK_KMAX_MON]
  int32_t tb_k, lds_si;
  // batch-lane tiles split by node range: the pipelined loop is not used.
  // pipe_k = 2: the pipelined loop takes two steps per pass (rows of
  // n_channel * T = 32 values, every delay >= 3, rings of H + 1 slots whose
  // last mirrors slot 0)
  int32_t no_pipe, pipe_k;
  // two-step loop: rows fewer than N / waves for the waves that carry an
  // extra integration pass
  int32_t row_shift, pad_rs;
};

// Node -> area reduction of one monitor's subsampled output (GlobalAverage,
// SpatialAverage): per instance b, samples j < count, voi v, area a.
struct AreaJob {
  int64_t raw_off, out_off;   // element offsets of instance b's data
  int64_t rsj, rsv, rsn;      // raw strides: sample, voi, node
  int64_t osj, osv, osa;      // output strides: sample, voi, area
  int32_t n_node, count;
};

struct InitArgs {
  const TileDesc* tiles;
  const ConnDesc* conns;
  void* ring;
  unsigned char* img;
  const void* params;
  const uint64_t* streams;
  int32_t prow, T, lds_state;
  uint32_t key0, key1;
  double lo[8], hi[8];
};

// The step kernel's loop (TVBK_LOOP_*) for a launch with arguments a whose
// coupling keeps nch history channels: stimulus and afferent monitors run
// in the sequential loop only; the time-blocked loop needs blocks of >= 2
// steps, no afferent monitor and no zero delay.
inline int step_loop(const StepArgs& a, int nch) {
  const bool pipe = a.uniform_rows && !a.zero_delay && nch * a.T <= 64 && !a.stim_tp && !a.has_aff && !a.no_pipe;
  if (pipe) return a.pipe_k == 2 ? TVBK_LOOP_PIPELINED2 : TVBK_LOOP_PIPELINED;
  return (a.tb_k > 1 && !a.has_aff && !a.zero_delay) ? TVBK_LOOP_TIME_BLOCKED : TVBK_LOOP_SEQUENTIAL;
}

hipError_t launch_step(int precision, int model, int cf, const StepArgs& a, int ntiles, int nthreads, size_t lds,
                       hipStream_t st);
hipError_t launch_init(int precision, int model, int cf, const InitArgs& a, int ntiles, hipStream_t st);
hipError_t launch_dense(int precision, int model, const StepArgs& a, int ntiles, int nthreads, size_t lds,
                        hipStream_t st);
// mode 0: mean over all nodes (numpy pairwise order); 1: sum over each
// area's nodes of weight[node] * x, in node order (area_of[node])
hipError_t launch_areas(int esize, int mode, const void* raw, void* out, const AreaJob* jobs, int n_jobs,
                        int max_count, int nvoi, int n_area, const int32_t* area_of, const double* weight,
                        hipStream_t st);
// split tiles: staged history of a block's steps into the ring slots
hipError_t launch_commit(int hsize, const TileDesc* tiles, const ConnDesc* conns, void* ring, const void* stage,
                         int T, int nch, int k0, int kn, int ntiles, hipStream_t st);
// Device address of the current device's static scratch array
// (TVBK_SCRATCH_BYTES, for the allocation-free entry points).
hipError_t scratch_device_ptr(void** ptr);
// Philox normals of one instance stream, z[k][sv][node] (float64)
hipError_t launch_debug_normals(bool f64, int nsv, int N, uint64_t stream, int64_t step0, int64_t n, uint32_t key0,
                                uint32_t key1, double* z, hipStream_t st);
hipError_t launch_debug_slots(const uint32_t* edges, int nnz, int H, int64_t step, int32_t* out, hipStream_t st);
// Balloon-Windkessel BOLD, one thread per series
hipError_t launch_balloon(const void* x, int x_elem_bytes, int64_t x_stride, int64_t n_samples, int64_t n_series,
                          const double* par, int linear, int classical, int start, int normalize,
                          double sample_period_ms, double* state, double* y, hipStream_t st);
// the gather-pattern calibration: 16 reads of 256-byte pieces per wave and
// batch in segments of k pieces at random offsets of buf (n_float floats),
// with wr one segment stored every k batches; out holds grid * 1024 floats
hipError_t launch_probe_gather(float* buf, size_t n_float, int k, bool wr, int grid, int iters, float* out,
                               hipStream_t st);
// scatter a compact history [H][nch][N] (history precision) / state
// [nsv][N] (state precision) into lane t of a tile, and the reverse gather.
hipError_t launch_scatter(int precision, bool gather, void* ring, unsigned char* img_state, void* hist,
                          void* state, int H, int N, int nch, int nsv, int T, int t, hipStream_t st);

}  // namespace tvbk
