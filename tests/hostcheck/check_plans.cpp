// Driver of the host check: builds plans shaped like the benchmark
// configurations and the edge cases of the C-ABI through tvbk.h, runs calls,
// reads monitors, uploads and exports histories, with the stand-in runtime
// of host_hip.cpp checking every kernel address.  Built with
// -fsanitize=address,undefined and -DTVBK_BOUNDS_CHECK; exit status 0 only
// when no sanitizer report, no bounds violation and no unexpected error.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "tvbk.h"

extern "C" int tvbk_hostcheck_violations();

namespace {

int g_fail = 0;

void expect(bool ok, const std::string& what) {
  if (!ok) {
    std::fprintf(stderr, "FAIL: %s (%s)\n", what.c_str(), tvbk_last_error(nullptr));
    ++g_fail;
  }
}

struct Conn {
  int n = 0, H = 0;
  std::vector<int32_t> rp, col, dl;
  std::vector<double> w;
  tvbk_connectome view() const {
    return tvbk_connectome{n, H, (int64_t)col.size(), rp.data(), col.data(), dl.data(), w.data()};
  }
};

// seeded synthetic connectome: symmetric pairs at `density`, delays in
// [dmin, dmax] steps, horizon dmax + 1
Conn synthetic(int n, double density, int dmin, int dmax, unsigned seed) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  std::vector<double> W((size_t)n * n, 0.0);
  std::vector<int> D((size_t)n * n, 0);
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      const int d = dmin + (int)(u(rng) * (dmax - dmin + 1));
      D[(size_t)i * n + j] = D[(size_t)j * n + i] = std::min(d, dmax);
      if (u(rng) < density) W[(size_t)i * n + j] = W[(size_t)j * n + i] = 0.01 + u(rng);
    }
  Conn c;
  c.n = n;
  c.H = dmax + 1;
  c.rp.push_back(0);
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j)
      if (W[(size_t)i * n + j] != 0.0) {
        c.col.push_back(j);
        c.dl.push_back(D[(size_t)i * n + j]);
        c.w.push_back(W[(size_t)i * n + j]);
      }
    c.rp.push_back((int32_t)c.col.size());
  }
  return c;
}

std::vector<double> hrf_taps(int S) {
  std::vector<double> t(S);
  for (int k = 0; k < S; ++k) t[k] = std::exp(-(double)k / 200.0) * std::sin((double)k / 100.0);
  return t;
}

struct Case {
  std::string name;
  std::vector<Conn> conns;
  std::vector<tvbk_connectome> views;
  std::vector<int32_t> inst_conn;
  std::vector<double> params, obs, taps;
  std::vector<int64_t> steps;
  tvbk_plan_desc d{};
};

int n_param(int model, int cfun, int scheme) {
  int nmp, ncp, nsv, nch;
  tvbk_model_info(model, cfun, &nmp, &ncp, &nsv, &nch);
  const bool sto = scheme == TVBK_EULER_STO || scheme == TVBK_HEUN_STO;
  return nmp + ncp + (sto ? nsv : 0);
}

void finish_desc(Case& c, int model, int cfun, int scheme, int prec, int noise, int B, int nobs, int nsv) {
  for (auto& cn : c.conns) c.views.push_back(cn.view());
  tvbk_plan_desc& d = c.d;
  d.abi_version = TVBK_ABI_VERSION;
  d.model = model; d.cfun = cfun; d.scheme = scheme; d.precision = prec; d.noise = noise;
  d.n_inst = B;
  d.n_conn = (int)c.views.size();
  d.conns = c.views.data();
  d.inst_conn = c.inst_conn.empty() ? nullptr : c.inst_conn.data();
  d.n_param = n_param(model, cfun, scheme);
  c.params.assign((size_t)B * d.n_param, 0.0);
  d.params = c.params.data();
  d.dt = 0.05;
  c.obs.assign((size_t)nobs * nsv, 0.0);
  for (int o = 0; o < nobs; ++o) c.obs[(size_t)o * nsv + o] = 1.0;
  d.n_obs = nobs;
  d.obs_coef = c.obs.data();
  d.seed = 42;
}

void set_params(Case& c, const std::vector<double>& base, std::mt19937_64& rng) {
  std::uniform_real_distribution<double> u(0.0, 1.0);
  const int np = c.d.n_param;
  for (int b = 0; b < c.d.n_inst; ++b)
    for (int k = 0; k < np; ++k) c.params[(size_t)b * np + k] = k < (int)base.size() ? base[k] : 1e-3;
  (void)u;
}

void add_tavg(Case& c, int istep, int nvoi) {
  tvbk_monitor& m = c.d.mon[c.d.n_mon++];
  m.kind = TVBK_MON_TAVG; m.istep = istep; m.n_voi = nvoi;
  for (int v = 0; v < nvoi; ++v) m.voi[v] = v;
}

void add_bold(Case& c, int istep, int ii, int nvoi, int keep) {
  c.taps = hrf_taps(5000);
  tvbk_monitor& m = c.d.mon[c.d.n_mon++];
  m.kind = TVBK_MON_BOLD; m.istep = istep; m.n_voi = nvoi; m.interim_istep = ii;
  m.n_taps = (int)c.taps.size(); m.taps = c.taps.data(); m.k1_v0 = 0.112; m.subtract_one = 1; m.keep_stock = keep;
  for (int v = 0; v < nvoi; ++v) m.voi[v] = v;
}

// create, run `calls` calls of `n` steps, read every monitor, export two
// instances, free; `lo`/`hi` initial state ranges
void exercise(Case& c, int64_t n, int calls, const std::vector<double>& lo, const std::vector<double>& hi) {
  std::fprintf(stderr, "case %s: %d instances\n", c.name.c_str(), c.d.n_inst);
  int64_t fp = 0;
  expect(tvbk_plan_footprint(&c.d, &fp) == TVBK_OK && fp > 0, c.name + ": footprint");
  tvbk_plan* p = nullptr;
  const int rc = tvbk_plan_create(&c.d, nullptr, &p);
  expect(rc == TVBK_OK, c.name + ": create");
  if (rc != TVBK_OK) return;
  int32_t kernel, tile, waves, lds, ntiles, split, nch;
  tvbk_plan_shape(p, &kernel, &tile, &waves, &lds);
  tvbk_plan_tiles(p, &ntiles, &split);
  tvbk_plan_channels(p, &nch);
  std::fprintf(stderr, "  kernel %d tile %d waves %d lds %d tiles %d split %d channels %d\n", kernel, tile, waves,
               lds, ntiles, split, nch);
  expect(tvbk_init_uniform(p, 7, lo.data(), hi.data()) == TVBK_OK, c.name + ": init");
  if (!c.steps.empty()) expect(tvbk_set_steps(p, c.steps.data()) == TVBK_OK, c.name + ": steps");
  int nmp, ncp, nsv, nch_model;
  tvbk_model_info(c.d.model, c.d.cfun, &nmp, &ncp, &nsv, &nch_model);
  std::vector<double> z;
  int maxn = 0;
  for (auto& cn : c.conns) maxn = std::max(maxn, cn.n);
  for (int k = 0; k < calls; ++k) {
    if (c.d.noise == TVBK_NOISE_INJECTED) {
      z.assign((size_t)n * nsv * maxn * c.d.n_inst, 0.5);
      expect(tvbk_set_noise(p, z.data(), n) == TVBK_OK, c.name + ": noise");
    }
    expect(tvbk_run_steps(p, n) == TVBK_OK, c.name + ": run");
    for (int m = 0; m < c.d.n_mon; ++m) {
      int32_t layout;
      int64_t tot;
      std::vector<int64_t> off(c.d.n_inst), cnt(c.d.n_inst), first(c.d.n_inst);
      expect(tvbk_monitor_layout(p, m, &layout, &tot, off.data(), cnt.data(), first.data()) == TVBK_OK,
             c.name + ": layout");
      std::vector<unsigned char> out((size_t)tot + 8);
      expect(tvbk_read_monitor(p, m, out.data(), tot) == TVBK_OK, c.name + ": read");
    }
  }
  std::vector<int32_t> flags(c.d.n_inst);
  expect(tvbk_read_flags(p, flags.data()) == TVBK_OK, c.name + ": flags");
  // history upload (both entry points) and export of the first and last instance
  for (int inst : {0, c.d.n_inst - 1}) {
    const int conn = c.inst_conn.empty() ? 0 : c.inst_conn[inst];
    const int H = c.conns[conn].H, N = c.conns[conn].n;
    std::vector<double> h((size_t)H * nch * N), s((size_t)nsv * N), h2(h.size()), s2(s.size());
    for (size_t q = 0; q < h.size(); ++q) h[q] = 0.25 + 1e-3 * (double)(q % 997);
    for (size_t q = 0; q < s.size(); ++q) s[q] = 0.5 + 1e-3 * (double)(q % 991);
    for (int noalloc = 0; noalloc < 2; ++noalloc) {
      const int r = noalloc ? tvbk_set_history_noalloc(p, inst, h.data(), 8, s.data(), 5)
                            : tvbk_set_history(p, inst, h.data(), 8, s.data(), 5);
      const size_t hb = ((size_t)H * nch * N * (c.d.precision == TVBK_F64 ? 8 : 4) + 255) / 256 * 256;
      const bool big = hb + s.size() * (c.d.precision == TVBK_F32 ? 4 : 8) > (size_t)TVBK_NOALLOC_MAX_BYTES;
      if (noalloc && big) { expect(r == TVBK_EINVAL, c.name + ": noalloc refuses a large history"); continue; }
      expect(r == TVBK_OK, c.name + ": set history");
      int64_t step = -1;
      expect(tvbk_export_state(p, inst, h2.data(), 8, s2.data(), &step) == TVBK_OK && step == 5,
             c.name + ": export");
      double eh = 0, es = 0;
      for (size_t q = 0; q < h.size(); ++q) eh = std::max(eh, std::fabs(h2[q] - h[q]));
      for (size_t q = 0; q < s.size(); ++q) es = std::max(es, std::fabs(s2[q] - s[q]));
      expect(eh < 1e-6 && es < 1e-6, c.name + ": history round trip");
    }
  }
  tvbk_plan_destroy(p);
}

}  // namespace

int main() {
  std::mt19937_64 rng(2024);
  const std::vector<double> mpr_base = {1.0, 0.0, 1.0, 15.0, -5.0, 0.0, 1.0, 0.0, 0.5, 0.0, 1e-3, 1e-3};
  const std::vector<double> g2d_base = {1.0, 0.0, -2.0, -10.0, 0.0, 0.02, 3.0, 1.0, 0.0, 1.0, 1.0, 1.0, 0.1, 0.0,
                                        1e-3, 1e-3};
  const std::vector<double> jr_base = {3.25, 22.0, 0.1, 0.05, 5.52, 0.0025, 0.56, 135.0, 1.0, 0.8, 0.25, 0.25,
                                       0.12, 0.32, 0.22, 0.0, 0.005, 6.0, 0.56, 1.0};
  const std::vector<double> mlo = {0.0, -2.0}, mhi = {2.0, 1.5};

  // C2: 68 nodes x 8192 instances, with and without the V-channel skip
  for (int skip = 0; skip < 2; ++skip) {
    Case c;
    c.name = skip ? "c2 skip" : "c2";
    c.conns.push_back(synthetic(68, 0.25, 40, 720, 1));
    finish_desc(c, TVBK_MODEL_MPR, TVBK_CFUN_LINEAR, TVBK_HEUN_STO, TVBK_MIXED, TVBK_NOISE_PHILOX, 8192, 2, 2);
    set_params(c, mpr_base, rng);
    add_tavg(c, 20, 2);
    add_bold(c, 40000, 80, 2, 0);
    c.d.max_steps = 2000;
    c.d.flags = skip ? TVBK_PLAN_SKIP_ZERO_COUPLING : 0;
    exercise(c, 2000, 2, mlo, mhi);
  }
  // C3: 1024 instances, own connectomes of 68 .. 998 nodes, uneven durations
  {
    Case c;
    c.name = "c3";
    const int sizes[5] = {68, 96, 192, 400, 998};
    for (int b = 0; b < 1024; ++b) {
      const int nn = sizes[rng() % 5];
      const double dens = 0.1 + 0.2 * (double)(rng() % 1000) / 1000.0;
      c.conns.push_back(synthetic(nn, dens, 4, 800, 100 + b));
      c.inst_conn.push_back(b);
      c.steps.push_back(200 + (int64_t)(rng() % 3000));
    }
    finish_desc(c, TVBK_MODEL_JR, TVBK_CFUN_SIGMOID_JR, TVBK_EULER_STO, TVBK_F32, TVBK_NOISE_PHILOX, 1024, 4, 6);
    c.d.dt = 0.1;
    set_params(c, jr_base, rng);
    add_tavg(c, 20, 4);
    add_bold(c, 20000, 40, 4, 0);
    c.d.max_steps = 2000;
    exercise(c, 2000, 2, {-0.01, -5, -0.5, -0.06, -0.2, -5}, {0.01, 5, 0.5, 0.06, 0.2, 5});
  }
  // C4-shaped requests: large and small connectomes, few and many instances
  const int req[4][2] = {{998, 512}, {300, 40}, {76, 1}, {150, 33}};
  for (auto& r : req) {
    Case c;
    c.name = "c4 n" + std::to_string(r[0]) + " b" + std::to_string(r[1]);
    c.conns.push_back(synthetic(r[0], 0.2, 33, 600, 7 + r[0]));
    finish_desc(c, TVBK_MODEL_G2D, TVBK_CFUN_LINEAR, TVBK_HEUN_STO, TVBK_MIXED, TVBK_NOISE_PHILOX, r[1], 1, 2);
    c.d.dt = 0.1;
    set_params(c, g2d_base, rng);
    add_tavg(c, 10, 1);
    c.d.max_steps = 1000;
    exercise(c, 1000, 2, {-2, -6}, {4, 6});
  }
  // edge cases: one node without edges; zero delays (dense kernel); forced
  // tiles; float64 batch; injected noise; afferent + area monitors; Bold stock
  {
    Case c;
    c.name = "one node";
    Conn one;
    one.n = 1; one.H = 1; one.rp = {0, 0};
    c.conns.push_back(one);
    finish_desc(c, TVBK_MODEL_MPR, TVBK_CFUN_LINEAR, TVBK_HEUN_DET, TVBK_F64, TVBK_NOISE_NONE, 3, 2, 2);
    set_params(c, mpr_base, rng);
    add_tavg(c, 20, 2);
    c.d.max_steps = 100;
    exercise(c, 100, 2, mlo, mhi);
  }
  {
    Case c;
    c.name = "zero delay dense";
    c.conns.push_back(synthetic(68, 0.25, 0, 0, 3));
    finish_desc(c, TVBK_MODEL_MPR, TVBK_CFUN_LINEAR, TVBK_HEUN_STO, TVBK_F32, TVBK_NOISE_PHILOX, 256, 2, 2);
    set_params(c, mpr_base, rng);
    add_tavg(c, 20, 2);
    c.d.max_steps = 400;
    exercise(c, 400, 2, mlo, mhi);
  }
  for (int tile : {16, 64}) {
    Case c;
    c.name = "tile " + std::to_string(tile);
    c.conns.push_back(synthetic(40, 0.5, 2, 300, 9));
    finish_desc(c, TVBK_MODEL_MPR, TVBK_CFUN_LINEAR, TVBK_HEUN_STO, TVBK_F32, TVBK_NOISE_INJECTED, 100, 2, 2);
    set_params(c, mpr_base, rng);
    add_tavg(c, 20, 2);
    add_bold(c, 400, 80, 2, 1);
    c.d.max_steps = 500;
    c.d.tile = tile;
    exercise(c, 500, 2, mlo, mhi);
  }
  {
    Case c;
    c.name = "f64 batch, afferent and area monitors";
    c.conns.push_back(synthetic(30, 0.4, 3, 200, 11));
    finish_desc(c, TVBK_MODEL_MPR, TVBK_CFUN_LINEAR, TVBK_HEUN_STO, TVBK_F64, TVBK_NOISE_PHILOX, 40, 2, 2);
    set_params(c, mpr_base, rng);
    static std::vector<int32_t> mask(30);
    for (int i = 0; i < 30; ++i) mask[i] = i % 3;
    add_tavg(c, 20, 2);
    tvbk_monitor& a = c.d.mon[c.d.n_mon++];
    a.kind = TVBK_MON_AFF_TAVG; a.istep = 10; a.n_voi = 2; a.voi[0] = 0; a.voi[1] = 1;
    tvbk_monitor& s = c.d.mon[c.d.n_mon++];
    s.kind = TVBK_MON_SAVG; s.istep = 5; s.n_voi = 1; s.voi[0] = 0; s.spatial_mask = mask.data(); s.n_area = 3;
    tvbk_monitor& g = c.d.mon[c.d.n_mon++];
    g.kind = TVBK_MON_GAVG; g.istep = 5; g.n_voi = 2; g.voi[0] = 0; g.voi[1] = 1;
    c.d.max_steps = 300;
    exercise(c, 300, 2, mlo, mhi);
  }
  const int v = tvbk_hostcheck_violations();
  std::fprintf(stderr, "host check: %d bounds violations, %d failures\n", v, g_fail);
  return (v || g_fail) ? 1 : 0;
}
