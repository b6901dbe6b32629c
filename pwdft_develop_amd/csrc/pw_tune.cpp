// Autotuner of the H*psi plans: measures candidate xy-slab FFT layouts and
// split-K tilings of the projector product on the device and keeps the
// fastest, with a text cache of the choices per shape.
//
// The reference's only tuning is FFTW's planner: Fourier::Initialize builds
// its plans with FFTW_MEASURE (src/SCF/fourier.cpp:62, 112-184), i.e. it
// times candidate algorithms for each transform shape once and keeps the
// fastest.  pw_autotune plays that part for this library's transforms and
// products; pw_batch_autotune does it for the batched persistent launch's
// work split.
//
// Cache file: one line per tuned shape, "<key> <choice...>", keys made of the
// device name and CU count, grid, sphere size, band count, precision and the
// stage ("slab" or "tn<nbeta>").  Lines are appended; the last line of a key
// wins.  A shape found in the cache is applied without timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "pw_internal.h"

namespace pw {

__global__ void k_tune_fill(float* __restrict__ x, long n, float scale) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    x[i] = scale * __sinf(0.7548776662f * (float)(i % 100003) + 0.1f);
}
__global__ void k_tune_fill_d(double* __restrict__ x, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    x[i] = sin(0.7548776662 * (double)(i % 100003) + 0.1);
}

static std::map<std::string, std::string> read_cache(const char* path) {
  std::map<std::string, std::string> m;
  if (!path) return m;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    const size_t sp = line.find(' ');
    if (sp == std::string::npos || line.empty() || line[0] == '#') continue;
    m[line.substr(0, sp)] = line.substr(sp + 1);
  }
  return m;
}

static void append_cache(const char* path, const std::string& key, const std::string& val) {
  if (!path) return;
  std::ofstream f(path, std::ios::app);
  f << key << " " << val << "\n";
}

static std::string device_tag(int device) {
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, device) != hipSuccess) return "dev";
  std::string name = p.gcnArchName;
  const size_t colon = name.find(':');
  if (colon != std::string::npos) name = name.substr(0, colon);
  return name + "-" + std::to_string(p.multiProcessorCount) + "cu";
}

// Median-free timing: the fastest of `reps` runs of body() after one warm run,
// by HIP events on the context's stream (ms).
template <typename Body>
static int time_min(hipStream_t s, int reps, Body body, double* ms) {
  hipEvent_t e0, e1;
  PW_HIP(hipEventCreate(&e0));
  PW_HIP(hipEventCreate(&e1));
  int rc = body();
  double best = 1e30;
  for (int r = 0; r < reps && rc == PW_OK; r++) {
    (void)hipEventRecord(e0, s);
    rc = body();
    (void)hipEventRecord(e1, s);
    if (hipEventSynchronize(e1) != hipSuccess) rc = fail(PW_ERR_HIP, "autotune: event wait failed");
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e0, e1);
    best = std::min(best, (double)t);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = best;
  return rc;
}

// reg: the register kernel with ly lanes per y column transform and sb
// bands per z plane (0: the default); else the LDS kernel's (cg, rp, threads)
struct SlabChoice {
  bool reg;
  int cg, rp, threads, ly, sb;
  bool persist = false;
  std::string str() const {
    if (reg) return "reg " + std::to_string(ly) + " " + std::to_string(sb) + (persist ? " p" : "");
    return "lds " + std::to_string(cg) + " " + std::to_string(rp) + " " + std::to_string(threads);
  }
};

static void apply_slab(pw_ctx::Plan& pl, const SlabChoice& ch) {
  pl.reg = ch.reg;
  pl.ly = ch.reg ? ch.ly : 0;
  pl.sb = ch.reg ? ch.sb : 0;
  pl.persist = ch.reg && ch.persist;
  pl.cg = ch.reg ? 0 : ch.cg;
  pl.rp = ch.reg ? 0 : ch.rp;
  pl.threads = ch.reg ? 0 : ch.threads;
}

static bool parse_slab(const std::string& v, SlabChoice* ch) {
  int cg = 0, rp = 0, th = 0, ly = 0, sb = 0;
  if (sscanf(v.c_str(), "reg %d %d", &ly, &sb) == 2 && (ly == 1 || ly == 2 || ly == 4) && sb >= 0) {
    *ch = SlabChoice{true, 0, 0, 0, ly, sb};
    ch->persist = v.size() > 2 && v.compare(v.size() - 2, 2, " p") == 0;
    return true;
  }
  if (sscanf(v.c_str(), "lds %d %d %d", &cg, &rp, &th) == 3 && cg >= 0 && rp > 0 &&
      (th == 256 || th == 512 || th == 1024)) {
    *ch = SlabChoice{false, cg, rp, th, 0, 0};
    return true;
  }
  return false;
}

}  // namespace pw

using namespace pw;

extern "C" int pw_autotune(pw_ctx* c, int dtype, int nb, const char* cache_path, char* summary, int len) {
  if (!c || nb < 1 || nb > c->max_bands) return fail(PW_ERR_ARG, "pw_autotune: bad argument");
  if (dtype != PW_F64 && dtype != PW_F32) return fail(PW_ERR_ARG, "unknown dtype %d", dtype);
  if (!c->have_vtot) return fail(PW_ERR_STATE, "vtot has not been set");
  const int esz = dtype == PW_F64 ? 16 : 8;
  pw_ctx::Plan& pl = c->plan[dtype];
  const std::string base = device_tag(c->device) + "|" + std::to_string(c->n1) + "x" + std::to_string(c->n2) + "x" +
                           std::to_string(c->n3) + "|npw" + std::to_string(c->npw) + "|nb" + std::to_string(nb) +
                           (dtype == PW_F64 ? "|f64" : "|f32");
  const std::string kslab = base + "|slab", ktn = base + "|tn" + std::to_string(c->nbeta);
  auto cache = read_cache(cache_path);
  std::string out;
  // blocks to time on (deterministic content)
  const long m = 2L * c->npw * nb;
  void *X = nullptr, *Y = nullptr, *W = nullptr;
  auto done = [&](int rc) {
    if (X) (void)hipFree(X);
    if (Y) (void)hipFree(Y);
    if (W) (void)hipFree(W);
    if (rc == PW_OK && summary && len > 0) snprintf(summary, (size_t)len, "%s", out.c_str());
    return rc;
  };
  const size_t es = dtype == PW_F64 ? 8 : 4;
  if (hipMalloc(&X, es * m) != hipSuccess || hipMalloc(&Y, es * m) != hipSuccess)
    return done(fail(PW_ERR_NOMEM, "pw_autotune: cannot allocate the timing blocks"));
  const int g = (int)std::min<long>((m + 255) / 256, 8192);
  if (dtype == PW_F64)
    hipLaunchKernelGGL(k_tune_fill_d, dim3(g), dim3(256), 0, c->stream, (double*)X, m);
  else
    hipLaunchKernelGGL(k_tune_fill, dim3(g), dim3(256), 0, c->stream, (float*)X, m, 1.f);
  // ---- xy slab layout ----
  SlabChoice best{false, 0, 0, 256, 0, 0};
  auto it = cache.find(kslab);
  if (it != cache.end() && parse_slab(it->second, &best) &&
      (!best.reg || has_reg_slab(c, dtype, best.ly, best.sb, best.persist))) {
    apply_slab(pl, best);
    out += "slab " + best.str() + " (cached)";
  } else {
    std::vector<SlabChoice> cand;
    for (int ly : {1, 2, 4})
      if (has_reg_slab(c, dtype, ly)) cand.push_back({true, 0, 0, 0, ly, 0});
    if (has_reg_slab(c, dtype, 1, 16)) cand.push_back({true, 0, 0, 0, 1, 16});
    if (has_reg_slab(c, dtype, 1, 0, true)) {
      SlabChoice p{true, 0, 0, 0, 1, 0};
      p.persist = true;
      cand.push_back(p);
    }
    const int nk = c->sp.kxmax + 1, npair = (c->n2 + 1) / 2;
    for (int rp : {1, 2, 4, 8, 16})
      if (rp <= npair && slab_lds_bytes(c, esz, rp, 0) <= 160 * 1024)
        for (int th : {256, 512}) cand.push_back({false, 0, rp, th, 0, 0});
    for (int cg : {nk, std::max(1, nk / 2), 8, 4}) {
      if (cg > nk) continue;
      for (int rp : {4, 8, 16, std::max(1, std::min(npair, nk * c->n2 / (2 * c->n1)))})
        if (rp <= npair && slab_lds_bytes(c, esz, rp, cg) <= 160 * 1024)
          for (int th : {512, 1024}) cand.push_back({false, cg, rp, th, 0, 0});
    }
    double tbest = 1e30;
    const pw_ctx::Plan saved = pl;
    int tried = 0;
    for (const SlabChoice& ch : cand) {
      apply_slab(pl, ch);
      double ms = 0;
      const int rc = time_min(c->stream, 3, [&] { return pw_mult_spinor(c, dtype, nb, X, Y); }, &ms);
      if (rc != PW_OK) {
        // a layout this device cannot launch: skip it
        continue;
      }
      tried++;
      if (ms < tbest) {
        tbest = ms;
        best = ch;
      }
    }
    if (tried == 0) {
      pl = saved;
      return done(fail(PW_ERR_HIP, "pw_autotune: no slab layout ran"));
    }
    apply_slab(pl, best);
    append_cache(cache_path, kslab, best.str());
    char b[96];
    snprintf(b, sizeof(b), " (%.4f ms over %d layouts)", tbest, tried);
    out += "slab " + best.str() + b;
  }
  // ---- projector product W = beta^T X: split-K tiling ----
  if (c->nbeta > 0) {
    int tm = 0, cap = 0;
    auto jt = cache.find(ktn);
    if (jt != cache.end() && sscanf(jt->second.c_str(), "%d %d", &tm, &cap) == 2 && (tm == 1 || tm == 2) && cap > 0) {
      pl.tn_tm = tm;
      pl.tn_cap = cap;
      out += "; tn " + std::to_string(tm) + " " + std::to_string(cap) + " (cached)";
    } else {
      if (hipMalloc(&W, es * (size_t)c->nbeta * nb) != hipSuccess)
        return done(fail(PW_ERR_NOMEM, "pw_autotune: cannot allocate the projection block"));
      const int K = 2 * c->npw;
      double tbest = 1e30;
      int btm = 1, bcap = 1024;
      for (int t : {1, 2})
        for (int cp : {256, 512, 1024, 2048}) {
          pl.tn_tm = t;
          pl.tn_cap = cp;
          double ms = 0;
          int rc;
          if (dtype == PW_F64)
            rc = time_min(c->stream, 3, [&] {
              return gemm_tn<double>(c, K, c->nbeta, nb, c->d_beta, K, (const double*)X, K, (double*)W, c->nbeta,
                                     nullptr, 0.0);
            }, &ms);
          else
            rc = time_min(c->stream, 3, [&] {
              return gemm_tn<float>(c, K, c->nbeta, nb, c->d_beta_f, K, (const float*)X, K, (float*)W, c->nbeta,
                                    nullptr, 0.f);
            }, &ms);
          if (rc == PW_OK && ms < tbest) {
            tbest = ms;
            btm = t;
            bcap = cp;
          }
        }
      pl.tn_tm = btm;
      pl.tn_cap = bcap;
      append_cache(cache_path, ktn, std::to_string(btm) + " " + std::to_string(bcap));
      char b[96];
      snprintf(b, sizeof(b), "; tn %d %d (%.4f ms)", btm, bcap, tbest);
      out += b;
    }
  }
  invalidate_graph(c);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return done(fail(PW_ERR_HIP, "pw_autotune: device error"));
  return done(PW_OK);
}
