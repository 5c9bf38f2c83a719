// extern "C" entry points (declared in include/ipeps_hip.h).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <memory>
#include <numeric>
#include <sstream>
#include <string>
#include <thread>

#include "../../include/ipeps_hip.h"
#include "corr.h"
#include "ctmrg.h"

using namespace ipeps;

// A context is shared by its handle and by every environment made from it:
// the device resources go when the last of them is destroyed, so handles may
// be destroyed in any order.
struct ipeps_ctx {
  std::shared_ptr<Ctx> c;
};

struct ipeps_env {
  std::shared_ptr<Ctx> hold;  // declared first: outlives the tensors in env
  int dtype = IPEPS_F64;
  ctm::Env env;
  // captured rg_step (ipeps_env_run_graph); dropped whenever the options, A or
  // a tensor are set, and recaptured when the environment's buffers moved
  std::unique_ptr<ctm::StepGraph> graph;
};

// Independent environments advanced together by a pool of workers.  Each
// worker owns a context (HIP stream, allocator, solver handle) and the
// environments assigned to it; the workers run on their own host threads, so
// their streams keep the device busy concurrently while any one of them waits
// on a host-side truncation decision.
struct ipeps_batch {
  struct Inst {
    int worker = 0;
    int64_t chi = 0;
    double cost = 0.0;
    ctm::Env env;
    Tensor H1, H2;
    double E[2] = {0, 0};
    int64_t steps = 0;
    std::unique_ptr<ctm::StepGraph> graph;  // declared after env: destroyed first
    int fails = 0;                          // consecutive replays whose checks failed
    bool no_graph = false;                  // capture refused or replays kept failing
  };
  int device = 0;
  int dtype = IPEPS_F64;
  std::vector<std::shared_ptr<Ctx>> workers;  // declared before insts: outlives their tensors
  std::vector<Inst> insts;
  bool graphs = false;
  // captures, replays kept, replays undone (then run eagerly), eager steps
  std::atomic<long> n_capture{0}, n_replay{0}, n_undo{0}, n_eager{0};
};

namespace {

thread_local std::string g_err;

std::vector<int64_t> dims_of_graph(const ctm::StepGraph& g) {
  std::vector<int64_t> d;
  for (int i = 0; i < 4; ++i) {
    d.insert(d.end(), g.C[i].dims.begin(), g.C[i].dims.end());
    d.insert(d.end(), g.E[i].dims.begin(), g.E[i].dims.end());
  }
  return d;
}

template <typename F>
int guard(F&& f) {
  try {
    f();
    g_err.clear();
    return IPEPS_OK;
  } catch (const Error& e) {
    g_err = e.what();
    return e.code ? e.code : IPEPS_ERR_INTERNAL;
  } catch (const std::exception& e) {
    g_err = e.what();
    return IPEPS_ERR_INTERNAL;
  } catch (...) {
    g_err = "unknown error";
    return IPEPS_ERR_INTERNAL;
  }
}

DType host_dt(int dtype) {
  if (dtype == IPEPS_F64 || dtype == IPEPS_MIXED) return DType::F64;
  if (dtype == IPEPS_C128) return DType::C128;
  throw Error(IPEPS_ERR_ARG, "unknown dtype code " + std::to_string(dtype));
}

DType dev_dt(int dtype) {
  if (dtype == IPEPS_MIXED) return DType::F32;
  return host_dt(dtype);
}

Tensor upload(Ctx& c, DType hdt, DType ddt, const std::vector<int64_t>& dims, const void* host) {
  IPEPS_CHECK(host != nullptr, "null host pointer");
  Tensor t = c.empty(hdt, dims);
  to_device(c, t, host);
  if (ddt != hdt) t = op::to_dtype(c, t, ddt);
  return t;
}

void download(Ctx& c, const Tensor& t, DType hdt, void* host) {
  IPEPS_CHECK(host != nullptr, "null host pointer");
  Tensor h = t.dt == hdt ? t : op::to_dtype(c, t, hdt);
  to_host(c, h, host);
}

Tensor upload_H(Ctx& c, const double* H, int64_t hdim) {
  IPEPS_CHECK(H != nullptr && hdim > 0, "null Hamiltonian");
  Tensor t = c.empty(DType::C128, {hdim, hdim});
  to_device(c, t, H);
  return t;
}

void set_opt(ctm::Options& o, const std::string& key, double v) {
  if (key == "svd_mode") o.svd.mode = (int)v;
  else if (key == "oversample") o.svd.oversample = (int)v;
  else if (key == "power_iters") o.svd.power_iters = (int)v;
  else if (key == "full_below") o.svd.full_below = (int)v;
  else if (key == "seed") o.svd.seed = (uint64_t)v;
  else if (key == "eta") o.svd.eta = v;
  else if (key == "mixed_cutoff") o.svd.mixed_cutoff = v;
  else if (key == "fast_orth") o.svd.fast_orth = (int)v;
  else if (key == "fast_svd") o.svd.fast_svd = (int)v;
  else if (key == "gram_guard") o.svd.gram_guard = v;
  else if (key == "power_iters_max") o.svd.power_iters_max = (int)v;
  else if (key == "warm_start") o.svd.warm_start = v != 0.0;
  else if (key == "resid_tol") o.svd.resid_tol = v;
  else if (key == "complement_iters") {
    IPEPS_CHECK(v >= 0, "complement_iters must be >= 0");
    o.svd.complement_iters = (int)v;
  } else if (key == "complement_tol") o.svd.complement_tol = v;
  else if (key == "early_mixed") {
    int m = (int)v;
    IPEPS_CHECK(m == 0 || m == 1 || m == 3 || m == 6, "early_mixed must be 0, 1, 3 or 6");
    o.svd.early_mixed = m;
  } else if (key == "late_apps") {
    IPEPS_CHECK(v >= 0, "late_apps must be >= 0");
    o.svd.late_apps = (int)v;
  }
  else if (key == "canonical_edge") o.canonical_edge = v != 0.0;
  else if (key == "energy_fp64") o.energy_fp64 = v != 0.0;
  else if (key == "checkpoint") o.checkpoint = v != 0.0;
  else if (key == "fused_absorb") o.fused_absorb = v != 0.0;
  else if (key == "paired_moves") o.paired_moves = v != 0.0;
  else if (key == "graph_safe") o.graph_safe = v != 0.0;
  else if (key == "graph") o.graph = v != 0.0;
  else if (key == "energy_mixed") {
    int m = (int)v;
    IPEPS_CHECK(m == 1 || m == 3 || m == 6, "energy_mixed must be 1, 3 or 6");
    o.energy_mixed = (k::Mixed)m;
  } else if (key == "energy_mixed_back") {
    int m = (int)v;
    IPEPS_CHECK(m == 0 || m == 1 || m == 3 || m == 6, "energy_mixed_back must be 0, 1, 3 or 6");
    o.energy_mixed_back = m;
  } else if (key == "mixed_back" || key == "mixed_proj") {
    int m = (int)v;
    IPEPS_CHECK(m == 0 || m == 1 || m == 3 || m == 6, key + " must be 0, 1, 3 or 6");
    (key == "mixed_back" ? o.mixed_back : o.mixed_proj) = m;
  } else if (key == "mixed") {
    int m = (int)v;
    IPEPS_CHECK(m == 1 || m == 3 || m == 6, "mixed must be 1, 3 or 6");
    o.mixed = (k::Mixed)m;
  } else
    throw Error(IPEPS_ERR_ARG, "unknown option '" + key + "'");
}

void parse_opts(ctm::Options& o, const char* s) {
  if (!s) return;
  std::stringstream ss(s);
  std::string item;
  while (std::getline(ss, item, ';')) {
    if (item.empty()) continue;
    auto eq = item.find('=');
    IPEPS_CHECK(eq != std::string::npos, "options: expected key=value");
    set_opt(o, item.substr(0, eq), std::stod(item.substr(eq + 1)));
  }
}

Tensor& env_slot(ctm::Env& e, int which) {
  if (which >= 0 && which < 4) return e.C[which];
  if (which >= 4 && which < 8) return e.E[which - 4];
  if (which == 8) return e.A;
  throw Error(IPEPS_ERR_ARG, "env tensor index must be 0..8");
}

}  // namespace

extern "C" {

int ipeps_version(void) { return 10000; }

const char* ipeps_last_error(void) { return g_err.c_str(); }

int ipeps_ctx_create(int device, ipeps_ctx** out) {
  return guard([&] {
    IPEPS_CHECK(out, "null out");
    std::shared_ptr<Ctx> c(ctx_create(device), [](Ctx* p) { ctx_destroy(p); });
    auto* h = new ipeps_ctx();
    h->c = std::move(c);
    *out = h;
  });
}

int ipeps_ctx_destroy(ipeps_ctx* ctx) {
  return guard([&] {
    delete ctx;
  });
}

int ipeps_ctx_sync(ipeps_ctx* ctx) {
  return guard([&] {
    IPEPS_CHECK(ctx, "null ctx");
    ctx->c->sync();
  });
}

int ipeps_ctx_mem(ipeps_ctx* ctx, int64_t* live, int64_t* peak) {
  return guard([&] {
    IPEPS_CHECK(ctx, "null ctx");
    if (live) *live = (int64_t)ctx->c->pool->live_bytes();
    if (peak) *peak = (int64_t)ctx->c->pool->peak_bytes();
  });
}

int ipeps_ctx_set_timing(ipeps_ctx* ctx, int on) {
  return guard([&] {
    IPEPS_CHECK(ctx, "null ctx");
    ctx->c->resolve_timing();
    ctx->c->timing = on != 0;
    ctx->c->timing_detail = on >= 2;
  });
}

int ipeps_ctx_set_timing_filter(ipeps_ctx* ctx, const char* prefix) {
  return guard([&] {
    IPEPS_CHECK(ctx, "null ctx");
    ctx->c->resolve_timing();
    ctx->c->timing_filter = prefix ? prefix : "";
  });
}

int ipeps_tune_config(int mode, double min_flops) {
  return guard([&] { k::tune_config(mode, min_flops); });
}

int ipeps_tune_stats(int64_t out[3]) {
  return guard([&] {
    IPEPS_CHECK(out, "null argument");
    long v[3];
    k::tune_stats(v);
    for (int i = 0; i < 3; ++i) out[i] = v[i];
  });
}

int ipeps_tune_clear(void) {
  return guard([&] { k::tune_clear(); });
}

int ipeps_tune_save(const char* path) {
  return guard([&] {
    IPEPS_CHECK(path, "null path");
    k::tune_save(path);
  });
}

int ipeps_tune_load(const char* path) {
  return guard([&] {
    IPEPS_CHECK(path, "null path");
    k::tune_load(path);
  });
}

int ipeps_ctx_kernel_stats(ipeps_ctx* ctx, char* buf, int64_t cap, int reset) {
  return guard([&] {
    IPEPS_CHECK(ctx && buf && cap > 0, "bad arguments");
    Ctx& c = *ctx->c;
    c.resolve_timing();
    std::ostringstream s;
    s.precision(17);
    s << "{";
    bool first = true;
    for (auto& kv : c.kstats) {
      s << (first ? "" : ", ") << "\"" << kv.first << "\": [" << kv.second.launches << ", " << kv.second.ms << ", "
        << kv.second.flops << ", " << kv.second.bytes << "]";
      first = false;
    }
    s << "}";
    std::string str = s.str();
    IPEPS_CHECK((int64_t)str.size() < cap, "kernel_stats: buffer too small");
    std::memcpy(buf, str.c_str(), str.size() + 1);
    if (reset) c.kstats.clear();
  });
}

int ipeps_env_create(ipeps_ctx* ctx, int dtype, const int64_t dims[5], const void* A, ipeps_env** out) {
  return guard([&] {
    IPEPS_CHECK(ctx && out && dims, "null argument");
    for (int i = 0; i < 5; ++i) IPEPS_CHECK(dims[i] > 0, "A dims must be positive");
    IPEPS_CHECK(dims[0] == dims[1] && dims[1] == dims[2] && dims[2] == dims[3], "A must be (D,D,D,D,d)");
    Ctx& c = *ctx->c;
    auto* e = new ipeps_env();
    e->hold = ctx->c;
    e->dtype = dtype;
    e->env.dt = dev_dt(dtype);
    e->env.opt.mixed = k::Mixed::BF16X3;
    e->env.opt.svd.stats = std::make_shared<op::SvdStats>();
    try {
      TapeScope ts(nullptr);
      e->env.A = upload(c, host_dt(dtype), dev_dt(dtype), {dims[0], dims[1], dims[2], dims[3], dims[4]}, A);
      e->env.Ad = op::conj(c, e->env.A);
      ctm::init_ctm1(c, e->env);
    } catch (...) {
      delete e;
      throw;
    }
    *out = e;
  });
}

int ipeps_env_destroy(ipeps_env* env) {
  return guard([&] { delete env; });
}

int ipeps_env_set_option(ipeps_env* env, const char* key, double value) {
  return guard([&] {
    IPEPS_CHECK(env && key, "null argument");
    env->graph.reset();
    if (std::string(key) == "draws") {
      IPEPS_CHECK(value >= 0.0, "draws must be >= 0");
      env->env.draws = (uint64_t)value;
    } else {
      set_opt(env->env.opt, key, value);
    }
  });
}

int ipeps_env_get_draws(ipeps_env* env, int64_t* out) {
  return guard([&] {
    IPEPS_CHECK(env && out, "null argument");
    *out = (int64_t)env->env.draws;
  });
}

int ipeps_env_copy_state(ipeps_env* dst, const ipeps_env* src) {
  return guard([&] {
    IPEPS_CHECK(dst && src, "null argument");
    IPEPS_CHECK(dst->env.A.dims == src->env.A.dims, "copy_state: A shapes differ");
    Ctx& c = *dst->hold;
    auto cp = [&](const Tensor& t, DType dt) {
      Tensor r = c.empty(t.dt, t.dims);
      if (t.bytes()) HIP_CHECK(hipMemcpyAsync(r.data(), t.data(), t.bytes(), hipMemcpyDeviceToDevice, c.stream));
      return op::to_dtype(c, r, dt);
    };
    // the source's stream must have finished writing what is copied
    src->hold->sync();
    for (int i = 0; i < 4; ++i) {
      dst->env.C[i] = cp(src->env.C[i], dst->env.dt);
      dst->env.E[i] = cp(src->env.E[i], dst->env.dt);
      const DType wdt = dst->env.dt == DType::C128 ? DType::C128 : DType::F64;
      dst->env.warm[i] = src->env.warm[i].defined() && src->env.warm[i].dt == wdt ? cp(src->env.warm[i], wdt)
                                                                                   : Tensor();
    }
    dst->env.draws = src->env.draws;
    c.sync();
  });
}

int ipeps_env_projector_stats(ipeps_env* env, double out[14], int reset) {
  return guard([&] {
    IPEPS_CHECK(env && out, "null argument");
    auto& st = env->env.opt.svd.stats;
    if (!st) st = std::make_shared<op::SvdStats>();
    {
      std::lock_guard<std::mutex> g(st->mu);
      out[0] = (double)st->calls;
      out[1] = (double)st->power_iters;
      out[2] = (double)st->power_max;
      out[3] = st->resid_max;
      out[4] = (double)st->checks;
      out[5] = (double)st->cg_calls;
      out[6] = (double)st->cg_iters;
      out[7] = st->cg_resid_max;
      out[8] = (double)st->cg_fallbacks;
      out[9] = (double)st->orth_calls;
      out[10] = (double)st->orth_second;
      out[11] = (double)st->orth_householder;
      out[12] = (double)st->warm_starts;
      out[13] = (double)st->tri_fallbacks;
    }
    if (reset) st->reset();
  });
}

int ipeps_env_tensor_dims(ipeps_env* env, int which, int64_t dims[5], int* nd) {
  return guard([&] {
    IPEPS_CHECK(env && dims && nd, "null argument");
    Tensor& t = env_slot(env->env, which);
    *nd = t.nd();
    for (int i = 0; i < t.nd(); ++i) dims[i] = t.dims[i];
  });
}

int ipeps_env_get_tensor(ipeps_env* env, int which, void* host) {
  return guard([&] {
    IPEPS_CHECK(env, "null env");
    download(*env->hold, env_slot(env->env, which), host_dt(env->dtype), host);
  });
}

int ipeps_env_set_tensor(ipeps_env* env, int which, int nd, const int64_t* dims, const void* host) {
  return guard([&] {
    IPEPS_CHECK(env && dims && nd > 0 && nd <= 5, "bad arguments");
    std::vector<int64_t> d(dims, dims + nd);
    Ctx& c = *env->hold;
    env->graph.reset();
    Tensor t = upload(c, host_dt(env->dtype), dev_dt(env->dtype), d, host);
    env_slot(env->env, which) = t;
    if (which == 8) env->env.Ad = op::conj(c, t);
  });
}

int ipeps_env_run_graph(ipeps_env* env, int64_t chi, int steps, int64_t out[4]) {
  return guard([&] {
    IPEPS_CHECK(env && chi > 0 && steps >= 0, "bad arguments");
    Ctx& c = *env->hold;
    HIP_CHECK(hipSetDevice(c.device));
    TapeScope ts(nullptr);
    auto& e = env->env;
    int64_t st[4] = {0, 0, 0, 0};
    auto current = [&]() {
      const auto* g = env->graph.get();
      if (!g || g->chi != chi) return false;
      for (int i = 0; i < 4; ++i)
        if (g->C[i].data() != e.C[i].data() || g->E[i].data() != e.E[i].data()) return false;
      return true;
    };
    for (int s = 0; s < steps; ++s) {
      if (!current()) {
        env->graph.reset();
        env->graph = ctm::capture_step(c, e, chi);
        ++st[0];
      }
      ctm::replay_step(*env->graph, e);
      c.sync();
      if (ctm::finish_step(*env->graph, e, nullptr)) {
        ++st[1];
        continue;
      }
      ++st[2];
      ctm::rg_step(c, e, chi);
      ++st[3];
      ctm::adopt_step(*env->graph, e);
    }
    if (out)
      for (int i = 0; i < 4; ++i) out[i] = st[i];
  });
}

int ipeps_env_set_A(ipeps_env* env, const void* A) {
  return guard([&] {
    if (env) env->graph.reset();
    IPEPS_CHECK(env, "null env");
    Ctx& c = *env->hold;
    Tensor t = upload(c, host_dt(env->dtype), dev_dt(env->dtype), env->env.A.dims, A);
    env->env.A = t;
    env->env.Ad = op::conj(c, t);
  });
}

int ipeps_rg_step(ipeps_env* env, int64_t chi, double* s_out, int64_t s_cap, int64_t* s_len) {
  return guard([&] {
    IPEPS_CHECK(env && chi > 0, "bad arguments");
    TapeScope ts(nullptr);
    auto s = ctm::rg_step(*env->hold, env->env, chi);
    if (s_len) *s_len = (int64_t)s.size();
    if (s_out) {
      IPEPS_CHECK(s_cap >= (int64_t)s.size(), "rg_step: s_out too small");
      std::memcpy(s_out, s.data(), s.size() * sizeof(double));
    }
  });
}

int ipeps_run_ctm(ipeps_env* env, int64_t chi, int min_iter, int max_iter, double tol, int conv_mode,
                  const double* H1, const double* H2, int64_t hdim, int* iters, double* E_out, double* diffs_out,
                  int64_t diffs_cap, int64_t* n_diffs) {
  return guard([&] {
    IPEPS_CHECK(env && chi > 0 && max_iter >= 1, "bad arguments");
    Ctx& c = *env->hold;
    TapeScope ts(nullptr);
    Tensor h1, h2;
    if (conv_mode == 0) {
      h1 = upload_H(c, H1, hdim);
      h2 = upload_H(c, H2, hdim);
    }
    auto r = ctm::run_ctm(c, env->env, chi, min_iter, max_iter, tol, conv_mode, h1, h2);
    if (iters) *iters = r.iters;
    if (E_out) { E_out[0] = r.E[0]; E_out[1] = r.E[1]; }
    if (n_diffs) *n_diffs = (int64_t)r.diffs.size();
    if (diffs_out) {
      int64_t n = std::min<int64_t>(diffs_cap, (int64_t)r.diffs.size());
      std::memcpy(diffs_out, r.diffs.data(), n * sizeof(double));
    }
  });
}

int ipeps_gs_energy(ipeps_env* env, const double* H1, const double* H2, int64_t hdim, double* E_out, double* N_out) {
  return guard([&] {
    IPEPS_CHECK(env, "null env");
    Ctx& c = *env->hold;
    TapeScope ts(nullptr);
    Tensor out = ctm::energy(c, env->env, upload_H(c, H1, hdim), upload_H(c, H2, hdim));
    double h[8];
    to_host(c, out, h);
    if (E_out) { E_out[0] = h[0]; E_out[1] = h[1]; }
    if (N_out) { N_out[0] = h[2]; N_out[1] = h[3]; }
  });
}

int ipeps_gs_norm(ipeps_env* env, double* out) {
  return guard([&] {
    IPEPS_CHECK(env && out, "null argument");
    TapeScope ts(nullptr);
    ctm::gs_norm(*env->hold, env->env, out);
  });
}

int ipeps_step_adjoint(ipeps_env* env, int64_t chi, const double* H1, const double* H2, int64_t hdim, double* E_out,
                       void* gradA_host) {
  return guard([&] {
    IPEPS_CHECK(env && chi > 0, "bad arguments");
    Ctx& c = *env->hold;
    double E[2];
    Tensor g = ctm::step_and_adjoint(c, env->env, chi, upload_H(c, H1, hdim), upload_H(c, H2, hdim), E);
    if (E_out) { E_out[0] = E[0]; E_out[1] = E[1]; }
    if (gradA_host) download(c, g, host_dt(env->dtype), gradA_host);
    else c.sync();
  });
}

int ipeps_gs_energy_grad(ipeps_ctx* ctx, int dtype, const int64_t dims[5], const void* A0, const double* H1,
                         const double* H2, int64_t hdim, int64_t chi, int min_iter, int max_iter, int ad_max_iter,
                         double tol, const char* options, double* E_out, void* grad_out, int* iters_out) {
  return guard([&] {
    IPEPS_CHECK(ctx && dims, "null argument");
    Ctx& c = *ctx->c;
    ctm::Options o;
    parse_opts(o, options);
    Tensor A = upload(c, host_dt(dtype), host_dt(dtype), {dims[0], dims[1], dims[2], dims[3], dims[4]}, A0);
    auto r = ctm::gs_energy_grad(c, A, dev_dt(dtype), o, upload_H(c, H1, hdim), upload_H(c, H2, hdim), chi,
                                 min_iter, max_iter, ad_max_iter, tol);
    if (E_out) { E_out[0] = r.E[0]; E_out[1] = r.E[1]; }
    if (grad_out) download(c, r.grad, host_dt(dtype), grad_out);
    if (iters_out) { iters_out[0] = r.conv_iters; iters_out[1] = r.ad_iters; }
  });
}

int ipeps_gs_energy_4x4(ipeps_env* env, const int64_t* s_dims, const double* s1, const double* s2,
                        const double* s3, const double* s4, double* E_out, double* N_out) {
  return guard([&] {
    IPEPS_CHECK(env && s_dims && s1 && s2 && s3 && s4, "null argument");
    Ctx& c = *env->hold;
    // s_dims: (d, d, r1), (r1, d, d, r2), (r2, d, d, r3), (r3, d, d), as 4 x 4 int64 (unused trailing = 1)
    const int nds[4] = {3, 4, 4, 3};
    const double* ptrs[4] = {s1, s2, s3, s4};
    std::array<Tensor, 4> s;
    for (int i = 0; i < 4; ++i) {
      std::vector<int64_t> d(s_dims + 4 * i, s_dims + 4 * i + nds[i]);
      for (auto x : d) IPEPS_CHECK(x > 0, "gs_energy_4x4: factor dims must be positive");
      s[i] = upload(c, DType::C128, DType::C128, d, ptrs[i]);
    }
    const int64_t dphys = env->env.A.dims[4];
    IPEPS_CHECK(s[0].dims[0] == dphys && s[0].dims[1] == dphys && s[1].dims[0] == s[0].dims[2] &&
                    s[2].dims[0] == s[1].dims[3] && s[3].dims[0] == s[2].dims[3] && s[3].dims[1] == dphys,
                "gs_energy_4x4: factor shapes do not chain (s1 (d,d,r1), s2 (r1,d,d,r2), s3 (r2,d,d,r3), s4 (r3,d,d))");
    double out[4];
    ctm::energy_4x4(c, env->env, s, out);
    const double den = out[2] * out[2] + out[3] * out[3];
    IPEPS_CHECK(den > 0.0, "gs_energy_4x4: zero norm");
    if (E_out) {
      E_out[0] = (out[0] * out[2] + out[1] * out[3]) / den;
      E_out[1] = (out[1] * out[2] - out[0] * out[3]) / den;
    }
    if (N_out) { N_out[0] = out[2]; N_out[1] = out[3]; }
  });
}

int ipeps_run_gs_vjp(ipeps_env* env, const void* x, const double* H1, const double* H2, int64_t hdim, int64_t chi,
                     int min_iter, int ad_max_iter, double tol, double dE, double* E_out, void* grad_out,
                     int* steps_out) {
  return guard([&] {
    IPEPS_CHECK(env && x && chi > 0 && ad_max_iter >= 1, "bad arguments");
    Ctx& c = *env->hold;
    const DType hdt = host_dt(env->dtype);
    Tensor tx = upload(c, hdt, hdt, env->env.A.dims, x);
    auto r = ctm::run_gs_vjp(c, env->env, tx, upload_H(c, H1, hdim), upload_H(c, H2, hdim), chi, min_iter,
                             ad_max_iter, tol);
    if (E_out) { E_out[0] = r.E[0]; E_out[1] = r.E[1]; }
    if (grad_out) {
      Tensor gr = r.grad;
      if (dE != 1.0) {
        gr = op::clone(c, gr);
        k::scale_cplx(c, gr.dt, gr.numel(), dE, 0.0, gr.data(), gr.data());
      }
      download(c, gr, hdt, grad_out);
    }
    if (steps_out) *steps_out = r.ad_iters;
  });
}

struct ipeps_es_path {
  std::shared_ptr<Ctx> hold;  // the context whose pool holds the projectors
  std::shared_ptr<const ctm::ProjPath> path;
};

int ipeps_es_path_create(ipeps_env* env, int64_t chi, int steps, ipeps_es_path** out) {
  return guard([&] {
    IPEPS_CHECK(env && out && chi > 0 && steps >= 0, "bad arguments");
    IPEPS_CHECK(env->dtype == IPEPS_C128, "es_path: the environment must be complex (IPEPS_C128)");
    *out = nullptr;
    auto p = std::unique_ptr<ipeps_es_path>(new ipeps_es_path());
    p->hold = env->hold;
    p->path = ctm::nested_proj_path(*env->hold, env->env, chi, steps);
    *out = p.release();
  });
}

int ipeps_es_path_destroy(ipeps_es_path* p) {
  return guard([&] {
    if (!p) return;
    p->path.reset();
    delete p;
  });
}

int ipeps_es_path_moves(const ipeps_es_path* p, int64_t* moves) {
  return guard([&] {
    IPEPS_CHECK(p && moves, "null argument");
    *moves = (int64_t)p->path->P.size();
  });
}

static void es_grad_impl(ipeps_env* env, const ipeps_es_path* path, const double* H1, const double* H2, int64_t hdim,
                         double px, double py, const void* B, int64_t chi, int min_iter, int max_iter,
                         int ad_max_iter, double tol, void* gH_out, void* envB_out, double* E4_out, double* Nb_out) {
  IPEPS_CHECK(env && B && chi > 0, "bad arguments");
  IPEPS_CHECK(env->dtype == IPEPS_C128, "es_grad: the environment must be complex (IPEPS_C128)");
  Ctx& c = *env->hold;
  Tensor tB = upload(c, DType::C128, DType::C128, env->env.A.dims, B);
  const double pi = 3.14159265358979323846;
  auto r = ctm::es_grad(c, env->env, tB, upload_H(c, H1, hdim), upload_H(c, H2, hdim), px * pi, py * pi, chi,
                        min_iter, max_iter, ad_max_iter, tol, path ? path->path : nullptr);
  if (gH_out) download(c, r.gH, DType::C128, gH_out);
  if (envB_out) download(c, r.envB, DType::C128, envB_out);
  if (E4_out) *E4_out = r.E4;
  if (Nb_out) { Nb_out[0] = r.Nb[0]; Nb_out[1] = r.Nb[1]; }
}

int ipeps_es_grad_path(ipeps_env* env, const ipeps_es_path* path, const double* H1, const double* H2, int64_t hdim,
                       double px, double py, const void* B, int64_t chi, int min_iter, int max_iter,
                       int ad_max_iter, double tol, void* gH_out, void* envB_out, double* E4_out, double* Nb_out) {
  return guard([&] {
    IPEPS_CHECK(path, "null path");
    es_grad_impl(env, path, H1, H2, hdim, px, py, B, chi, min_iter, max_iter, ad_max_iter, tol, gH_out, envB_out,
                 E4_out, Nb_out);
  });
}

int ipeps_es_grad(ipeps_env* env, const double* H1, const double* H2, int64_t hdim, double px, double py,
                  const void* B, int64_t chi, int min_iter, int max_iter, int ad_max_iter, double tol, void* gH_out,
                  void* envB_out, double* E4_out, double* Nb_out) {
  return guard([&] {
    es_grad_impl(env, nullptr, H1, H2, hdim, px, py, B, chi, min_iter, max_iter, ad_max_iter, tol, gH_out, envB_out,
                 E4_out, Nb_out);
  });
}

int ipeps_tcon(ipeps_ctx* ctx, int dtype, int ndA, const int64_t* dimsA, const void* A, const int* la, int ndB,
               const int64_t* dimsB, const void* B, const int* lb, int64_t* dimsC, int* ndC, void* C_host,
               int64_t C_cap) {
  return guard([&] {
    IPEPS_CHECK(ctx && dimsA && dimsB && la && lb && ndA >= 0 && ndB >= 0, "bad arguments");
    IPEPS_CHECK(dtype == IPEPS_F64 || dtype == IPEPS_C128 || dtype == IPEPS_MIXED, "tcon: F64, C128 or MIXED");
    Ctx& c = *ctx->c;
    // MIXED: float host data, FP32 storage and FP32 products (the "6" setting)
    const DType dt = dtype == IPEPS_MIXED ? DType::F32 : host_dt(dtype);
    Tensor tA = upload(c, dt, dt, std::vector<int64_t>(dimsA, dimsA + ndA), A);
    Tensor tB = upload(c, dt, dt, std::vector<int64_t>(dimsB, dimsB + ndB), B);
    op::Prec saved = op::prec();
    if (dtype == IPEPS_MIXED) op::prec().mixed = k::Mixed::BF16X6;
    Tensor C;
    try {
      C = op::tcon_raw(c, tA, std::vector<int>(la, la + ndA), false, tB, std::vector<int>(lb, lb + ndB), false);
    } catch (...) {
      op::prec() = saved;
      throw;
    }
    op::prec() = saved;
    IPEPS_CHECK(C.nd() <= 16, "tcon: output rank > 16");
    if (ndC) *ndC = C.nd();
    if (dimsC) for (int i = 0; i < C.nd(); ++i) dimsC[i] = C.dims[i];
    if (C_host) {
      IPEPS_CHECK(C_cap >= C.numel(), "tcon: output buffer too small");
      to_host(c, C, C_host);
    }
  });
}

int ipeps_absorb(ipeps_ctx* ctx, int dtype, int fused, const int64_t* dimsX, const void* X, const int* lx,
                 const int64_t* dimsA, const void* A, const int* la, const void* Ad, const int* lad, const void* G,
                 void* Y_out, void* gX_out, void* gA_out, void* gAd_out) {
  return guard([&] {
    IPEPS_CHECK(ctx && dimsX && X && lx && dimsA && A && la && Ad && lad && Y_out, "bad arguments");
    IPEPS_CHECK(dtype == IPEPS_F64 || dtype == IPEPS_MIXED, "absorb: F64 or MIXED (float host data)");
    Ctx& c = *ctx->c;
    const DType dt = dtype == IPEPS_MIXED ? DType::F32 : DType::F64;
    const std::vector<int64_t> dx(dimsX, dimsX + 6), da(dimsA, dimsA + 5);
    Tensor tX = upload(c, dt, dt, dx, X), tA = upload(c, dt, dt, da, A), tAd = upload(c, dt, dt, da, Ad);
    op::Prec saved = op::prec();
    op::prec().fused_absorb = fused != 0;
    op::prec().mixed = k::Mixed::BF16X3;
    op::prec().back = k::Mixed::BF16X3;
    Tape tape;
    tape.active = G != nullptr;
    TapeScope ts(G ? &tape : nullptr);
    Tensor lX = tX, lA = tA, lAd = tAd;
    if (G) {
      lX = op::leaf(tX);
      lA = op::leaf(tA);
      lAd = op::leaf(tAd);
    }
    Tensor Y;
    try {
      Y = op::absorb(c, lX, std::vector<int>(lx, lx + 6), lA, std::vector<int>(la, la + 5), lAd,
                     std::vector<int>(lad, lad + 5));
    } catch (...) {
      op::prec() = saved;
      throw;
    }
    op::prec() = saved;
    to_host(c, Y, Y_out);
    if (G) {
      tape.active = false;
      Tensor tG = upload(c, dt, dt, Y.dims, G);
      IPEPS_CHECK(Y.node != nullptr, "absorb: no tape node");
      op::acc(c, Y.node, tG);
      op::backward(c, tape);
      auto out = [&](const Tensor& leafT, void* host) {
        if (!host) return;
        if (leafT.node->grad.defined()) to_host(c, leafT.node->grad, host);
        else std::memset(host, 0, leafT.bytes());
      };
      out(lX, gX_out);
      out(lA, gA_out);
      out(lAd, gAd_out);
    }
  });
}

int ipeps_absorb_noalloc(ipeps_ctx* ctx, const int64_t* dimsX, const float* X, const int* lx, const int64_t* dimsA,
                         const float* A, const int* la, const float* Ad, const int* lad, const float* G, float* Y,
                         float* gX, float* gA, float* gAd) {
  return guard([&] {
    IPEPS_CHECK(ctx && dimsX && lx && dimsA && la && lad, "bad arguments");
    Ctx& c = *ctx->c;
    HIP_CHECK(hipSetDevice(c.device));
    // no timing brackets: they create events and grow the event list
    struct TimingOff {
      Ctx& c;
      bool was;
      explicit TimingOff(Ctx& cc) : c(cc), was(cc.timing) { c.timing = false; }
      ~TimingOff() { c.timing = was; }
    } off(c);
    op::absorb_noalloc(c, std::vector<int64_t>(dimsX, dimsX + 6), X, std::vector<int>(lx, lx + 6),
                       std::vector<int64_t>(dimsA, dimsA + 5), A, std::vector<int>(la, la + 5), Ad,
                       std::vector<int>(lad, lad + 5), G, Y, gX, gA, gAd);
  });
}

int ipeps_gemm(ipeps_ctx* ctx, int dtype, int mixed, int transA, int conjA, int transB, int conjB, int64_t M,
               int64_t N, int64_t K, double alpha, const void* A, int64_t lda, const void* B, int64_t ldb,
               double beta, void* C, int64_t ldc) {
  return guard([&] {
    IPEPS_CHECK(ctx && A && B && C && M >= 0 && N >= 0 && K >= 0, "bad arguments");
    Ctx& c = *ctx->c;
    DType dt = dev_dt(dtype);
    const int64_t ar = transA ? K : M, ac = transA ? M : K, br = transB ? N : K, bc = transB ? K : N;
    IPEPS_CHECK(lda >= std::max<int64_t>(1, ar) && ldb >= std::max<int64_t>(1, br) && ldc >= std::max<int64_t>(1, M),
                "gemm: leading dimension too small");
    Tensor tA = c.empty(dt, {lda, std::max<int64_t>(1, ac)}), tB = c.empty(dt, {ldb, std::max<int64_t>(1, bc)});
    Tensor tC = c.empty(dt, {ldc, std::max<int64_t>(1, N)});
    to_device(c, tA, A);
    to_device(c, tB, B);
    to_device(c, tC, C);
    k::GemmArgs g;
    g.dt = dt;
    g.transA = transA; g.conjA = conjA; g.transB = transB; g.conjB = conjB;
    g.M = M; g.N = N; g.K = K; g.alpha = alpha; g.beta = beta;
    g.A = tA.data(); g.lda = lda;
    g.B = tB.data(); g.ldb = ldb;
    g.C = tC.data(); g.ldc = ldc;
    if (dtype == IPEPS_MIXED) {
      IPEPS_CHECK(mixed == 1 || mixed == 3 || mixed == 6, "gemm: mixed must be 1, 3 or 6");
      g.mixed = (k::Mixed)mixed;
    }
    k::gemm(c, g);
    to_host(c, tC, C);
  });
}

int ipeps_svd(ipeps_ctx* ctx, int dtype, int64_t m, int64_t n, const void* A, void* U, double* S, void* V) {
  return guard([&] {
    IPEPS_CHECK(ctx && m > 0 && n > 0, "bad arguments");
    IPEPS_CHECK(dtype == IPEPS_F64 || dtype == IPEPS_C128, "svd: F64 or C128");
    Ctx& c = *ctx->c;
    DType dt = host_dt(dtype);
    Tensor tA = upload(c, dt, dt, {m, n}, A);
    Tensor tU, tS, tV;
    op::svd_full_raw(c, tA, tU, tS, tV);
    download(c, tU, dt, U);
    to_host(c, tS, S);
    download(c, tV, dt, V);
  });
}

int ipeps_svd_thin(ipeps_ctx* ctx, int dtype, int method, double gram_guard, int64_t m, int64_t n, const void* A, void* U,
                   double* S, void* V, int* method_used) {
  return guard([&] {
    IPEPS_CHECK(ctx && m >= n && n > 0 && A && U && S && V, "bad arguments (m >= n > 0 required)");
    IPEPS_CHECK(dtype == IPEPS_F64 || dtype == IPEPS_C128, "svd_thin: F64 or C128");
    IPEPS_CHECK(method >= 0 && method <= 3, "svd_thin: method must be 0, 1, 2 or 3");
    IPEPS_CHECK(method != 2 || (dtype == IPEPS_F64 && n <= k::JAC_MAXN_HOST), "svd_thin: method 2 needs F64, n <= 112");
    Ctx& c = *ctx->c;
    DType dt = host_dt(dtype);
    Tensor tA = upload(c, dt, dt, {m, n}, A);
    Tensor tU, tS, tV;
    int used = op::svd_tall(c, tA, method, gram_guard, tU, tS, tV);
    download(c, tU, dt, U);
    to_host(c, tS, S);
    download(c, tV, dt, V);
    if (method_used) *method_used = used;
  });
}

int ipeps_orth(ipeps_ctx* ctx, int dtype, int method, int64_t m, int64_t n, const void* Y, void* Q) {
  return guard([&] {
    IPEPS_CHECK(ctx && m >= n && n > 0 && Y && Q, "bad arguments (m >= n > 0 required)");
    IPEPS_CHECK(dtype == IPEPS_F64 || dtype == IPEPS_C128, "orth: F64 or C128");
    IPEPS_CHECK(method >= 0 && method <= 2, "orth: method must be 0, 1 or 2");
    Ctx& c = *ctx->c;
    DType dt = host_dt(dtype);
    Tensor tY = upload(c, dt, dt, {m, n}, Y);
    download(c, op::orth(c, tY, method), dt, Q);
  });
}

int ipeps_svd_back(ipeps_ctx* ctx, int dtype, int64_t N, int64_t n, const void* U, const double* S, const void* V,
                   const void* dU, const double* dS, const void* dV, double eta, void* dA_out) {
  return guard([&] {
    IPEPS_CHECK(ctx && N > 0 && n > 0 && n <= N && U && S && V && dA_out, "bad arguments");
    IPEPS_CHECK(dtype == IPEPS_F64 || dtype == IPEPS_C128, "svd_back: F64 or C128");
    Ctx& c = *ctx->c;
    DType dt = host_dt(dtype);
    Tensor tU = upload(c, dt, dt, {N, N}, U), tV = upload(c, dt, dt, {N, N}, V);
    Tensor tS = upload(c, DType::F64, DType::F64, {N}, S);
    Tensor UdU, VdV, tdS;
    if (dU) UdU = op::matmul_raw(c, tU, true, true, upload(c, dt, dt, {N, n}, dU), false, false);
    if (dV) VdV = op::matmul_raw(c, tV, true, true, upload(c, dt, dt, {N, n}, dV), false, false);
    if (dS) tdS = upload(c, DType::F64, DType::F64, {n}, dS);
    Tensor G = c.empty(dt, {N, N});
    k::svd_gamma(c, dt, N, n, tS.ptr<double>(), UdU.defined() ? UdU.data() : nullptr,
                 VdV.defined() ? VdV.data() : nullptr, tdS.defined() ? tdS.ptr<double>() : nullptr, eta, G.data());
    Tensor R = op::matmul_raw(c, op::matmul_raw(c, tU, false, false, G, false, false), false, false, tV, true, true);
    download(c, R, dt, dA_out);
  });
}

// ---- transfer matrix and correlations -----------------------------------------

// E tensors of a mixed environment widened to FP64 for the transfer matrix
static ctm::Env tm_env(Ctx& c, const ctm::Env& e) {
  if (e.dt != DType::F32) return e;
  ctm::Env r = e;
  r.dt = DType::F64;
  for (int i = 0; i < 4; ++i) r.E[i] = op::to_dtype(c, e.E[i], DType::F64);
  return r;
}

int ipeps_correlation_spin(ipeps_env* env, const double* op1, const double* op2, int64_t d, int max_iter,
                           int direction, double* s2s_out, double* m1_out, double* m2_out) {
  return guard([&] {
    IPEPS_CHECK(env && op1 && op2 && s2s_out && max_iter >= 1, "bad arguments");
    Ctx& c = *env->hold;
    TapeScope ts(nullptr);
    std::vector<double> s2s;
    double m1[2], m2[2];
    ctm::correlation_spin(c, env->env, upload_H(c, op1, d), upload_H(c, op2, d), max_iter, direction, s2s, m1, m2);
    std::memcpy(s2s_out, s2s.data(), s2s.size() * sizeof(double));
    if (m1_out) { m1_out[0] = m1[0]; m1_out[1] = m1[1]; }
    if (m2_out) { m2_out[0] = m2[0]; m2_out[1] = m2[1]; }
  });
}

int ipeps_tm_dims(ipeps_env* env, int direction, int64_t* rows, int64_t* cols) {
  return guard([&] {
    IPEPS_CHECK(env && rows && cols, "bad arguments");
    ctm::tm_dims(env->env, direction, *rows, *cols);
  });
}

int ipeps_tm_dense(ipeps_env* env, int direction, void* TM_out) {
  return guard([&] {
    IPEPS_CHECK(env && TM_out, "bad arguments");
    Ctx& c = *env->hold;
    TapeScope ts(nullptr);
    Tensor TM = ctm::tm_dense(c, tm_env(c, env->env), direction);
    download(c, TM, host_dt(env->dtype), TM_out);
  });
}

int ipeps_tm_apply(ipeps_env* env, int direction, int64_t nvec, const void* X, void* Y) {
  return guard([&] {
    IPEPS_CHECK(env && X && Y && nvec >= 1, "bad arguments");
    Ctx& c = *env->hold;
    TapeScope ts(nullptr);
    int64_t rows, cols;
    ctm::tm_dims(env->env, direction, rows, cols);
    ctm::Env e = tm_env(c, env->env);
    const DType hdt = host_dt(env->dtype);
    Tensor tX = upload(c, hdt, e.E[0].dt, {cols, nvec}, X);
    Tensor tY = ctm::tm_apply(c, e, direction, tX);
    download(c, tY, hdt, Y);
  });
}

// ---- batched independent environments -------------------------------------

int ipeps_batch_create(int device, int n, int dtype, const int64_t* dims, const void* const* A, const int64_t* chi,
                       const double* const* H1, const double* const* H2, const int64_t* hdim, int workers,
                       const char* options, ipeps_batch** out) {
  return guard([&] {
    IPEPS_CHECK(out && n > 0 && dims && A && chi && H1 && H2 && hdim, "bad arguments");
    IPEPS_CHECK(workers >= 1 && workers <= 64, "workers must be 1..64");
    ctm::Options opt;
    opt.mixed = k::Mixed::BF16X3;
    opt.paired_moves = false;  // the worker pool already runs instances concurrently
    parse_opts(opt, options);
    IPEPS_CHECK(!opt.graph || !opt.paired_moves, "graph=1 needs paired_moves=0");
    std::unique_ptr<ipeps_batch> b(new ipeps_batch());
    b->device = device;
    b->dtype = dtype;
    b->graphs = opt.graph;
    b->insts.resize(n);
    // longest-processing-time assignment by the per-step cost estimate
    // chi^3 D^4 + (chi D^2)^2 (chi + 16) of the corner and projector terms
    for (int i = 0; i < n; ++i) {
      const int64_t* d = dims + 5 * i;
      for (int a = 0; a < 5; ++a) IPEPS_CHECK(d[a] > 0, "A dims must be positive");
      IPEPS_CHECK(d[0] == d[1] && d[1] == d[2] && d[2] == d[3], "A must be (D,D,D,D,d)");
      IPEPS_CHECK(chi[i] > 0 && A[i] && H1[i] && H2[i] && hdim[i] == d[4] * d[4], "bad instance arguments");
      const double D2 = (double)(d[0] * d[0]), x = (double)chi[i];
      b->insts[i].chi = chi[i];
      b->insts[i].cost = x * x * x * D2 * D2 + (x * D2) * (x * D2) * (x + 16.0);
    }
    const int nw = std::min(workers, n);
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return b->insts[x].cost > b->insts[y].cost; });
    std::vector<double> load(nw, 0.0);
    for (int i : order) {
      int w = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      b->insts[i].worker = w;
      load[w] += b->insts[i].cost;
    }
    for (int w = 0; w < nw; ++w)
      b->workers.emplace_back(std::shared_ptr<Ctx>(ctx_create(device), [](Ctx* p) { ctx_destroy(p); }));
    for (int i = 0; i < n; ++i) {
      auto& in = b->insts[i];
      Ctx& c = *b->workers[in.worker];
      const int64_t* d = dims + 5 * i;
      TapeScope ts(nullptr);
      in.env.dt = dev_dt(dtype);
      in.env.opt = opt;
      in.env.A = upload(c, host_dt(dtype), dev_dt(dtype), {d[0], d[1], d[2], d[3], d[4]}, A[i]);
      in.env.Ad = op::conj(c, in.env.A);
      ctm::init_ctm1(c, in.env);
      in.H1 = upload_H(c, H1[i], hdim[i]);
      in.H2 = upload_H(c, H2[i], hdim[i]);
    }
    *out = b.release();
  });
}

int ipeps_batch_destroy(ipeps_batch* b) {
  return guard([&] { delete b; });
}

int ipeps_batch_run(ipeps_batch* b, int steps, double* E_out, double* seconds_out) {
  return guard([&] {
    IPEPS_CHECK(b && steps >= 0, "bad arguments");
    const int nw = (int)b->workers.size();
    std::vector<std::string> errs(nw);
    auto t0 = std::chrono::steady_clock::now();
    auto work = [&](int w) {
      try {
        HIP_CHECK(hipSetDevice(b->device));
        Ctx& c = *b->workers[w];
        TapeScope ts(nullptr);
        auto dims_of = [](const ctm::Env& e) {
          std::vector<int64_t> d;
          for (int i = 0; i < 4; ++i) {
            d.insert(d.end(), e.C[i].dims.begin(), e.C[i].dims.end());
            d.insert(d.end(), e.E[i].dims.begin(), e.E[i].dims.end());
          }
          return d;
        };
        for (int s = 0; s < steps; ++s) {
          // captured steps are enqueued first (no host round trip), then the
          // eager ones; an eager step with unchanged shapes captures the
          // instance's step for the following rounds
          std::vector<ipeps_batch::Inst*> replayed;
          for (auto& in : b->insts)
            if (in.worker == w && in.graph) {
              ctm::replay_step(*in.graph, in.env);
              replayed.push_back(&in);
            }
          for (auto& in : b->insts)
            if (in.worker == w && !in.graph) {
              const auto before = dims_of(in.env);
              ctm::rg_step(c, in.env, in.chi);
              ++in.steps;
              ++b->n_eager;
              if (b->graphs && !in.no_graph && dims_of(in.env) == before) {
                try {
                  in.graph = ctm::capture_step(c, in.env, in.chi);
                  ++b->n_capture;
                } catch (const Error&) {
                  in.no_graph = true;  // this instance's step does not fit a graph: eager from here on
                }
              }
            }
          c.sync();
          for (auto* in : replayed) {
            if (ctm::finish_step(*in->graph, in->env, nullptr)) {
              ++in->steps;
              ++b->n_replay;
              in->fails = 0;
              continue;
            }
            // checks failed (the cutoff rule would keep another rank, a
            // factorisation failed): the step again, eagerly
            ++b->n_undo;
            ctm::rg_step(c, in->env, in->chi);
            ++in->steps;
            ++b->n_eager;
            if (++in->fails >= 3 || dims_of(in->env) != dims_of_graph(*in->graph)) {
              in->graph.reset();
              in->no_graph = true;
            } else {
              ctm::adopt_step(*in->graph, in->env);
            }
          }
        }
        for (auto& in : b->insts)
          if (in.worker == w) {
            Tensor e = ctm::energy(c, in.env, in.H1, in.H2);
            double h[8];
            to_host(c, e, h);
            in.E[0] = h[0];
            in.E[1] = h[1];
          }
        c.sync();
      } catch (const std::exception& e) {
        errs[w] = e.what();
      }
    };
    std::vector<std::thread> th;
    for (int w = 1; w < nw; ++w) th.emplace_back(work, w);
    work(0);
    for (auto& t : th) t.join();
    auto t1 = std::chrono::steady_clock::now();
    for (int w = 0; w < nw; ++w)
      if (!errs[w].empty()) throw Error(IPEPS_ERR_INTERNAL, "batch worker " + std::to_string(w) + ": " + errs[w]);
    if (seconds_out) *seconds_out = std::chrono::duration<double>(t1 - t0).count();
    if (E_out)
      for (size_t i = 0; i < b->insts.size(); ++i) {
        E_out[2 * i] = b->insts[i].E[0];
        E_out[2 * i + 1] = b->insts[i].E[1];
      }
  });
}

int ipeps_batch_stats(ipeps_batch* b, int64_t out[4]) {
  return guard([&] {
    IPEPS_CHECK(b && out, "bad arguments");
    out[0] = b->n_capture;
    out[1] = b->n_replay;
    out[2] = b->n_undo;
    out[3] = b->n_eager;
  });
}

// which as ipeps_env_get_tensor; dims (room for 5) and nd describe the tensor
int ipeps_batch_get_tensor(ipeps_batch* b, int inst, int which, int64_t* dims, int* nd, void* host) {
  return guard([&] {
    IPEPS_CHECK(b && inst >= 0 && inst < (int)b->insts.size(), "bad instance");
    auto& in = b->insts[inst];
    Ctx& c = *b->workers[in.worker];
    HIP_CHECK(hipSetDevice(b->device));
    const Tensor& t = env_slot(in.env, which);
    if (dims && nd) {
      *nd = t.nd();
      for (int i = 0; i < t.nd() && i < 5; ++i) dims[i] = t.dims[i];
    }
    if (host) download(c, t, host_dt(b->dtype), host);
  });
}

}  // extern "C"
