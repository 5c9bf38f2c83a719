// Context, caching allocator and host/device copies.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <sstream>

#include "common.h"
#include "kernels.h"

namespace ipeps {

Pool::~Pool() {
  trim();
  for (auto& kv : parked_) (void)hipFree(kv.second);
  parked_.clear();
}

void* Pool::alloc(size_t bytes) {
  std::lock_guard<std::mutex> g(mu_);
  size_t b = round(bytes);
  if (capturing_) {
    auto ic = cap_free_.find(b);
    if (ic != cap_free_.end()) {
      void* p = ic->second;
      cap_free_.erase(ic);
      live_ += b;
      if (live_ > peak_) peak_ = live_;
      return p;
    }
  }
  auto it = free_.find(b);
  void* p = nullptr;
  if (it != free_.end()) {
    p = it->second;
    free_.erase(it);
    cached_ -= b;
  } else {
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      // release the cache and retry once
      (void)hipGetLastError();
      trim();
      HIP_CHECK(hipMalloc(&p, b));
    }
  }
  live_ += b;
  if (live_ > peak_) peak_ = live_;
  if (capturing_) {
    cap_blocks_.emplace_back(b, p);
    owned_.insert(p);
    return p;
  }
  static const bool poison = [] {
    const char* v = std::getenv("IPEPS_POOL_POISON");
    return v && v[0] == '1';
  }();
  if (poison) HIP_CHECK(hipMemsetAsync(p, 0xFF, b, stream));
  return p;
}

void Pool::release(void* p, size_t bytes) {
  std::lock_guard<std::mutex> g(mu_);
  size_t b = round(bytes);
  if (owned_.count(p)) {  // a graph's block: reusable only by the capture
    live_ -= b;
    if (capturing_) cap_free_.emplace(b, p);
    return;
  }
  if (holds_ > 0)
    parked_.emplace_back(b, p);
  else
    free_.emplace(b, p);
  live_ -= b;
  cached_ += b;
}

void Pool::hold() {
  std::lock_guard<std::mutex> g(mu_);
  ++holds_;
}

void Pool::unhold() {
  std::lock_guard<std::mutex> g(mu_);
  if (holds_ > 0 && --holds_ == 0) {
    for (auto& kv : parked_) free_.emplace(kv.first, kv.second);
    parked_.clear();
  }
}

void Pool::begin_graph() {
  std::lock_guard<std::mutex> g(mu_);
  if (capturing_) throw Error(IPEPS_ERR_STATE, "Pool: graph capture already in progress");
  capturing_ = true;
  cap_free_.clear();
  cap_blocks_.clear();
}

std::vector<std::pair<size_t, void*>> Pool::end_graph() {
  std::lock_guard<std::mutex> g(mu_);
  capturing_ = false;
  cap_free_.clear();
  std::vector<std::pair<size_t, void*>> out;
  out.swap(cap_blocks_);
  return out;
}

void Pool::release_graph(const std::vector<std::pair<size_t, void*>>& blocks) {
  std::lock_guard<std::mutex> g(mu_);
  for (auto& kv : blocks) {
    owned_.erase(kv.second);
    free_.emplace(kv.first, kv.second);
    cached_ += kv.first;
  }
}

void Pool::trim() {
  // parked blocks may still be in use on another stream: they stay
  size_t parked = 0;
  for (auto& kv : parked_) parked += kv.first;
  for (auto& kv : free_) (void)hipFree(kv.second);
  free_.clear();
  cached_ = parked;
}

Tensor Ctx::empty(DType dt, const std::vector<int64_t>& dims) {
  Tensor t;
  t.dt = dt;
  t.dims = dims;
  size_t bytes = t.bytes();
  auto b = std::make_shared<Buf>();
  b->bytes = bytes ? bytes : 1;
  b->p = pool->alloc(b->bytes);
  b->pool = pool.get();
  t.buf = b;
  return t;
}

Tensor Ctx::zeros(DType dt, const std::vector<int64_t>& dims) {
  Tensor t = empty(dt, dims);
  if (t.bytes()) dev_zero(*this, t.data(), t.bytes());
  return t;
}

void dev_copy(Ctx& c, void* dst, const void* src, size_t bytes) {
  if (!bytes) return;
  if (c.graph_safe) k::copy_bytes(c, dst, src, bytes);
  else HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c.stream));
}

void dev_copy2d(Ctx& c, void* dst, size_t dpitch, const void* src, size_t spitch, size_t width, size_t rows) {
  if (!width || !rows) return;
  if (c.graph_safe) k::copy_bytes_2d(c, dst, dpitch, src, spitch, width, rows);
  else HIP_CHECK(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, rows, hipMemcpyDeviceToDevice, c.stream));
}

void dev_zero(Ctx& c, void* p, size_t bytes) {
  if (!bytes) return;
  if (c.graph_safe) k::zero_bytes_2d(c, p, bytes, bytes, 1);
  else HIP_CHECK(hipMemsetAsync(p, 0, bytes, c.stream));
}

void dev_zero2d(Ctx& c, void* p, size_t pitch, size_t width, size_t rows) {
  if (!width || !rows) return;
  if (c.graph_safe) k::zero_bytes_2d(c, p, pitch, width, rows);
  else HIP_CHECK(hipMemset2DAsync(p, pitch, 0, width, rows, c.stream));
}

void checks_begin(Ctx& c, int64_t cap) {
  if (!c.checks.defined() || c.checks.numel() < cap) c.checks = c.empty(DType::F64, {cap});
  c.checks_pos = 0;
  c.check_recs.clear();
}

void record_check(Ctx& c, int kind, const void* src, int64_t len, int64_t a, int64_t b, double x) {
  IPEPS_CHECK(c.checks.defined() && c.checks_pos + len <= c.checks.numel(), "graph-safe checks buffer full");
  dev_copy(c, c.checks.ptr<double>() + c.checks_pos, src, (size_t)len * 8);
  c.check_recs.push_back(Ctx::CheckRec{kind, c.checks_pos, len, a, b, x});
  c.checks_pos += len;
}

static int64_t cutoff_keep(const double* s, int64_t n, int64_t chi, double rel) {
  // src/utility/wrapper.jl:37-55
  double smax = 0.0;
  for (int64_t i = 0; i < n; ++i) smax = std::max(smax, s[i]);
  int64_t above = 0;
  for (int64_t i = 0; i < n; ++i) above += (s[i] / smax > rel);
  int64_t keep = std::min<int64_t>(chi, above);
  if (n > keep && keep >= 1 && std::fabs(s[keep] - s[keep - 1]) < 1e-10) keep += 1;
  return keep;
}

bool validate_checks(const std::vector<Ctx::CheckRec>& recs, const double* h, std::string* why) {
  auto ints_zero = [](const double* v, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
      int a, b;
      std::memcpy(&a, &v[i], sizeof(int));
      std::memcpy(&b, (const char*)&v[i] + sizeof(int), sizeof(int));
      if (a != 0 || b != 0) return false;
    }
    return true;
  };
  for (size_t r = 0; r < recs.size(); ++r) {
    const auto& c = recs[r];
    const double* v = h + c.off;
    bool ok = true;
    switch (c.kind) {
      case CHECK_INFO_PAIRS: ok = ints_zero(v, c.len); break;
      case CHECK_CHOLQR2: ok = ints_zero(v, 2) && v[4] <= 0.5; break;
      case CHECK_INFO: {
        int a;
        std::memcpy(&a, v, sizeof(int));
        ok = a == 0;
        break;
      }
      case CHECK_CUTOFF: ok = cutoff_keep(v, c.len, c.a, c.x) == c.b; break;
      case CHECK_GUARD: ok = c.len >= 1 && v[0] > 0.0 && v[c.len - 1] >= c.x * v[0]; break;
      case CHECK_AT_MOST: ok = c.len >= 1 && v[0] <= c.x; break;
      default: ok = false;
    }
    if (!ok) {
      if (why) *why = "check " + std::to_string(r) + " (kind " + std::to_string(c.kind) + ") failed";
      return false;
    }
  }
  return true;
}

void Ctx::resolve_timing() {
  HIP_CHECK(hipStreamSynchronize(stream));
  for (auto& e : kev) {
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, e.a, e.b));
    auto& s = kstats[e.name];
    s.launches += 1;
    s.ms += ms;
    s.flops += e.flops;
    s.bytes += e.bytes;
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
  kev.clear();
}

Ctx& Ctx::second_lane() {
  if (!lane) lane.reset(ctx_create(device));
  return *lane;
}

Ctx* ctx_create(int device) {
  int n = 0;
  HIP_CHECK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) throw Error(IPEPS_ERR_ARG, "ctx_create: no such HIP device");
  HIP_CHECK(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIP_CHECK(hipGetDeviceProperties(&prop, device));
  std::string arch = prop.gcnArchName;
  if (arch.rfind("gfx950", 0) != 0)
    throw Error(IPEPS_ERR_STATE, "ctx_create: this library is built for gfx950 only, device is " + arch);
  // the device's address of the allocation-free entries' scratch array
  k::absorb_scratch_init(device);
  auto* c = new Ctx();
  c->device = device;
  HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  c->pool.reset(new Pool(device));
  c->pool->stream = c->stream;
  rocblas_handle h;
  if (rocblas_create_handle(&h) != rocblas_status_success) {
    delete c;
    throw Error(IPEPS_ERR_SOLVER, "rocblas_create_handle failed");
  }
  rocblas_set_stream(h, c->stream);
  // deterministic results: no atomics-based reductions inside rocBLAS/rocSOLVER
  rocblas_set_atomics_mode(h, rocblas_atomics_not_allowed);
  c->solver = h;
  return c;
}

void ctx_destroy(Ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->lane) {
    ctx_destroy(c->lane.release());
  }
  (void)hipStreamSynchronize(c->stream);
  if (c->solver) rocblas_destroy_handle((rocblas_handle)c->solver);
  // tensors the context itself holds go back to its pool before the pool goes
  c->checks = Tensor();
  c->split_cache.clear();
  c->out_offsets.clear();
  c->check_recs.clear();
  c->pool.reset();
  (void)hipStreamDestroy(c->stream);
  delete c;
}

void to_device(Ctx& c, Tensor& t, const void* host) {
  if (t.bytes()) HIP_CHECK(hipMemcpyAsync(t.data(), host, t.bytes(), hipMemcpyHostToDevice, c.stream));
  c.sync();
}

void to_host(Ctx& c, const Tensor& t, void* host) {
  if (t.bytes()) HIP_CHECK(hipMemcpyAsync(host, t.data(), t.bytes(), hipMemcpyDeviceToHost, c.stream));
  c.sync();
}

Tensor reshape(const Tensor& t, const std::vector<int64_t>& dims) {
  Tensor r = t;
  r.node.reset();
  int64_t n = 1;
  for (auto d : dims) n *= d;
  IPEPS_CHECK(n == t.numel(), "reshape: element count mismatch " + dims_str(t.dims) + " -> " + dims_str(dims));
  r.dims = dims;
  return r;
}

std::string dims_str(const std::vector<int64_t>& d) {
  std::ostringstream s;
  s << "(";
  for (size_t i = 0; i < d.size(); ++i) s << (i ? "," : "") << d[i];
  s << ")";
  return s.str();
}

}  // namespace ipeps
