// Context, caching allocator and host/device copies.
#include <rocblas/rocblas.h>

#include <cstdlib>
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
  if (t.bytes()) HIP_CHECK(hipMemsetAsync(t.data(), 0, t.bytes(), stream));
  return t;
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
