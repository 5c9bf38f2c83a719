// Core runtime types of the CTMRG engine: dtypes, device buffers, tensors,
// the per-context stream / allocator, and error reporting.
//
// Tensors are dense and column-major (first index fastest), the layout of
// Julia arrays, so a Julia caller can hand its arrays over unchanged.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <functional>
#include <map>
#include <mutex>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <initializer_list>
#include <vector>

namespace ipeps {

enum class DType : int { F64 = 0, C128 = 1, F32 = 2 };

inline size_t dtype_size(DType t) {
  switch (t) {
    case DType::F64: return 8;
    case DType::C128: return 16;
    case DType::F32: return 4;
  }
  return 0;
}
inline bool is_complex(DType t) { return t == DType::C128; }

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Status codes returned through the C ABI.
enum Status : int {
  IPEPS_OK = 0,
  IPEPS_ERR_ARG = 1,
  IPEPS_ERR_HIP = 2,
  IPEPS_ERR_SOLVER = 3,
  IPEPS_ERR_STATE = 4,
  IPEPS_ERR_INTERNAL = 5,
};

#define IPEPS_CHECK(cond, msg)                                                     \
  do {                                                                             \
    if (!(cond)) throw ::ipeps::Error(::ipeps::IPEPS_ERR_ARG, std::string(msg));   \
  } while (0)

#define HIP_CHECK(expr)                                                            \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess)                                                          \
      throw ::ipeps::Error(::ipeps::IPEPS_ERR_HIP,                                 \
                           std::string(#expr) + ": " + hipGetErrorString(_e));     \
  } while (0)

// ---------------------------------------------------------------------------
// Caching device allocator.  hipMalloc synchronises the device, so freed
// blocks are kept per exact size class and reused.
// ---------------------------------------------------------------------------
class Pool {
 public:
  explicit Pool(int device) : device_(device) {}
  ~Pool();
  // With IPEPS_POOL_POISON=1 in the environment every block handed out is
  // first filled with 0xFF bytes (NaN in every floating type) on `stream`, so
  // a kernel that reads memory nothing wrote turns its output into NaN.
  hipStream_t stream = nullptr;
  void* alloc(size_t bytes);
  void release(void* p, size_t bytes);
  void trim();
  // Blocks are reused in stream order: a block released by the host becomes
  // reusable at once, which is safe only while every kernel that touches it
  // runs on this pool's stream.  While two streams share tensors (the two
  // lanes of a reverse sweep), hold() parks every released block until the
  // matching unhold(), which the caller issues only after the streams have
  // joined; holds nest.
  void hold();
  void unhold();
  // Graph capture: between begin_graph() and end_graph() every block handed
  // out joins the graph (a block released during the capture is reused only
  // by the capture, in stream order); end_graph() returns the graph's blocks,
  // which stay out of the free lists until release_graph(), since a captured
  // graph replays into the same addresses.
  void begin_graph();
  std::vector<std::pair<size_t, void*>> end_graph();
  void release_graph(const std::vector<std::pair<size_t, void*>>& blocks);
  size_t live_bytes() const { return live_; }
  size_t peak_bytes() const { return peak_; }

 private:
  static size_t round(size_t b) {
    if (b < 256) return 256;
    // 1/8-octave size classes bound the waste at 12.5%.
    size_t p = 1;
    while (p < b) p <<= 1;
    size_t step = p >> 3;
    if (step < 256) step = 256;
    return (b + step - 1) / step * step;
  }
  int device_;
  std::mutex mu_;  // blocks of a pool may be released by another thread
  std::multimap<size_t, void*> free_;
  std::vector<std::pair<size_t, void*>> parked_;  // released under hold()
  int holds_ = 0;
  bool capturing_ = false;
  std::multimap<size_t, void*> cap_free_;               // released during the capture
  std::vector<std::pair<size_t, void*>> cap_blocks_;    // handed out during the capture
  std::set<void*> owned_;                               // blocks of live graphs
  size_t live_ = 0, peak_ = 0, cached_ = 0;
};

struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
  Pool* pool = nullptr;
  ~Buf() {
    if (p && pool) pool->release(p, bytes);
  }
};

struct Node;  // reverse-mode record, see ad.h

struct Tensor {
  std::shared_ptr<Buf> buf;
  size_t offset = 0;  // in bytes
  DType dt = DType::F64;
  std::vector<int64_t> dims;
  std::shared_ptr<Node> node;  // set when the tensor is on an active tape

  int nd() const { return (int)dims.size(); }
  int64_t numel() const {
    int64_t n = 1;
    for (auto d : dims) n *= d;
    return n;
  }
  size_t bytes() const { return (size_t)numel() * dtype_size(dt); }
  void* data() const { return buf ? (char*)buf->p + offset : nullptr; }
  template <typename T>
  T* ptr() const { return reinterpret_cast<T*>(data()); }
  int64_t dim(int i) const { return dims.at(i); }
  bool defined() const { return (bool)buf; }
};

struct Ctx;

// A GPU context: device, one stream, allocator, solver handle and the
// last-error text of its thread.
struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::unique_ptr<Pool> pool;
  void* solver = nullptr;  // rocblas_handle
  // Kernel-time instrumentation: when `timing` is set, launches are bracketed
  // by hipEvents on this stream and resolved by kernel_stats().
  struct KEvent {
    std::string name;
    hipEvent_t a, b;
    double flops, bytes;
  };
  struct KStat {
    long launches = 0;
    double ms = 0, flops = 0, bytes = 0;
  };
  bool timing = false;
  bool timing_detail = false;  // key GEMM statistics by shape as well
  std::string timing_filter;   // bracket only launches whose name starts with it (empty: all)
  std::vector<KEvent> kev;
  std::map<std::string, KStat> kstats;
  void resolve_timing();

  // Graph-safe mode (set while a step is captured into a HIP graph, or
  // eagerly to run exactly what such a graph runs): no host round trips, zero
  // fills and device copies as kernels, truncation at the largest admissible
  // rank, and every check the host would make deferred: the values it needs
  // are copied into `checks` (FP64, from checks_pos on) and described by a
  // CheckRec, for validate_checks() after the step.
  struct CheckRec {
    int kind;            // CheckKind
    int64_t off, len;    // doubles in `checks`
    int64_t a = 0, b = 0;
    double x = 0.0;
  };
  bool graph_safe = false;
  Tensor checks;  // from this context's pool: ctx_destroy releases it before the pool

  // bf16 planes of FP32 GEMM operands that several products re-read (the
  // half-system matrices inside a projector), kept while a SplitCacheScope
  // registers the source tensors: products of a registered operand reuse
  // its planes instead of splitting it again.
  struct SplitEntry {
    const void* p;
    int64_t ld, R, K, kp;
    bool kv;
    int ns;
    std::shared_ptr<Buf> planes;
  };
  std::vector<std::pair<const char*, size_t>> split_src;  // registered [p, p + bytes)
  std::vector<SplitEntry> split_cache;
  // Output offset tables of products written straight into a permuted
  // layout, keyed by the legs' extents and strides (eager steps only; at most
  // OUT_OFFSETS_MAX layouts, then the cache starts over).
  static constexpr size_t OUT_OFFSETS_MAX = 256;
  std::map<std::vector<int64_t>, std::shared_ptr<Buf>> out_offsets;
  int64_t checks_pos = 0;
  std::vector<CheckRec> check_recs;

  // Second lane for the independent moves of a CTMRG step (own stream,
  // allocator and solver handle), created on first use.
  std::unique_ptr<Ctx> lane;
  Ctx& second_lane();
  Tensor empty(DType dt, const std::vector<int64_t>& dims);
  Tensor zeros(DType dt, const std::vector<int64_t>& dims);
  void sync() { HIP_CHECK(hipStreamSynchronize(stream)); }
};

Ctx* ctx_create(int device);
void ctx_destroy(Ctx* c);

// Registers tensors whose bf16 operand planes the context may keep for the
// scope (the tensors must stay alive and unchanged meanwhile); the planes go
// when the scope ends.  Nested scopes add to the registration.
struct SplitCacheScope {
  Ctx& c;
  size_t n_src, n_cache;
  SplitCacheScope(Ctx& cc, std::initializer_list<const Tensor*> ts) : c(cc), n_src(cc.split_src.size()), n_cache(cc.split_cache.size()) {
    for (const Tensor* t : ts)
      if (t && t->defined() && t->dt == DType::F32) c.split_src.emplace_back((const char*)t->data(), t->bytes());
  }
  ~SplitCacheScope() {
    c.split_src.resize(n_src);
    c.split_cache.resize(n_cache);
  }
};

// Deferred checks of graph-safe mode.
enum CheckKind : int {
  CHECK_INFO_PAIRS = 1,  // doubles each holding two ints (Cholesky infos): all zero
  CHECK_CHOLQR2 = 2,     // [info pair, info pair, |G2 - I|_max, -, -]: infos zero, |G2 - I| <= 1/2
  CHECK_INFO = 3,        // one double holding an int: zero
  CHECK_CUTOFF = 4,      // spectrum (len values): the cutoff rule with chi = a, rel = x keeps b
  CHECK_GUARD = 5,       // spectrum, descending: s_0 > 0 and s_last >= x s_0
  CHECK_AT_MOST = 6,     // one value: at most x (NaN fails)
};
// Starts a graph-safe step on c: checks buffer of `cap` doubles, no records.
void checks_begin(Ctx& c, int64_t cap);
// Appends `len` FP64 device values at src as a record of `kind`.
void record_check(Ctx& c, int kind, const void* src, int64_t len, int64_t a = 0, int64_t b = 0, double x = 0.0);
// Validates host copies of the checks buffer against the records; on failure
// `why` says which record failed.
bool validate_checks(const std::vector<Ctx::CheckRec>& recs, const double* host, std::string* why);
// Device copies and zero fills that become kernels in graph-safe mode
// (hipMemcpyAsync / hipMemsetAsync otherwise).
void dev_copy(Ctx& c, void* dst, const void* src, size_t bytes);
void dev_copy2d(Ctx& c, void* dst, size_t dpitch, const void* src, size_t spitch, size_t width, size_t rows);
void dev_zero(Ctx& c, void* p, size_t bytes);
void dev_zero2d(Ctx& c, void* p, size_t pitch, size_t width, size_t rows);

// RAII bracket of one launch for the timing instrumentation.
struct TimedLaunch {
  Ctx& c;
  bool on;
  Ctx::KEvent ev;
  TimedLaunch(Ctx& cc, const std::string& name, double flops, double bytes)
      : c(cc), on(cc.timing && !name.empty() && (cc.timing_filter.empty() || name.compare(0, cc.timing_filter.size(), cc.timing_filter) == 0)) {
    if (!on) return;
    ev.name = name;
    ev.flops = flops;
    ev.bytes = bytes;
    HIP_CHECK(hipEventCreate(&ev.a));
    HIP_CHECK(hipEventCreate(&ev.b));
    HIP_CHECK(hipEventRecord(ev.a, c.stream));
  }
  ~TimedLaunch() {
    if (!on) return;
    (void)hipEventRecord(ev.b, c.stream);
    c.kev.push_back(ev);
  }
};

// Copy helpers (host <-> device, on the context stream, synchronous).
void to_device(Ctx& c, Tensor& t, const void* host);
void to_host(Ctx& c, const Tensor& t, void* host);

// Tensor with the same buffer and new dims (column-major reshape).
Tensor reshape(const Tensor& t, const std::vector<int64_t>& dims);

std::string dims_str(const std::vector<int64_t>& d);

}  // namespace ipeps
