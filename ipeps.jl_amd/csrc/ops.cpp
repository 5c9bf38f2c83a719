// Differentiable device ops, the contraction planner and the SVD/QR wrappers.
#include <array>

#include "ops.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <map>
#include <thread>

namespace ipeps {

Tape*& current_tape() {
  thread_local Tape* t = nullptr;
  return t;
}

int& current_seg() {
  thread_local int s = 0;
  return s;
}

namespace op {

Prec& prec() {
  thread_local Prec p;
  return p;
}

bool recording() {
  Tape* t = current_tape();
  return t && t->active;
}

static bool needs(const Tensor& t) { return recording() && (bool)t.node; }

std::shared_ptr<Node> make_node(std::function<void(Ctx&, Node&)> back) {
  auto n = std::make_shared<Node>();
  n->back = std::move(back);
  n->seg = current_seg();
  current_tape()->push(n);
  return n;
}

Tensor leaf(Tensor t) {
  if (!recording()) return t;
  auto n = std::make_shared<Node>();
  n->leaf = true;
  n->seg = current_seg();
  current_tape()->push(n);
  t.node = n;
  return t;
}

Tensor clone(Ctx& c, const Tensor& t) {
  Tensor r = c.empty(t.dt, t.dims);
  dev_copy(c, r.data(), t.data(), t.bytes());
  return r;
}

namespace {
// While a lane sweeps its segment of a pair, cotangents for nodes outside the
// segment (recorded before the pair) are kept here and applied after the join.
struct Deferral {
  int seg = 0;
  std::vector<std::pair<std::shared_ptr<Node>, Tensor>> list;
};
Deferral*& deferral() {
  thread_local Deferral* d = nullptr;
  return d;
}
}  // namespace

void acc(Ctx& c, const std::shared_ptr<Node>& n, const Tensor& g) {
  if (!n) return;
  Tensor gg = g;
  gg.node.reset();
  if (Deferral* d = deferral()) {
    if (n->seg != d->seg) {
      d->list.emplace_back(n, gg);
      return;
    }
  }
  if (!n->grad.defined()) {
    n->grad = gg;
    return;
  }
  IPEPS_CHECK(n->grad.numel() == g.numel() && n->grad.dt == g.dt,
              "acc: cotangent mismatch " + dims_str(n->grad.dims) + " vs " + dims_str(g.dims));
  if (n->grad.buf.use_count() > 1) {
    // the held cotangent's buffer is shared: one out-of-place sum (no copy)
    Tensor sum = c.empty(n->grad.dt, n->grad.dims);
    k::add(c, g.dt, g.numel(), n->grad.data(), g.data(), sum.data());
    n->grad = sum;
    return;
  }
  k::axpy(c, g.dt, g.numel(), 1.0, g.data(), n->grad.data());
}

// The cotangent n already holds when a contribution may be added to it in
// place (same segment of a pair, its buffer not shared), else null.
static Tensor* acc_target(const std::shared_ptr<Node>& n) {
  if (!n || !n->grad.defined() || n->grad.buf.use_count() > 1) return nullptr;
  if (Deferral* d = deferral())
    if (n->seg != d->seg) return nullptr;
  return &n->grad;
}

// acc(c, n, tcon_raw(...)) with the product added to n's cotangent by its GEMM
// when acc would add it there anyway (no temporary, no separate sum pass).
static void acc_tcon(Ctx& c, const std::shared_ptr<Node>& n, const Tensor& A, const std::vector<int>& la, bool conjA,
                     const Tensor& B, const std::vector<int>& lb, bool conjB) {
  Tensor* into = acc_target(n);
  const void* held = into ? into->data() : nullptr;
  Tensor r = tcon_raw(c, A, la, conjA, B, lb, conjB, into);
  if (held && r.data() == held) return;
  acc(c, n, r);
}

static void back_node(Ctx& c, Node& n) {
  if (!n.back) return;
  bool any = n.grad.defined();
  for (auto& s : n.slots) any = any || s.defined();
  if (any) n.back(c, n);
  n.grad = Tensor();
  n.slots.clear();
  n.back = nullptr;
}

// Reverse sweep of a pair: the second lane's segment on the context's second
// lane (own stream, allocator, host thread), the main segment here; the
// cotangents each sends to earlier nodes are applied after the join, the
// second lane's first, each in its own order -- the order of the sequential
// sweep, which takes the second lane's nodes (recorded last) first.
// Forward tensors captured by the VJP closures are read by both streams, and
// either host thread may drop the last reference while the other stream's
// kernels that read the block are still queued; so both pools park every
// block released during the sweep and take them back only once the main
// stream waits for the lane's join event.
static void backward_pair(Ctx& c, Tape& t, const Tape::Pair& p) {
  Ctx& l = c.second_lane();
  struct PoolHold {
    Pool *a, *b;
    PoolHold(Pool* x, Pool* y) : a(x), b(y) {
      a->hold();
      b->hold();
    }
    void release() {
      if (a) a->unhold();
      if (b) b->unhold();
      a = b = nullptr;
    }
    ~PoolHold() { release(); }
  } hold(c.pool.get(), l.pool.get());
  hipEvent_t fork, join;
  HIP_CHECK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
  HIP_CHECK(hipEventCreateWithFlags(&join, hipEventDisableTiming));
  HIP_CHECK(hipEventRecord(fork, c.stream));
  HIP_CHECK(hipStreamWaitEvent(l.stream, fork, 0));
  l.timing = c.timing;
  l.timing_detail = c.timing_detail;
  l.timing_filter = c.timing_filter;
  Deferral dl, dm;
  dl.seg = p.slane;
  dm.seg = p.smain;
  const Prec prec0 = prec();
  std::exception_ptr lane_err, main_err;
  std::thread th([&] {
    try {
      HIP_CHECK(hipSetDevice(l.device));
      prec() = prec0;
      deferral() = &dl;
      for (size_t i = p.l1; i-- > p.m1;) back_node(l, *t.order[i]);
      deferral() = nullptr;
      HIP_CHECK(hipEventRecord(join, l.stream));
    } catch (...) {
      deferral() = nullptr;
      lane_err = std::current_exception();
      (void)hipEventRecord(join, l.stream);
    }
  });
  try {
    deferral() = &dm;
    for (size_t i = p.m1; i-- > p.m0;) back_node(c, *t.order[i]);
    deferral() = nullptr;
  } catch (...) {
    deferral() = nullptr;
    main_err = std::current_exception();
  }
  th.join();
  const hipError_t werr = hipStreamWaitEvent(c.stream, join, 0);
  (void)hipEventDestroy(fork);
  (void)hipEventDestroy(join);
  c.kev.insert(c.kev.end(), l.kev.begin(), l.kev.end());
  l.kev.clear();
  if (werr != hipSuccess) {
    // without the wait the parked blocks may still be in use: keep them parked
    // until both streams drain
    (void)hipDeviceSynchronize();
  }
  hold.release();
  if (main_err) std::rethrow_exception(main_err);
  if (lane_err) std::rethrow_exception(lane_err);
  for (auto& e : dl.list) acc(c, e.first, e.second);
  for (auto& e : dm.list) acc(c, e.first, e.second);
}

void backward(Ctx& c, Tape& t) {
  TapeScope ts(nullptr);
  std::map<size_t, const Tape::Pair*> by_end;  // l1 -> pair
  for (auto& p : t.pairs) by_end[p.l1] = &p;
  for (size_t i = t.order.size(); i-- > 0;) {
    auto it = by_end.find(i + 1);
    if (it != by_end.end() && it->second->m1 < it->second->l1) {
      const Tape::Pair& p = *it->second;
      backward_pair(c, t, p);
      i = p.m0;  // continue below the pair
      continue;
    }
    back_node(c, *t.order[i]);
  }
}

// ---------------------------------------------------------------------------
// permute / conj / reshape
// ---------------------------------------------------------------------------
Tensor permute_raw(Ctx& c, const Tensor& t, const std::vector<int>& perm, bool cj) {
  std::vector<int64_t> od(perm.size());
  for (size_t i = 0; i < perm.size(); ++i) od[i] = t.dims.at(perm[i]);
  Tensor r = c.empty(t.dt, od);
  k::permute(c, t.dt, t.dims, perm, t.data(), r.data(), cj);
  return r;
}

static std::vector<int> inverse_perm(const std::vector<int>& p) {
  std::vector<int> q(p.size());
  for (size_t i = 0; i < p.size(); ++i) q[p[i]] = (int)i;
  return q;
}

Tensor permute(Ctx& c, const Tensor& t, const std::vector<int>& perm) {
  IPEPS_CHECK((int)perm.size() == t.nd(), "permute: rank mismatch");
  Tensor r = permute_raw(c, t, perm, false);
  if (needs(t)) {
    auto in = t.node;
    auto ip = inverse_perm(perm);
    r.node = make_node([in, ip](Ctx& c, Node& n) { acc(c, in, permute_raw(c, n.grad, ip, false)); });
  }
  return r;
}

Tensor conj(Ctx& c, const Tensor& t) {
  if (t.dt != DType::C128) {
    Tensor r = t;  // real: identity
    if (needs(t)) {
      auto in = t.node;
      r.node = make_node([in](Ctx& c, Node& n) { acc(c, in, n.grad); });
    }
    return r;
  }
  Tensor r = c.empty(t.dt, t.dims);
  k::conj_copy(c, t.dt, t.numel(), t.data(), r.data());
  if (needs(t)) {
    auto in = t.node;
    r.node = make_node([in](Ctx& c, Node& n) {
      Tensor g = c.empty(n.grad.dt, n.grad.dims);
      k::conj_copy(c, g.dt, g.numel(), n.grad.data(), g.data());
      acc(c, in, g);
    });
  }
  return r;
}

Tensor reshape(Ctx& c, const Tensor& t, const std::vector<int64_t>& dims) {
  Tensor r = ipeps::reshape(t, dims);
  if (needs(t)) {
    auto in = t.node;
    auto od = t.dims;
    r.node = make_node([in, od](Ctx& c, Node& n) { acc(c, in, ipeps::reshape(n.grad, od)); });
  }
  return r;
}

Tensor to_dtype(Ctx& c, const Tensor& t, DType dt) {
  if (t.dt == dt) return t;
  Tensor r = c.empty(dt, t.dims);
  k::convert(c, t.dt, dt, t.numel(), t.data(), r.data());
  return r;
}

Tensor cast(Ctx& c, const Tensor& t, DType dt) {
  if (t.dt == dt) return t;
  Tensor r = to_dtype(c, t, dt);
  r.node.reset();
  if (needs(t)) {
    auto in = t.node;
    DType from = t.dt;
    r.node = make_node([in, from](Ctx& c, Node& n) { acc(c, in, to_dtype(c, n.grad, from)); });
  }
  return r;
}

Tensor add(Ctx& c, const Tensor& a, const Tensor& b) {
  IPEPS_CHECK(a.dt == b.dt && a.numel() == b.numel(), "add: operand mismatch");
  Tensor r = clone(c, a);
  r.dims = a.dims;
  k::axpy(c, a.dt, a.numel(), 1.0, b.data(), r.data());
  if (needs(a) || needs(b)) {
    auto na = a.node, nb = b.node;
    auto da = a.dims, db = b.dims;
    r.node = make_node([na, nb, da, db](Ctx& c, Node& n) {
      if (na) acc(c, na, ipeps::reshape(n.grad, da));
      if (nb) acc(c, nb, ipeps::reshape(n.grad, db));
    });
  }
  return r;
}

Tensor scale_cplx(Ctx& c, const Tensor& t, double zr, double zi) {
  Tensor r = c.empty(t.dt, t.dims);
  k::scale_cplx(c, t.dt, t.numel(), zr, zi, t.data(), r.data());
  if (needs(t)) {
    auto in = t.node;
    r.node = make_node([in, zr, zi](Ctx& c, Node& n) {
      Tensor g = c.empty(n.grad.dt, n.grad.dims);
      k::scale_cplx(c, g.dt, g.numel(), zr, -zi, n.grad.data(), g.data());  // conj(z) * ybar
      acc(c, in, g);
    });
  }
  return r;
}

// ---------------------------------------------------------------------------
// contraction planner
// ---------------------------------------------------------------------------
namespace {

bool layout_is_identity(const std::vector<int>& layout, const std::vector<int64_t>& dims) {
  int last = -1;
  for (int ax : layout) {
    if (dims[ax] == 1) continue;
    if (ax < last) return false;
    last = ax;
  }
  return true;
}

int64_t prod_of(const std::vector<int>& axes, const std::vector<int64_t>& dims) {
  int64_t p = 1;
  for (int a : axes) p *= dims[a];
  return p;
}

std::vector<int> cat(const std::vector<int>& a, const std::vector<int>& b) {
  std::vector<int> r(a);
  r.insert(r.end(), b.begin(), b.end());
  return r;
}

}  // namespace

Tensor matmul_raw(Ctx& c, const Tensor& A, bool tA, bool cA, const Tensor& B, bool tB, bool cB) {
  IPEPS_CHECK(A.nd() == 2 && B.nd() == 2, "matmul: matrices expected");
  IPEPS_CHECK(A.dt == B.dt, "matmul: dtype mismatch");
  int64_t M = tA ? A.dims[1] : A.dims[0];
  int64_t K = tA ? A.dims[0] : A.dims[1];
  int64_t K2 = tB ? B.dims[1] : B.dims[0];
  int64_t N = tB ? B.dims[0] : B.dims[1];
  IPEPS_CHECK(K == K2, "matmul: inner dimension mismatch " + dims_str(A.dims) + " x " + dims_str(B.dims));
  Tensor C = c.empty(A.dt, {M, N});
  k::GemmArgs g;
  g.dt = A.dt;
  g.transA = tA; g.conjA = cA; g.transB = tB; g.conjB = cB;
  g.M = M; g.N = N; g.K = K;
  g.A = A.data(); g.lda = std::max<int64_t>(1, A.dims[0]);
  g.B = B.data(); g.ldb = std::max<int64_t>(1, B.dims[0]);
  g.C = C.data(); g.ldc = std::max<int64_t>(1, M);
  g.mixed = prec().mixed;
  k::gemm(c, g);
  return C;
}

Tensor tcon_raw(Ctx& c, const Tensor& A, const std::vector<int>& la, bool conjA, const Tensor& B,
                const std::vector<int>& lb, bool conjB, Tensor* into) {
  const int na = A.nd(), nb = B.nd();
  IPEPS_CHECK((int)la.size() == na && (int)lb.size() == nb, "tcon: label count does not match tensor rank");
  IPEPS_CHECK(A.dt == B.dt, "tcon: dtype mismatch");
  std::vector<int> ca, cb;
  for (int p = 0; p < na; ++p) {
    if (la[p] > 0) {
      auto it = std::find(lb.begin(), lb.end(), la[p]);
      IPEPS_CHECK(it != lb.end(), "tcon: positive label appears in one tensor only");
      int q = (int)(it - lb.begin());
      IPEPS_CHECK(A.dims[p] == B.dims[q], "tcon: contracted dimensions differ");
      ca.push_back(p);
      cb.push_back(q);
    }
  }
  for (int q = 0; q < nb; ++q)
    if (lb[q] > 0) IPEPS_CHECK(std::find(la.begin(), la.end(), lb[q]) != la.end(), "tcon: unmatched label");
  int nout = 0;
  for (int l : la) nout += l < 0;
  for (int l : lb) nout += l < 0;
  std::vector<int64_t> odims(nout, -1);
  std::vector<int> osrc(nout, -1), opos(nout, -1);
  for (int p = 0; p < na; ++p)
    if (la[p] < 0) {
      int kk = -la[p] - 1;
      IPEPS_CHECK(kk < nout && osrc[kk] < 0, "tcon: bad output label");
      odims[kk] = A.dims[p]; osrc[kk] = 0; opos[kk] = p;
    }
  for (int q = 0; q < nb; ++q)
    if (lb[q] < 0) {
      int kk = -lb[q] - 1;
      IPEPS_CHECK(kk < nout && osrc[kk] < 0, "tcon: bad output label");
      odims[kk] = B.dims[q]; osrc[kk] = 1; opos[kk] = q;
    }
  std::vector<int> fa_nat, fb_nat, fa_out, fb_out;
  for (int p = 0; p < na; ++p) if (la[p] < 0) fa_nat.push_back(p);
  for (int q = 0; q < nb; ++q) if (lb[q] < 0) fb_nat.push_back(q);
  for (int kk = 0; kk < nout; ++kk) (osrc[kk] == 0 ? fa_out : fb_out).push_back(opos[kk]);

  // contracted orders: A's natural order, or B's natural order
  std::vector<int> ca_b(ca), cb_b(cb);
  {
    std::vector<int> idx(cb.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int i, int j) { return cb[i] < cb[j]; });
    for (size_t i = 0; i < idx.size(); ++i) { ca_b[i] = ca[idx[i]]; cb_b[i] = cb[idx[i]]; }
  }
  const int64_t M = prod_of(fa_nat, A.dims), N = prod_of(fb_nat, B.dims), K = prod_of(ca, A.dims);
  const int64_t nA = A.numel(), nB = B.numel(), nC = M * N;

  struct Choice {
    int64_t cost;
    std::vector<int> CA, CB, FA, FB;
    bool ct;
  } best{INT64_MAX, {}, {}, {}, {}, false};
  for (int co = 0; co < 2; ++co) {
    const auto& CA = co ? ca_b : ca;
    const auto& CB = co ? cb_b : cb;
    for (int fo = 0; fo < 4; ++fo) {
      const auto& FA = (fo & 1) ? fa_out : fa_nat;
      const auto& FB = (fo & 2) ? fb_out : fb_nat;
      int64_t costA = (layout_is_identity(cat(FA, CA), A.dims) || layout_is_identity(cat(CA, FA), A.dims)) ? 0 : nA;
      int64_t costB = (layout_is_identity(cat(CB, FB), B.dims) || layout_is_identity(cat(FB, CB), B.dims)) ? 0 : nB;
      for (int ct = 0; ct < 2; ++ct) {
        // result legs in output numbering
        std::vector<int> res;
        auto push = [&](const std::vector<int>& F, int src) {
          for (int pos : F)
            for (int kk = 0; kk < nout; ++kk)
              if (osrc[kk] == src && opos[kk] == pos) res.push_back(kk);
        };
        if (!ct) { push(FA, 0); push(FB, 1); }
        else { push(FB, 1); push(FA, 0); }
        // a result the GEMM can write straight into the output layout (real
        // data, its fastest row leg the output's fastest leg, >= 16 long:
        // see below) costs its offset tables, not a permute pass
        const bool real = A.dt == DType::F32 || A.dt == DType::F64;
        const int64_t rows = ct ? N : M;
        const bool directable = real && !res.empty() && res[0] == 0 && odims[0] >= 16 && rows > 1;
        int64_t costC = layout_is_identity(res, odims) ? 0 : directable ? nC / 8 : nC;
        int64_t cost = costA + costB + costC;
        if (cost < best.cost) best = Choice{cost, CA, CB, FA, FB, (bool)ct};
      }
    }
  }

  // materialise operands
  Tensor Am = A, Bm = B;
  bool a_kxm;  // A stored as K x M (else M x K)
  bool cjA = conjA, cjB = conjB;
  if (layout_is_identity(cat(best.FA, best.CA), A.dims)) a_kxm = false;
  else if (layout_is_identity(cat(best.CA, best.FA), A.dims)) a_kxm = true;
  else {
    Am = permute_raw(c, A, cat(best.FA, best.CA), conjA && A.dt == DType::C128);
    cjA = false;
    a_kxm = false;
  }
  bool b_kxn;  // B stored as K x N (else N x K)
  if (layout_is_identity(cat(best.CB, best.FB), B.dims)) b_kxn = true;
  else if (layout_is_identity(cat(best.FB, best.CB), B.dims)) b_kxn = false;
  else {
    Bm = permute_raw(c, B, cat(best.CB, best.FB), conjB && B.dt == DType::C128);
    cjB = false;
    b_kxn = true;
  }
  std::vector<int64_t> rdims;
  std::vector<int> res;
  auto push = [&](const std::vector<int>& F, int src) {
    for (int pos : F)
      for (int kk = 0; kk < nout; ++kk)
        if (osrc[kk] == src && opos[kk] == pos) { res.push_back(kk); rdims.push_back(odims[kk]); }
  };
  if (!best.ct) { push(best.FA, 0); push(best.FB, 1); }
  else { push(best.FB, 1); push(best.FA, 0); }
  // A result whose legs come out of the GEMM in another order is written by
  // the GEMM straight into the output layout through row / column offset
  // tables (real data, the output's fastest leg the GEMM's fastest row leg
  // and at least 16 long, so a wave's stores stay in 64-B runs), instead of a
  // separate permute pass.
  const bool ident = layout_is_identity(res, odims);
  int64_t nrow_legs = 0;  // R's row legs: res[0 .. nrow_legs)
  {
    int64_t p = 1;
    while (nrow_legs < (int64_t)res.size() && p < (best.ct ? N : M)) p *= odims[res[nrow_legs++]];
  }
  const bool direct = !ident && nC > 0 && K > 0 && (A.dt == DType::F32 || A.dt == DType::F64) && nrow_legs >= 1 &&
                      res[0] == 0 && odims[0] >= 16 && nC < ((int64_t)1 << 31);
  // accumulate into *into when the GEMM writes the result's own layout
  const bool acc_in = into && into->defined() && into->dt == A.dt && into->numel() == nC && nC > 0 && K > 0 &&
                      (ident || direct);
  Tensor R = acc_in ? *into : c.empty(A.dt, (ident || !direct) ? rdims : odims);
  std::shared_ptr<Buf> offs;
  if (nC > 0) {
    k::GemmArgs g;
    g.dt = A.dt;
    g.K = K;
    g.mixed = prec().mixed;
    if (!best.ct) {
      g.M = M; g.N = N;
      g.A = Am.data(); g.transA = a_kxm; g.lda = std::max<int64_t>(1, a_kxm ? K : M); g.conjA = cjA;
      g.B = Bm.data(); g.transB = !b_kxn; g.ldb = std::max<int64_t>(1, b_kxn ? K : N); g.conjB = cjB;
      g.C = R.data(); g.ldc = std::max<int64_t>(1, M);
    } else {
      g.M = N; g.N = M;
      g.A = Bm.data(); g.transA = b_kxn; g.lda = std::max<int64_t>(1, b_kxn ? K : N); g.conjA = cjB;
      g.B = Am.data(); g.transB = !a_kxm; g.ldb = std::max<int64_t>(1, a_kxm ? K : M); g.conjB = cjA;
      g.C = R.data(); g.ldc = std::max<int64_t>(1, N);
    }
    if (acc_in) g.beta = 1.0;
    if (direct) {
      // output strides (column-major odims) of R's legs
      std::vector<int64_t> ost(nout);
      int64_t st = 1;
      for (int kk = 0; kk < nout; ++kk) {
        ost[kk] = st;
        st *= odims[kk];
      }
      std::vector<int64_t> rext, rstr, cext, cstr;
      for (int64_t i = 0; i < (int64_t)res.size(); ++i) {
        (i < nrow_legs ? rext : cext).push_back(odims[res[i]]);
        (i < nrow_legs ? rstr : cstr).push_back(ost[res[i]]);
      }
      const int64_t gm = g.M, gn = g.N;
      IPEPS_CHECK((int64_t)std::accumulate(rext.begin(), rext.end(), (int64_t)1, std::multiplies<int64_t>()) == gm,
                  "tcon: output row legs do not match the GEMM rows");
      // the tables depend only on the legs' extents and output strides: kept
      // per context (a step reuses each layout many times)
      std::vector<int64_t> key;
      key.push_back((int64_t)rext.size());
      key.insert(key.end(), rext.begin(), rext.end());
      key.insert(key.end(), rstr.begin(), rstr.end());
      key.insert(key.end(), cext.begin(), cext.end());
      key.insert(key.end(), cstr.begin(), cstr.end());
      // (graph-safe steps build their own: a block handed out during a
      // capture belongs to the graph, and its kernels run only at replay)
      auto it = c.graph_safe ? c.out_offsets.end() : c.out_offsets.find(key);
      if (it != c.out_offsets.end()) {
        offs = it->second;
      } else {
        const size_t ob = (size_t)(gm + gn) * sizeof(int);
        offs.reset(new Buf{c.pool->alloc(ob), ob, c.pool.get()});
        int* om = (int*)offs->p;
        k::mixed_radix_offsets(c, (int)rext.size(), rext.data(), rstr.data(), gm, om);
        k::mixed_radix_offsets(c, (int)cext.size(), cext.data(), cstr.data(), gn, om + gm);
        if (!c.graph_safe) {
          if (c.out_offsets.size() >= Ctx::OUT_OFFSETS_MAX) c.out_offsets.clear();
          c.out_offsets.emplace(key, offs);
        }
      }
      g.offM = (const int*)offs->p;
      g.offN = (const int*)offs->p + gm;
    }
    if (K == 0) k::fill_zero(c, R.data(), R.bytes());
    else k::gemm(c, g);
  }
  if (acc_in) return *into;
  if (ident || direct) return ipeps::reshape(R, odims);
  // R has legs res[0..]; output wants leg kk at position kk
  std::vector<int> perm(nout);
  for (int kk = 0; kk < nout; ++kk) {
    auto it = std::find(res.begin(), res.end(), kk);
    perm[kk] = (int)(it - res.begin());
  }
  return permute_raw(c, R, perm, false);
}

Tensor tcon(Ctx& c, const Tensor& A, const std::vector<int>& la, const Tensor& B, const std::vector<int>& lb) {
  Tensor C = tcon_raw(c, A, la, false, B, lb, false);
  C.node.reset();
  if (needs(A) || needs(B)) {
    auto nA = A.node, nB = B.node;
    Tensor Ak = A, Bk = B;  // values kept for the VJP
    Ak.node.reset();
    Bk.node.reset();
    const int na = A.nd(), nb = B.nd();
    int nout = C.nd();
    // labels for the VJPs
    std::vector<int> lc_a(nout), lb_a(nb), la_b(na), lc_b(nout);
    const int BIG = 100000;
    for (int kk = 0; kk < nout; ++kk) {
      int lab = -(kk + 1);
      auto ia = std::find(la.begin(), la.end(), lab);
      if (ia != la.end()) {
        int p = (int)(ia - la.begin());
        lc_a[kk] = -(p + 1);
        lc_b[kk] = BIG + kk;
        la_b[p] = BIG + kk;
      } else {
        int q = (int)(std::find(lb.begin(), lb.end(), lab) - lb.begin());
        lc_a[kk] = BIG + kk;
        lb_a[q] = BIG + kk;
        lc_b[kk] = -(q + 1);
      }
    }
    for (int p = 0; p < na; ++p)
      if (la[p] > 0) {
        int q = (int)(std::find(lb.begin(), lb.end(), la[p]) - lb.begin());
        lb_a[q] = -(p + 1);
        la_b[p] = -(q + 1);
      }
    bool gA = needs(A), gB = needs(B);
    const k::Mixed mx = prec().back;
    C.node = make_node([=](Ctx& c, Node& n) {
      PrecScope ps(mx);
      if (gA) acc_tcon(c, nA, n.grad, lc_a, false, Bk, lb_a, true);
      if (gB) acc_tcon(c, nB, Ak, la_b, true, n.grad, lc_b, false);
    });
  }
  return C;
}

Tensor matmul(Ctx& c, const Tensor& A, const Tensor& B) { return tcon(c, A, {-1, 1}, B, {1, -2}); }

// ---------------------------------------------------------------------------
// fused site absorption
// ---------------------------------------------------------------------------
namespace {

struct AbsorbPlan {
  bool ok = false;
  int D = 0, P = 0;
  std::vector<int64_t> odims;       // output dims (labels -1, -2, ...)
  std::vector<int> permA, permAd;   // A -> [u1, u2, a1, a2, p], Ad -> [v1, v2, b1, b2, p]
  k::AbsorbLegs xl, yl;             // X: rows, (u, v); Y: rows, (a, b)
};

std::vector<int64_t> col_strides(const std::vector<int64_t>& d) {
  std::vector<int64_t> s(d.size());
  int64_t a = 1;
  for (size_t i = 0; i < d.size(); ++i) { s[i] = a; a *= d[i]; }
  return s;
}

// X (6 legs: 2 rows, 2 shared with A, 2 shared with Ad), A / Ad (D,D,D,D,P
// with 2 legs shared with X, 2 free, 1 shared with each other).
AbsorbPlan plan_absorb(const Tensor& X, const std::vector<int>& lx, const Tensor& A, const std::vector<int>& la,
                       const Tensor& Ad, const std::vector<int>& lad) {
  AbsorbPlan pl;
  if (X.nd() != 6 || A.nd() != 5 || Ad.nd() != 5 || lx.size() != 6 || la.size() != 5 || lad.size() != 5) return pl;
  if (X.dt != DType::F32 || A.dt != DType::F32 || Ad.dt != DType::F32) return pl;
  auto find = [](const std::vector<int>& l, int lab) {
    for (size_t i = 0; i < l.size(); ++i)
      if (l[i] == lab) return (int)i;
    return -1;
  };
  std::vector<int> rowsX, uX, vX, uA, vAd;  // X positions of row / u / v legs; partner positions
  for (int i = 0; i < 6; ++i) {
    if (lx[i] < 0) { rowsX.push_back(i); continue; }
    const int ja = find(la, lx[i]), jd = find(lad, lx[i]);
    if (ja >= 0 && jd < 0) { uX.push_back(i); uA.push_back(ja); }
    else if (jd >= 0 && ja < 0) { vX.push_back(i); vAd.push_back(jd); }
    else return pl;
  }
  if (rowsX.size() != 2 || uX.size() != 2 || vX.size() != 2) return pl;
  int pA = -1, pAd = -1;
  std::vector<int> aA, bAd;
  for (int j = 0; j < 5; ++j) {
    if (la[j] < 0) aA.push_back(j);
    else if (find(lx, la[j]) < 0) { if (pA >= 0) return pl; pA = j; pAd = find(lad, la[j]); if (pAd < 0) return pl; }
  }
  for (int j = 0; j < 5; ++j)
    if (lad[j] < 0) bAd.push_back(j);
  if (aA.size() != 2 || bAd.size() != 2 || pA < 0) return pl;
  const int64_t D = A.dims[0], P = A.dims[pA];
  for (int j = 0; j < 5; ++j) {
    if (j != pA && A.dims[j] != D) return pl;
    if (j != pAd && Ad.dims[j] != D) return pl;
  }
  if (Ad.dims[pAd] != P || !k::absorb_supported((int)D, (int)P)) return pl;
  for (int i = 0; i < 2; ++i)
    if (X.dims[uX[i]] != D || X.dims[vX[i]] != D) return pl;
  // u = i(first shared A leg) + D i(second); same for v, a, b
  auto order2 = [](std::vector<int>& partner, std::vector<int>* other) {
    if (partner[0] > partner[1]) {
      std::swap(partner[0], partner[1]);
      if (other) std::swap((*other)[0], (*other)[1]);
    }
  };
  order2(uA, &uX);
  order2(vAd, &vX);
  const int nout = 6;
  pl.odims.assign(nout, 0);
  std::vector<int> opos_rows(2), opos_a(2), opos_b(2);
  for (int r = 0; r < 2; ++r) {
    const int o = -lx[rowsX[r]] - 1;
    if (o < 0 || o >= nout) return pl;
    pl.odims[o] = X.dims[rowsX[r]];
    opos_rows[r] = o;
  }
  for (int r = 0; r < 2; ++r) {
    const int oa = -la[aA[r]] - 1, ob = -lad[bAd[r]] - 1;
    if (oa < 0 || oa >= nout || ob < 0 || ob >= nout) return pl;
    pl.odims[oa] = D;
    pl.odims[ob] = D;
    opos_a[r] = oa;
    opos_b[r] = ob;
  }
  for (auto d : pl.odims)
    if (d <= 0) return pl;
  const auto xs = col_strides(X.dims), ys = col_strides(pl.odims);
  int64_t ny = 1;
  for (auto d : pl.odims) ny *= d;
  if (X.numel() >= ((int64_t)1 << 31) || ny >= ((int64_t)1 << 31)) return pl;
  // rows: the leg with the smaller X stride is r1 (consecutive rows of a block)
  int r1 = 0, r2 = 1;
  if (xs[rowsX[1]] < xs[rowsX[0]]) std::swap(r1, r2);
  pl.xl.n1 = (int)X.dims[rowsX[r1]];
  pl.xl.n2 = (int)X.dims[rowsX[r2]];
  pl.xl.s1 = xs[rowsX[r1]];
  pl.xl.s2 = xs[rowsX[r2]];
  pl.xl.sx1 = xs[uX[0]];
  pl.xl.sx2 = xs[uX[1]];
  pl.xl.sy1 = xs[vX[0]];
  pl.xl.sy2 = xs[vX[1]];
  pl.yl.n1 = pl.xl.n1;
  pl.yl.n2 = pl.xl.n2;
  pl.yl.s1 = ys[opos_rows[r1]];
  pl.yl.s2 = ys[opos_rows[r2]];
  pl.yl.sx1 = ys[opos_a[0]];
  pl.yl.sx2 = ys[opos_a[1]];
  pl.yl.sy1 = ys[opos_b[0]];
  pl.yl.sy2 = ys[opos_b[1]];
  pl.permA = {uA[0], uA[1], aA[0], aA[1], pA};
  pl.permAd = {vAd[0], vAd[1], bAd[0], bAd[1], pAd};
  pl.D = (int)D;
  pl.P = (int)P;
  pl.ok = true;
  return pl;
}

// Element strides of a site tensor (dims t) for the legs (u1, u2, a1, a2, p)
// at positions perm = {u1, u2, a1, a2, p}; transposed: the u and a legs swap.
std::array<int, 5> site_strides(const Tensor& t, const std::vector<int>& perm, bool transposed) {
  const auto st = col_strides(t.dims);
  std::array<int, 5> s;
  for (int i = 0; i < 5; ++i) s[i] = (int)st[perm[i]];
  if (transposed) {
    std::swap(s[0], s[2]);
    std::swap(s[1], s[3]);
  }
  return s;
}

// raw fused launch: Y (dims ydims) from X through legs xl -> yl
Tensor absorb_raw(Ctx& c, const Tensor& X, const k::AbsorbLegs& xl, const Tensor& A, const std::array<int, 5>& sa,
                  const Tensor& Ad, const std::array<int, 5>& sd, const std::vector<int64_t>& ydims,
                  const k::AbsorbLegs& yl, int D, int P) {
  Tensor Y = c.empty(DType::F32, ydims);
  k::AbsorbP p;
  p.U = D * D;
  p.P = P;
  p.D = D;
  p.x = xl;
  p.y = yl;
  p.X = X.ptr<float>();
  p.Y = Y.ptr<float>();
  p.Am = A.ptr<float>();
  p.Adm = Ad.ptr<float>();
  for (int i = 0; i < 5; ++i) {
    p.sa[i] = sa[i];
    p.sd[i] = sd[i];
  }
  k::absorb_fwd(c, p);
  return Y;
}

}  // namespace

Tensor absorb(Ctx& c, const Tensor& X, const std::vector<int>& lx, const Tensor& A, const std::vector<int>& la,
              const Tensor& Ad, const std::vector<int>& lad) {
  AbsorbPlan pl;
  if (prec().fused_absorb) pl = plan_absorb(X, lx, A, la, Ad, lad);
  if (!pl.ok) {
    // two pairwise contractions: X with A over the u legs, then with Ad
    std::vector<int> l1x(lx.size()), l1a(la.size()), l2xa, l2d = lad;
    int next = 1;
    for (size_t i = 0; i < lx.size(); ++i) {
      bool shared_a = std::find(la.begin(), la.end(), lx[i]) != la.end() && lx[i] > 0;
      if (shared_a) l1x[i] = lx[i];
      else { l1x[i] = -(next++); l2xa.push_back(lx[i]); }
    }
    for (size_t j = 0; j < la.size(); ++j) {
      bool shared_x = la[j] > 0 && std::find(lx.begin(), lx.end(), la[j]) != lx.end();
      if (shared_x) l1a[j] = la[j];
      else { l1a[j] = -(next++); l2xa.push_back(la[j]); }
    }
    Tensor XA = tcon(c, X, l1x, A, l1a);
    return tcon(c, XA, l2xa, Ad, l2d);
  }
  const int U = pl.D * pl.D;
  // A and Ad are read in place (the kernels gather them through strides)
  const auto sa = site_strides(A, pl.permA, false), sd = site_strides(Ad, pl.permAd, false);
  Tensor Y = absorb_raw(c, X, pl.xl, A, sa, Ad, sd, pl.odims, pl.yl, pl.D, pl.P);
  if (needs(X) || needs(A) || needs(Ad)) {
    auto nX = X.node, nA = A.node, nD = Ad.node;
    Tensor Xk = X, Ak = A, Dk = Ad;
    Xk.node.reset();
    Ak.node.reset();
    Dk.node.reset();
    const AbsorbPlan plk = pl;
    Y.node = make_node([=](Ctx& c, Node& n) {
      const Tensor& G = n.grad;
      IPEPS_CHECK(G.dt == DType::F32, "absorb: FP32 cotangent expected");
      if (nX) {
        // Xbar_r = sum_p A_p G_r Ad_p^T: the forward kernel with A, Ad transposed
        acc(c, nX, absorb_raw(c, G, plk.yl, Ak, site_strides(Ak, plk.permA, true), Dk,
                              site_strides(Dk, plk.permAd, true), Xk.dims, plk.xl, plk.D, plk.P));
      }
      if (nA || nD) {
        Tensor out = c.empty(DType::F32, {2 * U * U * plk.P});
        k::AbsorbWP w;
        w.U = U;
        w.P = plk.P;
        w.D = plk.D;
        w.x = plk.xl;
        w.g = plk.yl;
        w.X = Xk.ptr<float>();
        w.G = G.ptr<float>();
        w.Am = Ak.ptr<float>();
        w.Adm = Dk.ptr<float>();
        const auto ska = site_strides(Ak, plk.permA, false), skd = site_strides(Dk, plk.permAd, false);
        for (int i = 0; i < 5; ++i) {
          w.sa[i] = ska[i];
          w.sd[i] = skd[i];
        }
        k::absorb_wgrad(c, w, out.ptr<float>());
        // the cotangents come out as (u1, u2, a1, a2, p) [resp. (v1, v2, b1, b2, p)];
        // back to the site tensor's own leg order
        auto back = [&](const Tensor& g, const std::vector<int>& perm, const std::vector<int64_t>& dims) {
          std::vector<int> inv(5);
          for (int i = 0; i < 5; ++i) inv[perm[i]] = i;
          Tensor t = g;
          t.dims = {dims[perm[0]], dims[perm[1]], dims[perm[2]], dims[perm[3]], dims[perm[4]]};
          return permute_raw(c, t, inv);
        };
        if (nA) {
          Tensor gA = out;  // first half of out
          acc(c, nA, back(gA, plk.permA, Ak.dims));
        }
        if (nD) {
          Tensor gD = out;
          gD.offset += (size_t)U * U * plk.P * sizeof(float);
          acc(c, nD, back(gD, plk.permAd, Dk.dims));
        }
      }
    });
  }
  return Y;
}

void absorb_noalloc(Ctx& c, const std::vector<int64_t>& dx, const float* X, const std::vector<int>& lx,
                    const std::vector<int64_t>& da, const float* A, const std::vector<int>& la, const float* Ad,
                    const std::vector<int>& lad, const float* G, float* Y, float* gX, float* gA, float* gAd) {
  // shape-only tensors for the planner (no device memory behind them)
  Tensor sX, sA, sAd;
  sX.dt = sA.dt = sAd.dt = DType::F32;
  sX.dims = dx;
  sA.dims = da;
  sAd.dims = da;
  const AbsorbPlan pl = plan_absorb(sX, lx, sA, la, sAd, lad);
  IPEPS_CHECK(pl.ok, "absorb (no allocation): the labels and shapes must fit the fused FP32 kernel "
                     "(X with two open legs and two legs shared with each of A and Ad, A and Ad (D,D,D,D,d) "
                     "sharing the physical leg, 4 <= D <= 6, d <= 2)");
  IPEPS_CHECK(X && A && Ad, "absorb (no allocation): X, A and Ad are required");
  IPEPS_CHECK(Y || (G && (gX || gA || gAd)), "absorb (no allocation): nothing to compute");
  IPEPS_CHECK(G || !(gX || gA || gAd), "absorb (no allocation): the cotangents need G");
  const int U = pl.D * pl.D;
  auto fwd = [&](const float* in, const k::AbsorbLegs& il, bool transposed, float* out, const k::AbsorbLegs& ol) {
    k::AbsorbP p;
    p.U = U;
    p.P = pl.P;
    p.D = pl.D;
    p.x = il;
    p.y = ol;
    p.X = in;
    p.Y = out;
    p.Am = A;
    p.Adm = Ad;
    const auto sa = site_strides(sA, pl.permA, transposed), sd = site_strides(sAd, pl.permAd, transposed);
    for (int i = 0; i < 5; ++i) {
      p.sa[i] = sa[i];
      p.sd[i] = sd[i];
    }
    k::absorb_fwd(c, p);
  };
  if (Y) fwd(X, pl.xl, false, Y, pl.yl);
  if (!G) return;
  if (gX) fwd(G, pl.yl, true, gX, pl.xl);  // Xbar_r = sum_p A_p G_r Ad_p^T
  if (gA || gAd) {
    k::AbsorbWP w;
    w.U = U;
    w.P = pl.P;
    w.D = pl.D;
    w.x = pl.xl;
    w.g = pl.yl;
    w.X = X;
    w.G = G;
    w.Am = A;
    w.Adm = Ad;
    const auto ska = site_strides(sA, pl.permA, false), skd = site_strides(sAd, pl.permAd, false);
    for (int i = 0; i < 5; ++i) {
      w.sa[i] = ska[i];
      w.sd[i] = skd[i];
    }
    const float* out = k::absorb_wgrad_scratch(c, w);
    // (u1, u2, a1, a2, p) resp. (v1, v2, b1, b2, p) -> the site tensor's leg order
    auto back = [&](const float* g, const std::vector<int>& perm, float* dst) {
      std::vector<int> inv(5);
      for (int i = 0; i < 5; ++i) inv[perm[i]] = i;
      std::vector<int64_t> gd = {da[perm[0]], da[perm[1]], da[perm[2]], da[perm[3]], da[perm[4]]};
      k::permute(c, DType::F32, gd, inv, g, dst, false);
    };
    if (gA) back(out, pl.permA, gA);
    if (gAd) back(out + (int64_t)U * U * pl.P, pl.permAd, gAd);
  }
}

Tensor adjoint_matmul(Ctx& c, const Tensor& A, const Tensor& B) {
  return tcon(c, conj(c, A), {1, -1}, B, {1, -2});
}

// ---------------------------------------------------------------------------
// renormalize / scalings
// ---------------------------------------------------------------------------
Tensor renormalize(Ctx& c, const Tensor& t) {
  auto sc = std::make_shared<Buf>();
  sc->bytes = 32;
  sc->p = c.pool->alloc(32);
  sc->pool = c.pool.get();
  double* m = (double*)sc->p;
  int64_t* idx = (int64_t*)((char*)sc->p + 8);
  k::maxabs(c, t.dt, t.numel(), t.data(), m, idx);
  Tensor r = c.empty(t.dt, t.dims);
  k::scale_by_dev(c, t.dt, t.numel(), t.data(), r.data(), m, 0);
  if (needs(t)) {
    auto in = t.node;
    Tensor x = t;
    x.node.reset();
    r.node = make_node([in, x, sc](Ctx& c, Node& n) {
      double* mv = (double*)sc->p;
      int64_t* iv = (int64_t*)((char*)sc->p + 8);
      double* dotv = (double*)((char*)sc->p + 16);
      k::re_dot(c, x.dt, x.numel(), n.grad.data(), x.data(), dotv);
      Tensor g = c.empty(x.dt, x.dims);
      k::renorm_back(c, x.dt, x.numel(), x.data(), n.grad.data(), mv, iv, dotv, g.data());
      acc(c, in, g);
    });
  }
  return r;
}

Tensor renormalize_by(Ctx& c, const Tensor& x, const Tensor& ref) {
  auto sc = std::make_shared<Buf>();
  sc->bytes = 32;
  sc->p = c.pool->alloc(32);
  sc->pool = c.pool.get();
  double* m = (double*)sc->p;
  int64_t* idx = (int64_t*)((char*)sc->p + 8);
  k::maxabs(c, ref.dt, ref.numel(), ref.data(), m, idx);
  Tensor r = c.empty(x.dt, x.dims);
  k::scale_by_dev(c, x.dt, x.numel(), x.data(), r.data(), m, 0);
  if (needs(x) || needs(ref)) {
    auto nx = x.node, nr = ref.node;
    Tensor xk = x, rk = ref;
    xk.node.reset();
    rk.node.reset();
    r.node = make_node([nx, nr, xk, rk, sc](Ctx& c, Node& n) {
      double* mv = (double*)sc->p;
      int64_t* iv = (int64_t*)((char*)sc->p + 8);
      double* dotv = (double*)((char*)sc->p + 16);
      if (nx) {
        Tensor g = c.empty(xk.dt, xk.dims);
        k::scale_by_dev(c, xk.dt, xk.numel(), n.grad.data(), g.data(), mv, 0);
        acc(c, nx, g);
      }
      if (nr) {
        k::re_dot(c, xk.dt, xk.numel(), n.grad.data(), xk.data(), dotv);
        Tensor z = c.zeros(rk.dt, rk.dims);
        Tensor g = c.empty(rk.dt, rk.dims);
        k::renorm_back(c, rk.dt, rk.numel(), rk.data(), z.data(), mv, iv, dotv, g.data());
        acc(c, nr, g);
      }
    });
  }
  return r;
}

Tensor scale_vec(Ctx& c, const Tensor& X, const Tensor& s, bool rowwise) {
  IPEPS_CHECK(X.nd() == 2 && s.dt == DType::F64, "scale_vec: matrix and FP64 vector expected");
  IPEPS_CHECK(s.numel() == (rowwise ? X.dims[0] : X.dims[1]), "scale_vec: length mismatch");
  Tensor Y = c.empty(X.dt, X.dims);
  k::scale_cols(c, X.dt, X.dims[0], X.dims[1], X.data(), s.ptr<double>(), Y.data(), rowwise);
  if (needs(X) || needs(s)) {
    auto nX = X.node, nS = s.node;
    Tensor Xk = X, sk = s;
    Xk.node.reset();
    sk.node.reset();
    Y.node = make_node([=](Ctx& c, Node& n) {
      if (nX) {
        Tensor g = c.empty(Xk.dt, Xk.dims);
        k::scale_cols(c, Xk.dt, Xk.dims[0], Xk.dims[1], n.grad.data(), sk.ptr<double>(), g.data(), rowwise);
        acc(c, nX, g);
      }
      if (nS) {
        Tensor g = c.empty(DType::F64, sk.dims);
        k::col_re_dot(c, Xk.dt, Xk.dims[0], Xk.dims[1], Xk.data(), n.grad.data(), g.ptr<double>(), rowwise);
        acc(c, nS, g);
      }
    });
  }
  return Y;
}

Tensor inv_sqrt(Ctx& c, const Tensor& s) {
  IPEPS_CHECK(s.dt == DType::F64, "inv_sqrt: FP64 vector expected");
  Tensor r = c.empty(DType::F64, s.dims);
  k::inv_sqrt(c, s.numel(), s.ptr<double>(), r.ptr<double>());
  if (needs(s)) {
    auto in = s.node;
    Tensor sk = s;
    sk.node.reset();
    r.node = make_node([in, sk](Ctx& c, Node& n) {
      Tensor g = c.zeros(DType::F64, sk.dims);
      k::inv_sqrt_back(c, sk.numel(), sk.ptr<double>(), n.grad.ptr<double>(), g.ptr<double>());
      acc(c, in, g);
    });
  }
  return r;
}

// ---------------------------------------------------------------------------
// rocSOLVER wrappers
// ---------------------------------------------------------------------------
static void solver_check(rocblas_status s, const char* what) {
  if (s != rocblas_status_success)
    throw Error(IPEPS_ERR_SOLVER, std::string(what) + " failed with rocblas_status " + std::to_string((int)s));
}

// Full SVD of a square/rectangular F64/C128 matrix: A = U diag(S) Vh, returns V = Vh^H.
// left: rocblas_svect_all or _singular.
static void gesvd(Ctx& c, const Tensor& Ain, Tensor& U, Tensor& S, Tensor& V, bool thin) {
  IPEPS_CHECK(Ain.nd() == 2, "gesvd: matrix expected");
  int64_t m = Ain.dims[0], n = Ain.dims[1];
  int64_t mn = std::min(m, n);
  Tensor A = clone(c, Ain);
  rocblas_handle h = (rocblas_handle)c.solver;
  int64_t ucols = thin ? mn : m;
  U = c.empty(Ain.dt, {m, ucols});
  S = c.empty(DType::F64, {mn});
  Tensor Vh = c.empty(Ain.dt, {n, n});
  Tensor E = c.empty(DType::F64, {std::max<int64_t>(1, mn)});
  Tensor info = c.empty(DType::F64, {1});
  rocblas_svect ls = thin ? rocblas_svect_singular : rocblas_svect_all;
  std::unique_ptr<TimedLaunch> tl(new TimedLaunch(c, thin ? "rocsolver_gesvd_thin" : "rocsolver_gesvd", 0.0, 0.0));
  if (Ain.dt == DType::F64) {
    solver_check(rocsolver_dgesvd(h, ls, rocblas_svect_all, (rocblas_int)m, (rocblas_int)n, A.ptr<double>(),
                                  (rocblas_int)m, S.ptr<double>(), U.ptr<double>(), (rocblas_int)m,
                                  Vh.ptr<double>(), (rocblas_int)n, E.ptr<double>(), rocblas_outofplace,
                                  (rocblas_int*)info.data()),
                 "rocsolver_dgesvd");
  } else if (Ain.dt == DType::C128) {
    solver_check(rocsolver_zgesvd(h, ls, rocblas_svect_all, (rocblas_int)m, (rocblas_int)n,
                                  (rocblas_double_complex*)A.data(), (rocblas_int)m, S.ptr<double>(),
                                  (rocblas_double_complex*)U.data(), (rocblas_int)m,
                                  (rocblas_double_complex*)Vh.data(), (rocblas_int)n, E.ptr<double>(),
                                  rocblas_outofplace, (rocblas_int*)info.data()),
                 "rocsolver_zgesvd");
  } else {
    throw Error(IPEPS_ERR_ARG, "gesvd: F64 or C128 only");
  }
  tl.reset();
  int hinfo = 0;
  HIP_CHECK(hipMemcpyAsync(&hinfo, info.data(), sizeof(int), hipMemcpyDeviceToHost, c.stream));
  c.sync();
  if (hinfo != 0) throw Error(IPEPS_ERR_SOLVER, "gesvd did not converge, info=" + std::to_string(hinfo));
  // V = Vh^H, keep the first mn columns when thin
  Tensor Vfull = permute_raw(c, Vh, {1, 0}, true);
  if (thin && mn < n) {
    V = c.empty(Ain.dt, {n, mn});
    HIP_CHECK(hipMemcpyAsync(V.data(), Vfull.data(), V.bytes(), hipMemcpyDeviceToDevice, c.stream));
  } else {
    V = Vfull;
  }
}

void svd_thin_raw(Ctx& c, const Tensor& X, Tensor& W, Tensor& s, Tensor& Z) { gesvd(c, X, W, s, Z, true); }

// Thin one-sided-Jacobi SVD of a tall matrix (rocSOLVER gesvdj): X = U diag(S) V^H.
static void gesvdj_thin(Ctx& c, const Tensor& Ain, Tensor& U, Tensor& S, Tensor& V) {
  IPEPS_CHECK(Ain.nd() == 2 && Ain.dims[0] >= Ain.dims[1], "gesvdj: tall matrix expected");
  const int64_t m = Ain.dims[0], n = Ain.dims[1];
  Tensor A = clone(c, Ain);
  U = c.empty(Ain.dt, {m, n});
  S = c.empty(DType::F64, {n});
  Tensor Vh = c.empty(Ain.dt, {n, n});
  Tensor scratch = c.empty(DType::F64, {4});  // residual, n_sweeps, info
  double* res = scratch.ptr<double>();
  rocblas_int* nsw = (rocblas_int*)(res + 1);
  rocblas_int* info = (rocblas_int*)(res + 2);
  rocblas_handle h = (rocblas_handle)c.solver;
  {
    TimedLaunch tl(c, "rocsolver_gesvdj", 0.0, 0.0);
    const double tol = 1e-15;
    if (Ain.dt == DType::F64)
      solver_check(rocsolver_dgesvdj(h, rocblas_svect_singular, rocblas_svect_singular, (rocblas_int)m, (rocblas_int)n,
                                     A.ptr<double>(), (rocblas_int)m, tol, res, 100, nsw, S.ptr<double>(),
                                     U.ptr<double>(), (rocblas_int)m, Vh.ptr<double>(), (rocblas_int)n, info),
                   "rocsolver_dgesvdj");
    else if (Ain.dt == DType::C128)
      solver_check(rocsolver_zgesvdj(h, rocblas_svect_singular, rocblas_svect_singular, (rocblas_int)m, (rocblas_int)n,
                                     (rocblas_double_complex*)A.data(), (rocblas_int)m, tol, res, 100, nsw,
                                     S.ptr<double>(), (rocblas_double_complex*)U.data(), (rocblas_int)m,
                                     (rocblas_double_complex*)Vh.data(), (rocblas_int)n, info),
                   "rocsolver_zgesvdj");
    else
      throw Error(IPEPS_ERR_ARG, "gesvdj: F64 or C128 only");
  }
  int hinfo = 0;
  HIP_CHECK(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  c.sync();
  if (hinfo != 0) throw Error(IPEPS_ERR_SOLVER, "gesvdj did not converge, info=" + std::to_string(hinfo));
  V = permute_raw(c, Vh, {1, 0}, true);
}
void svd_full_raw(Ctx& c, const Tensor& X, Tensor& U, Tensor& s, Tensor& V) { gesvd(c, X, U, s, V, false); }

namespace {
// counters of the projector being built on this thread (null: none)
SvdStats*& orth_stats() {
  thread_local SvdStats* s = nullptr;
  return s;
}
void orth_count(long SvdStats::*field) {
  if (SvdStats* s = orth_stats()) {
    std::lock_guard<std::mutex> g(s->mu);
    ++(s->*field);
  }
}
struct OrthStatsScope {
  SvdStats* prev;
  explicit OrthStatsScope(SvdStats* s) : prev(orth_stats()) { orth_stats() = s; }
  ~OrthStatsScope() { orth_stats() = prev; }
};
}  // namespace

static Tensor orth_householder(Ctx& c, const Tensor& Yin) {
  IPEPS_CHECK(Yin.nd() == 2 && Yin.dims[0] >= Yin.dims[1], "orth: tall matrix expected");
  int64_t m = Yin.dims[0], n = Yin.dims[1];
  Tensor Y = clone(c, Yin);
  Tensor tau = c.empty(Yin.dt, {n});
  rocblas_handle h = (rocblas_handle)c.solver;
  TimedLaunch tl(c, "rocsolver_geqrf_orgqr", 0.0, 0.0);
  if (Y.dt == DType::F64) {
    solver_check(rocsolver_dgeqrf(h, (rocblas_int)m, (rocblas_int)n, Y.ptr<double>(), (rocblas_int)m, tau.ptr<double>()),
                 "rocsolver_dgeqrf");
    solver_check(rocsolver_dorgqr(h, (rocblas_int)m, (rocblas_int)n, (rocblas_int)n, Y.ptr<double>(), (rocblas_int)m,
                                  tau.ptr<double>()),
                 "rocsolver_dorgqr");
  } else if (Y.dt == DType::C128) {
    solver_check(rocsolver_zgeqrf(h, (rocblas_int)m, (rocblas_int)n, (rocblas_double_complex*)Y.data(), (rocblas_int)m,
                                  (rocblas_double_complex*)tau.data()),
                 "rocsolver_zgeqrf");
    solver_check(rocsolver_zungqr(h, (rocblas_int)m, (rocblas_int)n, (rocblas_int)n, (rocblas_double_complex*)Y.data(),
                                  (rocblas_int)m, (rocblas_double_complex*)tau.data()),
                 "rocsolver_zungqr");
  } else {
    throw Error(IPEPS_ERR_ARG, "orth: F64 or C128 only");
  }
  return Y;
}

// Cholesky factor (upper) in place: G = R^H R; LAPACK-style info to a device int.
static void potrf_upper(Ctx& c, Tensor& G, int* info) {
  const rocblas_int n = (rocblas_int)G.dims[0];
  rocblas_handle h = (rocblas_handle)c.solver;
  TimedLaunch tl(c, "rocsolver_potrf", 0.0, 0.0);
  if (G.dt == DType::F64)
    solver_check(rocsolver_dpotrf(h, rocblas_fill_upper, n, G.ptr<double>(), n, (rocblas_int*)info), "rocsolver_dpotrf");
  else if (G.dt == DType::C128)
    solver_check(rocsolver_zpotrf(h, rocblas_fill_upper, n, (rocblas_double_complex*)G.data(), n, (rocblas_int*)info),
                 "rocsolver_zpotrf");
  else
    throw Error(IPEPS_ERR_ARG, "potrf: F64 or C128 only");
}

// X <- X R^-1 for an upper-triangular R (n x n) and X (m x n).
static void trsm_right_upper(Ctx& c, const Tensor& R, Tensor& X) {
  const rocblas_int m = (rocblas_int)X.dims[0], n = (rocblas_int)X.dims[1];
  rocblas_handle h = (rocblas_handle)c.solver;
  solver_check(rocblas_set_pointer_mode(h, rocblas_pointer_mode_host), "rocblas_set_pointer_mode");
  TimedLaunch tl(c, "rocblas_trsm", 0.0, 0.0);
  if (X.dt == DType::F64) {
    const double one = 1.0;
    solver_check(rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                               rocblas_diagonal_non_unit, m, n, &one, R.ptr<double>(), n, X.ptr<double>(), m),
                 "rocblas_dtrsm");
  } else if (X.dt == DType::C128) {
    const rocblas_double_complex one(1.0, 0.0);
    solver_check(rocblas_ztrsm(h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                               rocblas_diagonal_non_unit, m, n, &one, (const rocblas_double_complex*)R.data(), n,
                               (rocblas_double_complex*)X.data(), m),
                 "rocblas_ztrsm");
  } else {
    throw Error(IPEPS_ERR_ARG, "trsm: F64 or C128 only");
  }
}

// FP64 GEMM on sub-blocks: C = alpha op(A) op(B) + beta C (leading dimensions given).
static void gemm_f64(Ctx& c, bool tA, const double* A, int64_t lda, bool tB, const double* B, int64_t ldb, double* C,
                     int64_t ldc, int64_t M, int64_t N, int64_t K, double alpha, double beta) {
  k::GemmArgs g;
  g.dt = DType::F64;
  g.transA = tA;
  g.transB = tB;
  g.M = M; g.N = N; g.K = K;
  g.alpha = alpha;
  g.beta = beta;
  g.A = A; g.lda = lda;
  g.B = B; g.ldb = ldb;
  g.C = C; g.ldc = ldc;
  k::gemm(c, g);
}

// L^-1 of the Cholesky factor of an SPD n x n FP64 G (lower triangle read),
// n <= 2 SCHOL_MAXN_HOST, into Li (ldl; zeros above the diagonal): the
// one-workgroup kernel directly, or on the two diagonal blocks of a 2 x 2
// split,
//   G = [G11, G21^T; G21, G22],  L11 L11^T = G11,  X = G21 L11^-T (= L21),
//   L22 L22^T = G22 - X X^T,
//   L^-1 = [L11^-1, 0; -L22^-1 X L11^-1, L22^-1].
// info1 / info2: the blocks' kernel infos (pivot j + 1 of the block on failure).
static void chol_inv(Ctx& c, int n, const double* G, int64_t ldg, double* Li, int64_t ldl, int* info1, int* info2) {
  if (n <= k::SCHOL_MAXN_HOST) {
    k::small_chol_inv(c, n, G, ldg, Li, ldl, info1);
    return;
  }
  IPEPS_CHECK(n <= 2 * k::SCHOL_MAXN_HOST, "chol_inv: n out of range");
  const int n2 = n / 2, n1 = n - n2;
  double* L11i = Li;
  double* L22i = Li + n1 + (int64_t)n1 * ldl;
  k::small_chol_inv(c, n1, G, ldg, L11i, ldl, info1);
  Tensor X = c.empty(DType::F64, {n2, n1});
  gemm_f64(c, false, G + n1, ldg, true, L11i, ldl, X.ptr<double>(), n2, n2, n1, n1, 1.0, 0.0);
  Tensor S = c.empty(DType::F64, {n2, n2});
  dev_copy2d(c, S.data(), n2 * 8, G + n1 + (int64_t)n1 * ldg, ldg * 8, n2 * 8, n2);
  gemm_f64(c, false, X.ptr<double>(), n2, true, X.ptr<double>(), n2, S.ptr<double>(), n2, n2, n2, n1, -1.0, 1.0);
  k::small_chol_inv(c, n2, S.ptr<double>(), n2, L22i, ldl, info2);
  Tensor T = c.empty(DType::F64, {n2, n1});
  gemm_f64(c, false, X.ptr<double>(), n2, false, L11i, ldl, T.ptr<double>(), n2, n2, n1, n1, 1.0, 0.0);
  gemm_f64(c, false, L22i, ldl, false, T.ptr<double>(), n2, Li + n1, ldl, n2, n1, n2, -1.0, 0.0);
  dev_zero2d(c, Li + (int64_t)n1 * ldl, ldl * 8, n1 * 8, n2);
}

// CholeskyQR2: Q1 = Y R1^-1 with R1^H R1 = Y^H Y, then the same on Q1.  The
// second pass is trusted when its Gram is within 1/2 of I (then cond(Q1)^2 <= 3
// and Q is orthonormal to working precision).  Returns false when a Cholesky
// failed or that check did not hold; Q is then undefined.
// CholeskyQR2 for FP64 with n <= 160 on the one-workgroup blocked Cholesky +
// triangular inverse: Q1 = Y L1^-T, Q = Q1 L2^-T (GEMMs instead of the
// triangular solves).  Same acceptance test and second-pass skip as
// orth_cholqr2.
static bool orth_cholqr2_small(Ctx& c, const Tensor& Y, Tensor& Q) {
  const int64_t n = Y.dims[1];
  Tensor st = c.zeros(DType::F64, {5});  // [info of each block, pass 1 / pass 2] x 2, |G2 - I|_max
  auto info = [&](int i) { return (int*)(st.ptr<double>() + i); };
  auto failed = [](const double* h, int i0) {
    int a, b;
    std::memcpy(&a, &h[i0], sizeof(int));
    std::memcpy(&b, &h[i0 + 1], sizeof(int));
    return a != 0 || b != 0;
  };
  Tensor G = matmul_raw(c, Y, true, false, Y, false, false);
  Tensor Li = c.empty(DType::F64, {n, n});
  chol_inv(c, (int)n, G.ptr<double>(), n, Li.ptr<double>(), n, info(0), info(1));
  Tensor Q1 = matmul_raw(c, Y, false, false, Li, true, false);
  Tensor G2 = matmul_raw(c, Q1, true, false, Q1, false, false);
  k::eye_dev(c, DType::F64, n, G2.data(), st.ptr<double>() + 4);
  if (c.graph_safe) {
    // both passes, checked after the step
    chol_inv(c, (int)n, G2.ptr<double>(), n, Li.ptr<double>(), n, info(2), info(3));
    Q = matmul_raw(c, Q1, false, false, Li, true, false);
    record_check(c, CHECK_CHOLQR2, st.data(), 5);
    return true;
  }
  double hs[5];
  HIP_CHECK(hipMemcpyAsync(hs, st.data(), sizeof(hs), hipMemcpyDeviceToHost, c.stream));
  c.sync();
  if (failed(hs, 0) || !(hs[4] <= 0.5)) return false;
  if (hs[4] <= 1e-12) {
    Q = Q1;
    return true;
  }
  orth_count(&SvdStats::orth_second);
  chol_inv(c, (int)n, G2.ptr<double>(), n, Li.ptr<double>(), n, info(2), info(3));
  Q = matmul_raw(c, Q1, false, false, Li, true, false);
  HIP_CHECK(hipMemcpyAsync(hs, st.data(), sizeof(hs), hipMemcpyDeviceToHost, c.stream));
  c.sync();
  return !failed(hs, 2);
}

// One CholeskyQR pass (FP64, n <= 320) with no host synchronisation: Q = Y L^-T
// from the Cholesky factor L of Y^T Y; the blocks' infos go to info[0..1]
// for a later check.  For the intermediate bases of the power iterations,
// which only carry the span (orthogonality error ~ u cond(Y)^2).
static Tensor orth_cholqr1_nosync(Ctx& c, const Tensor& Y, int* info) {
  const int64_t n = Y.dims[1];
  Tensor G = matmul_raw(c, Y, true, false, Y, false, false);
  Tensor Li = c.empty(DType::F64, {n, n});
  chol_inv(c, (int)n, G.ptr<double>(), n, Li.ptr<double>(), n, info, info + 1);
  return matmul_raw(c, Y, false, false, Li, true, false);
}

// The second pass is skipped when the first already left Q orthonormal to
// |Q^H Q - I|_max <= 1e-12 (well-conditioned Y: cond(Y)^2 u well below that).
static bool orth_cholqr2(Ctx& c, const Tensor& Y, Tensor& Q, bool small_chol) {
  const int64_t n = Y.dims[1];
  if (small_chol && Y.dt == DType::F64 && n <= 2 * k::SCHOL_MAXN_HOST) return orth_cholqr2_small(c, Y, Q);
  Tensor st = c.zeros(DType::F64, {3});  // [info1, info2, |G2 - I|_max]
  int* info1 = (int*)st.ptr<double>();
  int* info2 = (int*)(st.ptr<double>() + 1);
  Tensor G = matmul_raw(c, Y, true, true, Y, false, false);
  potrf_upper(c, G, info1);
  Q = clone(c, Y);
  trsm_right_upper(c, G, Q);
  Tensor G2 = matmul_raw(c, Q, true, true, Q, false, false);
  k::eye_dev(c, Q.dt, n, G2.data(), st.ptr<double>() + 2);
  double hs[3];
  HIP_CHECK(hipMemcpyAsync(hs, st.data(), sizeof(hs), hipMemcpyDeviceToHost, c.stream));
  c.sync();
  int i1, i2;
  std::memcpy(&i1, &hs[0], sizeof(int));
  if (i1 != 0 || !(hs[2] <= 0.5)) return false;
  if (hs[2] <= 1e-12) return true;
  orth_count(&SvdStats::orth_second);
  potrf_upper(c, G2, info2);
  trsm_right_upper(c, G2, Q);
  HIP_CHECK(hipMemcpyAsync(hs, st.data(), sizeof(hs), hipMemcpyDeviceToHost, c.stream));
  c.sync();
  std::memcpy(&i2, &hs[1], sizeof(int));
  return i2 == 0;
}

Tensor orth(Ctx& c, const Tensor& Y, int fast) {
  IPEPS_CHECK(Y.nd() == 2 && Y.dims[0] >= Y.dims[1], "orth: tall matrix expected");
  Tensor Q;
  orth_count(&SvdStats::orth_calls);
  if (fast && orth_cholqr2(c, Y, Q, fast == 2)) return Q;
  orth_count(&SvdStats::orth_householder);
  return orth_householder(c, Y);
}

// Thin SVD of a tall B (m x n) = W diag(s) X^H from the eigenpairs of B^H B:
// s = sqrt(lambda) descending, X the eigenvectors, W = B X diag(1/s) followed
// by one Newton-Schulz step W <- W (3I - W^H W)/2.  Eigenvalues carry an
// absolute error ~ u lambda_max, so the route is taken only when
// s_min / s_max >= guard; returns false otherwise (outputs undefined).
static bool gram_svd(Ctx& c, const Tensor& B, double guard, Tensor& W, Tensor& s, Tensor& X, bool jacobi,
                     bool stedc) {
  const int64_t m = B.dims[0], n = B.dims[1];
  Tensor G = matmul_raw(c, B, true, true, B, false, false);
  Tensor lam = c.empty(DType::F64, {n});
  Tensor E = c.empty(DType::F64, {n});
  Tensor st = c.zeros(DType::F64, {2});  // [info, |Z Z^T - I|_max of the tridiagonal solver's vectors]
  Tensor lam_used = lam;
  Tensor Qt;  // the tridiagonal form's Q while the rocSOLVER fallback may still need it
  auto run_stedc = [&](Tensor& Q) {
    TimedLaunch tl(c, "rocsolver_stedc", 0.0, 0.0);
    solver_check(rocsolver_dstedc((rocblas_handle)c.solver, rocblas_evect_original, (rocblas_int)n, lam.ptr<double>(),
                                  E.ptr<double>(), Q.ptr<double>(), (rocblas_int)n, (rocblas_int*)st.data()),
                 "rocsolver_dstedc");
  };
  rocblas_handle h = (rocblas_handle)c.solver;
  if (jacobi) {
    // kernels only: the one-workgroup Jacobi eigensolver (graph-safe mode
    // defers its check)
    IPEPS_CHECK(B.dt == DType::F64 && n <= k::JAC_MAXN_HOST, "Jacobi small SVD: FP64 and k <= 112 only");
    Tensor Q = c.empty(DType::F64, {n, n});
    k::jacobi_eig(c, (int)n, G.ptr<double>(), n, lam.ptr<double>(), Q.ptr<double>(), n, (int*)st.data());
    if (c.graph_safe) record_check(c, CHECK_INFO, st.data(), 1);
    G = Q;
  } else if (B.dt == DType::F64 && n <= k::TRI_MAXN_HOST) {
    // Householder tridiagonalisation and Q on the device (one workgroup + one
    // wave per column), then the tridiagonal eigenpairs by bisection and
    // inverse iteration (kernels) or by rocSOLVER's divide and conquer
    Tensor Q = c.empty(DType::F64, {n, n});
    k::sytrd_q(c, (int)n, G.ptr<double>(), n, lam.ptr<double>(), E.ptr<double>(), Q.ptr<double>(), n);
    IPEPS_CHECK(!(stedc && c.graph_safe), "graph-safe small SVD: no rocSOLVER stedc");
    if (stedc) {
      run_stedc(Q);
      G = Q;
    } else {
      // T = Z diag(lam) Z^T with Z's columns the rows of Zt; two Newton-Schulz
      // steps Zt <- (3I - Zt Zt^T) Zt / 2 make the rows orthonormal (the first
      // Gram's distance from I, e, goes to st[1]: after the steps the distance
      // is ~ e^4, so e <= 1e-3 is accepted; otherwise rocSOLVER redoes it)
      lam_used = c.empty(DType::F64, {n});
      Tensor Zt = c.empty(DType::F64, {n, n});
      k::tri_eig(c, (int)n, lam.ptr<double>(), E.ptr<double>(), lam_used.ptr<double>(), Zt.ptr<double>());
      Tensor S = matmul_raw(c, Zt, false, false, Zt, true, false);
      k::eye_dev(c, DType::F64, n, S.data(), st.ptr<double>() + 1);
      k::ns_coeff(c, DType::F64, n, S.data());
      Tensor Z1 = matmul_raw(c, S, false, false, Zt, false, false);
      Tensor S2 = matmul_raw(c, Z1, false, false, Z1, true, false);
      k::ns_coeff(c, DType::F64, n, S2.data());
      Tensor Z2 = matmul_raw(c, S2, false, false, Z1, false, false);
      G = matmul_raw(c, Q, false, false, Z2, true, false);
      Qt = Q;
      // graph-safe: no fallback inside a captured step; a rejected set fails
      // the step's deferred checks (the replay is undone and run eagerly)
      if (c.graph_safe) record_check(c, CHECK_AT_MOST, st.ptr<double>() + 1, 1, 0, 0, 1e-3);
    }
  } else {
    IPEPS_CHECK(!c.graph_safe, "graph-safe small SVD: F64 and k <= 320 only");
    TimedLaunch tl(c, "rocsolver_syevd", 0.0, 0.0);
    if (B.dt == DType::F64)
      solver_check(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)n, G.ptr<double>(),
                                    (rocblas_int)n, lam.ptr<double>(), E.ptr<double>(), (rocblas_int*)st.data()),
                   "rocsolver_dsyevd");
    else if (B.dt == DType::C128)
      solver_check(rocsolver_zheevd(h, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)n,
                                    (rocblas_double_complex*)G.data(), (rocblas_int)n, lam.ptr<double>(),
                                    E.ptr<double>(), (rocblas_int*)st.data()),
                   "rocsolver_zheevd");
    else
      throw Error(IPEPS_ERR_ARG, "gram_svd: F64 or C128 only");
  }
  X = c.empty(B.dt, {n, n});
  s = c.empty(DType::F64, {n});
  Tensor sinv = c.empty(DType::F64, {n});
  k::eig_to_svd(c, B.dt, n, lam_used.ptr<double>(), G.data(), s.ptr<double>(), X.data(), sinv.ptr<double>());
  if (c.graph_safe) {
    if (guard > 0.0) record_check(c, CHECK_GUARD, s.data(), n, 0, 0, guard);
  } else {
  std::vector<double> sh(n);
  double hst[2] = {0.0, 0.0};
  HIP_CHECK(hipMemcpyAsync(sh.data(), s.data(), n * 8, hipMemcpyDeviceToHost, c.stream));
  HIP_CHECK(hipMemcpyAsync(hst, st.data(), 16, hipMemcpyDeviceToHost, c.stream));
  c.sync();
  if (Qt.data() && !(hst[1] <= 1e-3)) {
    // the inverse-iteration vectors were not independent enough: rocSOLVER's
    // divide and conquer on the same (d, e, Q)
    orth_count(&SvdStats::tri_fallbacks);
    run_stedc(Qt);
    k::eig_to_svd(c, B.dt, n, lam.ptr<double>(), Qt.data(), s.ptr<double>(), X.data(), sinv.ptr<double>());
    HIP_CHECK(hipMemcpyAsync(sh.data(), s.data(), n * 8, hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipMemcpyAsync(hst, st.data(), 8, hipMemcpyDeviceToHost, c.stream));
    c.sync();
  }
  int info;
  std::memcpy(&info, &hst[0], sizeof(int));
  if (info != 0 || !(sh[0] > 0.0) || !(sh[n - 1] >= guard * sh[0])) return false;
  }
  Tensor BX = matmul_raw(c, B, false, false, X, false, false);
  Tensor W0 = c.empty(B.dt, {m, n});
  k::scale_cols(c, B.dt, m, n, BX.data(), sinv.ptr<double>(), W0.data(), false);
  Tensor T = matmul_raw(c, W0, true, true, W0, false, false);
  k::ns_coeff(c, B.dt, n, T.data());
  W = matmul_raw(c, W0, false, false, T, false, false);
  return true;
}

int svd_tall(Ctx& c, const Tensor& B, int method, double guard, Tensor& W, Tensor& s, Tensor& X) {
  IPEPS_CHECK(B.nd() == 2 && B.dims[0] >= B.dims[1], "svd_tall: tall matrix expected");
  if (method >= 1 && gram_svd(c, B, guard, W, s, X, method == 2, method == 3)) return method;
  IPEPS_CHECK(!c.graph_safe, "graph-safe small SVD: the Gram route only");
  gesvdj_thin(c, B, W, s, X);
  return 0;
}

// ---------------------------------------------------------------------------
// truncated SVD of M = R1 R2 with the reference cutoff rule
// ---------------------------------------------------------------------------
static int64_t cutoff_rule(const std::vector<double>& s, int64_t chi, double rel) {
  // src/utility/wrapper.jl:37-55 (rel = 1e-12 there)
  double smax = 0.0;
  for (double v : s) smax = std::max(smax, v);
  int64_t above = 0;
  for (double v : s) above += (v / smax > rel);
  int64_t n = std::min<int64_t>(chi, above);
  if ((int64_t)s.size() > n && n >= 1 && std::fabs(s[n] - s[n - 1]) < 1e-10) n += 1;
  return n;
}

static Tensor first_cols(Ctx& c, const Tensor& X, int64_t n) {
  if (X.dims[1] == n) return X;
  Tensor r = c.empty(X.dt, {X.dims[0], n});
  dev_copy(c, r.data(), X.data(), r.bytes());
  return r;
}

static Tensor first_n(Ctx& c, const Tensor& s, int64_t n) {
  Tensor r = c.empty(s.dt, {n});
  dev_copy(c, r.data(), s.data(), r.bytes());
  return r;
}

// Discarded-subspace term of svd_back (src/utility/autodiff.jl:41-75) beyond
// the k computed triplets.  With the full factorisation and dU, dV zero past
// the n kept columns, the pairs (kept i, discarded j) contribute
//   sum_j u_j [s_i u_j^H du_i + s_j v_j^H dv_i] / (s_i^2 - s_j^2) v_i^H
//       + u_i ( sum_j v_j [s_j u_j^H du_i + s_i v_j^H dv_i] / (s_i^2 - s_j^2) )^H
// = x_i v_i^H + u_i y_i^H, and (x_i, y_i) solves
//   s_i x - Mp y = a_i,   s_i y - Mp^H x = b_i,
// Mp = (I - Ub Ub^H) M (I - Vb Vb^H), a = (I - Ub Ub^H) dU, b = (I - Vb Vb^H) dV
// (expand x, y over the discarded u_j, v_j: the coefficients above are the
// solution, term by term).  The block operator [s_i, -Mp; -Mp^H, s_i] is
// Hermitian with eigenvalues s_i -+ s_j, positive definite while s_i exceeds
// every singular value of Mp, so one conjugate-gradient solve per column runs
// on the whole block at once (M applied as R1 (R2 .), never formed).  X, Y
// hold the starting point (a / s_i, b / s_i: the thin-SVD terms) and receive
// the solution; on a non-positive curvature they are left at that start.
static void complement_cg(Ctx& c, DType dt, DType sdt, const Tensor& R1, const Tensor& R2, const Tensor& Ub,
                          const Tensor& Vb, const Tensor& Sn, const Tensor& a, const Tensor& b, Tensor& X,
                          Tensor& Y, int iters, double tol, const std::shared_ptr<SvdStats>& stats) {
  const int64_t m = a.dims[0], nc = b.dims[0], n = a.dims[1];
  auto proj_out = [&](const Tensor& B, const Tensor& Z) {  // (I - B B^H) Z
    Tensor t = matmul_raw(c, B, true, true, Z, false, false);
    Tensor r = clone(c, Z);
    Tensor BT = matmul_raw(c, B, false, false, t, false, false);
    k::axpy(c, sdt, r.numel(), -1.0, BT.data(), r.data());
    return r;
  };
  auto apply = [&](const Tensor& Xi, const Tensor& Yi, Tensor& AX, Tensor& AY) {
    Tensor Yp = to_dtype(c, proj_out(Vb, Yi), dt);
    Tensor MY = proj_out(Ub, to_dtype(c, matmul_raw(c, R1, false, false, matmul_raw(c, R2, false, false, Yp, false,
                                                                                      false), false, false), sdt));
    Tensor Xp = to_dtype(c, proj_out(Ub, Xi), dt);
    Tensor MX = proj_out(Vb, to_dtype(c, matmul_raw(c, R2, true, true, matmul_raw(c, R1, true, true, Xp, false,
                                                                                     false), false, false), sdt));
    AX = c.empty(sdt, {m, n});
    AY = c.empty(sdt, {nc, n});
    k::scale_cols(c, sdt, m, n, Xi.data(), Sn.ptr<double>(), AX.data(), false);
    k::scale_cols(c, sdt, nc, n, Yi.data(), Sn.ptr<double>(), AY.data(), false);
    k::axpy(c, sdt, AX.numel(), -1.0, MY.data(), AX.data());
    k::axpy(c, sdt, AY.numel(), -1.0, MX.data(), AY.data());
  };
  auto coldot = [&](const Tensor& P1, const Tensor& Q1, const Tensor& P2, const Tensor& Q2, Tensor& out) {
    Tensor t = c.empty(DType::F64, {n});
    k::col_re_dot(c, sdt, m, n, P1.data(), Q1.data(), out.ptr<double>(), false);
    k::col_re_dot(c, sdt, nc, n, P2.data(), Q2.data(), t.ptr<double>(), false);
    k::axpy(c, DType::F64, n, 1.0, t.data(), out.data());
  };
  Tensor X0 = clone(c, X), Y0 = clone(c, Y);
  Tensor AX, AY;
  apply(X, Y, AX, AY);
  Tensor Rx = clone(c, a), Ry = clone(c, b);
  k::axpy(c, sdt, Rx.numel(), -1.0, AX.data(), Rx.data());
  k::axpy(c, sdt, Ry.numel(), -1.0, AY.data(), Ry.data());
  Tensor Px = clone(c, Rx), Py = clone(c, Ry);
  Tensor rr = c.empty(DType::F64, {n}), rr_new = c.empty(DType::F64, {n}), pap = c.empty(DType::F64, {n});
  Tensor bb = c.empty(DType::F64, {n});
  coldot(Rx, Rx, Ry, Ry, rr);
  coldot(a, a, b, b, bb);
  std::vector<double> hb(n), hr(n), hp(n);
  HIP_CHECK(hipMemcpyAsync(hb.data(), bb.data(), n * 8, hipMemcpyDeviceToHost, c.stream));
  int it = 0;
  double rel = 0.0;
  bool bad = false;
  const int check_every = 4;
  while (it < iters) {
    Tensor APx, APy;
    apply(Px, Py, APx, APy);
    coldot(Px, APx, Py, APy, pap);
    k::col_axpy_ratio(c, sdt, m, n, rr.ptr<double>(), pap.ptr<double>(), 1.0, Px.data(), X.data());
    k::col_axpy_ratio(c, sdt, nc, n, rr.ptr<double>(), pap.ptr<double>(), 1.0, Py.data(), Y.data());
    k::col_axpy_ratio(c, sdt, m, n, rr.ptr<double>(), pap.ptr<double>(), -1.0, APx.data(), Rx.data());
    k::col_axpy_ratio(c, sdt, nc, n, rr.ptr<double>(), pap.ptr<double>(), -1.0, APy.data(), Ry.data());
    coldot(Rx, Rx, Ry, Ry, rr_new);
    k::col_xpby_ratio(c, sdt, m, n, rr_new.ptr<double>(), rr.ptr<double>(), Rx.data(), Px.data());
    k::col_xpby_ratio(c, sdt, nc, n, rr_new.ptr<double>(), rr.ptr<double>(), Ry.data(), Py.data());
    std::swap(rr, rr_new);
    ++it;
    if (it % check_every == 0 || it == iters) {
      HIP_CHECK(hipMemcpyAsync(hr.data(), rr.data(), n * 8, hipMemcpyDeviceToHost, c.stream));
      HIP_CHECK(hipMemcpyAsync(hp.data(), pap.data(), n * 8, hipMemcpyDeviceToHost, c.stream));
      c.sync();
      rel = 0.0;
      for (int64_t j = 0; j < n; ++j) {
        if (!(hp[j] > 0.0) && hr[j] > 0.0) bad = true;  // non-positive curvature
        if (!std::isfinite(hr[j])) bad = true;
        if (hb[j] > 0.0) rel = std::max(rel, std::sqrt(std::max(hr[j], 0.0) / hb[j]));
      }
      if (bad || rel < tol) break;
    }
  }
  if (bad) {
    HIP_CHECK(hipMemcpyAsync(X.data(), X0.data(), X.bytes(), hipMemcpyDeviceToDevice, c.stream));
    HIP_CHECK(hipMemcpyAsync(Y.data(), Y0.data(), Y.bytes(), hipMemcpyDeviceToDevice, c.stream));
  }
  if (stats) {
    std::lock_guard<std::mutex> g(stats->mu);
    ++stats->cg_calls;
    stats->cg_iters += it;
    if (bad) ++stats->cg_fallbacks;
    else stats->cg_resid_max = std::max(stats->cg_resid_max, rel);
  }
}

uint64_t draw_seed(uint64_t base, uint64_t i) {
  uint64_t z = base + (i + 1) * 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

static DType solve_dt(DType dt) { return dt == DType::F32 ? DType::F64 : dt; }

SVDResult svd_trunc(Ctx& c, const Tensor& R1, const Tensor& R2, int64_t chi, const SvdOpts& o) {
  IPEPS_CHECK(R1.nd() == 2 && R2.nd() == 2 && R1.dims[1] == R2.dims[0], "svd_trunc: R1 R2 shape mismatch");
  const int64_t m = R1.dims[0], inner = R1.dims[1], nc = R2.dims[1];
  const int64_t kmin = std::min(m, nc);
  // new_chi = min(chi, size(R1, 2)) (src/ctm/ctmrg.jl:255)
  const int64_t chi_eff = std::min<int64_t>(chi, inner);
  const DType dt = R1.dt, sdt = solve_dt(dt);
  SVDResult out;
  const int64_t ntar = std::min(chi_eff, kmin);
  const bool full = (o.mode == 0) || (kmin <= o.full_below) || (kmin <= ntar + o.oversample);
  Tensor Uf, Sf, Vf;  // thin factors of M = R1 R2: U m x k, S k, V nc x k
  if (full) {
    // LinearAlgebra.svd(M) (thin) as svd1 does (src/utility/autodiff.jl:22-25)
    Tensor M = matmul_raw(c, R1, false, false, R2, false, false);
    gesvd(c, to_dtype(c, M, sdt), Uf, Sf, Vf, true);
  } else {
    const int64_t kk = std::min<int64_t>(kmin, ntar + o.oversample);
    const bool warm = o.warm_start && o.warm_in && o.warm_in->defined() && o.warm_in->dt == sdt &&
                      o.warm_in->nd() == 2 && o.warm_in->dims[0] == m && o.warm_in->dims[1] == kk;
    Tensor Om;
    if (!warm) {
      Om = c.empty(dt, {nc, kk});
      if (o.draws_dev) k::randn_drawn(c, dt, nc * kk, o.seed_base, o.draws_dev, o.draw_off, Om.data());
      else k::randn(c, dt, nc * kk, o.seed, Om.data());
    }
    auto apply_M = [&](const Tensor& X) {  // M X = R1 (R2 X)
      return matmul_raw(c, R1, false, false, matmul_raw(c, R2, false, false, X, false, false), false, false);
    };
    auto apply_MH = [&](const Tensor& X) {  // M^H X = R2^H (R1^H X)
      return matmul_raw(c, R2, true, true, matmul_raw(c, R1, true, true, X, false, false), false, false);
    };
    // Applications 0 .. 2 q of the range finder (before B = M^H Q): the first
    // 2 q + 1 - late_apps with early_mixed products on FP32 engines.
    int app = 0;
    const int n_apps = 2 * o.power_iters + 1;
    auto app_prec = [&]() {
      const bool early = dt == DType::F32 && o.early_mixed != 0 && app < n_apps - o.late_apps;
      ++app;
      return early ? (k::Mixed)o.early_mixed : prec().mixed;
    };
    auto range_M = [&](const Tensor& X) {
      PrecScope ps(app_prec(), prec().back);
      return apply_M(X);
    };
    auto range_MH = [&](const Tensor& X) {
      PrecScope ps(app_prec(), prec().back);
      return apply_MH(X);
    };
    const int fo = o.fast_orth;
    OrthStatsScope oss(o.stats.get());
    // FP64 bases of at most 320 columns: the power iterations' bases by one
    // CholeskyQR pass each with no host round trip, their Cholesky infos
    // checked once at the end; the final basis by CholeskyQR2 with its checks.
    // If any pass failed the range finder is run again with checked bases.
    const bool lazy = sdt == DType::F64 && fo == 2 && kk <= 2 * k::SCHOL_MAXN_HOST && o.power_iters > 0;
    Tensor Q;
    int q = 0;
    bool lazy_ok = false;
    if (lazy) {
      Tensor infos = c.zeros(DType::F64, {2 * (2 * o.power_iters + 1)});
      int slot = 0;
      auto info = [&]() { return (int*)(infos.ptr<double>() + 2 * (slot++)); };
      Q = warm ? *o.warm_in : orth_cholqr1_nosync(c, to_dtype(c, range_M(Om), sdt), info());
      if (warm) ++app;
      for (; q < o.power_iters; ++q) {
        Tensor Z = orth_cholqr1_nosync(c, to_dtype(c, range_MH(to_dtype(c, Q, dt)), sdt), info());
        Q = orth_cholqr1_nosync(c, to_dtype(c, range_M(to_dtype(c, Z, dt)), sdt), info());
      }
      lazy_ok = true;
      if (c.graph_safe) {
        record_check(c, CHECK_INFO_PAIRS, infos.data(), infos.numel());
      } else {
        std::vector<double> hi(infos.numel());
        HIP_CHECK(hipMemcpyAsync(hi.data(), infos.data(), infos.bytes(), hipMemcpyDeviceToHost, c.stream));
        c.sync();
        for (double v : hi) {
          int a, b;
          std::memcpy(&a, &v, sizeof(int));
          std::memcpy(&b, (const char*)&v + sizeof(int), sizeof(int));
          lazy_ok = lazy_ok && a == 0 && b == 0;
        }
      }
      if (lazy_ok) {
        orth_count(&SvdStats::orth_calls);
        Q = orth(c, Q, fo);  // re-orthonormalise the last basis (checked)
      } else {
        orth_count(&SvdStats::orth_householder);
        q = 0;
      }
    }
    if (!lazy_ok) {
      app = 0;
      Q = warm ? orth(c, *o.warm_in, fo) : orth(c, to_dtype(c, range_M(Om), sdt), fo);
      if (warm) ++app;
      for (; q < o.power_iters; ++q) {
        Tensor Z = orth(c, to_dtype(c, range_MH(to_dtype(c, Q, dt)), sdt), fo);
        Q = orth(c, to_dtype(c, range_M(to_dtype(c, Z, dt)), sdt), fo);
      }
    }
    const int qmax = std::max(o.power_iters, o.power_iters_max);
    IPEPS_CHECK(!c.graph_safe || o.resid_tol <= 0.0 || qmax <= o.power_iters,
                "graph-safe projector: no adaptive power iterations");
    IPEPS_CHECK(!c.graph_safe || !warm, "graph-safe projector: no warm start");
    double resid = -1.0;
    for (;;) {
      Tensor Bt = to_dtype(c, apply_MH(to_dtype(c, Q, dt)), sdt);  // nc x kk = M^H Q
      Tensor W, s, X;
      // Bt = W diag(s) X^H  ->  M ~ (Q X) diag(s) W^H
      // FP32 data carry ~1e-7 relative rounding in M, far above the Gram
      // route's u (s_1 / s_i)^2 for every triplet above the mixed cutoff: the
      // guard applies to FP64 engines only in graph-safe mode
      const double guard = (c.graph_safe && dt == DType::F32) ? 0.0 : o.gram_guard;
      // graph-safe steps take the tridiagonal route with its check deferred
      // (fast_svd = 2 keeps the one-workgroup Jacobi eigensolver)
      svd_tall(c, Bt, c.graph_safe ? (o.fast_svd == 2 ? 2 : 1) : o.fast_svd, guard, W, s, X);
      Uf = matmul_raw(c, Q, false, false, X, false, false);
      Sf = s;
      Vf = W;
      if (o.resid_tol <= 0.0 || q >= qmax) break;
      // a-posteriori check of the triplets the cutoff keeps: M^H u_i = s_i v_i
      // by construction, so the residual is r_i = M v_i - s_i u_i
      std::vector<double> sh(kk);
      HIP_CHECK(hipMemcpyAsync(sh.data(), s.data(), kk * 8, hipMemcpyDeviceToHost, c.stream));
      c.sync();
      const double rel0 = dt == DType::F32 ? std::max(1e-12, o.mixed_cutoff) : 1e-12;
      const int64_t nk = std::max<int64_t>(1, std::min<int64_t>(cutoff_rule(sh, chi_eff, rel0), kk));
      Tensor Rm = to_dtype(c, apply_M(to_dtype(c, first_cols(c, Vf, nk), dt)), sdt);  // m x nk
      Tensor US = c.empty(sdt, {m, nk});
      k::scale_cols(c, sdt, m, nk, Uf.data(), s.ptr<double>(), US.data(), false);
      k::axpy(c, sdt, m * nk, -1.0, US.data(), Rm.data());
      Tensor rr = c.empty(DType::F64, {nk});
      k::col_re_dot(c, sdt, m, nk, Rm.data(), Rm.data(), rr.ptr<double>(), false);
      std::vector<double> rh(nk);
      HIP_CHECK(hipMemcpyAsync(rh.data(), rr.data(), nk * 8, hipMemcpyDeviceToHost, c.stream));
      c.sync();
      double rmax = 0.0;
      for (double v : rh) rmax = std::max(rmax, std::sqrt(std::max(v, 0.0)));
      resid = sh[0] > 0.0 ? rmax / sh[0] : 0.0;
      if (o.stats) {
        std::lock_guard<std::mutex> g(o.stats->mu);
        ++o.stats->checks;
      }
      if (!(resid > o.resid_tol)) break;
      // one more power iteration from Bt = M^H Q
      Tensor Z = orth(c, Bt, fo);
      Q = orth(c, to_dtype(c, apply_M(to_dtype(c, Z, dt)), sdt), fo);
      ++q;
    }
    if (o.stats) {
      std::lock_guard<std::mutex> g(o.stats->mu);
      ++o.stats->calls;
      o.stats->power_iters += q;
      o.stats->power_max = std::max<long>(o.stats->power_max, q);
      if (resid >= 0.0) o.stats->resid_max = std::max(o.stats->resid_max, resid);
      if (warm) ++o.stats->warm_starts;
    }
    if (o.warm_start && o.warm_out) *o.warm_out = Uf;
  }
  const int64_t ns = Sf.numel();
  // FP32 engines keep only triplets above their noise floor (s / s_max >
  // mixed_cutoff): P = R2 V S^-1/2 amplifies the rounding error of smaller ones.
  const double rel = dt == DType::F32 ? std::max(1e-12, o.mixed_cutoff) : 1e-12;
  int64_t n;
  if (c.graph_safe) {
    // the rank of the last eager step, else the largest the rule admits
    // without its degeneracy step; the rule itself is replayed on the
    // spectrum after the step
    IPEPS_CHECK(!full, "graph-safe projector: randomized mode only");
    n = std::min<int64_t>(o.fixed_n > 0 ? o.fixed_n : chi_eff, ns);
    record_check(c, CHECK_CUTOFF, Sf.data(), ns, chi_eff, n, rel);
    out.s_host.assign(ns, 0.0);
  } else {
    out.s_host.resize(ns);
    HIP_CHECK(hipMemcpyAsync(out.s_host.data(), Sf.data(), ns * 8, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    n = cutoff_rule(out.s_host, chi_eff, rel);
  }
  IPEPS_CHECK(n >= 1 && n <= ns, "svd_trunc: empty truncation");
  out.n = n;
  Tensor Un = first_cols(c, Uf, n), Vn = first_cols(c, Vf, n), Sn = first_n(c, Sf, n);
  out.U = to_dtype(c, Un, dt);
  out.V = to_dtype(c, Vn, dt);
  out.S = Sn;
  if (needs(R1) || needs(R2)) {
    auto base = make_node(nullptr);
    base->slots.resize(3);
    std::weak_ptr<Node> wb = base;
    auto mk_slot = [&](int i, Tensor& t) {
      t.node = make_node([wb, i](Ctx& c, Node& nd) {
        auto b = wb.lock();
        if (b) b->slots[i] = nd.grad;
      });
    };
    auto n1 = R1.node, n2 = R2.node;
    Tensor R1k = R1, R2k = R2;
    R1k.node.reset();
    R2k.node.reset();
    const double eta = o.eta;
    // The adjoint works with every computed triplet: all min(m, nc) thin
    // triplets in full mode (the reference svd_back on svd1's output), the k
    // oversampled ones in randomized mode, where the couplings F_ij between
    // kept and computed-but-discarded triplets are exact and only triplets
    // beyond k are treated as zero.
    Tensor Ub = Uf, Vb = Vf, Sb = Sf;
    const k::Mixed mx = prec().back;
    const int cg_iters = full ? 0 : o.complement_iters;
    const double cg_tol = o.complement_tol;
    const std::shared_ptr<SvdStats> stats = o.stats;
    base->back = [=](Ctx& c, Node& b) {
      PrecScope ps(mx);
      // svd_back (src/utility/autodiff.jl:41-75) for U (m x kb), V (nc x kb):
      //   Mbar = U G V^H + (dU - U U^H dU) Sinv V^H + U Sinv (dV - V V^H dV)^H
      // with the cotangents dU, dV (m x n, nc x n) zero beyond column n.
      const int64_t kb = Sb.numel();
      auto priv = [&](int slot, const std::vector<int64_t>& dims) {
        if (!b.slots[slot].defined()) return c.zeros(sdt, dims);
        Tensor t = to_dtype(c, b.slots[slot], sdt);
        return t.buf == b.slots[slot].buf ? clone(c, t) : t;
      };
      Tensor dU = priv(0, {m, n}), dV = priv(2, {nc, n});
      Tensor UdU = matmul_raw(c, Ub, true, true, dU, false, false);  // kb x n
      Tensor VdV = matmul_raw(c, Vb, true, true, dV, false, false);
      Tensor G = c.empty(sdt, {kb, kb});
      k::svd_gamma(c, sdt, kb, n, Sb.ptr<double>(), UdU.data(), VdV.data(),
                   b.slots[1].defined() ? b.slots[1].ptr<double>() : nullptr, eta, G.data());
      Tensor Sinv = c.empty(DType::F64, {n});
      k::svd_sinv(c, n, Sb.ptr<double>(), eta, Sinv.ptr<double>());
      Tensor Un_ = first_cols(c, Ub, n), Vn_ = first_cols(c, Vb, n);
      // X1 = U G V^H-part left factor and the non-square corrections
      Tensor X1 = matmul_raw(c, Ub, false, false, G, false, false);  // m x kb
      Tensor X2;                                                      // nc x n
      const bool sq_u = (m == kb), sq_v = (nc == kb);
      if (!sq_u) {
        Tensor T1 = matmul_raw(c, Ub, false, false, UdU, false, false);  // m x n
        k::axpy(c, sdt, T1.numel(), -1.0, T1.data(), dU.data());
        k::scale_cols(c, sdt, m, n, dU.data(), Sinv.ptr<double>(), dU.data(), false);
      }
      if (!sq_v) {
        Tensor T2 = matmul_raw(c, Vb, false, false, VdV, false, false);
        k::axpy(c, sdt, T2.numel(), -1.0, T2.data(), dV.data());
        X2 = c.empty(sdt, {nc, n});
        k::scale_cols(c, sdt, nc, n, dV.data(), Sinv.ptr<double>(), X2.data(), false);
      }
      if (!sq_u && !sq_v && cg_iters > 0) {
        // the couplings to the triplets beyond the computed k, exactly:
        // dU' and X2 (their s_j -> 0 limit) start the solve and are replaced
        // by its solution; dU still holds (I - Ub Ub^H) dU_orig Sinv
        Tensor a = c.empty(sdt, {m, n}), bb = c.empty(sdt, {nc, n});
        Tensor Sn = first_n(c, Sb, n);
        k::scale_cols(c, sdt, m, n, dU.data(), Sn.ptr<double>(), a.data(), false);
        k::scale_cols(c, sdt, nc, n, X2.data(), Sn.ptr<double>(), bb.data(), false);
        complement_cg(c, dt, sdt, R1k, R2k, Ub, Vb, Sn, a, bb, dU, X2, cg_iters, cg_tol, stats);
      }
      // Mbar = X1 Vb^H + [dU' Vn^H] + [Un X2^H] = L F^H with the factors side by
      // side, L = [X1 | dU' | Un] (m x K3), F = [Vb | Vn | X2] (nc x K3); applied
      // to R1 / R2 without forming Mbar: R1bar = L (R2 F)^H, R2bar = (R1^H L) F^H
      std::vector<Tensor> lp{X1}, fp{Vb};
      if (!sq_u) {
        lp.push_back(dU);
        fp.push_back(Vn_);
      }
      if (!sq_v) {
        lp.push_back(Un_);
        fp.push_back(X2);
      }
      auto side_by_side = [&](const std::vector<Tensor>& parts, int64_t rows) {
        int64_t cols = 0;
        for (auto& t : parts) cols += t.dims[1];
        Tensor r = c.empty(dt, {rows, cols});
        size_t off = 0;
        for (auto& t : parts) {
          Tensor td = to_dtype(c, t, dt);
          HIP_CHECK(hipMemcpyAsync((char*)r.data() + off, td.data(), td.bytes(), hipMemcpyDeviceToDevice, c.stream));
          off += td.bytes();
        }
        return r;
      };
      Tensor L = side_by_side(lp, m), F = side_by_side(fp, nc);
      if (n1) acc(c, n1, matmul_raw(c, L, false, false, matmul_raw(c, R2k, false, false, F, false, false), true, true));
      if (n2) acc(c, n2, matmul_raw(c, matmul_raw(c, R1k, true, true, L, false, false), false, false, F, true, true));
    };
    mk_slot(0, out.U);
    mk_slot(1, out.S);
    mk_slot(2, out.V);
  }
  return out;
}

}  // namespace op
}  // namespace ipeps
