// CTMRG on the device: the reference's contraction networks written with the
// engine's differentiable tcon (label lists follow src/ctm/ctmrg.jl and
// src/graph/evaluation.jl), plus the step-checkpointed gradient driver.
#include "ctmrg.h"

#include <cmath>
#include <cstring>
#include <exception>
#include <thread>

namespace ipeps {
namespace ctm {

using op::tcon;
using L = std::vector<int>;

namespace {

Tensor mat(Ctx& c, const Tensor& x, int rows_nd) {
  int64_t r = 1, q = 1;
  for (int i = 0; i < x.nd(); ++i) (i < rows_nd ? r : q) *= x.dims[i];
  return op::reshape(c, x, {r, q});
}

Tensor tr2(Ctx& c, const Tensor& x) { return op::permute(c, x, {1, 0}); }

struct Proj {
  Tensor P1, P2;
  std::vector<double> s;
};

// get_projector (src/ctm/ctmrg.jl:254-265): M = R1 R2, truncated SVD,
// P1 = R2 V S^-1/2, P2 = S^-1/2 U^H R1.
// `draw` is the position of this projector's test matrix in the environment's
// stream: e.draws + direction (rg_step advances e.draws by 4 per step).
Proj get_projector(Ctx& c, const Env& e, const Tensor& R1, const Tensor& R2, int64_t chi, uint64_t draw) {
  op::SvdOpts so = e.opt.svd;
  so.seed = op::draw_seed(e.opt.svd.seed, draw);
  const uint64_t dir = draw - e.draws;
  if (so.warm_start && dir < 4) {
    so.warm_in = &e.warm[dir];
    so.warm_out = &e.warm[dir];
  }
  if (c.graph_safe && e.draws_dev.defined()) {
    // captured: the seed comes from the device counter at replay time
    so.draws_dev = (const uint64_t*)e.draws_dev.data();
    so.seed_base = e.opt.svd.seed;
    so.draw_off = draw - e.draws;
  }
  if (c.graph_safe && dir < 4) so.fixed_n = e.kept[dir];
  op::PrecScope ps(e.opt.mixed_proj ? (k::Mixed)e.opt.mixed_proj : op::prec().mixed, op::prec().back);
  // R1 and R2 enter every product of the range finder and P1 / P2: their
  // bf16 planes are split once
  SplitCacheScope scs(c, {&R1, &R2});
  auto r = op::svd_trunc(c, R1, R2, chi, so);
  if (dir < 4) e.kept[dir] = r.n;
  Tensor isq = op::inv_sqrt(c, r.S);
  Proj p;
  p.P1 = op::scale_vec(c, op::matmul(c, R2, r.V), isq, false);
  p.P2 = op::scale_vec(c, op::adjoint_matmul(c, r.U, R1), isq, true);
  p.s.assign(r.s_host.begin(), r.s_host.begin() + r.n);
  return p;
}

// Edge renormalisation.  ia / iad are the reference label lists of A and Ad
// in E*A and (E*A)*Ad (src/ctm/ctmrg.jl:285-289 for the left move).
Tensor edge_update(Ctx& c, const Env& e, const Tensor& E, const L& ia, const L& iad, const Tensor& P, const Tensor& Pd) {
  const Tensor& A = e.A;
  const Tensor& Ad = e.Ad;
  if (!e.opt.canonical_edge) {
    // reference order: chi^2 D^6 intermediate, then Pd and P
    Tensor EA = tcon(c, E, {-1, -3, 1, -6}, A, ia);
    Tensor EAAd = tcon(c, EA, {-1, -2, -4, -5, -7, 1, 2}, Ad, iad);
    auto s = EAAd.dims;
    EAAd = op::reshape(c, EAAd, {s[0] * s[1] * s[2], s[3] * s[4] * s[5], s[6], s[7]});
    Tensor EPd = tcon(c, Pd, {-1, 1}, EAAd, {1, -2, -3, -4});
    return tcon(c, EPd, {-1, 1, -3, -4}, P, {1, -2});
  }
  // canonical order: Pd first (chi^3 D^4 + 2 chi^2 D^6 d + chi^3 D^4)
  auto pos = [](const L& l, int lab) {
    for (size_t i = 0; i < l.size(); ++i)
      if (l[i] == lab) return (int)i;
    throw Error(IPEPS_ERR_INTERNAL, "edge_update: label missing");
  };
  const int a_p2 = pos(ia, -2), a_E = pos(ia, 1), a_p1 = pos(ia, -4), a_open = pos(ia, -5), a_ph = pos(ia, -7);
  const int d_p2 = pos(iad, -3), d_E = pos(iad, 1), d_p1 = pos(iad, -6), d_open = pos(iad, -8), d_ph = pos(iad, 2);
  const int64_t chi_e = E.dims[0], D = A.dims[0];
  const int64_t chip = Pd.dims[0], chin = P.dims[1];
  Tensor P2r = op::reshape(c, Pd, {chip, chi_e, D, D});
  Tensor X1 = tcon(c, P2r, {-1, 1, -2, -3}, E, {1, -4, -5, -6});  // (i, ap2, adp2, E2, Eket, Ebra)
  // A and Ad absorbed over (ap2, Eket) / (adp2, Ebra) and the physical leg:
  // one site absorption (fused kernel on the mixed engine), its output with
  // the legs P1 contracts last, so that X3 enters the last product as a
  // plain M x K matrix
  L la(5), ld(5);
  la[a_p2] = 10; la[a_E] = 11; la[a_p1] = -5; la[a_open] = -2; la[a_ph] = 3;
  ld[d_p2] = 20; ld[d_E] = 21; ld[d_ph] = 3; ld[d_p1] = -6; ld[d_open] = -3;
  Tensor X3 = op::absorb(c, X1, {-1, 10, 20, -4, 11, 21}, A, la, Ad, ld);  // (i, Aopen, Adopen, E2, Ap1, Adp1)
  Tensor P1r = op::reshape(c, P, {E.dims[1], D, D, chin});
  return tcon(c, X3, {-1, -3, -4, 30, 31, 32}, P1r, {30, 31, 32, -2});
}

void finish(Ctx& c, Env& e, int ci, const Tensor& C_, int cj, const Tensor& Cb, int ek, const Tensor& E_) {
  e.C[ci] = op::renormalize(c, C_);
  e.E[ek] = op::renormalize(c, E_);
  e.C[cj] = op::renormalize(c, Cb);
}

std::vector<double> left_rg(Ctx& c, Env& e, int64_t chi, uint64_t draw, Proj* only = nullptr) {
  const Tensor &C1 = e.C[0], &C4 = e.C[3], &E1 = e.E[0], &E3 = e.E[2], &E4 = e.E[3], &A = e.A, &Ad = e.Ad;
  // get_projector_left (src/ctm/ctmrg.jl:137-160)
  // The (C E) E products come out in the GEMM's own leg order (no permute)
  // with the row leg that is stride 1 in the absorbed corner first; the
  // absorption's output order makes R1 without a transpose.
  Tensor C1E1 = tcon(c, C1, {-1, 1}, E1, {1, -2, -3, -4});                  // (c1d, e1r, e1k, e1b)
  Tensor CEE4 = tcon(c, C1E1, {1, -1, -2, -3}, E4, {1, -4, -5, -6});        // (e1r, e1k, e1b, e4d, e4k, e4b)
  Tensor LU = op::absorb(c, CEE4, {-1, 21, 11, -4, 22, 12}, A, {21, 22, -5, -2, 3}, Ad, {11, 12, -6, -3, 3});
  Tensor R1 = mat(c, LU, 3);
  Tensor C4E3 = tcon(c, C4, {-1, 1}, E3, {1, -2, -3, -4});                  // (c4u, e3r, e3k, e3b)
  Tensor CEE4b = tcon(c, C4E3, {1, -4, -5, -6}, E4, {-1, 1, -2, -3});       // (e4u, e4k, e4b, e3r, e3k, e3b)
  Tensor LB = op::absorb(c, CEE4b, {-1, 21, 11, -4, 22, 12}, A, {-2, 21, 22, -5, 3}, Ad, {-3, 11, 12, -6, 3});
  Tensor R2 = mat(c, LB, 3);
  Proj p = get_projector(c, e, R1, R2, chi, draw);
  if (only) {
    *only = p;
    return p.s;
  }
  // proj_left (src/ctm/ctmrg.jl:267-291)
  Tensor C1E1p = tcon(c, C1, {-2, 1}, E1, {1, -1, -3, -4});
  Tensor nC1 = tr2(c, op::matmul(c, op::reshape(c, C1E1p, {C1E1p.dims[0], C1E1p.numel() / C1E1p.dims[0]}), p.P1));
  Tensor C4E3p = tcon(c, C4, {-1, 1}, E3, {1, -4, -2, -3});
  int64_t lastd = C4E3p.dims[3];
  Tensor nC4 = op::matmul(c, p.P2, op::reshape(c, C4E3p, {C4E3p.numel() / lastd, lastd}));
  Tensor nE4 = edge_update(c, e, E4, {-2, 1, -4, -5, -7}, {-3, 1, -6, -8, 2}, p.P1, p.P2);
  finish(c, e, 0, nC1, 3, nC4, 3, nE4);
  return p.s;
}

void right_rg(Ctx& c, Env& e, int64_t chi, uint64_t draw, Proj* only = nullptr) {
  const Tensor &C2 = e.C[1], &C3 = e.C[2], &E1 = e.E[0], &E2 = e.E[1], &E3 = e.E[2], &A = e.A, &Ad = e.Ad;
  // get_projector_right (src/ctm/ctmrg.jl:176-198)
  Tensor C2E1 = tcon(c, C2, {1, -1}, E1, {-2, 1, -3, -4});                  // (c2d, e1l, e1k, e1b)
  Tensor CEE2 = tcon(c, C2E1, {1, -1, -2, -3}, E2, {1, -4, -5, -6});        // (e1l, e1k, e1b, e2d, e2k, e2b)
  Tensor RU = op::absorb(c, CEE2, {-1, 21, 11, -4, 22, 12}, A, {21, -2, -5, 22, 3}, Ad, {11, -3, -6, 12, 3});
  Tensor R1 = mat(c, RU, 3);
  Tensor C3E3 = tcon(c, C3, {-1, 1}, E3, {-2, 1, -3, -4});                  // (c3u, e3l, e3k, e3b)
  Tensor CEE2b = tcon(c, C3E3, {1, -4, -5, -6}, E2, {-1, 1, -2, -3});       // (e2u, e2k, e2b, e3l, e3k, e3b)
  Tensor RB = op::absorb(c, CEE2b, {-1, 22, 12, -4, 21, 11}, A, {-2, -5, 21, 22, 3}, Ad, {-3, -6, 11, 12, 3});
  Tensor R2 = mat(c, RB, 3);
  Proj p = get_projector(c, e, R1, R2, chi, draw);
  if (only) {
    *only = p;
    return;
  }
  // proj_right (src/ctm/ctmrg.jl:293-318)
  Tensor C2E1p = tcon(c, C2, {1, -2}, E1, {-1, 1, -3, -4});
  Tensor nC2 = op::matmul(c, op::reshape(c, C2E1p, {C2E1p.dims[0], C2E1p.numel() / C2E1p.dims[0]}), p.P1);
  Tensor C3E3p = tcon(c, C3, {-1, 1}, E3, {-4, 1, -2, -3});
  int64_t lastd = C3E3p.dims[3];
  Tensor nC3 = op::matmul(c, p.P2, op::reshape(c, C3E3p, {C3E3p.numel() / lastd, lastd}));
  Tensor nE2 = edge_update(c, e, E2, {-2, -5, -4, 1, -7}, {-3, -8, -6, 1, 2}, p.P1, p.P2);
  finish(c, e, 1, nC2, 2, nC3, 1, nE2);
}

void top_rg(Ctx& c, Env& e, int64_t chi, uint64_t draw, Proj* only = nullptr) {
  const Tensor &C1 = e.C[0], &C2 = e.C[1], &E1 = e.E[0], &E2 = e.E[1], &E4 = e.E[3], &A = e.A, &Ad = e.Ad;
  // get_projector_top (src/ctm/ctmrg.jl:214-236)
  Tensor C1E4 = tcon(c, C1, {1, -1}, E4, {1, -2, -3, -4});                  // (c1r, e4d, e4k, e4b)
  Tensor CEE1 = tcon(c, C1E4, {1, -1, -2, -3}, E1, {1, -4, -5, -6});        // (e4d, e4k, e4b, e1r, e1k, e1b)
  Tensor UL = op::absorb(c, CEE1, {-1, 22, 12, -4, 21, 11}, A, {21, 22, -2, -5, 3}, Ad, {11, 12, -3, -6, 3});
  Tensor R1 = mat(c, UL, 3);
  Tensor C2E2 = tcon(c, C2, {-1, 1}, E2, {1, -2, -3, -4});                  // (c2l, e2d, e2k, e2b)
  Tensor CEE1b = tcon(c, C2E2, {1, -4, -5, -6}, E1, {-1, 1, -2, -3});       // (e1l, e1k, e1b, e2d, e2k, e2b)
  Tensor UR = op::absorb(c, CEE1b, {-1, 21, 11, -4, 22, 12}, A, {21, -2, -5, 22, 3}, Ad, {11, -3, -6, 12, 3});
  Tensor R2 = mat(c, UR, 3);
  Proj p = get_projector(c, e, R1, R2, chi, draw);
  if (only) {
    *only = p;
    return;
  }
  // proj_top (src/ctm/ctmrg.jl:320-345)
  Tensor C1E4p = tcon(c, C1, {1, -2}, E4, {1, -1, -3, -4});
  Tensor nC1 = op::matmul(c, op::reshape(c, C1E4p, {C1E4p.dims[0], C1E4p.numel() / C1E4p.dims[0]}), p.P1);
  Tensor C2E2p = tcon(c, C2, {-1, 1}, E2, {1, -4, -2, -3});
  int64_t lastd = C2E2p.dims[3];
  Tensor nC2 = op::matmul(c, p.P2, op::reshape(c, C2E2p, {C2E2p.numel() / lastd, lastd}));
  Tensor nE1 = edge_update(c, e, E1, {1, -2, -5, -4, -7}, {1, -3, -8, -6, 2}, p.P1, p.P2);
  finish(c, e, 0, nC1, 1, nC2, 0, nE1);
}

void bottom_rg(Ctx& c, Env& e, int64_t chi, uint64_t draw, Proj* only = nullptr) {
  const Tensor &C3 = e.C[2], &C4 = e.C[3], &E2 = e.E[1], &E3 = e.E[2], &E4 = e.E[3], &A = e.A, &Ad = e.Ad;
  // get_projector_bottom (src/ctm/ctmrg.jl:238-252... via 226-250)
  Tensor C4E4 = tcon(c, C4, {1, -1}, E4, {-2, 1, -3, -4});                  // (c4r, e4u, e4k, e4b)
  Tensor CEE3 = tcon(c, C4E4, {1, -1, -2, -3}, E3, {1, -4, -5, -6});        // (e4u, e4k, e4b, e3r, e3k, e3b)
  Tensor BL = op::absorb(c, CEE3, {-1, 21, 11, -4, 22, 12}, A, {-2, 21, 22, -5, 3}, Ad, {-3, 11, 12, -6, 3});
  Tensor R1 = mat(c, BL, 3);
  Tensor C3E2 = tcon(c, C3, {1, -1}, E2, {-2, 1, -3, -4});                  // (c3l, e2u, e2k, e2b)
  Tensor CEE3b = tcon(c, C3E2, {1, -4, -5, -6}, E3, {-1, 1, -2, -3});       // (e3l, e3k, e3b, e2u, e2k, e2b)
  Tensor BR = op::absorb(c, CEE3b, {-1, 21, 11, -4, 22, 12}, A, {-5, -2, 21, 22, 3}, Ad, {-6, -3, 11, 12, 3});
  Tensor R2 = mat(c, BR, 3);
  Proj p = get_projector(c, e, R1, R2, chi, draw);
  if (only) {
    *only = p;
    return;
  }
  // proj_bottom (src/ctm/ctmrg.jl:347-375)
  Tensor C4E4p = tcon(c, C4, {1, -2}, E4, {-1, 1, -3, -4});
  Tensor nC4 = op::matmul(c, op::reshape(c, C4E4p, {C4E4p.dims[0], C4E4p.numel() / C4E4p.dims[0]}), p.P1);
  Tensor C3E2p = tcon(c, C3, {1, -1}, E2, {-4, 1, -2, -3});
  int64_t lastd = C3E2p.dims[3];
  Tensor nC3 = tr2(c, op::matmul(c, p.P2, op::reshape(c, C3E2p, {C3E2p.numel() / lastd, lastd})));
  Tensor nE3 = edge_update(c, e, E3, {-5, -2, 1, -4, -7}, {-8, -3, 1, -6, 2}, p.P1, p.P2);
  finish(c, e, 3, nC4, 2, nC3, 2, nE3);
}

// ---- density matrices (src/graph/evaluation.jl:124-248) ---------------------
// The reference builds rho_h from a left half (C1 E1 E4 C4 E3 A Ad) and a
// right half, and rho_v from an upper and a lower half, with the physical legs
// opened early (its dominant term is chi^3 D^5 d^2 per half).  Here the
// upper-left corner C1 E1 E4 A Ad and the lower-right corner C3 E3 E2 A Ad are
// built once, with both physical legs open (chi^2 D^4 d^2 elements each), and
// each corner is closed twice: UL with C4 E3 (left half of rho_h) and with
// C2 E2 (upper half of rho_v); LR with C2 E1 (right half of rho_h) and with
// C4 E4 (lower half of rho_v).  The halves come out in the reference's leg
// order, so the final rho contractions are the reference's
// (evaluation.jl:183,207,231,247); the dominant term is chi^3 D^4 d^2.
//
// Legs: C1 (down, right), C2 (left, down), C3 (up, left), C4 (up, right);
// E1/E3 (left, right, ket, bra), E2/E4 (up, down, ket, bra);
// A (up, left, down, right, phys).

// (e1r, e4d, A_d, A_r, p, Ad_d, Ad_r, p'): the last product's own output
// order (its first operand's free legs, then the second's), so no permute
Tensor corner_ul(Ctx& c, const Tensor& C1, const Tensor& E1, const Tensor& E4, const Tensor& A, const Tensor& Ad) {
  Tensor C1E1 = tcon(c, C1, {-1, 1}, E1, {1, -2, -3, -4});
  Tensor X1 = tcon(c, C1E1, {1, -1, -2, -3}, E4, {1, -4, -5, -6});
  Tensor X2 = tcon(c, X1, {-1, 1, -2, -3, 2, -4}, A, {1, 2, -5, -6, -7});
  return tcon(c, X2, {-1, 1, -2, 2, -3, -4, -5}, Ad, {1, 2, -6, -7, -8});
}

// (e3l, e2u, A_u, A_l, p, Ad_u, Ad_l, p'), likewise
Tensor corner_lr(Ctx& c, const Tensor& C3, const Tensor& E3, const Tensor& E2, const Tensor& A, const Tensor& Ad) {
  Tensor C3E3 = tcon(c, C3, {-1, 1}, E3, {-2, 1, -3, -4});
  Tensor Y1 = tcon(c, C3E3, {1, -1, -2, -3}, E2, {-4, 1, -5, -6});
  Tensor Y2 = tcon(c, Y1, {-1, 1, -2, -3, 2, -4}, A, {-5, -6, 1, 2, -7});
  return tcon(c, Y2, {-1, 1, -2, 2, -3, -4, -5}, Ad, {-6, -7, 1, 2, -8});
}

// rho from two halves (x1, x2, a, ad, p, p') sharing their first four legs.
// The trace over the chi^2 D^2 shared legs is a sum with heavy cancellation
// (the bond energy is a small difference of rho's entries): FP32 halves are
// cast to FP64 for it.
Tensor close_dm(Ctx& c, const Tensor& h1, const Tensor& h2) {
  const bool widen = h1.dt == DType::F32;
  const Tensor a = widen ? op::cast(c, h1, DType::F64) : h1;
  const Tensor b = widen ? op::cast(c, h2, DType::F64) : h2;
  Tensor r = tcon(c, a, {1, 2, 3, 4, -1, -3}, b, {1, 2, 3, 4, -2, -4});
  int64_t d = h1.dims[4];
  return op::reshape(c, r, {d * d, d * d});
}

}  // namespace

namespace {

// f on the context's stream and g on its second lane, concurrently (their
// inputs and outputs are disjoint parts of the environment).  The lane
// waits for everything the main stream enqueued before the fork; the main
// stream waits for the lane's work after the join.  Tensors the lane drops
// are returned to their pools only after that wait, and the lane's kernel
// timings are merged into the context's.
template <typename F, typename G>
void run_pair(Ctx& c, Env& e, F&& f, G&& g) {
  Ctx& l = c.second_lane();
  std::array<Tensor, 4> holdC = e.C, holdE = e.E;  // released after the join
  hipEvent_t fork, join;
  HIP_CHECK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
  HIP_CHECK(hipEventCreateWithFlags(&join, hipEventDisableTiming));
  HIP_CHECK(hipEventRecord(fork, c.stream));
  HIP_CHECK(hipStreamWaitEvent(l.stream, fork, 0));
  l.timing = c.timing;
  l.timing_detail = c.timing_detail;
  l.timing_filter = c.timing_filter;
  // The lane records into its own tape, appended after the main lane's nodes
  // at the join: the tape order is then the sequential schedule's, so the
  // reverse sweep accumulates shared cotangents in the same order.
  Tape* tape = current_tape();
  Tape lane_tape;
  lane_tape.active = tape && tape->active;
  const bool rec = tape && tape->active;
  // tape segments of the two lanes (the reverse sweep runs them concurrently)
  const int smain = rec ? tape->next_seg++ : 0, slane = rec ? tape->next_seg++ : 0;
  const size_t m0 = rec ? tape->order.size() : 0;
  const op::Prec prec = op::prec();
  std::exception_ptr lane_err;
  std::thread th([&] {
    try {
      HIP_CHECK(hipSetDevice(l.device));
      TapeScope ts(tape ? &lane_tape : nullptr);
      current_seg() = slane;
      op::prec() = prec;
      g(l);
      current_seg() = 0;
      HIP_CHECK(hipEventRecord(join, l.stream));
    } catch (...) {
      current_seg() = 0;
      lane_err = std::current_exception();
      (void)hipEventRecord(join, l.stream);
    }
  });
  std::exception_ptr main_err;
  const int seg0 = current_seg();
  try {
    current_seg() = smain;
    f(c);
  } catch (...) {
    main_err = std::current_exception();
  }
  current_seg() = seg0;
  th.join();
  if (tape) {
    const size_t m1 = tape->order.size();
    tape->order.insert(tape->order.end(), lane_tape.order.begin(), lane_tape.order.end());
    if (rec && !main_err && !lane_err) tape->pairs.push_back(Tape::Pair{m0, m1, tape->order.size(), smain, slane});
  }
  (void)hipStreamWaitEvent(c.stream, join, 0);
  (void)hipEventDestroy(fork);
  (void)hipEventDestroy(join);
  c.kev.insert(c.kev.end(), l.kev.begin(), l.kev.end());
  l.kev.clear();
  if (main_err) std::rethrow_exception(main_err);
  if (lane_err) std::rethrow_exception(lane_err);
}

}  // namespace

void get_dms(Ctx& c, const Env& e, Tensor& roh, Tensor& rov) {
  const auto& C = e.C;
  const auto& E = e.E;
  Tensor Lh, Uv, Rh, Bv;
  // the upper-left corner and the halves it closes
  auto ul_side = [&](Ctx& cc) {
    Tensor UL = corner_ul(cc, C[0], E[0], E[3], e.A, e.Ad);
    // left half of rho_h: UL closed with C4 E3 -> (e1r, e3r, A_r, Ad_r, p, p')
    Tensor C4E3 = tcon(cc, C[3], {-1, 1}, E[2], {1, -2, -3, -4});
    Lh = tcon(cc, UL, {-1, 1, 2, -3, -5, 3, -4, -6}, C4E3, {1, -2, 2, 3});
    // upper half of rho_v: UL closed with C2 E2 -> (e4d, e2d, A_d, Ad_d, p, p')
    Tensor C2E2 = tcon(cc, C[1], {-1, 1}, E[1], {1, -2, -3, -4});
    Uv = tcon(cc, UL, {1, -1, -3, 2, -5, -4, 3, -6}, C2E2, {1, -2, 2, 3});
  };
  // the lower-right corner and the halves it closes
  auto lr_side = [&](Ctx& cc) {
    Tensor LR = corner_lr(cc, C[2], E[2], E[1], e.A, e.Ad);
    // right half of rho_h: LR closed with C2 E1 -> (e1l, e3l, A_l, Ad_l, p, p')
    Tensor C2E1 = tcon(cc, C[1], {1, -2}, E[0], {-1, 1, -3, -4});
    Rh = tcon(cc, LR, {-2, 1, 2, -3, -5, 3, -4, -6}, C2E1, {-1, 1, 2, 3});
    // lower half of rho_v: LR closed with C4 E4 -> (e4u, e2u, A_u, Ad_u, p, p')
    Tensor C4E4 = tcon(cc, C[3], {1, -2}, E[3], {-1, 1, -3, -4});
    Bv = tcon(cc, LR, {1, -2, -3, 2, -5, -4, 3, -6}, C4E4, {-1, 1, 2, 3});
  };
  if (e.opt.paired_moves) {
    // the two sides share only their inputs: second lane for the lower right
    Env held = e;
    run_pair(c, held, ul_side, lr_side);
  } else {
    ul_side(c);
    lr_side(c);
  }
  roh = close_dm(c, Lh, Rh);
  rov = close_dm(c, Uv, Bv);
}

void projector(Ctx& c, const Env& e0, int dir, int64_t chi, Tensor& P1, Tensor& P2, std::vector<double>& s) {
  Env e = e0;  // shares the tensors; the *_rg functions stop after the projector
  Proj p;
  const uint64_t draw = e0.draws + (uint64_t)dir;
  switch (dir) {
    case 0: left_rg(c, e, chi, draw, &p); break;
    case 1: right_rg(c, e, chi, draw, &p); break;
    case 2: top_rg(c, e, chi, draw, &p); break;
    default: bottom_rg(c, e, chi, draw, &p); break;
  }
  P1 = p.P1;
  P2 = p.P2;
  s = p.s;
}

void init_ctm1(Ctx& c, Env& e) {
  const Tensor &A = e.A, &Ad = e.Ad;
  const int64_t D = A.dims[0];
  auto C = [&](const L& la, const L& lb) {
    return op::renormalize(c, op::reshape(c, tcon(c, A, la, Ad, lb), {D * D, D * D}));
  };
  auto E = [&](const L& la, const L& lb) {
    return op::renormalize(c, op::reshape(c, tcon(c, A, la, Ad, lb), {D * D, D * D, D, D}));
  };
  // src/ctm/ctm_tensor.jl:97-130
  e.C[0] = C({1, 2, -1, -3, 3}, {1, 2, -2, -4, 3});
  e.C[1] = C({1, -1, -3, 2, 3}, {1, -2, -4, 2, 3});
  e.C[2] = C({-1, -3, 1, 2, 3}, {-2, -4, 1, 2, 3});
  e.C[3] = C({-1, 1, 2, -3, 3}, {-2, 1, 2, -4, 3});
  e.E[0] = E({1, -1, -5, -3, 2}, {1, -2, -6, -4, 2});
  e.E[1] = E({-1, -5, -3, 1, 2}, {-2, -6, -4, 1, 2});
  e.E[2] = E({-5, -1, 1, -3, 2}, {-6, -2, 1, -4, 2});
  e.E[3] = E({-1, 1, -3, -5, 2}, {-2, 1, -4, -6, 2});
}


static std::vector<double> rg_step_impl(Ctx& c, Env& e, int64_t chi);

// Graph-safe mode on c for a scope (timing off: a capture records no events).
struct GraphSafeScope {
  Ctx& c;
  bool timing;
  explicit GraphSafeScope(Ctx& cc, int64_t cap) : c(cc), timing(cc.timing) {
    checks_begin(c, cap);
    c.timing = false;
    c.graph_safe = true;
  }
  ~GraphSafeScope() {
    c.graph_safe = false;
    c.timing = timing;
  }
};

constexpr int64_t STEP_CHECKS_CAP = 1 << 14;

std::vector<double> rg_step(Ctx& c, Env& e, int64_t chi) {
  if (!e.opt.graph_safe || c.graph_safe) return rg_step_impl(c, e, chi);
  IPEPS_CHECK(!e.opt.paired_moves, "graph_safe needs paired_moves=0");
  std::vector<double> s;
  std::vector<Ctx::CheckRec> recs;
  std::vector<double> h;
  {
    GraphSafeScope gs(c, STEP_CHECKS_CAP);
    s = rg_step_impl(c, e, chi);
    recs = c.check_recs;
    h.resize(c.checks_pos);
    if (!h.empty()) HIP_CHECK(hipMemcpyAsync(h.data(), c.checks.data(), h.size() * 8, hipMemcpyDeviceToHost, c.stream));
    c.sync();
  }
  std::string why;
  if (!validate_checks(recs, h.data(), &why)) throw Error(IPEPS_ERR_STATE, "graph-safe step: " + why);
  return s;
}

StepGraph::~StepGraph() {
  if (c) (void)hipStreamSynchronize(c->stream);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (c && !blocks.empty()) c->pool->release_graph(blocks);
  if (host_checks) (void)hipHostFree(host_checks);
}

std::unique_ptr<StepGraph> capture_step(Ctx& c, Env& e, int64_t chi) {
  IPEPS_CHECK(!e.opt.paired_moves, "step capture needs paired_moves=0");
  IPEPS_CHECK(e.opt.svd.mode == 1 && !e.opt.svd.warm_start, "step capture needs the randomized projector, no warm start");
  IPEPS_CHECK(!op::recording(), "step capture: no tape");
  {
    // the step once, eagerly, in graph-safe form on a copy of the handles:
    // a path that would synchronise the stream (the full SVD, host checks)
    // throws here rather than inside the capture, the shapes must come back
    // unchanged and the deferred checks must pass
    Env probe = e;
    std::vector<Ctx::CheckRec> recs;
    std::vector<double> h;
    {
      GraphSafeScope gs(c, STEP_CHECKS_CAP);
      rg_step_impl(c, probe, chi);
      for (int i = 0; i < 4; ++i)
        if (probe.C[i].dims != e.C[i].dims || probe.E[i].dims != e.E[i].dims)
          throw Error(IPEPS_ERR_STATE, "step capture: shapes not stable");
      recs = c.check_recs;
      h.resize(c.checks_pos);
      if (!h.empty()) HIP_CHECK(hipMemcpyAsync(h.data(), c.checks.data(), h.size() * 8, hipMemcpyDeviceToHost, c.stream));
      c.sync();
    }
    std::string why;
    if (!validate_checks(recs, h.data(), &why)) throw Error(IPEPS_ERR_STATE, "step capture: " + why);
  }
  std::unique_ptr<StepGraph> g(new StepGraph());
  g->c = &c;
  g->chi = chi;
  if (!e.draws_dev.defined()) e.draws_dev = c.empty(DType::F64, {1});
  HIP_CHECK(hipMemcpyAsync(e.draws_dev.data(), &e.draws, 8, hipMemcpyHostToDevice, c.stream));
  for (int i = 0; i < 4; ++i) {
    g->bakC[i] = c.empty(e.C[i].dt, e.C[i].dims);
    g->bakE[i] = c.empty(e.E[i].dt, e.E[i].dims);
  }
  c.sync();
  const std::array<Tensor, 4> inC = e.C, inE = e.E;
  g->C = inC;
  g->E = inE;
  const uint64_t d0 = e.draws;
  c.pool->begin_graph();
  bool capturing = false;
  try {
    GraphSafeScope gs(c, STEP_CHECKS_CAP);
    HIP_CHECK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeRelaxed));
    capturing = true;
    for (int i = 0; i < 4; ++i) {
      dev_copy(c, g->bakC[i].data(), inC[i].data(), inC[i].bytes());
      dev_copy(c, g->bakE[i].data(), inE[i].data(), inE[i].bytes());
    }
    rg_step_impl(c, e, chi);
    for (int i = 0; i < 4; ++i) {
      IPEPS_CHECK(e.C[i].dims == inC[i].dims && e.E[i].dims == inE[i].dims, "step capture: shapes not stable");
      dev_copy(c, inC[i].data(), e.C[i].data(), inC[i].bytes());
      dev_copy(c, inE[i].data(), e.E[i].data(), inE[i].bytes());
    }
    k::counter_add(c, (uint64_t*)e.draws_dev.data(), 4);
    e.C = inC;
    e.E = inE;
    e.draws = d0;
    HIP_CHECK(hipStreamEndCapture(c.stream, &g->graph));
    capturing = false;
    g->checks = c.checks;
    g->recs = c.check_recs;
    g->nchecks = c.checks_pos;
    c.checks = Tensor();
  } catch (...) {
    if (capturing) {
      hipGraph_t dead = nullptr;
      (void)hipStreamEndCapture(c.stream, &dead);
      if (dead) (void)hipGraphDestroy(dead);
      hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
      if (hipStreamIsCapturing(c.stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
        dead = nullptr;
        (void)hipStreamEndCapture(c.stream, &dead);
        if (dead) (void)hipGraphDestroy(dead);
      }
    }
    (void)hipGetLastError();
    e.C = inC;
    e.E = inE;
    e.draws = d0;
    c.pool->release_graph(c.pool->end_graph());
    throw;
  }
  g->blocks = c.pool->end_graph();
  HIP_CHECK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
  HIP_CHECK(hipHostMalloc((void**)&g->host_checks, std::max<int64_t>(1, g->nchecks) * 8, hipHostMallocDefault));
  return g;
}

void replay_step(StepGraph& g, Env& e) {
  Ctx& c = *g.c;
  HIP_CHECK(hipGraphLaunch(g.exec, c.stream));
  if (g.nchecks)
    HIP_CHECK(hipMemcpyAsync(g.host_checks, g.checks.data(), g.nchecks * 8, hipMemcpyDeviceToHost, c.stream));
  e.draws += 4;
}

bool finish_step(StepGraph& g, Env& e, std::string* why) {
  if (validate_checks(g.recs, g.host_checks, why)) return true;
  Ctx& c = *g.c;
  for (int i = 0; i < 4; ++i) {
    dev_copy(c, e.C[i].data(), g.bakC[i].data(), e.C[i].bytes());
    dev_copy(c, e.E[i].data(), g.bakE[i].data(), e.E[i].bytes());
  }
  e.draws -= 4;
  HIP_CHECK(hipMemcpyAsync(e.draws_dev.data(), &e.draws, 8, hipMemcpyHostToDevice, c.stream));
  c.sync();
  return false;
}

void adopt_step(StepGraph& g, Env& e) {
  Ctx& c = *g.c;
  for (int i = 0; i < 4; ++i)
    IPEPS_CHECK(e.C[i].dims == g.C[i].dims && e.E[i].dims == g.E[i].dims, "adopt_step: shapes changed");
  for (int i = 0; i < 4; ++i) {
    if (e.C[i].data() != g.C[i].data()) dev_copy(c, g.C[i].data(), e.C[i].data(), e.C[i].bytes());
    if (e.E[i].data() != g.E[i].data()) dev_copy(c, g.E[i].data(), e.E[i].data(), e.E[i].bytes());
    e.C[i] = g.C[i];
    e.E[i] = g.E[i];
  }
  HIP_CHECK(hipMemcpyAsync(e.draws_dev.data(), &e.draws, 8, hipMemcpyHostToDevice, c.stream));
  c.sync();
}

static std::vector<double> rg_step_impl(Ctx& c, Env& e, int64_t chi) {
  op::prec().mixed = e.opt.mixed;
  op::prec().back = e.opt.mixed_back ? (k::Mixed)e.opt.mixed_back : e.opt.mixed;
  op::prec().fused_absorb = e.opt.fused_absorb;
  const uint64_t d0 = e.draws;
  std::vector<double> s;
  if (e.opt.paired_moves) {
    // left || right, then top || bottom (src/ctm/ctmrg.jl:57-64: right reads
    // none of left's outputs, bottom none of top's)
    run_pair(c, e, [&](Ctx& cc) { s = left_rg(cc, e, chi, d0 + 0); }, [&](Ctx& cc) { right_rg(cc, e, chi, d0 + 1); });
    run_pair(c, e, [&](Ctx& cc) { top_rg(cc, e, chi, d0 + 2); }, [&](Ctx& cc) { bottom_rg(cc, e, chi, d0 + 3); });
  } else {
    s = left_rg(c, e, chi, d0 + 0);
    right_rg(c, e, chi, d0 + 1);
    top_rg(c, e, chi, d0 + 2);
    bottom_rg(c, e, chi, d0 + 3);
  }
  e.draws = d0 + 4;
  return s;
}

static Env as_dtype(Ctx& c, const Env& e, DType dt) {
  if (e.dt == dt) return e;
  Env r = e;
  r.dt = dt;
  r.A = op::cast(c, e.A, dt);
  r.Ad = op::cast(c, e.Ad, dt);
  for (int i = 0; i < 4; ++i) {
    r.C[i] = op::cast(c, e.C[i], dt);
    r.E[i] = op::cast(c, e.E[i], dt);
  }
  return r;
}

Tensor energy(Ctx& c, const Env& e0, const Tensor& H1, const Tensor& H2) {
  Env e = (e0.dt == DType::F32 && e0.opt.energy_fp64) ? as_dtype(c, e0, DType::F64) : e0;
  // the VJP of the contractions at energy_mixed_back (0: the step's VJP setting)
  const int step_back = e.opt.mixed_back ? e.opt.mixed_back : (int)e.opt.mixed;
  op::PrecScope ps(e.opt.energy_mixed, (k::Mixed)(e.opt.energy_mixed_back == 0 ? step_back : e.opt.energy_mixed_back));
  Tensor roh, rov;
  get_dms(c, e, roh, rov);
  if (roh.dt == DType::F32) {
    roh = op::cast(c, roh, DType::F64);
    rov = op::cast(c, rov, DType::F64);
  }
  const int64_t m = roh.dims[0];
  IPEPS_CHECK(H1.dims.size() == 2 && H1.dims[0] == m && H1.dims[1] == m, "energy: H dimension must be d^2");
  Tensor out = c.empty(DType::F64, {8});
  k::energy2(c, roh.dt, m, roh.data(), rov.data(), H1.ptr<double>(), H2.ptr<double>(), out.ptr<double>());
  if (op::recording() && (roh.node || rov.node)) {
    auto nh = roh.node, nv = rov.node;
    Tensor rh = roh, rv = rov;
    rh.node.reset();
    rv.node.reset();
    Tensor h1 = H1, h2 = H2;
    out.node = op::make_node([=](Ctx& c, Node&) {
      Tensor gh = c.empty(rh.dt, rh.dims), gv = c.empty(rv.dt, rv.dims);
      k::energy2_back(c, rh.dt, m, rh.data(), rv.data(), h1.ptr<double>(), h2.ptr<double>(), gh.data(), gv.data());
      if (nh) op::acc(c, nh, gh);
      if (nv) op::acc(c, nv, gv);
    });
  }
  return out;
}

void gs_norm(Ctx& c, const Env& e, double out[2]) {
  const auto& C = e.C;
  const auto& E = e.E;
  Tensor C1E1 = tcon(c, C[0], {-1, 1}, E[0], {1, -2, -3, -4});
  Tensor CEC2 = tcon(c, C1E1, {-1, 1, -3, -4}, C[1], {1, -2});
  Tensor CECE4 = tcon(c, CEC2, {1, -2, -3, -5}, E[3], {1, -1, -4, -6});
  Tensor C4E3 = tcon(c, C[3], {-1, 1}, E[2], {1, -2, -3, -4});
  Tensor CEC3 = tcon(c, C4E3, {-1, 1, -3, -4}, C[2], {-2, 1});
  Tensor CECE2 = tcon(c, CEC3, {-1, 1, -3, -5}, E[1], {-2, 1, -4, -6});
  Tensor ndm = tcon(c, CECE4, {1, 2, -1, -2, -5, -6}, CECE2, {1, 2, -3, -4, -7, -8});
  Tensor ndmAd = tcon(c, ndm, {-1, -2, -3, -4, 1, 2, 3, 4}, e.Ad, {1, 2, 3, 4, -5});
  Tensor nrm = tcon(c, ndmAd, {1, 2, 3, 4, 5}, e.A, {1, 2, 3, 4, 5});
  Tensor n64 = op::to_dtype(c, nrm, nrm.dt == DType::F32 ? DType::F64 : nrm.dt);
  if (n64.dt == DType::C128) {
    to_host(c, n64, out);
  } else {
    to_host(c, n64, out);
    out[1] = 0.0;
  }
}

static void read_energy(Ctx& c, const Tensor& t, double E[2]) {
  double h[8];
  to_host(c, t, h);
  E[0] = h[0];
  E[1] = h[1];
}

RunResult run_ctm(Ctx& c, Env& e, int64_t chi, int min_iter, int max_iter, double tol, int conv_mode,
                  const Tensor& H1, const Tensor& H2, bool) {
  // src/ctm/ctmrg.jl:1-55: diffs starts as [1.0], conv as the scalar 1.0
  RunResult r;
  r.diffs.push_back(1.0);
  std::vector<double> conv = {1.0, 0.0};  // complex scalar or spectrum
  bool conv_is_scalar = true;
  for (int i = 1; i <= max_iter; ++i) {
    r.iters = i;
    std::vector<double> s = rg_step(c, e, chi);
    r.s_last = s;
    std::vector<double> old = conv;
    bool old_scalar = conv_is_scalar;
    // conv only matters when it can decide the stop (i >= min_iter) or feeds
    // the next diff (i == min_iter - 1); skipping the rest changes nothing.
    const bool need = (conv_mode != 0) || (i >= min_iter - 1 && i < max_iter) || i == max_iter;
    if (conv_mode == 0) {
      if (need) {
        TapeScope ts(nullptr);
        Tensor Et = energy(c, e, H1, H2);
        double E[2];
        read_energy(c, Et, E);
        conv = {E[0], E[1]};
        conv_is_scalar = true;
        r.E[0] = E[0];
        r.E[1] = E[1];
      } else {
        conv = {NAN, NAN};
        conv_is_scalar = true;
      }
    } else {
      conv.clear();
      for (double v : s) { conv.push_back(v); conv.push_back(0.0); }
      conv_is_scalar = s.size() == 1;
    }
    // lengths equal -> diff = norm(conv - old)
    size_t len_new = conv.size() / 2, len_old = old.size() / 2;
    (void)old_scalar;
    if (len_new == len_old) {
      double acc = 0.0;
      for (size_t q = 0; q < conv.size(); ++q) acc += (conv[q] - old[q]) * (conv[q] - old[q]);
      r.diffs.push_back(std::sqrt(acc));
    }
    if (i >= min_iter && r.diffs.back() < tol) break;
  }
  return r;
}

namespace {

struct Checkpoint {
  std::array<Tensor, 4> C, E, warm;
  uint64_t draws = 0;
};

Checkpoint save(const Env& e) {
  Checkpoint k;
  k.draws = e.draws;
  k.warm = e.warm;
  for (int i = 0; i < 4; ++i) {
    k.C[i] = e.C[i];
    k.E[i] = e.E[i];
    k.C[i].node.reset();
    k.E[i].node.reset();
  }
  return k;
}

void load(Env& e, const Checkpoint& k) {
  e.draws = k.draws;
  e.warm = k.warm;
  for (int i = 0; i < 4; ++i) {
    e.C[i] = k.C[i];
    e.E[i] = k.E[i];
  }
}

}  // namespace

// Reverse sweep over checkpointed steps: cotC/cotE hold the cotangents of the
// environment after the last step; A's cotangent accumulates into gA.
static void backprop_steps(Ctx& c, Env& e, const std::vector<Checkpoint>& ck, int64_t chi, std::array<Tensor, 4> cotC,
                           std::array<Tensor, 4> cotE, Tensor& gA) {
  for (int t = (int)ck.size() - 1; t >= 0; --t) {
    Tape tape;
    tape.active = true;
    TapeScope ts(&tape);
    load(e, ck[t]);
    Tensor A = op::leaf(e.A), Ad;
    e.A = A;
    e.Ad = op::conj(c, A);
    for (int i = 0; i < 4; ++i) {
      e.C[i] = op::leaf(e.C[i]);
      e.E[i] = op::leaf(e.E[i]);
    }
    std::array<std::shared_ptr<Node>, 4> inC, inE;
    for (int i = 0; i < 4; ++i) {
      inC[i] = e.C[i].node;
      inE[i] = e.E[i].node;
    }
    rg_step(c, e, chi);
    for (int i = 0; i < 4; ++i) {
      if (cotC[i].defined()) op::acc(c, e.C[i].node, cotC[i]);
      if (cotE[i].defined()) op::acc(c, e.E[i].node, cotE[i]);
    }
    tape.active = false;
    op::backward(c, tape);
    if (A.node->grad.defined()) {
      if (!gA.defined()) gA = op::clone(c, A.node->grad);
      else k::axpy(c, gA.dt, gA.numel(), 1.0, A.node->grad.data(), gA.data());
    }
    for (int i = 0; i < 4; ++i) {
      cotC[i] = inC[i]->grad;
      cotE[i] = inE[i]->grad;
    }
    e.A.node.reset();
    e.Ad.node.reset();
    for (int i = 0; i < 4; ++i) {
      e.C[i].node.reset();
      e.E[i].node.reset();
    }
  }
}

// Leaves on the tape for A (Ad = conj(A) is derived) and, optionally, the
// environment; returns the A leaf.
static Tensor mark_leaves(Ctx& c, Env& e, bool env_leaves) {
  Tensor A = op::leaf(e.A);
  e.A = A;
  e.Ad = op::conj(c, A);
  if (env_leaves)
    for (int i = 0; i < 4; ++i) {
      e.C[i] = op::leaf(e.C[i]);
      e.E[i] = op::leaf(e.E[i]);
    }
  return A;
}

static void strip_nodes(Env& e) {
  e.A.node.reset();
  e.Ad.node.reset();
  for (int i = 0; i < 4; ++i) {
    e.C[i].node.reset();
    e.E[i].node.reset();
  }
}

GradResult gs_energy_grad(Ctx& c, const Tensor& A0, DType env_dt, const Options& opt, const Tensor& H1,
                          const Tensor& H2, int64_t chi, int min_iter, int max_iter, int ad_max_iter, double tol) {
  TapeScope none(nullptr);
  // x = renormalize(A0) outside the AD (src/optimize/optim_gs.jl:55)
  Tensor x = op::renormalize(c, A0);
  Env e;
  e.opt = opt;
  e.dt = env_dt;
  e.A = op::cast(c, x, env_dt);
  e.Ad = op::conj(c, e.A);
  init_ctm1(c, e);
  auto r = run_ctm(c, e, chi, min_iter, max_iter, tol, 0, H1, H2);
  GradResult g = run_gs_vjp(c, e, e.A, H1, H2, chi, min_iter, ad_max_iter, tol);
  g.conv_iters = r.iters;
  return g;
}

GradResult run_gs_vjp(Ctx& c, const Env& env0, const Tensor& xin, const Tensor& H1, const Tensor& H2, int64_t chi,
                      int min_iter, int ad_max_iter, double tol) {
  GradResult g;
  TapeScope none(nullptr);
  const Options& opt = env0.opt;
  const DType env_dt = env0.dt;
  // setproperties(ts, A = x, Ad = conj(x)) on a copy: env0's tensors are
  // never written (every step replaces them), so it is held constant
  Env e = env0;
  e.A = op::cast(c, xin, env_dt);
  e.Ad = op::conj(c, e.A);
  // Differentiated run (run_gs): the same stopping rule with max_iter =
  // ad_max_iter from the converged environment, which is held constant.
  // Without checkpointing one tape spans every step (memory ~ steps x the
  // per-step intermediates); with checkpointing only (C, E) per step is
  // kept and each step is recomputed on its own tape in the reverse sweep.
  Tape tape;
  tape.active = !opt.checkpoint;
  TapeScope ts(&tape);
  Tensor Aleaf;
  if (!opt.checkpoint) Aleaf = mark_leaves(c, e, false);
  std::vector<Checkpoint> ck;
  std::vector<double> conv = {1.0, 0.0};
  std::vector<double> diffs = {1.0};
  int it = 0;
  for (int i = 1; i <= ad_max_iter; ++i) {
    it = i;
    if (opt.checkpoint) ck.push_back(save(e));
    rg_step(c, e, chi);
    const bool need = (i >= min_iter - 1 && i < ad_max_iter);
    if (need) {
      TapeScope off(nullptr);  // conv_fun runs under Zygote.@ignore (src/ctm/ctmrg.jl:22)
      double E[2];
      read_energy(c, energy(c, e, H1, H2), E);
      std::vector<double> nc = {E[0], E[1]};
      diffs.push_back(std::hypot(nc[0] - conv[0], nc[1] - conv[1]));
      conv = nc;
    } else {
      conv = {NAN, NAN};
      diffs.push_back(NAN);
    }
    if (i >= min_iter && diffs.back() < tol) break;
  }
  g.ad_iters = it;
  Tensor gA;
  if (!opt.checkpoint) {
    Tensor Et = energy(c, e, H1, H2);
    op::acc(c, Et.node, c.empty(DType::F64, {8}));  // the energy node takes dL/dRe(E) = 1
    tape.active = false;
    read_energy(c, Et, g.E);
    op::backward(c, tape);
    if (Aleaf.node->grad.defined()) gA = Aleaf.node->grad;
    else gA = c.zeros(Aleaf.dt, Aleaf.dims);
    strip_nodes(e);
  } else {
    std::array<Tensor, 4> cotC, cotE;
    Tape t2;
    t2.active = true;
    TapeScope ts2(&t2);
    Tensor A = mark_leaves(c, e, true);
    Tensor Et = energy(c, e, H1, H2);
    op::acc(c, Et.node, c.empty(DType::F64, {8}));
    t2.active = false;
    read_energy(c, Et, g.E);
    op::backward(c, t2);
    for (int i = 0; i < 4; ++i) {
      cotC[i] = e.C[i].node->grad;
      cotE[i] = e.E[i].node->grad;
    }
    if (A.node->grad.defined()) gA = op::clone(c, A.node->grad);
    strip_nodes(e);
    backprop_steps(c, e, ck, chi, cotC, cotE, gA);
  }
  DType gdt = env_dt == DType::F32 ? DType::F64 : env_dt;
  g.grad = op::to_dtype(c, gA, gdt);
  return g;
}

Tensor step_and_adjoint(Ctx& c, Env& e, int64_t chi, const Tensor& H1, const Tensor& H2, double E_out[2]) {
  TapeScope none(nullptr);
  if (e.opt.checkpoint) {
    std::vector<Checkpoint> ck = {save(e)};
    rg_step(c, e, chi);
    std::array<Tensor, 4> cotC, cotE;
    Tensor gA;
    {
      Tape tape;
      tape.active = true;
      TapeScope ts(&tape);
      Tensor A = mark_leaves(c, e, true);
      Tensor Et = energy(c, e, H1, H2);
      op::acc(c, Et.node, c.empty(DType::F64, {8}));
      tape.active = false;
      op::backward(c, tape);
      for (int i = 0; i < 4; ++i) {
        cotC[i] = e.C[i].node->grad;
        cotE[i] = e.E[i].node->grad;
      }
      if (A.node->grad.defined()) gA = op::clone(c, A.node->grad);
      read_energy(c, Et, E_out);
      strip_nodes(e);
    }
    Checkpoint after = save(e);
    backprop_steps(c, e, ck, chi, cotC, cotE, gA);
    load(e, after);  // leave the environment advanced by one step
    e.Ad = op::conj(c, e.A);
    return gA;
  }
  // one tape over the step and the energy
  Tape tape;
  tape.active = true;
  TapeScope ts(&tape);
  Tensor A = mark_leaves(c, e, false);
  rg_step(c, e, chi);
  Tensor Et = energy(c, e, H1, H2);
  op::acc(c, Et.node, c.empty(DType::F64, {8}));
  tape.active = false;
  op::backward(c, tape);
  Tensor gA = A.node->grad.defined() ? A.node->grad : c.zeros(A.dt, A.dims);
  read_energy(c, Et, E_out);
  strip_nodes(e);
  return gA;
}

}  // namespace ctm
}  // namespace ipeps
