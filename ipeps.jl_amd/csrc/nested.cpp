// Excited-state (nested tensor) CTMRG on the device.
//
// A nested tensor is [X, X_B, X_Bd, X_BB] (src/nested_tensor/nested_tensor.jl),
// an undefined component standing for EmptyT.  Contractions follow the
// product rule of nested_tensor.jl:48-63, the moves use projectors built from
// the plain parts (src/ctm/ctmrg.jl:122-252), the new tensors are scaled by
// the plain component's max-abs and shifted by the momentum phase
// (ctmrg.jl:377-491).
#include <cmath>
#include <cstring>

#include "ctmrg.h"

namespace ipeps {
namespace ctm {

using L = std::vector<int>;

namespace {

NT plain(const Tensor& t) {
  NT r;
  r.c[0] = t;
  r.nested = false;
  return r;
}

NT nt(const Tensor& a, const Tensor& b, const Tensor& cc, const Tensor& d) {
  NT r;
  r.c = {a, b, cc, d};
  return r;
}

Tensor tc(Ctx& c, const Tensor& a, const L& la, const Tensor& b, const L& lb) {
  if (!a.defined() || !b.defined()) return Tensor();
  return op::tcon(c, a, la, b, lb);
}

Tensor addt(Ctx& c, const Tensor& a, const Tensor& b) {
  if (!a.defined()) return b;
  if (!b.defined()) return a;
  return op::add(c, a, b);
}

// nested_tensor.jl:48-63 (and the NestedTensor x Array methods below it)
NT ntcon(Ctx& c, const NT& A, const L& la, const NT& B, const L& lb) {
  NT r;
  if (A.nested && B.nested) {
    r.c[0] = tc(c, A.c[0], la, B.c[0], lb);
    r.c[1] = addt(c, tc(c, A.c[0], la, B.c[1], lb), tc(c, A.c[1], la, B.c[0], lb));
    r.c[2] = addt(c, tc(c, A.c[0], la, B.c[2], lb), tc(c, A.c[2], la, B.c[0], lb));
    r.c[3] = addt(c, addt(c, tc(c, A.c[0], la, B.c[3], lb), tc(c, A.c[3], la, B.c[0], lb)),
                  addt(c, tc(c, A.c[1], la, B.c[2], lb), tc(c, A.c[2], la, B.c[1], lb)));
  } else if (A.nested) {
    for (int i = 0; i < 4; ++i) r.c[i] = tc(c, A.c[i], la, B.c[0], lb);
  } else if (B.nested) {
    for (int i = 0; i < 4; ++i) r.c[i] = tc(c, A.c[0], la, B.c[i], lb);
  } else {
    r.c[0] = tc(c, A.c[0], la, B.c[0], lb);
    r.nested = false;
  }
  return r;
}

NT nreshape(Ctx& c, const NT& x, const std::vector<int64_t>& dims) {
  NT r = x;
  for (auto& t : r.c)
    if (t.defined()) t = op::reshape(c, t, dims);
  return r;
}

NT ntranspose(Ctx& c, const NT& x) {
  NT r = x;
  for (auto& t : r.c)
    if (t.defined()) t = op::permute(c, t, {1, 0});
  return r;
}

NT nmatmul(Ctx& c, const NT& a, const NT& b) { return ntcon(c, a, {-1, 1}, b, {1, -2}); }

// nested_tensor.jl:41-46: every component divided by max|X_1|
NT nrenormalize(Ctx& c, const NT& x) {
  NT r = x;
  r.c[0] = op::renormalize(c, x.c[0]);
  for (int i = 1; i < 4; ++i)
    if (x.c[i].defined()) r.c[i] = op::renormalize_by(c, x.c[i], x.c[0]);
  return r;
}

// nested_tensor.jl:65
NT shift(Ctx& c, const NT& x, double phi) {
  if (!x.nested || phi == 0.0) return x;
  NT r = x;
  if (x.c[1].defined()) r.c[1] = op::scale_cplx(c, x.c[1], std::cos(phi), std::sin(phi));
  if (x.c[2].defined()) r.c[2] = op::scale_cplx(c, x.c[2], std::cos(phi), -std::sin(phi));
  return r;
}

NT all_A(const NestedEnv& e) { return nt(e.g.A, e.B, Tensor(), Tensor()); }
NT all_Ad(const NestedEnv& e) { return nt(e.g.Ad, Tensor(), e.Bd, Tensor()); }
NT all_C(const NestedEnv& e, int i) { return nt(e.g.C[i], e.BC[i], e.BdC[i], e.BBC[i]); }
NT all_E(const NestedEnv& e, int i) { return nt(e.g.E[i], e.BE[i], e.BdE[i], e.BBE[i]); }

void store_C(NestedEnv& e, int i, const NT& x) {
  e.g.C[i] = x.c[0];
  e.BC[i] = x.c[1];
  e.BdC[i] = x.c[2];
  e.BBC[i] = x.c[3];
}
void store_E(NestedEnv& e, int i, const NT& x) {
  e.g.E[i] = x.c[0];
  e.BE[i] = x.c[1];
  e.BdE[i] = x.c[2];
  e.BBE[i] = x.c[3];
}

// reference order of proj_* (ctmrg.jl:285-289)
NT edge(Ctx& c, const NT& E, const NT& A, const NT& Ad, const L& ia, const L& iad, const NT& P, const NT& Pd) {
  NT EA = ntcon(c, E, {-1, -3, 1, -6}, A, ia);
  NT EAAd = ntcon(c, EA, {-1, -2, -4, -5, -7, 1, 2}, Ad, iad);
  auto s = EAAd.c[0].dims;
  EAAd = nreshape(c, EAAd, {s[0] * s[1] * s[2], s[3] * s[4] * s[5], s[6], s[7]});
  NT EPd = ntcon(c, Pd, {-1, 1}, EAAd, {1, -2, -3, -4});
  return ntcon(c, EPd, {-1, 1, -3, -4}, P, {1, -2});
}

std::vector<int64_t> rows_rest(const NT& x) {
  const Tensor& t = x.c[0];
  return {t.dims[0], t.numel() / t.dims[0]};
}
std::vector<int64_t> rest_last(const NT& x) {
  const Tensor& t = x.c[0];
  int64_t l = t.dims[3];
  return {t.numel() / l, l};
}

std::vector<double> move(Ctx& c, NestedEnv& e, int dir, int64_t chi) {
  Tensor P1, P2;
  std::vector<double> s;
  const int64_t t = e.moves++;
  if (e.path && t < (int64_t)e.path->P.size()) {
    P1 = e.path->P[t][0];
    P2 = e.path->P[t][1];
    s = e.path->s[t];
  } else {
    projector(c, e.g, dir, chi, P1, P2, s);
  }
  if (e.record) {
    e.record->P.push_back({P1, P2});
    e.record->s.push_back(s);
  }
  const NT P = plain(P1), Pd = plain(P2), A = all_A(e), Ad = all_Ad(e);
  if (dir == 0) {  // proj_left + up_left
    NT C1E1 = ntcon(c, all_C(e, 0), {-2, 1}, all_E(e, 0), {1, -1, -3, -4});
    NT nC1 = ntranspose(c, nmatmul(c, nreshape(c, C1E1, rows_rest(C1E1)), P));
    NT C4E3 = ntcon(c, all_C(e, 3), {-1, 1}, all_E(e, 2), {1, -4, -2, -3});
    NT nC4 = nmatmul(c, Pd, nreshape(c, C4E3, rest_last(C4E3)));
    NT nE4 = edge(c, all_E(e, 3), A, Ad, {-2, 1, -4, -5, -7}, {-3, 1, -6, -8, 2}, P, Pd);
    store_C(e, 0, shift(c, nrenormalize(c, nC1), -e.px));
    store_E(e, 3, shift(c, nrenormalize(c, nE4), -e.px));
    store_C(e, 3, shift(c, nrenormalize(c, nC4), -e.px));
  } else if (dir == 1) {  // proj_right + up_right
    NT C2E1 = ntcon(c, all_C(e, 1), {1, -2}, all_E(e, 0), {-1, 1, -3, -4});
    NT nC2 = nmatmul(c, nreshape(c, C2E1, rows_rest(C2E1)), P);
    NT C3E3 = ntcon(c, all_C(e, 2), {-1, 1}, all_E(e, 2), {-4, 1, -2, -3});
    NT nC3 = nmatmul(c, Pd, nreshape(c, C3E3, rest_last(C3E3)));
    NT nE2 = edge(c, all_E(e, 1), A, Ad, {-2, -5, -4, 1, -7}, {-3, -8, -6, 1, 2}, P, Pd);
    store_C(e, 1, shift(c, nrenormalize(c, nC2), e.px));
    store_E(e, 1, shift(c, nrenormalize(c, nE2), e.px));
    store_C(e, 2, shift(c, nrenormalize(c, nC3), e.px));
  } else if (dir == 2) {  // proj_top + up_top
    NT C1E4 = ntcon(c, all_C(e, 0), {1, -2}, all_E(e, 3), {1, -1, -3, -4});
    NT nC1 = nmatmul(c, nreshape(c, C1E4, rows_rest(C1E4)), P);
    NT C2E2 = ntcon(c, all_C(e, 1), {-1, 1}, all_E(e, 1), {1, -4, -2, -3});
    NT nC2 = nmatmul(c, Pd, nreshape(c, C2E2, rest_last(C2E2)));
    NT nE1 = edge(c, all_E(e, 0), A, Ad, {1, -2, -5, -4, -7}, {1, -3, -8, -6, 2}, P, Pd);
    store_C(e, 0, shift(c, nrenormalize(c, nC1), -e.py));
    store_E(e, 0, shift(c, nrenormalize(c, nE1), -e.py));
    store_C(e, 1, shift(c, nrenormalize(c, nC2), -e.py));
  } else {  // proj_bottom + up_bottom
    NT C4E4 = ntcon(c, all_C(e, 3), {1, -2}, all_E(e, 3), {-1, 1, -3, -4});
    NT nC4 = nmatmul(c, nreshape(c, C4E4, rows_rest(C4E4)), P);
    NT C3E2 = ntcon(c, all_C(e, 2), {1, -1}, all_E(e, 1), {-4, 1, -2, -3});
    NT nC3 = ntranspose(c, nmatmul(c, Pd, nreshape(c, C3E2, rest_last(C3E2))));
    NT nE3 = edge(c, all_E(e, 2), A, Ad, {-5, -2, 1, -4, -7}, {-8, -3, 1, -6, 2}, P, Pd);
    store_C(e, 3, shift(c, nrenormalize(c, nC4), e.py));
    store_E(e, 2, shift(c, nrenormalize(c, nE3), e.py));
    store_C(e, 2, shift(c, nrenormalize(c, nC3), e.py));
  }
  return s;
}

// evaluation.jl:169-200 on nested tensors
NT dm_h(Ctx& c, const NT& C1, const NT& C2, const NT& C3, const NT& C4, const NT& E1l, const NT& E1r, const NT& E2,
        const NT& E3l, const NT& E3r, const NT& E4, const NT& Al, const NT& Ar, const NT& Adl, const NT& Adr) {
  NT C1E1 = ntcon(c, C1, {-1, 1}, E1l, {1, -2, -3, -4});
  NT LU = ntcon(c, C1E1, {-1, -2, -3, 1}, Adl, {1, -4, -5, -6, -7});
  NT C4E3 = ntcon(c, C4, {-1, 1}, E3l, {1, -2, -3, -4});
  NT CEE4 = ntcon(c, E4, {-1, 1, -3, -5}, C4E3, {1, -2, -4, -6});
  NT LB = ntcon(c, CEE4, {-1, -2, 1, 2, -4, -5}, Al, {-3, 1, 2, -6, -7});
  NT Lt = ntcon(c, LB, {1, -2, 2, 3, 4, -3, -5}, LU, {1, -1, 2, 3, 4, -4, -6});
  NT C2E1 = ntcon(c, C2, {1, -2}, E1r, {-1, 1, -3, -4});
  NT RU = ntcon(c, C2E1, {-1, -2, 1, -3}, Ar, {1, -4, -5, -6, -7});
  NT C3E3 = ntcon(c, C3, {-1, 1}, E3r, {-2, 1, -3, -4});
  NT CEE2 = ntcon(c, C3E3, {1, -2, -3, -5}, E2, {-1, 1, -4, -6});
  NT RB = ntcon(c, CEE2, {-1, -2, -3, -4, 1, 2}, Adr, {-5, -6, 1, 2, -7});
  NT Rt = ntcon(c, RU, {-1, 1, 2, -3, 3, 4, -5}, RB, {1, -2, 3, 4, 2, -4, -6});
  NT roh = ntcon(c, Lt, {1, 2, 3, 4, -1, -3}, Rt, {1, 2, 3, 4, -2, -4});
  int64_t d = Al.c[0].dims[4];
  return nreshape(c, roh, {d * d, d * d});
}

// evaluation.jl:217-248 on nested tensors
NT dm_v(Ctx& c, const NT& C1, const NT& C2, const NT& C3, const NT& C4, const NT& E1, const NT& E2u, const NT& E2d,
        const NT& E3, const NT& E4u, const NT& E4d, const NT& Au, const NT& Ad_, const NT& Adu, const NT& Add) {
  NT C1E1 = ntcon(c, C1, {-1, 1}, E1, {1, -2, -3, -4});
  NT CEE4 = ntcon(c, C1E1, {1, -2, -3, -5}, E4u, {1, -1, -4, -6});
  NT UL = ntcon(c, CEE4, {-1, -2, 1, 2, -3, -4}, Au, {1, 2, -5, -6, -7});
  NT C2E2 = ntcon(c, C2, {-1, 1}, E2u, {1, -2, -3, -4});
  NT UR = ntcon(c, C2E2, {-1, -2, -3, 1}, Adu, {-4, -5, -6, 1, -7});
  NT U = ntcon(c, UL, {-1, 1, 3, 4, -3, 2, -5}, UR, {1, -2, 2, 3, 4, -4, -6});
  NT C4E3 = ntcon(c, C4, {-1, 1}, E3, {1, -2, -3, -4});
  NT CEE4b = ntcon(c, C4E3, {1, -2, -4, -6}, E4d, {-1, 1, -3, -5});
  NT BL = ntcon(c, CEE4b, {-1, -2, -3, -4, 1, 2}, Add, {-5, 1, 2, -6, -7});
  NT C3E2 = ntcon(c, C3, {1, -2}, E2d, {-1, 1, -3, -4});
  NT BR = ntcon(c, C3E2, {-1, -2, 1, -3}, Ad_, {-4, -5, -6, 1, -7});
  NT Bm = ntcon(c, BL, {-1, 1, 2, 3, -4, 4, -6}, BR, {-2, 1, 4, -3, 2, 3, -5});
  NT rov = ntcon(c, U, {1, 2, 3, 4, -1, -3}, Bm, {1, 2, 3, 4, -2, -4});
  int64_t d = Au.c[0].dims[4];
  return nreshape(c, rov, {d * d, d * d});
}

NT single_dm(Ctx& c, const NestedEnv& e) {
  NT C1E1 = ntcon(c, all_C(e, 0), {-1, 1}, all_E(e, 0), {1, -2, -3, -4});
  NT CEC2 = ntcon(c, C1E1, {-1, 1, -3, -4}, all_C(e, 1), {1, -2});
  NT CECE4 = ntcon(c, CEC2, {1, -2, -3, -5}, all_E(e, 3), {1, -1, -4, -6});
  NT C4E3 = ntcon(c, all_C(e, 3), {-1, 1}, all_E(e, 2), {1, -2, -3, -4});
  NT CEC3 = ntcon(c, C4E3, {-1, 1, -3, -4}, all_C(e, 2), {-2, 1});
  NT CECE2 = ntcon(c, CEC3, {-1, 1, -3, -5}, all_E(e, 1), {-2, 1, -4, -6});
  return ntcon(c, CECE4, {1, 2, -1, -2, -5, -6}, CECE2, {1, 2, -3, -4, -7, -8});
}

}  // namespace

std::vector<double> nested_rg_step(Ctx& c, NestedEnv& e, int64_t chi) {
  op::prec().mixed = e.g.opt.mixed;
  op::prec().back = e.g.opt.mixed;
  IPEPS_CHECK(e.g.dt == DType::C128 && (!e.B.defined() || e.B.dt == DType::C128), "nested CTMRG runs in complex FP64");
  auto s = move(c, e, 0, chi);
  move(c, e, 1, chi);
  move(c, e, 2, chi);
  move(c, e, 3, chi);
  e.g.draws += 4;  // the projectors of the four moves used draws + 0..3
  return s;
}

Tensor es_energy(Ctx& c, const NestedEnv& e, const Tensor& H1, const Tensor& H2) {
  // get_dms(ts; only_gs=false) (evaluation.jl:138-150)
  NT C1 = all_C(e, 0), C2 = all_C(e, 1), C3 = all_C(e, 2), C4 = all_C(e, 3);
  NT E1 = all_E(e, 0), E2 = all_E(e, 1), E3 = all_E(e, 2), E4 = all_E(e, 3);
  NT A = all_A(e), Ad = all_Ad(e);
  const double px = e.px, py = e.py;
  NT roh = dm_h(c, C1, shift(c, C2, px), shift(c, C3, px), C4, E1, shift(c, E1, px), shift(c, E2, px), E3,
                shift(c, E3, px), E4, A, shift(c, A, px), Ad, shift(c, Ad, px));
  NT rov = dm_v(c, C1, C2, shift(c, C3, py), shift(c, C4, py), E1, E2, shift(c, E2, py), shift(c, E3, py), E4,
                shift(c, E4, py), A, shift(c, A, py), Ad, shift(c, Ad, py));
  IPEPS_CHECK(roh.c[3].defined() && rov.c[3].defined(), "es_energy: BB component missing");
  const int64_t m = roh.c[0].dims[0];
  Tensor out = c.empty(DType::F64, {4});
  k::es_energy(c, m, roh.c[0].data(), rov.c[0].data(), roh.c[3].data(), rov.c[3].data(), H1.ptr<double>(),
               H2.ptr<double>(), out.ptr<double>());
  if (op::recording() && (roh.c[3].node || rov.c[3].node)) {
    auto nh = roh.c[3].node, nv = rov.c[3].node;
    Tensor h1 = H1, h2 = H2, o = out;
    std::vector<int64_t> dims = roh.c[3].dims;
    out.node = op::make_node([=](Ctx& c, Node&) {
      Tensor gh = c.empty(DType::C128, dims), gv = c.empty(DType::C128, dims);
      k::es_energy_back(c, m, h1.ptr<double>(), h2.ptr<double>(), o.ptr<double>(), gh.data(), gv.data());
      if (nh) op::acc(c, nh, gh);
      if (nv) op::acc(c, nv, gv);
    });
  }
  return out;
}

void all_norm(Ctx& c, const NestedEnv& e, double Nb[2], Tensor& envB) {
  NT ndm = single_dm(c, e);
  NT ndmA = ntcon(c, ndm, {1, 2, 3, 4, -1, -2, -3, -4}, all_A(e), {1, 2, 3, 4, -5});
  NT alln = ntcon(c, ndmA, {1, 2, 3, 4, 5}, all_Ad(e), {1, 2, 3, 4, 5});
  IPEPS_CHECK(ndmA.c[1].defined(), "all_norm: B component missing");
  double n0[2];
  to_host(c, alln.c[0], n0);
  double nb[2];
  Tensor nbt = op::tcon(c, ndmA.c[1], {1, 2, 3, 4, 5}, e.Bd, {1, 2, 3, 4, 5});
  to_host(c, nbt, nb);
  // complex division by all_norm[1][1]
  const double den = n0[0] * n0[0] + n0[1] * n0[1];
  const double ir = n0[0] / den, ii = -n0[1] / den;  // 1 / n0
  Nb[0] = nb[0] * ir - nb[1] * ii;
  Nb[1] = nb[0] * ii + nb[1] * ir;
  envB = op::scale_cplx(c, ndmA.c[1], ir, ii);
}

namespace {

bool same_opts(const Options& a, const Options& b) {
  const op::SvdOpts &x = a.svd, &y = b.svd;
  return x.mode == y.mode && x.oversample == y.oversample && x.power_iters == y.power_iters &&
         x.full_below == y.full_below && x.seed == y.seed && x.mixed_cutoff == y.mixed_cutoff &&
         x.fast_orth == y.fast_orth && x.fast_svd == y.fast_svd && x.gram_guard == y.gram_guard &&
         x.power_iters_max == y.power_iters_max && x.resid_tol == y.resid_tol && x.warm_start == y.warm_start &&
         x.early_mixed == y.early_mixed && x.late_apps == y.late_apps &&
         a.canonical_edge == b.canonical_edge && a.mixed == b.mixed && a.mixed_back == b.mixed_back &&
         a.mixed_proj == b.mixed_proj && a.fused_absorb == b.fused_absorb && a.paired_moves == b.paired_moves;
}

bool same_tensor(Ctx& c, const Tensor& a, const Tensor& b) {
  if (a.defined() != b.defined()) return false;
  if (!a.defined()) return true;
  if (a.dt != b.dt || a.dims != b.dims) return false;
  if (a.data() == b.data() || a.bytes() == 0) return true;
  std::vector<char> ha(a.bytes()), hb(b.bytes());
  HIP_CHECK(hipMemcpyAsync(ha.data(), a.data(), a.bytes(), hipMemcpyDeviceToHost, c.stream));
  HIP_CHECK(hipMemcpyAsync(hb.data(), b.data(), b.bytes(), hipMemcpyDeviceToHost, c.stream));
  c.sync();
  return std::memcmp(ha.data(), hb.data(), ha.size()) == 0;
}

}  // namespace

std::shared_ptr<const ProjPath> nested_proj_path(Ctx& c, const Env& g, int64_t chi, int steps) {
  IPEPS_CHECK(chi > 0 && steps >= 0, "nested_proj_path: bad arguments");
  IPEPS_CHECK(g.dt == DType::C128, "nested CTMRG runs in complex FP64");
  IPEPS_CHECK(!g.opt.svd.warm_start, "nested_proj_path: warm-started projectors depend on the run's history");
  TapeScope none(nullptr);
  auto p = std::make_shared<ProjPath>();
  p->start = g;
  p->chi = chi;
  NestedEnv e;
  e.g = g;
  e.record = p.get();
  for (int i = 0; i < steps; ++i) nested_rg_step(c, e, chi);
  c.sync();  // other contexts' streams read the projectors
  return p;
}

bool path_matches(Ctx& c, const ProjPath& path, const Env& env, int64_t chi) {
  const Env& s = path.start;
  if (path.chi != chi || s.dt != env.dt || s.draws != env.draws || !same_opts(s.opt, env.opt)) return false;
  if (!same_tensor(c, s.A, env.A) || !same_tensor(c, s.Ad, env.Ad)) return false;
  for (int i = 0; i < 4; ++i)
    if (!same_tensor(c, s.C[i], env.C[i]) || !same_tensor(c, s.E[i], env.E[i])) return false;
  return true;
}

EsGrad es_grad(Ctx& c, const Env& g, const Tensor& B, const Tensor& H1, const Tensor& H2, double px, double py,
               int64_t chi, int min_iter, int max_iter, int ad_max_iter, double tol,
               std::shared_ptr<const ProjPath> path) {
  TapeScope none(nullptr);
  IPEPS_CHECK(!path || path_matches(c, *path, g, chi),
              "es_grad: the projector path was made from another environment state, chi or options");
  EsGrad r;
  NestedEnv e;
  e.g = g;
  e.path = std::move(path);
  e.B = B;
  e.Bd = op::conj(c, B);
  e.px = px;
  e.py = py;
  auto conv_of = [&](NestedEnv& x, double out[2]) {
    // conv_fun = get_es_energy / get_all_norm[1] (src/optimize/optim_es.jl:175)
    TapeScope off(nullptr);
    double h[4];
    to_host(c, es_energy(c, x, H1, H2), h);
    double nb[2];
    Tensor envB;
    all_norm(c, x, nb, envB);
    const double den = nb[0] * nb[0] + nb[1] * nb[1];
    out[0] = h[0] * nb[0] / den;
    out[1] = -h[0] * nb[1] / den;
  };
  // converge the nested environment (run_ctm, ctmrg.jl:1-55)
  double conv[2] = {1.0, 0.0};
  std::vector<double> diffs = {1.0};
  for (int i = 1; i <= max_iter; ++i) {
    r.conv_iters = i;
    nested_rg_step(c, e, chi);
    if (i >= min_iter - 1 && i < max_iter) {
      double nc[2];
      conv_of(e, nc);
      diffs.push_back(std::hypot(nc[0] - conv[0], nc[1] - conv[1]));
      conv[0] = nc[0];
      conv[1] = nc[1];
    } else {
      conv[0] = conv[1] = NAN;
      diffs.push_back(NAN);
    }
    if (i >= min_iter && diffs.back() < tol) break;
  }
  // differentiated run (run_es, optim_es.jl:196-210)
  Tape tape;
  tape.active = true;
  TapeScope ts(&tape);
  Tensor Bl = op::leaf(B);
  e.B = Bl;
  e.Bd = op::conj(c, Bl);
  conv[0] = 1.0;
  conv[1] = 0.0;
  diffs = {1.0};
  for (int i = 1; i <= ad_max_iter; ++i) {
    nested_rg_step(c, e, chi);
    if (i >= min_iter - 1 && i < ad_max_iter) {
      double nc[2];
      conv_of(e, nc);
      diffs.push_back(std::hypot(nc[0] - conv[0], nc[1] - conv[1]));
      conv[0] = nc[0];
      conv[1] = nc[1];
    } else {
      diffs.push_back(NAN);
    }
    if (i >= min_iter && diffs.back() < tol) break;
  }
  Tensor Et = es_energy(c, e, H1, H2);
  op::acc(c, Et.node, c.empty(DType::F64, {4}));
  tape.active = false;
  double h[4];
  to_host(c, Et, h);
  r.E4 = h[0];
  op::backward(c, tape);
  r.gH = Bl.node->grad.defined() ? Bl.node->grad : c.zeros(B.dt, B.dims);
  // strip tape references before the norm
  e.B = B;
  e.Bd = op::conj(c, B);
  auto strip = [](Tensor& t) { t.node.reset(); };
  for (int i = 0; i < 4; ++i) {
    strip(e.g.C[i]); strip(e.g.E[i]); strip(e.BC[i]); strip(e.BdC[i]); strip(e.BBC[i]);
    strip(e.BE[i]); strip(e.BdE[i]); strip(e.BBE[i]);
  }
  {
    TapeScope off(nullptr);
    all_norm(c, e, r.Nb, r.envB);
  }
  return r;
}

}  // namespace ctm
}  // namespace ipeps
