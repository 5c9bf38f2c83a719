// The 2 x 2 plaquette energy of a four-site Hamiltonian given as a product of
// local factors (src/graph/evaluation.jl:340-355, 443-519): the four
// enlarged corners of the plaquette, each contracted once with A, Ad and once
// with A, op_Ad (Ad with its factor s_i absorbed), joined into
// energy = tr(L_e R_e) and norm = tr(L_n R_n).
#include "ctmrg.h"

namespace ipeps {
namespace ctm {

namespace {

Tensor mat3(Ctx& c, const Tensor& x) {  // (d1 d2 d3) x rest, column-major
  const int64_t r = x.dims[0] * x.dims[1] * x.dims[2];
  return op::reshape(c, x, {r, x.numel() / r});
}

// get_op_Ad4 (src/graph/evaluation.jl:504-519): s_i absorbed into Ad
std::array<Tensor, 4> op_ad4(Ctx& c, const std::array<Tensor, 4>& s, const Tensor& Ad) {
  std::array<Tensor, 4> o;
  Tensor t = op::tcon(c, s[0], {-6, 1, -4}, Ad, {-1, -2, -3, -5, 1});
  auto d = t.dims;
  o[0] = op::reshape(c, t, {d[0], d[1], d[2], d[3] * d[4], d[5]});
  t = op::tcon(c, s[1], {-2, -7, 1, -4}, Ad, {-1, -3, -5, -6, 1});
  d = t.dims;
  o[1] = op::reshape(c, t, {d[0], d[1] * d[2], d[3] * d[4], d[5], d[6]});
  t = op::tcon(c, s[2], {-1, -7, 1, -3}, Ad, {-2, -4, -5, -6, 1});
  d = t.dims;
  o[2] = op::reshape(c, t, {d[0] * d[1], d[2] * d[3], d[4], d[5], d[6]});
  t = op::tcon(c, s[3], {-4, -6, 1}, Ad, {-1, -2, -3, -5, 1});
  d = t.dims;
  o[3] = op::reshape(c, t, {d[0], d[1], d[2], d[3] * d[4], d[5]});
  return o;
}

}  // namespace

void energy_4x4(Ctx& c, const Env& e0, const std::array<Tensor, 4>& s_in, double out[4]) {
  TapeScope none(nullptr);
  // one complex (or real) FP64 working type for the environment and the factors
  bool cplx = e0.dt == DType::C128;
  for (auto& s : s_in) cplx = cplx || s.dt == DType::C128;
  const DType w = cplx ? DType::C128 : DType::F64;
  auto cv = [&](const Tensor& t) { return op::to_dtype(c, t, w); };
  std::array<Tensor, 4> C, E, s;
  for (int i = 0; i < 4; ++i) {
    C[i] = cv(e0.C[i]);
    E[i] = cv(e0.E[i]);
    s[i] = cv(s_in[i]);
  }
  const Tensor A = cv(e0.A), Ad = cv(e0.Ad);
  const auto o = op_ad4(c, s, Ad);
  using op::tcon;
  // upper left (A1), lower left (A4), upper right (A2), lower right (A3)
  Tensor C1E1 = tcon(c, C[0], {-1, 1}, E[0], {1, -2, -3, -4});
  Tensor CEE4 = tcon(c, C1E1, {1, -2, -5, -3}, E[3], {1, -1, -6, -4});
  Tensor CEEA1 = tcon(c, CEE4, {-1, -2, -3, -4, 1, 2}, A, {1, 2, -5, -6, -7});
  Tensor eLU = mat3(c, tcon(c, CEEA1, {-1, -4, 1, 2, -2, -5, 3}, o[0], {1, 2, -3, -6, 3}));
  Tensor nLU = mat3(c, tcon(c, CEEA1, {-1, -4, 1, 2, -2, -5, 3}, Ad, {1, 2, -3, -6, 3}));

  Tensor C4E3 = tcon(c, C[3], {-1, 1}, E[2], {1, -2, -3, -4});
  Tensor CEE4b = tcon(c, C4E3, {1, -2, -6, -4}, E[3], {-1, 1, -5, -3});
  Tensor CEEA4 = tcon(c, CEE4b, {-1, -2, -3, -4, 1, 2}, A, {-5, 1, 2, -6, -7});
  Tensor eLB = mat3(c, tcon(c, CEEA4, {-1, -4, 1, 2, -2, -5, 3}, o[3], {-3, 1, 2, -6, 3}));
  Tensor nLB = mat3(c, tcon(c, CEEA4, {-1, -4, 1, 2, -2, -5, 3}, Ad, {-3, 1, 2, -6, 3}));

  Tensor C2E1 = tcon(c, C[1], {1, -2}, E[0], {-1, 1, -3, -4});
  Tensor CEE2 = tcon(c, C2E1, {-1, 1, -5, -3}, E[1], {1, -2, -6, -4});
  Tensor CEEA2 = tcon(c, CEE2, {-1, -2, -3, -4, 1, 2}, A, {1, -5, -6, 2, -7});
  Tensor eRU = mat3(c, tcon(c, CEEA2, {-1, -4, 1, 2, -2, -5, 3}, o[1], {1, -3, -6, 2, 3}));
  Tensor nRU = mat3(c, tcon(c, CEEA2, {-1, -4, 1, 2, -2, -5, 3}, Ad, {1, -3, -6, 2, 3}));

  Tensor C3E3 = tcon(c, C[2], {-1, 1}, E[2], {-2, 1, -3, -4});
  Tensor CEE2b = tcon(c, C3E3, {1, -2, -5, -3}, E[1], {-1, 1, -6, -4});
  Tensor CEEA3 = tcon(c, CEE2b, {-1, -2, -5, -6, 1, 2}, A, {-3, -4, 1, 2, -7});
  Tensor eRB = mat3(c, tcon(c, CEEA3, {-1, -4, -2, -5, 1, 2, 3}, o[2], {-3, -6, 1, 2, 3}));
  Tensor nRB = mat3(c, tcon(c, CEEA3, {-1, -4, -2, -5, 1, 2, 3}, Ad, {-3, -6, 1, 2, 3}));

  // left = LU^T LB, right = RU RB, value = sum_ij left_ij right_ij
  auto join = [&](const Tensor& LU, const Tensor& LB, const Tensor& RU, const Tensor& RB, double v[2]) {
    Tensor left = tcon(c, LU, {1, -1}, LB, {1, -2});
    Tensor right = tcon(c, RU, {-1, 1}, RB, {1, -2});
    Tensor val = tcon(c, left, {1, 2}, right, {1, 2});
    double h[2] = {0.0, 0.0};
    to_host(c, val, h);
    v[0] = h[0];
    v[1] = val.dt == DType::C128 ? h[1] : 0.0;
  };
  join(eLU, eLB, eRU, eRB, out);
  join(nLU, nLB, nRU, nRB, out + 2);
}

}  // namespace ctm
}  // namespace ipeps
