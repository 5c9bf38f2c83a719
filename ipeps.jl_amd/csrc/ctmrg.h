// CTMRG environment, directional moves, energy, and the step-checkpointed
// adjoint of optim_gs's gradient.
#pragma once

#include <array>
#include <memory>
#include <vector>

#include "ops.h"

namespace ipeps {
namespace ctm {

// Engine options (settable per environment through the C ABI).
struct Options {
  op::SvdOpts svd;             // projector SVD (full = reference semantics, or randomized)
  bool canonical_edge = true;  // P2-first edge update (chi^3 D^4) instead of the chi^2 D^6 intermediate
  k::Mixed mixed = k::Mixed::BF16X3;
  int mixed_back = 0;          // products of the step's VJPs: 1, 3, 6, or 0 = `mixed`
  int mixed_proj = 0;          // products of the projector (range finder, P1 / P2): 1, 3, 6, or 0 = `mixed`
  // Energy contractions of an F32 environment: FP64 (energy_fp64) or split-bf16
  // with `energy_mixed` products; the trace and the ratios are FP64 either way.
  bool energy_fp64 = false;
  k::Mixed energy_mixed = k::Mixed::BF16X6;
  int energy_mixed_back = 0;   // products of the energy's VJP: 1, 3, 6, or 0 = `mixed`
  bool checkpoint = false;     // recompute each step in the reverse sweep instead of keeping its tape
  bool fused_absorb = true;    // fused FP32-MFMA site absorption in the corners (mixed engine, 4 <= D <= 6)
  bool paired_moves = true;    // left || right and top || bottom on two streams / host threads
  // rg_step in graph-safe mode (what a captured step runs), its deferred
  // checks validated after each step (an error if one fails)
  bool graph_safe = false;
  // batches only: each instance's rg_step captured into a HIP graph once its
  // shapes are stable and replayed (see StepGraph)
  bool graph = false;
};

// Plain ground-state environment (src/ctm/ctm_tensor.jl:3-18).  Storage dtype
// is F64, C128, or F32 (the mixed-precision engine).
struct Env {
  DType dt = DType::F64;
  Tensor A, Ad;                 // (D, D, D, D, d): legs (up, left, down, right, phys)
  std::array<Tensor, 4> C;      // C1 (down, right), C2 (left, down), C3 (up, left), C4 (up, right)
  std::array<Tensor, 4> E;      // E1 (l, r, ket, bra), E2 (u, d, ., .), E3 (l, r, ., .), E4 (u, d, ., .)
  Options opt;
  // Randomized projectors draw their test matrices from a counter-based
  // stream keyed by (opt.svd.seed, draws + direction); rg_step advances
  // draws by 4.  Two environments in the same state with the same seed draw
  // the same matrices whatever their dtype or move schedule, and a
  // recomputed step replays its forward draws.
  mutable uint64_t draws = 0;
  // Warm start of the randomized projectors (option warm_start): each move's
  // last left Ritz basis (m x k, FP64 / complex FP64), indexed by direction
  // (left, right, top, bottom); part of the environment's state.
  mutable std::array<Tensor, 4> warm;
  // Device copy of `draws` (one uint64 in an F64 tensor) while a captured
  // step is attached: the graph forms its test-matrix seeds from it.
  mutable Tensor draws_dev;
  // Rank each direction's projector kept in the last step (left, right, top,
  // bottom): the rank a graph-safe step truncates at (the cutoff rule is
  // replayed on its spectrum afterwards).
  mutable std::array<int64_t, 4> kept{{0, 0, 0, 0}};
};

// Boundary initialisation C_i, E_i from A and Ad (src/ctm/ctm_tensor.jl:97-130).
void init_ctm1(Ctx& c, Env& e);

// Projectors (P1, P2) and spectrum of one move from the plain environment:
// dir 0 left, 1 right, 2 top, 3 bottom (src/ctm/ctmrg.jl:122-265).
void projector(Ctx& c, const Env& e, int dir, int64_t chi, Tensor& P1, Tensor& P2, std::vector<double>& s);

// One CTMRG step left -> right -> top -> bottom (src/ctm/ctmrg.jl:57-64).
// Returns the left move's truncated spectrum.
std::vector<double> rg_step(Ctx& c, Env& e, int64_t chi);

// A captured rg_step of one environment (HIP graph).  The graph reads the
// environment's C / E buffers, saves them into backups, runs the step in
// graph-safe mode and writes the new C / E back into the same buffers, so a
// replay advances the environment in place; the step's deferred checks land
// in `checks` and are copied to pinned host memory after each replay.  A
// replay whose checks fail is undone from the backups, and the step is then
// run eagerly (the caller's decision).  Capture needs stable shapes (chi
// saturated), unpaired moves and the randomized projector with k <= 112.
struct StepGraph {
  Ctx* c = nullptr;
  int64_t chi = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<std::pair<size_t, void*>> blocks;  // pool blocks the graph replays into
  std::array<Tensor, 4> C, E;        // the environment buffers the graph reads and rewrites
  std::array<Tensor, 4> bakC, bakE;
  Tensor checks;
  std::vector<Ctx::CheckRec> recs;
  double* host_checks = nullptr;  // pinned
  int64_t nchecks = 0;
  ~StepGraph();
};
// Captures one step of e (nothing runs; e is unchanged).
std::unique_ptr<StepGraph> capture_step(Ctx& c, Env& e, int64_t chi);
// Enqueues one replay (e.draws advances).
void replay_step(StepGraph& g, Env& e);
// After the stream has finished a replay: validates its checks; on failure
// restores the pre-step environment, returns false and says why.
bool finish_step(StepGraph& g, Env& e, std::string* why);
// After an eager step of e with a graph attached: the new C / E go back into
// the graph's buffers (dims must match) and the device draw counter follows.
void adopt_step(StepGraph& g, Env& e);

// Two-site density matrices rho_h, rho_v (d^2 x d^2)
// (src/graph/evaluation.jl:124-248, ground-state branch).
void get_dms(Ctx& c, const Env& e, Tensor& roh, Tensor& rov);

// E = tr(H1 roh)/tr(roh) + tr(H2 rov)/tr(rov), Na = (tr roh + tr rov)/2
// (src/graph/evaluation.jl:15-29).  H1/H2: device C128 (m x m).  Returns a
// device FP64[8] tensor [ReE, ImE, ReNa, ImNa, tr roh (re, im), tr rov (re, im)];
// when recording, its node expects the seed dL/dRe(E) = 1.
Tensor energy(Ctx& c, const Env& e, const Tensor& H1, const Tensor& H2);

// <Ad|single-site environment|A> (src/graph/evaluation.jl:47-59, 110-122).
void gs_norm(Ctx& c, const Env& e, double out[2]);

// 2 x 2 plaquette energy_norm_4x4 with A1..A4 = A (src/graph/evaluation.jl:
// 340-355, 443-519) for the four-site Hamiltonian factors s1 (d,d,r1),
// s2 (r1,d,d,r2), s3 (r2,d,d,r3), s4 (r3,d,d) of get_local_h: out = [Re, Im]
// of energy, then of norm (get_gs_energy_4x4 = energy / norm).  Computed in
// FP64 (complex when the environment or a factor is).
void energy_4x4(Ctx& c, const Env& e, const std::array<Tensor, 4>& s, double out[4]);

struct RunResult {
  int iters = 0;
  std::vector<double> diffs;
  double E[2] = {0, 0};
  std::vector<double> s_last;
};

// run_ctm with the reference stopping rule (src/ctm/ctmrg.jl:1-55).
// conv_mode 0: energy (conv_fun = get_gs_energy), 1: left spectrum.
RunResult run_ctm(Ctx& c, Env& e, int64_t chi, int min_iter, int max_iter, double tol, int conv_mode,
                  const Tensor& H1, const Tensor& H2, bool differentiable_tail = false);

// Gradient contract of optim_gs's fg! (src/optimize/optim_gs.jl:53-91, 120-143):
// x = renormalize(A0); converge from init_ctm1(x); then run ad steps from the
// converged environment held constant and differentiate Re E w.r.t. x.
// Step-level checkpointing keeps one step's tape alive at a time.
struct GradResult {
  double E[2] = {0, 0};
  int conv_iters = 0, ad_iters = 0;
  Tensor grad;  // same dims/dtype as A (F64 or C128; F32 env returns F64 grad)
};
GradResult gs_energy_grad(Ctx& c, const Tensor& A0, DType env_dt, const Options& opt, const Tensor& H1,
                          const Tensor& H2, int64_t chi, int min_iter, int max_iter, int ad_max_iter, double tol);
// run_gs and its pullback (src/optimize/optim_gs.jl:133-143): from env0
// (held constant, not modified) with A = x, Ad = conj(x), run_ctm's stopping
// rule with max_iter = ad_max_iter, E = get_gs_energy; grad = dRe(E)/dx.
// gs_energy_grad is renormalize + init_ctm1 + run_ctm + this.
GradResult run_gs_vjp(Ctx& c, const Env& env0, const Tensor& x, const Tensor& H1, const Tensor& H2, int64_t chi,
                      int min_iter, int ad_max_iter, double tol);

// ---------------------------------------------------------------------------
// Excited states: nested tensors [X, X_B, X_Bd, X_BB]
// (src/nested_tensor/nested_tensor.jl, src/ctm/ctm_tensor.jl:20-55)
// ---------------------------------------------------------------------------
struct NT {
  std::array<Tensor, 4> c;  // undefined Tensor = EmptyT
  bool nested = true;
};

// Projectors along nested runs.  A product's first component involves only
// first components (nested_tensor.jl:48-63), so the plain parts of a nested
// environment evolve as they would with B = EmptyT, whatever B and the
// momentum are, and move t (= 4 step + dir) of every nested run started from
// the same environment uses the same projector (ctmrg.jl:122-265).  A path
// holds the projectors of the first moves, computed once by a run with
// B = EmptyT; runs that carry it take theirs from it.
struct ProjPath {
  Env start;                                // the environment the runs start from
  int64_t chi = 0;
  std::vector<std::array<Tensor, 2>> P;     // (P1, P2) of move t
  std::vector<std::vector<double>> s;       // truncated spectrum of move t
};

struct NestedEnv {
  Env g;                        // plain (ground-state) parts: A, Ad, C, E
  Tensor B, Bd;                 // tangent tensor and its conjugate
  std::array<Tensor, 4> BC, BdC, BBC, BE, BdE, BBE;
  double px = 0.0, py = 0.0;    // momentum phases (already multiplied by pi)
  int64_t moves = 0;            // moves made since the start
  std::shared_ptr<const ProjPath> path;  // projectors of moves < path->P.size()
  ProjPath* record = nullptr;   // when set, each move's projector is appended to it
};

// The projectors of the first `steps` steps of every nested run from g at
// bond dimension chi (a run with B = EmptyT).  The path reads g's tensors
// and is valid for environments whose state equals g's (path_matches).
std::shared_ptr<const ProjPath> nested_proj_path(Ctx& c, const Env& g, int64_t chi, int steps);
// Whether nested runs from env at bond dimension chi may take their
// projectors from path: same tensors (identical storage or bitwise equal),
// test-matrix position and engine options.
bool path_matches(Ctx& c, const ProjPath& path, const Env& env, int64_t chi);

// Nested rg_step: projectors from the plain parts, nested renormalisation and
// the momentum shift of up_* (src/ctm/ctmrg.jl:66-120, 267-491).
std::vector<double> nested_rg_step(Ctx& c, NestedEnv& e, int64_t chi);
// get_es_energy (src/graph/evaluation.jl:31-45): Re E[4]; when recording the
// returned FP64[4] tensor's node expects the seed dL/dRe(E4) = 1.
Tensor es_energy(Ctx& c, const NestedEnv& e, const Tensor& H1, const Tensor& H2);
// get_all_norm (src/graph/evaluation.jl:61-77): Nb (re, im) and envB.
void all_norm(Ctx& c, const NestedEnv& e, double Nb[2], Tensor& envB);

struct EsGrad {
  Tensor gH, envB;   // (D,D,D,D,d) C128: d Re E4 / dB and the B environment
  double E4 = 0.0, Nb[2] = {0, 0};
  int conv_iters = 0;
};
// get_es_grad (src/optimize/optim_es.jl:167-194) from a converged, normalised
// ground-state environment g and one tangent tensor B.
// With `path` (path_matches(g, chi) must hold) the moves it covers take its
// projectors instead of recomputing them.
EsGrad es_grad(Ctx& c, const Env& g, const Tensor& B, const Tensor& H1, const Tensor& H2, double px, double py,
               int64_t chi, int min_iter, int max_iter, int ad_max_iter, double tol,
               std::shared_ptr<const ProjPath> path = nullptr);

// One differentiated step from a given environment: forward rg_step on the
// tape, the energy and its VJP, then the step VJP; the A cotangent is
// returned (the benchmark unit "step + adjoint").
Tensor step_and_adjoint(Ctx& c, Env& e, int64_t chi, const Tensor& H1, const Tensor& H2, double E_out[2]);

}  // namespace ctm
}  // namespace ipeps
