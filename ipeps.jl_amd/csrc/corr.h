// Transfer-matrix spectrum and spin correlations (src/graph/expectation.jl:193-361).
#pragma once

#include <vector>

#include "ctmrg.h"

namespace ipeps {
namespace ctm {

// correlation_spin (expectation.jl:246-271) for direction 0 (h) or 1 (v):
// s2s receives max_iter complex values (re, im interleaved), m1 / m2 the
// one-operator contractions.  op1 / op2: d x d C128 on the device.  Runs in
// complex FP64 whatever the environment's dtype.
void correlation_spin(Ctx& c, const Env& e, const Tensor& op1, const Tensor& op2, int max_iter, int dir,
                      std::vector<double>& s2s, double m1[2], double m2[2]);

// Transfer matrix of correlation_TM (expectation.jl:193-220): TM reshaped to
// (rows, cols) column-major, its matrix-free product with X (cols x nvec), and
// its shape.  The product and the dense matrix use the E tensors' dtype (F64
// for mixed environments).
void tm_dims(const Env& e, int dir, int64_t& rows, int64_t& cols);
Tensor tm_dense(Ctx& c, const Env& e, int dir);
Tensor tm_apply(Ctx& c, const Env& e, int dir, const Tensor& X);

}  // namespace ctm
}  // namespace ipeps
