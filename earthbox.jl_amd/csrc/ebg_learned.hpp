// Learned coarse-grid / smoother correction applied inside the FGMRES
// preconditioner (see ebg_learned.hip).
#pragma once
#include <hip/hip_runtime.h>

struct ebg_ctx;

namespace ebg {
struct LearnedCorrection;
// z += network(r, log viscosity) in the preconditioner's output units
void learned_apply(LearnedCorrection* lc, ebg_ctx* ctx, const double* r, double* z, hipStream_t s);
void learned_release(LearnedCorrection* lc);
}  // namespace ebg
