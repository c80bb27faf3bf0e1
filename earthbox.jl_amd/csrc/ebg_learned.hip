// Learned correction network: placeholder until the network kernels land.
#include "ebg_common.hpp"
#include "ebg_learned.hpp"

namespace ebg {
struct LearnedCorrection {
    int unused;
};
void learned_apply(LearnedCorrection*, ebg_ctx*, const double*, double*, hipStream_t) {}
void learned_release(LearnedCorrection* lc) { delete lc; }
}  // namespace ebg
