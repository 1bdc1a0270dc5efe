// Fused unbounded pipeline (placeholder until the fused kernels land).
#include <hip/hip_runtime.h>

#include "solver.hpp"

namespace flups {

struct FastPlan {};
FastPlan* make_fast_plan(SolverImpl& s) {
    (void)s;
    return nullptr;
}
void destroy_fast_plan(FastPlan* f) { delete f; }
void fast_solve(FastPlan* f, double* field, const double* rhs, const View& uin, const View& uout, hipStream_t s) {
    (void)f; (void)field; (void)rhs; (void)uin; (void)uout; (void)s;
}

}  // namespace flups
