// C entry point over the reference's Util_LinEqs (LibSrc/Util/Util_LinEqs.h),
// so tests can compare the oracle's Gauss-Jordan restatement with the
// reference's own compiled code.  Test infrastructure only.
#include "Util_LinEqs.h"

extern "C" int ref_lin_eqs(float* coeff, float* vals, int dim) {
    Util_LinEqs solver;
    return solver.DoIt(coeff, vals, dim) ? 1 : 0;
}
