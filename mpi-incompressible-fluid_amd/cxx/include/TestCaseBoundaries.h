// Lid-driven cavity boundary data of the two driver test cases
// (include/TestCaseBoundaries.h:18-56 of the reference): v = 1 on the lid
// plane x = 1 (case 1) or x = -0.5 (case 2), zero elsewhere, within a 1e-12
// tolerance.  Named functions, so the velocity shim recognises them and the
// device evaluates the same data (MIF_FN_LID_1 / MIF_FN_LID_2).
#ifndef MIF_SHIM_TEST_CASE_BOUNDARIES_H
#define MIF_SHIM_TEST_CASE_BOUNDARIES_H

#include "Real.h"

namespace mif {

constexpr Real exact_solution_precision = 1e-12;

Real exact_u_t1(Real t, Real x, Real y, Real z);
Real exact_v_t1(Real t, Real x, Real y, Real z);
Real exact_w_t1(Real t, Real x, Real y, Real z);
Real exact_p_initial_t1(Real x, Real y, Real z);
Real exact_u_t2(Real t, Real x, Real y, Real z);
Real exact_v_t2(Real t, Real x, Real y, Real z);
Real exact_w_t2(Real t, Real x, Real y, Real z);
Real exact_p_initial_t2(Real x, Real y, Real z);

}  // namespace mif

#endif  // MIF_SHIM_TEST_CASE_BOUNDARIES_H
