// 2D convection kernels, power-of-two sizes (the rows also serve the velocity-form
// Navier-Stokes term, whose products are Convection's).
#include "k_step2d.hpp"

namespace tfsm {

hipError_t launch_col2d_conv(const Col2Args& a, int nb, int stage, int v, int var, hipStream_t st) {
  constexpr int TERM = TERM_CONV;
  TFSM_COL_P2(TFSM_COL_CASE)
  return hipErrorInvalidValue;
}

hipError_t launch_row2d_conv(const Row2Args& a, int var, int wgt, hipStream_t st) {
  constexpr int TERM = TERM_CONV;
  TFSM_ROW_P2(TFSM_ROW_CASE)
  return hipErrorInvalidValue;
}

hipError_t launch_run2d_conv(const Col2Args& ca, const Row2Args& ra, const Run2Args& r, hipStream_t st) {
  constexpr int TERM = TERM_CONV;
  TFSM_RUN_SIZES(TFSM_RUN_CASE)
  return hipErrorInvalidValue;
}

}  // namespace tfsm
