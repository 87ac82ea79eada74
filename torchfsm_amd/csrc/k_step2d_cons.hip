// 2D ConservativeConvection kernels, power-of-two sizes.
#include "k_step2d.hpp"

namespace tfsm {

hipError_t launch_col2d_cons(const Col2Args& a, int nb, int stage, int v, int var, hipStream_t st) {
  constexpr int TERM = TERM_CONS;
  TFSM_COL_P2(TFSM_COL_CASE)
  return hipErrorInvalidValue;
}

hipError_t launch_row2d_cons(const Row2Args& a, int var, int wgt, hipStream_t st) {
  constexpr int TERM = TERM_CONS;
  TFSM_ROW_P2(TFSM_ROW_CASE)
  return hipErrorInvalidValue;
}

}  // namespace tfsm
