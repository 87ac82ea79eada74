// 2D velocity-form Navier-Stokes column kernels, power-of-two sizes.
#include "k_step2d.hpp"

namespace tfsm {

hipError_t launch_col2d_nsp(const Col2Args& a, int nb, int stage, int v, int var, hipStream_t st) {
  constexpr int TERM = TERM_NSP;
  TFSM_COL_P2(TFSM_COL_CASE)
  return hipErrorInvalidValue;
}

}  // namespace tfsm
