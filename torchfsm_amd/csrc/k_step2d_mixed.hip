// 2D kernels of every term at the mixed-radix sizes.
#include "k_step2d.hpp"

namespace tfsm {

template <int TERM>
static hipError_t col_mixed(const Col2Args& a, int nb, int stage, int v, int var, hipStream_t st) {
  TFSM_COL_MX(TFSM_COL_CASE)
  return hipErrorInvalidValue;
}

template <int TERM>
static hipError_t row_mixed(const Row2Args& a, int var, int wgt, hipStream_t st) {
  TFSM_ROW_MX(TFSM_ROW_CASE)
  return hipErrorInvalidValue;
}

hipError_t launch_col2d_mixed(int term, const Col2Args& a, int nb, int stage, int v, int var, hipStream_t st) {
  if (term == TERM_VC) return col_mixed<TERM_VC>(a, nb, stage, v, var, st);
  if (term == TERM_CONV) return col_mixed<TERM_CONV>(a, nb, stage, v, var, st);
  if (term == TERM_NSP) return col_mixed<TERM_NSP>(a, nb, stage, v, var, st);
  if (term == TERM_KS) return col_mixed<TERM_KS>(a, nb, stage, v, var, st);
  if (term == TERM_CONS) return col_mixed<TERM_CONS>(a, nb, stage, v, var, st);
  return hipErrorInvalidValue;
}

hipError_t launch_row2d_mixed(int term, const Row2Args& a, int var, int wgt, hipStream_t st) {
  if (term == TERM_VC) return row_mixed<TERM_VC>(a, var, wgt, st);
  if (term == TERM_CONV || term == TERM_NSP) return row_mixed<TERM_CONV>(a, var, wgt, st);
  if (term == TERM_KS) return row_mixed<TERM_KS>(a, var, wgt, st);
  if (term == TERM_CONS) return row_mixed<TERM_CONS>(a, var, wgt, st);
  return hipErrorInvalidValue;
}

}  // namespace tfsm
