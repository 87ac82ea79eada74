// Dispatch of the 2D fused ETD evaluation kernels (k_step2d.hpp) over term, size and launch shape.
#include "k_step2d.hpp"

namespace tfsm {

hipError_t launch_col2d(int term, const Col2Args& a, int nb, int stage, int v, int var, hipStream_t st) {
  if (is_mixed_size_2d(a.N0)) return launch_col2d_mixed(term, a, nb, stage, v, var, st);
  if (term == TERM_VC) return launch_col2d_vc(a, nb, stage, v, var, st);
  if (term == TERM_CONV) return launch_col2d_conv(a, nb, stage, v, var, st);
  if (term == TERM_NSP) return launch_col2d_nsp(a, nb, stage, v, var, st);
  if (term == TERM_KS) return launch_col2d_ks(a, nb, stage, v, var, st);
  if (term == TERM_CONS) return launch_col2d_cons(a, nb, stage, v, var, st);
  return hipErrorInvalidValue;
}

// the velocity-form Navier-Stokes term has the products of Convection (its projection is applied
// in the column kernel), so it shares Convection's row kernels
hipError_t launch_row2d(int term, const Row2Args& a, int var, int wgt, hipStream_t st) {
  if (is_mixed_size_2d(a.N1)) return launch_row2d_mixed(term, a, var, wgt, st);
  if (term == TERM_VC) return launch_row2d_vc(a, var, wgt, st);
  if (term == TERM_CONV || term == TERM_NSP) return launch_row2d_conv(a, var, wgt, st);
  if (term == TERM_KS) return launch_row2d_ks(a, var, wgt, st);
  if (term == TERM_CONS) return launch_row2d_cons(a, var, wgt, st);
  return hipErrorInvalidValue;
}

hipError_t launch_run2d(int term, const Col2Args& ca, const Row2Args& ra, const Run2Args& r, hipStream_t st) {
  if (term == TERM_VC) return launch_run2d_vc(ca, ra, r, st);
  if (term == TERM_CONV) return launch_run2d_conv(ca, ra, r, st);
  return hipErrorInvalidValue;
}

bool run2d_supported(int term, int n0, int n1, int pf) {
  // fields at the pitch of the 2/3 rule or narrower (k2_run holds them in LDS)
  return (term == TERM_VC || term == TERM_CONV) && n0 == n1 && (n0 == 32 || n0 == 64) &&
         pf <= ((n0 / 3 + 1 + 15) / 16) * 16;
}

bool col2d_supported(int n0, int v, int var) {
  if (!(v == 8 || v == 16)) return false;
  if (n0 == 16 || n0 == 32 || is_mixed_size_2d(n0)) return var == 0;
  if (n0 == 1024) return v == 8 && (var == 0 || var == 1);  // 16 columns of 1024 points exceed the LDS
  return (n0 == 64 || n0 == 128 || n0 == 256 || n0 == 512) && (var == 0 || var == 1);
}
bool row2d_supported(int n1, int var, int wgt) {
  if (n1 == 16 || n1 == 32 || is_mixed_size_2d(n1)) return var == 0 && wgt == 256;
  if (n1 == 1024) return var == 0 && (wgt == 64 || wgt == 256);
  return (n1 == 64 || n1 == 128 || n1 == 256 || n1 == 512) && (var == 0 || var == 1) && (wgt == 64 || wgt == 256);
}

}  // namespace tfsm
