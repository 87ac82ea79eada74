// Unfused nonlinear evaluation and stage combine (layouts in engine.hpp, GenArgs).
//
// Used where no fused kernel covers the plan: any mesh size (the transforms between these launches
// run the fixed-size kernels where an axis has them and the general-length FFTs elsewhere),
// ConservativeConvection in 3D, and a complex linear symbol together with a product term in 2D / 3D.
// The terms follow the reference's closures: dealias the input (_base.py:359-406), then
//   Convection            sum_j u_j d_j u_c                        (_convection.py:18-46)
//   VorticityConvection   u . grad w via the streamfunction        (_navier_stokes.py:27-46)
//   NSPressureConvection  grad p - conv, p = invlap div conv       (_navier_stokes.py:221-262)
//   KSConvection          1/2 |grad phi|^2 (minus its mean)        (_ks_convection.py:18-38)
//   ConservativeConvection sum_i d_i (u_i u_c)                     (_conservative_convection.py:17-27)
// and the stage combine is the fused kernels' per-mode one (stage_combine / stage_combine_c).
#include "engine.hpp"

namespace tfsm {
namespace {

unsigned gen_grid(size_t total) {
  size_t g = (total + 255) / 256;
  if (g > 65536) g = 65536;
  return (unsigned)(g ? g : 1);
}

__device__ __forceinline__ int gsk(int i, int n) { return i < (n + 1) / 2 ? i : i - n; }

// mode m of the half spectrum: axis wavenumbers a_d, and whether the mode is a stored frequency
// (last index < H) and inside the dealiasing box
struct GMode {
  float av[3];
  bool ny[3];  // the index along axis d is that axis' Nyquist frequency (n_d even, index n_d / 2)
  bool stored, keep;
};

__device__ __forceinline__ GMode gen_mode(const GenArgs& a, size_t m, size_t blk) {
  GMode g;
  const int il = (int)(m % a.P);
  const size_t r = m / a.P;
  int i0 = 0, i1 = 0;
  if (a.D == 2) i0 = (int)r;
  if (a.D == 3) {
    i1 = (int)(r % a.n[1]);
    i0 = (int)(r / a.n[1]);
  }
  g.stored = il < a.H;
  const int ilc = g.stored ? il : 0;
  g.keep = g.stored && il <= a.kmax[a.D - 1];
  g.av[0] = g.av[1] = g.av[2] = 0.f;
  g.ny[0] = g.ny[1] = g.ny[2] = false;
  const int nl = a.n[a.D - 1];
  g.ny[a.D - 1] = (nl % 2 == 0) && ilc == nl / 2;
  if (a.D == 1) {
    g.av[0] = a.kv[0][blk * a.kv_bstride + ilc];
  } else if (a.D == 2) {
    g.av[0] = a.kv[0][i0];
    g.av[1] = a.kv[1][ilc];
    g.ny[0] = (a.n[0] % 2 == 0) && i0 == a.n[0] / 2;
    g.keep = g.keep && abs(gsk(i0, a.n[0])) <= a.kmax[0];
  } else {
    g.av[0] = a.kv[0][i0];
    g.av[1] = a.kv[1][i1];
    g.av[2] = a.kv[2][ilc];
    g.ny[0] = (a.n[0] % 2 == 0) && i0 == a.n[0] / 2;
    g.ny[1] = (a.n[1] % 2 == 0) && i1 == a.n[1] / 2;
    g.keep = g.keep && abs(gsk(i0, a.n[0])) <= a.kmax[0] && abs(gsk(i1, a.n[1])) <= a.kmax[1];
  }
  // a mode that is its own conjugate partner keeps its full value (nsp_project, engine.hpp)
  const bool own = (ilc == 0 || g.ny[a.D - 1]) && (a.D < 2 || i0 == 0 || g.ny[0]) && (a.D < 3 || i1 == 0 || g.ny[1]);
  if (own) g.ny[0] = g.ny[1] = g.ny[2] = false;
  return g;
}

// i a z
__device__ __forceinline__ float2 gik(float a, float2 z) { return make_float2(-mulr(a, z.y), mulr(a, z.x)); }

__global__ void k_gen_fields(GenArgs a, float2* __restrict__ S, float2* __restrict__ F) {
  const size_t total = (size_t)a.B * a.modes;
  const float2 z = make_float2(0.f, 0.f);
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t b = e / a.modes, m = e % a.modes;
    const GMode g = gen_mode(a, m, b * a.C);
    float2* Fb = F + b * a.NF * a.modes + m;
    float2* Sb = S + b * a.C * a.modes + m;
    if (!g.stored) {
      for (int f = 0; f < a.NF; ++f) Fb[(size_t)f * a.modes] = z;
      continue;
    }
    // the forced NSPressureConvection core masks the integrator's stage tensor in place
    // (_navier_stokes.py:249-250)
    if (a.mask_in_place && !g.keep)
      for (int c = 0; c < a.C; ++c) Sb[(size_t)c * a.modes] = z;
    switch (a.term) {
      case TERM_CONV:
      case TERM_NSP:
        for (int c = 0; c < a.C; ++c) {
          const float2 d = g.keep ? Sb[(size_t)c * a.modes] : z;
          Fb[(size_t)c * a.modes] = d;
          for (int j = 0; j < a.D; ++j) Fb[(size_t)(a.C + c * a.D + j) * a.modes] = gik(g.av[j], d);
        }
        break;
      case TERM_VC: {
        // psi = -w invlap = w / |k|^2; u = (d_y psi, -d_x psi)
        const float2 d = g.keep ? Sb[0] : z;
        const float k2 = addr(mulr(g.av[0], g.av[0]), mulr(g.av[1], g.av[1]));
        const float2 psi = k2 > 0.f ? make_float2(__fdiv_rn(d.x, k2), __fdiv_rn(d.y, k2)) : z;
        Fb[0] = gik(g.av[1], psi);
        Fb[a.modes] = gik(-g.av[0], psi);
        Fb[2 * a.modes] = gik(g.av[0], d);
        Fb[3 * a.modes] = gik(g.av[1], d);
        break;
      }
      case TERM_KS: {
        const float2 d = g.keep ? Sb[0] : z;
        for (int j = 0; j < a.D; ++j) Fb[(size_t)j * a.modes] = gik(g.av[j], d);
        break;
      }
      default:  // TERM_CONS
        for (int c = 0; c < a.C; ++c) Fb[(size_t)c * a.modes] = g.keep ? Sb[(size_t)c * a.modes] : z;
        break;
    }
  }
}

__global__ void k_gen_products(GenArgs a, const float* __restrict__ R, float* __restrict__ Q) {
  const size_t total = (size_t)a.B * a.npts;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t b = e / a.npts, x = e % a.npts;
    const float* Rb = R + b * a.NF * a.npts + x;
    float* Qb = Q + b * a.NP * a.npts + x;
    switch (a.term) {
      case TERM_CONV:
      case TERM_NSP:
        for (int c = 0; c < a.C; ++c) {
          float s = mulr(Rb[0], Rb[(size_t)(a.C + c * a.D) * a.npts]);
          for (int j = 1; j < a.D; ++j) s = addr(s, mulr(Rb[(size_t)j * a.npts], Rb[(size_t)(a.C + c * a.D + j) * a.npts]));
          Qb[(size_t)c * a.npts] = s;
        }
        break;
      case TERM_VC:
        Qb[0] = addr(mulr(Rb[0], Rb[2 * a.npts]), mulr(Rb[a.npts], Rb[3 * a.npts]));
        break;
      case TERM_KS: {
        float s = mulr(Rb[0], Rb[0]);
        for (int j = 1; j < a.D; ++j) s = addr(s, mulr(Rb[(size_t)j * a.npts], Rb[(size_t)j * a.npts]));
        Qb[0] = mulr(0.5f, s);
        break;
      }
      default: {  // TERM_CONS: u_i u_j, i <= j, row-major over the upper triangle
        int p = 0;
        for (int i = 0; i < a.C; ++i)
          for (int j = i; j < a.C; ++j) Qb[(size_t)(p++) * a.npts] = mulr(Rb[(size_t)i * a.npts], Rb[(size_t)j * a.npts]);
        break;
      }
    }
  }
}

__device__ __forceinline__ int cons_pair(int i, int j, int C) {
  if (i > j) {
    const int t = i;
    i = j;
    j = t;
  }
  return i * C - i * (i - 1) / 2 + (j - i);
}

__global__ void k_gen_stage(GenArgs a, GenStageArgs s) {
  const size_t total = (size_t)a.B * a.modes;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t b = e / a.modes, m = e % a.modes;
    if (s.mask.skip((int)b)) continue;
    const GMode g = gen_mode(a, m, b * a.C);
    if (!g.stored) continue;
    const float2* Qb = s.Qs + b * a.NP * a.modes + m;
    float2 nv[3];
    switch (a.term) {
      case TERM_CONV:
        for (int c = 0; c < a.C; ++c) nv[c] = Qb[(size_t)c * a.modes];
        break;
      case TERM_NSP:
        // the projection of nsp_project (engine.hpp), with the Nyquist-plane parity; an external
        // force f enters as conv - f before the projection and + f after (_navier_stokes.py:251-260)
        if (s.force) {
          const float2* fb = s.force + b * s.force_bstride + m;
          for (int c = 0; c < a.C; ++c) nv[c] = rsub(Qb[(size_t)c * a.modes], fb[(size_t)c * s.force_cstride]);
        } else {
          for (int c = 0; c < a.C; ++c) nv[c] = Qb[(size_t)c * a.modes];
        }
        if (a.C == 2) {
          float2 cv[2] = {nv[0], nv[1]};
          const float av[2] = {g.av[0], g.av[1]};
          const bool ny[2] = {g.ny[0], g.ny[1]};
          nsp_project<2>(cv, av, ny);
          nv[0] = cv[0];
          nv[1] = cv[1];
        } else {
          float2 cv[3] = {nv[0], nv[1], nv[2]};
          const float av[3] = {g.av[0], g.av[1], g.av[2]};
          const bool ny[3] = {g.ny[0], g.ny[1], g.ny[2]};
          nsp_project<3>(cv, av, ny);
          nv[0] = cv[0];
          nv[1] = cv[1];
          nv[2] = cv[2];
        }
        if (s.force) {
          const float2* fb = s.force + b * s.force_bstride + m;
          for (int c = 0; c < a.C; ++c) nv[c] = radd(nv[c], fb[(size_t)c * s.force_cstride]);
        }
        break;
      case TERM_KS:
        nv[0] = Qb[0];
        if (m == 0 && s.dc) nv[0] = rsub(nv[0], s.dc[0]);
        break;
      case TERM_VC:
        nv[0] = Qb[0];
        break;
      default:  // TERM_CONS: N_c = sum_i (i a_i) (u_i u_c)^
        // on axis i's Nyquist plane the Hermitian part of (i a_i) X is 0 (nsp_project, engine.hpp)
        for (int c = 0; c < a.C; ++c) {
          float2 acc = gik(g.ny[0] ? 0.f : g.av[0], Qb[(size_t)cons_pair(0, c, a.C) * a.modes]);
          for (int i = 1; i < a.D; ++i)
            acc = radd(acc, gik(g.ny[i] ? 0.f : g.av[i], Qb[(size_t)cons_pair(i, c, a.C) * a.modes]));
          nv[c] = acc;
        }
        break;
    }
    for (int c = 0; c < a.C; ++c) {
      const size_t blk = b * a.C + c;
      const size_t ei = blk * a.modes + m;
      const float2* src = s.src ? s.src + b * s.src_bstride + c * s.src_cstride : nullptr;
      const float2 n = nl_value(a.coef_t, nv[c], src, m);
      const size_t cb = (s.coef_per_block ? blk : b) * s.coef_bstride;
      if (s.masked_stage && !g.keep && (s.stage == ST_R2_E1 || s.stage == ST_R4_E2)) {
        // the reference reuses the stage spectrum its forced core masked in place: outside the
        // box the recomputed stage value a = E U + c1 A1 (ST_R2_E1) or E2 U + c1 A1 (ST_R4_E2)
        // is zero there (U itself was zeroed by this step's first evaluation)
        const float* cf = s.coef + cb;
        const float2 n0 = s.A1[ei];
        if (s.stage == ST_R2_E1) {
          const float2 un = rscale(cf[(size_t)CS_C2 * a.modes + m], rsub(n, n0));
          s.U[ei] = un;
          s.S[ei] = un;
        } else {
          const float2 u = s.U[ei], na = s.A2[ei];
          const float c1 = cf[(size_t)CS_C1 * a.modes + m];
          const float c5x2 = 2.0f * cf[(size_t)CS_C5 * a.modes + m];
          const float2 two_nb = make_float2(2.0f * n.x, 2.0f * n.y);
          s.A2[ei] = radd(radd(rscale(cf[(size_t)CS_E * a.modes + m], u), rscale(cf[(size_t)CS_C4 * a.modes + m], n0)),
                          rscale(c5x2, radd(na, n)));
          s.S[ei] = rscale(c1, rsub(two_nb, n0));
        }
        continue;
      }
      if (s.coefc)
        s.S[ei] = stage_combine_c(s.stage, n, s.U, s.A1, s.A2, ei, s.coefc + cb, a.modes, m);
      else
        s.S[ei] = stage_combine(s.stage, n, s.U, s.A1, s.A2, ei, s.coef + cb, a.modes, m, true);
    }
  }
}

}  // namespace

hipError_t launch_gen_fields(const GenArgs& a, float2* S, float2* F, hipStream_t st) {
  hipLaunchKernelGGL(k_gen_fields, dim3(gen_grid((size_t)a.B * a.modes)), dim3(256), 0, st, a, S, F);
  return hipGetLastError();
}

hipError_t launch_gen_products(const GenArgs& a, const float* R, float* Q, hipStream_t st) {
  hipLaunchKernelGGL(k_gen_products, dim3(gen_grid((size_t)a.B * a.npts)), dim3(256), 0, st, a, R, Q);
  return hipGetLastError();
}

hipError_t launch_gen_stage(const GenArgs& a, const GenStageArgs& s, hipStream_t st) {
  if (a.C > 3) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_gen_stage, dim3(gen_grid((size_t)a.B * a.modes)), dim3(256), 0, st, a, s);
  return hipGetLastError();
}

}  // namespace tfsm
