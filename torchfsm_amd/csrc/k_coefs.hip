// Setup kernels: linear symbol L per mode (float32, the reference's op order), ETD coefficient
// tables (SETDRK: float64 contour means; ETDRK: the reference's complex64 closed forms; complex L:
// complex tables), and the accumulation of explicit source spectra.
#include "engine.hpp"

namespace tfsm {



// L = sum_t coef_t * sym_t in generator order (_base.py:339-357), each product and sum rounded in
// float32: lap = (-(a0 a0)) + (-(a1 a1)) [+ (-(a2 a2))] (mesh.py:406-426), biharm = lap * lap,
// unit = 1 (_source.py:9-17), SpatialDerivative (i a_d)^n (mesh.py:399-404: n = 1 the value,
// n = 2 and 3 complex products, higher n from double).  The reference's symbols are complex64;
// only odd derivative orders give them a nonzero imaginary part, so L is written as its real part
// (and, for complex-L plans, as complex).  Padding modes (k >= H) get 0.
__device__ __forceinline__ float2 deriv_symbol(float a, int n) {
  if (n == 1) return make_float2(0.f, a);
  if (n == 2) return make_float2(-mulr(a, a), 0.f);
  if (n == 3) return make_float2(0.f, -mulr(mulr(a, a), a));
  const double p = pow((double)a, (double)n);
  switch (n & 3) {
    case 0: return make_float2((float)p, 0.f);
    case 1: return make_float2(0.f, (float)p);
    case 2: return make_float2((float)-p, 0.f);
    default: return make_float2(0.f, (float)-p);
  }
}

__global__ void k_linear_symbol(LinArgs a) {
  const size_t total = a.modes * a.Bc;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t mode = e % a.modes;
    const int b = (int)(e / a.modes);
    const int k = (int)(mode % a.P);
    size_t r = mode / a.P;
    int idx[3] = {0, 0, 0};
    idx[a.D - 1] = k;
    if (a.D == 3) {
      idx[1] = (int)(r % a.n[1]);
      r /= a.n[1];
    }
    if (a.D >= 2) idx[0] = (int)r;
    float2 out = make_float2(0.f, 0.f);
    if (k < a.H) {
      float lap = 0.f;
      for (int d = 0; d < a.D; ++d) {
        const float av = a.kv[d][(size_t)b * a.kv_bstride + idx[d]];
        const float g = -mulr(av, av);
        lap = (d == 0) ? g : addr(lap, g);
      }
      for (int t = 0; t < a.nterm; ++t) {
        const float c = a.coef ? a.coef[t * a.Bc + b] : a.cval[t];
        float2 sym = make_float2(1.f, 0.f);
        if (a.kind[t] == LIN_LAP) sym.x = lap;
        else if (a.kind[t] == LIN_BIHARM) sym.x = mulr(lap, lap);
        else if (a.kind[t] == LIN_DERIV)
          sym = deriv_symbol(a.kv[a.dim[t]][(size_t)b * a.kv_bstride + idx[a.dim[t]]], a.order[t]);
        const float2 v = make_float2(mulr(c, sym.x), mulr(c, sym.y));
        out = (t == 0) ? v : radd(out, v);
      }
    }
    a.L[e] = out.x;
    if (a.Lc) a.Lc[e] = out;
  }
}

// ---------------------------------------------------------------- complex double helpers
struct cd {
  double x, y;
};
__device__ __forceinline__ cd cd_add(cd a, cd b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cd cd_sub(cd a, cd b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cd cd_mul(cd a, cd b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ cd cd_scale(cd a, double s) { return {a.x * s, a.y * s}; }
__device__ __forceinline__ cd cd_div(cd a, cd b) {
  const double d = b.x * b.x + b.y * b.y;
  return {(a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d};
}
__device__ __forceinline__ cd cd_exp(cd a) {
  const double e = exp(a.x);
  double s, c;
  sincos(a.y, &s, &c);
  return {e * c, e * s};
}



// SETDRK1-4: Kassam-Trefethen contour means in float64 (_uncached.py:34-43, 71-72, 102-104,
// 135-156, 191-211).  z = fl32(L * fl32(dt)) is the reference's float32 centre; the contour points
// lr_j = r e^{i theta_j} + z with theta_j = 2 pi (j - 1/2) / M, the mean is taken in double and
// rounded once.  E = exp(z), E2 = exp(fl32(L * fl32(dt/2))) in double, rounded once.
__global__ void k_setdrk_coefs(CoefArgs a) {
  const size_t total = a.modes * a.Bc;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t mode = e % a.modes;
    const size_t b = e / a.modes;
    const float l = a.L[e];
    const double z = (double)mulr(l, a.dt32);
    const double z2 = (double)mulr(l, a.hdt32);
    float* out = a.coef + b * (size_t)NCOEF * a.modes + mode;
    out[CS_E * a.modes] = (float)exp(z);
    if (a.scheme == SETDRK3 || a.scheme == SETDRK4) out[CS_E2 * a.modes] = (float)exp(z2);
    cd s1 = {0, 0}, s2 = {0, 0}, s3 = {0, 0}, s4 = {0, 0}, s5 = {0, 0};
    for (int jj = 1; jj <= a.M; ++jj) {
      const double th = 2.0 * M_PI * (jj - 0.5) / a.M;
      cd lr = {a.radius * cos(th) + z, a.radius * sin(th)};
      const cd one = {1.0, 0.0};
      const cd el = cd_exp(lr);
      const cd lr2 = cd_mul(lr, lr);
      const cd lr3 = cd_mul(lr2, lr);
      switch (a.scheme) {
        case SETDRK1:
          s1 = cd_add(s1, cd_div(cd_sub(el, one), lr));
          break;
        case SETDRK2:
          s1 = cd_add(s1, cd_div(cd_sub(el, one), lr));
          s2 = cd_add(s2, cd_div(cd_sub(cd_sub(el, one), lr), lr2));
          break;
        case SETDRK3: {
          const cd eh = cd_exp(cd_scale(lr, 0.5));
          s1 = cd_add(s1, cd_div(cd_sub(eh, one), lr));
          s2 = cd_add(s2, cd_div(cd_sub(el, one), lr));
          // (-4 - lr + e^lr (4 - 3 lr + lr^2)) / lr^3
          cd t = cd_add(cd_sub({4.0, 0.0}, cd_scale(lr, 3.0)), lr2);
          s3 = cd_add(s3, cd_div(cd_add(cd_sub({-4.0, 0.0}, lr), cd_mul(el, t)), lr3));
          // 4 (2 + lr + e^lr (-2 + lr)) / lr^3
          t = cd_add({-2.0, 0.0}, lr);
          s4 = cd_add(s4, cd_div(cd_scale(cd_add(cd_add({2.0, 0.0}, lr), cd_mul(el, t)), 4.0), lr3));
          // (-4 - 3 lr - lr^2 + e^lr (4 - lr)) / lr^3
          t = cd_sub({4.0, 0.0}, lr);
          s5 = cd_add(s5, cd_div(cd_add(cd_sub(cd_sub({-4.0, 0.0}, cd_scale(lr, 3.0)), lr2), cd_mul(el, t)), lr3));
          break;
        }
        case SETDRK4: {
          const cd eh = cd_exp(cd_scale(lr, 0.5));
          s1 = cd_add(s1, cd_div(cd_sub(eh, one), lr));
          cd t = cd_add(cd_sub({4.0, 0.0}, cd_scale(lr, 3.0)), lr2);
          s2 = cd_add(s2, cd_div(cd_add(cd_sub({-4.0, 0.0}, lr), cd_mul(el, t)), lr3));
          t = cd_add({-2.0, 0.0}, lr);
          s3 = cd_add(s3, cd_div(cd_add(cd_add({2.0, 0.0}, lr), cd_mul(el, t)), lr3));
          t = cd_sub({4.0, 0.0}, lr);
          s4 = cd_add(s4, cd_div(cd_add(cd_sub(cd_sub({-4.0, 0.0}, cd_scale(lr, 3.0)), lr2), cd_mul(el, t)), lr3));
          break;
        }
        default:
          break;
      }
    }
    const double w = a.dt / a.M;
    switch (a.scheme) {
      case SETDRK1:
        out[CS_C1 * a.modes] = (float)(w * s1.x);
        break;
      case SETDRK2:
        out[CS_C1 * a.modes] = (float)(w * s1.x);
        out[CS_C2 * a.modes] = (float)(w * s2.x);
        break;
      case SETDRK3:
        out[CS_C1 * a.modes] = (float)(w * s1.x);
        out[CS_C2 * a.modes] = (float)(w * s2.x);
        out[CS_C3 * a.modes] = (float)(w * s3.x);
        out[CS_C4 * a.modes] = (float)(w * s4.x);
        out[CS_C5 * a.modes] = (float)(w * s5.x);
        break;
      case SETDRK4:
        out[CS_C1 * a.modes] = (float)(w * s1.x);
        out[CS_C4 * a.modes] = (float)(w * s2.x);
        out[CS_C5 * a.modes] = (float)(w * s3.x);
        out[CS_C6 * a.modes] = (float)(w * s4.x);
        break;
      default:
        break;
    }
  }
}

// ETDRK0/1/2: the reference's complex64 closed forms (_etdrk.py:21, 43-45, 66-70) with a real L:
// E = exp(fl32(dt L)) (double exp, rounded once: the reference's float32 exp is correctly rounded
// on these inputs), c1 = where(L == 0, dt, (E - 1) * (1 / L)), c2 = where(L == 0, dt/2,
// ((E - 1) - L dt) * (1 / (L^2 dt))).  complex64 division by a value with zero imaginary part is a
// multiplication by the reciprocal in torch, kept here; every op is rounded in float32, so the
// float32 cancellation of c2 at small |dt L| is reproduced, not repaired.
__global__ void k_etdrk_coefs(CoefArgs a) {
  const size_t total = a.modes * a.Bc;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t mode = e % a.modes;
    const size_t b = e / a.modes;
    const float l = a.L[e];
    const float ldt = mulr(l, a.dt32);
    const float E = (float)exp((double)ldt);
    float* out = a.coef + b * (size_t)NCOEF * a.modes + mode;
    out[CS_E * a.modes] = E;
    if (a.scheme == ETDRK0) continue;
    const float em1 = subr(E, 1.0f);
    out[CS_C1 * a.modes] = (l == 0.f) ? a.dt32 : mulr(em1, __fdiv_rn(1.0f, l));
    if (a.scheme == ETDRK2) {
      const float num = subr(em1, ldt);
      const float den = mulr(mulr(l, l), a.dt32);
      out[CS_C2 * a.modes] = (l == 0.f) ? (float)(a.dt / 2) : mulr(num, __fdiv_rn(1.0f, den));
    }
  }
}

// Complex-L plans (odd-order SpatialDerivative, e.g. KdV): complex E = exp(fl32(dt L)) and
// E2 = exp(fl32(dt/2 L)) (double, rounded once); SETDRK c's are the real parts of the contour
// means, as the reference takes .real of them (_uncached.py:72, 103-104, 137-156, 193-211);
// ETDRK c1 = where(L == 0, dt, (E - 1) / L), c2 = where(L == 0, dt/2, ((E - 1) - L dt) / ((L L) dt))
// in complex64 ops (_etdrk.py:43-45, 66-70).
__device__ __forceinline__ float2 cexp_rounded(float2 z) {
  const double e = exp((double)z.x);
  double s, c;
  sincos((double)z.y, &s, &c);
  return make_float2((float)(e * c), (float)(e * s));
}

__global__ void k_coefs_c(CoefArgs a) {
  const size_t total = a.modes * a.Bc;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t mode = e % a.modes;
    const size_t b = e / a.modes;
    const float2 l = a.Lc[e];
    const float2 z = make_float2(mulr(l.x, a.dt32), mulr(l.y, a.dt32));
    float2* out = a.coefc + b * (size_t)NCOEF * a.modes + mode;
    const float2 E = cexp_rounded(z);
    out[CS_E * a.modes] = E;
    if (a.scheme == ETDRK0) continue;
    const bool l0 = l.x == 0.f && l.y == 0.f;
    if (a.scheme == ETDRK1 || a.scheme == ETDRK2) {
      const float2 em1 = make_float2(subr(E.x, 1.0f), E.y);
      out[CS_C1 * a.modes] = l0 ? make_float2(a.dt32, 0.f) : cdivr(em1, l);
      if (a.scheme == ETDRK2) {
        const float2 num = rsub(em1, z);
        const float2 ll = cmulr(l, l);
        const float2 den = make_float2(mulr(ll.x, a.dt32), mulr(ll.y, a.dt32));
        out[CS_C2 * a.modes] = l0 ? make_float2((float)(a.dt / 2), 0.f) : cdivr(num, den);
      }
      continue;
    }
    const float2 z2 = make_float2(mulr(l.x, a.hdt32), mulr(l.y, a.hdt32));
    if (a.scheme == SETDRK3 || a.scheme == SETDRK4) out[CS_E2 * a.modes] = cexp_rounded(z2);
    cd s1 = {0, 0}, s2 = {0, 0}, s3 = {0, 0}, s4 = {0, 0}, s5 = {0, 0};
    const cd one = {1.0, 0.0};
    for (int jj = 1; jj <= a.M; ++jj) {
      const double th = 2.0 * M_PI * (jj - 0.5) / a.M;
      const cd lr = {a.radius * cos(th) + (double)z.x, a.radius * sin(th) + (double)z.y};
      const cd el = cd_exp(lr);
      const cd lr2 = cd_mul(lr, lr);
      const cd lr3 = cd_mul(lr2, lr);
      const cd eh = cd_exp(cd_scale(lr, 0.5));
      cd t;
      switch (a.scheme) {
        case SETDRK1:
          s1 = cd_add(s1, cd_div(cd_sub(el, one), lr));
          break;
        case SETDRK2:
          s1 = cd_add(s1, cd_div(cd_sub(el, one), lr));
          s2 = cd_add(s2, cd_div(cd_sub(cd_sub(el, one), lr), lr2));
          break;
        case SETDRK3:
          s1 = cd_add(s1, cd_div(cd_sub(eh, one), lr));
          s2 = cd_add(s2, cd_div(cd_sub(el, one), lr));
          t = cd_add(cd_sub({4.0, 0.0}, cd_scale(lr, 3.0)), lr2);
          s3 = cd_add(s3, cd_div(cd_add(cd_sub({-4.0, 0.0}, lr), cd_mul(el, t)), lr3));
          t = cd_add({-2.0, 0.0}, lr);
          s4 = cd_add(s4, cd_div(cd_scale(cd_add(cd_add({2.0, 0.0}, lr), cd_mul(el, t)), 4.0), lr3));
          t = cd_sub({4.0, 0.0}, lr);
          s5 = cd_add(s5, cd_div(cd_add(cd_sub(cd_sub({-4.0, 0.0}, cd_scale(lr, 3.0)), lr2), cd_mul(el, t)), lr3));
          break;
        default:  // SETDRK4
          s1 = cd_add(s1, cd_div(cd_sub(eh, one), lr));
          t = cd_add(cd_sub({4.0, 0.0}, cd_scale(lr, 3.0)), lr2);
          s2 = cd_add(s2, cd_div(cd_add(cd_sub({-4.0, 0.0}, lr), cd_mul(el, t)), lr3));
          t = cd_add({-2.0, 0.0}, lr);
          s3 = cd_add(s3, cd_div(cd_add(cd_add({2.0, 0.0}, lr), cd_mul(el, t)), lr3));
          t = cd_sub({4.0, 0.0}, lr);
          s4 = cd_add(s4, cd_div(cd_add(cd_sub(cd_sub({-4.0, 0.0}, cd_scale(lr, 3.0)), lr2), cd_mul(el, t)), lr3));
          break;
      }
    }
    const double w = a.dt / a.M;
    auto re = [&](cd v) { return make_float2((float)(w * v.x), 0.f); };
    switch (a.scheme) {
      case SETDRK1: out[CS_C1 * a.modes] = re(s1); break;
      case SETDRK2: out[CS_C1 * a.modes] = re(s1); out[CS_C2 * a.modes] = re(s2); break;
      case SETDRK3:
        out[CS_C1 * a.modes] = re(s1);
        out[CS_C2 * a.modes] = re(s2);
        out[CS_C3 * a.modes] = re(s3);
        out[CS_C4 * a.modes] = re(s4);
        out[CS_C5 * a.modes] = re(s5);
        break;
      default:
        out[CS_C1 * a.modes] = re(s1);
        out[CS_C4 * a.modes] = re(s2);
        out[CS_C5 * a.modes] = re(s3);
        out[CS_C6 * a.modes] = re(s4);
        break;
    }
  }
}

// Pack the slot tables into the per-mode layout read by the fused column kernel (see engine.hpp).
__global__ void k_pack_coefs(CoefArgs a) {
  const size_t total = a.modes * a.Bc;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t mode = e % a.modes;
    const size_t b = e / a.modes;
    const float* t = a.coef + b * (size_t)NCOEF * a.modes + mode;
    auto S = [&](int slot) { return t[(size_t)slot * a.modes]; };
    float4 h0 = make_float4(0.f, 0.f, 0.f, 0.f), h1 = h0;
    switch (a.scheme) {
      case ETDRK0: h0 = make_float4(S(CS_E), 0.f, 0.f, 0.f); break;
      case ETDRK1: case SETDRK1: h0 = make_float4(S(CS_E), S(CS_C1), 0.f, 0.f); break;
      case ETDRK2: case SETDRK2: h0 = make_float4(S(CS_E), S(CS_C1), S(CS_C2), 0.f); break;
      case SETDRK3:
        h0 = make_float4(S(CS_E2), S(CS_C1), S(CS_E), S(CS_C2));
        h1 = make_float4(S(CS_E), S(CS_C3), S(CS_C4), S(CS_C5));
        break;
      case SETDRK4:  // the third evaluation forms the partial sum (E U + c4 N0) + 2 c5 (Na + Nb) too
        h0 = make_float4(S(CS_E2), S(CS_C1), S(CS_E), S(CS_C4));
        h1 = make_float4(S(CS_E), S(CS_C4), S(CS_C5), S(CS_C6));
        break;
      default: break;
    }
    a.cpk[cpk_index(e, 0)] = h0;
    a.cpk[cpk_index(e, 1)] = h1;
    if (a.scheme == SETDRK4) {
      a.c5t[e] = S(CS_C5);
      if (a.ck01) a.ck01[e] = make_float2(S(CS_E2), S(CS_C1));
      if (a.c6t) a.c6t[e] = S(CS_C6);
    }
  }
}

// S <- S + coef * X  (first source: S = coef * X); X and S are half spectra of the same shape.
__global__ void k_accum_source(float2* __restrict__ S, const float2* __restrict__ X, float coef, size_t total,
                               int first) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const float2 v = rscale(coef, X[e]);
    S[e] = first ? v : radd(S[e], v);
  }
}

// out = L * U + N   (Operator.__call__ via _build_operator, _base.py:408-439); L per mode, real.
// have_lin == 2: out = where(L == 0, 1, 1 / L) * U   (LinearOperator.solve, _base.py:250-258).
__global__ void k_apply_linear(const float2* __restrict__ U, const float2* __restrict__ Nv, const float* __restrict__ L,
                               size_t l_bstride, size_t modes, int C, size_t total, int have_lin, int have_nl,
                               float2* __restrict__ out) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t mode = e % modes;
    const size_t b = e / modes / C;
    float2 r = make_float2(0.f, 0.f);
    if (have_lin == 2) {  // solve: where(L == 0, 1, 1 / L) * U
      const float l = L[b * l_bstride + mode];
      out[e] = rscale(l == 0.f ? 1.0f : __fdiv_rn(1.0f, l), U[e]);
      continue;
    }
    if (have_lin) r = rscale(L[b * l_bstride + mode], U[e]);
    if (have_nl) r = have_lin ? radd(r, Nv[e]) : Nv[e];
    out[e] = r;
  }
}

// complex-L plans: out = L * U + N, or (solve) where(L == 0, 1, 1 / L) * U with complex64 ops
__global__ void k_apply_linear_c(const float2* __restrict__ U, const float2* __restrict__ Nv,
                                 const float2* __restrict__ L, size_t l_bstride, size_t modes, int C, size_t total,
                                 int have_lin, int have_nl, float2* __restrict__ out) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t mode = e % modes;
    const size_t b = e / modes / C;
    const float2 l = L[b * l_bstride + mode];
    if (have_lin == 2) {
      const float2 inv = (l.x == 0.f && l.y == 0.f) ? make_float2(1.f, 0.f) : cdivr(make_float2(1.f, 0.f), l);
      out[e] = cmulr(inv, U[e]);
      continue;
    }
    float2 r = make_float2(0.f, 0.f);
    if (have_lin) r = cmulr(l, U[e]);
    if (have_nl) r = have_lin ? radd(r, Nv[e]) : Nv[e];
    out[e] = r;
  }
}

static unsigned grid_for(size_t total) {
  size_t b = (total + 255) / 256;
  if (b > 65535u * 4) b = 65535u * 4;
  return (unsigned)(b ? b : 1);
}

hipError_t launch_linear_symbol(const LinArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(k_linear_symbol, dim3(grid_for(a.modes * a.Bc)), dim3(256), 0, st, a);
  return hipGetLastError();
}

// axis wavenumbers of one axis table row: a_i = fl32(2 pi) * (fl32(k_i) * s), s = fl32(1 / (n d)),
// the same float32 operations as the host table (fftfreq, mesh.py:185-191; grad, mesh.py:399-404)
__global__ void k_axis_row(float* kv, int n, float s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = i < (n + 1) / 2 ? i : i - n;
  const float f = (float)k * s;
  kv[i] = (float)(2.0 * M_PI) * f;
}

hipError_t launch_axis_row(float* kv, int n, float s, hipStream_t st) {
  hipLaunchKernelGGL(k_axis_row, dim3((n + 255) / 256), dim3(256), 0, st, kv, n, s);
  return hipGetLastError();
}

hipError_t launch_apply_linear_c(const float2* U, const float2* Nv, const float2* L, size_t l_bstride, size_t modes,
                                 int C, size_t total, int have_lin, int have_nl, float2* out, hipStream_t st) {
  hipLaunchKernelGGL(k_apply_linear_c, dim3(grid_for(total)), dim3(256), 0, st, U, Nv, L, l_bstride, modes, C, total,
                     have_lin, have_nl, out);
  return hipGetLastError();
}

hipError_t launch_coefs(const CoefArgs& a, hipStream_t st) {
  if (a.Lc)
    hipLaunchKernelGGL(k_coefs_c, dim3(grid_for(a.modes * a.Bc)), dim3(256), 0, st, a);
  else if (a.scheme >= SETDRK1)
    hipLaunchKernelGGL(k_setdrk_coefs, dim3(grid_for(a.modes * a.Bc)), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL(k_etdrk_coefs, dim3(grid_for(a.modes * a.Bc)), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_pack_coefs(const CoefArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(k_pack_coefs, dim3(grid_for(a.modes * a.Bc)), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_accum_source(float2* S, const float2* X, float coef, size_t total, int first, hipStream_t st) {
  hipLaunchKernelGGL(k_accum_source, dim3(grid_for(total)), dim3(256), 0, st, S, X, coef, total, first);
  return hipGetLastError();
}

// Runge-Kutta stage combination out = x + dt * ((c0 k0 + c1 k1) + ...), in the order of the
// reference's python sum (_rk.py:43-58: sum([a_i * k ...]) starts from 0, then dt * ..., then x + ...)
__global__ void k_rk_combine(RkArgs a) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < a.total; e += (size_t)gridDim.x * blockDim.x) {
    float2 acc = make_float2(0.f, 0.f);
    for (int i = 0; i < a.nk; ++i) {
      const float2 t = rscale(a.c[i], a.k[i][e]);
      acc = i == 0 ? t : radd(acc, t);
    }
    a.out[e] = radd(a.x[e], rscale(a.dt, acc));
  }
}

// batch mean of the product DC (one workgroup, fixed summation order: deterministic)
__global__ void __launch_bounds__(256) k_ks_dc(const float2* __restrict__ P, int B, size_t bstride, int nx,
                                               size_t xstride, float2* __restrict__ out) {
  __shared__ float2 red[256];
  float2 acc = make_float2(0.f, 0.f);
  const int total = B * nx;
  for (int i = threadIdx.x; i < total; i += 256) {
    const int b = i / nx, x = i % nx;
    acc = radd(acc, P[(size_t)b * bstride + (size_t)x * xstride]);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = radd(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = make_float2(__fdiv_rn(red[0].x, (float)B), __fdiv_rn(red[0].y, (float)B));
}

hipError_t launch_ks_dc(const float2* P, int B, size_t bstride, int nx, size_t xstride, float2* out, hipStream_t st) {
  hipLaunchKernelGGL(k_ks_dc, dim3(1), dim3(256), 0, st, P, B, bstride, nx, xstride, out);
  return hipGetLastError();
}

hipError_t launch_rk_combine(const RkArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(k_rk_combine, dim3(grid_for(a.total)), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_apply_linear(const float2* U, const float2* Nv, const float* L, size_t l_bstride, size_t modes, int C,
                               size_t total, int have_lin, int have_nl, float2* out, hipStream_t st) {
  hipLaunchKernelGGL(k_apply_linear, dim3(grid_for(total)), dim3(256), 0, st, U, Nv, L, l_bstride, modes, C, total,
                     have_lin, have_nl, out);
  return hipGetLastError();
}

}  // namespace tfsm
