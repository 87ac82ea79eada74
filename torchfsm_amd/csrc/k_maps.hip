// Spectral maps of the diagnostic operators (single applications, no time stepping): Div, Grad,
// Curl, Vorticity2Velocity, the pressure of Vorticity2Pressure / Velocity2Pressure, the
// divergence of ConservativeConvection's products, and the dealias mask; plus the physical-space
// products u_i u_j of ConservativeConvection.  Per mode, in the reference's complex64 op order
// (multiplication by i a is exact: (0 + i a)(x + i y) = (-a y) + i (a x)).
#include "engine.hpp"

namespace tfsm {

// (i a) z
__device__ __forceinline__ float2 ia_mul(float a, float2 z) { return make_float2(-mulr(a, z.y), mulr(a, z.x)); }

__global__ void k_spectral_map(MapArgs a) {
  const size_t total = (size_t)a.B * a.modes;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t b = e / a.modes, mode = e % a.modes;
    const int k = (int)(mode % a.P);
    size_t r = mode / a.P;
    int idx[3] = {0, 0, 0};
    idx[a.D - 1] = k;
    if (a.D == 3) {
      idx[1] = (int)(r % a.n[1]);
      r /= a.n[1];
    }
    if (a.D >= 2) idx[0] = (int)r;
    float av[3] = {0.f, 0.f, 0.f};
    bool keep = true;
    for (int d = 0; d < a.D; ++d) {
      av[d] = a.kv[d][idx[d]];
      const int ks = d == a.D - 1 ? idx[d] : (idx[d] <= a.n[d] / 2 ? idx[d] : a.n[d] - idx[d]);
      keep = keep && ks <= a.kmax[d];
    }
    const bool valid = k < a.H;
    const float2* in = a.in + (size_t)b * a.cin * a.modes + mode;
    float2* out = a.out + (size_t)b * a.cout * a.modes + mode;
    auto IN = [&](int c) { return in[(size_t)c * a.modes]; };
    auto OUT = [&](int c, float2 v) {
      out[(size_t)c * a.modes] = valid ? make_float2(mulr(a.coef, v.x), mulr(a.coef, v.y)) : make_float2(0.f, 0.f);
    };
    // -(1 / lap) (1 at lap = 0), lap = sum_d -(a_d a_d) in the reference's order (mesh.py:406-419)
    auto neg_invlap = [&]() {
      float lap = -mulr(av[0], av[0]);
      for (int d = 1; d < a.D; ++d) lap = addr(lap, -mulr(av[d], av[d]));
      return -(lap == 0.f ? 1.0f : __fdiv_rn(1.0f, lap));
    };
    switch (a.op) {
      case MAP_DEALIAS:  // u^ * low_pass_filter (mesh.py:443-461)
        for (int c = 0; c < a.cin; ++c) OUT(c, keep ? IN(c) : make_float2(0.f, 0.f));
        break;
      case MAP_DIV: {  // sum_i (i a_i) u_i   (_div.py:17-28)
        float2 s = ia_mul(av[0], IN(0));
        for (int i = 1; i < a.D; ++i) s = radd(s, ia_mul(av[i], IN(i)));
        OUT(0, s);
        break;
      }
      case MAP_GRAD:  // (i a_i) p   (_grad.py:12-15)
        for (int i = 0; i < a.D; ++i) OUT(i, ia_mul(av[i], IN(0)));
        break;
      case MAP_CURL:
        if (a.D == 2) {  // (i a_0) u_1 - (i a_1) u_0   (_curl.py:18-27)
          OUT(0, rsub(ia_mul(av[0], IN(1)), ia_mul(av[1], IN(0))));
        } else {  // (_curl.py:39-56)
          OUT(0, rsub(ia_mul(av[1], IN(2)), ia_mul(av[2], IN(1))));
          OUT(1, rsub(ia_mul(av[2], IN(0)), ia_mul(av[0], IN(2))));
          OUT(2, rsub(ia_mul(av[0], IN(1)), ia_mul(av[1], IN(0))));
        }
        break;
      case MAP_V2V: {  // w * (-invlap) [i a_1, -(i a_0)]   (_navier_stokes.py:75-91)
        const float ni = neg_invlap();
        const float2 w = IN(0);
        const float c0 = mulr(ni, av[1]), c1 = mulr(ni, -av[0]);
        OUT(0, make_float2(-mulr(w.y, c0), mulr(w.x, c0)));
        OUT(1, make_float2(-mulr(w.y, c1), mulr(w.x, c1)));
        break;
      }
      case MAP_PRESSURE: {  // (-1 * sum_i (i a_i) conv_i) * invlap   (_navier_stokes.py:140-144, 202-206)
        float2 s = ia_mul(av[0], IN(0));
        for (int i = 1; i < a.D; ++i) s = radd(s, ia_mul(av[i], IN(i)));
        const float il = -neg_invlap();
        OUT(0, make_float2(mulr(-s.x, il), mulr(-s.y, il)));
        break;
      }
      case MAP_CONSDIV:  // sum_i (i a_i) X[i][c]   (_conservative_convection.py:23-27)
        for (int c = 0; c < a.cout; ++c) {
          float2 s = ia_mul(av[0], IN(c));
          for (int i = 1; i < a.D; ++i) s = radd(s, ia_mul(av[i], IN(i * a.cout + c)));
          OUT(c, s);
        }
        break;
      default:
        break;
    }
  }
}

// out[b][i][j][x] = u[b][i][x] * u[b][j][x]   (_conservative_convection.py:25)
__global__ void k_pointwise_products(const float* __restrict__ u, float* __restrict__ out, int B, int C,
                                     size_t npts) {
  const size_t total = (size_t)B * C * C * npts;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t x = e % npts;
    size_t r = e / npts;
    const int j = (int)(r % C);
    r /= C;
    const int i = (int)(r % C);
    const size_t b = r / C;
    out[e] = mulr(u[(b * C + i) * npts + x], u[(b * C + j) * npts + x]);
  }
}

static unsigned grid_of(size_t total) {
  size_t g = (total + 255) / 256;
  if (g > 65535u * 4) g = 65535u * 4;
  return (unsigned)(g ? g : 1);
}

hipError_t launch_spectral_map(const MapArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(k_spectral_map, dim3(grid_of((size_t)a.B * a.modes)), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_pointwise_products(const float* u, float* out, int B, int C, size_t npts, hipStream_t st) {
  hipLaunchKernelGGL(k_pointwise_products, dim3(grid_of((size_t)B * C * C * npts)), dim3(256), 0, st, u, out, B, C,
                     npts);
  return hipGetLastError();
}

}  // namespace tfsm
