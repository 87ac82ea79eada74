// 2D fused ETD evaluation kernels (one nonlinear evaluation = one row kernel + one column kernel).
// Templates only: the k_step2d_*.hip units instantiate them per term and size.
//
// Column kernel (axis 0, strided columns, tile of V adjacent ky):
//   forward x-FFT of the y-transformed products  ->  N^ = coef_t * F + S
//   -> stage combine (reads/writes U, A1, A2)    ->  s = next stage spectrum
//   -> dealiased fields of s (VC: psi, kx psi, kx w, w;  Conv C=2: u0, kx u0, u1, kx u1)
//   -> inverse x-FFT of each field (only tiles with ky <= kmax1)  ->  G
// Row kernel (axis 1, contiguous rows):
//   packed C2R of the fields (two real fields per complex transform, symbols i ky / i applied
//   here)  ->  physical product  ->  R2C  ->  products P (all H1 modes; the output is not masked)
//
// Reference path replaced: the nonlinear closure _base.py:375-403 with _VorticityConvectionCore
// (_navier_stokes.py:27-46), _ConvectionCore (_convection.py:18-46), _NSPressureConvectionCore
// (_navier_stokes.py:221-262) or _KSConvectionCore (_ks_convection.py:8-38), and the stage combines of
// _setdrk_step.py / _etdrk.py.
#pragma once
#include "engine.hpp"

namespace tfsm {

// physical-space pair packing of two Hermitian half spectra A, B into one full spectrum
// Z = A + iB (positions <= N/2) or conj(A) + i conj(B) (mirror positions); at k = 0 and N/2
// only the real parts of A and B enter (C2R semantics: the reference takes .real afterwards).
__device__ __forceinline__ float2 pack_pair(float2 A, float2 Bv, bool cj, bool edge) {
  if (edge) {
    A.y = 0.f;
    Bv.y = 0.f;
  }
  return cj ? make_float2(A.x + Bv.y, Bv.x - A.y) : make_float2(A.x - Bv.y, A.y + Bv.x);
}

template <int TERM>
struct Term2;
template <>
struct Term2<TERM_VC> {
  static constexpr int NC = 1, NF = 4;
};
template <>
struct Term2<TERM_CONV> {
  static constexpr int NC = 2, NF = 4;
};
template <>
struct Term2<TERM_NSP> {  // velocity-form Navier-Stokes: Convection fields, projected spectrum
  static constexpr int NC = 2, NF = 4;
};
template <>
struct Term2<TERM_KS> {  // KSConvection: fields phi^, a_x phi^; one product 1/2 |grad phi|^2
  static constexpr int NC = 1, NF = 2;
};
template <>
struct Term2<TERM_CONS> {  // ConservativeConvection, C = 2: fields u0, u1; product planes per
  static constexpr int NC = 2, NF = 2;  // channel c: F_y(u0 u_c) and i a_y F_y(u1 u_c)
};

// Stage combine of one channel's column (see combine_column in engine.hpp): column base ky,
// element stride P1.  With two channels the other channel's values stay live, so those kernels load
// one mode's operands at a time (BM = 1) to stay within 128 VGPRs.
template <int N0, int VAR, int BM>
__device__ __forceinline__ void combine_column2(const Col2Args& a, int stage, size_t base, const float2* src,
                                                const float4* cpk, const float* c5b, int cko, int ky, bool colv, int j,
                                                const float2 (&v)[Fft<N0, VAR>::RMAX],
                                                float2 (&s)[Fft<N0, VAR>::RMAX]) {
  using F = Fft<N0, VAR>;
  const size_t cb = base + ky;
  combine_column<F::RMAX, F::NT, BM>(stage, a.U + cb, a.A1 + cb, a.A2 + cb, src ? src + ky : nullptr,
                                     cpk + (2 * ky + cko), c5b, (unsigned)ky, a.coef_t, (unsigned)a.P1, j, colv, v,
                                     s);
}

// Velocity-form Navier-Stokes (NSPressureConvection): both channels' convection spectra are
// transformed first, projected mode by mode (nsp_project), then combined channel by channel.
template <int N0, int V, int VAR>
__device__ __forceinline__ void col_nsp_combine(const Col2Args& a, int stage, float2 (&s)[2][Fft<N0, VAR>::RMAX],
                                                float2* s_buf, const float2* s_tw, int b, int ky, bool colv, int j,
                                                int vi) {
  using F = Fft<N0, VAR>;
  constexpr int RMAX = F::RMAX, NT = F::NT;
  const size_t plane = (size_t)N0 * a.P1;
  const float2 z = make_float2(0.f, 0.f);
  if (stage == ST_PRO) {
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int m = 0; m < RMAX; ++m) {
        const float2 t = a.U[((size_t)b * 2 + c) * plane + (size_t)(j + m * NT) * a.P1 + ky];
        s[c][m] = colv ? t : z;
      }
    return;
  }
  float2 v[2][RMAX];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
#pragma unroll
    for (int m = 0; m < RMAX; ++m) {
      const float2 t = a.P[((size_t)b * 2 + c) * plane + (size_t)(j + m * NT) * a.P1 + ky];
      v[c][m] = colv ? t : z;
    }
    fft_wg<N0, -1, V, VAR>(v[c], s_buf, s_tw, j, vi);
  }
  const float ayv = colv ? a.ay[ky] : 0.f;
#pragma unroll
  for (int m = 0; m < RMAX; ++m) {
    float2 cv[2] = {v[0][m], v[1][m]};
    const float av[2] = {a.ax[j + m * NT], ayv};
    nsp_project<2>(cv, av);
    v[0][m] = cv[0];
    v[1][m] = cv[1];
  }
  const float4* cpk = a.cpk + (size_t)b * a.cpk_bstride;
  const float* c5b = a.c5t + (size_t)b * (a.cpk_bstride / 2);
  // cpk_index(ci, half) = 2 ci + cko: every mode of this lane has ci & 15 == ky & 15
  const int cko = (stage_coef_half(stage) << 4) - (ky & 15);
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const float2* src = a.src ? a.src + (size_t)b * a.src_bstride + (size_t)c * a.src_cstride : nullptr;
    combine_column2<N0, VAR, 1>(a, stage, ((size_t)b * 2 + c) * plane, src, cpk, c5b, cko, ky, colv, j, v[c], s[c]);
  }
}

template <int N0, int V, int TERM, int VAR>
__global__ void __launch_bounds__((Fft<N0, VAR>::NT * V)) k2_col(Col2Args a, int stage) {
  using F = Fft<N0, VAR>;
  constexpr int RMAX = F::RMAX, NT = F::NT;
  constexpr int NC = Term2<TERM>::NC, NF = Term2<TERM>::NF;
  __shared__ float2 s_tw[Fft<N0, VAR>::tw_size()];
  __shared__ float2 s_buf[LdsSize<N0, V>::slots];
  __shared__ float s_ax[N0];  // x wavenumbers: read per mode by the field pass without a memory wait
  load_twiddles<N0, VAR>(s_tw, a.tw0);
  for (int i = threadIdx.x; i < N0; i += NT * V) s_ax[i] = a.ax[i];
  __syncthreads();
  const int vi = threadIdx.x % V, j = threadIdx.x / V;
  // Block -> (tile, sample).  Workgroups are dispatched round-robin over the 8 XCDs (block v on
  // XCD v % 8); with nb % 8 == 0 each XCD owns a group of nb/8 samples and walks the tiles in
  // order, so a coefficient tile is fetched once into that XCD's L2 and reused by its samples, and
  // the heavy tiles (those with kept columns, ky <= kmax1) are dispatched before the light ones.
  const int ntiles = (a.H1 + V - 1) / V;
  int tile, bl;
  if ((a.nb & 7) == 0) {
    const int x = blockIdx.x & 7, sidx = blockIdx.x >> 3, g = a.nb >> 3;
    tile = sidx / g;
    bl = x * g + sidx % g;
  } else {
    tile = blockIdx.x % ntiles;
    bl = blockIdx.x / ntiles;
  }
  const int b = a.b0 + bl;
  if (a.mask.skip(b)) return;  // uniform per workgroup: this sample's steps are done
  const int ky = tile * V + vi;
  const bool colv = ky < a.H1;
  const size_t plane = (size_t)N0 * a.P1;
  const float4* cpk = a.cpk + (size_t)b * a.cpk_bstride;
  const float* c5b = a.c5t + (size_t)b * (a.cpk_bstride / 2);
  // cpk_index(ci, half) = 2 ci + cko: every mode of this lane has ci & 15 == ky & 15
  const int cko = (stage_coef_half(stage) << 4) - (ky & 15);
  float2 s[NC][RMAX];
  if constexpr (TERM == TERM_NSP) {
    col_nsp_combine<N0, V, VAR>(a, stage, s, s_buf, s_tw, b, ky, colv, j, vi);
  } else {
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const size_t base = ((size_t)b * NC + c) * plane;
    if (stage == ST_PRO) {
#pragma unroll
      for (int m = 0; m < RMAX; ++m) {
        const float2 t = a.U[base + (size_t)(j + m * NT) * a.P1 + ky];
        s[c][m] = colv ? t : make_float2(0.f, 0.f);
      }
    } else {
      float2 v[RMAX];
      if constexpr (TERM == TERM_CONS) {
        // N_c = (i a_x) F_x(F_y(u0 u_c)) + F_x(i a_y F_y(u1 u_c))   (_conservative_convection.py:23-27)
        const size_t pb = ((size_t)b * 4 + 2 * c) * plane;
#pragma unroll
        for (int m = 0; m < RMAX; ++m) {
          const float2 t = a.P[pb + (size_t)(j + m * NT) * a.P1 + ky];
          v[m] = colv ? t : make_float2(0.f, 0.f);
        }
        fft_wg<N0, -1, V, VAR>(v, s_buf, s_tw, j, vi);
#pragma unroll
        for (int m = 0; m < RMAX; ++m) {
          const float axm = a.ax[j + m * NT];
          v[m] = make_float2(-mulr(axm, v[m].y), mulr(axm, v[m].x));
        }
        float2 w[RMAX];
#pragma unroll
        for (int m = 0; m < RMAX; ++m) {
          const float2 t = a.P[pb + plane + (size_t)(j + m * NT) * a.P1 + ky];
          w[m] = colv ? t : make_float2(0.f, 0.f);
        }
        fft_wg<N0, -1, V, VAR>(w, s_buf, s_tw, j, vi);
#pragma unroll
        for (int m = 0; m < RMAX; ++m) v[m] = radd(v[m], w[m]);
      } else {
#pragma unroll
        for (int m = 0; m < RMAX; ++m) {
          const float2 t = a.P[base + (size_t)(j + m * NT) * a.P1 + ky];
          v[m] = colv ? t : make_float2(0.f, 0.f);
        }
        fft_wg<N0, -1, V, VAR>(v, s_buf, s_tw, j, vi);
      }
      if constexpr (TERM == TERM_KS) {
        // remove_mean: fft(p - mean(p)) differs from fft(p) only at k = 0, by the batch mean of the DC
        if (a.dc && ky == 0 && j == 0) v[0] = rsub(v[0], a.dc[0]);
      }
      const float2* src = a.src ? a.src + (size_t)b * a.src_bstride + (size_t)c * a.src_cstride : nullptr;
      combine_column2<N0, VAR, (NC > 1 || RMAX >= 16) ? 1 : 4>(a, stage, base, src, cpk, c5b, cko, ky, colv, j, v, s[c]);
    }
  }
  }
  if (tile * V > a.kmax1) return;  // uniform per workgroup: no kept column in this tile
  const float ay = a.ay[ky];  // ky < P1 <= N1: in the table
  (void)ay;
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    float2 v[RMAX];
#pragma unroll
    for (int m = 0; m < RMAX; ++m) {
      const int kx = j + m * NT;
      const int ks = kx <= N0 / 2 ? kx : N0 - kx;
      const bool keep = colv && ky <= a.kmax1 && ks <= a.kmax0;
      float2 val = make_float2(0.f, 0.f);
      if (keep) {
        const float ax = s_ax[kx];
        if constexpr (TERM == TERM_VC) {
          const float2 w = s[0][m];
          if (f == 2) val = make_float2(mulr(ax, w.x), mulr(ax, w.y));
          else if (f == 3) val = w;
          else {
            // psi = -w * invlap,  lap = -(ax^2) + -(ay^2),  invlap(0) = 1   (mesh.py:406-419)
            const float lap = addr(-mulr(ax, ax), -mulr(ay, ay));
            const float il = (lap == 0.f) ? 1.0f : __fdiv_rn(1.0f, lap);
            float2 psi = make_float2(mulr(-w.x, il), mulr(-w.y, il));
            val = (f == 0) ? psi : make_float2(mulr(ax, psi.x), mulr(ax, psi.y));
          }
        } else if constexpr (TERM == TERM_CONS) {
          val = s[f][m];
        } else {
          const float2 u = s[f >> 1][m];
          val = (f & 1) ? make_float2(mulr(ax, u.x), mulr(ax, u.y)) : u;
        }
        val = make_float2(val.x * a.scale, val.y * a.scale);  // power-of-two scale: exact
      }
      v[m] = val;
    }
    fft_wg<N0, +1, V, VAR>(v, s_buf, s_tw, j, vi);
    if (ky < a.PF) {
      float2* g = a.G + (((size_t)b * NF + f) * N0) * a.PF + ky;
#pragma unroll
      for (int m = 0; m < RMAX; ++m) g[(size_t)(j + m * NT) * a.PF] = v[m];
    }
  }
}

template <int N1, int ROWS, int TERM, int VAR>
__global__ void __launch_bounds__((Fft<N1, VAR>::NT * ROWS)) k2_row(Row2Args a) {
  using F = Fft<N1, VAR>;
  constexpr int RMAX = F::RMAX, NT = F::NT;
  constexpr int SL = LdsSize<N1, 1>::slots;
  constexpr int NF = Term2<TERM>::NF;
  __shared__ float2 s_tw[Fft<N1, VAR>::tw_size()];
  __shared__ float2 s_buf[ROWS * SL];
  load_twiddles<N1, VAR>(s_tw, a.tw1);
  __syncthreads();
  const int g = threadIdx.x / NT, j = threadIdx.x % NT;
  int row = blockIdx.x * ROWS + g;
  // with NT a multiple of 64 a wavefront holds one row: the row (and every row pointer) is uniform,
  // so the loads and stores take a scalar base and a 32-bit lane offset
  if constexpr (NT % 64 == 0) row = __builtin_amdgcn_readfirstlane(row);
  const bool in_range = row < a.B * a.N0;
  const int b = a.b0 + (in_range ? row / a.N0 : 0);
  const bool rv = in_range && !a.mask.skip(b);  // rows of samples whose steps are done stay idle
  const int x = rv ? row % a.N0 : 0;
  float2* lds = s_buf + g * SL;
  const size_t fs = (size_t)a.N0 * a.PF;
  const float2* gb = a.G + ((size_t)b * NF) * fs + (size_t)x * a.PF;

  if constexpr (TERM == TERM_VC) {
    float2 z1[RMAX], z2[RMAX];
#pragma unroll
    for (int m = 0; m < RMAX; ++m) {
      const int pos = j + m * NT;
      const bool cj = pos > N1 / 2;
      const int k = cj ? N1 - pos : pos;
      const bool edge = (k == 0) || (k == N1 / 2);
      float2 f0 = make_float2(0.f, 0.f), f1 = f0, f2 = f0, f3 = f0;
      float ay = 0.f;
      if (rv && k <= a.kmax1) {
        f0 = gb[k];
        f1 = gb[fs + k];
        f2 = gb[2 * fs + k];
        f3 = gb[3 * fs + k];
        ay = a.ay[k];
      }
      // ux = C2R(i ay psi), uy = C2R(-i kx psi), wx = C2R(i kx w), wy = C2R(i ay w)
      float2 A = make_float2(-ay * f0.y, ay * f0.x);
      float2 Bv = make_float2(f1.y, -f1.x);
      z1[m] = pack_pair(A, Bv, cj, edge);
      A = make_float2(-f2.y, f2.x);
      Bv = make_float2(-ay * f3.y, ay * f3.x);
      z2[m] = pack_pair(A, Bv, cj, edge);
    }
    fft_wg<N1, +1, 1, VAR, RowSync<N1, VAR>::WS>(z1, lds, s_tw, j, 0);
    fft_wg<N1, +1, 1, VAR, RowSync<N1, VAR>::WS>(z2, lds, s_tw, j, 0);
    float2 v[RMAX];
#pragma unroll
    for (int m = 0; m < RMAX; ++m) v[m] = make_float2(z1[m].x * z2[m].x + z1[m].y * z2[m].y, 0.f);
    fft_wg<N1, -1, 1, VAR, RowSync<N1, VAR>::WS>(v, lds, s_tw, j, 0);
    if (rv) {
      float2* p = a.P + ((size_t)b * a.N0 + x) * a.P1;
#pragma unroll
      for (int m = 0; m < RMAX; ++m) {
        const int pos = j + m * NT;
        if (pos <= N1 / 2) p[pos] = v[m];
      }
    }
  } else if constexpr (TERM == TERM_CONS) {
    // ConservativeConvection, C = D = 2: fields F0 = u0, F1 = u1 (x-inverse done); products
    // u0 u0, u0 u1, u1 u1; planes [F(u0 u0), i a_y F(u1 u0), F(u0 u1), i a_y F(u1 u1)]
    float2 z1[RMAX];
#pragma unroll
    for (int m = 0; m < RMAX; ++m) {
      const int pos = j + m * NT;
      const bool cj = pos > N1 / 2;
      const int k = cj ? N1 - pos : pos;
      const bool edge = (k == 0) || (k == N1 / 2);
      float2 f0 = make_float2(0.f, 0.f), f1 = f0;
      if (rv && k <= a.kmax1) {
        f0 = gb[k];
        f1 = gb[fs + k];
      }
      z1[m] = pack_pair(f0, f1, cj, edge);  // (u0, u1)
    }
    fft_wg<N1, +1, 1, VAR, RowSync<N1, VAR>::WS>(z1, lds, s_tw, j, 0);
    float2 v[RMAX], w[RMAX];
#pragma unroll
    for (int m = 0; m < RMAX; ++m) {
      v[m] = make_float2(z1[m].x * z1[m].x, z1[m].x * z1[m].y);  // u0 u0 + i u0 u1
      w[m] = make_float2(z1[m].y * z1[m].y, 0.f);                // u1 u1
    }
    fft_wg<N1, -1, 1, VAR, RowSync<N1, VAR>::WS>(v, lds, s_tw, j, 0);
    // unpack two real spectra: F00 = (Z[k] + conj Z[N-k]) / 2, F01 = (Z[k] - conj Z[N-k]) / (2i)
#pragma unroll
    for (int m = 0; m < RMAX; ++m) lds[lds_slot<1>(j + m * NT, 0)] = v[m];
    lds_sync<RowSync<N1, VAR>::WS>();
    float2 f00[RMAX], f01[RMAX];
#pragma unroll
    for (int m = 0; m < RMAX; ++m) {
      const int pos = j + m * NT;
      const float2 zk = v[m];
      const float2 zm = lds[lds_slot<1>(pos == 0 ? 0 : N1 - pos, 0)];
      f00[m] = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
      const float2 d = make_float2(zk.x - zm.x, zk.y + zm.y);
      f01[m] = make_float2(0.5f * d.y, -0.5f * d.x);
    }
    lds_sync<RowSync<N1, VAR>::WS>();
    fft_wg<N1, -1, 1, VAR, RowSync<N1, VAR>::WS>(w, lds, s_tw, j, 0);
    if (rv) {
      const size_t pl = (size_t)a.N0 * a.P1;  // one product plane of one sample
      float2* p = a.P + (size_t)b * 4 * pl + (size_t)x * a.P1;
#pragma unroll
      for (int m = 0; m < RMAX; ++m) {
        const int pos = j + m * NT;
        if (pos <= N1 / 2) {
          const float ay = a.ay[pos];
          p[pos] = f00[m];
          p[pl + pos] = make_float2(-mulr(ay, f01[m].y), mulr(ay, f01[m].x));
          p[2 * pl + pos] = f01[m];
          p[3 * pl + pos] = make_float2(-mulr(ay, w[m].y), mulr(ay, w[m].x));
        }
      }
    }
  } else if constexpr (TERM == TERM_KS) {
    // KSConvection, C = 1: fields F0 = phi, F1 = kx phi (x-inverse done);
    // 1/2 (d_x phi^2 + d_y phi^2), d_x phi = C2R(i F1), d_y phi = C2R(i a_y F0)  (_ks_convection.py:27-38)
    float2 z1[RMAX];
#pragma unroll
    for (int m = 0; m < RMAX; ++m) {
      const int pos = j + m * NT;
      const bool cj = pos > N1 / 2;
      const int k = cj ? N1 - pos : pos;
      const bool edge = (k == 0) || (k == N1 / 2);
      float2 f0 = make_float2(0.f, 0.f), f1 = f0;
      float ay = 0.f;
      if (rv && k <= a.kmax1) {
        f0 = gb[k];
        f1 = gb[fs + k];
        ay = a.ay[k];
      }
      z1[m] = pack_pair(make_float2(-f1.y, f1.x), make_float2(-ay * f0.y, ay * f0.x), cj, edge);
    }
    fft_wg<N1, +1, 1, VAR, RowSync<N1, VAR>::WS>(z1, lds, s_tw, j, 0);
    float2 v[RMAX];
#pragma unroll
    for (int m = 0; m < RMAX; ++m)
      v[m] = make_float2(mulr(0.5f, addr(mulr(z1[m].x, z1[m].x), mulr(z1[m].y, z1[m].y))), 0.f);
    fft_wg<N1, -1, 1, VAR, RowSync<N1, VAR>::WS>(v, lds, s_tw, j, 0);
    if (rv) {
      float2* p = a.P + ((size_t)b * a.N0 + x) * a.P1;
#pragma unroll
      for (int m = 0; m < RMAX; ++m) {
        const int pos = j + m * NT;
        if (pos <= N1 / 2) p[pos] = v[m];
      }
    }
  } else {
    // Convection, C = D = 2: fields F0 = u0, F1 = kx u0, F2 = u1, F3 = kx u1 (x-inverse done)
    float2 z1[RMAX], z2[RMAX], z3[RMAX];
#pragma unroll
    for (int m = 0; m < RMAX; ++m) {
      const int pos = j + m * NT;
      const bool cj = pos > N1 / 2;
      const int k = cj ? N1 - pos : pos;
      const bool edge = (k == 0) || (k == N1 / 2);
      float2 f0 = make_float2(0.f, 0.f), f1 = f0, f2 = f0, f3 = f0;
      float ay = 0.f;
      if (rv && k <= a.kmax1) {
        f0 = gb[k];
        f1 = gb[fs + k];
        f2 = gb[2 * fs + k];
        f3 = gb[3 * fs + k];
        ay = a.ay[k];
      }
      z1[m] = pack_pair(f0, f2, cj, edge);                                                   // u0, u1
      z2[m] = pack_pair(make_float2(-f1.y, f1.x), make_float2(-ay * f0.y, ay * f0.x), cj, edge);  // d0u0, d1u0
      z3[m] = pack_pair(make_float2(-f3.y, f3.x), make_float2(-ay * f2.y, ay * f2.x), cj, edge);  // d0u1, d1u1
    }
    fft_wg<N1, +1, 1, VAR, RowSync<N1, VAR>::WS>(z1, lds, s_tw, j, 0);
    fft_wg<N1, +1, 1, VAR, RowSync<N1, VAR>::WS>(z2, lds, s_tw, j, 0);
    fft_wg<N1, +1, 1, VAR, RowSync<N1, VAR>::WS>(z3, lds, s_tw, j, 0);
    float2 v[RMAX];
#pragma unroll
    for (int m = 0; m < RMAX; ++m) {
      const float p0 = z1[m].x * z2[m].x + z1[m].y * z2[m].y;
      const float p1 = z1[m].x * z3[m].x + z1[m].y * z3[m].y;
      v[m] = make_float2(p0, p1);
    }
    fft_wg<N1, -1, 1, VAR, RowSync<N1, VAR>::WS>(v, lds, s_tw, j, 0);
    // unpack two real spectra: P0 = (Z[k] + conj Z[N-k]) / 2, P1 = (Z[k] - conj Z[N-k]) / (2i)
#pragma unroll
    for (int m = 0; m < RMAX; ++m) lds[lds_slot<1>(j + m * NT, 0)] = v[m];
    lds_sync<RowSync<N1, VAR>::WS>();
    if (rv) {
      float2* p0 = a.P + (((size_t)b * 2 + 0) * a.N0 + x) * a.P1;
      float2* p1 = a.P + (((size_t)b * 2 + 1) * a.N0 + x) * a.P1;
#pragma unroll
      for (int m = 0; m < RMAX; ++m) {
        const int pos = j + m * NT;
        if (pos <= N1 / 2) {
          const float2 zk = v[m];
          const float2 zm = lds[lds_slot<1>(pos == 0 ? 0 : N1 - pos, 0)];
          p0[pos] = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
          const float2 d = make_float2(zk.x - zm.x, zk.y + zm.y);
          p1[pos] = make_float2(0.5f * d.y, -0.5f * d.x);
        }
      }
    }
  }
}

// ---------------------------------------------------------------- host launchers
// Tuning knobs (chosen per shape class by the autotuner): column tile V in {8, 16}, radix-plan
// variant of the column and row transforms (0: N/RMAX threads per transform; 1: half as many),
// and the row-kernel workgroup size (64 or 256 threads).
template <int N0, int V, int TERM, int VAR>
inline hipError_t launch_col_t(const Col2Args& a, int nb, int stage, hipStream_t st) {
  const int tiles = (a.H1 + V - 1) / V;
  Col2Args b = a;
  b.nb = nb;
  hipLaunchKernelGGL((k2_col<N0, V, TERM, VAR>), dim3(tiles * nb), dim3(Fft<N0, VAR>::NT * V), 0, st, b, stage);
  return hipGetLastError();
}

template <int N1, int TERM, int VAR, int WGT>
inline hipError_t launch_row_t(const Row2Args& a, hipStream_t st) {
  constexpr int NT = Fft<N1, VAR>::NT;
  constexpr int ROWS = (NT >= WGT) ? 1 : WGT / NT;
  const int rows = a.B * a.N0;
  hipLaunchKernelGGL((k2_row<N1, ROWS, TERM, VAR>), dim3((rows + ROWS - 1) / ROWS), dim3(NT * ROWS), 0, st, a);
  return hipGetLastError();
}


// size lists of the column (N0, tile V, radix variant) and row (N1, variant, workgroup size)
// instantiations, split over the k_step2d_*.hip units to keep each compile short
#define TFSM_COL_P2(X)                                                                                        \
  X(16, 8, 0) X(16, 16, 0) X(32, 8, 0) X(32, 16, 0) X(64, 8, 0) X(64, 16, 0) X(64, 8, 1) X(64, 16, 1)          \
  X(128, 8, 0) X(128, 16, 0) X(128, 8, 1) X(128, 16, 1) X(256, 8, 0) X(256, 16, 0) X(256, 8, 1) X(256, 16, 1) \
  X(512, 8, 0) X(512, 16, 0) X(512, 8, 1) X(512, 16, 1)
#define TFSM_COL_MX(X)                                                                                       \
  X(48, 8, 0) X(48, 16, 0) X(96, 8, 0) X(96, 16, 0) X(100, 8, 0) X(100, 16, 0) X(112, 8, 0) X(112, 16, 0)    \
  X(192, 8, 0) X(192, 16, 0) X(200, 8, 0) X(200, 16, 0)
#define TFSM_ROW_P2(X)                                                                                        \
  X(16, 0, 256) X(32, 0, 256) X(64, 0, 256) X(64, 1, 256) X(64, 0, 64) X(64, 1, 64) X(128, 0, 256)            \
  X(128, 1, 256) X(128, 0, 64) X(128, 1, 64) X(256, 0, 256) X(256, 1, 256) X(256, 0, 64) X(256, 1, 64)        \
  X(512, 0, 256) X(512, 1, 256) X(512, 0, 64) X(512, 1, 64)
#define TFSM_ROW_MX(X) X(48, 0, 256) X(96, 0, 256) X(100, 0, 256) X(112, 0, 256) X(192, 0, 256) X(200, 0, 256)

#define TFSM_COL_CASE(N, VV, VR) \
  if (a.N0 == N && v == VV && var == VR) return launch_col_t<N, VV, TERM, VR>(a, nb, stage, st);
#define TFSM_ROW_CASE(N, VR, W) \
  if (a.N1 == N && var == VR && wgt == W) return launch_row_t<N, TERM, VR, W>(a, st);

// per-unit entry points (k_step2d_*.hip)
hipError_t launch_col2d_vc(const Col2Args& a, int nb, int stage, int v, int var, hipStream_t st);
hipError_t launch_col2d_conv(const Col2Args& a, int nb, int stage, int v, int var, hipStream_t st);
hipError_t launch_col2d_nsp(const Col2Args& a, int nb, int stage, int v, int var, hipStream_t st);
hipError_t launch_col2d_mixed(int term, const Col2Args& a, int nb, int stage, int v, int var, hipStream_t st);
hipError_t launch_row2d_vc(const Row2Args& a, int var, int wgt, hipStream_t st);
hipError_t launch_row2d_conv(const Row2Args& a, int var, int wgt, hipStream_t st);
hipError_t launch_row2d_mixed(int term, const Row2Args& a, int var, int wgt, hipStream_t st);
hipError_t launch_col2d_ks(const Col2Args& a, int nb, int stage, int v, int var, hipStream_t st);
hipError_t launch_col2d_cons(const Col2Args& a, int nb, int stage, int v, int var, hipStream_t st);
hipError_t launch_row2d_cons(const Row2Args& a, int var, int wgt, hipStream_t st);
hipError_t launch_row2d_ks(const Row2Args& a, int var, int wgt, hipStream_t st);

}  // namespace tfsm
