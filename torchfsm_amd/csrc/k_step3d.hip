// 3D fused ETD evaluation kernels for Convection with C = D = 3 (3D vector Burgers), for the
// velocity-form Navier-Stokes term (NSPressureConvection: the same fields and products, the
// projection applied to the convection spectrum in the x-column kernel), and for KSConvection
// (C = 1: fields phi^, a_x phi^ -> phi, a_y phi, a_x phi; one product 1/2 |grad phi|^2).
//
// One nonlinear evaluation = four passes over the half spectrum [B][3][N0][N1][P2]:
//   k3_row   z rows: 12 C2R (packed in pairs) of the y-inverse fields -> the three products
//            p_c = u_0 d_x u_c + u_1 d_y u_c + u_2 d_z u_c -> R2C (p0 + i p1 packed, p2 alone)
//   cols_c2c y columns, forward, in place on the products (generic strided FFT pass)
//   k3_col   x columns: forward FFT of the products -> N^ -> stage combine -> the dealiased fields
//            u^_c and a_x u^_c of the next stage -> inverse x FFT (kept (ky, kz) only) -> G1
//   k3_yinv  y columns: from G1 the nine fields u^_c, a_y u^_c, a_x u^_c -> inverse y FFT -> G2
// Reference path replaced: _ConvectionCore (_convection.py:18-46) inside the closure
// (_base.py:375-403) and the ETDRK2 / SETDRK2 combines (_etdrk.py:72-82, _setdrk_step.py:14-25).
//
// Field layouts (kz <= kmax2 only, pitch PF2; ky compressed to the K1 = 2 kmax1 + 1 kept values):
//   G1 [B][6][N0][K1][PF2]   f = 2c: u^_c,  f = 2c + 1: a_x u^_c          (x inverse done)
//   G2 [B][9][N0][N1][PF2]   f = 3c: u_c,   3c + 1: a_y u_c,  3c + 2: a_x u_c  (x, y inverse done)
#include "engine.hpp"

namespace tfsm {

__device__ __forceinline__ float2 pack_pair3(float2 A, float2 Bv, bool cj, bool edge) {
  if (edge) {
    A.y = 0.f;
    Bv.y = 0.f;
  }
  return cj ? make_float2(A.x + Bv.y, Bv.x - A.y) : make_float2(A.x - Bv.y, A.y + Bv.x);
}

// kept ky -> compressed index, or -1
__device__ __forceinline__ int ky_compress(int ky, int N1, int kmax1) {
  if (ky <= kmax1) return ky;
  if (ky >= N1 - kmax1) return ky - (N1 - 2 * kmax1 - 1);
  return -1;
}

// ---------------------------------------------------------------- x columns (fused combine)
// Fields f0 = u^_c and f0 + 1 = a_x u^_c of one channel's combined spectrum s (kept modes, scaled),
// inverse x FFT, stored to G1 at the compressed ky index kyc.
template <int N0, int V, int VAR>
__device__ __forceinline__ void col3_fields(const Col3Args& a, const float2 (&s)[Fft<N0, VAR>::RMAX], int f0, int b,
                                            int nc, int kyc, int kz, bool colv, float2* s_buf, const float2* s_tw,
                                            const float* s_ax, int j, int vi) {
  using F = Fft<N0, VAR>;
  constexpr int RMAX = F::RMAX, NT = F::NT;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    float2 v[RMAX];
#pragma unroll
    for (int m = 0; m < RMAX; ++m) {
      const int kx = j + m * NT;
      const int ks = kx <= N0 / 2 ? kx : N0 - kx;
      const bool keep = colv && kz <= a.kmax2 && ks <= a.kmax0;
      float2 val = make_float2(0.f, 0.f);
      if (keep) {
        const float2 u = s[m];
        if (h) {
          const float ax = s_ax[kx];
          val = make_float2(mulr(ax, u.x), mulr(ax, u.y));
        } else {
          val = u;
        }
        val = make_float2(val.x * a.scale, val.y * a.scale);
      }
      v[m] = val;
    }
    fft_wg<N0, +1, V, VAR>(v, s_buf, s_tw, j, vi);
    if (kz < a.PF2) {
      float2* g = a.G1 + ((((size_t)b * 2 * nc + f0 + h) * N0) * a.K1 + kyc) * a.PF2 + kz;
#pragma unroll
      for (int m = 0; m < RMAX; ++m) g[(size_t)(j + m * NT) * a.K1 * a.PF2] = v[m];
    }
  }
}

// ST >= 0: the stage is a compile-time constant (only that stage's combine is compiled, which
// lowers the register count: 128-point Convection columns 193 -> 133 / 148 VGPRs for ETDRK2's
// evaluations, three waves per SIMD instead of two); ST = -1: the stage comes at run time.
// (Holding the stage kernels to 128 VGPRs, four waves, spills a few values and measured slower.)
template <int N0, int V, int VAR, int TERM, int ST = -1>
__global__ void __launch_bounds__((Fft<N0, VAR>::NT * V)) k3_col(Col3Args a, int stage_rt) {
  const int stage = ST >= 0 ? ST : stage_rt;
  using F = Fft<N0, VAR>;
  constexpr int RMAX = F::RMAX, NT = F::NT;
  __shared__ float2 s_tw[Fft<N0, VAR>::tw_size()];
  __shared__ float2 s_buf[LdsSize<N0, V>::slots];
  __shared__ float s_ax[N0];  // x wavenumbers: read per mode by the field pass without a memory wait
  load_twiddles<N0, VAR>(s_tw, a.tw0);
  for (int i = threadIdx.x; i < N0; i += NT * V) s_ax[i] = a.ax[i];
  __syncthreads();
  const int vi = threadIdx.x % V, j = threadIdx.x / V;
  // tiles (ky, kz-tile) go round-robin to the XCDs (block v on XCD v % 8); the nb samples of one
  // tile run back to back on the same XCD, so each coefficient tile is fetched once per launch
  const int nzt = a.P2 / V;
  const int ntiles = a.N1 * nzt;
  const int x8 = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int tile = x8 + 8 * (slot / a.nb);
  const int bl = slot % a.nb;
  if (tile >= ntiles) return;
  const int b = a.b0 + bl;
  if (a.mask.skip(b)) return;
  const int ky = tile / nzt;
  const int kz = (tile % nzt) * V + vi;
  const bool colv = kz < a.H2;
  const size_t inner = (size_t)a.N1 * a.P2;  // x stride
  const size_t plane = (size_t)N0 * inner;   // modes
  const size_t coff = (size_t)ky * a.P2 + kz;
  const float4* cpk = a.cpk + (size_t)b * a.cpk_bstride;
  const size_t tb = (size_t)b * (a.cpk_bstride / 2);  // per-mode tables (c5t, ck01, c6t) of this sample
  // cpk_index(coff + xo, half) = 2 (coff + xo) + cko: (coff + xo) & 15 == kz & 15
  const int cko = (stage_coef_half(stage) << 4) - (kz & 15);
  constexpr int NC3 = TERM == TERM_KS ? 1 : 3;  // channels
  // combine batch (modes whose operands are loaded together): fewer than in 2D, the NSP kernel keeps
  // all three channels' values live
  constexpr int CBM = RMAX >= 16 ? 1 : (RMAX >= 8 ? 2 : 4);
  const int kyc = ky_compress(ky, a.N1, a.kmax1);
  const bool kept = kyc >= 0 && (tile % nzt) * V <= a.kmax2;  // uniform: the tile has kept columns
  if constexpr (TERM == TERM_NSP) {
    float2 s[3][RMAX];
    if (stage == ST_PRO) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const size_t base = ((size_t)b * 3 + c) * plane + coff;
#pragma unroll
        for (int m = 0; m < RMAX; ++m) {
          const float2 t = a.U[base + (size_t)(j + m * NT) * inner];
          s[c][m] = colv ? t : make_float2(0.f, 0.f);
        }
      }
    } else {
      // velocity-form Navier-Stokes: the three convection spectra first, projected mode by mode
      float2 v3[3][RMAX];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const size_t base = ((size_t)b * 3 + c) * plane + coff;
#pragma unroll
        for (int m = 0; m < RMAX; ++m) {
          const float2 t = a.P[base + (size_t)(j + m * NT) * inner];
          v3[c][m] = colv ? t : make_float2(0.f, 0.f);
        }
        fft_wg<N0, -1, V, VAR>(v3[c], s_buf, s_tw, j, vi);
      }
      const float ayv = a.ay[ky], azv = colv ? a.az[kz] : 0.f;
      // Nyquist-plane parity of nsp_project (every fused 3D size is even)
      const bool nyy = 2 * ky == a.N1, nyz = 2 * kz == a.N2, oyz = (ky == 0 || nyy) && (kz == 0 || nyz);
#pragma unroll
      for (int m = 0; m < RMAX; ++m) {
        float2 cv[3] = {v3[0][m], v3[1][m], v3[2][m]};
        const int x = j + m * NT;
        const float av[3] = {a.ax[x], ayv, azv};
        const bool nyx = 2 * x == N0, own = (x == 0 || nyx) && oyz;
        const bool ny[3] = {nyx && !own, nyy && !own, nyz && !own};
        nsp_project<3>(cv, av, ny);
#pragma unroll
        for (int c = 0; c < 3; ++c) v3[c][m] = cv[c];
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const size_t base = ((size_t)b * 3 + c) * plane + coff;
        const float2* src = a.src ? a.src + (size_t)b * a.src_bstride + (size_t)c * a.src_cstride + coff : nullptr;
        combine_column<RMAX, NT, CBM, (ST >= 0)>(stage, a.U + base, a.A1 + base, a.A2 + base, src, cpk + (2 * coff + cko),
                                      a.c5t + tb, (unsigned)coff, a.coef_t, (unsigned)inner, j, colv, v3[c], s[c],
                                      a.ck01 ? a.ck01 + tb : nullptr, a.c6t ? a.c6t + tb : nullptr);
      }
    }
    if (!kept) return;
#pragma unroll
    for (int c = 0; c < 3; ++c)
      col3_fields<N0, V, VAR>(a, s[c], 2 * c, b, 3, kyc, kz, colv, s_buf, s_tw, s_ax, j, vi);
  } else {
    // one channel at a time: its spectrum, combine and two fields (u^_c, a_x u^_c), so that only
    // that channel's values are live (the fields of channel c need nothing of the others)
#pragma unroll
    for (int c = 0; c < NC3; ++c) {
      const size_t base = ((size_t)b * NC3 + c) * plane + coff;
      float2 s[RMAX];
      if (stage == ST_PRO) {
#pragma unroll
        for (int m = 0; m < RMAX; ++m) {
          const float2 t = a.U[base + (size_t)(j + m * NT) * inner];
          s[m] = colv ? t : make_float2(0.f, 0.f);
        }
      } else {
        float2 v[RMAX];
#pragma unroll
        for (int m = 0; m < RMAX; ++m) {
          const float2 t = a.P[base + (size_t)(j + m * NT) * inner];
          v[m] = colv ? t : make_float2(0.f, 0.f);
        }
        fft_wg<N0, -1, V, VAR>(v, s_buf, s_tw, j, vi);
        if constexpr (TERM == TERM_KS) {
          // remove_mean: fft(p - mean(p)) differs from fft(p) only at k = 0, by the batch mean of the DC
          if (a.dc && ky == 0 && kz == 0 && j == 0) v[0] = rsub(v[0], a.dc[0]);
        }
        const float2* src = a.src ? a.src + (size_t)b * a.src_bstride + (size_t)c * a.src_cstride + coff : nullptr;
        combine_column<RMAX, NT, CBM, (ST >= 0)>(stage, a.U + base, a.A1 + base, a.A2 + base, src, cpk + (2 * coff + cko),
                                      a.c5t + tb, (unsigned)coff, a.coef_t, (unsigned)inner, j, colv, v, s,
                                      a.ck01 ? a.ck01 + tb : nullptr, a.c6t ? a.c6t + tb : nullptr);
      }
      if (kept) col3_fields<N0, V, VAR>(a, s, 2 * c, b, NC3, kyc, kz, colv, s_buf, s_tw, s_ax, j, vi);
    }
  }
}

// ---------------------------------------------------------------- y columns (inverse, 6 -> 9 fields)
template <int N1, int V, int VAR, int NCH>
__global__ void __launch_bounds__((Fft<N1, VAR>::NT * V)) k3_yinv(Yinv3Args a) {
  using F = Fft<N1, VAR>;
  constexpr int RMAX = F::RMAX, NT = F::NT;
  __shared__ float2 s_tw[Fft<N1, VAR>::tw_size()];
  __shared__ float2 s_buf[LdsSize<N1, V>::slots];
  load_twiddles<N1, VAR>(s_tw, a.tw1);
  __syncthreads();
  const int vi = threadIdx.x % V, j = threadIdx.x / V;
  const int nzt = a.PF2 / V;
  const int kz = (blockIdx.x % nzt) * V + vi;
  const int x = (blockIdx.x / nzt) % a.N0;
  const int b = a.b0 + blockIdx.x / (nzt * a.N0);
  if (a.mask.skip(b)) return;
  const bool kv = kz <= a.kmax2;
#pragma unroll 1
  for (int c = 0; c < NCH; ++c) {
    const float2* g0 = a.G1 + ((((size_t)b * 2 * NCH + 2 * c) * a.N0 + x) * a.K1) * a.PF2 + kz;
    const float2* g1 = a.G1 + ((((size_t)b * 2 * NCH + 2 * c + 1) * a.N0 + x) * a.K1) * a.PF2 + kz;
    float2 f0[RMAX], f1[RMAX];
#pragma unroll
    for (int m = 0; m < RMAX; ++m) {
      const int ky = j + m * NT;
      const int kc = ky_compress(ky, N1, a.kmax1);
      const bool ld = kv && kc >= 0;
      f0[m] = ld ? g0[(size_t)kc * a.PF2] : make_float2(0.f, 0.f);
      f1[m] = ld ? g1[(size_t)kc * a.PF2] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float2 v[RMAX];
#pragma unroll
      for (int m = 0; m < RMAX; ++m) {
        if (k == 0) v[m] = f0[m];
        else if (k == 2) v[m] = f1[m];
        else {
          const float ay = a.ay[j + m * NT];
          v[m] = make_float2(mulr(ay, f0[m].x), mulr(ay, f0[m].y));
        }
      }
      fft_wg<N1, +1, V, VAR>(v, s_buf, s_tw, j, vi);
      if (kz < a.PF2) {
        float2* g = a.G2 + ((((size_t)b * 3 * NCH + 3 * c + k) * a.N0 + x) * N1) * a.PF2 + kz;
#pragma unroll
        for (int m = 0; m < RMAX; ++m) g[(size_t)(j + m * NT) * a.PF2] = v[m];
      }
    }
  }
}

// ---------------------------------------------------------------- z rows (products)
// ROWS consecutive rows of one sample per workgroup (ROWS divides N0 * N1).  The kept kz of the
// nine G2 fields of those rows are contiguous per field: they are staged into LDS (dynamic, 9 x
// ROWS x PF2 complex) with coalesced 16-byte loads first, so all loads are in flight together,
// and the twelve inverse transforms then read LDS only.
template <int N2, int ROWS, int VAR, int TERM>
__global__ void __launch_bounds__((Fft<N2, VAR>::NT * ROWS)) k3_row(Row3Args a) {
  using F = Fft<N2, VAR>;
  constexpr int RMAX = F::RMAX, NT = F::NT;
  constexpr int NFS = TERM == TERM_KS ? 3 : 9;  // staged G2 fields
  constexpr int SL = LdsSize<N2, 1>::slots;
  __shared__ float2 s_tw[Fft<N2, VAR>::tw_size()];
  __shared__ float2 s_buf[ROWS * SL];
  extern __shared__ float4 s_in4[];
  float2* s_in = reinterpret_cast<float2*>(s_in4);
  load_twiddles<N2, VAR>(s_tw, a.tw2);
  const int g = threadIdx.x / NT, j = threadIdx.x % NT;
  const size_t row0 = (size_t)blockIdx.x * ROWS;
  const int nxy = a.N0 * a.N1;
  const int bl = (int)(row0 / nxy);
  const int xy0 = (int)(row0 % nxy);
  const int b = a.b0 + bl;
  if (a.mask.skip(b)) return;  // whole workgroup (its rows belong to one sample)
  const int PF2 = a.PF2;
  const size_t fs = (size_t)nxy * PF2;  // one G2 field of one sample
  {
    const float4* src = reinterpret_cast<const float4*>(a.G2 + (size_t)b * NFS * fs + (size_t)xy0 * PF2);
    const int per = ROWS * PF2 / 2;      // float4 per field
    const size_t fs4 = fs / 2;
    for (int i = threadIdx.x; i < NFS * per; i += NT * ROWS) {
      const int f = i / per, r = i - f * per;
      s_in4[i] = src[f * fs4 + r];
    }
  }
  __syncthreads();
  float2* lds = s_buf + g * SL;
  const float2* in = s_in + (size_t)g * PF2;  // field f of this row at in[f * ROWS * PF2 + k]
  // z wavenumbers of this lane's modes, loaded once (k <= N2/2 is in the table)
  float azr[RMAX];
#pragma unroll
  for (int m = 0; m < RMAX; ++m) {
    const int pos = j + m * NT;
    azr[m] = a.az[pos > N2 / 2 ? N2 - pos : pos];
  }
  // packed inverse transform of two fields: A = ma(fa), B = mb(fb) where the multiplier is
  // 0: 1, 1: i, 2: i a_z  (u_c = C2R(G[3c]), d_y u_c = C2R(i G[3c+1]), d_x u_c = C2R(i G[3c+2]),
  // d_z u_c = C2R(i a_z G[3c]))
  auto inv_pair = [&](float2 (&z)[RMAX], int fa, int ma, int fb, int mb) {
#pragma unroll
    for (int m = 0; m < RMAX; ++m) {
      const int pos = j + m * NT;
      const bool cj = pos > N2 / 2;
      const int k = cj ? N2 - pos : pos;
      const bool edge = (k == 0) || (k == N2 / 2);
      float2 A = make_float2(0.f, 0.f), Bv = A;
      if (k <= a.kmax2) {
        const float az = azr[m];
        const float2 x1 = in[fa * ROWS * PF2 + k], x2 = in[fb * ROWS * PF2 + k];
        A = ma == 0 ? x1 : (ma == 1 ? make_float2(-x1.y, x1.x) : make_float2(-az * x1.y, az * x1.x));
        Bv = mb == 0 ? x2 : (mb == 1 ? make_float2(-x2.y, x2.x) : make_float2(-az * x2.y, az * x2.x));
      }
      z[m] = pack_pair3(A, Bv, cj, edge);
    }
    fft_wg<N2, +1, 1, VAR, RowSync<N2, VAR>::WS>(z, lds, s_tw, j, 0);
  };
  if constexpr (TERM == TERM_KS) {
    // KSConvection, C = 1: G2 fields 0: phi, 1: a_y phi, 2: a_x phi;
    // p = 1/2 ((d_x phi^2 + d_y phi^2) + d_z phi^2)   (_ks_convection.py:27-38)
    float2 z[RMAX], w[RMAX];
    inv_pair(z, 2, 1, 1, 1);  // (d_x phi, d_y phi)
    inv_pair(w, 0, 2, 0, 0);  // (d_z phi, phi)
#pragma unroll
    for (int m = 0; m < RMAX; ++m)
      z[m] = make_float2(
          mulr(0.5f, addr(addr(mulr(z[m].x, z[m].x), mulr(z[m].y, z[m].y)), mulr(w[m].x, w[m].x))), 0.f);
    fft_wg<N2, -1, 1, VAR, RowSync<N2, VAR>::WS>(z, lds, s_tw, j, 0);
    float2* pb = a.P + (size_t)b * nxy * a.P2 + (size_t)(xy0 + g) * a.P2;
#pragma unroll
    for (int m = 0; m < RMAX; ++m) {
      const int pos = j + m * NT;
      if (pos <= N2 / 2) pb[pos] = z[m];
    }
    return;
  }
  float u0[RMAX], u1[RMAX], u2[RMAX], p0[RMAX], p1[RMAX], p2[RMAX];
  float2 z[RMAX];
  inv_pair(z, 0, 0, 3, 0);  // (u0, u1)
#pragma unroll
  for (int m = 0; m < RMAX; ++m) {
    u0[m] = z[m].x;
    u1[m] = z[m].y;
  }
  inv_pair(z, 6, 0, 2, 1);  // (u2, d_x u0)
#pragma unroll
  for (int m = 0; m < RMAX; ++m) {
    u2[m] = z[m].x;
    p0[m] = u0[m] * z[m].y;
  }
  inv_pair(z, 1, 1, 0, 2);  // (d_y u0, d_z u0)
#pragma unroll
  for (int m = 0; m < RMAX; ++m) p0[m] = (p0[m] + u1[m] * z[m].x) + u2[m] * z[m].y;
  inv_pair(z, 5, 1, 4, 1);  // (d_x u1, d_y u1)
#pragma unroll
  for (int m = 0; m < RMAX; ++m) p1[m] = u0[m] * z[m].x + u1[m] * z[m].y;
  inv_pair(z, 3, 2, 8, 1);  // (d_z u1, d_x u2)
#pragma unroll
  for (int m = 0; m < RMAX; ++m) {
    p1[m] = p1[m] + u2[m] * z[m].x;
    p2[m] = u0[m] * z[m].y;
  }
  inv_pair(z, 7, 1, 6, 2);  // (d_y u2, d_z u2)
#pragma unroll
  for (int m = 0; m < RMAX; ++m) p2[m] = (p2[m] + u1[m] * z[m].x) + u2[m] * z[m].y;
  // forward: (p0 + i p1) packed, then p2
#pragma unroll
  for (int m = 0; m < RMAX; ++m) z[m] = make_float2(p0[m], p1[m]);
  fft_wg<N2, -1, 1, VAR, RowSync<N2, VAR>::WS>(z, lds, s_tw, j, 0);
#pragma unroll
  for (int m = 0; m < RMAX; ++m) lds[lds_slot<1>(j + m * NT, 0)] = z[m];
  lds_sync<RowSync<N2, VAR>::WS>();
  const size_t cs = (size_t)nxy * a.P2;  // one channel of the products
  float2* pb = a.P + (size_t)b * 3 * cs + (size_t)(xy0 + g) * a.P2;
#pragma unroll
  for (int m = 0; m < RMAX; ++m) {
    const int pos = j + m * NT;
    if (pos <= N2 / 2) {
      const float2 zk = z[m];
      const float2 zm = lds[lds_slot<1>(pos == 0 ? 0 : N2 - pos, 0)];
      pb[pos] = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
      const float2 d = make_float2(zk.x - zm.x, zk.y + zm.y);
      pb[cs + pos] = make_float2(0.5f * d.y, -0.5f * d.x);
    }
  }
  lds_sync<RowSync<N2, VAR>::WS>();
#pragma unroll
  for (int m = 0; m < RMAX; ++m) z[m] = make_float2(p2[m], 0.f);
  fft_wg<N2, -1, 1, VAR, RowSync<N2, VAR>::WS>(z, lds, s_tw, j, 0);
#pragma unroll
  for (int m = 0; m < RMAX; ++m) {
    const int pos = j + m * NT;
    if (pos <= N2 / 2) pb[2 * cs + pos] = z[m];
  }
}

// ---------------------------------------------------------------- host launchers
#define TFSM_SWITCH3(n, ...)                           \
  switch (n) {                                         \
    case 16: { constexpr int NN = 16; __VA_ARGS__ }    \
    case 32: { constexpr int NN = 32; __VA_ARGS__ }    \
    case 64: { constexpr int NN = 64; __VA_ARGS__ }    \
    case 128: { constexpr int NN = 128; __VA_ARGS__ }  \
    case 48: { constexpr int NN = 48; __VA_ARGS__ }  \
    case 96: { constexpr int NN = 96; __VA_ARGS__ }  \
    case 100: { constexpr int NN = 100; __VA_ARGS__ }  \
    default: return hipErrorInvalidValue;              \
  }

// Convection column kernels also compiled per ETDRK2 / SETDRK2 and SETDRK4 stage (and for the
// field pass of the initial state): the 64- and 128-point x columns with the default radix plan
template <int NN, int VR>
struct Col3StageKernels {
  static constexpr bool value = (NN == 64 || NN == 128) && VR == (NN == 128 ? 2 : 0);
};

template <int NN, int VR>
static hipError_t launch_col3d_t(const Col3Args& a, int nb, int stage, hipStream_t st) {
  constexpr int V = 16;
  const int ntiles = a.N1 * (a.P2 / V);
  const int slots = (ntiles + 7) / 8;
  Col3Args b = a;
  b.nb = nb;
  if (a.term == TERM_NSP)
    TFSM_LAUNCH((k3_col<NN, V, VR, TERM_NSP>), dim3(slots * 8 * nb), dim3(Fft<NN, VR>::NT * V), 0, st, b, stage);
  else if (a.term == TERM_KS)
    TFSM_LAUNCH((k3_col<NN, V, VR, TERM_KS>), dim3(slots * 8 * nb), dim3(Fft<NN, VR>::NT * V), 0, st, b, stage);
  else if constexpr (Col3StageKernels<NN, VR>::value) {
    const dim3 grid(slots * 8 * nb), block(Fft<NN, VR>::NT * V);
    switch (stage) {
      case ST_PRO: TFSM_LAUNCH((k3_col<NN, V, VR, TERM_CONV, ST_PRO>), grid, block, 0, st, b, stage); break;
      case ST_R2_E0: TFSM_LAUNCH((k3_col<NN, V, VR, TERM_CONV, ST_R2_E0>), grid, block, 0, st, b, stage); break;
      case ST_R2_E1: TFSM_LAUNCH((k3_col<NN, V, VR, TERM_CONV, ST_R2_E1>), grid, block, 0, st, b, stage); break;
      case ST_R4_E0: TFSM_LAUNCH((k3_col<NN, V, VR, TERM_CONV, ST_R4_E0>), grid, block, 0, st, b, stage); break;
      case ST_R4_E1: TFSM_LAUNCH((k3_col<NN, V, VR, TERM_CONV, ST_R4_E1>), grid, block, 0, st, b, stage); break;
      case ST_R4_E2: TFSM_LAUNCH((k3_col<NN, V, VR, TERM_CONV, ST_R4_E2>), grid, block, 0, st, b, stage); break;
      case ST_R4_E3: TFSM_LAUNCH((k3_col<NN, V, VR, TERM_CONV, ST_R4_E3>), grid, block, 0, st, b, stage); break;
      default: TFSM_LAUNCH((k3_col<NN, V, VR, TERM_CONV>), grid, block, 0, st, b, stage);
    }
  } else
    TFSM_LAUNCH((k3_col<NN, V, VR, TERM_CONV>), dim3(slots * 8 * nb), dim3(Fft<NN, VR>::NT * V), 0, st, b,
                       stage);
  return hipGetLastError();
}

template <int NN, int VR>
static hipError_t launch_yinv3d_t(const Yinv3Args& a, hipStream_t st) {
  constexpr int V = 16;
  const int nzt = a.PF2 / V;
  if (a.nc == 1)
    TFSM_LAUNCH((k3_yinv<NN, V, VR, 1>), dim3(nzt * a.N0 * a.nb), dim3(Fft<NN, VR>::NT * V), 0, st, a);
  else
    TFSM_LAUNCH((k3_yinv<NN, V, VR, 3>), dim3(nzt * a.N0 * a.nb), dim3(Fft<NN, VR>::NT * V), 0, st, a);
  return hipGetLastError();
}

// col_var 0: the 3D default radix plan of the axis (var3), 1: radix-4 plan (4 values per thread)
hipError_t launch_col3d(const Col3Args& a, int nb, int stage, int col_var, hipStream_t st) {
  TFSM_SWITCH3(a.N0, {
    if constexpr (!is_mixed_size(NN)) {
      if (col_var == 1) return launch_col3d_t<NN, 3>(a, nb, stage, st);
    }
    return launch_col3d_t<NN, (NN == 128 ? 2 : 0)>(a, nb, stage, st);
  })
}

hipError_t launch_yinv3d(const Yinv3Args& a, int col_var, hipStream_t st) {
  TFSM_SWITCH3(a.N1, {
    if constexpr (!is_mixed_size(NN)) {
      if (col_var == 1) return launch_yinv3d_t<NN, 3>(a, st);
    }
    return launch_yinv3d_t<NN, (NN == 128 ? 2 : 0)>(a, st);
  })
}

// rows per workgroup of the 3D row kernel: about 256 threads, a divisor of NN^2 (the row count of
// one sample of a cubic mesh) whose staged fields fit 96 KiB of LDS at the largest field pitch
// (kmax <= NN/3)
template <int NN, int NT>
constexpr int rows3d() {
  int r = NT >= 256 ? 1 : 256 / NT;
  const int pf = ((NN / 3 + 1) + 15) / 16 * 16;
  while (r > 1 && ((NN * NN) % r != 0 || 9 * r * pf * 8 > 96 * 1024)) --r;
  return r;
}

template <int NN, int VR>
static hipError_t launch_row3d_t(const Row3Args& a, hipStream_t st) {
  constexpr int NT = Fft<NN, VR>::NT;
  constexpr int R = rows3d<NN, NT>();
  const size_t rows = (size_t)a.nb * a.N0 * a.N1;
  // whole workgroups of one sample each; staged fields in dynamic LDS
  if ((a.N0 * a.N1) % R != 0 || (a.PF2 % 2) != 0) return hipErrorInvalidValue;
  const bool ks = a.term == TERM_KS;
  const size_t lds = (size_t)(ks ? 3 : 9) * R * a.PF2 * sizeof(float2);
  if (lds > 96 * 1024) return hipErrorInvalidValue;
  if (ks)
    TFSM_LAUNCH((k3_row<NN, R, VR, TERM_KS>), dim3((unsigned)(rows / R)), dim3(NT * R), lds, st, a);
  else
    TFSM_LAUNCH((k3_row<NN, R, VR, TERM_CONV>), dim3((unsigned)(rows / R)), dim3(NT * R), lds, st, a);
  return hipGetLastError();
}

// row_var 0: the 3D default radix plan of the axis (var3), 1: radix-4 plan (4 values per thread)
hipError_t launch_row3d(const Row3Args& a, int row_var, hipStream_t st) {
  TFSM_SWITCH3(a.N2, {
    if constexpr (!is_mixed_size(NN)) {
      if (row_var == 1) return launch_row3d_t<NN, 3>(a, st);
    }
    return launch_row3d_t<NN, (NN == 128 ? 2 : 0)>(a, st);
  })
}

// powers of two 16..128 per axis, or a cubic mesh of one of the mixed-radix 3D sizes
bool step3d_supported(int n0, int n1, int n2) {
  auto ok = [](int n) { return n == 16 || n == 32 || n == 64 || n == 128; };
  if (ok(n0) && ok(n1) && ok(n2)) return true;
  return n0 == n1 && n1 == n2 && is_mixed_size_3d(n0);
}

}  // namespace tfsm
