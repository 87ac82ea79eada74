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
// rs: read the source spectrum src in this column (src_col); callers look it up before their
// forward transform, so that its load does not stall the combine.
template <int N0, int VAR, int BM, bool FENCE = false, int PRE = 0>
__device__ __forceinline__ void combine_column2(const Col2Args& a, int stage, size_t base, const float2* src,
                                                const float4* cpk, const float* c5b, size_t tb, int cko, int ky,
                                                bool colv, int j,
                                                const float2 (&v)[Fft<N0, VAR>::RMAX],
                                                float2 (&s)[Fft<N0, VAR>::RMAX], bool rs,
                                                const ColPre<Fft<N0, VAR>::RMAX>* pre = nullptr) {
  using F = Fft<N0, VAR>;
  const size_t cb = base + ky;
  combine_column<F::RMAX, F::NT, BM, FENCE, PRE>(stage, a.U + cb, a.A1 + cb, a.A2 + cb, rs ? src + ky : nullptr,
                                     cpk + (2 * ky + cko), c5b, (unsigned)ky, a.coef_t, (unsigned)a.P1, j, colv, v,
                                     s, a.ck01 ? a.ck01 + tb : nullptr, a.c6t ? a.c6t + tb : nullptr, pre);
}

// Column transforms of the tiled column kernel.  PP: two LDS exchange buffers used alternately
// (fft_wg_pp, one workgroup barrier per exchange instead of two), for the 1024-thread tiles, which
// hold one workgroup per CU anyway and have the LDS for a second buffer.
template <int N0, int V, int VAR>
struct ColPP {
  static constexpr bool value = Fft<N0, VAR>::NT * V == 1024 &&
                                (2 * LdsSize<N0, V>::slots + Fft<N0, VAR>::tw_size()) * 8 + N0 * 4 <= 152 * 1024;
};

template <int N0, int DIR, int V, int VAR, bool PP>
__device__ __forceinline__ void col_fft(float2 (&v)[(Fft<N0, VAR>::RMAX)], float2* s_buf, const float2* s_tw, int j,
                                        int vi) {
  if constexpr (PP) fft_wg_pp<N0, DIR, V, VAR>(v, s_buf, s_buf + LdsSize<N0, V>::slots, s_tw, j, vi);
  else fft_wg<N0, DIR, V, VAR>(v, s_buf, s_tw, j, vi);
}

// A source spectrum is read only in the columns where it has a nonzero mode (src_cols).
__device__ __forceinline__ bool src_col(const Col2Args& a, int ky) {
  return a.src && (!a.src_cols || a.src_cols[ky]);
}
// The same for lane vi of a tile of V in {8, 16} columns starting at c0 (a multiple of V, so the
// tile's flags lie in one 32-bit word): from the bits in the kernel arguments when present.
template <int V>
__device__ __forceinline__ bool src_col_tile(const Col2Args& a, int c0, int vi) {
  static_assert(V == 8 || V == 16, "a tile's flags in one word");
  if (!a.src) return false;
  if (a.src_bits_ok) return (a.src_bits[c0 >> 5] >> ((c0 & 31) + vi)) & 1u;
  return !a.src_cols || a.src_cols[c0 + vi];
}

// Twiddles and x wavenumbers of a column kernel's LDS, staged through registers: load() issues the
// global loads (before the tile's own loads), store() writes LDS once the tile's loads are in
// flight, so that one memory latency covers both instead of two in a row.
template <int N0, int VAR, int NTH>
struct LdsFill {
  static constexpr int NTW = Fft<N0, VAR>::tw_size();
  static constexpr int KT = (NTW + NTH - 1) / NTH, KA = (N0 + NTH - 1) / NTH;
  float2 t[KT];
  float x[KA];
  __device__ __forceinline__ void load(const float2* tw, const float* ax) {
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      const int i = (int)threadIdx.x + k * NTH;
      if (i < NTW) t[k] = tw[i];
    }
#pragma unroll
    for (int k = 0; k < KA; ++k) {
      const int i = (int)threadIdx.x + k * NTH;
      if (i < N0) x[k] = ax[i];
    }
  }
  __device__ __forceinline__ void store(float2* s_tw, float* s_ax) const {
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      const int i = (int)threadIdx.x + k * NTH;
      if (i < NTW) s_tw[i] = t[k];
    }
#pragma unroll
    for (int k = 0; k < KA; ++k) {
      const int i = (int)threadIdx.x + k * NTH;
      if (i < N0) s_ax[i] = x[k];
    }
  }
};

// Velocity-form Navier-Stokes (NSPressureConvection): both channels' convection spectra are
// transformed first, projected mode by mode (nsp_project), then combined channel by channel.
template <int N0, int V, int VAR, bool PP = false>
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
  // per-column values the projection and the combine need, loaded before the transforms
  const float ayv = colv ? a.ay[ky] : 0.f;
  const bool rs = src_col(a, ky);
  float2 v[2][RMAX];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
#pragma unroll
    for (int m = 0; m < RMAX; ++m) {
      const float2 t = a.P[((size_t)b * 2 + c) * plane + (size_t)(j + m * NT) * a.P1 + ky];
      v[c][m] = colv ? t : z;
    }
    col_fft<N0, -1, V, VAR, PP>(v[c], s_buf, s_tw, j, vi);
  }
  // Nyquist-plane parity of nsp_project (N0 is even for every fused size)
  const bool nyy = 2 * ky == a.N1, zy = ky == 0;
#pragma unroll
  for (int m = 0; m < RMAX; ++m) {
    float2 cv[2] = {v[0][m], v[1][m]};
    const int x = j + m * NT;
    const float av[2] = {a.ax[x], ayv};
    const bool nyx = 2 * x == N0, own = (x == 0 || nyx) && (zy || nyy);
    const bool ny[2] = {nyx && !own, nyy && !own};
    nsp_project<2>(cv, av, ny);
    v[0][m] = cv[0];
    v[1][m] = cv[1];
  }
  const float4* cpk = a.cpk + (size_t)b * a.cpk_bstride;
  const size_t tb = (size_t)b * (a.cpk_bstride / 2);
  const float* c5b = a.c5t + tb;
  // cpk_index(ci, half) = 2 ci + cko: every mode of this lane has ci & 15 == ky & 15
  const int cko = (stage_coef_half(stage) << 4) - (ky & 15);
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const float2* src = a.src ? a.src + (size_t)b * a.src_bstride + (size_t)c * a.src_cstride : nullptr;
    combine_column2<N0, VAR, 1>(a, stage, ((size_t)b * 2 + c) * plane, src, cpk, c5b, tb, cko, ky, colv, j, v[c],
                                s[c], rs);
  }
}

// Stage combine, fields and inverse x-transforms of column tile `tile` of sample b: the body of
// k2_col (and of the column pass of k2_run).  ST >= 0: the stage is a compile-time constant.
// FILL: the caller has issued the LDS twiddle and wavenumber loads (fill) and this function writes
// them to LDS, with a barrier, right before its first LDS use; otherwise they are in LDS already.
template <int N0, int V, int TERM, int VAR, int ST, bool CLAMP = false, bool FILL = false, bool PP = false>
__device__ __forceinline__ void col2_tile(const Col2Args& a, int stage, int tile, int b, float2* s_buf,
                                          float2* s_tw, float* s_ax, int j, int vi,
                                          const LdsFill<N0, VAR, Fft<N0, VAR>::NT * V>* fill = nullptr) {
  using F = Fft<N0, VAR>;
  constexpr int RMAX = F::RMAX, NT = F::NT;
  constexpr int NC = Term2<TERM>::NC, NF = Term2<TERM>::NF;
  const int ky = tile * V + vi;
  const bool colv = ky < a.H1;
  // the column the lane's loads address: its own, or (CLAMP, tiles wider than the pitch P1) the last
  // valid column for the lanes past it, whose results are discarded.  Without CLAMP the tiles stay
  // within the pitch (V <= 16 and P1 a multiple of 16), whose padded columns every array holds.
  const int kl = (CLAMP && !colv) ? a.H1 - 1 : ky;
  const size_t plane = (size_t)N0 * a.P1;
  const float4* cpk = a.cpk + (size_t)b * a.cpk_bstride;
  const size_t tb = (size_t)b * (a.cpk_bstride / 2);
  const float* c5b = a.c5t + tb;
  // cpk_index(ci, half) = 2 ci + cko: every mode of this lane has ci & 15 == kl & 15
  const int cko = (stage_coef_half(stage) << 4) - (kl & 15);
  // per-column values of the combine and the field pass, loaded before the transforms so that
  // their latency overlaps the product's loads (kl < P1 <= N1: in the wavenumber table)
  bool rs;
  if constexpr (!CLAMP && (V == 8 || V == 16)) rs = src_col_tile<V>(a, tile * V, vi);
  else rs = src_col(a, kl);
  const float ay = a.ay[kl];
  auto fill_lds = [&]() {
    if constexpr (FILL) {
      fill->store(s_tw, s_ax);
      __syncthreads();
    }
  };
  float2 s[NC][RMAX];
  if constexpr (TERM == TERM_NSP) {
    fill_lds();
    col_nsp_combine<N0, V, VAR, PP>(a, stage, s, s_buf, s_tw, b, ky, colv, j, vi);
  } else {
  if constexpr (TERM == TERM_CONS) fill_lds();
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const size_t base = ((size_t)b * NC + c) * plane;
    if (stage == ST_PRO) {
#pragma unroll
      for (int m = 0; m < RMAX; ++m) {
        const float2 t = a.U[base + (size_t)(j + m * NT) * a.P1 + kl];
        s[c][m] = colv ? t : make_float2(0.f, 0.f);
      }
    } else {
      float2 v[RMAX];
      // stage kernels of one-channel products whose combine reads one state array (U, or A2 in the
      // last SETDRK4 evaluation): that array and the compact coefficients are loaded before the
      // forward transform (PRE)
      constexpr int PRE = (NC == 1 && TERM != TERM_CONS)
                              ? ((ST == ST_R4_E0 || ST == ST_R4_E1) ? 1 : (ST == ST_R4_E3 ? 2 : 0))
                              : 0;
      ColPre<RMAX> pre;
      if constexpr (TERM == TERM_CONS) {
        // N_c = (i a_x) F_x(F_y(u0 u_c)) + F_x(i a_y F_y(u1 u_c))   (_conservative_convection.py:23-27)
        const size_t pb = ((size_t)b * 4 + 2 * c) * plane;
#pragma unroll
        for (int m = 0; m < RMAX; ++m) {
          const float2 t = a.P[pb + (size_t)(j + m * NT) * a.P1 + kl];
          v[m] = colv ? t : make_float2(0.f, 0.f);
        }
        col_fft<N0, -1, V, VAR, PP>(v, s_buf, s_tw, j, vi);
        // Nyquist-plane parity (nsp_project, engine.hpp): (i a_x) drops out on the x-Nyquist row,
        // (i a_y), applied by the row kernel, on the y-Nyquist column, except at modes that are
        // their own conjugate partner
        const bool nyy = 2 * ky == a.N1, zy = ky == 0;
#pragma unroll
        for (int m = 0; m < RMAX; ++m) {
          const int x = j + m * NT;
          const bool own = (x == 0 || 2 * x == N0) && (zy || nyy);
          const float axm = (2 * x == N0 && !own) ? 0.f : a.ax[x];
          v[m] = make_float2(-mulr(axm, v[m].y), mulr(axm, v[m].x));
        }
        float2 w[RMAX];
#pragma unroll
        for (int m = 0; m < RMAX; ++m) {
          const float2 t = a.P[pb + plane + (size_t)(j + m * NT) * a.P1 + kl];
          w[m] = colv ? t : make_float2(0.f, 0.f);
        }
        col_fft<N0, -1, V, VAR, PP>(w, s_buf, s_tw, j, vi);
#pragma unroll
        for (int m = 0; m < RMAX; ++m) {
          const int x = j + m * NT;
          if (nyy && !(x == 0 || 2 * x == N0)) w[m] = make_float2(0.f, 0.f);
          v[m] = radd(v[m], w[m]);
        }
      } else {
#pragma unroll
        for (int m = 0; m < RMAX; ++m) {
          const float2 t = a.P[base + (size_t)(j + m * NT) * a.P1 + kl];
          v[m] = colv ? t : make_float2(0.f, 0.f);
        }
        if constexpr (PRE > 0) {
          // the combine's state operand and compact coefficients, issued behind the product's loads
          // and before the forward transform, whose first use waits only for the product
          const unsigned o0 = (unsigned)j * a.P1 + kl, ostep = (unsigned)NT * a.P1;
          const float2* opb = (PRE == 2 ? a.A2 : a.U) + base;
#pragma unroll
          for (int m = 0; m < RMAX; ++m) pre.op[m] = opb[o0 + (unsigned)m * ostep];
          // (the stage kernels run only with both compact tables present: launch_col_t)
          if constexpr (PRE == 2) {
#pragma unroll
            for (int m = 0; m < RMAX; ++m) pre.c6[m] = a.c6t[tb + o0 + (unsigned)m * ostep];
          } else {
#pragma unroll
            for (int m = 0; m < RMAX; ++m) pre.k01[m] = a.ck01[tb + o0 + (unsigned)m * ostep];
          }
        }
        if (c == 0) fill_lds();
        col_fft<N0, -1, V, VAR, PP>(v, s_buf, s_tw, j, vi);
      }
      if constexpr (TERM == TERM_KS) {
        // remove_mean: fft(p - mean(p)) differs from fft(p) only at k = 0, by the batch mean of the DC
        if (a.dc && ky == 0 && j == 0) v[0] = rsub(v[0], a.dc[0]);
      }
      const float2* src = a.src ? a.src + (size_t)b * a.src_bstride + (size_t)c * a.src_cstride : nullptr;
      combine_column2<N0, VAR, (NC > 1 || RMAX >= 16) ? 1 : 4, (ST >= 0), PRE>(a, stage, base, src, cpk, c5b, tb,
                                                                                  cko, kl, colv, j, v, s[c], rs,
                                                                                  &pre);
    }
  }
  }
  if (tile * V > a.kmax1) return;  // uniform per workgroup: no kept column in this tile
  if (stage == ST_PRO && TERM != TERM_NSP && TERM != TERM_CONS) fill_lds();  // no transform so far
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
    col_fft<N0, +1, V, VAR, PP>(v, s_buf, s_tw, j, vi);
    if (ky <= a.kmax1) {  // kept columns only: the row kernel reads no other
      float2* g = a.G + (((size_t)b * NF + f) * N0) * a.PF + ky;
#pragma unroll
      for (int m = 0; m < RMAX; ++m) g[(size_t)(j + m * NT) * a.PF] = v[m];
    }
  }
}



// ST >= 0: the stage is a compile-time constant (only that stage's combine is compiled, which keeps
// the code a launch executes small); ST = -1: the stage comes at run time.
template <int N0, int V, int TERM, int VAR, int ST = -1>
__global__ void __launch_bounds__((Fft<N0, VAR>::NT * V)) k2_col(Col2Args a, int stage_rt) {
  const int stage = ST >= 0 ? ST : stage_rt;
  using F = Fft<N0, VAR>;
  constexpr int NT = F::NT;
  // (the run-time-stage kernel keeps one buffer: the second one's addressing would make it spill)
  constexpr bool PP = ColPP<N0, V, VAR>::value && ST >= 0;
  __shared__ float2 s_tw[Fft<N0, VAR>::tw_size()];
  __shared__ float2 s_buf[(PP ? 2 : 1) * LdsSize<N0, V>::slots];
  __shared__ float s_ax[N0];  // x wavenumbers: read per mode by the field pass without a memory wait
  // the LDS tables' loads go out first; col2_tile writes them to LDS once its own loads are issued
  LdsFill<N0, VAR, NT * V> fill;
  fill.load(a.tw0, a.ax);
  const int vi = threadIdx.x % V, j = threadIdx.x / V;
  // Block -> (tile, sample).  Workgroups are dispatched round-robin over the 8 XCDs (block v on
  // XCD v % 8); with nb % 8 == 0 each XCD owns a group of nb/8 samples and walks the tiles in
  // order, so a coefficient tile is fetched once into that XCD's L2 and reused by its samples, and
  // the heavy tiles (those with kept columns, ky <= kmax1) are dispatched before the light ones.
  // (A tile-major order, each XCD taking an eighth of the tiles for all samples so that every
  // coefficient tile is fetched by one or two XCDs only, made the 512-point columns 35 % slower: at
  // any moment every CU then reads the same band of columns, at one stride across the rows.)
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
  col2_tile<N0, V, TERM, VAR, ST, false, true, PP>(a, stage, tile, b, s_buf, s_tw, s_ax, j, vi, &fill);
}

// Row x of sample b (rv false: idle row that only takes part in the transforms' synchronisation):
// packed C2R of the fields, physical product, R2C into the products P.  The body of k2_row (and of
// the row pass of k2_run); `lds` is the row's own LDS region.
// PAIR (VorticityConvection): rows x and x + 1 (x even, same sample) by the same threads, one after
// the other up to their real products p_x, p_x+1, which then share one forward transform,
// Z = F(p_x + i p_x+1), unpacked as F(p_x)[k] = (Z[k] + conj Z[N-k]) / 2 and
// F(p_x+1)[k] = (Z[k] - conj Z[N-k]) / (2i): five transforms per two rows instead of six.
template <int N1, int TERM, int VAR, bool PAIR = false>
__device__ __forceinline__ void row2_body(const Row2Args& a, int b, int x, bool rv, float2* lds, const float2* s_tw,
                                          int j) {
  using F = Fft<N1, VAR>;
  constexpr int RMAX = F::RMAX, NT = F::NT;
  constexpr int NF = Term2<TERM>::NF;
  const size_t fs = (size_t)a.N0 * a.PF;
  const float2* gb = a.G + ((size_t)b * NF) * fs + (size_t)x * a.PF;

  if constexpr (TERM == TERM_VC) {
    constexpr bool WS = RowSync<N1, VAR>::WS;
    float2 v[RMAX];
    // one loop body for both rows of a pair (unrolled, the second row's loads and addresses are
    // hoisted above the first row's transforms: both rows' fields live at once, beyond the
    // register budget)
#pragma unroll 1
    for (int h = 0; h < (PAIR ? 2 : 1); ++h) {
      const float2* gr = gb + (size_t)h * a.PF;  // row x + h
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
          f0 = gr[k];
          f1 = gr[fs + k];
          f2 = gr[2 * fs + k];
          f3 = gr[3 * fs + k];
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
      fft_wg<N1, +1, 1, VAR, WS>(z1, lds, s_tw, j, 0);
      fft_wg<N1, +1, 1, VAR, WS>(z2, lds, s_tw, j, 0);
#pragma unroll
      for (int m = 0; m < RMAX; ++m) {
        const float pr = z1[m].x * z2[m].x + z1[m].y * z2[m].y;
        if (h == 0) v[m] = make_float2(pr, 0.f);
        else v[m].y = pr;
      }
    }
    fft_wg<N1, -1, 1, VAR, WS>(v, lds, s_tw, j, 0);
    if constexpr (PAIR) {
      // Z[N - k] of every position this lane holds, through the row's LDS region (the transform's
      // last stage read it before its closing synchronisation)
#pragma unroll
      for (int m = 0; m < RMAX; ++m) lds[lds_slot<1>(j + m * NT, 0)] = v[m];
      lds_sync<WS>();
      float2 pa[RMAX], pb[RMAX];
#pragma unroll
      for (int m = 0; m < RMAX; ++m) {
        const int pos = j + m * NT;
        const float2 zk = v[m];
        const float2 zm = lds[lds_slot<1>(pos == 0 ? 0 : N1 - pos, 0)];
        pa[m] = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
        const float2 d = make_float2(zk.x - zm.x, zk.y + zm.y);
        pb[m] = make_float2(0.5f * d.y, -0.5f * d.x);
      }
      if (rv) {
        float2* p = a.P + ((size_t)b * a.N0 + x) * a.P1;
#pragma unroll
        for (int m = 0; m < RMAX; ++m) {
          const int pos = j + m * NT;
          if (pos <= N1 / 2) {
            p[pos] = pa[m];
            p[a.P1 + pos] = pb[m];
          }
        }
      }
    } else if (rv) {
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


// row kernels that take rows in pairs sharing one forward transform (row2_body<..., PAIR>):
// 512-point VorticityConvection rows with the default radix plan (126 VGPRs against 94 unpaired;
// the other sizes' paired kernels need 120-250)
template <int TERM, int N1, int VAR>
struct Row2Pair {
  static constexpr bool value = TERM == TERM_VC && N1 == 512 && VAR == 0;
};

template <int N1, int ROWS, int TERM, int VAR>
__global__ void __launch_bounds__((Fft<N1, VAR>::NT * ROWS)) k2_row(Row2Args a) {
  using F = Fft<N1, VAR>;
  constexpr int NT = F::NT;
  constexpr int SL = LdsSize<N1, 1>::slots;
  __shared__ float2 s_tw[Fft<N1, VAR>::tw_size()];
  __shared__ float2 s_buf[ROWS * SL];
  load_twiddles<N1, VAR>(s_tw, a.tw1);
  __syncthreads();
  const int g = threadIdx.x / NT, j = threadIdx.x % NT;
  // VorticityConvection: each thread group takes a pair of rows (row2_body<..., PAIR>); N0 is even,
  // so the two rows belong to the same sample
  constexpr int RP = Row2Pair<TERM, N1, VAR>::value ? 2 : 1;
  int row = (blockIdx.x * ROWS + g) * RP;
  // with NT a multiple of 64 a wavefront holds one row: the row (and every row pointer) is uniform,
  // so the loads and stores take a scalar base and a 32-bit lane offset
  if constexpr (NT % 64 == 0) row = __builtin_amdgcn_readfirstlane(row);
  const bool in_range = row < a.B * a.N0;
  const int b = a.b0 + (in_range ? row / a.N0 : 0);
  const bool rv = in_range && !a.mask.skip(b);  // rows of samples whose steps are done stay idle
  const int x = rv ? row % a.N0 : 0;
  row2_body<N1, TERM, VAR, (RP == 2)>(a, b, x, rv, s_buf + g * SL, s_tw, j);
}

// ---------------------------------------------------------------- persistent per-sample stepper
// k2_run: for small square meshes (N = 32, 64) one workgroup per sample runs every step of a
// tfsm_step call, with the sample's fields G and products P held in LDS for the whole run: each
// evaluation's row pass (all N rows at once, one row per NT threads) reads G from LDS and writes P
// to LDS, and its column pass (one tile of N columns) reads P and writes G, with a barrier between
// them; only the state, the stage arrays and the coefficient tables are read from global memory.
// One launch covers all steps instead of two launches per evaluation.  Same device code as the
// tiled kernels (col2_tile, row2_body, with G and P addressed through flat pointers into LDS);
// compiled into another kernel, the transforms' multiply-adds may fuse differently, so results
// agree with the tiled kernels to float32 rounding rather than bit for bit.  The column tile is
// wider than the pitch, so the column pass clamps its load column (col2_tile<..., CLAMP>).
// (At 128^2 G and P do not fit LDS; through L2 the stepper took 160 us per SETDRK4 step against
// 90-130 us for the tiled kernels, and 128^2 keeps them.)
template <int N, int TERM>
struct Run2Shape {
  using F = Fft<N, 0>;
  static constexpr int NT = F::NT;               // threads per transform (rows and columns)
  static constexpr int V = N;                    // one column tile holds all N/2 + 1 columns
  static constexpr int THREADS = NT * V;         // 128 (N = 32), 512 (N = 64)
  static constexpr int ROWS = THREADS / NT;      // = N: every row in one pass
  static constexpr int PDR = F::PD < 16 ? F::PD : 16;
  static constexpr int SLR = N + N / PDR + 1;    // LDS slots of one row (padding of lds_slot<1, PD>)
  static constexpr int COL_SLOTS = LdsSize<N, V>::slots;
  static constexpr int SLOTS = COL_SLOTS > ROWS * SLR ? COL_SLOTS : ROWS * SLR;
  static constexpr int P1 = (N / 2 + 1 + 15) / 16 * 16;  // pitch of the half spectrum
  static constexpr int PF = (N / 3 + 1 + 15) / 16 * 16;  // field pitch at the 2/3 rule (kmax = N/3)
  static constexpr int NF = Term2<TERM>::NF;
  static constexpr int NP = Term2<TERM>::NC;     // product planes: VorticityConvection 1, Convection 2
  static constexpr int G_SLOTS = NF * N * PF;
  static constexpr int P_SLOTS = NP * N * P1;
  static_assert(ROWS == N, "every row in one pass");
};

template <int N, int TERM>
__global__ void __launch_bounds__((Run2Shape<N, TERM>::THREADS)) k2_run(Col2Args ca, Row2Args ra, Run2Args r) {
  static_assert(TERM == TERM_VC || TERM == TERM_CONV, "the persistent stepper's column loads are clamped for "
                                                      "Convection and VorticityConvection only");
  using S = Run2Shape<N, TERM>;
  constexpr int NT = S::NT, V = S::V;
  __shared__ float2 s_tw[Fft<N, 0>::tw_size()];
  __shared__ float2 s_buf[S::SLOTS];
  __shared__ float2 s_G[S::G_SLOTS];
  __shared__ float2 s_P[S::P_SLOTS];
  __shared__ float s_ax[N];
  const int b = r.b0 + blockIdx.x;
  int ns = r.nsteps;
  if (r.act) ns = r.act[b] < ns ? r.act[b] : ns;  // per-sample step count (uniform per workgroup)
  if (ns <= 0 && !r.rec_slot) return;
  // per-sample views with sample index 0: global arrays offset to sample b, G and P in LDS
  const size_t per = r.per_sample;
  Col2Args c = ca;
  c.U += (size_t)b * per;
  c.A1 += (size_t)b * per;
  c.A2 += (size_t)b * per;
  c.cpk += (size_t)b * ca.cpk_bstride;
  c.c5t += (size_t)b * (ca.cpk_bstride / 2);
  if (c.ck01) c.ck01 += (size_t)b * (ca.cpk_bstride / 2);
  if (c.c6t) c.c6t += (size_t)b * (ca.cpk_bstride / 2);
  c.coef += (size_t)b * ca.coef_bstride;
  if (c.src) c.src += (size_t)b * ca.src_bstride;
  c.G = s_G;
  c.P = s_P;
  c.b0 = 0;
  Row2Args w = ra;
  w.G = s_G;
  w.P = s_P;
  w.b0 = 0;
  // rows and columns share the plan: a square mesh with the same radix variant on both axes
  load_twiddles<N, 0>(s_tw, ca.tw0);
  for (int i = threadIdx.x; i < N; i += S::THREADS) s_ax[i] = ca.ax[i];
  __syncthreads();
  const int nev = scheme_evals(r.scheme);
  auto record = [&](int i) {  // the state before step i (i = ns: the final state), as k_record
    const int sl = r.rec_slot[i];
    if (sl < 0 || sl >= r.T) return;
    float2* dst = r.traj + ((size_t)b * r.T + sl) * per;
    for (size_t q = threadIdx.x; q < per; q += S::THREADS) dst[q] = c.U[q];
  };
  // pass k = 0: fields of the state (ST_PRO); k = 1 + i nev + e: evaluation e of step i
  const int npass = ns > 0 ? 1 + ns * nev : 0;
#pragma unroll 1
  for (int k = 0; k < npass; ++k) {
    // the lane indices are re-derived in every pass behind an opaque copy: otherwise the compiler
    // hoists every mode's addresses out of the step loop and keeps them live across both passes
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int vi = tid % V, jc = tid / V;  // column pass: column vi
    const int g = tid / NT, jr = tid % NT;  // row pass: row g
    const int e = k > 0 ? (k - 1) % nev : 0;
    if (k > 0) {
      __syncthreads();  // fields G of the previous column pass (and the state) complete
      if (e == 0 && r.rec_slot) record((k - 1) / nev);
      row2_body<N, TERM, 0>(w, 0, g, true, s_buf + g * S::SLR, s_tw, jr);
    }
    __syncthreads();  // products P complete; the LDS exchange buffer is reused
    col2_tile<N, V, TERM, 0, -1, true>(c, k > 0 ? scheme_stage(r.scheme, e) : ST_PRO, 0, 0, s_buf, s_tw, s_ax, jc,
                                       vi);
  }
  if (r.rec_slot) {
    __syncthreads();
    record(ns > 0 ? ns : 0);
  }
}

// ---------------------------------------------------------------- host launchers
// Tuning knobs (chosen per shape class by the autotuner): column tile V in {8, 16}, radix-plan
// variant of the column and row transforms (0: N/RMAX threads per transform; 1: half as many),
// and the row-kernel workgroup size (64 or 256 threads).
// Shapes whose column kernels are also compiled per SETDRK4 stage (and for the field pass of the
// initial state): the 512-point VorticityConvection columns of the Navier-Stokes workloads with 8
// values per thread.  One stage's kernel holds a quarter of the combine code of the run-time-stage
// kernel (13-16 KB of code instead of 21 KB) and 87-128 VGPRs; with 16 values per thread the
// stage kernels need 133-164 VGPRs (occupancy 3) and the run-time-stage kernel is kept.
template <int N0, int V, int TERM, int VAR>
struct ColStageKernels {
  static constexpr bool value = TERM == TERM_VC && N0 == 512 && VAR == 0;
};

template <int N0, int V, int TERM, int VAR>
inline hipError_t launch_col_t(const Col2Args& a, int nb, int stage, hipStream_t st) {
  const int tiles = (a.H1 + V - 1) / V;
  Col2Args b = a;
  b.nb = nb;
  const dim3 grid(tiles * nb), block(Fft<N0, VAR>::NT * V);
  // the stage kernels read SETDRK4's compact coefficient tables unconditionally
  if constexpr (ColStageKernels<N0, V, TERM, VAR>::value) {
    const bool tables = a.ck01 && a.c6t;
    switch (tables ? stage : -1) {
      case ST_PRO: TFSM_LAUNCH((k2_col<N0, V, TERM, VAR, ST_PRO>), grid, block, 0, st, b, stage); break;
      case ST_R4_E0: TFSM_LAUNCH((k2_col<N0, V, TERM, VAR, ST_R4_E0>), grid, block, 0, st, b, stage); break;
      case ST_R4_E1: TFSM_LAUNCH((k2_col<N0, V, TERM, VAR, ST_R4_E1>), grid, block, 0, st, b, stage); break;
      case ST_R4_E2: TFSM_LAUNCH((k2_col<N0, V, TERM, VAR, ST_R4_E2>), grid, block, 0, st, b, stage); break;
      case ST_R4_E3: TFSM_LAUNCH((k2_col<N0, V, TERM, VAR, ST_R4_E3>), grid, block, 0, st, b, stage); break;
      default: TFSM_LAUNCH((k2_col<N0, V, TERM, VAR>), grid, block, 0, st, b, stage);
    }
  } else {
    TFSM_LAUNCH((k2_col<N0, V, TERM, VAR>), grid, block, 0, st, b, stage);
  }
  return hipGetLastError();
}

template <int N1, int TERM, int VAR, int WGT>
inline hipError_t launch_row_t(const Row2Args& a, hipStream_t st) {
  constexpr int NT = Fft<N1, VAR>::NT;
  constexpr int ROWS = (NT >= WGT) ? 1 : WGT / NT;
  const int rows = a.B * a.N0 / (Row2Pair<TERM, N1, VAR>::value ? 2 : 1);  // thread groups (row pairs)
  TFSM_LAUNCH((k2_row<N1, ROWS, TERM, VAR>), dim3((rows + ROWS - 1) / ROWS), dim3(NT * ROWS), 0, st, a);
  return hipGetLastError();
}


// size lists of the column (N0, tile V, radix variant) and row (N1, variant, workgroup size)
// instantiations, split over the k_step2d_*.hip units to keep each compile short
#define TFSM_COL_P2(X)                                                                                        \
  X(16, 8, 0) X(16, 16, 0) X(32, 8, 0) X(32, 16, 0) X(64, 8, 0) X(64, 16, 0) X(64, 8, 1) X(64, 16, 1)          \
  X(128, 8, 0) X(128, 16, 0) X(128, 8, 1) X(128, 16, 1) X(256, 8, 0) X(256, 16, 0) X(256, 8, 1) X(256, 16, 1) \
  X(512, 8, 0) X(512, 16, 0) X(512, 8, 1) X(512, 16, 1) X(1024, 8, 0) X(1024, 8, 1)
#define TFSM_COL_MX(X)                                                                                       \
  X(48, 8, 0) X(48, 16, 0) X(96, 8, 0) X(96, 16, 0) X(100, 8, 0) X(100, 16, 0) X(112, 8, 0) X(112, 16, 0)    \
  X(192, 8, 0) X(192, 16, 0) X(200, 8, 0) X(200, 16, 0)
#define TFSM_ROW_P2(X)                                                                                        \
  X(16, 0, 256) X(32, 0, 256) X(64, 0, 256) X(64, 1, 256) X(64, 0, 64) X(64, 1, 64) X(128, 0, 256)            \
  X(128, 1, 256) X(128, 0, 64) X(128, 1, 64) X(256, 0, 256) X(256, 1, 256) X(256, 0, 64) X(256, 1, 64)        \
  X(512, 0, 256) X(512, 1, 256) X(512, 0, 64) X(512, 1, 64) X(1024, 0, 256) X(1024, 0, 64)
#define TFSM_ROW_MX(X) X(48, 0, 256) X(96, 0, 256) X(100, 0, 256) X(112, 0, 256) X(192, 0, 256) X(200, 0, 256)

#define TFSM_COL_CASE(N, VV, VR) \
  if (a.N0 == N && v == VV && var == VR) return launch_col_t<N, VV, TERM, VR>(a, nb, stage, st);
#define TFSM_ROW_CASE(N, VR, W) \
  if (a.N1 == N && var == VR && wgt == W) return launch_row_t<N, TERM, VR, W>(a, st);

template <int N, int TERM>
inline hipError_t launch_run_t(const Col2Args& ca, const Row2Args& ra, const Run2Args& r, hipStream_t st) {
  // the kernel's LDS holds G at the field pitch of the 2/3 rule and P at the mesh's pitch
  if (ca.PF > Run2Shape<N, TERM>::PF || ca.P1 != Run2Shape<N, TERM>::P1) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k2_run<N, TERM>), dim3(r.nb), dim3(Run2Shape<N, TERM>::THREADS), 0, st, ca, ra, r);
  return hipGetLastError();
}
#define TFSM_RUN_SIZES(X) X(32) X(64)
#define TFSM_RUN_CASE(NN) \
  if (ca.N0 == NN && ca.N1 == NN) return launch_run_t<NN, TERM>(ca, ra, r, st);

// per-unit entry points (k_step2d_*.hip)
hipError_t launch_run2d_vc(const Col2Args& ca, const Row2Args& ra, const Run2Args& r, hipStream_t st);
hipError_t launch_run2d_conv(const Col2Args& ca, const Row2Args& ra, const Run2Args& r, hipStream_t st);
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
