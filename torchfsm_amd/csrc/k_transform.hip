// Batched multi-dimensional FFT passes (FourierMesh.fft / .ifft, mesh.py:481-491), spectrum
// conversions at the API boundary, the exact linear stepper (ETDRK0, _etdrk.py:10-27) and the 1D
// persistent stepper.
#include "engine.hpp"

namespace tfsm {

// ---------------------------------------------------------------- contiguous rows
// real rows [nrows][in_ld] -> half spectrum rows [nrows][out_ld] (k = 0..N/2), times scale
template <int N, int ROWS>
__global__ void __launch_bounds__(Fft<N>::NT* ROWS)
    k_rows_r2c(const float* __restrict__ in, size_t in_ld, float2* __restrict__ out, size_t out_ld, int nrows,
               float scale, const float2* tw) {
  using F = Fft<N>;
  constexpr int RMAX = F::RMAX, NT = F::NT, SL = LdsSize<N, 1>::slots;
  __shared__ float2 s_tw[Fft<N>::tw_size()];
  __shared__ float2 s_buf[ROWS * SL];
  load_twiddles<N>(s_tw, tw);
  __syncthreads();
  const int g = threadIdx.x / NT, j = threadIdx.x % NT;
  const int row = blockIdx.x * ROWS + g;
  const bool rv = row < nrows;
  float2 v[RMAX];
#pragma unroll
  for (int m = 0; m < RMAX; ++m) v[m] = make_float2(rv ? in[(size_t)row * in_ld + j + m * NT] : 0.f, 0.f);
  fft_wg<N, -1, 1, 0, RowSync<N, 0>::WS>(v, s_buf + g * SL, s_tw, j, 0);
  if (rv) {
#pragma unroll
    for (int m = 0; m < RMAX; ++m) {
      const int pos = j + m * NT;
      if (pos <= N / 2) out[(size_t)row * out_ld + pos] = make_float2(v[m].x * scale, v[m].y * scale);
    }
  }
}

// half spectrum rows -> real rows, times scale (C2R: imaginary parts of k = 0 and N/2 ignored)
template <int N, int ROWS>
__global__ void __launch_bounds__(Fft<N>::NT* ROWS)
    k_rows_c2r(const float2* __restrict__ in, size_t in_ld, float* __restrict__ out, size_t out_ld, int nrows,
               float scale, const float2* tw) {
  using F = Fft<N>;
  constexpr int RMAX = F::RMAX, NT = F::NT, SL = LdsSize<N, 1>::slots;
  __shared__ float2 s_tw[Fft<N>::tw_size()];
  __shared__ float2 s_buf[ROWS * SL];
  load_twiddles<N>(s_tw, tw);
  __syncthreads();
  const int g = threadIdx.x / NT, j = threadIdx.x % NT;
  const int row = blockIdx.x * ROWS + g;
  const bool rv = row < nrows;
  float2 v[RMAX];
#pragma unroll
  for (int m = 0; m < RMAX; ++m) {
    const int pos = j + m * NT;
    const bool cj = pos > N / 2;
    const int k = cj ? N - pos : pos;
    float2 x = rv ? in[(size_t)row * in_ld + k] : make_float2(0.f, 0.f);
    if (k == 0 || k == N / 2) x.y = 0.f;
    v[m] = cj ? make_float2(x.x, -x.y) : x;
  }
  fft_wg<N, +1, 1, 0, RowSync<N, 0>::WS>(v, s_buf + g * SL, s_tw, j, 0);
  if (rv) {
#pragma unroll
    for (int m = 0; m < RMAX; ++m) out[(size_t)row * out_ld + j + m * NT] = v[m].x * scale;
  }
}

// complex rows in place (length N, pitch ld)
template <int N, int ROWS, int DIR>
__global__ void __launch_bounds__(Fft<N>::NT* ROWS)
    k_rows_c2c(const float2* in, size_t in_ld, float2* out, size_t out_ld, int nrows, float scale,
               const float2* tw) {
  using F = Fft<N>;
  constexpr int RMAX = F::RMAX, NT = F::NT, SL = LdsSize<N, 1>::slots;
  __shared__ float2 s_tw[Fft<N>::tw_size()];
  __shared__ float2 s_buf[ROWS * SL];
  load_twiddles<N>(s_tw, tw);
  __syncthreads();
  const int g = threadIdx.x / NT, j = threadIdx.x % NT;
  const int row = blockIdx.x * ROWS + g;
  const bool rv = row < nrows;
  float2 v[RMAX];
#pragma unroll
  for (int m = 0; m < RMAX; ++m) v[m] = rv ? in[(size_t)row * in_ld + j + m * NT] : make_float2(0.f, 0.f);
  fft_wg<N, DIR, 1, 0, RowSync<N, 0>::WS>(v, s_buf + g * SL, s_tw, j, 0);
  if (rv) {
#pragma unroll
    for (int m = 0; m < RMAX; ++m) out[(size_t)row * out_ld + j + m * NT] = make_float2(v[m].x * scale, v[m].y * scale);
  }
}

// ---------------------------------------------------------------- strided columns
// Element (o, p, i) at base + o*outer_stride + p*col_stride + i, p along the transformed axis,
// i in [0, ninner) contiguous.  A workgroup transforms V adjacent i of one o.  Out of place
// (in may equal out).
template <int N, int V, int DIR>
__global__ void __launch_bounds__(Fft<N>::NT* V)
    k_cols_c2c(const float2* in, float2* out, size_t outer_stride, size_t col_stride, int ninner, float scale,
               const float2* tw, StepMask mask, int b0, unsigned o0, unsigned outer_per_sample) {
  using F = Fft<N>;
  constexpr int RMAX = F::RMAX, NT = F::NT;
  __shared__ float2 s_tw[Fft<N>::tw_size()];
  __shared__ float2 s_buf[LdsSize<N, V>::slots];
  if (mask.act && mask.skip(b0 + (int)((o0 + blockIdx.y) / outer_per_sample))) return;
  load_twiddles<N>(s_tw, tw);
  __syncthreads();
  const int vi = threadIdx.x % V, j = threadIdx.x / V;
  const int i = blockIdx.x * V + vi;
  const bool iv = i < ninner;
  const size_t base = (size_t)blockIdx.y * outer_stride + i;
  float2 v[RMAX];
#pragma unroll
  for (int m = 0; m < RMAX; ++m) v[m] = iv ? in[base + (size_t)(j + m * NT) * col_stride] : make_float2(0.f, 0.f);
  fft_wg<N, DIR, V>(v, s_buf, s_tw, j, vi);
  if (iv) {
#pragma unroll
    for (int m = 0; m < RMAX; ++m)
      out[base + (size_t)(j + m * NT) * col_stride] = make_float2(v[m].x * scale, v[m].y * scale);
  }
}

// ---------------------------------------------------------------- spectrum conversions


// full [nblk][N0..N_{D-1}] -> half [nblk][N0..][P]: Hermitian projection (x + conj(x_mirror)) / 2
__global__ void k_full_to_half(const float2* __restrict__ full, float2* __restrict__ half, ConvArgs a) {
  const int nl = a.n[a.D - 1];
  const size_t inner_half = (a.D == 1 ? 1 : (size_t)a.n[0] * (a.D == 3 ? a.n[1] : 1));
  const size_t total = a.nblk * inner_half * a.P;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(e % a.P);
    size_t r = e / a.P;
    int i1 = 0, i0 = 0;
    if (a.D == 3) {
      i1 = (int)(r % a.n[1]);
      r /= a.n[1];
    }
    if (a.D >= 2) {
      i0 = (int)(r % a.n[0]);
      r /= a.n[0];
    }
    const size_t blk = r;
    float2 val = make_float2(0.f, 0.f);
    if (k < a.H) {
      size_t fb = blk;
      size_t fi = 0, fm = 0;
      const int mk = (nl - k) % nl;
      if (a.D == 1) {
        fi = fb * nl + k;
        fm = fb * nl + mk;
      } else if (a.D == 2) {
        const int m0 = (a.n[0] - i0) % a.n[0];
        fi = (fb * a.n[0] + i0) * nl + k;
        fm = (fb * a.n[0] + m0) * nl + mk;
      } else {
        const int m0 = (a.n[0] - i0) % a.n[0], m1 = (a.n[1] - i1) % a.n[1];
        fi = ((fb * a.n[0] + i0) * a.n[1] + i1) * nl + k;
        fm = ((fb * a.n[0] + m0) * a.n[1] + m1) * nl + mk;
      }
      const float2 x = full[fi], y = full[fm];
      val = make_float2(0.5f * (x.x + y.x), 0.5f * (x.y - y.y));
    }
    half[e] = val;
  }
}

// half -> full with Hermitian fill X[-i][-k] = conj(X[i][k])
__global__ void k_half_to_full(const float2* __restrict__ half, float2* __restrict__ full, ConvArgs a) {
  const int nl = a.n[a.D - 1];
  size_t cnt = a.nblk * nl;
  if (a.D >= 2) cnt *= a.n[0];
  if (a.D == 3) cnt *= a.n[1];
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < cnt; e += (size_t)gridDim.x * blockDim.x) {
    const int il = (int)(e % nl);
    size_t r = e / nl;
    int i1 = 0, i0 = 0;
    if (a.D == 3) {
      i1 = (int)(r % a.n[1]);
      r /= a.n[1];
    }
    if (a.D >= 2) {
      i0 = (int)(r % a.n[0]);
      r /= a.n[0];
    }
    const size_t blk = r;
    const bool cj = il > nl / 2;
    const int k = cj ? nl - il : il;
    const int j0 = (cj && a.D >= 2) ? (a.n[0] - i0) % a.n[0] : i0;
    const int j1 = (cj && a.D == 3) ? (a.n[1] - i1) % a.n[1] : i1;
    size_t hi;
    if (a.D == 1) hi = blk * a.P + k;
    else if (a.D == 2) hi = (blk * a.n[0] + j0) * a.P + k;
    else hi = ((blk * a.n[0] + j0) * a.n[1] + j1) * a.P + k;
    float2 x = half[hi];
    full[e] = cj ? make_float2(x.x, -x.y) : x;
  }
}

// ---------------------------------------------------------------- exact linear stepping (ETDRK0)
// U <- E U, nsteps times, recording frames before step i where rec_slot[i] is in [0, T) (and after).
__global__ void k_linear_steps(float2* __restrict__ U, const float* __restrict__ E, size_t coef_bstride,
                               size_t modes, int C, size_t total, int nsteps, const int* __restrict__ rec_slot,
                               float2* __restrict__ traj, int T) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t mode = e % modes;
    const size_t bc = e / modes;
    const size_t b = bc / C, c = bc % C;
    const float ev = E[b * coef_bstride + mode];
    float2 u = U[e];
    for (int s = 0; s <= nsteps; ++s) {
      if (rec_slot) {
        const int sl = rec_slot[s];
        if (sl >= 0 && sl < T) traj[((b * T + sl) * C + c) * modes + mode] = u;
      }
      if (s < nsteps) u = rscale(ev, u);
    }
    U[e] = u;
  }
}


// complex-L plans: U <- E U with complex E (complex64 product)
__global__ void k_linear_steps_c(float2* __restrict__ U, const float2* __restrict__ E, size_t coef_bstride,
                                 size_t modes, int C, size_t total, int nsteps, const int* __restrict__ rec_slot,
                                 float2* __restrict__ traj, int T) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t mode = e % modes;
    const size_t bc = e / modes;
    const size_t b = bc / C, c = bc % C;
    const float2 ev = E[b * coef_bstride + mode];
    float2 u = U[e];
    for (int s = 0; s <= nsteps; ++s) {
      if (rec_slot) {
        const int sl = rec_slot[s];
        if (sl >= 0 && sl < T) traj[((b * T + sl) * C + c) * modes + mode] = u;
      }
      if (s < nsteps) u = cmulr(ev, u);
    }
    U[e] = u;
  }
}

// copy the state U [B][per_sample] to traj[b][rec_slot[step]] when that slot is in [0, T)
// (_TrajRecorder.record before step i, traj_recorder.py:46-55); the schedule is read on the
// device so a captured step sequence does not depend on it.
__global__ void k_record(const float2* __restrict__ U, float2* __restrict__ traj, const int* __restrict__ rec_slot,
                         int step, int B, size_t per_sample, int T) {
  const int sl = rec_slot[step];
  if (sl < 0 || sl >= T) return;
  const size_t total = (size_t)B * per_sample;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t b = e / per_sample, r = e % per_sample;
    traj[(b * T + sl) * per_sample + r] = U[e];
  }
}

// ---------------------------------------------------------------- 1D persistent stepper
// V samples per workgroup (interleaved in LDS); every sample's state, stage spectrum and
// auxiliary arrays stay owned by fixed threads for the whole run: thread j owns modes k = j + m*NT.
// Convection (C = D = 1): N = coef_t * fft(u * u_x) [+ S], u from the dealiased stage spectrum;
// ConservativeConvection: N = coef_t * (i a) fft(u * u) [+ S].
// CPLX: complex-L plan (complex coefficient tables, stage_combine_c).
// The state U, the stage arrays A1, A2 and (real plans) the per-mode coefficient table stay in LDS for
// the whole run when they fit (LSTATE, LCOEF): the stage combine then waits on LDS instead of
// on one memory round trip per mode, which at small N is most of an evaluation.  Each lane only
// ever touches its own modes k = j + m NT of these copies, so they need no barriers of their own.
template <int N, int V, bool CPLX>
__global__ void __launch_bounds__(Fft<N>::NT* V) k1_run(Run1Args a) {
  using F = Fft<N>;
  constexpr int RMAX = F::RMAX, NT = F::NT, H = N / 2 + 1;
  constexpr bool LSTATE = (size_t)V * H * 3 * sizeof(float2) <= 16384;
  constexpr bool LCOEF = !CPLX && (size_t)V * H * NCOEF * sizeof(float) <= 20480;
  __shared__ float2 s_tw[Fft<N>::tw_size()];
  __shared__ float2 s_buf[LdsSize<N, V>::slots];
  __shared__ float2 s_stage[V * H];
  __shared__ float2 s_st[LSTATE ? 3 * V * H : 1];
  __shared__ float s_cf[LCOEF ? V * NCOEF * H : 1];
  constexpr bool LAX = (size_t)V * H * sizeof(float) <= 8192;  // wavenumbers k = 0..N/2 in LDS too
  __shared__ float s_ax[LAX ? V * H : 1];
  load_twiddles<N>(s_tw, a.tw);
  const int vi = threadIdx.x % V, j = threadIdx.x / V;
  const int smp = blockIdx.x * V + vi;
  const bool valid = smp < a.B;
  int nmax = 0;
  for (int w = 0; w < V; ++w) {
    const int sw = blockIdx.x * V + w;
    if (sw < a.B) {
      const int ns = a.nsteps_per ? a.nsteps_per[sw] : a.nsteps;
      nmax = ns > nmax ? ns : nmax;
    }
  }
  const int nst = valid ? (a.nsteps_per ? a.nsteps_per[smp] : a.nsteps) : 0;
  const size_t so = (size_t)(valid ? smp : 0) * a.P;
  float2* U = a.U + so;
  float2* A1 = a.A1 + so;
  float2* A2 = a.A2 + so;
  const float* cf = a.coef + (size_t)(valid ? smp : 0) * a.coef_bstride;
  const float2* cfc = CPLX ? a.coefc + (size_t)(valid ? smp : 0) * a.coef_bstride : nullptr;
  const float* ax = a.ax + (size_t)(valid ? smp : 0) * a.ax_bstride;
  const float2* src = a.src ? a.src + (size_t)(valid ? smp : 0) * a.src_bstride : nullptr;
  float2* stg = s_stage + vi * H;
  for (int k = j; k < H; k += NT) stg[k] = valid ? U[k] : make_float2(0.f, 0.f);
  // working copies of U, A1, A2 (written back at the end: U after stepping, A1 after an
  // evaluation-only call)
  float2* Uw = U;
  float2* A1w = A1;
  float2* A2w = A2;
  if constexpr (LSTATE) {
    Uw = s_st + vi * H;
    A1w = s_st + (V + vi) * H;
    A2w = s_st + (2 * V + vi) * H;
    for (int k = j; k < H; k += NT) Uw[k] = valid ? U[k] : make_float2(0.f, 0.f);
  }
  const float* axw = ax;
  if constexpr (LAX) {
    float* t = s_ax + (size_t)vi * H;
    for (int k = j; k < H; k += NT) t[k] = ax[k];
    axw = t;
  }
  const float* cfw = cf;
  size_t cstride = a.P;
  if constexpr (LCOEF) {
    float* t = s_cf + (size_t)vi * NCOEF * H;
    for (int sl = 0; sl < NCOEF; ++sl)
      for (int k = j; k < H; k += NT) t[sl * H + k] = cf[(size_t)sl * a.P + k];
    cfw = t;
    cstride = H;
  }
  __syncthreads();
  const int nev = a.eval_only ? 1 : scheme_evals(a.scheme);
  const int nloop = a.eval_only ? 1 : nmax;
  for (int step = 0; step < nloop; ++step) {
    const bool active = valid && step < nst;
    if (a.rec_slot && active) {
      const int sl = a.rec_slot[step];
      if (sl >= 0 && sl < a.T)
        for (int k = j; k < H; k += NT) a.traj[((size_t)smp * a.T + sl) * a.P + k] = Uw[k];
    }
    for (int e = 0; e < nev; ++e) {
      const int stage = a.eval_only ? ST_EVAL : scheme_stage(a.scheme, e);
      float2 v[RMAX];
#pragma unroll
      for (int m = 0; m < RMAX; ++m) {
        const int pos = j + m * NT;
        const bool cj = pos > N / 2;
        const int k = cj ? N - pos : pos;
        const bool edge = (k == 0) || (k == N / 2);
        float2 val = make_float2(0.f, 0.f);
        float axk = 0.f;
        if (k <= a.kmax) {
          val = stg[k];
          val = make_float2(val.x * a.scale, val.y * a.scale);
          axk = a.cons ? 0.f : axw[k];  // u alone: u_x in the imaginary part would add its round-off to u
        }
        float2 A = val;
        float2 Bv = make_float2(-axk * val.y, axk * val.x);  // i a_k u^
        if (edge) {
          A.y = 0.f;
          Bv.y = 0.f;
        }
        v[m] = cj ? make_float2(A.x + Bv.y, Bv.x - A.y) : make_float2(A.x - Bv.y, A.y + Bv.x);
      }
      fft_wg<N, +1, V>(v, s_buf, s_tw, j, vi);
      // Convection u u_x, or ConservativeConvection's u u (its derivative follows the transform)
#pragma unroll
      for (int m = 0; m < RMAX; ++m) v[m] = make_float2(a.cons ? v[m].x * v[m].x : v[m].x * v[m].y, 0.f);
      fft_wg<N, -1, V>(v, s_buf, s_tw, j, vi);
      if (a.cons) {
#pragma unroll
        for (int m = 0; m < RMAX; ++m) {
          const int pos = j + m * NT;
          const float ap = pos <= N / 2 ? axw[pos] : 0.f;
          v[m] = make_float2(-mulr(ap, v[m].y), mulr(ap, v[m].x));  // (i a) fft(u u)
        }
      }
#pragma unroll
      for (int m = 0; m < RMAX; ++m) {
        const int pos = j + m * NT;
        if (pos <= N / 2) {
          float2 sn = make_float2(0.f, 0.f);
          if (active || (a.eval_only && valid)) {
            float2 nv = nl_value(a.coef_t, v[m], src, pos);
            if constexpr (CPLX) sn = stage_combine_c(stage, nv, Uw, A1w, A2w, pos, cfc, a.P, pos);
            else sn = stage_combine(stage, nv, Uw, A1w, A2w, pos, cfw, cstride, pos, true);
          }
          stg[pos] = sn;
        }
      }
      __syncthreads();
    }
  }
  if (a.rec_slot && valid && !a.eval_only) {
    const int sl = a.rec_slot[nst];
    if (sl >= 0 && sl < a.T)
      for (int k = j; k < H; k += NT) a.traj[((size_t)smp * a.T + sl) * a.P + k] = Uw[k];
  }
  if (LSTATE && valid) {
    if (a.eval_only)
      for (int k = j; k < H; k += NT) A1[k] = A1w[k];
    else
      for (int k = j; k < H; k += NT) U[k] = Uw[k];
  }
}

// ---------------------------------------------------------------- 1D KSConvection, one stage per launch
// KSConvection's remove_mean subtracts the mean over the whole batch, so the stages of different
// samples are coupled through the DC mode: each stage is evaluated for every sample (raw product
// spectrum), the batch mean of the DC is reduced, then the stage combine runs.
// Nraw = fft(1/2 (d_x phi)^2) of the dealiased stage spectrum S (_ks_convection.py:27-38).
template <int N, int V>
__global__ void __launch_bounds__(Fft<N>::NT* V) k1_eval_ks(Run1Args a, const float2* S, float2* Nraw) {
  using F = Fft<N>;
  constexpr int RMAX = F::RMAX, NT = F::NT;
  __shared__ float2 s_tw[Fft<N>::tw_size()];
  __shared__ float2 s_buf[LdsSize<N, V>::slots];
  load_twiddles<N>(s_tw, a.tw);
  __syncthreads();
  const int vi = threadIdx.x % V, j = threadIdx.x / V;
  const int smp = blockIdx.x * V + vi;
  const bool valid = smp < a.B;
  const float2* Sb = S + (size_t)(valid ? smp : 0) * a.P;
  const float* ax = a.ax + (size_t)(valid ? smp : 0) * a.ax_bstride;
  float2 v[RMAX];
#pragma unroll
  for (int m = 0; m < RMAX; ++m) {
    const int pos = j + m * NT;
    const bool cj = pos > N / 2;
    const int k = cj ? N - pos : pos;
    float2 g = make_float2(0.f, 0.f);
    if (valid && k <= a.kmax) {
      const float2 val = make_float2(Sb[k].x * a.scale, Sb[k].y * a.scale);
      const float axk = ax[k];
      g = make_float2(-axk * val.y, axk * val.x);  // i a_k phi^
    }
    if (k == 0 || k == N / 2) g.y = 0.f;
    v[m] = cj ? make_float2(g.x, -g.y) : g;
  }
  fft_wg<N, +1, V>(v, s_buf, s_tw, j, vi);
#pragma unroll
  for (int m = 0; m < RMAX; ++m) v[m] = make_float2(mulr(0.5f, mulr(v[m].x, v[m].x)), 0.f);
  fft_wg<N, -1, V>(v, s_buf, s_tw, j, vi);
  if (valid) {
#pragma unroll
    for (int m = 0; m < RMAX; ++m) {
      const int pos = j + m * NT;
      if (pos <= N / 2) Nraw[(size_t)smp * a.P + pos] = v[m];
    }
  }
}

// stage combine of every (sample, k <= N/2): N = coef_t (Nraw - [k = 0] dc) + S_src, then the
// stage's update of U / A1 / A2; S receives the next stage spectrum
__global__ void k1_stage(Run1Args a, int stage, const float2* __restrict__ Nraw, const float2* __restrict__ dc,
                         float2* __restrict__ S) {
  const size_t total = (size_t)a.B * a.H;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t smp = e / a.H;
    const int pos = (int)(e % a.H);
    const size_t so = smp * a.P;
    float2 v = Nraw[so + pos];
    if (pos == 0 && dc) v = rsub(v, dc[0]);
    const float2* src = a.src ? a.src + smp * a.src_bstride : nullptr;
    const float2 nv = nl_value(a.coef_t, v, src, pos);
    if (a.coefc)
      S[so + pos] = stage_combine_c(stage, nv, a.U + so, a.A1 + so, a.A2 + so, pos, a.coefc + smp * a.coef_bstride,
                                    a.P, pos);
    else
      S[so + pos] = stage_combine(stage, nv, a.U + so, a.A1 + so, a.A2 + so, pos, a.coef + smp * a.coef_bstride, a.P,
                                  pos, true);
  }
}

// ---------------------------------------------------------------- host launchers
template <int N>
static constexpr int rows_per_wg() {
  return Fft<N>::NT >= 256 ? 1 : 256 / Fft<N>::NT;
}

#define TFSM_SWITCH_N(n, BODY)  \
  switch (n) {                  \
    case 16: { constexpr int NN = 16; BODY; } \
    case 32: { constexpr int NN = 32; BODY; } \
    case 64: { constexpr int NN = 64; BODY; } \
    case 128: { constexpr int NN = 128; BODY; } \
    case 256: { constexpr int NN = 256; BODY; } \
    case 512: { constexpr int NN = 512; BODY; } \
    case 1024: { constexpr int NN = 1024; BODY; } \
    case 2048: { constexpr int NN = 2048; BODY; } \
    case 4096: { constexpr int NN = 4096; BODY; } \
    case 48: { constexpr int NN = 48; BODY; } \
    case 96: { constexpr int NN = 96; BODY; } \
    case 100: { constexpr int NN = 100; BODY; } \
    case 112: { constexpr int NN = 112; BODY; } \
    case 192: { constexpr int NN = 192; BODY; } \
    case 200: { constexpr int NN = 200; BODY; } \
    case 384: { constexpr int NN = 384; BODY; } \
    case 400: { constexpr int NN = 400; BODY; } \
    case 768: { constexpr int NN = 768; BODY; } \
    case 800: { constexpr int NN = 800; BODY; } \
    default: return hipErrorInvalidValue; \
  }

hipError_t launch_rows_r2c(int n, const float* in, size_t in_ld, float2* out, size_t out_ld, int nrows, float scale,
                           const float2* tw, hipStream_t st) {
  if (nrows <= 0) return hipSuccess;
  TFSM_SWITCH_N(n, {
    constexpr int R = rows_per_wg<NN>();
    hipLaunchKernelGGL((k_rows_r2c<NN, R>), dim3((nrows + R - 1) / R), dim3(Fft<NN>::NT * R), 0, st, in, in_ld, out,
                       out_ld, nrows, scale, tw);
    return hipGetLastError();
  })
}

hipError_t launch_rows_c2r(int n, const float2* in, size_t in_ld, float* out, size_t out_ld, int nrows, float scale,
                           const float2* tw, hipStream_t st) {
  if (nrows <= 0) return hipSuccess;
  TFSM_SWITCH_N(n, {
    constexpr int R = rows_per_wg<NN>();
    hipLaunchKernelGGL((k_rows_c2r<NN, R>), dim3((nrows + R - 1) / R), dim3(Fft<NN>::NT * R), 0, st, in, in_ld, out,
                       out_ld, nrows, scale, tw);
    return hipGetLastError();
  })
}

hipError_t launch_rows_c2c(int n, int dir, const float2* in, size_t in_ld, float2* out, size_t out_ld, int nrows,
                           float scale, const float2* tw, hipStream_t st) {
  if (nrows <= 0) return hipSuccess;
  TFSM_SWITCH_N(n, {
    constexpr int R = rows_per_wg<NN>();
    if (dir < 0)
      hipLaunchKernelGGL((k_rows_c2c<NN, R, -1>), dim3((nrows + R - 1) / R), dim3(Fft<NN>::NT * R), 0, st, in, in_ld,
                         out, out_ld, nrows, scale, tw);
    else
      hipLaunchKernelGGL((k_rows_c2c<NN, R, 1>), dim3((nrows + R - 1) / R), dim3(Fft<NN>::NT * R), 0, st, in, in_ld,
                         out, out_ld, nrows, scale, tw);
    return hipGetLastError();
  })
}

hipError_t launch_cols_c2c(int n, int dir, const float2* in, float2* out, size_t nouter, size_t outer_stride,
                           size_t col_stride, int ninner, float scale, const float2* tw, hipStream_t st,
                           StepMask mask, int b0, size_t outer_per_sample) {
  if (nouter == 0 || ninner <= 0) return hipSuccess;
  if (n > 512) return hipErrorInvalidValue;
  if (mask.act && (outer_per_sample == 0 || nouter > 0xffffffffu)) return hipErrorInvalidValue;
  TFSM_SWITCH_N(n, {
    constexpr int V = (NN > 512) ? 1 : ((Fft<NN>::NT * 16 <= 1024) ? 16 : 8);
    // grid.y is limited to 65535: launch the outer blocks in chunks
    for (size_t o0 = 0; o0 < nouter; o0 += 65535) {
      const size_t no = (nouter - o0) < 65535 ? (nouter - o0) : 65535;
      dim3 grid((ninner + V - 1) / V, (unsigned)no);
      const float2* ib = in + o0 * outer_stride;
      float2* ob = out + o0 * outer_stride;
      const unsigned ops = (unsigned)(outer_per_sample ? outer_per_sample : 1);
      if (dir < 0)
        TFSM_LAUNCH((k_cols_c2c<NN, V, -1>), grid, dim3(Fft<NN>::NT * V), 0, st, ib, ob, outer_stride,
                           col_stride, ninner, scale, tw, mask, b0, (unsigned)o0, ops);
      else
        TFSM_LAUNCH((k_cols_c2c<NN, V, 1>), grid, dim3(Fft<NN>::NT * V), 0, st, ib, ob, outer_stride,
                           col_stride, ninner, scale, tw, mask, b0, (unsigned)o0, ops);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  })
}

hipError_t launch_full_to_half(const float2* full, float2* half, const ConvArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(k_full_to_half, dim3(1024), dim3(256), 0, st, full, half, a);
  return hipGetLastError();
}

hipError_t launch_half_to_full(const float2* half, float2* full, const ConvArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(k_half_to_full, dim3(1024), dim3(256), 0, st, half, full, a);
  return hipGetLastError();
}

hipError_t launch_linear_steps(float2* U, const float* E, size_t coef_bstride, size_t modes, int C, size_t total,
                               int nsteps, const int* rec_slot, float2* traj, int T, hipStream_t st) {
  unsigned blocks = (unsigned)((total + 255) / 256);
  if (blocks > 65535u * 4) blocks = 65535u * 4;
  hipLaunchKernelGGL(k_linear_steps, dim3(blocks), dim3(256), 0, st, U, E, coef_bstride, modes, C, total, nsteps,
                     rec_slot, traj, T);
  return hipGetLastError();
}

hipError_t launch_linear_steps_c(float2* U, const float2* E, size_t coef_bstride, size_t modes, int C, size_t total,
                                 int nsteps, const int* rec_slot, float2* traj, int T, hipStream_t st) {
  unsigned blocks = (unsigned)((total + 255) / 256);
  if (blocks > 65535u * 4) blocks = 65535u * 4;
  hipLaunchKernelGGL(k_linear_steps_c, dim3(blocks), dim3(256), 0, st, U, E, coef_bstride, modes, C, total, nsteps,
                     rec_slot, traj, T);
  return hipGetLastError();
}

hipError_t launch_record(const float2* U, float2* traj, const int* rec_slot, int step, int B, size_t per_sample,
                         int T, hipStream_t st) {
  size_t total = (size_t)B * per_sample;
  unsigned blocks = (unsigned)((total + 255) / 256);
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(k_record, dim3(blocks), dim3(256), 0, st, U, traj, rec_slot, step, B, per_sample, T);
  return hipGetLastError();
}

template <int N>
static hipError_t launch_run1_t(const Run1Args& a, hipStream_t st) {
  constexpr int NT = Fft<N>::NT;
  constexpr int V = NT >= 64 ? 1 : 64 / NT;
  if (a.coefc)
    hipLaunchKernelGGL((k1_run<N, V, true>), dim3((a.B + V - 1) / V), dim3(NT * V), 0, st, a);
  else
    hipLaunchKernelGGL((k1_run<N, V, false>), dim3((a.B + V - 1) / V), dim3(NT * V), 0, st, a);
  return hipGetLastError();
}

template <int N>
static hipError_t launch_eval_ks_t(const Run1Args& a, const float2* S, float2* Nraw, hipStream_t st) {
  constexpr int NT = Fft<N>::NT;
  constexpr int V = NT >= 64 ? 1 : 64 / NT;
  hipLaunchKernelGGL((k1_eval_ks<N, V>), dim3((a.B + V - 1) / V), dim3(NT * V), 0, st, a, S, Nraw);
  return hipGetLastError();
}

hipError_t launch_k1_eval_ks(const Run1Args& a, const float2* S, float2* Nraw, hipStream_t st) {
  TFSM_SWITCH_N(a.N, { return launch_eval_ks_t<NN>(a, S, Nraw, st); })
}

hipError_t launch_k1_stage(const Run1Args& a, int stage, const float2* Nraw, const float2* dc, float2* S,
                           hipStream_t st) {
  const size_t total = (size_t)a.B * a.H;
  unsigned blocks = (unsigned)((total + 255) / 256);
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k1_stage, dim3(blocks), dim3(256), 0, st, a, stage, Nraw, dc, S);
  return hipGetLastError();
}

hipError_t launch_run1d(const Run1Args& a, hipStream_t st) {
  switch (a.N) {
    case 32: return launch_run1_t<32>(a, st);
    case 64: return launch_run1_t<64>(a, st);
    case 128: return launch_run1_t<128>(a, st);
    case 256: return launch_run1_t<256>(a, st);
    case 512: return launch_run1_t<512>(a, st);
    case 1024: return launch_run1_t<1024>(a, st);
    case 2048: return launch_run1_t<2048>(a, st);
    case 4096: return launch_run1_t<4096>(a, st);
    case 48: return launch_run1_t<48>(a, st);
    case 96: return launch_run1_t<96>(a, st);
    case 100: return launch_run1_t<100>(a, st);
    case 112: return launch_run1_t<112>(a, st);
    case 192: return launch_run1_t<192>(a, st);
    case 200: return launch_run1_t<200>(a, st);
    case 384: return launch_run1_t<384>(a, st);
    case 400: return launch_run1_t<400>(a, st);
    case 768: return launch_run1_t<768>(a, st);
    case 800: return launch_run1_t<800>(a, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace tfsm
