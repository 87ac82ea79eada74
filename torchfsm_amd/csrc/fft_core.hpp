// Workgroup-cooperative Stockham FFT building blocks for gfx950 (CDNA4).
//
// A transform of length N is carried by NT = N / RMAX threads.  Thread j holds, in registers,
// the RMAX elements at positions j + m*NT (m = 0..RMAX-1).  Every Stockham stage reads its R inputs
// at stride N/R and writes its outputs to expand(jb) + r*Ns; with the register layout above the
// first stage reads straight from registers and the last stage writes straight back to them in
// natural order, so a transform of S stages needs S-1 LDS exchanges and no bit reversal.
//
// Several transforms can share a workgroup in two ways:
//   * V interleaved transforms (strided columns of an array): position p of transform vi lives at
//     LDS slot p*(V+PAD) + vi, so the V lanes of one position are contiguous and a wavefront moves
//     whole 128-byte rows of a column tile (the HBM layout of the tile, no transpose needed);
//   * G independent groups (contiguous rows), each with its own LDS region.
//
// Lengths may mix radices 2, 4, 8, 16 with 3, 5, 7 (Stockham needs no particular order of them).
// Twiddles come from per-stage tables (Fft<N>::tw_offset, computed in double on the host, rounded
// once), staged in LDS at kernel start.
#pragma once
#include <hip/hip_runtime.h>

namespace tfsm {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by +i (DIR>0) or -i (DIR<0)
template <int DIR>
__device__ __forceinline__ float2 mul_i(float2 a) {
  if constexpr (DIR > 0) return make_float2(-a.y, a.x);
  else return make_float2(a.y, -a.x);
}

// ---------------------------------------------------------------- in-register DFTs (natural order)
// DIR = -1: forward (exp(-2 pi i ...)), DIR = +1: inverse (unnormalised).
template <int DIR>
__device__ __forceinline__ void dft2(float2* x) {
  float2 t = x[0];
  x[0] = cadd(t, x[1]);
  x[1] = csub(t, x[1]);
}

template <int DIR>
__device__ __forceinline__ void dft4(float2& x0, float2& x1, float2& x2, float2& x3) {
  float2 t0 = cadd(x0, x2), t1 = csub(x0, x2);
  float2 t2 = cadd(x1, x3), t3 = mul_i<DIR>(csub(x1, x3));
  x0 = cadd(t0, t2);
  x2 = csub(t0, t2);
  x1 = cadd(t1, t3);
  x3 = csub(t1, t3);
}

template <int DIR>
__device__ __forceinline__ void dft4(float2* x) { dft4<DIR>(x[0], x[1], x[2], x[3]); }

template <int DIR>
__device__ __forceinline__ void dft8(float2* x) {
  constexpr float c = 0.70710678118654752440f;
  float2 e0 = x[0], e1 = x[2], e2 = x[4], e3 = x[6];
  float2 o0 = x[1], o1 = x[3], o2 = x[5], o3 = x[7];
  dft4<DIR>(e0, e1, e2, e3);
  dft4<DIR>(o0, o1, o2, o3);
  // W8^1 = (c, DIR c), W8^2 = DIR i, W8^3 = (-c, DIR c)
  float2 w1 = make_float2(c * (o1.x - DIR * o1.y), c * (o1.y + DIR * o1.x));
  float2 w2 = mul_i<DIR>(o2);
  float2 w3 = make_float2(-c * (o3.x + DIR * o3.y), c * (DIR * o3.x - o3.y));
  x[0] = cadd(e0, o0);
  x[4] = csub(e0, o0);
  x[1] = cadd(e1, w1);
  x[5] = csub(e1, w1);
  x[2] = cadd(e2, w2);
  x[6] = csub(e2, w2);
  x[3] = cadd(e3, w3);
  x[7] = csub(e3, w3);
}

template <int DIR>
__device__ __forceinline__ void dft16(float2* x) {
  float2 e[8], o[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    e[i] = x[2 * i];
    o[i] = x[2 * i + 1];
  }
  dft8<DIR>(e);
  dft8<DIR>(o);
  // W16^k = cos(2 pi k/16) + DIR i sin(2 pi k/16)
  constexpr float C[8] = {1.0f, 0.92387953251128675613f, 0.70710678118654752440f, 0.38268343236508977173f,
                          0.0f, -0.38268343236508977173f, -0.70710678118654752440f, -0.92387953251128675613f};
  constexpr float S[8] = {0.0f, 0.38268343236508977173f, 0.70710678118654752440f, 0.92387953251128675613f,
                          1.0f, 0.92387953251128675613f, 0.70710678118654752440f, 0.38268343236508977173f};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    float2 w = make_float2(C[k], DIR * S[k]);
    float2 t = (k == 0) ? o[0] : cmul(o[k], w);
    x[k] = cadd(e[k], t);
    x[k + 8] = csub(e[k], t);
  }
}

// Odd prime radix R in {3, 5, 7}, direct from the definition folded over the conjugate pairs
// (n, R-n):  X_k = x_0 + sum_{n=1}^{(R-1)/2} [ (x_n + x_{R-n}) cos(2 pi n k/R)
//                                            + DIR i (x_n - x_{R-n}) sin(2 pi n k/R) ],
// X_{R-k} the same with the sine part negated.
template <int R>
struct OddRoots;
template <>
struct OddRoots<3> {
  static constexpr float C[3] = {1.0f, -0.5f, -0.5f};
  static constexpr float S[3] = {0.0f, 0.86602540378443864676f, -0.86602540378443864676f};
};
template <>
struct OddRoots<5> {
  static constexpr float C[5] = {1.0f, 0.30901699437494742410f, -0.80901699437494742410f, -0.80901699437494742410f,
                                 0.30901699437494742410f};
  static constexpr float S[5] = {0.0f, 0.95105651629515357212f, 0.58778525229247312917f, -0.58778525229247312917f,
                                 -0.95105651629515357212f};
};
template <>
struct OddRoots<7> {
  static constexpr float C[7] = {1.0f,
                                 0.62348980185873353053f,
                                 -0.22252093395631440429f,
                                 -0.90096886790241912624f,
                                 -0.90096886790241912624f,
                                 -0.22252093395631440429f,
                                 0.62348980185873353053f};
  static constexpr float S[7] = {0.0f,
                                 0.78183148246802980871f,
                                 0.97492791218182360702f,
                                 0.43388373911755812048f,
                                 -0.43388373911755812048f,
                                 -0.97492791218182360702f,
                                 -0.78183148246802980871f};
};

template <int R, int DIR>
__device__ __forceinline__ void dft_odd(float2* x) {
  constexpr int H = (R - 1) / 2;
  float2 t[H + 1], d[H + 1];
#pragma unroll
  for (int n = 1; n <= H; ++n) {
    t[n] = cadd(x[n], x[R - n]);
    d[n] = csub(x[n], x[R - n]);
  }
  float2 out[R];
  float2 s0 = x[0];
#pragma unroll
  for (int n = 1; n <= H; ++n) s0 = cadd(s0, t[n]);
  out[0] = s0;
#pragma unroll
  for (int k = 1; k <= H; ++k) {
    float2 a = x[0], b = make_float2(0.f, 0.f);
#pragma unroll
    for (int n = 1; n <= H; ++n) {
      const float c = OddRoots<R>::C[(n * k) % R], sn = OddRoots<R>::S[(n * k) % R];
      a = make_float2(a.x + c * t[n].x, a.y + c * t[n].y);
      b = make_float2(b.x + sn * d[n].x, b.y + sn * d[n].y);
    }
    const float2 ib = mul_i<DIR>(b);
    out[k] = cadd(a, ib);
    out[R - k] = csub(a, ib);
  }
#pragma unroll
  for (int r = 0; r < R; ++r) x[r] = out[r];
}

template <int R, int DIR>
__device__ __forceinline__ void dft(float2* x) {
  if constexpr (R == 2) dft2<DIR>(x);
  else if constexpr (R == 4) dft4<DIR>(x);
  else if constexpr (R == 8) dft8<DIR>(x);
  else if constexpr (R == 16) dft16<DIR>(x);
  else if constexpr (R == 3 || R == 5 || R == 7) dft_odd<R, DIR>(x);
  else static_assert(R == 2, "unsupported radix");
}

// ---------------------------------------------------------------- radix plans
// FftPlan<N>: the radix sequence used for a transform of length N, RMAX elements per thread.
template <int N, int VAR = 0>
struct FftPlan;
// VAR 0: default plans
template <> struct FftPlan<16, 0> { static constexpr int RMAX = 4; static constexpr int NST = 2; static constexpr int R[5] = {4, 4, 0, 0, 0}; };
template <> struct FftPlan<32, 0> { static constexpr int RMAX = 8; static constexpr int NST = 2; static constexpr int R[5] = {8, 4, 0, 0, 0}; };
template <> struct FftPlan<64, 0> { static constexpr int RMAX = 8; static constexpr int NST = 2; static constexpr int R[5] = {8, 8, 0, 0, 0}; };
template <> struct FftPlan<128, 0> { static constexpr int RMAX = 16; static constexpr int NST = 2; static constexpr int R[5] = {16, 8, 0, 0, 0}; };
template <> struct FftPlan<256, 0> { static constexpr int RMAX = 16; static constexpr int NST = 2; static constexpr int R[5] = {16, 16, 0, 0, 0}; };
template <> struct FftPlan<512, 0> { static constexpr int RMAX = 8; static constexpr int NST = 3; static constexpr int R[5] = {8, 8, 8, 0, 0}; };
template <> struct FftPlan<1024, 0> { static constexpr int RMAX = 16; static constexpr int NST = 3; static constexpr int R[5] = {16, 8, 8, 0, 0}; };
template <> struct FftPlan<2048, 0> { static constexpr int RMAX = 16; static constexpr int NST = 3; static constexpr int R[5] = {16, 16, 8, 0, 0}; };
template <> struct FftPlan<4096, 0> { static constexpr int RMAX = 16; static constexpr int NST = 3; static constexpr int R[5] = {16, 16, 16, 0, 0}; };
// Mixed-radix lengths (even N with factors 3, 5, 7): RMAX is a common multiple of the radices,
// so every stage's R divides the values a thread holds.
template <> struct FftPlan<48, 0> { static constexpr int RMAX = 12; static constexpr int NST = 3; static constexpr int R[5] = {4, 4, 3, 0, 0}; };
template <> struct FftPlan<96, 0> { static constexpr int RMAX = 12; static constexpr int NST = 4; static constexpr int R[5] = {4, 4, 2, 3, 0}; };
template <> struct FftPlan<100, 0> { static constexpr int RMAX = 20; static constexpr int NST = 3; static constexpr int R[5] = {4, 5, 5, 0, 0}; };
template <> struct FftPlan<112, 0> { static constexpr int RMAX = 28; static constexpr int NST = 3; static constexpr int R[5] = {4, 4, 7, 0, 0}; };
template <> struct FftPlan<192, 0> { static constexpr int RMAX = 12; static constexpr int NST = 4; static constexpr int R[5] = {4, 4, 4, 3, 0}; };
template <> struct FftPlan<200, 0> { static constexpr int RMAX = 20; static constexpr int NST = 4; static constexpr int R[5] = {4, 2, 5, 5, 0}; };
template <> struct FftPlan<384, 0> { static constexpr int RMAX = 12; static constexpr int NST = 5; static constexpr int R[5] = {4, 4, 4, 2, 3}; };
template <> struct FftPlan<400, 0> { static constexpr int RMAX = 20; static constexpr int NST = 4; static constexpr int R[5] = {4, 4, 5, 5, 0}; };
template <> struct FftPlan<768, 0> { static constexpr int RMAX = 12; static constexpr int NST = 5; static constexpr int R[5] = {4, 4, 4, 4, 3}; };
template <> struct FftPlan<800, 0> { static constexpr int RMAX = 20; static constexpr int NST = 5; static constexpr int R[5] = {4, 4, 2, 5, 5}; };
// VAR 1: twice the elements per thread (half the threads per transform), same radix sequence
// where possible; lets a column tile of the same width run with half the workgroup size.
template <> struct FftPlan<64, 1> { static constexpr int RMAX = 16; static constexpr int NST = 2; static constexpr int R[5] = {8, 8, 0, 0, 0}; };
template <> struct FftPlan<128, 1> { static constexpr int RMAX = 16; static constexpr int NST = 3; static constexpr int R[5] = {8, 4, 4, 0, 0}; };
template <> struct FftPlan<256, 1> { static constexpr int RMAX = 8; static constexpr int NST = 3; static constexpr int R[5] = {8, 8, 4, 0, 0}; };
template <> struct FftPlan<512, 1> { static constexpr int RMAX = 16; static constexpr int NST = 3; static constexpr int R[5] = {8, 8, 8, 0, 0}; };
// (1024: half the elements per thread of VAR 0, for 2D columns of 1024 points)
template <> struct FftPlan<1024, 1> { static constexpr int RMAX = 8; static constexpr int NST = 4; static constexpr int R[5] = {8, 8, 4, 4, 0}; };
// VAR 2: 8 elements per thread where VAR 0 holds 16 (lower register use in the 3D kernels)
template <> struct FftPlan<128, 2> { static constexpr int RMAX = 8; static constexpr int NST = 3; static constexpr int R[5] = {8, 4, 4, 0, 0}; };
// VAR 3: 4 elements per thread (radix 4, one radix-2 stage where needed): the fewest registers
template <> struct FftPlan<16, 3> { static constexpr int RMAX = 4; static constexpr int NST = 2; static constexpr int R[5] = {4, 4, 0, 0, 0}; };
template <> struct FftPlan<32, 3> { static constexpr int RMAX = 4; static constexpr int NST = 3; static constexpr int R[5] = {4, 4, 2, 0, 0}; };
template <> struct FftPlan<64, 3> { static constexpr int RMAX = 4; static constexpr int NST = 3; static constexpr int R[5] = {4, 4, 4, 0, 0}; };
template <> struct FftPlan<128, 3> { static constexpr int RMAX = 4; static constexpr int NST = 4; static constexpr int R[5] = {4, 4, 4, 2, 0}; };

template <int N, int VAR = 0>
struct Fft {
  static constexpr int RMAX = FftPlan<N, VAR>::RMAX;
  static constexpr int NT = N / RMAX;  // threads per transform
  static constexpr int NST = FftPlan<N, VAR>::NST;
  static constexpr int radix(int s) { return FftPlan<N, VAR>::R[s]; }
  static constexpr int ns_before(int s) {
    int ns = 1;
    for (int i = 0; i < s; ++i) ns *= FftPlan<N, VAR>::R[i];
    return ns;
  }
  // Twiddle table layout: stage s >= 1 holds W^(r k), r < R_s, k < Ns_s, at tw_offset(s) + r*Ns + k
  // with W = exp(-2 pi i / (Ns R)).  Lanes of one stage read consecutive k (or broadcast), so the
  // table reads are free of LDS bank conflicts.
  static constexpr int tw_offset(int s) {
    int off = 0;
    for (int i = 1; i < s; ++i) off += radix(i) * ns_before(i);
    return off;
  }
  static constexpr int tw_size() { return tw_offset(NST) > 0 ? tw_offset(NST) : 1; }
  static constexpr int PD = FftPlan<N, VAR>::R[0];  // LDS padding period of single transforms
  static constexpr bool plan_ok() {
    int p = 1;
    for (int i = 0; i < NST; ++i) {
      if (RMAX % radix(i) != 0) return false;
      p *= radix(i);
    }
    return p == N && NST <= 5;
  }
  static_assert(plan_ok(), "radix plan must multiply to N and each radix must divide RMAX");
};

// LDS slot of position p of interleaved transform vi (V transforms interleaved).  A single
// transform (V = 1) is padded by one slot per PD positions, PD = the plan's first radix: the
// first stage writes at stride PD and the second at stride 1 within groups of PD, and with this
// padding both fill all banks without conflicts (PD is a power of two).
template <int V, int PD = 16>
__device__ __forceinline__ int lds_slot(int p, int vi) {
  if constexpr (V == 1) return p + (int)((unsigned)p / PD);
  else return p * (V + 1) + vi;
}
// slots for any plan (PD >= 4)
template <int N, int V>
struct LdsSize {
  static constexpr int slots = (V == 1) ? (N + N / 4) : N * (V + 1);
};

// Synchronisation between the LDS phases of a transform.  WS = false: the transform spans
// several wavefronts (workgroup barrier).  WS = true: every transform lives inside one wavefront
// (rows with N / RMAX <= 64 threads, 64 a multiple of it): LDS operations of one wavefront execute
// in order, so a compiler fence at wavefront scope is all that is needed.
template <bool WS>
__device__ __forceinline__ void lds_sync() {
  if constexpr (WS) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// One Stockham stage (radix R, Ns = product of earlier radices).  See header comment.
template <int N, int S, int DIR, int V, int VAR = 0, bool WS = false>
__device__ __forceinline__ void fft_stage(float2 (&v)[(Fft<N, VAR>::RMAX)], float2* lds, const float2* tw, int j,
                                          int vi) {
  using F = Fft<N, VAR>;
  constexpr int RMAX = F::RMAX;
  constexpr int NT = F::NT;
  constexpr int R = F::radix(S);
  constexpr int NS = F::ns_before(S);
  constexpr int Q = RMAX / R;
  constexpr bool FIRST = (S == 0);
  constexpr bool LAST = (S == F::NST - 1);
  constexpr int TWO = F::tw_offset(S);
  if constexpr (!FIRST) {
#pragma unroll
    for (int m = 0; m < RMAX; ++m) v[m] = lds[lds_slot<V, F::PD>(j + m * NT, vi)];
    lds_sync<WS>();
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int jb = j + q * NT;
    float2 x[R];
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = v[q + r * Q];
    if constexpr (NS > 1) {
      const int k = jb % NS;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 w = tw[TWO + r * NS + k];
        x[r] = (DIR < 0) ? cmul(x[r], w) : cmulc(x[r], w);
      }
    }
    dft<R, DIR>(x);
    if constexpr (LAST) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[q + r * Q] = x[r];
    } else {
      const int idxd = (jb / NS) * NS * R + (jb % NS);
#pragma unroll
      for (int r = 0; r < R; ++r) lds[lds_slot<V, F::PD>(idxd + r * NS, vi)] = x[r];
    }
  }
  if constexpr (!LAST) lds_sync<WS>();
}

// Full transform: registers (positions j + m*NT) -> registers (same positions, natural order).
// All threads of the workgroup must call it (it contains barriers).
template <int N, int DIR, int V, int VAR = 0, bool WS = false>
__device__ __forceinline__ void fft_wg(float2 (&v)[(Fft<N, VAR>::RMAX)], float2* lds, const float2* tw, int j,
                                       int vi) {
  using F = Fft<N, VAR>;
  static_assert(!WS || (V == 1 && F::NT <= 64 && 64 % F::NT == 0), "wave-local transforms are rows of one wavefront");
  fft_stage<N, 0, DIR, V, VAR, WS>(v, lds, tw, j, vi);
  if constexpr (F::NST > 1) fft_stage<N, 1, DIR, V, VAR, WS>(v, lds, tw, j, vi);
  if constexpr (F::NST > 2) fft_stage<N, 2, DIR, V, VAR, WS>(v, lds, tw, j, vi);
  if constexpr (F::NST > 3) fft_stage<N, 3, DIR, V, VAR, WS>(v, lds, tw, j, vi);
  if constexpr (F::NST > 4) fft_stage<N, 4, DIR, V, VAR, WS>(v, lds, tw, j, vi);
}

// One Stockham stage reading the previous stage's values from `in` and writing this stage's to
// `out` (two LDS buffers used alternately): no barrier is needed between the reads and the writes
// of one stage, only after the writes.
template <int N, int S, int DIR, int V, int VAR = 0>
__device__ __forceinline__ void fft_stage_pp(float2 (&v)[(Fft<N, VAR>::RMAX)], const float2* in, float2* out,
                                             const float2* tw, int j, int vi) {
  using F = Fft<N, VAR>;
  constexpr int RMAX = F::RMAX;
  constexpr int NT = F::NT;
  constexpr int R = F::radix(S);
  constexpr int NS = F::ns_before(S);
  constexpr int Q = RMAX / R;
  constexpr bool FIRST = (S == 0);
  constexpr bool LAST = (S == F::NST - 1);
  constexpr int TWO = F::tw_offset(S);
  if constexpr (!FIRST) {
#pragma unroll
    for (int m = 0; m < RMAX; ++m) v[m] = in[lds_slot<V, F::PD>(j + m * NT, vi)];
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int jb = j + q * NT;
    float2 x[R];
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = v[q + r * Q];
    if constexpr (NS > 1) {
      const int k = jb % NS;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 w = tw[TWO + r * NS + k];
        x[r] = (DIR < 0) ? cmul(x[r], w) : cmulc(x[r], w);
      }
    }
    dft<R, DIR>(x);
    if constexpr (LAST) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[q + r * Q] = x[r];
    } else {
      const int idxd = (jb / NS) * NS * R + (jb % NS);
#pragma unroll
      for (int r = 0; r < R; ++r) out[lds_slot<V, F::PD>(idxd + r * NS, vi)] = x[r];
    }
  }
  if constexpr (!LAST) __syncthreads();
}

// fft_wg over two LDS buffers: stage s writes buffer s % 2 and stage s + 1 reads it, so each
// exchange costs one workgroup barrier instead of two.  The last stage reads buffer (NST - 2) % 2;
// with an odd number of stages that is lds1, and the next call's first write (lds0) is already
// ordered after every read of lds0 by this call's barriers; with an even number a trailing barrier
// orders it.  Same arithmetic as fft_wg, so the same results.
template <int N, int DIR, int V, int VAR = 0>
__device__ __forceinline__ void fft_wg_pp(float2 (&v)[(Fft<N, VAR>::RMAX)], float2* lds0, float2* lds1,
                                          const float2* tw, int j, int vi) {
  using F = Fft<N, VAR>;
  fft_stage_pp<N, 0, DIR, V, VAR>(v, lds1, lds0, tw, j, vi);
  if constexpr (F::NST > 1) fft_stage_pp<N, 1, DIR, V, VAR>(v, lds0, lds1, tw, j, vi);
  if constexpr (F::NST > 2) fft_stage_pp<N, 2, DIR, V, VAR>(v, lds1, lds0, tw, j, vi);
  if constexpr (F::NST > 3) fft_stage_pp<N, 3, DIR, V, VAR>(v, lds0, lds1, tw, j, vi);
  if constexpr (F::NST > 4) fft_stage_pp<N, 4, DIR, V, VAR>(v, lds1, lds0, tw, j, vi);
  if constexpr (F::NST % 2 == 0) __syncthreads();
}

// Supported lengths.  Powers of two, and the mixed-radix lengths above (each also needs its
// kernels instantiated, which the size lists below drive).
#define TFSM_MIXED_SIZES_1D(X) X(48) X(96) X(100) X(112) X(192) X(200) X(384) X(400) X(768) X(800)
#define TFSM_MIXED_SIZES_2D(X) X(48) X(96) X(100) X(112) X(192) X(200)
#define TFSM_MIXED_SIZES_3D(X) X(48) X(96) X(100)
constexpr bool is_mixed_size(int n) {
#define TFSM_MS_EQ(N) n == N ||
  return TFSM_MIXED_SIZES_1D(TFSM_MS_EQ) false;
#undef TFSM_MS_EQ
}
constexpr bool is_mixed_size_2d(int n) {
#define TFSM_MS_EQ(N) n == N ||
  return TFSM_MIXED_SIZES_2D(TFSM_MS_EQ) false;
#undef TFSM_MS_EQ
}
constexpr bool is_mixed_size_3d(int n) {
#define TFSM_MS_EQ(N) n == N ||
  return TFSM_MIXED_SIZES_3D(TFSM_MS_EQ) false;
#undef TFSM_MS_EQ
}

// row transforms (V = 1): wave-local synchronisation whenever a row fits one wavefront
template <int N, int VAR>
struct RowSync {
  static constexpr bool WS = Fft<N, VAR>::NT <= 64 && 64 % Fft<N, VAR>::NT == 0;
};

// Cooperative copy of the twiddle table into LDS (call before the first fft_wg, then barrier).
template <int N, int VAR = 0>
__device__ __forceinline__ void load_twiddles(float2* tw_lds, const float2* tw_glb) {
  for (int i = threadIdx.x; i < Fft<N, VAR>::tw_size(); i += blockDim.x) tw_lds[i] = tw_glb[i];
}

}  // namespace tfsm
