// General-length batched FFTs for the axes the fixed-size kernels do not cover (mesh.py:481-491
// transforms any length).  A length n that factors into 2, 3, 5 and 7 runs a Stockham FFT with
// runtime radices (4, 2, 3, 5, 7) in one workgroup's LDS; any other n runs Bluestein's algorithm,
// a circular convolution of length M = 2^k >= 2n - 1.  Transforms longer than GFFT_LDS_MAX points
// split L = L1 * L2 into two LDS passes (four-step: L1-point FFTs, twiddles, L2-point FFTs).
//
// A batch of lines is described by GLines: line (o, i) starts at element o * os + i * is and its
// t-th element sits at + t * es.  Modes: C2C (complex n -> complex n), R2C (real n -> the n/2 + 1
// non-negative frequencies), C2R (n/2 + 1 Hermitian-half values -> real n; the imaginary parts of
// the zero and Nyquist frequencies are ignored).  dir -1 is exp(-2 pi i jk/n), +1 exp(+2 pi i jk/n)
// without the 1/n; every output value is multiplied by `scale`.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

namespace tfsm {

constexpr int GFFT_LDS_MAX = 4096;  // longest FFT held in one workgroup's LDS (2 x 32 KB ping-pong)
constexpr int GFFT_MAX_RADICES = 16;
constexpr int GFFT_THREADS = 256;

enum GfftMode : int { GM_C2C = 0, GM_R2C = 1, GM_C2R = 2 };

// radix sequence and twiddles exp(-2 pi i t / L), t < L, of one in-LDS FFT length
struct GRadix {
  int L = 0;
  int nr = 0;
  int r[GFFT_MAX_RADICES] = {};
  int pow2 = 0;  // every radix is 2 or 4: stage indices by shifts and masks
  const float2* tw = nullptr;
};

// tables of one transform length (built on the current device by gaxis_build)
struct GAxis {
  int n = 0;
  bool blu = false;   // Bluestein
  int L = 0;          // n, or M for Bluestein
  bool four = false;  // L > GFFT_LDS_MAX: L = g1.L * g2.L
  GRadix g, g1, g2;
  const float2* tw4 = nullptr;  // four-step: exp(-2 pi i t / L), t < L (twiddles between the passes)
  const float2* chirp[2] = {nullptr, nullptr};  // [n] exp(-/+ i pi (t^2 mod 2n) / n): forward, inverse
  const float2* bhat[2] = {nullptr, nullptr};   // [M] FFT_M of the conjugate chirp kernel, times 1/M
};

struct GLines {
  size_t n_o = 1, n_i = 1;
  size_t os = 0, is = 0, es = 1;
};

// true when n factors into 2, 3, 5 and 7
bool gfft_smooth(int n);
// build the tables of length n (n >= 1); device allocations are appended to `allocs`
hipError_t gaxis_build(GAxis& a, int n, std::vector<void*>& allocs);
// complex scratch elements gfft_run needs for `lines` lines of this axis (0 unless four-step)
size_t gfft_tmp_elems(const GAxis& a, size_t lines);
// transform every line; in == out is allowed when both describe the same lines (C2C);
// tmp: gfft_tmp_elems(a, n_o * n_i) complex elements, or nullptr when that is 0
hipError_t gfft_run(const GAxis& a, int mode, int dir, const void* in, const GLines& li, void* out, const GLines& lo,
                    float scale, float2* tmp, hipStream_t st);

}  // namespace tfsm
