// General-length batched FFTs (declarations and conventions in gfft.hpp).
//
// One workgroup of 256 threads transforms 256 / T lines, T threads per line (T = the power of two
// nearest above L / 4, within 16..256).  A line is gathered from its strided layout into LDS (real
// inputs promoted, Hermitian halves mirrored, Bluestein chirp applied), transformed by Stockham
// autosort stages with runtime radices (a ping-pong pair of L-point buffers), and scattered back
// with the output's own index map (first n/2 + 1 values for R2C, real parts for C2R).  Four-step
// and Bluestein transforms chain these launches through a scratch buffer.
#include <cmath>
#include <vector>

#include "gfft.hpp"

namespace tfsm {
namespace {

// One launch: line index = (o * n_i + i) * n_s + s.  Loaded element j of a line is logical input
// index q = j * qj + s * qs; stored element k is logical output index r = k * rk + s * rs.
struct GLaunch {
  GRadix g;
  int dir;               // sign of the exponent of the (first) FFT
  const float2* bmid;    // one-pass Bluestein: multiply by bmid, then the +1 FFT of the same length
  long long n_o, n_i, n_s;
  const void* in;
  long long ios, iis, ies, qj, qs;
  int in_mode;           // 0 complex, 1 real, 2 Hermitian half (hbins values of an nfull-point line)
  long long nload, nfull, hbins;
  const float2* pre;     // input multiplier per logical q (Bluestein chirp)
  void* out;
  long long oos, ois, oes, rk, rs;
  int out_mode;          // 0 complex, 1 real part
  long long nstore;      // logical outputs r < nstore are stored
  const float2* post;    // output multiplier per logical r (Bluestein chirp or kernel spectrum)
  long long twn;         // four-step twiddle exp(twdir 2 pi i k s / twn) (0: none)
  int twdir;
  const float2* twt;     // its table exp(-2 pi i t / twn), t < twn (null: computed in float64)
  float scale;
  int T;
  // the workgroup's lines are consecutive s (n_s > 1) or i (n_s == 1) of one (o, i) / o, and
  // adjacent lines are adjacent in memory on input (ilv_in) / output (ilv_out): the threads then
  // walk those lines interleaved, so neighbouring threads touch neighbouring addresses
  int ilv_in, ilv_out;
};

__device__ __forceinline__ float2 gmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ float2 gtw(const float2* tw, int t, int dir) {
  const float2 w = tw[t];
  return dir < 0 ? w : make_float2(w.x, -w.y);
}

// in-register DFT of R values; w_R = exp(dir 2 pi i / R) read from the L-point table
template <int R>
__device__ __forceinline__ void gbfly(float2 (&v)[7], const float2* tw, int L, int dir) {
  if constexpr (R == 2) {
    const float2 a = v[0], b = v[1];
    v[0] = make_float2(a.x + b.x, a.y + b.y);
    v[1] = make_float2(a.x - b.x, a.y - b.y);
  } else if constexpr (R == 4) {
    const float2 t0 = make_float2(v[0].x + v[2].x, v[0].y + v[2].y);
    const float2 t1 = make_float2(v[0].x - v[2].x, v[0].y - v[2].y);
    const float2 t2 = make_float2(v[1].x + v[3].x, v[1].y + v[3].y);
    const float2 d = make_float2(v[1].x - v[3].x, v[1].y - v[3].y);
    const float2 t3 = dir < 0 ? make_float2(d.y, -d.x) : make_float2(-d.y, d.x);  // d * (-i) or d * i
    v[0] = make_float2(t0.x + t2.x, t0.y + t2.y);
    v[2] = make_float2(t0.x - t2.x, t0.y - t2.y);
    v[1] = make_float2(t1.x + t3.x, t1.y + t3.y);
    v[3] = make_float2(t1.x - t3.x, t1.y - t3.y);
  } else {
    float2 w[R];
#pragma unroll
    for (int m = 0; m < R; ++m) w[m] = gtw(tw, m * (L / R), dir);
    float2 y[R];
#pragma unroll
    for (int m = 0; m < R; ++m) {
      float2 acc = v[0];
#pragma unroll
      for (int q = 1; q < R; ++q) {
        const float2 p = gmul(v[q], w[(q * m) % R]);
        acc.x += p.x;
        acc.y += p.y;
      }
      y[m] = acc;
    }
#pragma unroll
    for (int m = 0; m < R; ++m) v[m] = y[m];
  }
}

// one Stockham stage: Ns points already combined; butterfly b reads b + q L/R, twiddles by
// w^(q (b mod Ns)) with w = exp(dir 2 pi i / (Ns R)), writes (b / Ns) Ns R + (b mod Ns) + q Ns
template <int R>
__device__ void gstage(const float2* src, float2* dst, const float2* tw, int L, int Ns, int nsl, int dir, int t,
                       int T) {
  const int nb = L / R, span = L / (Ns * R);
  for (int b = t; b < nb; b += T) {
    // nsl = log2(Ns) for power-of-two plans (no integer division), else -1
    const int k = nsl >= 0 ? (b & (Ns - 1)) : b % Ns;
    const int grp = nsl >= 0 ? (b >> nsl) : b / Ns;
    float2 v[7];
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = src[b + q * nb];
    if (Ns > 1) {
#pragma unroll
      for (int q = 1; q < R; ++q) v[q] = gmul(v[q], gtw(tw, q * k * span, dir));
    }
    gbfly<R>(v, tw, L, dir);
    const int d0 = grp * Ns * R + k;
#pragma unroll
    for (int q = 0; q < R; ++q) dst[d0 + q * Ns] = v[q];
  }
}

// every stage of g; the result ends in *src (every thread of the workgroup takes part)
__device__ void gfft_lds(const GRadix& g, int dir, float2*& src, float2*& dst, int t, int T) {
  int Ns = 1, nsl = g.pow2 ? 0 : -1;
  for (int s = 0; s < g.nr; ++s) {
    const int R = g.r[s];
    switch (R) {
      case 4: gstage<4>(src, dst, g.tw, g.L, Ns, nsl, dir, t, T); break;
      case 2: gstage<2>(src, dst, g.tw, g.L, Ns, nsl, dir, t, T); break;
      case 3: gstage<3>(src, dst, g.tw, g.L, Ns, nsl, dir, t, T); break;
      case 5: gstage<5>(src, dst, g.tw, g.L, Ns, nsl, dir, t, T); break;
      default: gstage<7>(src, dst, g.tw, g.L, Ns, nsl, dir, t, T); break;
    }
    __syncthreads();
    float2* x = src;
    src = dst;
    dst = x;
    Ns *= R;
    if (nsl >= 0) nsl += R == 4 ? 2 : 1;
  }
}

__global__ void __launch_bounds__(GFFT_THREADS) k_gfft(GLaunch a) {
  extern __shared__ float2 g_lds[];
  const int T = a.T, L = a.g.L;
  const int lpw = GFFT_THREADS / T;
  // T and lpw are powers of two
  const int tsh = __ffs(T) - 1, lsh = __ffs(lpw) - 1;
  const int sub = threadIdx.x >> tsh, t = threadIdx.x & (T - 1);
  float2* src = g_lds + (size_t)sub * 2 * L;
  float2* dst = src + L;
  const long long nlines = a.n_o * a.n_i * a.n_s;
  const long long line = (long long)blockIdx.x * lpw + sub;
  const bool valid = line < nlines;
  long long s = 0, i = 0, o = 0;
  if (valid) {
    s = line % a.n_s;
    const long long r = line / a.n_s;
    i = r % a.n_i;
    o = r / a.n_i;
  }
  const long long L0 = (long long)blockIdx.x * lpw;  // first line of the workgroup
  long long s0 = 0, i0 = 0, o0 = 0;
  if (a.ilv_in || a.ilv_out) {
    s0 = L0 % a.n_s;
    const long long r0 = L0 / a.n_s;
    i0 = r0 % a.n_i;
    o0 = r0 / a.n_i;
  }
  if (a.ilv_in) {
    for (int e = threadIdx.x; e < lpw * L; e += GFFT_THREADS) {
      const int ls = e & (lpw - 1), j = e >> lsh;
      float2 v = make_float2(0.f, 0.f);
      const long long sl = a.n_s > 1 ? s0 + ls : 0, il = a.n_s > 1 ? i0 : i0 + ls;
      const long long q = (long long)j * a.qj + sl * a.qs;
      if (L0 + ls < nlines && q < a.nload) {
        const long long ibl = o0 * a.ios + il * a.iis;
        if (a.in_mode == 1) {
          v.x = ((const float*)a.in)[ibl + q * a.ies];
        } else if (a.in_mode == 2) {
          const float2* x = (const float2*)a.in;
          if (q < a.hbins) {
            v = x[ibl + q * a.ies];
          } else {
            const float2 m = x[ibl + (a.nfull - q) * a.ies];
            v = make_float2(m.x, -m.y);
          }
          if (q == 0 || 2 * q == a.nfull) v.y = 0.f;
        } else {
          v = ((const float2*)a.in)[ibl + q * a.ies];
        }
        if (a.pre) v = gmul(v, a.pre[q]);
      }
      g_lds[(size_t)ls * 2 * L + j] = v;
    }
  }
  const long long ib = o * a.ios + i * a.iis;
  for (int j = t; j < L && !a.ilv_in; j += T) {
    float2 v = make_float2(0.f, 0.f);
    const long long q = (long long)j * a.qj + s * a.qs;
    if (valid && q < a.nload) {
      if (a.in_mode == 1) {
        v.x = ((const float*)a.in)[ib + q * a.ies];
      } else if (a.in_mode == 2) {
        const float2* x = (const float2*)a.in;
        if (q < a.hbins) {
          v = x[ib + q * a.ies];
        } else {
          const float2 m = x[ib + (a.nfull - q) * a.ies];
          v = make_float2(m.x, -m.y);
        }
        if (q == 0 || 2 * q == a.nfull) v.y = 0.f;
      } else {
        v = ((const float2*)a.in)[ib + q * a.ies];
      }
      if (a.pre) v = gmul(v, a.pre[q]);
    }
    src[j] = v;
  }
  __syncthreads();
  gfft_lds(a.g, a.dir, src, dst, t, T);
  if (a.bmid) {
    for (int k = t; k < L; k += T) src[k] = gmul(src[k], a.bmid[k]);
    __syncthreads();
    gfft_lds(a.g, +1, src, dst, t, T);
  }
  if (a.ilv_out) {
    // every line ends in the same buffer of its pair: this thread's offset holds for all of them
    const int fo = (int)(src - (g_lds + (size_t)sub * 2 * L));
    for (int e = threadIdx.x; e < lpw * L; e += GFFT_THREADS) {
      const int ls = e & (lpw - 1), k = e >> lsh;
      if (L0 + ls >= nlines) continue;
      const long long sl = a.n_s > 1 ? s0 + ls : 0, il = a.n_s > 1 ? i0 : i0 + ls;
      const long long r = (long long)k * a.rk + sl * a.rs;
      if (r >= a.nstore) continue;
      float2 v = g_lds[(size_t)ls * 2 * L + fo + k];
      if (a.twn) {
        long long e2 = (long long)k * sl;  // k < L1, s < L2: below twn = L1 L2
        if (e2 >= a.twn) e2 %= a.twn;
        if (a.twt) {
          v = gmul(v, gtw(a.twt, (int)e2, a.twdir));
        } else {
          double sn, cs;
          sincos((double)a.twdir * 2.0 * M_PI * (double)e2 / (double)a.twn, &sn, &cs);
          v = gmul(v, make_float2((float)cs, (float)sn));
        }
      }
      if (a.post) v = gmul(v, a.post[r]);
      v.x *= a.scale;
      v.y *= a.scale;
      const long long obl = o0 * a.oos + il * a.ois;
      if (a.out_mode == 1)
        ((float*)a.out)[obl + r * a.oes] = v.x;
      else
        ((float2*)a.out)[obl + r * a.oes] = v;
    }
    return;
  }
  if (!valid) return;
  const long long ob = o * a.oos + i * a.ois;
  for (int k = t; k < L; k += T) {
    const long long r = (long long)k * a.rk + s * a.rs;
    if (r >= a.nstore) continue;
    float2 v = src[k];
    if (a.twn) {
      long long e = (long long)k * s;  // k < L1, s < L2: below twn = L1 L2
      if (e >= a.twn) e %= a.twn;
      if (a.twt) {
        v = gmul(v, gtw(a.twt, (int)e, a.twdir));
      } else {
        double sn, cs;
        sincos((double)a.twdir * 2.0 * M_PI * (double)e / (double)a.twn, &sn, &cs);
        v = gmul(v, make_float2((float)cs, (float)sn));
      }
    }
    if (a.post) v = gmul(v, a.post[r]);
    v.x *= a.scale;
    v.y *= a.scale;
    if (a.out_mode == 1)
      ((float*)a.out)[ob + r * a.oes] = v.x;
    else
      ((float2*)a.out)[ob + r * a.oes] = v;
  }
}

int threads_per_line(int L) {
  int t = 16;
  while (t < GFFT_THREADS && t * 4 < L) t *= 2;
  return t;
}

hipError_t launch_g(GLaunch a, hipStream_t st) {
  a.T = threads_per_line(a.g.L);
  // lines adjacent in memory (strided transforms: columns, four-step passes): up to 16 lines per
  // workgroup, walked interleaved, so that their loads and stores coalesce
  const long long run = a.n_s > 1 ? a.n_s : a.n_i;
  const long long din = a.n_s > 1 ? a.qs * a.ies : a.iis, dout = a.n_s > 1 ? a.rs * a.oes : a.ois;
  a.ilv_in = a.ilv_out = 0;
  if (din == 1 || dout == 1) {
    int l = 1;
    while (l < 16 && 2 * l * a.g.L <= GFFT_LDS_MAX && run % (2 * l) == 0) l *= 2;
    if (l > 1) {
      a.T = GFFT_THREADS / l;
      a.ilv_in = din == 1;
      a.ilv_out = dout == 1;
    }
  }
  const long long lpw = GFFT_THREADS / a.T;
  const long long nlines = a.n_o * a.n_i * a.n_s;
  if (nlines <= 0) return hipSuccess;
  const long long blocks = (nlines + lpw - 1) / lpw;
  if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  const size_t lds = (size_t)lpw * 2 * a.g.L * sizeof(float2);
  hipLaunchKernelGGL(k_gfft, dim3((unsigned)blocks), dim3(GFFT_THREADS), lds, st, a);
  return hipGetLastError();
}

bool radices(int L, GRadix& g) {
  g.L = L;
  g.nr = 0;
  g.pow2 = (L & (L - 1)) == 0 ? 1 : 0;
  int n = L;
  while (n % 4 == 0 && g.nr < GFFT_MAX_RADICES) {
    g.r[g.nr++] = 4;
    n /= 4;
  }
  const int rs[4] = {2, 3, 5, 7};
  for (int r : rs)
    while (n % r == 0 && g.nr < GFFT_MAX_RADICES) {
      g.r[g.nr++] = r;
      n /= r;
    }
  return n == 1;
}

// largest L1 <= GFFT_LDS_MAX dividing L with L / L1 <= GFFT_LDS_MAX (0: none)
int split_four(int L) {
  for (int l1 = GFFT_LDS_MAX; l1 >= 2; --l1)
    if (L % l1 == 0 && L / l1 <= GFFT_LDS_MAX) return l1;
  return 0;
}

hipError_t upload(const std::vector<float2>& h, const float2** out, std::vector<void*>& allocs) {
  void* d = nullptr;
  hipError_t e = hipMalloc(&d, h.size() * sizeof(float2));
  if (e != hipSuccess) return e;
  allocs.push_back(d);
  e = hipMemcpy(d, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice);
  *out = (const float2*)d;
  return e;
}

hipError_t build_radix(GRadix& g, int L, std::vector<void*>& allocs) {
  if (!radices(L, g)) return hipErrorInvalidValue;
  std::vector<float2> h(L);
  for (int t = 0; t < L; ++t) {
    const double ang = -2.0 * M_PI * (double)t / (double)L;
    h[t] = make_float2((float)std::cos(ang), (float)std::sin(ang));
  }
  return upload(h, &g.tw, allocs);
}

// in-place radix-2 FFT in double (host), exponent sign -1
void host_fft_pow2(std::vector<double>& re, std::vector<double>& im) {
  const size_t n = re.size();
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      std::swap(re[i], re[j]);
      std::swap(im[i], im[j]);
    }
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    const double ang = -2.0 * M_PI / (double)len;
    for (size_t i = 0; i < n; i += len)
      for (size_t k = 0; k < len / 2; ++k) {
        const double wr = std::cos(ang * (double)k), wi = std::sin(ang * (double)k);
        const size_t a = i + k, b = i + k + len / 2;
        const double xr = re[b] * wr - im[b] * wi, xi = re[b] * wi + im[b] * wr;
        re[b] = re[a] - xr;
        im[b] = im[a] - xi;
        re[a] += xr;
        im[a] += xi;
      }
  }
}

}  // namespace

bool gfft_smooth(int n) {
  if (n < 1) return false;
  for (int r : {2, 3, 5, 7})
    while (n % r == 0) n /= r;
  return n == 1;
}

hipError_t gaxis_build(GAxis& a, int n, std::vector<void*>& allocs) {
  if (n < 1) return hipErrorInvalidValue;
  a = GAxis();
  a.n = n;
  if (gfft_smooth(n) && (n <= GFFT_LDS_MAX || split_four(n))) {
    a.L = n;
  } else {
    long long m = 1;
    while (m < 2LL * n - 1) m <<= 1;
    if (m > (long long)GFFT_LDS_MAX * GFFT_LDS_MAX) return hipErrorInvalidValue;
    a.blu = true;
    a.L = (int)m;
  }
  hipError_t e;
  if (a.L <= GFFT_LDS_MAX) {
    e = build_radix(a.g, a.L, allocs);
  } else {
    a.four = true;
    const int l1 = split_four(a.L);
    if (!l1) return hipErrorInvalidValue;
    e = build_radix(a.g1, l1, allocs);
    if (e == hipSuccess) e = build_radix(a.g2, a.L / l1, allocs);
    if (e == hipSuccess) {
      // the L-point twiddle table of the four-step (float32 of float64 values, 8 L bytes)
      std::vector<float2> h(a.L);
      for (int t = 0; t < a.L; ++t) {
        const double ang = -2.0 * M_PI * (double)t / (double)a.L;
        h[t] = make_float2((float)std::cos(ang), (float)std::sin(ang));
      }
      e = upload(h, &a.tw4, allocs);
    }
  }
  if (e != hipSuccess || !a.blu) return e;
  // chirp c_d[t] = exp(-/+ i pi (t^2 mod 2n) / n); kernel b[m] = conj(c_d[|m|]) for |m| < n on the
  // circle of M points; bhat = FFT_M(b) / M
  const int M = a.L;
  for (int d = 0; d < 2; ++d) {
    const double sg = d == 0 ? -1.0 : 1.0;
    std::vector<float2> c(n);
    std::vector<double> br(M, 0.0), bi(M, 0.0);
    for (int t = 0; t < n; ++t) {
      const long long q = ((long long)t * t) % (2LL * n);
      const double ang = sg * M_PI * (double)q / (double)n;
      const double cr = std::cos(ang), ci = std::sin(ang);
      c[t] = make_float2((float)cr, (float)ci);
      br[t] = cr;
      bi[t] = -ci;
      if (t > 0) {
        br[M - t] = cr;
        bi[M - t] = -ci;
      }
    }
    host_fft_pow2(br, bi);
    std::vector<float2> bh(M);
    for (int k = 0; k < M; ++k) bh[k] = make_float2((float)(br[k] / M), (float)(bi[k] / M));
    e = upload(c, &a.chirp[d], allocs);
    if (e == hipSuccess) e = upload(bh, &a.bhat[d], allocs);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

size_t gfft_tmp_elems(const GAxis& a, size_t lines) { return a.four ? 2 * lines * (size_t)a.L : 0; }

hipError_t gfft_run(const GAxis& a, int mode, int dir, const void* in, const GLines& li, void* out, const GLines& lo,
                    float scale, float2* tmp, hipStream_t st) {
  if (a.n < 1 || (dir != -1 && dir != 1)) return hipErrorInvalidValue;
  if (li.n_o != lo.n_o || li.n_i != lo.n_i) return hipErrorInvalidValue;
  const int d = dir < 0 ? 0 : 1;
  const long long n = a.n, H = n / 2 + 1;
  GLaunch base{};
  base.n_o = (long long)li.n_o;
  base.n_i = (long long)li.n_i;
  base.n_s = 1;
  base.in = in;
  base.ios = (long long)li.os;
  base.iis = (long long)li.is;
  base.ies = (long long)li.es;
  base.qj = 1;
  base.qs = 0;
  base.in_mode = mode == GM_R2C ? 1 : (mode == GM_C2R ? 2 : 0);
  base.nload = n;
  base.nfull = n;
  base.hbins = H;
  base.pre = a.blu ? a.chirp[d] : nullptr;
  base.out = out;
  base.oos = (long long)lo.os;
  base.ois = (long long)lo.is;
  base.oes = (long long)lo.es;
  base.rk = 1;
  base.rs = 0;
  base.out_mode = mode == GM_C2R ? 1 : 0;
  base.nstore = mode == GM_R2C ? H : n;
  base.post = a.blu ? a.chirp[d] : nullptr;
  base.scale = scale;
  if (!a.four) {
    GLaunch x = base;
    x.g = a.g;
    x.dir = a.blu ? -1 : dir;
    x.bmid = a.blu ? a.bhat[d] : nullptr;
    return launch_g(x, st);
  }
  if (!tmp) return hipErrorInvalidValue;
  const long long L = a.L, L1 = a.g1.L, L2 = a.g2.L;
  const long long lines = base.n_o * base.n_i;
  float2* t1 = tmp;
  float2* t2 = tmp + (size_t)lines * L;
  // pass A: L1-point FFTs over j1 of x[j1 L2 + j2], twiddled by w_L^(k1 j2), into t1[line][k1 L2 + j2]
  GLaunch A = base;
  A.g = a.g1;
  A.dir = a.blu ? -1 : dir;
  A.n_s = L2;
  A.qj = L2;
  A.qs = 1;
  A.out = t1;
  A.oos = base.n_i * L;
  A.ois = L;
  A.oes = 1;
  A.rk = L2;
  A.rs = 1;
  A.out_mode = 0;
  A.nstore = L;
  A.post = nullptr;
  A.twn = L;
  A.twdir = A.dir;
  A.twt = a.tw4;
  A.scale = 1.0f;
  // pass B: L2-point FFTs over j2 of t1[k1 L2 + j2]; result X[k1 + L1 k2]
  GLaunch B = base;
  B.g = a.g2;
  B.dir = A.dir;
  B.n_s = L1;
  B.in = t1;
  B.ios = base.n_i * L;
  B.iis = L;
  B.ies = 1;
  B.qj = 1;
  B.qs = L2;
  B.in_mode = 0;
  B.nload = L;
  B.pre = nullptr;
  B.rk = L1;
  B.rs = 1;
  hipError_t e = launch_g(A, st);
  if (e != hipSuccess) return e;
  if (!a.blu) return launch_g(B, st);
  // Bluestein: X * bhat into t2 (natural order), then the inverse four-step into the output
  GLaunch B1 = B;
  B1.out = t2;
  B1.oos = base.n_i * L;
  B1.ois = L;
  B1.oes = 1;
  B1.out_mode = 0;
  B1.nstore = L;
  B1.post = a.bhat[d];
  B1.scale = 1.0f;
  e = launch_g(B1, st);
  if (e != hipSuccess) return e;
  GLaunch C = A;
  C.dir = +1;
  C.twdir = +1;
  C.in = t2;
  C.ios = base.n_i * L;
  C.iis = L;
  C.ies = 1;
  C.in_mode = 0;
  C.nload = L;
  C.pre = nullptr;
  e = launch_g(C, st);
  if (e != hipSuccess) return e;
  GLaunch D = B;
  D.dir = +1;
  return launch_g(D, st);
}

}  // namespace tfsm
