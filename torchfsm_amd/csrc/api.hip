// C ABI of the ETD engine: plans (geometry, symbols, coefficient tables), transforms, stepping,
// graph capture.  Kernels live in k_*.hip; this file holds host logic only.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "engine.hpp"
#include "gfft.hpp"
#include "tfsm.h"

using namespace tfsm;

namespace tfsm {
thread_local hipEvent_t g_launch_ev[2] = {nullptr, nullptr};
}

// Workspace of tfsm_step_noalloc: a device array of fixed size, one copy per GPU, loaded with the
// code object and held until the process ends.  Host code reaches it only through the address
// hipGetSymbolAddress returns for the current device (scratch_on_device below); the name itself is
// a host-side stand-in and never a device address.
static __device__ __align__(256) unsigned char scratch[TFSM_NOALLOC_SCRATCH_BYTES];

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    return fail(TFSM_E_OOM, std::string(what) + ": out of device memory (" + hipGetErrorString(e) + ")");
  return fail(TFSM_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// status of a HIP call as a TFSM code (0 on success), for code paths with a single exit
int launch_status(hipError_t e, const char* what) { return e == hipSuccess ? TFSM_OK : hip_fail(e, what); }

#define HIP_TRY(expr, what)               \
  do {                                    \
    hipError_t _e = (expr);               \
    if (_e != hipSuccess) return hip_fail(_e, what); \
  } while (0)

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
size_t roundup(size_t x, size_t m) { return (x + m - 1) / m * m; }

// fftfreq order of integer wavenumbers
int signed_k(int i, int n) { return i < (n + 1) / 2 ? i : i - n; }

// axis wavenumbers exactly as the reference forms them in float32:
// f = fl32(k) * fl32(1 / (n d)), d = length / n  (torch.fft.fftfreq, mesh.py:185-191)
// a = fl32(fl32(2 pi) * f)                        ((2j*pi*bf)^1, mesh.py:399-404)
void axis_table(int n, double length, std::vector<float>& a, std::vector<float>* f_out = nullptr) {
  const double d = length / n;
  const float s = (float)(1.0 / (n * d));
  const float two_pi = (float)(2.0 * M_PI);
  a.resize(n);
  if (f_out) f_out->resize(n);
  for (int i = 0; i < n; ++i) {
    const float f = (float)signed_k(i, n) * s;
    a[i] = two_pi * f;
    if (f_out) (*f_out)[i] = f;
  }
}

// low_pass_filter (mesh.py:443-461): keep |f| <= max|f| * rate, in float32; returns max kept |k|
int dealias_kmax(int n, double length, double rate) {
  std::vector<float> a, f;
  axis_table(n, length, a, &f);
  float fmax = 0.f;
  for (float v : f) fmax = std::fabs(v) > fmax ? std::fabs(v) : fmax;
  const float thr = fmax * (float)rate;
  int kmax = -1;
  for (int i = 0; i < n; ++i) {
    const int k = std::abs(signed_k(i, n));
    const bool masked = std::fabs(f[i]) > thr;
    if (!masked && k > kmax) kmax = k;
  }
  // the kept set must be {|k| <= kmax} for the engine's index-range mask
  for (int i = 0; i < n; ++i) {
    const int k = std::abs(signed_k(i, n));
    const bool masked = std::fabs(f[i]) > thr;
    if ((k <= kmax) == masked) return -2;
  }
  return kmax;
}

}  // namespace

struct tfsm_plan {
  int device = 0;  // HIP device current at tfsm_plan_create (its tables live there)
  int D = 0, n[3] = {1, 1, 1}, B = 0, C = 0;
  double len[3] = {0, 0, 0};
  int H = 0, P = 0;
  size_t modes = 0;  // complex elements of one [N...] block in half-spectrum layout
  int term = TERM_NONE;
  bool ks_mean = false;  // KSConvection(remove_mean=True): the batch mean of the product is removed
  float coef_t = 0.f;
  double rate = 2.0 / 3.0;
  int kmax[3] = {0, 0, 0};
  int nlin = 0;
  int lin_kind[4] = {0, 0, 0, 0};
  int lin_dim[4] = {0, 0, 0, 0}, lin_order[4] = {0, 0, 0, 0};
  bool cplx = false;        // complex L (odd-order SpatialDerivative): complex symbol and tables
  float2* Lc = nullptr;     // [Bc][modes]
  float2* coefc = nullptr;  // [Bc][NCOEF][modes]
  bool per_len = false;  // 1D plan with per-sample lengths (axis tables per row)
  int Bc = 1;
  float2* tw[3][4] = {};  // [axis][radix-plan variant]
  float* kv[3] = {nullptr, nullptr, nullptr};
  size_t kv_bstride = 0;
  float* L = nullptr;
  float* coef = nullptr;
  float4* cpk = nullptr;  // packed per-mode coefficients of the current scheme
  float* c5t = nullptr;   // SETDRK4's c5 per mode (read alone by the third evaluation)
  float2* ck01 = nullptr; // SETDRK4's [E2, c1] per mode (first and second evaluations)
  float* c6t = nullptr;   // SETDRK4's c6 per mode (last evaluation)
  int scheme = -1;
  double dt = 0.0;
  int n_points = 16;
  double radius = 1.0;
  float2* src = nullptr;
  int srcB = 0, srcC = 0;
  unsigned char* src_cols = nullptr;  // 2D: [P] 1 where a source mode of last-axis column k is nonzero
  uint32_t src_bits[Col2Args::SRC_BIT_WORDS] = {};  // the same flags as bits (P <= 544)
  bool src_bits_ok = false;
  int V = 16;        // column tile of the fused 2D column kernel
  int col_var = 0;   // radix-plan variant of the column transforms
  int row_var = 0;   // radix-plan variant of the row transforms
  int row_wg = 256;  // row-kernel workgroup size
  int sub_batch = 0; // samples stepped together through all steps (0 = whole batch)
  int nstreams = 1;  // > 1: sub-batch k runs on stream k % nstreams (0: the caller's, i: aux[i - 1])
  bool persistent = false;  // 2D: step with the persistent per-sample kernel k2_run
  hipStream_t aux[TFSM_MAX_STREAMS - 1] = {};
  hipEvent_t ev_fork = nullptr, ev_join[TFSM_MAX_STREAMS - 1] = {};
  int PF = 0, NF = 0;
  int K1 = 0;  // 3D: kept ky values (2 kmax1 + 1)
  size_t off_A1 = 0, off_A2 = 0, off_P = 0, off_G = 0, off_G2 = 0, ws_bytes = 0;
  size_t off_DC = 0;              // KSConvection: batch mean of the product DC (one complex value)
  size_t off_S = 0, off_N = 0;    // 1D KSConvection: stage spectrum and raw product spectrum
  std::vector<void*> allocs;
  // adjoint (tfsm_adjoint_prepare)
  float* dcoef = nullptr;  // [Bc][3][modes]
  float* sym = nullptr;    // [modes]
  int adj_term = -2;       // linear term whose coefficient gets a gradient (-1 none; -2 not prepared)
  int adj_scheme = -1;
  // unfused path (k_general.hip): stepping without the fused kernels
  bool general = false;    // the plan steps through the unfused path
  bool fused_ok = false;   // the fused kernels cover this plan (the unfused path is then optional)
  GAxis gax[3];            // general-length FFT tables of the axes without fixed-size kernels
  float2* gtmp = nullptr;  // four-step scratch of those axes
  float2* force = nullptr;  // NSPressureConvection's external force spectrum [forceB][forceC][modes]
  int forceB = 0, forceC = 0;
  size_t ws_fused = 0;     // workspace of the fused path (the unfused buffers follow it)
  size_t off_gS = 0, off_gF = 0, off_gR = 0, off_gQ = 0;

  size_t state_elems() const { return (size_t)B * C * modes; }
  size_t src_bstride() const { return srcB > 1 ? (size_t)srcC * modes : 0; }
  size_t src_cstride() const { return srcC > 1 ? modes : 0; }
  size_t coef_bstride() const { return Bc > 1 ? (size_t)NCOEF * modes : 0; }
};

struct tfsm_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

template <class T>
int dev_alloc(tfsm_plan* p, T** ptr, size_t count, const char* what) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, count * sizeof(T) > 0 ? count * sizeof(T) : 16);
  if (e != hipSuccess) return hip_fail(e, what);
  p->allocs.push_back(q);
  *ptr = static_cast<T*>(q);
  return TFSM_OK;
}

// per-stage twiddle tables in the layout of Fft<N, VAR>::tw_offset: W^(r k), W = exp(-2 pi i/(Ns R))
template <int N, int VAR>
void twiddle_table_t(std::vector<float2>& h) {
  using F = Fft<N, VAR>;
  h.assign(F::tw_size(), make_float2(1.f, 0.f));
  for (int s = 1; s < F::NST; ++s) {
    const int R = F::radix(s), ns = F::ns_before(s), off = F::tw_offset(s);
    for (int r = 0; r < R; ++r)
      for (int k = 0; k < ns; ++k) {
        const double ang = -2.0 * M_PI * (double)r * (double)k / ((double)ns * R);
        h[off + r * ns + k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
      }
  }
}

bool twiddle_table(int n, int var, std::vector<float2>& h) {
  if (var == 0) {
    switch (n) {
      case 16: twiddle_table_t<16, 0>(h); return true;
      case 32: twiddle_table_t<32, 0>(h); return true;
      case 64: twiddle_table_t<64, 0>(h); return true;
      case 128: twiddle_table_t<128, 0>(h); return true;
      case 256: twiddle_table_t<256, 0>(h); return true;
      case 512: twiddle_table_t<512, 0>(h); return true;
      case 1024: twiddle_table_t<1024, 0>(h); return true;
      case 2048: twiddle_table_t<2048, 0>(h); return true;
      case 4096: twiddle_table_t<4096, 0>(h); return true;
      case 48: twiddle_table_t<48, 0>(h); return true;
      case 96: twiddle_table_t<96, 0>(h); return true;
      case 100: twiddle_table_t<100, 0>(h); return true;
      case 112: twiddle_table_t<112, 0>(h); return true;
      case 192: twiddle_table_t<192, 0>(h); return true;
      case 200: twiddle_table_t<200, 0>(h); return true;
      case 384: twiddle_table_t<384, 0>(h); return true;
      case 400: twiddle_table_t<400, 0>(h); return true;
      case 768: twiddle_table_t<768, 0>(h); return true;
      case 800: twiddle_table_t<800, 0>(h); return true;
      default: return false;
    }
  }
  if (var == 2) {
    if (n == 128) {
      twiddle_table_t<128, 2>(h);
      return true;
    }
    return false;
  }
  if (var == 3) {
    switch (n) {
      case 16: twiddle_table_t<16, 3>(h); return true;
      case 32: twiddle_table_t<32, 3>(h); return true;
      case 64: twiddle_table_t<64, 3>(h); return true;
      case 128: twiddle_table_t<128, 3>(h); return true;
      default: return false;
    }
  }
  switch (n) {
    case 64: twiddle_table_t<64, 1>(h); return true;
    case 128: twiddle_table_t<128, 1>(h); return true;
    case 256: twiddle_table_t<256, 1>(h); return true;
    case 512: twiddle_table_t<512, 1>(h); return true;
    case 1024: twiddle_table_t<1024, 1>(h); return true;
    default: return false;
  }
}

int build_twiddles(tfsm_plan* p, int d) {
  for (int var = 0; var < 4; ++var) {
    std::vector<float2> h;
    if (!twiddle_table(p->n[d], var, h)) {
      p->tw[d][var] = nullptr;
      continue;
    }
    int rc = dev_alloc(p, &p->tw[d][var], h.size(), "twiddle table");
    if (rc) return rc;
    HIP_TRY(hipMemcpy(p->tw[d][var], h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice), "twiddle upload");
  }
  return TFSM_OK;
}

// ---------------------------------------------------------------- transforms per axis
// Fixed-size kernels where the axis has them (rows: every length of the twiddle list; columns: up
// to 512 points), the general-length FFTs (gfft_run) elsewhere.
bool fast_rows(const tfsm_plan* p, int d) { return p->tw[d][0] != nullptr; }
bool fast_cols(const tfsm_plan* p, int d) { return p->tw[d][0] != nullptr && p->n[d] <= 512; }
bool axis_needs_gfft(const tfsm_plan* p, int d) { return d == p->D - 1 ? !fast_rows(p, d) : !fast_cols(p, d); }

// nrows rows of the last axis: GM_R2C (real in, ld in floats), GM_C2R, GM_C2C
hipError_t xf_rows(tfsm_plan* p, int mode, int dir, const void* in, size_t in_ld, void* out, size_t out_ld,
                   size_t nrows, float scale, hipStream_t st) {
  const int d = p->D - 1, n = p->n[d];
  if (nrows == 0) return hipSuccess;
  if (fast_rows(p, d) && nrows <= 0x7fffffffu) {
    if (mode == GM_R2C)
      return launch_rows_r2c(n, (const float*)in, in_ld, (float2*)out, out_ld, (int)nrows, scale, p->tw[d][0], st);
    if (mode == GM_C2R)
      return launch_rows_c2r(n, (const float2*)in, in_ld, (float*)out, out_ld, (int)nrows, scale, p->tw[d][0], st);
    return launch_rows_c2c(n, dir, (const float2*)in, in_ld, (float2*)out, out_ld, (int)nrows, scale, p->tw[d][0], st);
  }
  if (p->gax[d].n != n) return hipErrorInvalidValue;
  GLines li, lo;
  li.n_o = lo.n_o = nrows;
  li.os = in_ld;
  lo.os = out_ld;
  return gfft_run(p->gax[d], mode, dir, in, li, out, lo, scale, p->gtmp, st);
}

// complex transform along axis d < D - 1: nouter blocks of [n_d][...] with elements k * col_stride
// + i, i < ninner, of block o at o * outer_stride
hipError_t xf_cols(tfsm_plan* p, int d, int dir, const float2* in, float2* out, size_t nouter, size_t outer_stride,
                   size_t col_stride, size_t ninner, hipStream_t st) {
  if (nouter == 0 || ninner == 0) return hipSuccess;
  if (fast_cols(p, d) && ninner <= 0x7fffffffu)
    return launch_cols_c2c(p->n[d], dir, in, out, nouter, outer_stride, col_stride, (int)ninner, 1.0f, p->tw[d][0], st);
  if (p->gax[d].n != p->n[d]) return hipErrorInvalidValue;
  GLines l;
  l.n_o = nouter;
  l.n_i = ninner;
  l.os = outer_stride;
  l.is = 1;
  l.es = col_stride;
  return gfft_run(p->gax[d], GM_C2C, dir, in, l, out, l, 1.0f, p->gtmp, st);
}

// nblk real blocks [N0..N_{D-1}] -> half spectra [N0..][P]
int xf_forward_blocks(tfsm_plan* p, const float* u, float2* out, size_t nblk, hipStream_t st) {
  const int D = p->D, nl = p->n[D - 1];
  size_t rows = nblk;
  for (int d = 0; d < D - 1; ++d) rows *= p->n[d];
  HIP_TRY(xf_rows(p, GM_R2C, -1, u, (size_t)nl, out, (size_t)p->P, rows, 1.0f, st), "r2c launch");
  if (D == 2)
    HIP_TRY(xf_cols(p, 0, -1, out, out, nblk, p->modes, (size_t)p->P, (size_t)p->H, st), "column fft launch");
  if (D == 3) {
    const size_t plane = (size_t)p->n[1] * p->P;
    HIP_TRY(xf_cols(p, 1, -1, out, out, nblk * p->n[0], plane, (size_t)p->P, (size_t)p->H, st), "y fft launch");
    HIP_TRY(xf_cols(p, 0, -1, out, out, nblk, p->modes, plane, plane, st), "x fft launch");
  }
  return TFSM_OK;
}

// nblk half spectra -> real blocks, scaled by `scale`; `work` receives the partial transforms
// (it may be `in`, which is then overwritten; unused in 1D)
int xf_inverse_blocks(tfsm_plan* p, const float2* in, float2* work, float* u, size_t nblk, float scale,
                      hipStream_t st) {
  const int D = p->D, nl = p->n[D - 1];
  size_t rows = nblk;
  for (int d = 0; d < D - 1; ++d) rows *= p->n[d];
  if (D == 2) {
    HIP_TRY(xf_cols(p, 0, +1, in, work, nblk, p->modes, (size_t)p->P, (size_t)p->H, st), "column ifft launch");
    in = work;
  }
  if (D == 3) {
    const size_t plane = (size_t)p->n[1] * p->P;
    HIP_TRY(xf_cols(p, 0, +1, in, work, nblk, p->modes, plane, plane, st), "x ifft launch");
    HIP_TRY(xf_cols(p, 1, +1, work, work, nblk * p->n[0], plane, (size_t)p->P, (size_t)p->H, st), "y ifft launch");
    in = work;
  }
  HIP_TRY(xf_rows(p, GM_C2R, +1, in, (size_t)p->P, u, (size_t)nl, rows, scale, st), "c2r launch");
  return TFSM_OK;
}

// ---------------------------------------------------------------- unfused stepping
size_t plan_npts(const tfsm_plan* p) {
  size_t n = 1;
  for (int d = 0; d < p->D; ++d) n *= p->n[d];
  return n;
}

GenArgs gen_args(const tfsm_plan* p) {
  GenArgs g;
  g.term = p->term;
  g.D = p->D;
  g.B = p->B;
  g.C = p->C;
  for (int d = 0; d < 3; ++d) {
    g.n[d] = p->n[d];
    g.kmax[d] = p->kmax[d];
    g.kv[d] = p->kv[d];
  }
  g.H = p->H;
  g.P = p->P;
  g.modes = p->modes;
  g.npts = plan_npts(p);
  g.kv_bstride = p->kv_bstride;
  g.NF = gen_nf(p->term, p->C, p->D);
  g.NP = gen_np(p->term, p->C);
  g.coef_t = p->coef_t;
  return g;
}

// unfused buffers after the fused workspace: S (stage spectrum), F (fields, later product spectra),
// R (real fields), Q (real products)
void gen_layout(tfsm_plan* p) {
  const GenArgs g = gen_args(p);
  const size_t sb = p->state_elems() * sizeof(float2);
  const size_t nfa = (size_t)(g.NF > g.NP ? g.NF : g.NP);
  p->off_gS = roundup(p->ws_fused, 256);
  p->off_gF = p->off_gS + roundup(sb, 256);
  p->off_gR = p->off_gF + roundup((size_t)p->B * nfa * p->modes * sizeof(float2), 256);
  p->off_gQ = p->off_gR + roundup((size_t)p->B * g.NF * g.npts * sizeof(float), 256);
  p->ws_bytes = p->off_gQ + roundup((size_t)p->B * g.NP * g.npts * sizeof(float), 256);
}

// general-length tables of the axes without fixed-size kernels, and their four-step scratch
int build_gaxes(tfsm_plan* p) {
  size_t tmp = 0;
  // blocks per transform: state channels, fields and products of both paths, adjoint fields
  int nb = p->C;
  const int counts[4] = {gen_nf(p->term, p->C, p->D), gen_np(p->term, p->C), 2 * p->C + p->C * p->D, 5};
  for (int c : counts) nb = c > nb ? c : nb;
  const size_t blk = (size_t)p->B * nb;
  for (int d = 0; d < p->D; ++d) {
    if (!axis_needs_gfft(p, d)) continue;
    hipError_t e = gaxis_build(p->gax[d], p->n[d], p->allocs);
    if (e == hipErrorInvalidValue) return fail(TFSM_E_INVALID, "mesh axis too long for the general-length FFT");
    HIP_TRY(e, "general FFT tables");
    // lines of one transform along axis d (full complex rows of tfsm_c2c included)
    size_t lines = blk;
    for (int e2 = 0; e2 < p->D - 1; ++e2)
      if (e2 != d) lines *= p->n[e2];
    if (d != p->D - 1) lines *= (size_t)p->n[p->D - 1] > (size_t)p->P ? (size_t)p->n[p->D - 1] : (size_t)p->P;
    const size_t t = gfft_tmp_elems(p->gax[d], lines);
    if (t > tmp) tmp = t;
  }
  if (tmp) return dev_alloc(p, &p->gtmp, tmp, "general FFT scratch");
  return TFSM_OK;
}

// N(S) -> product spectra in F, and the batch mean of their DC for KSConvection; mask_in_place:
// S is zeroed outside the dealiasing box (the forced NSPressureConvection core does that)
int gen_eval(tfsm_plan* p, float2* S, char* ws, hipStream_t st, bool mask_in_place = false) {
  GenArgs g = gen_args(p);
  g.mask_in_place = mask_in_place ? 1 : 0;
  float2* F = (float2*)(ws + p->off_gF);
  float* R = (float*)(ws + p->off_gR);
  float* Q = (float*)(ws + p->off_gQ);
  HIP_TRY(launch_gen_fields(g, S, F, st), "field launch");
  int rc = xf_inverse_blocks(p, F, F, R, (size_t)p->B * g.NF, (float)(1.0 / (double)g.npts), st);
  if (rc) return rc;
  HIP_TRY(launch_gen_products(g, R, Q, st), "product launch");
  rc = xf_forward_blocks(p, Q, F, (size_t)p->B * g.NP, st);
  if (rc) return rc;
  if (p->ks_mean)
    HIP_TRY(launch_ks_dc(F, p->B, (size_t)g.NP * p->modes, 1, 0, (float2*)(ws + p->off_DC), st), "batch-mean launch");
  return TFSM_OK;
}

// evaluation from `in` and the stage combine `stage` (S receives the next stage spectrum).
// stepping: a forced NSPressureConvection masks the stage tensor it evaluates in place, as the
// reference's core does, except the state of a one-evaluation scheme, whose E u the reference
// forms before the evaluation (_etdrk.py:47-51, _setdrk_step.py:5-11)
int gen_stage(tfsm_plan* p, float2* in, float2* U, char* ws, int stage, StepMask m, hipStream_t st,
              bool stepping = false) {
  const bool forced = p->force != nullptr;
  int rc = gen_eval(p, in, ws, st, forced && stepping && stage != ST_R1_E0);
  if (rc) return rc;
  GenStageArgs s;
  s.stage = stage;
  s.Qs = (const float2*)(ws + p->off_gF);
  s.dc = p->ks_mean ? (const float2*)(ws + p->off_DC) : nullptr;
  s.U = U;
  s.A1 = (float2*)(ws + p->off_A1);
  s.A2 = (float2*)(ws + p->off_A2);
  s.S = (float2*)(ws + p->off_gS);
  s.coef = p->coef;
  s.coefc = p->cplx ? p->coefc : nullptr;
  s.coef_bstride = p->coef_bstride();
  s.coef_per_block = p->per_len ? 1 : 0;
  s.src = p->src;
  s.src_bstride = p->src_bstride();
  s.src_cstride = p->src_cstride();
  s.force = p->force;
  s.force_bstride = p->forceB > 1 ? (size_t)p->forceC * p->modes : 0;
  s.force_cstride = p->forceC > 1 ? p->modes : 0;
  s.masked_stage = forced && stepping ? 1 : 0;
  s.mask = m;
  HIP_TRY(launch_gen_stage(gen_args(p), s, st), "stage launch");
  return TFSM_OK;
}

int gen_step(tfsm_plan* p, float2* U, int nsteps, const int* rec_slot, void* traj, int T, char* ws, hipStream_t st,
             const int* act) {
  const int nev = scheme_evals(p->scheme);
  const size_t per = (size_t)p->C * p->modes;
  float2* S = (float2*)(ws + p->off_gS);
  StepMask m;
  m.act = act;
  int rc = TFSM_OK;
  for (int i = 0; i < nsteps && !rc; ++i) {
    m.it = i;
    if (rec_slot) rc = launch_status(launch_record(U, (float2*)traj, rec_slot, i, p->B, per, T, st), "record launch");
    for (int e = 0; e < nev && !rc; ++e)
      rc = gen_stage(p, e == 0 ? U : S, U, ws, scheme_stage(p->scheme, e), m, st, true);
  }
  if (rec_slot && !rc) rc = launch_status(launch_record(U, (float2*)traj, rec_slot, nsteps, p->B, per, T, st), "record launch");
  return rc;
}

int check_stream_errors(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, what);
  return TFSM_OK;
}

Row2Args row_args(tfsm_plan* p, char* ws, int b0, int nb) {
  Row2Args r;
  r.B = nb;
  r.N0 = p->n[0];
  r.N1 = p->n[1];
  r.H1 = p->H;
  r.P1 = p->P;
  r.PF = p->PF;
  r.kmax1 = p->kmax[1];
  r.NF = p->NF;
  r.tw1 = p->tw[1][p->row_var];
  r.ay = p->kv[1];
  r.G = (const float2*)(ws + p->off_G);
  r.P = (float2*)(ws + p->off_P);
  r.b0 = b0;
  return r;
}

Col2Args col_args(tfsm_plan* p, float2* U, char* ws) {
  Col2Args c;
  c.B = p->B;
  c.C = p->C;
  c.N0 = p->n[0];
  c.N1 = p->n[1];
  c.H1 = p->H;
  c.P1 = p->P;
  c.PF = p->PF;
  c.kmax0 = p->kmax[0];
  c.kmax1 = p->kmax[1];
  c.NF = p->NF;
  c.coef_t = p->coef_t;
  c.scale = (float)(1.0 / ((double)p->n[0] * p->n[1]));
  c.tw0 = p->tw[0][p->col_var];
  c.ax = p->kv[0];
  c.ay = p->kv[1];
  c.P = (const float2*)(ws + p->off_P);
  c.U = U;
  c.A1 = (float2*)(ws + p->off_A1);
  c.A2 = (float2*)(ws + p->off_A2);
  c.coef = p->coef;
  c.coef_bstride = p->coef_bstride();
  c.cpk = p->cpk;
  c.cpk_bstride = p->Bc > 1 ? 2 * p->modes : 0;
  c.c5t = p->c5t;
  c.ck01 = p->scheme == SETDRK4 ? p->ck01 : nullptr;
  c.c6t = p->scheme == SETDRK4 ? p->c6t : nullptr;
  c.src = p->src;
  c.src_bstride = p->src_bstride();
  c.src_cstride = p->src_cstride();
  c.src_cols = p->src_cols;
  if (p->src_bits_ok) {
    for (int w = 0; w < Col2Args::SRC_BIT_WORDS; ++w) c.src_bits[w] = p->src_bits[w];
    c.src_bits_ok = 1;
  }
  c.G = (float2*)(ws + p->off_G);
  c.dc = p->ks_mean ? (const float2*)(ws + p->off_DC) : nullptr;
  c.b0 = 0;
  c.nb = p->B;
  return c;
}

int launch_row(tfsm_plan* p, char* ws, int b0, int nb, hipStream_t st, StepMask m = StepMask()) {
  Row2Args r = row_args(p, ws, b0, nb);
  r.mask = m;
  HIP_TRY(launch_row2d(p->term, r, p->row_var, p->row_wg, st), "row kernel launch");
  return TFSM_OK;
}

int launch_col(tfsm_plan* p, float2* U, char* ws, int stage, int b0, int nb, hipStream_t st, StepMask m = StepMask()) {
  Col2Args c = col_args(p, U, ws);
  c.b0 = b0;
  c.mask = m;
  HIP_TRY(launch_col2d(p->term, c, nb, stage, p->V, p->col_var, st), "column kernel launch");
  return TFSM_OK;
}

// one nonlinear-evaluation pair of 2D kernels for samples [b0, b0 + nb)
int run2d_eval(tfsm_plan* p, float2* U, char* ws, int stage, int b0, int nb, hipStream_t st, StepMask m) {
  if (stage != ST_PRO) {
    int rc = launch_row(p, ws, b0, nb, st, m);
    if (rc) return rc;
    // KSConvection's mean couples the batch (the plan never splits it into sub-batches)
    if (p->ks_mean)
      HIP_TRY(launch_ks_dc((const float2*)(ws + p->off_P), p->B, p->modes, p->n[0], (size_t)p->P,
                           (float2*)(ws + p->off_DC), st),
              "batch-mean launch");
  }
  return launch_col(p, U, ws, stage, b0, nb, st, m);
}

// ---------------------------------------------------------------- 3D evaluation
Col3Args col3_args(tfsm_plan* p, float2* U, char* ws, int b0, int nb) {
  Col3Args c;
  c.B = p->B;
  c.N0 = p->n[0];
  c.N1 = p->n[1];
  c.N2 = p->n[2];
  c.H2 = p->H;
  c.P2 = p->P;
  c.PF2 = p->PF;
  c.K1 = p->K1;
  c.kmax0 = p->kmax[0];
  c.kmax1 = p->kmax[1];
  c.kmax2 = p->kmax[2];
  c.term = p->term;
  c.nc = p->C;
  c.coef_t = p->coef_t;
  c.scale = (float)(1.0 / ((double)p->n[0] * p->n[1] * p->n[2]));
  c.tw0 = p->tw[0][p->col_var == 1 ? 3 : var3(p->n[0])];
  c.ax = p->kv[0];
  c.ay = p->kv[1];
  c.az = p->kv[2];
  c.P = (const float2*)(ws + p->off_P);
  c.U = U;
  c.A1 = (float2*)(ws + p->off_A1);
  c.A2 = (float2*)(ws + p->off_A2);
  c.cpk = p->cpk;
  c.cpk_bstride = p->Bc > 1 ? 2 * p->modes : 0;
  c.c5t = p->c5t;
  c.ck01 = p->scheme == SETDRK4 ? p->ck01 : nullptr;
  c.c6t = p->scheme == SETDRK4 ? p->c6t : nullptr;
  c.src = p->src;
  c.src_bstride = p->src_bstride();
  c.src_cstride = p->src_cstride();
  c.G1 = (float2*)(ws + p->off_G);
  c.dc = p->ks_mean ? (const float2*)(ws + p->off_DC) : nullptr;
  c.b0 = b0;
  c.nb = nb;
  return c;
}

// one nonlinear evaluation of the 3D convection term for samples [b0, b0 + nb):
// z rows (products) -> y columns forward -> x columns (combine, fields) -> y columns inverse
// `parts` selects kernels (bit 0 z rows, 1 y forward, 2 x columns, 3 y inverse) for profiling.
int run3d_eval(tfsm_plan* p, float2* U, char* ws, int stage, int b0, int nb, hipStream_t st, int parts = 15,
               StepMask m = StepMask()) {
  const int N0 = p->n[0], N1 = p->n[1];
  float2* Pp = (float2*)(ws + p->off_P);
  if (stage != ST_PRO && (parts & 1)) {
    Row3Args r;
    r.term = p->term;
    r.N0 = N0;
    r.N1 = N1;
    r.N2 = p->n[2];
    r.P2 = p->P;
    r.PF2 = p->PF;
    r.kmax2 = p->kmax[2];
    r.tw2 = p->tw[2][p->row_var == 1 ? 3 : var3(p->n[2])];
    r.az = p->kv[2];
    r.G2 = (const float2*)(ws + p->off_G2);
    r.P = Pp;
    r.b0 = b0;
    r.nb = nb;
    r.mask = m;
    HIP_TRY(launch_row3d(r, p->row_var, st), "3D row kernel launch");
  }
  if (stage != ST_PRO && (parts & 2)) {
    float2* Pb = Pp + (size_t)b0 * p->C * p->modes;
    HIP_TRY(launch_cols_c2c(N1, -1, Pb, Pb, (size_t)nb * p->C * N0, (size_t)N1 * p->P, (size_t)p->P, p->H, 1.0f,
                            p->tw[1][0], st, m, b0, (size_t)p->C * N0),
            "3D y-forward launch");
    if (p->ks_mean)
      HIP_TRY(launch_ks_dc(Pp, p->B, p->modes, N0, (size_t)N1 * p->P, (float2*)(ws + p->off_DC), st),
              "batch-mean launch");
  }
  if (parts & 4) {
    Col3Args c = col3_args(p, U, ws, b0, nb);
    c.mask = m;
    HIP_TRY(launch_col3d(c, nb, stage, p->col_var, st), "3D column kernel launch");
  }
  if (!(parts & 8)) return TFSM_OK;
  Yinv3Args y;
  y.N0 = N0;
  y.N1 = N1;
  y.PF2 = p->PF;
  y.K1 = p->K1;
  y.kmax1 = p->kmax[1];
  y.kmax2 = p->kmax[2];
  y.nc = p->C;
  y.tw1 = p->tw[1][p->col_var == 1 ? 3 : var3(N1)];
  y.ay = p->kv[1];
  y.G1 = (const float2*)(ws + p->off_G);
  y.G2 = (float2*)(ws + p->off_G2);
  y.b0 = b0;
  y.nb = nb;
  y.mask = m;
  HIP_TRY(launch_yinv3d(y, p->col_var, st), "3D y-inverse launch");
  return TFSM_OK;
}

int run_eval(tfsm_plan* p, float2* U, char* ws, int stage, int b0, int nb, hipStream_t st, StepMask m = StepMask()) {
  return p->D == 3 ? run3d_eval(p, U, ws, stage, b0, nb, st, 15, m) : run2d_eval(p, U, ws, stage, b0, nb, st, m);
}

Run1Args run1_args(tfsm_plan* p, float2* U, char* ws) {
  Run1Args a;
  a.B = p->B * p->C;
  a.N = p->n[0];
  a.H = p->H;
  a.P = p->P;
  a.kmax = p->kmax[0];
  a.coef_t = p->coef_t;
  a.scale = (float)(1.0 / p->n[0]);
  a.scheme = p->scheme;
  a.tw = p->tw[0][0];
  a.ax = p->kv[0];
  a.ax_bstride = p->kv_bstride;
  a.U = U;
  a.A1 = (float2*)(ws + p->off_A1);
  a.A2 = (float2*)(ws + p->off_A2);
  a.coef = p->coef;
  a.coef_bstride = p->coef_bstride();
  a.coefc = p->cplx ? p->coefc : nullptr;
  a.cons = p->term == TERM_CONS ? 1 : 0;
  a.src = p->src;
  a.src_bstride = p->src_bstride();
  a.nsteps = 0;
  a.nsteps_per = nullptr;
  a.rec_slot = nullptr;
  a.traj = nullptr;
  a.T = 0;
  a.eval_only = 0;
  return a;
}

// act: optional per-sample step counts (device int32 [B], 2D / 3D): sample b advances by
// min(act[b], nsteps) steps and is not touched after that
int step_impl(tfsm_plan* p, void* uh, int nsteps, const int* rec_slot, void* traj, int T, void* ws, hipStream_t st,
              const int* act = nullptr) {
  float2* U = (float2*)uh;
  if (p->scheme == ETDRK0) {
    if (p->cplx)
      HIP_TRY(launch_linear_steps_c(U, p->coefc + CS_E * p->modes, p->coef_bstride(), p->modes, p->C,
                                    p->state_elems(), nsteps, rec_slot, (float2*)traj, T, st),
              "linear step launch");
    else
      HIP_TRY(launch_linear_steps(U, p->coef + CS_E * p->modes, p->coef_bstride(), p->modes, p->C, p->state_elems(),
                                  nsteps, rec_slot, (float2*)traj, T, st),
              "linear step launch");
    return TFSM_OK;
  }
  if (!ws) return fail(TFSM_E_INVALID, "workspace pointer is null");
  if (p->general) return gen_step(p, U, nsteps, rec_slot, traj, T, (char*)ws, st, act);
  if (p->D == 1 && p->term == TERM_KS) {
    // one launch per stage (the batch mean couples the samples): raw product -> mean -> combine
    Run1Args a = run1_args(p, U, (char*)ws);
    float2* S = (float2*)((char*)ws + p->off_S);
    float2* Nr = (float2*)((char*)ws + p->off_N);
    float2* dc = p->ks_mean ? (float2*)((char*)ws + p->off_DC) : nullptr;
    const int nev = scheme_evals(p->scheme);
    const size_t per = (size_t)p->C * p->modes;
    for (int i = 0; i < nsteps; ++i) {
      if (rec_slot) HIP_TRY(launch_record(U, (float2*)traj, rec_slot, i, p->B, per, T, st), "record launch");
      for (int e = 0; e < nev; ++e) {
        HIP_TRY(launch_k1_eval_ks(a, e == 0 ? U : S, Nr, st), "1D product launch");
        if (dc) HIP_TRY(launch_ks_dc(Nr, a.B, (size_t)p->P, 1, 0, dc, st), "batch-mean launch");
        HIP_TRY(launch_k1_stage(a, scheme_stage(p->scheme, e), Nr, dc, S, st), "1D stage launch");
      }
    }
    if (rec_slot) HIP_TRY(launch_record(U, (float2*)traj, rec_slot, nsteps, p->B, per, T, st), "record launch");
    return TFSM_OK;
  }
  if (p->D == 1) {
    Run1Args a = run1_args(p, U, (char*)ws);
    a.nsteps = nsteps;
    a.rec_slot = rec_slot;
    a.traj = (float2*)traj;
    a.T = T;
    HIP_TRY(launch_run1d(a, st), "1D stepper launch");
    return TFSM_OK;
  }
  if (p->D == 2 && p->persistent && run2d_supported(p->term, p->n[0], p->n[1], p->PF) && !p->cplx) {
    // one workgroup per sample runs every step (small meshes: launch-bound otherwise)
    Col2Args ca = col_args(p, U, (char*)ws);
    Row2Args ra = row_args(p, (char*)ws, 0, p->B);
    // k2_run runs the default radix plan (variant 0) on both axes, whatever the tiled kernels' tuning
    ca.tw0 = p->tw[0][0];
    ra.tw1 = p->tw[1][0];
    Run2Args r;
    r.b0 = 0;
    r.nb = p->B;
    r.scheme = p->scheme;
    r.nsteps = nsteps;
    r.act = act;
    r.rec_slot = rec_slot;
    r.traj = (float2*)traj;
    r.T = T;
    r.per_sample = (size_t)p->C * p->modes;
    HIP_TRY(launch_run2d(p->term, ca, ra, r, st), "persistent 2D stepper launch");
    return TFSM_OK;
  }
  if (p->D >= 2) {
    // Temporal blocking: a sub-batch of samples goes through many steps before the next one, so
    // its state, auxiliary stages, products and fields stay resident in the last-level cache
    // across the 2 * evals launches of every step.  Samples are independent: same results.
    // With several streams, sub-batch k runs on stream k % ns (fork from and join back into the
    // caller's stream with events, which also holds under graph capture): their kernels overlap
    // and fill each other's latency.  Sub-batches use disjoint parts of the workspace.
    const int nev = scheme_evals(p->scheme);
    const int sb = (p->sub_batch > 0 && p->sub_batch < p->B) ? p->sub_batch : p->B;
    const size_t per = (size_t)p->C * p->modes;
    // streams in use: no more than there are sub-batches
    const int nsub = (p->B + sb - 1) / sb;
    const int ns = p->nstreams < nsub ? p->nstreams : nsub;
    if (ns > 1) {
      HIP_TRY(hipEventRecord(p->ev_fork, st), "fork event");
      for (int i = 0; i + 1 < ns; ++i) HIP_TRY(hipStreamWaitEvent(p->aux[i], p->ev_fork, 0), "fork wait");
    }
    // On several streams the steps are queued in rounds (every sub-batch gets its next steps
    // before the following round starts): a stream's hardware queue holds a bounded number of
    // packets and the host blocks once one is full, so queuing all steps of one sub-batch first
    // would leave the other streams idle until that sub-batch is nearly done.  A first round of
    // one step starts every stream early.  One stream: a single round, every step of a
    // sub-batch before the next.
    const int round0 = ns > 1 ? 1 : nsteps, round = ns > 1 ? 16 : nsteps;
    const hipStream_t caller = st;
    int rc = TFSM_OK;
    for (int s0 = 0, r = 0; !rc; ++r) {
      const int len = r == 0 ? round0 : round;
      const int s1 = nsteps - s0 < len ? nsteps : s0 + len;
      for (int b0 = 0, k = 0; b0 < p->B && !rc; b0 += sb, ++k) {
        st = (k % ns) ? p->aux[k % ns - 1] : caller;
        const int nb = (p->B - b0) < sb ? (p->B - b0) : sb;
        float2* Ub = U + (size_t)b0 * per;
        float2* Tb = traj ? (float2*)traj + (size_t)b0 * T * per : nullptr;
        StepMask m;
        m.act = act;
        m.it = s0;
        if (s0 == 0) rc = run_eval(p, U, (char*)ws, ST_PRO, b0, nb, st, m);
        for (int i = s0; i < s1 && !rc; ++i) {
          m.it = i;
          if (rec_slot) rc = launch_status(launch_record(Ub, Tb, rec_slot, i, nb, per, T, st), "record launch");
          for (int e = 0; e < nev && !rc; ++e)
            rc = run_eval(p, U, (char*)ws, scheme_stage(p->scheme, e), b0, nb, st, m);
        }
        if (s1 == nsteps && rec_slot && !rc)
          rc = launch_status(launch_record(Ub, Tb, rec_slot, nsteps, nb, per, T, st), "record launch");
      }
      if (s1 == nsteps) break;
      s0 = s1;
    }
    // the joins are recorded on every exit, also after a failed launch: work already queued on the
    // auxiliary streams is then still ordered before anything the caller queues next
    for (int i = 0; i + 1 < ns; ++i) {
      const int jr = launch_status(hipEventRecord(p->ev_join[i], p->aux[i]), "join event");
      const int jw = jr ? jr : launch_status(hipStreamWaitEvent(caller, p->ev_join[i], 0), "join wait");
      if (!rc) rc = jw;
    }
    return rc;
  }
  return fail(TFSM_E_INVALID, "stepping is not available for this dimension");
}

// ---------------------------------------------------------------- adjoint helpers
AdjArgs adj_args(tfsm_plan* p) {
  AdjArgs a;
  a.D = p->D;
  a.C = p->C;
  a.B = p->B;
  a.npts = 1;
  for (int d = 0; d < 3; ++d) {
    a.n[d] = p->n[d];
    a.kmax[d] = p->kmax[d];
    a.kv[d] = p->kv[d];
    if (d < p->D) a.npts *= p->n[d];
  }
  a.H = p->H;
  a.P = p->P;
  a.modes = p->modes;
  a.coef_t = p->coef_t;
  a.term = p->term;
  return a;
}

struct AdjLayout {
  size_t Aa, N0s, T1, Jv, Ab, Nb, Na, Ac, Nc, Cb, Bb, Abar, gpart, F, Fs, X, Q, total;
};

// fields and products of the transposed Jacobian (k_adj_fields / k_adj_products)
int adj_nfa(const tfsm_plan* p) { return p->term == TERM_VC ? 5 : 2 * p->C + p->C * p->D; }
int adj_nq(const tfsm_plan* p) { return p->term == TERM_VC ? 4 : p->C + p->C * p->D; }

AdjLayout adj_layout(const tfsm_plan* p) {
  const int nfa = adj_nfa(p), nq = adj_nq(p);
  size_t npts = 1;
  for (int d = 0; d < p->D; ++d) npts *= p->n[d];
  const size_t S = roundup(p->state_elems() * sizeof(float2), 256);
  AdjLayout l;
  size_t o = 0;
  l.Aa = o; o += S;
  l.N0s = o; o += S;
  l.T1 = o; o += S;
  l.Jv = o; o += S;
  l.Ab = o; o += S;
  l.Nb = o; o += S;
  // SETDRK4 only: Na, c, Nc and the stage adjoints
  const bool four = p->scheme == SETDRK4;
  l.Na = o; o += four ? S : 0;
  l.Ac = o; o += four ? S : 0;
  l.Nc = o; o += four ? S : 0;
  l.Cb = o; o += four ? S : 0;
  l.Bb = o; o += four ? S : 0;
  l.Abar = o; o += four ? S : 0;
  l.gpart = o; o += roundup(p->state_elems() * sizeof(float), 256);
  l.F = o; o += roundup((size_t)p->B * nfa * p->modes * sizeof(float2), 256);
  l.Fs = o; o += roundup((size_t)p->B * (nfa > nq ? nfa : nq) * p->modes * sizeof(float2), 256);
  l.X = o; o += roundup((size_t)p->B * nfa * npts * sizeof(float), 256);
  l.Q = o; o += roundup((size_t)p->B * nq * npts * sizeof(float), 256);
  l.total = o;
  return l;
}

// W = J_N(U)^T V (Convection, 1D / 2D)
int adj_jt(tfsm_plan* p, const float2* U, const float2* V, float2* W, char* wsa, hipStream_t st) {
  const AdjLayout l = adj_layout(p);
  const AdjArgs a = adj_args(p);
  const int nfa = adj_nfa(p), nq = adj_nq(p);
  float2* F = (float2*)(wsa + l.F);
  float2* Fs = (float2*)(wsa + l.Fs);
  float* X = (float*)(wsa + l.X);
  float* Q = (float*)(wsa + l.Q);
  HIP_TRY(launch_adj_fields(a, U, V, F, st), "adjoint fields launch");
  const float scale = (float)(1.0 / (double)a.npts);
  int rc = xf_inverse_blocks(p, F, Fs, X, (size_t)p->B * nfa, scale, st);
  if (rc) return rc;
  HIP_TRY(launch_adj_products(a, X, Q, st), "adjoint products launch");
  rc = xf_forward_blocks(p, Q, Fs, (size_t)p->B * nq, st);
  if (rc) return rc;
  HIP_TRY(launch_adj_combine(a, Fs, W, st), "adjoint combine launch");
  return TFSM_OK;
}

// N(U) -> workspace A1
int eval_nonlinear(tfsm_plan* p, float2* U, char* ws, hipStream_t st) {
  if (p->general) return gen_stage(p, U, U, ws, ST_EVAL, StepMask(), st);
  if (p->D == 1) {
    Run1Args r = run1_args(p, U, ws);
    r.eval_only = 1;
    r.nsteps = 1;
    HIP_TRY(launch_run1d(r, st), "1D evaluation launch");
    return TFSM_OK;
  }
  int rc = run_eval(p, U, ws, ST_PRO, 0, p->B, st);
  if (!rc) rc = run_eval(p, U, ws, ST_EVAL, 0, p->B, st);
  return rc;
}

AdjLinArgs lin_args(tfsm_plan* p, float2* Y) {
  AdjLinArgs a;
  a.B = p->B;
  a.C = p->C;
  a.P = p->P;
  a.H = p->H;
  a.modes = p->modes;
  a.coef = p->coef;
  a.coef_bstride = p->coef_bstride();
  a.nt = 0;
  a.Y = Y;
  return a;
}

void lin_term(AdjLinArgs& a, const float2* X, int slot, float s) {
  a.X[a.nt] = X;
  a.slot[a.nt] = slot;
  a.s[a.nt] = s;
  ++a.nt;
}

AdjGradArgs grad_args(tfsm_plan* p, const float2* Xbar, float* gpart, int accumulate) {
  AdjGradArgs g;
  g.B = p->B;
  g.C = p->C;
  g.P = p->P;
  g.H = p->H;
  g.nlast = p->n[p->D - 1];
  g.modes = p->modes;
  g.dcoef = p->dcoef;
  g.dcoef_bstride = p->Bc > 1 ? (size_t)ADJ_NDCOEF * p->modes : 0;
  g.sym = p->sym;
  g.Xbar = Xbar;
  g.nt = 0;
  g.gpart = gpart;
  g.accumulate = accumulate;
  return g;
}

void grad_term(AdjGradArgs& g, const float2* Y, int dslot, float s) {
  g.Y[g.nt] = Y;
  g.dslot[g.nt] = dslot;
  g.s[g.nt] = s;
  ++g.nt;
}

// One adjoint SETDRK4 step (see k_adjoint.hip): recompute N0, a, Na, b, Nb, c (and Nc for the
// coefficient gradient) from u, then the transposed stages from c back to u.
int adjoint_step4(tfsm_plan* p, float2* U, const float2* Ubn, float2* Ubo, double* grad, char* W, char* A,
                  hipStream_t st) {
  enum { D_E = 0, D_E2 = 1, D_C1 = 2, D_C4 = 3, D_C5 = 4, D_C6 = 5 };
  const AdjLayout l = adj_layout(p);
  float2* A1 = (float2*)(W + p->off_A1);
  float2* N0 = (float2*)(A + l.N0s);
  float2* Sa = (float2*)(A + l.Aa);
  float2* Na = (float2*)(A + l.Na);
  float2* Sb = (float2*)(A + l.Ab);
  float2* Nb = (float2*)(A + l.Nb);
  float2* Sc = (float2*)(A + l.Ac);
  float2* Nc = (float2*)(A + l.Nc);
  float2* T = (float2*)(A + l.T1);
  float2* Jv = (float2*)(A + l.Jv);
  float2* Cb = (float2*)(A + l.Cb);
  float2* Bb = (float2*)(A + l.Bb);
  float2* Ab = (float2*)(A + l.Abar);
  float* gp = (grad && p->adj_term >= 0) ? (float*)(A + l.gpart) : nullptr;
  const size_t sbytes = p->state_elems() * sizeof(float2);
  auto nonlin = [&](float2* X, float2* out) -> int {
    int rc = eval_nonlinear(p, X, W, st);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(out, A1, sbytes, hipMemcpyDeviceToDevice, st), "adjoint copy");
    return TFSM_OK;
  };
  auto comb = [&](AdjLinArgs& a) -> int {
    HIP_TRY(launch_adj_lincomb(a, st), "adjoint stage launch");
    return TFSM_OK;
  };
  int rc = nonlin(U, N0);
  if (rc) return rc;
  AdjLinArgs a = lin_args(p, Sa);  // a = E2 u + c1 N0
  lin_term(a, U, CS_E2, 1.f);
  lin_term(a, N0, CS_C1, 1.f);
  if ((rc = comb(a)) || (rc = nonlin(Sa, Na))) return rc;
  a = lin_args(p, Sb);  // b = E2 u + c1 Na
  lin_term(a, U, CS_E2, 1.f);
  lin_term(a, Na, CS_C1, 1.f);
  if ((rc = comb(a)) || (rc = nonlin(Sb, Nb))) return rc;
  a = lin_args(p, Sc);  // c = E2 a + c1 (2 Nb - N0)
  lin_term(a, Sa, CS_E2, 1.f);
  lin_term(a, Nb, CS_C1, 2.f);
  lin_term(a, N0, CS_C1, -1.f);
  if ((rc = comb(a))) return rc;
  if (gp && (rc = nonlin(Sc, Nc))) return rc;
  // backward
  a = lin_args(p, T);  // Nc bar = c6 u'bar
  lin_term(a, Ubn, CS_C6, 1.f);
  if ((rc = comb(a)) || (rc = adj_jt(p, Sc, T, Cb, A, st))) return rc;
  a = lin_args(p, T);  // Nb bar = 2 c5 u'bar + 2 c1 cbar
  lin_term(a, Ubn, CS_C5, 2.f);
  lin_term(a, Cb, CS_C1, 2.f);
  if ((rc = comb(a)) || (rc = adj_jt(p, Sb, T, Bb, A, st))) return rc;
  a = lin_args(p, T);  // Na bar = 2 c5 u'bar + c1 bbar
  lin_term(a, Ubn, CS_C5, 2.f);
  lin_term(a, Bb, CS_C1, 1.f);
  if ((rc = comb(a)) || (rc = adj_jt(p, Sa, T, Jv, A, st))) return rc;
  a = lin_args(p, Ab);  // abar = E2 cbar + J(a)^T Na bar
  lin_term(a, Cb, CS_E2, 1.f);
  lin_term(a, Jv, -1, 1.f);
  if ((rc = comb(a))) return rc;
  a = lin_args(p, T);  // N0 bar = c4 u'bar - c1 cbar + c1 abar
  lin_term(a, Ubn, CS_C4, 1.f);
  lin_term(a, Cb, CS_C1, -1.f);
  lin_term(a, Ab, CS_C1, 1.f);
  if ((rc = comb(a)) || (rc = adj_jt(p, U, T, Jv, A, st))) return rc;
  a = lin_args(p, Ubo);  // ubar = E u'bar + E2 bbar + E2 abar + J(u)^T N0 bar
  lin_term(a, Ubn, CS_E, 1.f);
  lin_term(a, Bb, CS_E2, 1.f);
  lin_term(a, Ab, CS_E2, 1.f);
  lin_term(a, Jv, -1, 1.f);
  if ((rc = comb(a))) return rc;
  if (!gp) return TFSM_OK;
  AdjGradArgs g = grad_args(p, Ubn, gp, 0);
  grad_term(g, U, D_E, 1.f);
  grad_term(g, N0, D_C4, 1.f);
  grad_term(g, Na, D_C5, 2.f);
  grad_term(g, Nb, D_C5, 2.f);
  grad_term(g, Nc, D_C6, 1.f);
  HIP_TRY(launch_adj_gacc(g, st), "adjoint gradient launch");
  g = grad_args(p, Cb, gp, 1);
  grad_term(g, Sa, D_E2, 1.f);
  grad_term(g, Nb, D_C1, 2.f);
  grad_term(g, N0, D_C1, -1.f);
  HIP_TRY(launch_adj_gacc(g, st), "adjoint gradient launch");
  g = grad_args(p, Bb, gp, 1);
  grad_term(g, U, D_E2, 1.f);
  grad_term(g, Na, D_C1, 1.f);
  HIP_TRY(launch_adj_gacc(g, st), "adjoint gradient launch");
  g = grad_args(p, Ab, gp, 1);
  grad_term(g, U, D_E2, 1.f);
  grad_term(g, N0, D_C1, 1.f);
  HIP_TRY(launch_adj_gacc(g, st), "adjoint gradient launch");
  HIP_TRY(launch_adj_reduce(gp, (size_t)p->C * p->modes, p->B, grad, st), "adjoint reduce launch");
  return TFSM_OK;
}

AdjStageArgs stage_args(tfsm_plan* p, int op) {
  AdjStageArgs s;
  s.op = op;
  s.two_stage = (p->scheme == ETDRK2 || p->scheme == SETDRK2) ? 1 : 0;
  s.B = p->B;
  s.C = p->C;
  s.P = p->P;
  s.H = p->H;
  s.nlast = p->n[p->D - 1];
  s.modes = p->modes;
  s.coef = p->coef;
  s.coef_bstride = p->coef_bstride();
  s.dcoef = p->dcoef;
  s.dcoef_bstride = p->Bc > 1 ? (size_t)ADJ_NDCOEF * p->modes : 0;
  s.sym = p->sym;
  s.X0 = s.X1 = s.X2 = s.X3 = nullptr;
  s.Y0 = s.Y1 = nullptr;
  s.gpart = nullptr;
  return s;
}

}  // namespace

// Device address of `scratch` on device `dev`, resolved once per device with that device current.
static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

static int scratch_on_device(int dev, void** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(TFSM_E_INVALID, "device index out of range for the static scratch");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* q = nullptr;
    const hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "address of the static scratch");
    if (!q) return fail(TFSM_E_DEVICE, "address of the static scratch is null");
    g_scratch_ptr[dev] = q;
  }
  *out = g_scratch_ptr[dev];
  return TFSM_OK;
}

// nbatch blocks of the plan's [C][N...] shape: at least one, and few enough that every launch's row
// count fits the kernels' 32-bit row index
static int check_nbatch(const tfsm_plan* p, int32_t nbatch) {
  if (nbatch < 1) return fail(TFSM_E_INVALID, "nbatch must be >= 1");
  size_t rows = (size_t)nbatch * p->C;
  for (int d = 0; d < p->D - 1; ++d) rows *= p->n[d];
  if (rows * p->n[p->D - 1] > (size_t)INT32_MAX) return fail(TFSM_E_INVALID, "nbatch too large for one call");
  return TFSM_OK;
}

extern "C" {

const char* tfsm_last_error(void) { return g_err.c_str(); }

int tfsm_version(void) { return (0 << 16) | 1; }

int tfsm_plan_create(const tfsm_geometry* geom, const tfsm_terms* terms, void* stream, tfsm_plan** out) {
  if (!geom || !terms || !out) return fail(TFSM_E_INVALID, "null argument");
  *out = nullptr;
  const int D = geom->ndim;
  if (D < 1 || D > 3) return fail(TFSM_E_INVALID, "the engine supports 1D, 2D and 3D meshes");
  if (geom->batch < 1 || geom->channels < 1) return fail(TFSM_E_INVALID, "batch and channels must be >= 1");
  // sizes the fused kernels cover; any other size (>= 2 points per axis) runs the unfused path
  bool fused_sizes = true;
  for (int d = 0; d < D; ++d) {
    if (!(geom->length[d] > 0)) return fail(TFSM_E_INVALID, "mesh lengths must be positive");
    const int n = geom->n[d];
    if (n < 2) return fail(TFSM_E_INVALID, "mesh axes need at least 2 points");
    const bool ok = D == 1   ? ((is_pow2(n) && n >= 32 && n <= 4096) || is_mixed_size(n))
                    : D == 2 ? ((is_pow2(n) && n >= 16 && n <= 1024) || is_mixed_size_2d(n))
                             : ((is_pow2(n) && n >= 16 && n <= 128) || is_mixed_size_3d(n));
    fused_sizes = fused_sizes && ok;
  }
  const bool ks_keep = terms->nonlinear == TFSM_TERM_KS_KEEP_MEAN;
  const int term = ks_keep ? TERM_KS : terms->nonlinear;
  if (term != TERM_NONE && term != TERM_CONV && term != TERM_VC && term != TERM_NSP && term != TERM_KS &&
      term != TERM_CONS)
    return fail(TFSM_E_INVALID, "unknown nonlinear term");
  if (term == TERM_CONS && geom->channels != D)
    return fail(TFSM_E_INVALID, "convection operator only works for vector field with the same dimension as mesh");
  if (term == TERM_KS && geom->channels != 1) return fail(TFSM_E_INVALID, "KSConvection only supports scalar field");
  if ((term == TERM_CONV || term == TERM_NSP) && geom->channels != D)
    return fail(TFSM_E_INVALID, "convection operator only works for vector field with the same dimension as mesh");
  if (term == TERM_NSP && D == 1)
    return fail(TFSM_E_INVALID, "term not supported by engine: NSPressureConvection on a 1D mesh");
  if (term == TERM_VC && (D != 2 || geom->channels != 1))
    return fail(TFSM_E_INVALID, "Only vorticity in 2Dmesh is supported");
  if (terms->n_linear < 0 || terms->n_linear > 4) return fail(TFSM_E_INVALID, "at most 4 linear terms");
  bool cplx = false;
  for (int t = 0; t < terms->n_linear; ++t) {
    const int kind = terms->linear_kind[t];
    if (kind < TFSM_LIN_LAP || kind > TFSM_LIN_DERIV) return fail(TFSM_E_INVALID, "unknown linear term");
    if (kind != TFSM_LIN_DERIV) continue;
    if (terms->linear_dim[t] < 0 || terms->linear_dim[t] >= D)
      return fail(TFSM_E_INVALID, "SpatialDerivative axis out of range");
    if (terms->linear_order[t] < 1 || terms->linear_order[t] > 8)
      return fail(TFSM_E_INVALID, "SpatialDerivative order must be 1..8");
    if (geom->channels != 1) return fail(TFSM_E_INVALID, "The SpatialDerivative operator only supports scalar field.");
    if (terms->linear_order[t] & 1) cplx = true;
  }
  // the fused kernels cover every product term on their sizes, except ConservativeConvection in 3D
  // and a complex L with a product term in 2D / 3D; the unfused path covers the rest
  bool fused = fused_sizes && !(term == TERM_CONS && D == 3) && !(cplx && term != TERM_NONE && D > 1);
  if (fused && D == 3 && term != TERM_NONE && term != TERM_CONS)
    fused = step3d_supported(geom->n[0], geom->n[1], geom->n[2]);
  if (terms->n_linear > 0 && !terms->linear_coef) return fail(TFSM_E_INVALID, "linear coefficients missing");
  const int Bc = terms->coef_batch < 1 ? 1 : terms->coef_batch;
  if (Bc != 1 && Bc != geom->batch) return fail(TFSM_E_INVALID, "coef_batch must be 1 or the batch size");
  if (geom->length_per_sample && D != 1) return fail(TFSM_E_INVALID, "per-sample lengths are 1D only");

  hipStream_t st = (hipStream_t)stream;
  int device = 0;
  HIP_TRY(hipGetDevice(&device), "current device");
  tfsm_plan* p = new tfsm_plan();
  p->device = device;
  p->D = D;
  p->B = geom->batch;
  p->C = geom->channels;
  for (int d = 0; d < D; ++d) {
    p->n[d] = geom->n[d];
    p->len[d] = geom->length[d];
  }
  p->H = p->n[D - 1] / 2 + 1;
  p->P = (int)roundup(p->H, 16);
  p->modes = (size_t)p->P;
  for (int d = 0; d < D - 1; ++d) p->modes *= p->n[d];
  p->term = term;
  p->ks_mean = term == TERM_KS && !ks_keep;
  p->coef_t = terms->nonlinear_coef;
  p->rate = terms->dealias_rate > 0 ? terms->dealias_rate : 2.0 / 3.0;
  p->nlin = terms->n_linear;
  for (int t = 0; t < p->nlin; ++t) {
    p->lin_kind[t] = terms->linear_kind[t];
    // the axis and order fields are read only for the terms that have them
    if (p->lin_kind[t] == TFSM_LIN_DERIV) {
      p->lin_dim[t] = terms->linear_dim[t];
      p->lin_order[t] = terms->linear_order[t];
    }
  }
  p->cplx = cplx;
  p->fused_ok = fused;
  p->general = !fused && term != TERM_NONE;
  p->per_len = D == 1 && geom->length_per_sample;
  p->Bc = Bc;
  int rc = TFSM_OK;
  for (int d = 0; d < D; ++d) {
    p->kmax[d] = dealias_kmax(p->n[d], p->len[d], p->rate);
    if (p->kmax[d] < -1) {
      tfsm_plan_destroy(p);
      return fail(TFSM_E_INVALID, "dealias mask is not an index range");
    }
    rc = build_twiddles(p, d);
    if (rc) break;
  }
  if (!rc) rc = build_gaxes(p);
  // axis wavenumber tables
  if (!rc) {
    for (int d = 0; d < D && !rc; ++d) {
      if (D == 1 && geom->length_per_sample) {
        std::vector<float> all((size_t)p->B * p->C * p->n[0]), a;
        for (int b = 0; b < p->B; ++b) {
          axis_table(p->n[0], geom->length_per_sample[b], a);
          for (int c = 0; c < p->C; ++c)
            std::memcpy(&all[((size_t)b * p->C + c) * p->n[0]], a.data(), a.size() * sizeof(float));
        }
        rc = dev_alloc(p, &p->kv[0], all.size(), "axis table");
        if (!rc) {
          hipError_t e = hipMemcpy(p->kv[0], all.data(), all.size() * sizeof(float), hipMemcpyHostToDevice);
          if (e != hipSuccess) rc = hip_fail(e, "axis table upload");
        }
        p->kv_bstride = (size_t)p->n[0];
      } else {
        std::vector<float> a;
        axis_table(p->n[d], p->len[d], a);
        rc = dev_alloc(p, &p->kv[d], a.size(), "axis table");
        if (!rc) {
          hipError_t e = hipMemcpy(p->kv[d], a.data(), a.size() * sizeof(float), hipMemcpyHostToDevice);
          if (e != hipSuccess) rc = hip_fail(e, "axis table upload");
        }
      }
    }
  }
  // linear symbol L per (coefficient sample, mode)
  const int Lb = (D == 1 && geom->length_per_sample) ? p->B * p->C : Bc;
  if (D == 1 && geom->length_per_sample) p->Bc = Lb;
  if (!rc) rc = dev_alloc(p, &p->L, (size_t)Lb * p->modes, "linear symbol");
  if (!rc) rc = dev_alloc(p, &p->coef, (size_t)Lb * NCOEF * p->modes, "coefficient table");
  if (!rc) rc = dev_alloc(p, &p->cpk, (size_t)Lb * 2 * p->modes, "packed coefficient table");
  if (!rc) rc = dev_alloc(p, &p->c5t, (size_t)Lb * p->modes, "c5 table");
  if (!rc) rc = dev_alloc(p, &p->ck01, (size_t)Lb * p->modes, "SETDRK4 [E2, c1] table");
  if (!rc) rc = dev_alloc(p, &p->c6t, (size_t)Lb * p->modes, "SETDRK4 c6 table");
  if (!rc && cplx) rc = dev_alloc(p, &p->Lc, (size_t)Lb * p->modes, "complex linear symbol");
  if (!rc && cplx) rc = dev_alloc(p, &p->coefc, (size_t)Lb * NCOEF * p->modes, "complex coefficient table");
  if (!rc) {
    hipError_t e = hipMemsetAsync(p->coef, 0, (size_t)Lb * NCOEF * p->modes * sizeof(float), st);
    if (e != hipSuccess) rc = hip_fail(e, "coefficient table clear");
  }
  if (!rc) {
    LinArgs la;
    la.D = D;
    for (int d = 0; d < 3; ++d) {
      la.n[d] = p->n[d];
      la.kv[d] = p->kv[d];
    }
    la.H = p->H;
    la.P = p->P;
    la.modes = p->modes;
    la.kv_bstride = p->kv_bstride;
    la.nterm = p->nlin;
    for (int t = 0; t < 4; ++t) {
      la.kind[t] = t < p->nlin ? terms->linear_kind[t] : 0;
      la.dim[t] = p->lin_dim[t];
      la.order[t] = p->lin_order[t];
    }
    la.Bc = Lb;
    la.L = p->L;
    la.Lc = p->Lc;
    float* dcoef = nullptr;
    std::vector<float> hc((size_t)(p->nlin > 0 ? p->nlin : 1) * Lb, 0.f);
    for (int t = 0; t < p->nlin; ++t)
      for (int b = 0; b < Lb; ++b) {
        const int src_b = (Bc == 1) ? 0 : (D == 1 && geom->length_per_sample ? b / p->C : b);
        hc[(size_t)t * Lb + b] = terms->linear_coef[(size_t)t * Bc + src_b];
      }
    rc = dev_alloc(p, &dcoef, hc.size(), "linear coefficients");
    if (!rc) {
      hipError_t e = hipMemcpy(dcoef, hc.data(), hc.size() * sizeof(float), hipMemcpyHostToDevice);
      if (e != hipSuccess) rc = hip_fail(e, "linear coefficient upload");
    }
    la.coef = dcoef;
    for (int t = 0; t < 4; ++t) la.cval[t] = 0.f;
    if (!rc) {
      hipError_t e = launch_linear_symbol(la, st);
      if (e != hipSuccess) rc = hip_fail(e, "linear symbol launch");
    }
  }
  // workspace layout
  if (!rc) {
    const size_t sb = p->state_elems() * sizeof(float2);
    p->off_A1 = 0;
    p->off_A2 = roundup(sb, 256);
    if (!p->fused_ok) {
      p->ws_bytes = p->off_A2 + roundup(sb, 256);
    } else if (D == 3) {
      // products P [B][3][N0][N1][P2]; fields G1 [B][6][N0][K1][PF2]; G2 [B][9][N0][N1][PF2]
      p->PF = (int)roundup((size_t)p->kmax[2] + 1, 16);
      p->K1 = 2 * p->kmax[1] + 1;
      p->off_P = p->off_A2 + roundup(sb, 256);
      p->off_G = p->off_P + roundup(sb, 256);
      p->off_G2 = p->off_G + roundup((size_t)p->B * 6 * p->n[0] * p->K1 * p->PF * sizeof(float2), 256);
      p->ws_bytes = p->off_G2 + roundup((size_t)p->B * 9 * p->n[0] * p->n[1] * p->PF * sizeof(float2), 256);
    } else if (D == 2) {
      p->NF = (term == TERM_KS || term == TERM_CONS) ? 2 : 4;
      p->PF = (int)roundup((size_t)p->kmax[1] + 1, 16);
      p->off_P = p->off_A2 + roundup(sb, 256);
      // ConservativeConvection keeps two product planes per channel
      p->off_G = p->off_P + roundup(term == TERM_CONS ? 2 * sb : sb, 256);
      p->ws_bytes = p->off_G + roundup((size_t)p->B * p->NF * p->n[0] * p->PF * sizeof(float2), 256);
    } else {
      p->ws_bytes = p->off_A2 + roundup(sb, 256);
      if (term == TERM_KS) {
        p->off_S = p->ws_bytes;
        p->off_N = p->off_S + roundup(sb, 256);
        p->ws_bytes = p->off_N + roundup(sb, 256);
      }
    }
    p->off_DC = p->ws_bytes;
    p->ws_bytes += 256;
    p->ws_fused = p->ws_bytes;
    if (p->general) gen_layout(p);
  }
  // default launch shape (the autotuner may refine it): 2D columns of 512 points run the
  // half-thread radix plan for one-channel products (the two-channel ones keep a second column in
  // registers and run faster with 8 values per thread); temporal blocking uses the smallest
  // sub-batch (multiple of 8) that still gives >= 512 column workgroups (two per CU) while its
  // working set stays under 256 MB
  if (!rc && D == 2 && p->fused_ok) {
    const bool one_channel = term == TERM_VC || term == TERM_KS;
    const int n0 = p->n[0], n1 = p->n[1];
    // measured per axis size (SETDRK4, B = 16..32): 512 keeps 16-column tiles; 256 runs 8-column
    // tiles with the 8-values-per-thread plan (26 % faster than 16 columns of 16 values); 64 and
    // 128 run 8-column tiles and one row per 64-thread row group (5-9 %)
    p->col_var = n0 == 256 ? 1 : 0;
    if (n0 == 64 || n0 == 128 || n0 == 256) p->V = 8;
    // 1024-point columns: 8 columns per tile (LDS), the 8-values-per-thread plan (the 16-value
    // plan holds 2 waves per SIMD and spills for NSPressureConvection and ConservativeConvection)
    if (n0 == 1024) {
      p->V = 8;
      p->col_var = 1;
    }
    if (n1 == 64 || n1 == 128) p->row_wg = 64;
    const size_t per_sample = 4 * (size_t)p->C * p->modes * sizeof(float2) +
                              (size_t)p->NF * p->n[0] * p->PF * sizeof(float2);
    const int tiles = (p->H + p->V - 1) / p->V;
    int sb = 8;
    while (sb < p->B && (size_t)tiles * sb < 512) sb += 8;
    if (sb >= p->B || (size_t)sb * per_sample > ((size_t)256 << 20) || p->ks_mean) sb = 0;
    p->sub_batch = sb;
    // 512-point one-channel columns (VorticityConvection) with 32 or more samples: sub-batches of 16
    // on two streams, so one sub-batch's VALU-bound row pass overlaps the other's memory-bound
    // column pass (NS 512^2, B = 64: 0.630 ms/step against 0.662 for one stream and 32-sample
    // sub-batches, and 0.676 for the 16-values-per-thread columns it replaces)
    if (n0 == 512 && term == TERM_VC && p->B >= 32) {
      p->sub_batch = 16;
      rc = tfsm_plan_set_streams(p, 2);
    }
    // the same overlap for the other square shape classes where sweeps showed it (SETDRK4, 200
    // steps): 512^2 Convection and NSPressureConvection on two streams, sub-batches of 16 from
    // B = 32 (0.606 -> 0.544 and 0.734 -> 0.682 ms/step), of 8 at B = 16..31 (Convection 0.285 ->
    // 0.248); 256^2 VorticityConvection and Convection in 16-sample sub-batches, on two streams
    // from B = 32 (0.130 -> 0.119, 0.169 -> 0.157) and on four from B = 64 (0.259 -> 0.164, 0.337
    // -> 0.213)
    const bool square = p->n[0] == p->n[1];
    if (!rc && square && n0 == 512 && (term == TERM_CONV || term == TERM_NSP) && p->B >= 16) {
      p->sub_batch = p->B >= 32 ? 16 : 8;
      rc = tfsm_plan_set_streams(p, 2);
    }
    if (!rc && square && n0 == 256 && (term == TERM_VC || term == TERM_CONV) && p->B >= 32) {
      p->sub_batch = 16;
      rc = tfsm_plan_set_streams(p, p->B >= 64 ? 4 : 2);
    }
  }
  // small square meshes step with the persistent per-sample kernel (one launch per call)
  if (!rc && D == 2 && p->fused_ok && run2d_supported(term, p->n[0], p->n[1], p->PF) && !p->cplx)
    p->persistent = true;
  if (!rc && D == 3) p->row_var = (p->n[2] >= 64 && is_pow2(p->n[2])) ? 1 : 0;  // radix-4 z rows (fewer registers)
  // 3D Convection and NSPressureConvection at 128^3 with 4 or more samples: sub-batches of 2 on two
  // streams, so one sub-batch's VALU-bound z-row pass overlaps the other's memory-bound column
  // passes (Burgers, B = 8: ETDRK2 1.52 -> 1.34 ms/step, SETDRK4 3.04 -> 2.71; NavierStokes SETDRK4,
  // B = 4: 1.74 -> 1.66)
  const bool conv3 = D == 3 && p->fused_ok && (term == TERM_CONV || term == TERM_NSP);
  if (!rc && conv3 && p->n[0] == 128 && p->n[1] == 128 && p->n[2] == 128 && p->B >= 4) {
    p->sub_batch = 2;
    rc = tfsm_plan_set_streams(p, 2);
  }
  // 64^3 with 8 or more samples: sub-batches of 2 on four streams (a 2-sample pass is too short to
  // fill the GPU alone; four in flight do): Burgers SETDRK4, B = 8: 0.41 -> 0.37 ms/step, B = 16:
  // 0.93 -> 0.78; NavierStokes SETDRK4, B = 16: 1.09 -> 0.91
  if (!rc && conv3 && p->n[0] == 64 && p->n[1] == 64 && p->n[2] == 64 && p->B >= 8) {
    p->sub_batch = 2;
    rc = tfsm_plan_set_streams(p, 4);
  }
  if (rc) {
    tfsm_plan_destroy(p);
    return rc;
  }
  *out = p;
  return TFSM_OK;
}

int tfsm_plan_destroy(tfsm_plan* plan) {
  if (!plan) return TFSM_OK;
  for (int i = 0; i + 1 < TFSM_MAX_STREAMS; ++i) {
    if (plan->aux[i]) {
      (void)hipStreamSynchronize(plan->aux[i]);
      (void)hipStreamDestroy(plan->aux[i]);
    }
    if (plan->ev_join[i]) (void)hipEventDestroy(plan->ev_join[i]);
  }
  if (plan->ev_fork) (void)hipEventDestroy(plan->ev_fork);
  for (void* q : plan->allocs) (void)hipFree(q);
  delete plan;
  return TFSM_OK;
}

int64_t tfsm_plan_pitch(const tfsm_plan* plan) { return plan ? plan->P : 0; }

int64_t tfsm_plan_state_elems(const tfsm_plan* plan) { return plan ? (int64_t)plan->state_elems() : 0; }

int64_t tfsm_plan_workspace_bytes(const tfsm_plan* plan) { return plan ? (int64_t)plan->ws_bytes : 0; }

int64_t tfsm_plan_inverse_scratch_bytes(const tfsm_plan* plan, int32_t nbatch) {
  if (!plan) return 0;
  if (plan->D == 1) return 0;
  return (int64_t)((size_t)nbatch * plan->C * plan->modes * sizeof(float2));
}

int tfsm_plan_set_tuning(tfsm_plan* plan, int32_t col_tile, int32_t col_var, int32_t row_var, int32_t row_wg,
                         int32_t sub_batch) {
  if (!plan) return fail(TFSM_E_INVALID, "null plan");
  if (!plan->fused_ok) return fail(TFSM_E_INVALID, "no fused kernels for this plan: it has no launch shape");
  if (plan->D == 3 && (row_var < 0 || row_var > 1 || col_var < 0 || col_var > 1))
    return fail(TFSM_E_INVALID, "unsupported 3D tuning");
  if (plan->D == 3 && (col_var == 1 || row_var == 1) &&
      (!is_pow2(plan->n[0]) || !is_pow2(plan->n[1]) || !is_pow2(plan->n[2])))
    return fail(TFSM_E_INVALID, "the radix-4 3D plans need power-of-two axes");
  if (plan->D == 2) {
    if (!col2d_supported(plan->n[0], col_tile, col_var)) return fail(TFSM_E_INVALID, "unsupported column tuning");
    if ((plan->P % col_tile) != 0) return fail(TFSM_E_INVALID, "column tile must divide the pitch");
    if (!row2d_supported(plan->n[1], row_var, row_wg)) return fail(TFSM_E_INVALID, "unsupported row tuning");
  }
  if (sub_batch > 0 && sub_batch < plan->B && plan->ks_mean)
    return fail(TFSM_E_INVALID, "KSConvection couples the batch through its mean: no sub-batches");
  plan->V = col_tile;
  plan->col_var = col_var;
  plan->row_var = row_var;
  plan->row_wg = row_wg;
  plan->sub_batch = sub_batch < 0 ? 0 : sub_batch;
  return TFSM_OK;
}

int tfsm_plan_get_tuning(const tfsm_plan* plan, int32_t* out) {
  if (!plan || !out) return fail(TFSM_E_INVALID, "null argument");
  out[0] = plan->V;
  out[1] = plan->col_var;
  out[2] = plan->row_var;
  out[3] = plan->row_wg;
  out[4] = plan->sub_batch;
  out[5] = plan->nstreams;
  return TFSM_OK;
}

int tfsm_plan_set_persistent(tfsm_plan* plan, int32_t on) {
  if (!plan) return fail(TFSM_E_INVALID, "null plan");
  if (on && plan->general) return fail(TFSM_E_INVALID, "the plan steps through the unfused path");
  if (on && !(plan->D == 2 && run2d_supported(plan->term, plan->n[0], plan->n[1], plan->PF) && !plan->cplx))
    return fail(TFSM_E_INVALID, "the persistent 2D stepper covers square 32^2 and 64^2 meshes with Convection or "
                                "VorticityConvection (2/3 dealiasing or coarser)");
  plan->persistent = on != 0;
  return TFSM_OK;
}

int32_t tfsm_plan_persistent(const tfsm_plan* plan) { return plan && plan->persistent ? 1 : 0; }

int tfsm_plan_set_general(tfsm_plan* plan, int32_t on) {
  if (!plan) return fail(TFSM_E_INVALID, "null plan");
  if (on) {
    if (plan->term == TERM_NONE) return fail(TFSM_E_INVALID, "a plan without a product term has no nonlinear evaluation");
    plan->general = true;
    plan->persistent = false;
    gen_layout(plan);
    return TFSM_OK;
  }
  if (!plan->fused_ok) return fail(TFSM_E_INVALID, "no fused kernels for this plan: it steps through the unfused path");
  plan->general = false;
  plan->ws_bytes = plan->ws_fused;
  return TFSM_OK;
}

int32_t tfsm_plan_general(const tfsm_plan* plan) { return plan && plan->general ? 1 : 0; }

int tfsm_plan_set_streams(tfsm_plan* plan, int32_t nstreams) {
  if (!plan) return fail(TFSM_E_INVALID, "null plan");
  if (nstreams < 1 || nstreams > TFSM_MAX_STREAMS) return fail(TFSM_E_INVALID, "nstreams must be 1 to 4");
  if (nstreams > 1 && plan->ks_mean)
    return fail(TFSM_E_INVALID, "KSConvection couples the batch through its mean: one stream only");
  if (nstreams > 1 && !plan->ev_fork)
    HIP_TRY(hipEventCreateWithFlags(&plan->ev_fork, hipEventDisableTiming), "fork event");
  for (int i = 0; i + 1 < nstreams; ++i) {
    if (plan->aux[i]) continue;
    // default priority, non-blocking with respect to the null stream
    HIP_TRY(hipStreamCreateWithFlags(&plan->aux[i], hipStreamNonBlocking), "auxiliary stream");
    HIP_TRY(hipEventCreateWithFlags(&plan->ev_join[i], hipEventDisableTiming), "join event");
  }
  plan->nstreams = nstreams;
  return TFSM_OK;
}

int tfsm_plan_set_scheme(tfsm_plan* plan, int32_t scheme, double dt, int32_t n_points, double radius, void* stream) {
  if (!plan) return fail(TFSM_E_INVALID, "null plan");
  if (scheme < ETDRK0 || scheme > SETDRK4) return fail(TFSM_E_INVALID, "unknown scheme");
  const bool nonlinear = plan->term != TERM_NONE || plan->src != nullptr;
  if (scheme == ETDRK0 && nonlinear) return fail(TFSM_E_INVALID, "The ETDRK0 integrator only supports linear term");
  if (scheme != ETDRK0 && plan->term == TERM_NONE)
    return fail(TFSM_E_INVALID, "term not supported by engine: a nonlinear scheme needs a product term");
  if (n_points < 1) return fail(TFSM_E_INVALID, "n_integration_points must be >= 1");
  CoefArgs a;
  a.scheme = scheme;
  a.dt = dt;
  a.dt32 = (float)dt;
  a.hdt32 = (float)(0.5 * dt);
  a.M = n_points;
  a.radius = radius;
  a.modes = plan->modes;
  a.Bc = plan->Bc;
  a.L = plan->L;
  a.coef = plan->coef;
  a.cpk = plan->cpk;
  a.c5t = plan->c5t;
  a.ck01 = plan->ck01;
  a.c6t = plan->c6t;
  a.Lc = plan->Lc;
  a.coefc = plan->coefc;
  HIP_TRY(launch_coefs(a, (hipStream_t)stream), "coefficient launch");
  if (!plan->cplx) HIP_TRY(launch_pack_coefs(a, (hipStream_t)stream), "coefficient pack launch");
  plan->scheme = scheme;
  plan->dt = dt;
  plan->n_points = n_points;
  plan->radius = radius;
  return TFSM_OK;
}

int tfsm_plan_set_sample(tfsm_plan* p, int32_t b, const float* linear_coef, double length, void* stream) {
  if (!p) return fail(TFSM_E_INVALID, "null plan");
  if (b < 0 || b >= p->B) return fail(TFSM_E_INVALID, "sample index out of range");
  const int rows_per = p->per_len ? p->C : 1;
  if (p->Bc != p->B * rows_per) return fail(TFSM_E_INVALID, "the plan has shared coefficients (coef_batch 1)");
  if (length > 0 && !p->per_len) return fail(TFSM_E_INVALID, "per-sample lengths need a 1D plan created with them");
  hipStream_t st = (hipStream_t)stream;
  for (int r = b * rows_per; r < (b + 1) * rows_per; ++r) {
    if (length > 0) {
      const int n = p->n[0];
      const double d = length / n;
      HIP_TRY(launch_axis_row(p->kv[0] + (size_t)r * p->kv_bstride, n, (float)(1.0 / (n * d)), st), "axis row launch");
    }
    if (linear_coef || length > 0) {
      if (!linear_coef) return fail(TFSM_E_INVALID, "a new length needs the sample's linear coefficients");
      LinArgs la;
      la.D = p->D;
      for (int d = 0; d < 3; ++d) {
        la.n[d] = p->n[d];
        la.kv[d] = p->kv[d];
      }
      if (p->per_len) la.kv[0] = p->kv[0] + (size_t)r * p->kv_bstride;
      la.H = p->H;
      la.P = p->P;
      la.modes = p->modes;
      la.kv_bstride = 0;
      la.nterm = p->nlin;
      for (int t = 0; t < 4; ++t) {
        la.kind[t] = t < p->nlin ? p->lin_kind[t] : 0;
        la.dim[t] = p->lin_dim[t];
        la.order[t] = p->lin_order[t];
        la.cval[t] = t < p->nlin ? linear_coef[t] : 0.f;
      }
      la.coef = nullptr;
      la.Bc = 1;
      la.L = p->L + (size_t)r * p->modes;
      la.Lc = p->cplx ? p->Lc + (size_t)r * p->modes : nullptr;
      HIP_TRY(launch_linear_symbol(la, st), "linear symbol launch");
    }
    if (p->scheme >= 0) {
      CoefArgs a;
      a.scheme = p->scheme;
      a.dt = p->dt;
      a.dt32 = (float)p->dt;
      a.hdt32 = (float)(0.5 * p->dt);
      a.M = p->n_points;
      a.radius = p->radius;
      a.modes = p->modes;
      a.Bc = 1;
      a.L = p->L + (size_t)r * p->modes;
      a.coef = p->coef + (size_t)r * NCOEF * p->modes;
      a.cpk = p->cpk + (size_t)r * 2 * p->modes;
      a.c5t = p->c5t + (size_t)r * p->modes;
      a.ck01 = p->ck01 + (size_t)r * p->modes;
      a.c6t = p->c6t + (size_t)r * p->modes;
      a.Lc = p->cplx ? p->Lc + (size_t)r * p->modes : nullptr;
      a.coefc = p->cplx ? p->coefc + (size_t)r * NCOEF * p->modes : nullptr;
      HIP_TRY(launch_coefs(a, st), "coefficient launch");
      if (!p->cplx) HIP_TRY(launch_pack_coefs(a, st), "coefficient pack launch");
    }
  }
  return TFSM_OK;
}

int tfsm_plan_coefficients(tfsm_plan* p, int32_t slot, float* out, void* stream) {
  if (!p || !out) return fail(TFSM_E_INVALID, "null argument");
  if (slot < 0 || slot >= NCOEF) return fail(TFSM_E_INVALID, "coefficient slot out of range");
  if (p->scheme < 0) return fail(TFSM_E_STATE, "no scheme set on the plan");
  const size_t bytes = p->modes * sizeof(float);
  for (int b = 0; b < p->Bc; ++b)
    HIP_TRY(hipMemcpyAsync(out + (size_t)b * p->modes, p->coef + ((size_t)b * NCOEF + slot) * p->modes, bytes,
                           hipMemcpyDeviceToDevice, (hipStream_t)stream),
            "coefficient copy");
  return TFSM_OK;
}

int tfsm_forward(tfsm_plan* p, const float* u, void* uh, int32_t nbatch, void* stream) {
  if (!p || !u || !uh) return fail(TFSM_E_INVALID, "null argument");
  if (int rc = check_nbatch(p, nbatch)) return rc;
  return xf_forward_blocks(p, u, (float2*)uh, (size_t)nbatch * p->C, (hipStream_t)stream);
}

int tfsm_inverse(tfsm_plan* p, const void* uh, float* u, int32_t nbatch, void* scratch, void* stream) {
  if (!p || !u || !uh) return fail(TFSM_E_INVALID, "null argument");
  if (int rc = check_nbatch(p, nbatch)) return rc;
  if (p->D > 1 && !scratch) return fail(TFSM_E_INVALID, "inverse transform needs scratch");
  const float scale = (float)(1.0 / (double)plan_npts(p));
  return xf_inverse_blocks(p, (const float2*)uh, (float2*)scratch, u, (size_t)nbatch * p->C, scale,
                           (hipStream_t)stream);
}

int tfsm_c2c(tfsm_plan* p, const void* in, void* out, int32_t nbatch, int32_t dir, void* stream) {
  if (!p || !in || !out) return fail(TFSM_E_INVALID, "null argument");
  if (dir != -1 && dir != 1) return fail(TFSM_E_INVALID, "dir must be -1 or +1");
  if (int rc = check_nbatch(p, nbatch)) return rc;
  hipStream_t st = (hipStream_t)stream;
  const size_t blocks = (size_t)nbatch * p->C;
  const int nl = p->n[p->D - 1];
  const float scale = dir > 0 ? (float)(1.0 / (double)plan_npts(p)) : 1.0f;
  size_t rows = blocks;
  for (int d = 0; d < p->D - 1; ++d) rows *= p->n[d];
  HIP_TRY(xf_rows(p, GM_C2C, dir, in, (size_t)nl, out, (size_t)nl, rows, scale, st), "row c2c launch");
  float2* o = (float2*)out;
  if (p->D == 2)
    HIP_TRY(xf_cols(p, 0, dir, o, o, blocks, (size_t)p->n[0] * nl, (size_t)nl, (size_t)nl, st), "column c2c launch");
  if (p->D == 3) {
    const size_t plane = (size_t)p->n[1] * nl;
    HIP_TRY(xf_cols(p, 1, dir, o, o, blocks * p->n[0], plane, (size_t)nl, (size_t)nl, st), "y c2c launch");
    HIP_TRY(xf_cols(p, 0, dir, o, o, blocks, (size_t)p->n[0] * plane, plane, plane, st), "x c2c launch");
  }
  return TFSM_OK;
}

int tfsm_half_to_full(tfsm_plan* p, const void* uh, void* full, int32_t nbatch, void* stream) {
  if (!p || !uh || !full) return fail(TFSM_E_INVALID, "null argument");
  if (int rc = check_nbatch(p, nbatch)) return rc;
  ConvArgs a;
  a.D = p->D;
  for (int d = 0; d < 3; ++d) a.n[d] = p->n[d];
  a.H = p->H;
  a.P = p->P;
  a.nblk = (size_t)nbatch * p->C;
  HIP_TRY(launch_half_to_full((const float2*)uh, (float2*)full, a, (hipStream_t)stream), "half_to_full launch");
  return TFSM_OK;
}

int tfsm_full_to_half(tfsm_plan* p, const void* full, void* uh, int32_t nbatch, void* stream) {
  if (!p || !uh || !full) return fail(TFSM_E_INVALID, "null argument");
  if (int rc = check_nbatch(p, nbatch)) return rc;
  ConvArgs a;
  a.D = p->D;
  for (int d = 0; d < 3; ++d) a.n[d] = p->n[d];
  a.H = p->H;
  a.P = p->P;
  a.nblk = (size_t)nbatch * p->C;
  HIP_TRY(launch_full_to_half((const float2*)full, (float2*)uh, a, (hipStream_t)stream), "full_to_half launch");
  return TFSM_OK;
}

int tfsm_plan_add_source(tfsm_plan* p, const float* src, int32_t src_batch, int32_t src_channels, float coef,
                         void* stream) {
  if (!p || !src) return fail(TFSM_E_INVALID, "null argument");
  if (!(src_batch == 1 || src_batch == p->B) || !(src_channels == 1 || src_channels == p->C))
    return fail(TFSM_E_INVALID, "source batch/channels must be 1 or match the state");
  if (p->src && (src_batch != p->srcB || src_channels != p->srcC))
    return fail(TFSM_E_INVALID, "all explicit sources of one operator must have the same shape");
  hipStream_t st = (hipStream_t)stream;
  const size_t total = (size_t)src_batch * src_channels * p->modes;
  float2* spec = nullptr;
  HIP_TRY(hipMalloc(&spec, total * sizeof(float2)), "source spectrum scratch");
  // forward transform of the source with a plan-compatible batch of src_batch blocks of src_channels
  const int saveC = p->C;
  p->C = src_channels;
  int rc = tfsm_forward(p, src, spec, src_batch, stream);
  p->C = saveC;
  const bool first = p->src == nullptr;
  if (!rc && first) rc = dev_alloc(p, &p->src, total, "source spectrum");
  if (!rc) {
    hipError_t e = launch_accum_source(p->src, spec, coef, total, first ? 1 : 0, st);
    if (e != hipSuccess) rc = hip_fail(e, "source accumulate launch");
  }
  (void)hipStreamSynchronize(st);
  (void)hipFree(spec);
  if (rc) return rc;
  p->srcB = src_batch;
  p->srcC = src_channels;
  if (p->D == 2) {
    // columns (last-axis frequencies) where the accumulated source has a nonzero mode: the fused
    // column kernels skip the source loads elsewhere (adding an exact zero changes nothing)
    std::vector<float2> h(total);
    HIP_TRY(hipMemcpy(h.data(), p->src, total * sizeof(float2), hipMemcpyDeviceToHost), "source spectrum download");
    std::vector<unsigned char> cols(p->P, 0);
    for (size_t e = 0; e < total; ++e)
      if (h[e].x != 0.f || h[e].y != 0.f) cols[e % p->P] = 1;
    if (!p->src_cols) rc = dev_alloc(p, &p->src_cols, (size_t)p->P, "source column flags");
    if (rc) return rc;
    HIP_TRY(hipMemcpy(p->src_cols, cols.data(), cols.size(), hipMemcpyHostToDevice), "source column flags upload");
    p->src_bits_ok = p->P <= 32 * Col2Args::SRC_BIT_WORDS;
    for (int w = 0; w < Col2Args::SRC_BIT_WORDS; ++w) p->src_bits[w] = 0;
    if (p->src_bits_ok)
      for (int k = 0; k < p->P; ++k)
        if (cols[k]) p->src_bits[k >> 5] |= 1u << (k & 31);
  }
  return TFSM_OK;
}

int tfsm_plan_add_force(tfsm_plan* p, const float* force, int32_t force_batch, int32_t force_channels, float coef,
                        void* stream) {
  if (!p || !force) return fail(TFSM_E_INVALID, "null argument");
  if (p->term != TERM_NSP)
    return fail(TFSM_E_INVALID, "an external force belongs to NSPressureConvection (velocity-form Navier-Stokes)");
  if (!(force_batch == 1 || force_batch == p->B) || !(force_channels == 1 || force_channels == p->C))
    return fail(TFSM_E_INVALID, "force batch/channels must be 1 or match the state");
  if (p->force && (force_batch != p->forceB || force_channels != p->forceC))
    return fail(TFSM_E_INVALID, "all explicit sources of one force must have the same shape");
  hipStream_t st = (hipStream_t)stream;
  const size_t total = (size_t)force_batch * force_channels * p->modes;
  float2* spec = nullptr;
  HIP_TRY(hipMalloc(&spec, total * sizeof(float2)), "force spectrum scratch");
  const int saveC = p->C;
  p->C = force_channels;
  int rc = tfsm_forward(p, force, spec, force_batch, stream);
  p->C = saveC;
  const bool first = p->force == nullptr;
  if (!rc && first) rc = dev_alloc(p, &p->force, total, "force spectrum");
  if (!rc) {
    hipError_t e = launch_accum_source(p->force, spec, coef, total, first ? 1 : 0, st);
    if (e != hipSuccess) rc = hip_fail(e, "force accumulate launch");
  }
  (void)hipStreamSynchronize(st);
  (void)hipFree(spec);
  if (rc) return rc;
  p->forceB = force_batch;
  p->forceC = force_channels;
  // the forced core masks the integrator's stage tensors in place: only the unfused path does that
  p->fused_ok = false;
  p->persistent = false;
  if (!p->general) {
    p->general = true;
    gen_layout(p);
  }
  return TFSM_OK;
}

int tfsm_step(tfsm_plan* p, void* uh, int32_t nsteps, const int32_t* rec_slot, void* traj, int32_t T, void* ws,
              void* stream) {
  if (!p || !uh) return fail(TFSM_E_INVALID, "null argument");
  if (p->scheme < 0) return fail(TFSM_E_STATE, "no scheme set on the plan");
  if (nsteps < 0) return fail(TFSM_E_INVALID, "nsteps must be >= 0");
  if (rec_slot && (!traj || T < 1)) return fail(TFSM_E_INVALID, "recording needs a trajectory buffer");
  return step_impl(p, uh, nsteps, rec_slot, traj, T, ws, (hipStream_t)stream);
}

int tfsm_step_noalloc(tfsm_plan* p, void* uh, int32_t nsteps, const int32_t* rec_slot, void* traj, int32_t T,
                      void* stream) {
  if (!p || !uh) return fail(TFSM_E_INVALID, "null argument");
  if (p->scheme < 0) return fail(TFSM_E_STATE, "no scheme set on the plan");
  if (nsteps < 0) return fail(TFSM_E_INVALID, "nsteps must be >= 0");
  if (rec_slot && (!traj || T < 1)) return fail(TFSM_E_INVALID, "recording needs a trajectory buffer");
  if (p->ws_bytes > (size_t)TFSM_NOALLOC_SCRATCH_BYTES)
    return fail(TFSM_E_INVALID, "the plan's workspace (" + std::to_string(p->ws_bytes) +
                                    " bytes) exceeds the static scratch of tfsm_step_noalloc (" +
                                    std::to_string((long long)TFSM_NOALLOC_SCRATCH_BYTES) +
                                    " bytes): use tfsm_step with a caller workspace or a smaller batch");
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev), "current device");
  if (dev != p->device) return fail(TFSM_E_INVALID, "the plan's device must be the current device");
  void* ws = nullptr;
  const int rc = scratch_on_device(dev, &ws);
  if (rc) return rc;
  return step_impl(p, uh, nsteps, rec_slot, traj, T, ws, (hipStream_t)stream);
}

int tfsm_step_varying(tfsm_plan* p, void* uh, const int32_t* nsteps_per_sample, int32_t max_steps, void* ws,
                      void* stream) {
  if (!p || !uh || !nsteps_per_sample || !ws) return fail(TFSM_E_INVALID, "null argument");
  if (p->scheme < 0) return fail(TFSM_E_STATE, "no scheme set on the plan");
  if (max_steps < 0) return fail(TFSM_E_INVALID, "max_steps must be >= 0");
  if (p->scheme == ETDRK0) return fail(TFSM_E_INVALID, "per-sample step counts need a nonlinear plan");
  if (p->term == TERM_KS && (p->D == 1 || p->ks_mean))
    return fail(TFSM_E_INVALID, "per-sample step counts are not available for KSConvection in 1D or with remove_mean "
                                "(its mean couples the samples)");
  if (p->D == 1) {
    Run1Args a = run1_args(p, (float2*)uh, (char*)ws);
    a.nsteps = max_steps;
    a.nsteps_per = nsteps_per_sample;
    HIP_TRY(launch_run1d(a, (hipStream_t)stream), "1D stepper launch");
    return TFSM_OK;
  }
  return step_impl(p, uh, max_steps, nullptr, nullptr, 0, ws, (hipStream_t)stream, nsteps_per_sample);
}

int tfsm_apply(tfsm_plan* p, const void* uh, void* out, void* ws, void* stream) {
  if (!p || !uh || !out) return fail(TFSM_E_INVALID, "null argument");
  hipStream_t st = (hipStream_t)stream;
  const bool have_nl = p->term != TERM_NONE;
  const bool have_lin = p->nlin > 0;
  float2* U = (float2*)uh;
  if (have_nl) {
    if (!ws) return fail(TFSM_E_INVALID, "workspace pointer is null");
    if (p->general) {
      if (int rc = gen_stage(p, U, U, (char*)ws, ST_EVAL, StepMask(), st)) return rc;
    } else if (p->D == 1 && p->term == TERM_KS) {
      Run1Args a = run1_args(p, U, (char*)ws);
      float2* Nr = (float2*)((char*)ws + p->off_N);
      float2* dc = p->ks_mean ? (float2*)((char*)ws + p->off_DC) : nullptr;
      HIP_TRY(launch_k1_eval_ks(a, U, Nr, st), "1D product launch");
      if (dc) HIP_TRY(launch_ks_dc(Nr, a.B, (size_t)p->P, 1, 0, dc, st), "batch-mean launch");
      HIP_TRY(launch_k1_stage(a, ST_EVAL, Nr, dc, (float2*)((char*)ws + p->off_S), st), "1D stage launch");
    } else if (p->D == 1) {
      Run1Args a = run1_args(p, U, (char*)ws);
      a.eval_only = 1;
      a.nsteps = 1;
      HIP_TRY(launch_run1d(a, st), "1D evaluation launch");
    } else {
      int rc = run_eval(p, U, (char*)ws, ST_PRO, 0, p->B, st);
      if (!rc) rc = run_eval(p, U, (char*)ws, ST_EVAL, 0, p->B, st);
      if (rc) return rc;
    }
  }
  const float2* Nv = have_nl ? (const float2*)((char*)ws + p->off_A1) : nullptr;
  if (p->cplx)
    HIP_TRY(launch_apply_linear_c(U, Nv, p->Lc, p->Bc > 1 ? p->modes : 0, p->modes, p->C, p->state_elems(), have_lin,
                                  have_nl, (float2*)out, st),
            "apply launch");
  else
    HIP_TRY(launch_apply_linear(U, Nv, p->L, p->Bc > 1 ? p->modes : 0, p->modes, p->C, p->state_elems(), have_lin,
                                have_nl, (float2*)out, st),
            "apply launch");
  return TFSM_OK;
}

int tfsm_spectral_map(tfsm_plan* p, int32_t op, float coef, const void* in, void* out, void* stream) {
  if (!p || !in || !out) return fail(TFSM_E_INVALID, "null argument");
  MapArgs a;
  a.op = op;
  a.D = p->D;
  a.B = p->B;
  switch (op) {
    case MAP_DEALIAS: a.cin = p->C; a.cout = p->C; break;
    case MAP_DIV: a.cin = p->D; a.cout = 1; break;
    case MAP_GRAD: a.cin = 1; a.cout = p->D; break;
    case MAP_CURL:
      if (p->D < 2) return fail(TFSM_E_INVALID, "curl operator only works for 2D or 3D vector field");
      a.cin = p->D; a.cout = p->D == 2 ? 1 : 3;
      break;
    case MAP_V2V:
      if (p->D != 2) return fail(TFSM_E_INVALID, "Only vorticity in 2Dmesh is supported");
      a.cin = 1; a.cout = 2;
      break;
    case MAP_PRESSURE: a.cin = p->D; a.cout = 1; break;
    case MAP_CONSDIV: a.cin = p->D * p->D; a.cout = p->D; break;
    default: return fail(TFSM_E_INVALID, "unknown spectral map");
  }
  for (int d = 0; d < 3; ++d) {
    a.n[d] = p->n[d];
    a.kv[d] = p->kv[d];
    a.kmax[d] = p->kmax[d];
  }
  if (p->kv_bstride) return fail(TFSM_E_INVALID, "spectral maps need shared lengths");
  a.H = p->H;
  a.P = p->P;
  a.modes = p->modes;
  a.coef = coef;
  a.in = (const float2*)in;
  a.out = (float2*)out;
  HIP_TRY(launch_spectral_map(a, (hipStream_t)stream), "spectral map launch");
  return TFSM_OK;
}

int tfsm_pointwise_products(tfsm_plan* p, const float* u, float* out, void* stream) {
  if (!p || !u || !out) return fail(TFSM_E_INVALID, "null argument");
  size_t npts = 1;
  for (int d = 0; d < p->D; ++d) npts *= p->n[d];
  HIP_TRY(launch_pointwise_products(u, out, p->B, p->C, npts, (hipStream_t)stream), "products launch");
  return TFSM_OK;
}

int tfsm_rk_combine(tfsm_plan* p, void* out, const void* x, const void* const* ks, const float* coefs, int32_t nk,
                    float dt, void* stream) {
  if (!p || !out || !x || (nk > 0 && (!ks || !coefs))) return fail(TFSM_E_INVALID, "null argument");
  if (nk < 0 || nk > RK_MAX_TERMS) return fail(TFSM_E_INVALID, "at most 8 stage terms");
  RkArgs a;
  a.x = (const float2*)x;
  for (int i = 0; i < RK_MAX_TERMS; ++i) {
    a.k[i] = i < nk ? (const float2*)ks[i] : nullptr;
    a.c[i] = i < nk ? coefs[i] : 0.f;
    if (i < nk && !ks[i]) return fail(TFSM_E_INVALID, "null stage pointer");
  }
  a.nk = nk;
  a.dt = dt;
  a.out = (float2*)out;
  a.total = p->state_elems();
  HIP_TRY(launch_rk_combine(a, (hipStream_t)stream), "rk combine launch");
  return TFSM_OK;
}

int tfsm_solve(tfsm_plan* p, const void* uh, void* out, void* stream) {
  if (!p || !uh || !out) return fail(TFSM_E_INVALID, "null argument");
  if (p->nlin == 0) return fail(TFSM_E_INVALID, "solve needs a linear operator");
  if (p->cplx)
    HIP_TRY(launch_apply_linear_c((const float2*)uh, nullptr, p->Lc, p->Bc > 1 ? p->modes : 0, p->modes, p->C,
                                  p->state_elems(), 2, 0, (float2*)out, (hipStream_t)stream),
            "solve launch");
  else
    HIP_TRY(launch_apply_linear((const float2*)uh, nullptr, p->L, p->Bc > 1 ? p->modes : 0, p->modes, p->C,
                                p->state_elems(), 2, 0, (float2*)out, (hipStream_t)stream),
            "solve launch");
  return TFSM_OK;
}

int tfsm_launch_pass(tfsm_plan* p, void* uh, void* ws, int32_t pass_id, int32_t stage, int32_t sub, void* stream) {
  if (!p || !uh || !ws) return fail(TFSM_E_INVALID, "null argument");
  if (p->general) return fail(TFSM_E_INVALID, "the plan steps through the unfused path: no fused passes");
  hipStream_t st = (hipStream_t)stream;
  // the launches cover sub-batch `sub`, exactly as tfsm_step launches them
  const int sb = (p->sub_batch > 0 && p->sub_batch < p->B) ? p->sub_batch : p->B;
  if (sub < 0 || (int64_t)sub * sb >= p->B) return fail(TFSM_E_INVALID, "sub-batch index out of range");
  const int b0 = sub * sb;
  const int nb = (p->B - b0) < sb ? (p->B - b0) : sb;
  if (p->D == 2) {
    if (p->term == TERM_NONE) return fail(TFSM_E_INVALID, "no product term");
    if (pass_id == 0) return launch_row(p, (char*)ws, b0, nb, st);
    if (pass_id == 1) {
      if (stage < ST_PRO || stage > ST_EVAL) return fail(TFSM_E_INVALID, "bad stage");
      if (stage != ST_PRO && p->scheme < 0) return fail(TFSM_E_STATE, "no scheme set on the plan");
      return launch_col(p, (float2*)uh, (char*)ws, stage, b0, nb, st);
    }
  }
  if (p->D == 3) {
    if (p->term == TERM_NONE) return fail(TFSM_E_INVALID, "no product term");
    if (stage < ST_PRO || stage > ST_EVAL) return fail(TFSM_E_INVALID, "bad stage");
    if (stage != ST_PRO && p->scheme < 0) return fail(TFSM_E_STATE, "no scheme set on the plan");
    // one whole evaluation (4 kernels) per call; pass ids 0/1 are not split for 3D
    if (pass_id == 3) return run3d_eval(p, (float2*)uh, (char*)ws, stage, b0, nb, st);
    if (pass_id >= 4 && pass_id <= 7)
      return run3d_eval(p, (float2*)uh, (char*)ws, stage, b0, nb, st, 1 << (pass_id - 4));
  }
  if (p->D == 1 && pass_id == 2) {
    if (p->scheme < 0) return fail(TFSM_E_STATE, "no scheme set on the plan");
    if (p->term == TERM_KS) return fail(TFSM_E_INVALID, "1D KSConvection steps stage by stage (tfsm_step)");
    Run1Args a = run1_args(p, (float2*)uh, (char*)ws);
    a.nsteps = 1;
    HIP_TRY(launch_run1d(a, st), "1D stepper launch");
    return TFSM_OK;
  }
  return fail(TFSM_E_INVALID, "unknown pass for this plan");
}

int tfsm_launch_pass_timed(tfsm_plan* p, void* uh, void* ws, int32_t pass_id, int32_t stage, int32_t sub,
                           void* ev_start, void* ev_stop, void* stream) {
  if (!ev_start || !ev_stop) return fail(TFSM_E_INVALID, "null event");
  g_launch_ev[0] = (hipEvent_t)ev_start;
  g_launch_ev[1] = (hipEvent_t)ev_stop;
  const int rc = tfsm_launch_pass(p, uh, ws, pass_id, stage, sub, stream);
  g_launch_ev[0] = g_launch_ev[1] = nullptr;
  return rc;
}

int64_t tfsm_plan_adjoint_workspace_bytes(const tfsm_plan* plan) {
  return plan ? (int64_t)adj_layout(plan).total : 0;
}

int tfsm_adjoint_prepare(tfsm_plan* p, int32_t grad_term, void* stream) {
  if (!p) return fail(TFSM_E_INVALID, "null plan");
  if (p->scheme < 0) return fail(TFSM_E_STATE, "no scheme set on the plan");
  if (p->scheme != ETDRK1 && p->scheme != ETDRK2 && p->scheme != SETDRK1 && p->scheme != SETDRK2 &&
      p->scheme != SETDRK4)
    return fail(TFSM_E_INVALID, "term not supported by engine: the adjoint covers ETDRK1/2 and SETDRK1/2/4");
  if (!(p->term == TERM_CONV || (p->term == TERM_VC && p->D == 2)))
    return fail(TFSM_E_INVALID,
                "term not supported by engine: the adjoint covers Convection (1D, 2D, 3D) and VorticityConvection");
  // explicit sources are constant: they enter the recomputed states but no Jacobian
  if (p->cplx) return fail(TFSM_E_INVALID, "term not supported by engine: the adjoint needs a real linear symbol");
  if (p->force) return fail(TFSM_E_INVALID, "term not supported by engine: the adjoint of a forced NSPressureConvection");
  if (p->per_len) return fail(TFSM_E_INVALID, "term not supported by engine: the adjoint needs shared lengths");
  if (grad_term < -1 || grad_term >= p->nlin) return fail(TFSM_E_INVALID, "grad_term out of range");
  hipStream_t st = (hipStream_t)stream;
  int rc = TFSM_OK;
  if (!p->dcoef) rc = dev_alloc(p, &p->dcoef, (size_t)p->Bc * ADJ_NDCOEF * p->modes, "adjoint coefficient table");
  if (!rc && !p->sym) rc = dev_alloc(p, &p->sym, p->modes, "adjoint symbol");
  if (rc) return rc;
  if (p->scheme == SETDRK4)
    HIP_TRY(launch_adj_dcoefs4(p->L, p->modes, p->Bc, p->dt, p->n_points, p->radius, p->dcoef, st),
            "adjoint coefficient launch");
  else
    HIP_TRY(launch_adj_dcoefs(p->L, p->modes, p->Bc, p->dt, p->dcoef, st), "adjoint coefficient launch");
  if (grad_term >= 0) HIP_TRY(launch_adj_sym(adj_args(p), p->lin_kind[grad_term], p->sym, st), "adjoint symbol launch");
  p->adj_term = grad_term;
  p->adj_scheme = p->scheme;
  return TFSM_OK;
}

int tfsm_adjoint_step(tfsm_plan* p, void* u, const void* ubar_next, void* ubar, double* grad, void* ws, void* wsa,
                      void* stream) {
  if (!p || !u || !ubar_next || !ubar || !ws || !wsa) return fail(TFSM_E_INVALID, "null argument");
  if (p->adj_term < -1 || p->adj_scheme != p->scheme) return fail(TFSM_E_STATE, "tfsm_adjoint_prepare not called");
  if (ubar == ubar_next || ubar == u) return fail(TFSM_E_INVALID, "ubar must not alias the inputs");
  hipStream_t st = (hipStream_t)stream;
  char* W = (char*)ws;
  char* A = (char*)wsa;
  const AdjLayout l = adj_layout(p);
  float2* U = (float2*)u;
  const float2* Ubn = (const float2*)ubar_next;
  float2* Ubo = (float2*)ubar;
  float2* A1 = (float2*)(W + p->off_A1);
  float2* Aa = (float2*)(A + l.Aa);
  float2* N0s = (float2*)(A + l.N0s);
  float2* T1 = (float2*)(A + l.T1);
  float2* Jv = (float2*)(A + l.Jv);
  float2* Ab = (float2*)(A + l.Ab);
  float2* Nb = (float2*)(A + l.Nb);
  float* gp = (grad && p->adj_term >= 0) ? (float*)(A + l.gpart) : nullptr;
  const size_t sbytes = p->state_elems() * sizeof(float2);
  if (p->scheme == SETDRK4) return adjoint_step4(p, U, Ubn, Ubo, grad, W, A, st);
  int rc = eval_nonlinear(p, U, W, st);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(N0s, A1, sbytes, hipMemcpyDeviceToDevice, st), "adjoint copy");
  AdjStageArgs s = stage_args(p, ADJ_PRE);
  if (s.two_stage) {
    s.op = ADJ_STAGE_A;
    s.X0 = U;
    s.X1 = N0s;
    s.Y0 = Aa;
    HIP_TRY(launch_adj_stage(s, st), "adjoint stage launch");
    rc = eval_nonlinear(p, Aa, W, st);  // Na -> A1
    if (rc) return rc;
    s = stage_args(p, ADJ_PRE);
    s.X0 = Ubn;
    s.Y0 = T1;
    HIP_TRY(launch_adj_stage(s, st), "adjoint stage launch");
    rc = adj_jt(p, Aa, T1, Jv, A, st);
    if (rc) return rc;
    s = stage_args(p, ADJ_MID);
    s.X0 = Ubn;
    s.X1 = Jv;
    s.X2 = A1;
    s.X3 = N0s;
    s.Y0 = Ab;
    s.Y1 = Nb;
    s.gpart = gp;
    HIP_TRY(launch_adj_stage(s, st), "adjoint stage launch");
    rc = adj_jt(p, U, Nb, Jv, A, st);
    if (rc) return rc;
    s = stage_args(p, ADJ_POST);
    s.X0 = Ab;
  } else {
    s.X0 = Ubn;
    s.Y0 = T1;
    HIP_TRY(launch_adj_stage(s, st), "adjoint stage launch");
    rc = adj_jt(p, U, T1, Jv, A, st);
    if (rc) return rc;
    s = stage_args(p, ADJ_POST);
    s.X0 = Ubn;
  }
  s.X1 = Jv;
  s.X2 = U;
  s.X3 = N0s;
  s.Y0 = Ubo;
  s.gpart = gp;
  HIP_TRY(launch_adj_stage(s, st), "adjoint stage launch");
  if (gp) HIP_TRY(launch_adj_reduce(gp, (size_t)p->C * p->modes, p->B, grad, st), "adjoint reduce launch");
  return TFSM_OK;
}

}  // extern "C"

namespace {
// nsteps of step_impl (optionally with per-sample counts) captured into an instantiated graph
int graph_capture(tfsm_plan* p, void* uh, int nsteps, const int* act, void* ws, void* stream, tfsm_graph** out) {
  *out = nullptr;
  hipStream_t cap = nullptr;
  HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "capture stream");
  // order the capture stream after the caller's stream so that setup work is visible
  hipEvent_t ev = nullptr;
  HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "capture event");
  HIP_TRY(hipEventRecord(ev, (hipStream_t)stream), "capture event record");
  HIP_TRY(hipStreamWaitEvent(cap, ev, 0), "capture wait");
  HIP_TRY(hipStreamSynchronize(cap), "capture sync");
  (void)hipEventDestroy(ev);
  tfsm_graph* g = new tfsm_graph();
  hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    delete g;
    (void)hipStreamDestroy(cap);
    return hip_fail(e, "begin capture");
  }
  int rc = step_impl(p, uh, nsteps, nullptr, nullptr, 0, ws, cap, act);
  e = hipStreamEndCapture(cap, &g->graph);
  if (rc == TFSM_OK && e != hipSuccess) rc = hip_fail(e, "end capture");
  if (rc == TFSM_OK) {
    e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) rc = hip_fail(e, "graph instantiate");
  }
  (void)hipStreamDestroy(cap);
  if (rc) {
    tfsm_graph_destroy(g);
    return rc;
  }
  *out = g;
  return TFSM_OK;
}
}  // namespace

extern "C" {

int tfsm_graph_create(tfsm_plan* p, void* uh, int32_t nsteps, void* ws, void* stream, tfsm_graph** out) {
  if (!p || !uh || !out) return fail(TFSM_E_INVALID, "null argument");
  if (p->scheme < 0) return fail(TFSM_E_STATE, "no scheme set on the plan");
  if (nsteps < 1) return fail(TFSM_E_INVALID, "a graph needs nsteps >= 1");
  return graph_capture(p, uh, nsteps, nullptr, ws, stream, out);
}

int tfsm_graph_create_varying(tfsm_plan* p, void* uh, const int32_t* nsteps_per_sample, int32_t max_steps, void* ws,
                              void* stream, tfsm_graph** out) {
  if (!p || !uh || !nsteps_per_sample || !ws || !out) return fail(TFSM_E_INVALID, "null argument");
  if (p->scheme < 0) return fail(TFSM_E_STATE, "no scheme set on the plan");
  if (max_steps < 1) return fail(TFSM_E_INVALID, "a graph needs max_steps >= 1");
  if (p->D < 2 || p->scheme == ETDRK0) return fail(TFSM_E_INVALID, "per-sample step graphs need a nonlinear 2D or 3D plan");
  if (p->term == TERM_KS && p->ks_mean)
    return fail(TFSM_E_INVALID, "per-sample step counts are not available for KSConvection with remove_mean "
                                "(its mean couples the samples)");
  return graph_capture(p, uh, max_steps, nsteps_per_sample, ws, stream, out);
}

int tfsm_graph_launch(tfsm_graph* g, void* stream) {
  if (!g || !g->exec) return fail(TFSM_E_INVALID, "null graph");
  HIP_TRY(hipGraphLaunch(g->exec, (hipStream_t)stream), "graph launch");
  return TFSM_OK;
}

int tfsm_graph_destroy(tfsm_graph* g) {
  if (!g) return TFSM_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  return TFSM_OK;
}

}  // extern "C"
