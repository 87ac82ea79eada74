// Shared definitions of the ETD engine: data layout, stage codes, kernel argument blocks and the
// per-mode stage combine.
//
// Data layout in HBM (all complex values are interleaved float2):
//   half spectrum  [B][C][N0]...[N_{D-2}][P]  last axis holds k = 0..N_{D-1}/2 (H values), P = pitch
//                  (H rounded up to a multiple of 16 so that column tiles start on 128-byte lines)
//   coefficient    [Bc][8][modes] float, one real table per slot (E, E2, c1..c6); Bc = 1 or B
//   2D fields      [B][NF][N0][PF] float2: inverse-x-transformed fields, ky <= kmax1 only
//   2D products    [B][NP][N0][P]  float2: forward-y-transformed physical products
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_core.hpp"

namespace tfsm {

// Kernel timing (tfsm_launch_pass_timed): when set, the next pass launch records these events at
// its kernel's start and end (hipExtLaunchKernelGGL, no launch gap inside the interval) and clears
// them; every other launch is a plain one.
extern thread_local hipEvent_t g_launch_ev[2];
#define TFSM_LAUNCH(K, G, B, SH, ST, ...)                                                  \
  do {                                                                                     \
    if (::tfsm::g_launch_ev[0]) {                                                           \
      hipEvent_t e0_ = ::tfsm::g_launch_ev[0], e1_ = ::tfsm::g_launch_ev[1];                \
      ::tfsm::g_launch_ev[0] = ::tfsm::g_launch_ev[1] = nullptr;                            \
      hipExtLaunchKernelGGL(K, G, B, SH, ST, e0_, e1_, 0, __VA_ARGS__);                     \
    } else {                                                                               \
      hipLaunchKernelGGL(K, G, B, SH, ST, __VA_ARGS__);                                     \
    }                                                                                      \
  } while (0)

enum TermKind : int { TERM_NONE = 0, TERM_CONV = 1, TERM_VC = 2, TERM_NSP = 3, TERM_KS = 4, TERM_CONS = 6 };

enum SchemeId : int { ETDRK0 = 0, ETDRK1 = 1, ETDRK2 = 2, SETDRK1 = 3, SETDRK2 = 4, SETDRK3 = 5, SETDRK4 = 6 };

// coefficient slots
enum CoefSlot : int { CS_E = 0, CS_E2 = 1, CS_C1 = 2, CS_C2 = 3, CS_C3 = 4, CS_C4 = 5, CS_C5 = 6, CS_C6 = 7, NCOEF = 8 };

// Stage codes of the fused combine.  s = the spectrum the next nonlinear evaluation needs.
enum Stage : int {
  ST_PRO = 0,   // s = U (fields of the current state only)
  ST_R1_E0,     // U' = E U + c1 N                                   (_etdrk.py:47-51, _setdrk_step.py:5-11)
  ST_R2_E0,     // A1 = N0;  s = E U + c1 N0                         (_setdrk_step.py:14-25)
  ST_R2_E1,     // U' = (E U + c1 A1) + c2 (N - A1)
  ST_R3_E0,     // A1 = N0;  s = E2 U + c1 N0                        (_setdrk_step.py:28-52)
  ST_R3_E1,     // A2 = Na;  s = E U + c2 (2 Na - A1)
  ST_R3_E2,     // U' = E U + c3 A1 + c4 A2 + c5 N
  ST_R4_E0,     // A1 = N0;  s = E2 U + c1 N0                        (_setdrk_step.py:55-82)
  ST_R4_E1,     // A2 = Na;  s = E2 U + c1 Na          (coef_2 is coef_1, _uncached.py:194)
  ST_R4_E2,     // a = E2 U + c1 A1; s = E2 a + c1 (2 N - A1); A2 = (E U + c4 A1) + (2 c5) (A2 + N)
  ST_R4_E3,     // U' = A2 + c6 N   (= ((E U + c4 N0) + (2 c5)(Na + Nb)) + c6 Nc, in that order)
  ST_EVAL,      // A1 = N (one nonlinear evaluation, no time step; Operator.__call__)
};

// number of nonlinear evaluations per step and the stage code of each
__host__ __device__ inline int scheme_evals(int scheme) {
  switch (scheme) {
    case ETDRK1: case SETDRK1: return 1;
    case ETDRK2: case SETDRK2: return 2;
    case SETDRK3: return 3;
    case SETDRK4: return 4;
    default: return 0;
  }
}
__host__ __device__ inline int scheme_stage(int scheme, int e) {
  switch (scheme) {
    case ETDRK1: case SETDRK1: return ST_R1_E0;
    case ETDRK2: case SETDRK2: return e == 0 ? ST_R2_E0 : ST_R2_E1;
    case SETDRK3: return ST_R3_E0 + e;
    case SETDRK4: return ST_R4_E0 + e;
    default: return ST_PRO;
  }
}

// ---------------------------------------------------------------- non-contracted FP32 arithmetic
// The reference evaluates every product and sum as a separate float32 op (torch elementwise
// kernels); the combine below keeps that rounding by forbidding FMA contraction.
// (HIP's __fmul_rn / __fadd_rn are plain operators and would still be fused into FMAs under the
// default -ffp-contract=fast; the pragma removes the contract flag from these operations.)
__device__ __forceinline__ float mulr(float a, float b) {
#pragma clang fp contract(off)
  return a * b;
}
__device__ __forceinline__ float addr(float a, float b) {
#pragma clang fp contract(off)
  return a + b;
}
__device__ __forceinline__ float subr(float a, float b) {
#pragma clang fp contract(off)
  return a - b;
}
__device__ __forceinline__ float2 rscale(float c, float2 z) { return make_float2(mulr(c, z.x), mulr(c, z.y)); }
__device__ __forceinline__ float2 radd(float2 a, float2 b) { return make_float2(addr(a.x, b.x), addr(a.y, b.y)); }
__device__ __forceinline__ float2 rsub(float2 a, float2 b) { return make_float2(subr(a.x, b.x), subr(a.y, b.y)); }
// complex64 product as torch computes it: (ac - bd) + (ad + bc) i, each product rounded
__device__ __forceinline__ float2 cmulr(float2 c, float2 z) {
  return make_float2(subr(mulr(c.x, z.x), mulr(c.y, z.y)), addr(mulr(c.x, z.y), mulr(c.y, z.x)));
}
// complex64 quotient a / b with Smith's scaling (c10::complex operator/)
__device__ __forceinline__ float2 cdivr(float2 a, float2 b) {
  const float ac = fabsf(b.x), ad = fabsf(b.y);
  if (ac >= ad) {
    const float rat = __fdiv_rn(b.y, b.x);
    const float scl = __fdiv_rn(1.0f, addr(b.x, mulr(b.y, rat)));
    return make_float2(mulr(addr(a.x, mulr(a.y, rat)), scl), mulr(subr(a.y, mulr(a.x, rat)), scl));
  }
  const float rat = __fdiv_rn(b.x, b.y);
  const float scl = __fdiv_rn(1.0f, addr(b.y, mulr(b.x, rat)));
  return make_float2(mulr(addr(mulr(a.x, rat), a.y), scl), mulr(subr(mulr(a.y, rat), a.x), scl));
}

// Per-mode stage combine.  Arrays are indexed by the same element index; coefficient pointer
// `cf` points at slot 0 of this sample's table, `cstride` = modes per table.  Returns s.
__device__ __forceinline__ float2 stage_combine(int stage, float2 nval, float2* __restrict__ U,
                                                float2* __restrict__ A1, float2* __restrict__ A2, size_t ei,
                                                const float* __restrict__ cf, size_t cstride, size_t ci,
                                                bool write) {
  auto C = [&](int slot) { return cf[(size_t)slot * cstride + ci]; };
  switch (stage) {
    case ST_R1_E0: {
      float2 u = U[ei];
      float2 un = radd(rscale(C(CS_E), u), rscale(C(CS_C1), nval));
      if (write) U[ei] = un;
      return un;
    }
    case ST_R2_E0: {
      float2 u = U[ei];
      if (write) A1[ei] = nval;
      return radd(rscale(C(CS_E), u), rscale(C(CS_C1), nval));
    }
    case ST_R2_E1: {
      float2 u = U[ei], n0 = A1[ei];
      float2 a = radd(rscale(C(CS_E), u), rscale(C(CS_C1), n0));
      float2 un = radd(a, rscale(C(CS_C2), rsub(nval, n0)));
      if (write) U[ei] = un;
      return un;
    }
    case ST_R3_E0: {
      float2 u = U[ei];
      if (write) A1[ei] = nval;
      return radd(rscale(C(CS_E2), u), rscale(C(CS_C1), nval));
    }
    case ST_R3_E1: {
      float2 u = U[ei], n0 = A1[ei];
      if (write) A2[ei] = nval;
      float2 two_na = make_float2(2.0f * nval.x, 2.0f * nval.y);
      return radd(rscale(C(CS_E), u), rscale(C(CS_C2), rsub(two_na, n0)));
    }
    case ST_R3_E2: {
      float2 u = U[ei], n0 = A1[ei], na = A2[ei];
      float2 un = radd(radd(radd(rscale(C(CS_E), u), rscale(C(CS_C3), n0)), rscale(C(CS_C4), na)),
                       rscale(C(CS_C5), nval));
      if (write) U[ei] = un;
      return un;
    }
    case ST_R4_E0: {
      float2 u = U[ei];
      if (write) A1[ei] = nval;
      return radd(rscale(C(CS_E2), u), rscale(C(CS_C1), nval));
    }
    case ST_R4_E1: {
      float2 u = U[ei];
      if (write) A2[ei] = nval;
      return radd(rscale(C(CS_E2), u), rscale(C(CS_C1), nval));
    }
    case ST_R4_E2: {
      float2 u = U[ei], n0 = A1[ei], na = A2[ei];
      const float e2 = C(CS_E2), c1 = C(CS_C1);
      float2 a = radd(rscale(e2, u), rscale(c1, n0));
      float2 two_nb = make_float2(2.0f * nval.x, 2.0f * nval.y);
      if (write) {
        const float c5x2 = 2.0f * C(CS_C5);
        A2[ei] = radd(radd(rscale(C(CS_E), u), rscale(C(CS_C4), n0)), rscale(c5x2, radd(na, nval)));
      }
      return radd(rscale(e2, a), rscale(c1, rsub(two_nb, n0)));
    }
    case ST_R4_E3: {
      float2 un = radd(A2[ei], rscale(C(CS_C6), nval));
      if (write) U[ei] = un;
      return un;
    }
    case ST_EVAL: {
      if (write) A1[ei] = nval;
      return nval;
    }
    default:
      return U[ei];
  }
}

// Complex-coefficient stage combine (complex-L plans: E, E2 complex; ETDRK c1, c2 complex; SETDRK
// c's real, stored with a zero imaginary part).  Same stage formulas and op order as
// stage_combine, with every coefficient product a complex64 product.
__device__ __forceinline__ float2 stage_combine_c(int stage, float2 nval, float2* __restrict__ U,
                                                  float2* __restrict__ A1, float2* __restrict__ A2, size_t ei,
                                                  const float2* __restrict__ cf, size_t cstride, size_t ci) {
  auto C = [&](int slot) { return cf[(size_t)slot * cstride + ci]; };
  switch (stage) {
    case ST_R1_E0: {
      const float2 un = radd(cmulr(C(CS_E), U[ei]), cmulr(C(CS_C1), nval));
      U[ei] = un;
      return un;
    }
    case ST_R2_E0:
      A1[ei] = nval;
      return radd(cmulr(C(CS_E), U[ei]), cmulr(C(CS_C1), nval));
    case ST_R2_E1: {
      const float2 n0 = A1[ei];
      const float2 a = radd(cmulr(C(CS_E), U[ei]), cmulr(C(CS_C1), n0));
      const float2 un = radd(a, cmulr(C(CS_C2), rsub(nval, n0)));
      U[ei] = un;
      return un;
    }
    case ST_R3_E0:
      A1[ei] = nval;
      return radd(cmulr(C(CS_E2), U[ei]), cmulr(C(CS_C1), nval));
    case ST_R3_E1: {
      const float2 n0 = A1[ei];
      A2[ei] = nval;
      const float2 two_na = make_float2(2.0f * nval.x, 2.0f * nval.y);
      return radd(cmulr(C(CS_E), U[ei]), cmulr(C(CS_C2), rsub(two_na, n0)));
    }
    case ST_R3_E2: {
      const float2 un = radd(radd(radd(cmulr(C(CS_E), U[ei]), cmulr(C(CS_C3), A1[ei])), cmulr(C(CS_C4), A2[ei])),
                             cmulr(C(CS_C5), nval));
      U[ei] = un;
      return un;
    }
    case ST_R4_E0:
      A1[ei] = nval;
      return radd(cmulr(C(CS_E2), U[ei]), cmulr(C(CS_C1), nval));
    case ST_R4_E1:
      A2[ei] = nval;
      return radd(cmulr(C(CS_E2), U[ei]), cmulr(C(CS_C1), nval));
    case ST_R4_E2: {
      const float2 e2 = C(CS_E2), c1 = C(CS_C1), u = U[ei], n0 = A1[ei], na = A2[ei];
      const float2 a = radd(cmulr(e2, u), cmulr(c1, n0));
      const float2 two_nb = make_float2(2.0f * nval.x, 2.0f * nval.y);
      const float2 c5 = C(CS_C5);
      const float2 c5x2 = make_float2(2.0f * c5.x, 2.0f * c5.y);
      A2[ei] = radd(radd(cmulr(C(CS_E), u), cmulr(C(CS_C4), n0)), cmulr(c5x2, radd(na, nval)));
      return radd(cmulr(e2, a), cmulr(c1, rsub(two_nb, n0)));
    }
    case ST_R4_E3: {
      const float2 un = radd(A2[ei], cmulr(C(CS_C6), nval));
      U[ei] = un;
      return un;
    }
    case ST_EVAL:
      A1[ei] = nval;
      return nval;
    default:
      return U[ei];
  }
}

// ---------------------------------------------------------------- packed coefficients
// Per mode, the coefficients a scheme's stages need are packed into 8 floats (two float4 halves)
// so that each stage reads them with one 16-byte load:
//   ETDRK1/SETDRK1: [E, c1, -, - | -]          ETDRK2/SETDRK2: [E, c1, c2, - | -]
//   SETDRK3:        [E2, c1, E, c2 | E, c3, c4, c5]
//   SETDRK4:        [E2, c1, -, - | E, c4, c5, c6]
// Within a table the halves are interleaved in blocks of 16 modes: [block][half][16] float4, so a
// column tile's 16 consecutive modes of one stage read 256 contiguous bytes (tables hold a multiple
// of 16 modes: the pitch is).
__host__ __device__ __forceinline__ size_t cpk_index(size_t mode, int half) {
  return 2 * mode - (mode & 15) + ((size_t)half << 4);
}
__host__ __device__ inline int stage_coef_half(int stage) {
  return (stage == ST_R3_E2 || stage == ST_R4_E3) ? 1 : 0;
}
__host__ __device__ inline bool stage_reads_a1(int stage) {
  return stage == ST_R2_E1 || stage == ST_R3_E1 || stage == ST_R3_E2 || stage == ST_R4_E2;
}
// the state U: every stage but the last of SETDRK4 (whose partial sum is in A2) and ST_EVAL
__host__ __device__ inline bool stage_reads_u(int stage) { return stage != ST_R4_E3 && stage != ST_EVAL; }
__host__ __device__ inline bool stage_reads_a2(int stage) {
  return stage == ST_R3_E2 || stage == ST_R4_E2 || stage == ST_R4_E3;
}
// 0: the stage writes U, 1: A1, 2: A2
__host__ __device__ inline int stage_target(int stage) {
  switch (stage) {
    case ST_R2_E0: case ST_R3_E0: case ST_R4_E0: case ST_EVAL: return 1;
    case ST_R3_E1: case ST_R4_E1: case ST_R4_E2: return 2;
    default: return 0;
  }
}

// Stage combine on values already in registers (same arithmetic and rounding as stage_combine).
// Returns s; *w is the value the stage stores to its target array (stage_target).
__device__ __forceinline__ float2 stage_combine_v(int stage, float2 nval, float2 u, float2 a1, float2 a2, float4 c,
                                                  float c5, float2* w) {
  switch (stage) {
    case ST_R1_E0: {  // c = [E, c1, ...]
      float2 un = radd(rscale(c.x, u), rscale(c.y, nval));
      *w = un;
      return un;
    }
    case ST_R2_E0:  // [E, c1, c2]
      *w = nval;
      return radd(rscale(c.x, u), rscale(c.y, nval));
    case ST_R2_E1: {
      float2 a = radd(rscale(c.x, u), rscale(c.y, a1));
      float2 un = radd(a, rscale(c.z, rsub(nval, a1)));
      *w = un;
      return un;
    }
    case ST_R3_E0:  // half 0: [E2, c1, E, c2]
      *w = nval;
      return radd(rscale(c.x, u), rscale(c.y, nval));
    case ST_R3_E1: {
      *w = nval;
      float2 two_na = make_float2(2.0f * nval.x, 2.0f * nval.y);
      return radd(rscale(c.z, u), rscale(c.w, rsub(two_na, a1)));
    }
    case ST_R3_E2: {  // half 1: [E, c3, c4, c5]
      float2 un = radd(radd(radd(rscale(c.x, u), rscale(c.y, a1)), rscale(c.z, a2)), rscale(c.w, nval));
      *w = un;
      return un;
    }
    case ST_R4_E0:  // half 0: [E2, c1]
    case ST_R4_E1:
      *w = nval;
      return radd(rscale(c.x, u), rscale(c.y, nval));
    case ST_R4_E2: {  // half 0: [E2, c1, E, c4]; c5 from half 1
      float2 a = radd(rscale(c.x, u), rscale(c.y, a1));
      float2 two_nb = make_float2(2.0f * nval.x, 2.0f * nval.y);
      *w = radd(radd(rscale(c.z, u), rscale(c.w, a1)), rscale(2.0f * c5, radd(a2, nval)));
      return radd(rscale(c.x, a), rscale(c.y, rsub(two_nb, a1)));
    }
    case ST_R4_E3: {  // half 1: [E, c4, c5, c6]
      float2 un = radd(a2, rscale(c.w, nval));
      *w = un;
      return un;
    }
    case ST_EVAL:
      *w = nval;
      return nval;
    default:
      *w = u;
      return u;
  }
}

// Stage combine of one channel's column of RMAX modes per lane into s, storing the stage target
// (the column kernels, 2D and 3D).  Mode m of lane j is element (j + m NT) * stride of the column
// whose first element U, A1, A2, src point at; cpk points at the packed coefficient pair of that
// first element, so mode m's is cpk[2 (j + m NT) stride]; that element's c5 is c5b[c5o].  Operands are loaded in batches of BM
// modes and every load of a batch is issued before its first use, so a batch costs one memory
// latency instead of one per operand.  BM defaults to 4 for plans of up to 12 values per thread and
// to 1 above, where the transforms already hold most of the 128 VGPRs that keep two 512-thread
// workgroups per CU.  The loads are unconditional: a lane past the last column (colv false) reads the padded
// columns of the same rows, which every array and table holds, and its result is discarded.
// Offsets within one channel are 32-bit (every supported mesh has fewer than 2^28 modes).
// SETDRK4's first, second and last evaluations need 2, 2 and 1 of the packed entry's floats: with
// k01b ([E2, c1] per mode) and c6b (c6 per mode), tables laid out like c5b, those stages read 8 and
// 4 bytes per mode instead of the 16-byte entry.
// PRE (kernels compiled for one SETDRK4 stage): operands the caller loaded before its forward
// transform, so that their memory latency overlaps the transform.  PRE = 1 (first and second
// evaluations): pre.op holds the state U and, when k01b is set, pre.k01 this lane's k01b entries;
// PRE = 2 (last evaluation): pre.op holds A2 and, when c6b is set, pre.c6 this lane's c6b entries.
template <int RMAX>
struct ColPre {
  float2 op[RMAX];
  float2 k01[RMAX];
  float c6[RMAX];
};

template <int RMAX, int NT, int BM = (RMAX >= 16 ? 1 : 4), bool FENCE = false, int PRE = 0>
__device__ __forceinline__ void combine_column(int stage, float2* U, float2* A1, float2* A2, const float2* src,
                                               const float4* cpk, const float* c5b, unsigned c5o, float coef_t,
                                               unsigned stride, int j, bool colv, const float2 (&v)[RMAX],
                                               float2 (&s)[RMAX], const float2* k01b = nullptr,
                                               const float* c6b = nullptr, const ColPre<RMAX>* pre = nullptr) {
  static_assert(RMAX % BM == 0, "combine batch must divide the values per thread");
  const float2 z = make_float2(0.f, 0.f);
  const bool rdu = stage_reads_u(stage), rd1 = stage_reads_a1(stage), rd2 = stage_reads_a2(stage);
  const bool rd5 = stage == ST_R4_E2;  // c5, from its own table
  const bool rk01 = k01b && (stage == ST_R4_E0 || stage == ST_R4_E1);
  const bool rk6 = c6b && stage == ST_R4_E3;
  const int tgt = stage_target(stage);
  float2* dst = tgt == 0 ? U : (tgt == 1 ? A1 : A2);
  const unsigned o0 = (unsigned)j * stride, ostep = (unsigned)NT * stride;
#pragma unroll
  for (int m0 = 0; m0 < RMAX; m0 += BM) {
    // FENCE (kernels compiled for one stage): keep each batch's loads in their batch; with the stage
    // known at compile time the scheduler would otherwise issue every mode's operand loads at once,
    // far beyond the register budget
    if constexpr (FENCE) __builtin_amdgcn_sched_barrier(0);
    float2 uu[BM], q1[BM], q2[BM], sv[BM];
    float4 cc[BM];
    float c5[BM];
#pragma unroll
    for (int i = 0; i < BM; ++i) {
      const unsigned o = o0 + (unsigned)(m0 + i) * ostep;
      if (rk01) {
        const float2 t = PRE == 1 ? pre->k01[m0 + i] : k01b[c5o + o];
        cc[i] = make_float4(t.x, t.y, 0.f, 0.f);
      } else if (rk6) {
        cc[i] = make_float4(0.f, 0.f, 0.f, PRE == 2 ? pre->c6[m0 + i] : c6b[c5o + o]);
      } else {
        cc[i] = cpk[2 * o];
      }
      uu[i] = z;
      q1[i] = z;
      q2[i] = z;
      sv[i] = z;
      c5[i] = 0.f;
    }
    if (rdu) {
#pragma unroll
      for (int i = 0; i < BM; ++i) uu[i] = PRE == 1 ? pre->op[m0 + i] : U[o0 + (unsigned)(m0 + i) * ostep];
    }
    if (rd1) {
#pragma unroll
      for (int i = 0; i < BM; ++i) q1[i] = A1[o0 + (unsigned)(m0 + i) * ostep];
    }
    if (rd2) {
#pragma unroll
      for (int i = 0; i < BM; ++i) q2[i] = PRE == 2 ? pre->op[m0 + i] : A2[o0 + (unsigned)(m0 + i) * ostep];
    }
    if (src) {
#pragma unroll
      for (int i = 0; i < BM; ++i) sv[i] = src[o0 + (unsigned)(m0 + i) * ostep];
    }
    // last: the compiler forms 2 c5 right after this load, and the wait for it must not hold back
    // the loads above
    if (rd5) {
#pragma unroll
      for (int i = 0; i < BM; ++i) c5[i] = c5b[c5o + o0 + (unsigned)(m0 + i) * ostep];
    }
#pragma unroll
    for (int i = 0; i < BM; ++i) {
      const int m = m0 + i;
      float2 nv = rscale(coef_t, v[m]);
      if (src) nv = radd(nv, sv[i]);
      float2 w;
      const float2 r = stage_combine_v(stage, nv, uu[i], q1[i], q2[i], cc[i], c5[i], &w);
      s[m] = colv ? r : z;
      if (colv) dst[o0 + (unsigned)m * ostep] = w;
    }
  }
}

// NSPressureConvection of one mode from the convection spectra conv_c (_navier_stokes.py:244-262),
// in the reference's op order: t_i = (i a_i) conv_i, p = invlap * sum_i t_i,
// N_c = (i a_c) p - conv_c, with invlap = 1 / lap (1 at lap = 0), lap = sum_i -(a_i a_i).
template <int C>
__device__ __forceinline__ void nsp_project(float2 (&cv)[C], const float (&av)[C]) {
  float lap = -mulr(av[0], av[0]);
#pragma unroll
  for (int i = 1; i < C; ++i) lap = addr(lap, -mulr(av[i], av[i]));
  const float il = (lap == 0.f) ? 1.0f : __fdiv_rn(1.0f, lap);
  float2 sum = make_float2(-mulr(av[0], cv[0].y), mulr(av[0], cv[0].x));
#pragma unroll
  for (int i = 1; i < C; ++i) sum = radd(sum, make_float2(-mulr(av[i], cv[i].y), mulr(av[i], cv[i].x)));
  const float2 p = make_float2(mulr(il, sum.x), mulr(il, sum.y));
#pragma unroll
  for (int c = 0; c < C; ++c) cv[c] = rsub(make_float2(-mulr(av[c], p.y), mulr(av[c], p.x)), cv[c]);
}

// Hermitian part on Nyquist planes.  The reference steps the full complex spectrum.  Where a
// derivative is applied after a product's FFT (the projection of NSPressureConvection,
// ConservativeConvection), its value on the Nyquist plane of an even axis d (index n_d / 2, where
// a_d is even in k instead of odd) is not Hermitian, and its real output keeps only the Hermitian
// part, which the half spectrum stores: a product of symbols survives there only when it has an
// even number of factors a_d of Nyquist axes.  ny[d] flags the Nyquist axes of a mode; callers
// clear every flag at modes that are their own conjugate partner (every index 0 or Nyquist), whose
// full value the half spectrum holds.
template <int C>
__device__ __forceinline__ void nsp_project(float2 (&cv)[C], const float (&av)[C], const bool (&ny)[C]) {
  float lap = -mulr(av[0], av[0]);
#pragma unroll
  for (int i = 1; i < C; ++i) lap = addr(lap, -mulr(av[i], av[i]));
  const float il = (lap == 0.f) ? 1.0f : __fdiv_rn(1.0f, lap);
  float2 sum[2] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
  bool have[2] = {false, false};
#pragma unroll
  for (int i = 0; i < C; ++i) {
    const float2 t = make_float2(-mulr(av[i], cv[i].y), mulr(av[i], cv[i].x));
    const int q = ny[i] ? 1 : 0;
    sum[q] = have[q] ? radd(sum[q], t) : t;
    have[q] = true;
  }
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int q = ny[c] ? 1 : 0;
    const float2 p = make_float2(mulr(il, sum[q].x), mulr(il, sum[q].y));
    cv[c] = rsub(make_float2(-mulr(av[c], p.y), mulr(av[c], p.x)), cv[c]);
  }
}

// Nonlinear value of one mode: N = coef_t * F + S   (S = sum of coef_s * source spectra)
__device__ __forceinline__ float2 nl_value(float coef_t, float2 f, const float2* __restrict__ src, size_t si) {
  float2 n = rscale(coef_t, f);
  if (src) n = radd(n, src[si]);
  return n;
}

// ---------------------------------------------------------------- kernel argument blocks
// Per-sample step counts (continuous batching of 2D / 3D requests): a launch belongs to step `it`
// of its tfsm_step_varying call, and sample b takes part only while it < act[b]; its workgroups
// return before touching memory otherwise.  act = null: every sample steps.
struct StepMask {
  const int* act = nullptr;
  int it = 0;
  __device__ __forceinline__ bool skip(int b) const { return act && it >= act[b]; }
};

struct Col2Args {
  int B, C, N0, N1, H1, P1, PF;
  int kmax0, kmax1;
  int NF;                     // number of field arrays
  float coef_t;               // coefficient of the product term
  float scale;                // 1/(N0 N1)
  const float2* tw0;          // twiddles for N0
  const float* ax;            // axis-0 wavenumbers (2 pi f), float32
  const float* ay;            // axis-1 wavenumbers
  const float2* P;            // products [B][NP][N0][P1]
  float2* U;                  // state [B][C][N0][P1]
  float2* A1;
  float2* A2;
  const float* coef;          // [Bc][NCOEF][N0*P1]
  size_t coef_bstride;        // 0 or NCOEF*N0*P1
  const float4* cpk;          // packed coefficients [Bc][N0*P1][2]
  size_t cpk_bstride;         // 0 or 2*N0*P1 (float4 units)
  const float* c5t;           // SETDRK4's c5 [Bc][N0*P1] (batch stride cpk_bstride / 2)
  const float2* ck01 = nullptr;  // SETDRK4's [E2, c1] per mode (same layout as c5t), or null
  const float* c6t = nullptr;    // SETDRK4's c6 per mode, or null
  const float2* src;          // [Bs][Cs][N0][P1] or null
  size_t src_bstride, src_cstride;
  const unsigned char* src_cols = nullptr;  // [P1]: 1 where some source mode of column ky is nonzero
  // the same flags as bits (column ky: bit ky & 31 of word ky >> 5), when src_bits_ok: a tile of 8
  // or 16 columns reads its flags from the kernel arguments instead of memory
  static constexpr int SRC_BIT_WORDS = 17;  // P1 <= 544
  uint32_t src_bits[SRC_BIT_WORDS] = {};
  int src_bits_ok = 0;
  float2* G;                  // fields [B][NF][N0][PF]
  const float2* dc = nullptr;  // KSConvection: batch mean of the product DC (subtracted at k = 0), or null
  int b0;                     // first sample handled by this launch
  int nb;                     // samples handled by this launch (set by the launcher)
  StepMask mask;
};

struct Row2Args {
  int B, N0, N1, H1, P1, PF, kmax1, NF;
  const float2* tw1;
  const float* ay;
  const float2* G;
  float2* P;
  int b0;
  StepMask mask;
};

// persistent per-sample 2D stepper (k2_run): samples [b0, b0 + nb), one workgroup each
struct Run2Args {
  int b0, nb;
  int scheme, nsteps;
  const int* act;           // per-sample step counts (device) or null
  const int* rec_slot;      // device [nsteps + 1]: frame recorded before step i, or -1; or null
  float2* traj;             // [B][T][per_sample]
  int T;
  size_t per_sample;        // C * modes
};

struct Col3Args {
  int B, N0, N1, N2, H2, P2, PF2, K1;
  int kmax0, kmax1, kmax2;
  int term;                   // TERM_CONV, TERM_NSP or TERM_KS
  int nc;                     // channels: 3 (Convection, NSP) or 1 (KSConvection)
  float coef_t, scale;
  const float2* tw0;
  const float* ax;
  const float* ay;
  const float* az;
  const float2* P;            // y-forward products [B][3][N0][N1][P2]
  float2* U;
  float2* A1;
  float2* A2;
  const float4* cpk;
  size_t cpk_bstride;
  const float* c5t;           // SETDRK4's c5 [Bc][modes] (batch stride cpk_bstride / 2)
  const float2* ck01 = nullptr;  // SETDRK4's [E2, c1] per mode (same layout as c5t), or null
  const float* c6t = nullptr;    // SETDRK4's c6 per mode, or null
  const float2* src;
  size_t src_bstride, src_cstride;
  float2* G1;                 // [B][2 nc][N0][K1][PF2]
  const float2* dc = nullptr;  // KSConvection: batch mean of the product DC, or null
  int b0, nb;
  StepMask mask;
};

struct Yinv3Args {
  int N0, N1, PF2, K1, kmax1, kmax2;
  int nc;                     // channels (fields 2 nc -> 3 nc)
  const float2* tw1;
  const float* ay;
  const float2* G1;
  float2* G2;                 // [B][3 nc][N0][N1][PF2]
  int b0, nb;
  StepMask mask;
};

struct Row3Args {
  int term;                   // TERM_CONV / TERM_NSP (Convection products) or TERM_KS
  int N0, N1, N2, P2, PF2, kmax2;
  const float2* tw2;
  const float* az;
  const float2* G2;
  float2* P;                  // products [B][nc][N0][N1][P2]
  int b0, nb;
  StepMask mask;
};

struct Run1Args {
  int B, N, H, P, kmax;
  float coef_t, scale;
  int scheme;
  const float2* tw;
  const float* ax;            // [Bk][N] axis wavenumbers
  size_t ax_bstride;
  float2* U;
  float2* A1;
  float2* A2;
  const float* coef;
  size_t coef_bstride;
  const float2* coefc = nullptr;  // complex-L plans: complex coefficient tables (same strides), else null
  const float2* src;
  size_t src_bstride;
  int nsteps;
  const int* nsteps_per;      // optional per-sample step counts (device), null = nsteps for all
  const int* rec_slot;        // device [nsteps+1]: frame index recorded before step i, or -1
  float2* traj;               // [B][T][P] recorded half spectra
  int T;
  int eval_only;              // 1: one nonlinear evaluation into A1, no stepping
  int cons = 0;               // 1: ConservativeConvection (N = i a fft(u u)), else Convection
};


enum LinKind : int { LIN_LAP = 1, LIN_BIHARM = 2, LIN_UNIT = 3, LIN_DERIV = 4 };

struct LinArgs {
  int D;
  int n[3];
  int H, P;
  size_t modes;            // prod(n[:-1]) * P
  const float* kv[3];      // axis wavenumbers a_d = fl32(fl32(2 pi) * f_d)
  size_t kv_bstride;       // per-sample axis tables (1D serving); 0 = shared
  int nterm;
  int kind[4];
  int dim[4] = {0, 0, 0, 0}, order[4] = {0, 0, 0, 0};  // LIN_DERIV parameters
  const float* coef;       // [nterm][Bc]; null: use cval (one row, coefficients by value)
  float cval[4];
  int Bc;
  float* L;                // [Bc][modes] real part
  float2* Lc = nullptr;    // [Bc][modes] complex L (complex-L plans), or null
};

struct CoefArgs {
  float4* cpk;             // packed coefficients [Bc][modes][2], written after the slot tables
  float* c5t;              // SETDRK4's c5 alone [Bc][modes] (the third evaluation's fifth coefficient)
  float2* ck01 = nullptr;  // SETDRK4's [E2, c1] [Bc][modes] (first and second evaluations)
  float* c6t = nullptr;    // SETDRK4's c6 [Bc][modes] (last evaluation)
  int scheme;
  double dt;
  float dt32, hdt32;       // fl32(dt), fl32(0.5 * dt)
  int M;                   // contour points
  double radius;
  size_t modes;
  int Bc;
  const float* L;          // [Bc][modes]
  float* coef;             // [Bc][NCOEF][modes]
  const float2* Lc = nullptr;  // complex-L plans: [Bc][modes], and the complex tables below
  float2* coefc = nullptr;     // [Bc][NCOEF][modes]
};

struct ConvArgs {
  int D;
  int n[3];
  int H, P;
  size_t nblk;  // number of [N...] blocks (batch * channels)
};

// ---------------------------------------------------------------- host launchers (defined in *.hip)
hipError_t launch_col2d(int term, const Col2Args& a, int nb, int stage, int v, int var, hipStream_t st);
hipError_t launch_row2d(int term, const Row2Args& a, int var, int wgt, hipStream_t st);
bool col2d_supported(int n0, int v, int var);
// persistent per-sample stepper (k2_run): square 32^2 and 64^2 meshes, Convection or
// VorticityConvection, field pitch pf at most that of the 2/3 rule
hipError_t launch_run2d(int term, const Col2Args& ca, const Row2Args& ra, const Run2Args& r, hipStream_t st);
bool run2d_supported(int term, int n0, int n1, int pf);
bool row2d_supported(int n1, int var, int wgt);
hipError_t launch_rows_r2c(int n, const float* in, size_t in_ld, float2* out, size_t out_ld, int nrows, float scale,
                           const float2* tw, hipStream_t st);
hipError_t launch_rows_c2r(int n, const float2* in, size_t in_ld, float* out, size_t out_ld, int nrows, float scale,
                           const float2* tw, hipStream_t st);
hipError_t launch_rows_c2c(int n, int dir, const float2* in, size_t in_ld, float2* out, size_t out_ld, int nrows,
                           float scale, const float2* tw, hipStream_t st);
// with a mask, outer block o of the launch belongs to sample b0 + o / outer_per_sample
hipError_t launch_cols_c2c(int n, int dir, const float2* in, float2* out, size_t nouter, size_t outer_stride,
                           size_t col_stride, int ninner, float scale, const float2* tw, hipStream_t st,
                           StepMask mask = StepMask(), int b0 = 0, size_t outer_per_sample = 1);
hipError_t launch_full_to_half(const float2* full, float2* half, const ConvArgs& a, hipStream_t st);
hipError_t launch_half_to_full(const float2* half, float2* full, const ConvArgs& a, hipStream_t st);
hipError_t launch_linear_steps(float2* U, const float* E, size_t coef_bstride, size_t modes, int C, size_t total,
                               int nsteps, const int* rec_slot, float2* traj, int T, hipStream_t st);
hipError_t launch_linear_steps_c(float2* U, const float2* E, size_t coef_bstride, size_t modes, int C, size_t total,
                                 int nsteps, const int* rec_slot, float2* traj, int T, hipStream_t st);
hipError_t launch_record(const float2* U, float2* traj, const int* rec_slot, int step, int B, size_t per_sample,
                         int T, hipStream_t st);
hipError_t launch_run1d(const Run1Args& a, hipStream_t st);
// KSConvection in 1D, one launch per stage: the raw product spectrum of stage spectrum S, then the
// stage combine (with the batch-mean correction of the DC mode) into S, U, A1, A2
hipError_t launch_k1_eval_ks(const Run1Args& a, const float2* S, float2* Nraw, hipStream_t st);
hipError_t launch_k1_stage(const Run1Args& a, int stage, const float2* Nraw, const float2* dc, float2* S,
                           hipStream_t st);
// out = (1/B) sum_b sum_x P[b * bstride + x * xstride]: the DC of the product spectra averaged over
// the batch (KSConvection's remove_mean subtracts the mean over batch and space)
hipError_t launch_ks_dc(const float2* P, int B, size_t bstride, int nx, size_t xstride, float2* out, hipStream_t st);
hipError_t launch_col3d(const Col3Args& a, int nb, int stage, int col_var, hipStream_t st);
hipError_t launch_yinv3d(const Yinv3Args& a, int col_var, hipStream_t st);
hipError_t launch_row3d(const Row3Args& a, int row_var, hipStream_t st);
bool step3d_supported(int n0, int n1, int n2);
// radix-plan variant used by the 3D kernels for an axis of length n
inline int var3(int n) { return n == 128 ? 2 : 0; }
hipError_t launch_linear_symbol(const LinArgs& a, hipStream_t st);
hipError_t launch_axis_row(float* kv, int n, float s, hipStream_t st);
hipError_t launch_coefs(const CoefArgs& a, hipStream_t st);
hipError_t launch_pack_coefs(const CoefArgs& a, hipStream_t st);
hipError_t launch_accum_source(float2* S, const float2* X, float coef, size_t total, int first, hipStream_t st);
// ---------------------------------------------------------------- adjoint (k_adjoint.hip)
struct AdjArgs {
  int term = TERM_CONV;     // TERM_CONV or TERM_VC
  int D, C, B;
  int n[3];
  int H, P;
  size_t modes;             // prod(n[:-1]) * P
  size_t npts;              // prod(n)
  int kmax[3];
  const float* kv[3];
  float coef_t;
};
enum AdjOp : int { ADJ_STAGE_A = 0, ADJ_PRE = 1, ADJ_MID = 2, ADJ_POST = 3 };
struct AdjStageArgs {
  int op, two_stage;
  int B, C, P, H, nlast;
  size_t modes;
  const float* coef;        // [Bc][NCOEF][modes]
  size_t coef_bstride;
  const float* dcoef;       // [Bc][3][modes]: dE/dL, dc1/dL, dc2/dL
  size_t dcoef_bstride;
  const float* sym;         // dL/dtheta [modes]
  const float2* X0;
  const float2* X1;
  const float2* X2;
  const float2* X3;
  float2* Y0;
  float2* Y1;
  float* gpart;             // [B][C][modes] per-mode gradient terms, or null
};
constexpr int ADJ_NDCOEF = 6;  // coefficient-derivative slots: 2-stage dE, dc1, dc2; SETDRK4 dE, dE2, dc1, dc4, dc5, dc6
constexpr int ADJ_MAXT = 5;
struct AdjLinArgs {
  int B, C, P, H;
  size_t modes;
  const float* coef;        // [Bc][NCOEF][modes]
  size_t coef_bstride;
  int nt;
  const float2* X[ADJ_MAXT];
  int slot[ADJ_MAXT];       // coefficient slot (CoefSlot) or -1 for 1
  float s[ADJ_MAXT];
  float2* Y;
};
struct AdjGradArgs {
  int B, C, P, H, nlast;
  size_t modes;
  const float* dcoef;       // [Bc][ADJ_NDCOEF][modes]
  size_t dcoef_bstride;
  const float* sym;         // dL/dtheta [modes]
  const float2* Xbar;
  int nt;
  const float2* Y[ADJ_MAXT];
  int dslot[ADJ_MAXT];
  float s[ADJ_MAXT];
  float* gpart;
  int accumulate;
};
hipError_t launch_adj_fields(const AdjArgs& a, const float2* U, const float2* V, float2* F, hipStream_t st);
hipError_t launch_adj_dcoefs4(const float* L, size_t modes, int Bc, double dt, int M, double r, float* dcoef,
                              hipStream_t st);
hipError_t launch_adj_lincomb(const AdjLinArgs& a, hipStream_t st);
hipError_t launch_adj_gacc(const AdjGradArgs& a, hipStream_t st);
hipError_t launch_adj_products(const AdjArgs& a, const float* X, float* Q, hipStream_t st);
hipError_t launch_adj_combine(const AdjArgs& a, const float2* R, float2* W, hipStream_t st);
hipError_t launch_adj_stage(const AdjStageArgs& a, hipStream_t st);
hipError_t launch_adj_reduce(const float* gpart, size_t per, int B, double* out, hipStream_t st);
hipError_t launch_adj_dcoefs(const float* L, size_t modes, int Bc, double dt, float* dcoef, hipStream_t st);
hipError_t launch_adj_sym(const AdjArgs& a, int kind, float* S, hipStream_t st);

// ---------------------------------------------------------------- unfused path (k_general.hip)
// Meshes and term combinations without fused kernels (any transform length, ConservativeConvection
// in 3D, a complex L with a product term in 2D / 3D) evaluate N(S) in separate launches:
//   fields   F[b][f]  = mask(S) times the field symbol          (half spectrum, [B][NFA][modes])
//   inverse  R[b][f]  = real fields                              ([B][NF][npts] float)
//   products Q[b][p]  = pointwise products of the fields         ([B][NP][npts] float)
//   forward  Qs[b][p] = spectra of the products                  (in F's memory)
//   stage    N = the term's combination of Qs, then the stage combine of U / A1 / A2 -> S
struct GenArgs {
  int term, D, B, C;
  int n[3];
  int H, P;
  size_t modes, npts;
  int kmax[3];
  const float* kv[3];
  size_t kv_bstride;  // 1D per-sample lengths: axis table stride per (b, c) block, else 0
  int NF, NP;
  float coef_t;
  int mask_in_place = 0;  // the evaluation zeroes its input outside the dealiasing box (forced NSP)
};
// fields and products per term (ConservativeConvection: the C (C + 1) / 2 products u_i u_j, i <= j)
inline int gen_nf(int term, int C, int D) {
  switch (term) {
    case TERM_CONV: case TERM_NSP: return C + C * D;
    case TERM_VC: return 4;
    case TERM_KS: return D;
    case TERM_CONS: return C;
    default: return 0;
  }
}
inline int gen_np(int term, int C) {
  switch (term) {
    case TERM_CONV: case TERM_NSP: return C;
    case TERM_CONS: return C * (C + 1) / 2;
    default: return 1;
  }
}
struct GenStageArgs {
  int stage;
  const float2* Qs;
  const float2* dc;      // KSConvection with the batch mean removed: (1/B) sum_b Qs_b(k = 0)
  float2 *U, *A1, *A2, *S;
  const float* coef;     // [Bc][NCOEF][modes]
  const float2* coefc;   // complex-L plans
  size_t coef_bstride;
  int coef_per_block;    // tables per (b, c) block (1D per-sample lengths) instead of per sample
  const float2* src;
  size_t src_bstride, src_cstride;
  // NSPressureConvection with an external force: force spectrum [Bf][Cf][modes] (null: none); with
  // masked_stage set, a stage that recomputes a masked stage spectrum from U and A1 (ST_R2_E1,
  // ST_R4_E2) takes it as zero outside the dealiasing box
  const float2* force = nullptr;
  size_t force_bstride = 0, force_cstride = 0;
  int masked_stage = 0;
  StepMask mask;
};
hipError_t launch_gen_fields(const GenArgs& a, float2* S, float2* F, hipStream_t st);
hipError_t launch_gen_products(const GenArgs& a, const float* R, float* Q, hipStream_t st);
hipError_t launch_gen_stage(const GenArgs& a, const GenStageArgs& s, hipStream_t st);

// ---------------------------------------------------------------- diagnostic maps (k_maps.hip)
enum MapOp : int { MAP_DEALIAS = 0, MAP_DIV = 1, MAP_GRAD = 2, MAP_CURL = 3, MAP_V2V = 4, MAP_PRESSURE = 5,
                   MAP_CONSDIV = 6 };
struct MapArgs {
  int op, D, B, cin, cout;
  int n[3];
  int H, P;
  size_t modes;
  const float* kv[3];
  int kmax[3];
  float coef;
  const float2* in;  // [B][cin][modes]
  float2* out;       // [B][cout][modes]
};
hipError_t launch_spectral_map(const MapArgs& a, hipStream_t st);
hipError_t launch_pointwise_products(const float* u, float* out, int B, int C, size_t npts, hipStream_t st);

constexpr int RK_MAX_TERMS = 8;
struct RkArgs {
  const float2* x;
  const float2* k[RK_MAX_TERMS];
  float c[RK_MAX_TERMS];
  int nk;
  float dt;
  float2* out;
  size_t total;
};
hipError_t launch_rk_combine(const RkArgs& a, hipStream_t st);

hipError_t launch_apply_linear(const float2* U, const float2* Nv, const float* L, size_t l_bstride, size_t modes, int C,
                               size_t total, int have_lin, int have_nl, float2* out, hipStream_t st);
hipError_t launch_apply_linear_c(const float2* U, const float2* Nv, const float2* L, size_t l_bstride, size_t modes,
                                 int C, size_t total, int have_lin, int have_nl, float2* out, hipStream_t st);

}  // namespace tfsm
