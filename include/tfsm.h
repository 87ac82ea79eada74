/* tfsm.h -- C ABI of the MI355X ETD time-stepping engine (libtfsm.so).
 *
 * Plain C types only: device buffers are raw pointers (typically torch tensor data_ptr()), the
 * stream is a hipStream_t passed as void*, sizes are integers.  Every function returns 0 on
 * success or a negative TFSM_E* code; tfsm_last_error() returns a thread-local message for the
 * last failure.  All work is asynchronous on the given stream; nothing synchronises the host.
 *
 * The reference (qiauil/torchfsm) is pure Python over PyTorch and binds no FFI; each entry point
 * below names the reference Python interface whose work it replaces.
 *
 * Layout: a half spectrum of a batch is [nbatch][C][N0]...[N_{D-2}][P] complex64 (interleaved
 * float2), the last axis holding k = 0..N_{D-1}/2 and P = tfsm_plan_pitch() >= N_{D-1}/2+1.
 * Real fields are [nbatch][C][N0]...[N_{D-1}] float32, contiguous.
 */
#ifndef TFSM_H
#define TFSM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TFSM_OK 0
#define TFSM_E_INVALID -1     /* bad argument or unsupported configuration (ValueError) */
#define TFSM_E_STATE -2       /* call out of order, e.g. stepping before a scheme is set */
#define TFSM_E_DEVICE -3      /* HIP runtime error (launch, copy, allocation) */
#define TFSM_E_OOM -4         /* device allocation failed (RuntimeError wrapping OOM) */

/* nonlinear term of the operator (the dealiased NonlinearFunc of the closure) */
#define TFSM_TERM_NONE 0
#define TFSM_TERM_CONV 1      /* Convection: sum_i u_i d_i u_c, C == D   (_convection.py:18-46) */
#define TFSM_TERM_VC 2        /* VorticityConvection, D == 2, C == 1     (_navier_stokes.py:27-46) */
#define TFSM_TERM_NSP 3       /* NSPressureConvection without force: -(I - k k^T/|k|^2) applied to
                                 Convection, C == D in {2, 3}       (_navier_stokes.py:221-262) */
#define TFSM_TERM_KS 4        /* KSConvection(remove_mean=True): 1/2 |grad phi|^2 minus its mean over
                                 batch and space, C == 1, D in {1, 2, 3} (_ks_convection.py:8-38) */
#define TFSM_TERM_KS_KEEP_MEAN 5  /* KSConvection(remove_mean=False)                             */
#define TFSM_TERM_CONS 6      /* ConservativeConvection: sum_i d_i (u_i u_c), C == D in {1, 2}
                                 (_conservative_convection.py:17-27)                          */

/* linear symbols (LinearCoef cores) */
#define TFSM_LIN_LAP 1        /* Laplacian       (_laplacian.py:12-15)  */
#define TFSM_LIN_BIHARM 2     /* Biharmonic      (_biharmonic.py:13-16) */
#define TFSM_LIN_UNIT 3       /* ImplicitSource() unit symbol (_source.py:9-17) */
#define TFSM_LIN_DERIV 4      /* SpatialDerivative(dim, order): (i a_dim)^order (_spatial_derivative.py:7-20,
                                 mesh.py:399-404); an odd order makes L imaginary (complex-L plan) */

/* time integration schemes (ETDRKIntegrator / SETDRKIntegrator members) */
#define TFSM_ETDRK0 0
#define TFSM_ETDRK1 1
#define TFSM_ETDRK2 2
#define TFSM_SETDRK1 3
#define TFSM_SETDRK2 4
#define TFSM_SETDRK3 5
#define TFSM_SETDRK4 6

typedef struct tfsm_plan tfsm_plan;
typedef struct tfsm_graph tfsm_graph;

/* Periodic box and batch: MeshGrid / FourierMesh (mesh.py:9-146, 294-344) plus the (B, C) of the
 * state registered by OperatorLike.register_mesh (_base.py:581-624). */
typedef struct tfsm_geometry {
  int32_t ndim;                      /* 1, 2 or 3 */
  int32_t n[3];                      /* points per axis, >= 2 (any length; see tfsm_plan_create) */
  double length[3];                  /* end - start per axis */
  int32_t batch;                     /* B */
  int32_t channels;                  /* C */
  const double* length_per_sample;   /* 1D only, optional host [batch]: per-sample domain length */
} tfsm_geometry;

/* Operator composition after register_mesh (_base.py:339-406): L = sum coef_t * symbol_t in
 * generator order, N = nonlinear_coef * term(dealiased u^) (+ explicit sources, added with
 * tfsm_plan_add_source). */
typedef struct tfsm_terms {
  int32_t nonlinear;                 /* TFSM_TERM_* */
  float nonlinear_coef;
  double dealias_rate;               /* 2/3 by default (_base.py:317) */
  int32_t n_linear;                  /* 0..4 */
  int32_t linear_kind[4];            /* TFSM_LIN_* */
  const float* linear_coef;          /* host [n_linear][coef_batch] */
  int32_t coef_batch;                /* 1, or batch for per-sample coefficients (batched nu) */
  int32_t linear_dim[4];             /* TFSM_LIN_DERIV: axis index */
  int32_t linear_order[4];           /* TFSM_LIN_DERIV: derivative order (1..8) */
} tfsm_terms;

/* Message of the last failure on this thread ("" if none). */
const char* tfsm_last_error(void);

/* Library version, (major << 16) | minor. */
int tfsm_version(void);

/* Compile an operator on a mesh into a plan: symbols, dealias limits, twiddle tables, L.
 * Any mesh of >= 2 points per axis is accepted (the reference transforms any length,
 * mesh.py:481-491).  The fused step kernels cover 1D powers of two 32..4096 and 48, 96, 100, 112,
 * 192, 200, 384, 400, 768, 800 points; 2D axes of powers of two 16..1024 and 48, 96, 100, 112, 192,
 * 200; 3D axes of powers of two 16..128 and cubic 48, 96, 100 meshes.  Other plans (other sizes,
 * ConservativeConvection in 3D, a complex L with a product term in 2D / 3D, NSPressureConvection
 * with an external force, tfsm_plan_add_force) step through the
 * unfused path: separate field, product and stage launches with general-length FFTs (Stockham with
 * radices 2..7 in LDS, Bluestein for other lengths, four-step beyond 4096 points).
 * Replaces OperatorLike.register_mesh + _build_linear_coefs + _build_nonlinear_funcs
 * (_base.py:339-406, 581-624) and the FourierMesh symbol caches (mesh.py:399-461). */
int tfsm_plan_create(const tfsm_geometry* geom, const tfsm_terms* terms, void* stream, tfsm_plan** out);

/* Free a plan and its tables.  (Reference: garbage collection of _state_dict, _base.py:307-315.) */
int tfsm_plan_destroy(tfsm_plan* plan);

/* Pitch P of the last spectral axis (complex elements). */
int64_t tfsm_plan_pitch(const tfsm_plan* plan);

/* Complex elements of one half-spectrum state [batch][C][...][P]. */
int64_t tfsm_plan_state_elems(const tfsm_plan* plan);

/* Bytes of caller-provided workspace that tfsm_step / tfsm_apply need. */
int64_t tfsm_plan_workspace_bytes(const tfsm_plan* plan);

/* Bytes of scratch tfsm_inverse needs for nbatch blocks of [C][N...]. */
int64_t tfsm_plan_inverse_scratch_bytes(const tfsm_plan* plan, int32_t nbatch);

/* Launch shape of the fused 2D kernels (chosen per shape class by the autotuner): column tile
 * width (8 or 16 columns), radix-plan variant of the column and row transforms (0 or 1: variant 1
 * carries a transform with half as many threads), row-kernel workgroup size (64 or 256 threads),
 * and the sub-batch of samples stepped together (through all steps of a tfsm_step call on one
 * stream, in rounds of 16 steps on several; 0 = the whole batch; smaller sub-batches keep their
 * working set in the last-level cache).
 * (No reference counterpart: launch configuration of the engine's kernels.) */
int tfsm_plan_set_tuning(tfsm_plan* plan, int32_t col_tile, int32_t col_var, int32_t row_var, int32_t row_wg,
                         int32_t sub_batch);

/* Current launch shape: out[6] = column tile, column variant, row variant, row workgroup,
 * sub-batch, streams.  A new plan starts from a default derived from its shape (2D: column tile
 * and radix variant per axis size and product term, 64-thread row groups for 64- and 128-point
 * rows, sub-batch sized to keep >= 512 column workgroups per launch, and sub-batches on two
 * streams for 512^2 VorticityConvection (16 samples, B >= 32), Convection and
 * NSPressureConvection (16 samples from B = 32, 8 at B = 16..31) and 256^2 VorticityConvection
 * and Convection (16 samples, B >= 32; four streams from B = 64); 3D: radix-4
 * z rows from 64 points, and for Convection and NSPressureConvection sub-batches of 2 on two
 * streams at 128^3 with 4 or more samples, on four streams at 64^3 with 8 or more).
 * (No reference counterpart: launch configuration.) */
int tfsm_plan_get_tuning(const tfsm_plan* plan, int32_t* out);

/* Spread the sub-batches of tfsm_step (see tfsm_plan_set_tuning) over nstreams streams
 * (1 <= nstreams <= TFSM_MAX_STREAMS): sub-batch k runs on the caller's stream when
 * k % nstreams == 0 and otherwise on one of the plan's auxiliary streams (default priority),
 * forked from and joined back into the caller's stream with events, so that their kernels
 * overlap; steps are queued in rounds over the sub-batches so that every stream stays fed.  1
 * (the default, except for the multi-stream defaults of tfsm_plan_get_tuning) keeps one stream.
 * Results are identical.  (No reference counterpart: launch configuration.) */
#define TFSM_MAX_STREAMS 4
int tfsm_plan_set_streams(tfsm_plan* plan, int32_t nstreams);

/* Step 2D plans of square 32^2 or 64^2 meshes with Convection or VorticityConvection with the
 * persistent per-sample kernel (on = 1, the default for those plans: one workgroup per sample
 * runs every step of a tfsm_step call, one launch per call) or with the tiled row and column
 * kernels (on = 0).  Same arithmetic; the two kernels contract multiply-adds differently, so
 * results agree to float32 rounding (relative 1e-6), not bit for bit.  TFSM_E_INVALID for other
 * plans.  (No reference counterpart: launch configuration.) */
int tfsm_plan_set_persistent(tfsm_plan* plan, int32_t on);

/* 1 if tfsm_step runs the persistent per-sample 2D kernel for this plan, else 0. */
int32_t tfsm_plan_persistent(const tfsm_plan* plan);

/* Step through the unfused path (on = 1: separate field, product, transform and stage launches)
 * or the fused kernels (on = 0, TFSM_E_INVALID where none cover the plan).  Plans the fused
 * kernels do not cover start on the unfused path.  Switching grows or shrinks
 * tfsm_plan_workspace_bytes.  (No reference counterpart: choice of kernels.) */
int tfsm_plan_set_general(tfsm_plan* plan, int32_t on);

/* 1 if tfsm_step runs the unfused path for this plan, else 0. */
int32_t tfsm_plan_general(const tfsm_plan* plan);

/* ETD coefficients on the device: SETDRK1-4 as float64 contour means of n_points points on a
 * circle of radius `radius`; ETDRK0-2 with the reference's complex64 closed forms.
 * Replaces OperatorLike._build_integrator (_base.py:441-526) with ETDRK0/1/2.__init__
 * (_etdrk.py:15-70) and UnCachedSETDRK1-4.__init__ (_uncached.py:20-211). */
int tfsm_plan_set_scheme(tfsm_plan* plan, int32_t scheme, double dt, int32_t n_points, double radius, void* stream);

/* Re-target sample b of a plan created with per-sample coefficients (coef_batch = batch): new
 * linear coefficients (host float[n_linear], in generator order) and, for 1D plans created with
 * length_per_sample, a new domain length (> 0; <= 0 keeps it).  Recomputes that sample's axis
 * table, L and, when a scheme is set, its ETD coefficients, asynchronously on the stream; other
 * samples are untouched.  Used by the serving loop to admit a request into a free slot.
 * Replaces register_mesh + _build_integrator (_base.py:441-526, 581-624) for one sample of a
 * batched-parameter operator. */
int tfsm_plan_set_sample(tfsm_plan* plan, int32_t b, const float* linear_coef, double length, void* stream);

/* Copy coefficient table `slot` (0 E, 1 E2, 2..7 coef_1..coef_6) into out: float32
 * [coef_batch][modes] with modes = prod(N[:-1]) * P.  Inspection of the integrator's
 * _exp_term / _half_exp_term / _coef_i attributes (_etdrk.py:21-70, _uncached.py:30-211). */
int tfsm_plan_coefficients(tfsm_plan* plan, int32_t slot, float* out, void* stream);

/* Add an explicit source term coef * fftn(src) to every nonlinear evaluation.  src: device
 * float32 [src_batch][src_channels][N...] (src_batch in {1, B}, src_channels in {1, C}); the
 * spectrum is computed here on the device.  Replaces _ExplicitSourceCore (_base.py:994-1015). */
int tfsm_plan_add_source(tfsm_plan* plan, const float* src, int32_t src_batch, int32_t src_channels, float coef,
                         void* stream);

/* External force f of NSPressureConvection (NavierStokes(Re, force), pde.py:85-100): adds
 * coef * fftn(force) to the plan's force spectrum; force is [force_batch][force_channels][N...]
 * float32 on the device, force_batch in {1, batch}, force_channels in {1, channels}.  The forced
 * term is N = grad p - (conv - f) + f with p = invlap div (conv - f), and, as in the reference, it
 * zeroes the stage spectra it evaluates outside the dealiasing box in place; such plans step on the
 * unfused path.  TFSM_E_INVALID for plans without NSPressureConvection.
 * Replaces _NSPressureConvectionCore with external_force (_navier_stokes.py:221-262) for forces
 * made of explicit sources (_base.py:994-1028). */
int tfsm_plan_add_force(tfsm_plan* p, const float* force, int32_t force_batch, int32_t force_channels, float coef,
                        void* stream);

/* Real fields -> half spectra, unnormalised: nbatch blocks of [C][N...].
 * Replaces FourierMesh.fft (mesh.py:481-485) for real input. */
int tfsm_forward(tfsm_plan* plan, const float* u, void* uh, int32_t nbatch, void* stream);

/* Half spectra -> real fields, scaled by 1 / prod(N): ifftn(.).real (mesh.py:487-491, and the
 * output path of OperatorLike.integrate, _base.py:751).  scratch >= tfsm_plan_inverse_scratch_bytes. */
int tfsm_inverse(tfsm_plan* plan, const void* uh, float* u, int32_t nbatch, void* scratch, void* stream);

/* Complex fields -> complex spectra (dir = -1, unnormalised) or back (dir = +1, scaled by
 * 1 / prod(N)) over the last D axes of contiguous [nbatch][C][N...] complex64 arrays; out may
 * equal in.  Replaces FourierMesh.fft / .ifft for complex input (mesh.py:481-491). */
int tfsm_c2c(tfsm_plan* plan, const void* in, void* out, int32_t nbatch, int32_t dir, void* stream);

/* Half -> full complex spectrum with Hermitian fill (return_in_fourier, _base.py:748-749). */
int tfsm_half_to_full(tfsm_plan* plan, const void* uh, void* full, int32_t nbatch, void* stream);

/* Full complex spectrum -> half spectrum (Hermitian part; exact for spectra of real fields).
 * Entry of u_0_fft into OperatorLike.integrate (_base.py:727-728). */
int tfsm_full_to_half(tfsm_plan* plan, const void* full, void* uh, int32_t nbatch, void* stream);

/* Advance the state uh (in place) by nsteps steps of the plan's scheme.  If rec_slot (device
 * int32 [nsteps + 1]) is given, the state before step i is copied to traj[b][rec_slot[i]] when
 * 0 <= rec_slot[i] < T (i = nsteps: the final state; other slot values record nothing); traj is
 * [B][T][C][...][P].
 * Replaces the hot loop of OperatorLike.integrate (_base.py:729-746) with integrator.step
 * (_setdrk_step.py:5-82, _etdrk.py:23-82), the nonlinear closure (_base.py:375-403) and
 * _TrajRecorder.record (traj_recorder.py:46-55). */
int tfsm_step(tfsm_plan* plan, void* uh, int32_t nsteps, const int32_t* rec_slot, void* traj, int32_t T, void* ws,
              void* stream);

/* Bytes of the static device scratch that tfsm_step_noalloc uses as its workspace (64 MiB). */
#define TFSM_NOALLOC_SCRATCH_BYTES (64ll << 20)

/* tfsm_step without a caller workspace and without any device allocation: the workspace is a
 * static device array of TFSM_NOALLOC_SCRATCH_BYTES per GPU (loaded with the library's code
 * object, held until the process ends; its address is resolved once per device on the first
 * call with that device current).  For callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.  The plan's device must
 * be current, and a plan whose tfsm_plan_workspace_bytes exceeds the scratch is rejected
 * (TFSM_E_INVALID).  Every call on one GPU shares that GPU's scratch: calls that overlap on one GPU
 * (two streams, two threads) race and must be serialised by the caller.  The scratch keeps the
 * last call's contents; the step reads no workspace value it has not written in the same call, so
 * nothing is cleared.  Results are identical to tfsm_step's.
 * (No reference counterpart: an allocation-free variant of tfsm_step for callers that cannot
 * allocate.) */
int tfsm_step_noalloc(tfsm_plan* plan, void* uh, int32_t nsteps, const int32_t* rec_slot, void* traj, int32_t T,
                      void* stream);

/* Like tfsm_step (no recording), with a per-sample step count (device int32 [B]): sample b advances
 * by min(nsteps_per_sample[b], max_steps) steps and is not touched after that (a count of 0 leaves
 * it as it is).  Continuous batching of requests: a slot whose request finishes inside a chunk
 * stops there, an empty slot is not stepped.  1D plans run one persistent launch; 2D / 3D plans
 * launch their kernels max_steps times and every workgroup of a finished sample returns at once.
 * Not available for ETDRK0, for 1D KSConvection, or with KSConvection's remove_mean (its mean
 * couples the samples).  Replaces, per request, the hot loop of OperatorLike.integrate
 * (_base.py:729-746). */
int tfsm_step_varying(tfsm_plan* plan, void* uh, const int32_t* nsteps_per_sample, int32_t max_steps, void* ws,
                      void* stream);

/* out = L uh + N(uh) (half spectra), or only the linear / nonlinear part when the operator has
 * only that part.  Replaces OperatorLike.__call__ (_base.py:753-790) via _build_operator
 * (_base.py:408-439). */
int tfsm_apply(tfsm_plan* plan, const void* uh, void* out, void* ws, void* stream);

/* Runge-Kutta stage combination on half spectra of the plan's state size:
 * out = x + dt * ((coefs[0] ks[0] + coefs[1] ks[1]) + ...), nk <= 8 (host arrays of device
 * pointers and coefficients), in the order of the reference's stage sums.  With tfsm_apply as the
 * right-hand side L u + N(u) this runs the explicit Runge-Kutta integrators.  Replaces the stage
 * and update sums of _RKBase._rk_step (_rk.py:43-58). */
/* Diagnostic spectral maps (single applications of operators that change the channel count), on
 * half spectra of the plan's mesh and batch; out = coef * map(in):
 *   TFSM_MAP_DEALIAS   C -> C   u^ * low_pass_filter (mesh.py:443-461), C = the plan's channels
 *   TFSM_MAP_DIV       D -> 1   sum_i (i a_i) u_i                       (_div.py:17-28)
 *   TFSM_MAP_GRAD      1 -> D   (i a_i) p                               (_grad.py:12-15)
 *   TFSM_MAP_CURL      2 -> 1 (2D), 3 -> 3 (3D)                         (_curl.py:18-56)
 *   TFSM_MAP_V2V       1 -> 2   -invlap [i a_1, -i a_0] w, 2D           (_navier_stokes.py:75-91)
 *   TFSM_MAP_PRESSURE  D -> 1   -(sum_i (i a_i) conv_i) invlap          (_navier_stokes.py:140-144, 202-206)
 *   TFSM_MAP_CONSDIV   D*D -> D sum_i (i a_i) X[i][c]                  (_conservative_convection.py:23-27) */
#define TFSM_MAP_DEALIAS 0
#define TFSM_MAP_DIV 1
#define TFSM_MAP_GRAD 2
#define TFSM_MAP_CURL 3
#define TFSM_MAP_V2V 4
#define TFSM_MAP_PRESSURE 5
#define TFSM_MAP_CONSDIV 6
int tfsm_spectral_map(tfsm_plan* plan, int32_t op, float coef, const void* in, void* out, void* stream);

/* Physical-space products out[b][i][j] = u[b][i] * u[b][j] of a [batch][C][N...] float field of the
 * plan's mesh and batch (ConservativeConvection's u u^T, _conservative_convection.py:25). */
int tfsm_pointwise_products(tfsm_plan* plan, const float* u, float* out, void* stream);

int tfsm_rk_combine(tfsm_plan* plan, void* out, const void* x, const void* const* ks, const float* coefs, int32_t nk,
                    float dt, void* stream);

/* out = uh * where(L == 0, 1, 1 / L) (half spectra).  Replaces LinearOperator.solve
 * (_base.py:217-262). */
int tfsm_solve(tfsm_plan* plan, const void* uh, void* out, void* stream);

/* Launch one pass of the fused step on the current state/workspace, for profiling, over sub-batch
 * `sub` (0, 1, ...: samples [sub * sub_batch, ...) of tfsm_plan_set_tuning's sub-batch; the whole
 * batch when none is set), exactly as tfsm_step launches it: pass 0 = the
 * row kernel (2D), pass 1 = the column kernel with stage code `stage` (0 prologue, 7..10 the four
 * SETDRK4 evaluations), pass 2 = the 1D persistent stepper for one step, pass 3 = one whole 3D
 * evaluation (z-row, y-forward, x-column and y-inverse kernels) with stage code `stage`, passes
 * 4..7 = one of those four kernels alone.  (No reference
 * counterpart: instrumentation of the engine's own kernels.) */
int tfsm_launch_pass(tfsm_plan* plan, void* uh, void* ws, int32_t pass_id, int32_t stage, int32_t sub,
                     void* stream);

/* tfsm_launch_pass with its (first) kernel timed: ev_start and ev_stop (hipEvent_t, created by the
 * caller with timing enabled) are recorded at that kernel's start and end by the dispatch itself
 * (hipExtLaunchKernelGGL), so hipEventElapsedTime gives the kernel's own duration, without the
 * launch gap an event recorded between launches includes.  (No reference counterpart: profiling.) */
int tfsm_launch_pass_timed(tfsm_plan* p, void* uh, void* ws, int32_t pass_id, int32_t stage, int32_t sub,
                           void* ev_start, void* ev_stop, void* stream);

/* Bytes of the adjoint workspace tfsm_adjoint_step needs (stage arrays, transposed-Jacobian
 * fields and products). */
int64_t tfsm_plan_adjoint_workspace_bytes(const tfsm_plan* plan);

/* Prepare the adjoint of the plan's current scheme (ETDRK1/2, SETDRK1/2; Convection in 1D/2D):
 * float64 tables of dE/dL, dc1/dL, dc2/dL and, when grad_term >= 0, the symbol dL/dtheta of linear
 * term grad_term (its coefficient theta receives a gradient).  Call again after tfsm_plan_set_scheme.
 * Replaces what torch autograd records through integrator.step (_etdrk.py:72-82,
 * _setdrk_step.py:14-25) and the coefficient constructors (_etdrk.py:54-70, _uncached.py:91-104). */
int tfsm_adjoint_prepare(tfsm_plan* plan, int32_t grad_term, void* stream);

/* One reverse step: given the state u before a step and the adjoint ubar_next of the state after
 * it, write the adjoint ubar of u (real inner product over the full spectrum; half spectra) and,
 * if grad is non-null (device float64 [batch]), add each sample's derivative of the loss with
 * respect to the prepared term's coefficient.  Recomputes N(u), the stage a and N(a) from u.
 * ws: tfsm_plan_workspace_bytes, wsa: tfsm_plan_adjoint_workspace_bytes.
 * Replaces torch autograd's backward through one integrator.step call. */
int tfsm_adjoint_step(tfsm_plan* plan, void* u, const void* ubar_next, void* ubar, double* grad, void* ws, void* wsa,
                      void* stream);

/* Capture nsteps of tfsm_step (no recording) on uh/ws into a HIP graph for repeated launch.
 * (No reference counterpart: the reference launches ~116 kernels per step from Python.) */
int tfsm_graph_create(tfsm_plan* plan, void* uh, int32_t nsteps, void* ws, void* stream, tfsm_graph** out);

/* Capture max_steps of tfsm_step_varying (2D / 3D plans) into a HIP graph.  The kernels read the
 * per-sample counts from nsteps_per_sample (device int32 [B]) at every launch of the graph, so the
 * caller rewrites that array (on the launch stream) before each replay.
 * (No reference counterpart: launch batching of the serving loop.) */
int tfsm_graph_create_varying(tfsm_plan* plan, void* uh, const int32_t* nsteps_per_sample, int32_t max_steps,
                              void* ws, void* stream, tfsm_graph** out);
int tfsm_graph_launch(tfsm_graph* graph, void* stream);
int tfsm_graph_destroy(tfsm_graph* graph);

#ifdef __cplusplus
}
#endif

#endif /* TFSM_H */
