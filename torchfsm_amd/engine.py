"""Python handle of a compiled engine plan (tfsm_plan) and its device buffers.

A plan is one operator on one mesh for one batch size: wavenumber and twiddle tables, the linear
symbol L, ETD coefficient tables, accumulated explicit-source spectra, and the workspace layout of
the fused step kernels.  All device memory for states and workspaces comes from the torch caching
allocator; pointers and the current torch stream are passed through the C ABI.
"""

import ctypes
from typing import Optional, Sequence, Tuple

import torch

from . import _lib
from ._lib import check, lib

SCHEME_IDS = {
    "ETDRK0": _lib.ETDRK0,
    "ETDRK1": _lib.ETDRK1,
    "ETDRK2": _lib.ETDRK2,
    "SETDRK1": _lib.SETDRK1,
    "SETDRK2": _lib.SETDRK2,
    "SETDRK3": _lib.SETDRK3,
    "SETDRK4": _lib.SETDRK4,
}


def _stream(device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _ptr(t: Optional[torch.Tensor]) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr() if t is not None else 0)


def require_gpu_tensor(t: torch.Tensor, what="input"):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{what} must be a torch.Tensor")
    if t.device.type != "cuda":
        raise RuntimeError(
            f"{what} is on {t.device}: the MI355X engine runs on the GPU only (move the tensor and mesh to 'cuda')"
        )


class SpectralPlan:
    """tfsm_plan wrapper.

    Args:
        mesh_info: sequence of (start, end, n) per axis.
        batch, channels: B and C of the state.
        nonlinear: None, "conv", "vc", "nsp", "ks", "ks_keep_mean" or "cons" (the dealiased product term)
            and its coefficient.
        linear: list of (kind, coef) with kind in {"lap", "biharm", "unit"} or ("deriv", dim, order)
            (SpatialDerivative); coef a float or a per-sample sequence of length `batch`.
        device: CUDA device.
        dealias_rate: 2/3 by default.
        lengths_per_sample: 1D only, per-sample domain lengths (serving of mixed 1D requests).
    """

    KIND = {"lap": _lib.LIN_LAP, "biharm": _lib.LIN_BIHARM, "unit": _lib.LIN_UNIT}
    TERM = {None: _lib.TERM_NONE, "conv": _lib.TERM_CONV, "vc": _lib.TERM_VC, "nsp": _lib.TERM_NSP,
            "ks": _lib.TERM_KS, "ks_keep_mean": _lib.TERM_KS_KEEP_MEAN, "cons": _lib.TERM_CONS}

    def __init__(self, mesh_info, batch, channels, nonlinear=None, nonlinear_coef=0.0, linear=(), device="cuda",
                 dealias_rate=2 / 3, lengths_per_sample: Optional[Sequence[float]] = None):
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise RuntimeError("the MI355X engine needs a CUDA (HIP) device")
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.mesh_info = [tuple(m) for m in mesh_info]
        self.ndim = len(self.mesh_info)
        self.n = tuple(int(m[2]) for m in self.mesh_info)
        self.batch = int(batch)
        self.channels = int(channels)
        self.nonlinear = nonlinear
        geom = _lib.Geometry()
        geom.ndim = self.ndim
        for d, (s, e, n) in enumerate(self.mesh_info):
            geom.n[d] = int(n)
            geom.length[d] = float(e - s)
        geom.batch = self.batch
        geom.channels = self.channels
        self._lens = None
        if lengths_per_sample is not None:
            self._lens = (ctypes.c_double * self.batch)(*[float(x) for x in lengths_per_sample])
            geom.length_per_sample = ctypes.cast(self._lens, ctypes.POINTER(ctypes.c_double))
        terms = _lib.Terms()
        terms.nonlinear = self.TERM[nonlinear]
        terms.nonlinear_coef = float(nonlinear_coef)
        terms.dealias_rate = float(dealias_rate)
        lin = list(linear)
        if len(lin) > 4:
            raise ValueError("term not supported by engine: more than 4 linear terms")
        terms.n_linear = len(lin)
        per_sample = any(not isinstance(c, (int, float)) for _, c in lin)
        cb = self.batch if per_sample else 1
        flat = []
        for t, (kind, c) in enumerate(lin):
            if isinstance(kind, tuple) and kind[0] == "deriv":
                terms.linear_kind[t] = _lib.LIN_DERIV
                terms.linear_dim[t] = int(kind[1])
                terms.linear_order[t] = int(kind[2])
            elif kind in self.KIND:
                terms.linear_kind[t] = self.KIND[kind]
            else:
                raise ValueError(f"term not supported by engine: linear kind {kind}")
            if isinstance(c, (int, float)):
                flat += [float(c)] * cb
            else:
                vals = [float(v) for v in c]
                if len(vals) == 1:
                    vals = vals * cb
                if len(vals) != cb:
                    raise ValueError("per-sample linear coefficients must match the batch size")
                flat += vals
        self._coefs = (ctypes.c_float * max(1, len(flat)))(*flat)
        terms.linear_coef = ctypes.cast(self._coefs, ctypes.POINTER(ctypes.c_float))
        terms.coef_batch = cb
        self._coef_batch = cb if lengths_per_sample is None else self.batch
        self._plan = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(lib.tfsm_plan_create(ctypes.byref(geom), ctypes.byref(terms), _stream(self.device),
                                       ctypes.byref(self._plan)), "plan")
        self.pitch = int(lib.tfsm_plan_pitch(self._plan))
        self.state_elems = int(lib.tfsm_plan_state_elems(self._plan))
        self.ws_bytes = int(lib.tfsm_plan_workspace_bytes(self._plan))
        self.spec_shape = (self.batch, self.channels) + self.n[:-1] + (self.pitch,)
        self._ws = None
        self.scheme = None
        self.dt = None
        cur = (ctypes.c_int32 * 6)()
        check(lib.tfsm_plan_get_tuning(self._plan, ctypes.cast(cur, ctypes.c_void_p)), "tuning")
        self.tuning = tuple(int(x) for x in cur[:5])
        self.streams = int(cur[5])
        self.persistent_supported = self.persistent  # on by default wherever the plan supports it

    # ------------------------------------------------------------------ lifecycle
    def close(self):
        """Destroy the native plan (idempotent).  The handle is cleared before the destroy call and
        without touching module globals: at interpreter shutdown this can run from a finalizer after
        the module's globals are gone, and a handle left set would be destroyed a second time."""
        if "_step_graph" in self.__dict__:
            self._drop_step_graph()
        h = self.__dict__.get("_plan")
        self._plan = None
        if h is not None and h.value and lib is not None:
            lib.tfsm_plan_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._plan

    def workspace(self) -> torch.Tensor:
        if self._ws is None:
            self._ws = torch.zeros(max(self.ws_bytes, 256), dtype=torch.uint8, device=self.device)
        return self._ws

    def set_tuning(self, col_tile=16, col_var=0, row_var=0, row_wg=256, sub_batch=0):
        """Launch shape of the fused 2D kernels and the temporal-blocking sub-batch (see
        tfsm_plan_set_tuning)."""
        self._drop_step_graph()
        check(lib.tfsm_plan_set_tuning(self._plan, int(col_tile), int(col_var), int(row_var), int(row_wg),
                                       int(sub_batch)), "tuning")
        self.tuning = (int(col_tile), int(col_var), int(row_var), int(row_wg), int(sub_batch))

    @property
    def persistent(self) -> bool:
        """True when 2D stepping runs the persistent per-sample kernel (small square meshes)."""
        return bool(lib.tfsm_plan_persistent(self._plan))

    def set_persistent(self, on: bool):
        """Persistent per-sample 2D stepper (True) or the tiled row and column kernels (False)."""
        self._drop_step_graph()
        check(lib.tfsm_plan_set_persistent(self._plan, 1 if on else 0), "persistent")

    @property
    def general(self) -> bool:
        """True when stepping runs the unfused path (separate field, product, transform and stage
        launches): meshes and terms the fused kernels do not cover, or set_general(True)."""
        return bool(lib.tfsm_plan_general(self._plan))

    def set_general(self, on: bool):
        """Unfused path (True) or the fused kernels (False, where they cover the plan)."""
        self._drop_step_graph()
        check(lib.tfsm_plan_set_general(self._plan, 1 if on else 0), "general")
        ws = int(lib.tfsm_plan_workspace_bytes(self._plan))
        if ws != self.ws_bytes:
            self.ws_bytes = ws
            self._ws = None

    def set_streams(self, nstreams: int):
        """1: one stream; 2-4: sub-batch k runs on stream k % nstreams, overlapping (tfsm_plan_set_streams)."""
        self._drop_step_graph()
        check(lib.tfsm_plan_set_streams(self._plan, int(nstreams)), "streams")
        self.streams = int(nstreams)

    def _drop_step_graph(self):
        """A graph captured by integrate() bakes the scheme's stages and the launch shape in."""
        cache = self.__dict__.pop("_step_graph", None)
        if cache is not None:
            cache[1].close()

    def set_scheme(self, scheme: str, dt: float, n_points: int = 16, radius: float = 1.0):
        self._drop_step_graph()
        with torch.cuda.device(self.device):
            check(lib.tfsm_plan_set_scheme(self._plan, SCHEME_IDS[scheme], float(dt), int(n_points), float(radius),
                                           _stream(self.device)), "scheme")
        self.scheme = scheme
        self.dt = dt

    def set_sample(self, b: int, linear_coefs: Sequence[float], length: float = 0.0):
        """Re-target slot b of a per-sample plan: its linear coefficients (generator order) and,
        for 1D plans with per-sample lengths, its domain length (tfsm_plan_set_sample)."""
        vals = (ctypes.c_float * max(1, len(linear_coefs)))(*[float(c) for c in linear_coefs])
        with torch.cuda.device(self.device):
            check(lib.tfsm_plan_set_sample(self._plan, int(b), ctypes.cast(vals, ctypes.c_void_p), float(length),
                                           _stream(self.device)), "set_sample")

    COEF_SLOTS = {"E": 0, "E2": 1, "c1": 2, "c2": 3, "c3": 4, "c4": 5, "c5": 6, "c6": 7}

    def coefficients(self, name: str) -> torch.Tensor:
        """Coefficient table (E, E2, c1..c6) as float32 [coef_batch, C?=1, N..., P] (valid k < N/2+1)."""
        nb = self.batch if self._coef_batch > 1 else 1
        out = torch.empty((nb, 1) + self.n[:-1] + (self.pitch,), dtype=torch.float32, device=self.device)
        with torch.cuda.device(self.device):
            check(lib.tfsm_plan_coefficients(self._plan, self.COEF_SLOTS[name], _ptr(out), _stream(self.device)),
                  "coefficients")
        return out

    def add_source(self, src: torch.Tensor, coef: float):
        """Explicit source coef * fftn(src); src is (Bs, Cs, N...) with Bs in {1, B}, Cs in {1, C}."""
        self._drop_step_graph()
        if src.dim() != self.ndim + 2:
            raise ValueError(f"explicit source must have shape (B, C, N...), got {tuple(src.shape)}")
        s = src.detach().to(device=self.device, dtype=torch.float32).contiguous()
        with torch.cuda.device(self.device):
            check(lib.tfsm_plan_add_source(self._plan, _ptr(s), int(s.shape[0]), int(s.shape[1]), float(coef),
                                           _stream(self.device)), "source")
            torch.cuda.current_stream(self.device).synchronize()

    def add_force(self, force: torch.Tensor, coef: float):
        """External force coef * fftn(force) of NSPressureConvection (tfsm_plan_add_force); the plan
        then steps on the unfused path, whose workspace is larger."""
        self._drop_step_graph()
        if force.dim() != self.ndim + 2:
            raise ValueError(f"external force must have shape (B, C, N...), got {tuple(force.shape)}")
        f = force.detach().to(device=self.device, dtype=torch.float32).contiguous()
        with torch.cuda.device(self.device):
            check(lib.tfsm_plan_add_force(self._plan, _ptr(f), int(f.shape[0]), int(f.shape[1]), float(coef),
                                          _stream(self.device)), "force")
            torch.cuda.current_stream(self.device).synchronize()
        ws = int(lib.tfsm_plan_workspace_bytes(self._plan))
        if ws != self.ws_bytes:
            self.ws_bytes = ws
            self._ws = None

    # ------------------------------------------------------------------ argument checks
    def _check(self, t: torch.Tensor, trailing, dtype, what: str, nb: Optional[int] = None) -> int:
        """Every tensor handed to the C ABI is a contiguous tensor of this plan's device and dtype whose
        shape is (nb, *trailing): the kernels index it from its shape alone, so a mismatch would read or
        write outside it.  Returns its leading size."""
        require_gpu_tensor(t, what)
        if t.device != self.device:
            raise ValueError(f"{what} is on {t.device}, the plan on {self.device}")
        if t.dtype != dtype:
            raise TypeError(f"{what} must be {dtype}, got {t.dtype}")
        if not t.is_contiguous():
            raise ValueError(f"{what} must be contiguous")
        if t.dim() != 1 + len(trailing) or tuple(t.shape[1:]) != tuple(trailing) or t.shape[0] < 1:
            raise ValueError(f"{what} has shape {tuple(t.shape)}, expected (n, {', '.join(map(str, trailing))})")
        if nb is not None and t.shape[0] != nb:
            raise ValueError(f"{what} holds {t.shape[0]} samples, expected {nb}")
        return int(t.shape[0])

    def _check_state(self, t: torch.Tensor, what: str = "state"):
        self._check(t, self.spec_shape[1:], torch.complex64, what, self.batch)

    def _map_channels(self, op: int):
        """(input, output) channel counts of a spectral map, as tfsm_spectral_map forms them."""
        D, C = self.ndim, self.channels
        table = {_lib.MAP_DEALIAS: (C, C), _lib.MAP_DIV: (D, 1), _lib.MAP_GRAD: (1, D),
                 _lib.MAP_CURL: (D, 1 if D == 2 else 3), _lib.MAP_V2V: (1, 2), _lib.MAP_PRESSURE: (D, 1),
                 _lib.MAP_CONSDIV: (D * D, D)}
        if op not in table:
            raise ValueError(f"unknown spectral map {op}")
        return table[op]

    # ------------------------------------------------------------------ buffers
    def empty_spectrum(self, nbatch=None) -> torch.Tensor:
        nb = self.batch if nbatch is None else nbatch
        return torch.zeros((nb, self.channels) + self.n[:-1] + (self.pitch,), dtype=torch.complex64,
                           device=self.device)

    # ------------------------------------------------------------------ transforms
    def forward(self, u: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        """Real (nb, C, N...) float32 -> half spectrum (nb, C, ..., P)."""
        require_gpu_tensor(u, "forward input")
        u = u.to(torch.float32).contiguous()
        nb = self._check(u, (self.channels,) + self.n, torch.float32, "forward input")
        out = self.empty_spectrum(nb) if out is None else out
        self._check(out, self.spec_shape[1:], torch.complex64, "forward output", nb)
        with torch.cuda.device(self.device):
            check(lib.tfsm_forward(self._plan, _ptr(u), _ptr(out), int(nb), _stream(self.device)), "forward")
        return out

    def inverse(self, uh: torch.Tensor, nbatch: Optional[int] = None) -> torch.Tensor:
        """Half spectrum (nb, C, ..., P) -> real (nb, C, N...) = ifftn(.).real."""
        have = self._check(uh, self.spec_shape[1:], torch.complex64, "inverse input")
        nb = have if nbatch is None else int(nbatch)
        if not 1 <= nb <= have:
            raise ValueError(f"nbatch {nb} outside 1..{have}")
        out = torch.empty((nb, self.channels) + self.n, dtype=torch.float32, device=self.device)
        sb = int(lib.tfsm_plan_inverse_scratch_bytes(self._plan, int(nb)))
        scratch = torch.empty(max(sb, 8), dtype=torch.uint8, device=self.device)
        with torch.cuda.device(self.device):
            check(lib.tfsm_inverse(self._plan, _ptr(uh), _ptr(out), int(nb), _ptr(scratch), _stream(self.device)),
                  "inverse")
        return out

    def half_to_full(self, uh: torch.Tensor, nbatch: Optional[int] = None) -> torch.Tensor:
        have = self._check(uh, self.spec_shape[1:], torch.complex64, "half spectrum")
        nb = have if nbatch is None else int(nbatch)
        if not 1 <= nb <= have:
            raise ValueError(f"nbatch {nb} outside 1..{have}")
        out = torch.empty((nb, self.channels) + self.n, dtype=torch.complex64, device=self.device)
        with torch.cuda.device(self.device):
            check(lib.tfsm_half_to_full(self._plan, _ptr(uh), _ptr(out), int(nb), _stream(self.device)), "half_to_full")
        return out

    def full_to_half(self, full: torch.Tensor) -> torch.Tensor:
        require_gpu_tensor(full, "full spectrum")
        full = full.to(torch.complex64).contiguous()
        nb = self._check(full, (self.channels,) + self.n, torch.complex64, "full spectrum")
        out = self.empty_spectrum(nb)
        with torch.cuda.device(self.device):
            check(lib.tfsm_full_to_half(self._plan, _ptr(full), _ptr(out), int(nb), _stream(self.device)), "full_to_half")
        return out

    def c2c(self, x: torch.Tensor, direction: int) -> torch.Tensor:
        """Complex (nb, C, N...) -> complex, over the last D axes (direction -1 forward, +1 inverse)."""
        require_gpu_tensor(x, "c2c input")
        x = x.to(torch.complex64).contiguous()
        nb = self._check(x, (self.channels,) + self.n, torch.complex64, "c2c input")
        out = torch.empty_like(x)
        with torch.cuda.device(self.device):
            check(lib.tfsm_c2c(self._plan, _ptr(x), _ptr(out), int(nb), int(direction), _stream(self.device)), "c2c")
        return out

    # ------------------------------------------------------------------ stepping
    def step(self, uh: torch.Tensor, nsteps: int, rec_slot: Optional[torch.Tensor] = None,
             traj: Optional[torch.Tensor] = None, rec_slot_host: Optional[Sequence[int]] = None):
        """Advance uh in place by nsteps (optionally recording frames into traj [B][T][C][...][P]).

        Every recorded slot must lie in -1..T-1 (-1: no frame at that step).  The check reads
        rec_slot_host, the caller's host copy of the schedule, when given, and otherwise copies the
        device schedule to the host (one synchronisation)."""
        if self.scheme is None:
            raise RuntimeError("set_scheme must be called before stepping")
        self._check_state(uh)
        if int(nsteps) < 0:
            raise ValueError("nsteps must be >= 0")
        if (rec_slot is None) != (traj is None):
            raise ValueError("recording needs both the slot schedule and the trajectory buffer")
        T = 0
        if traj is not None:
            if rec_slot.dim() != 1 or rec_slot.numel() < int(nsteps) + 1:
                raise ValueError(f"record slots: one per step and one after the last ({int(nsteps) + 1})")
            self._check(rec_slot, (), torch.int32, "record slots")
            if traj.dim() < 3:
                raise ValueError("trajectory buffer must be [B][T][C][...][P]")
            T = int(traj.shape[1])
            self._check(traj, (T,) + tuple(self.spec_shape[1:]), torch.complex64, "trajectory buffer", self.batch)
            sched = (list(rec_slot_host)[: int(nsteps) + 1] if rec_slot_host is not None
                     else rec_slot[: int(nsteps) + 1].cpu().tolist())
            if len(sched) < int(nsteps) + 1:
                raise ValueError(f"record slots: one per step and one after the last ({int(nsteps) + 1})")
            bad = [v for v in sched if not -1 <= int(v) < T]
            if bad:
                raise ValueError(f"record slots must lie in -1..{T - 1} (trajectory of {T} frames); got {bad[0]}")
        ws = self.workspace()
        with torch.cuda.device(self.device):
            check(lib.tfsm_step(self._plan, _ptr(uh), int(nsteps), _ptr(rec_slot), _ptr(traj), T, _ptr(ws),
                                _stream(self.device)), "step")
        return uh

    def step_varying(self, uh: torch.Tensor, nsteps_per_sample: torch.Tensor, max_steps: int):
        """Advance sample b of uh by min(nsteps_per_sample[b], max_steps) steps (device int32 [B])."""
        self._check_state(uh)
        self._check(nsteps_per_sample, (), torch.int32, "steps per sample", self.batch)
        if int(max_steps) < 0:
            raise ValueError("max_steps must be >= 0")
        ws = self.workspace()
        with torch.cuda.device(self.device):
            check(lib.tfsm_step_varying(self._plan, _ptr(uh), _ptr(nsteps_per_sample), int(max_steps), _ptr(ws),
                                        _stream(self.device)), "step_varying")
        return uh

    def apply(self, uh: torch.Tensor) -> torch.Tensor:
        """L uh + N(uh) (half spectra)."""
        self._check_state(uh)
        out = torch.empty_like(uh)
        ws = self.workspace()
        with torch.cuda.device(self.device):
            check(lib.tfsm_apply(self._plan, _ptr(uh), _ptr(out), _ptr(ws), _stream(self.device)), "apply")
        return out

    def rk_combine(self, x: torch.Tensor, ks, coefs, dt: float, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        """x + dt * sum_i coefs[i] ks[i] (half spectra), the reference's stage-sum order."""
        if len(ks) != len(coefs):
            raise ValueError("one coefficient per stage")
        self._check_state(x, "stage base")
        for k in ks:
            self._check_state(k, "stage value")
        out = torch.empty_like(x) if out is None else out
        self._check_state(out, "stage sum")
        ptrs = (ctypes.c_void_p * max(1, len(ks)))(*[k.data_ptr() for k in ks])
        cs = (ctypes.c_float * max(1, len(coefs)))(*[float(c) for c in coefs])
        with torch.cuda.device(self.device):
            check(lib.tfsm_rk_combine(self._plan, _ptr(out), _ptr(x), ctypes.cast(ptrs, ctypes.c_void_p),
                                      ctypes.cast(cs, ctypes.c_void_p), len(ks), float(dt), _stream(self.device)),
                  "rk_combine")
        return out

    def spectral_map(self, op: int, x: torch.Tensor, cout: int, coef: float = 1.0) -> torch.Tensor:
        """coef * map(x) for a diagnostic operator (tfsm_spectral_map); x and the result are half
        spectra of this plan's mesh and batch with the map's channel counts."""
        cin, cmap = self._map_channels(int(op))
        if cout != cmap:
            raise ValueError(f"spectral map {op} gives {cmap} channels, not {cout}")
        x = x.contiguous()
        self._check(x, (cin,) + tuple(self.spec_shape[2:]), torch.complex64, "spectral map input", self.batch)
        out = torch.empty((self.batch, cout) + tuple(self.spec_shape[2:]), dtype=torch.complex64, device=self.device)
        with torch.cuda.device(self.device):
            check(lib.tfsm_spectral_map(self._plan, int(op), float(coef), _ptr(x), _ptr(out),
                                        _stream(self.device)), "spectral map")
        return out

    def pointwise_products(self, u: torch.Tensor) -> torch.Tensor:
        """u_i u_j of a physical field [B][C][N...] (tfsm_pointwise_products) -> [B][C*C][N...]."""
        u = u.contiguous()
        self._check(u, (self.channels,) + self.n, torch.float32, "product input", self.batch)
        out = torch.empty((self.batch, self.channels * self.channels) + tuple(u.shape[2:]), dtype=torch.float32,
                          device=self.device)
        with torch.cuda.device(self.device):
            check(lib.tfsm_pointwise_products(self._plan, _ptr(u), _ptr(out), _stream(self.device)), "products")
        return out

    def solve(self, uh: torch.Tensor) -> torch.Tensor:
        """uh / L with 1/0 := 1 (half spectra)."""
        self._check_state(uh)
        out = torch.empty_like(uh)
        with torch.cuda.device(self.device):
            check(lib.tfsm_solve(self._plan, _ptr(uh), _ptr(out), _stream(self.device)), "solve")
        return out

    def launch_pass(self, uh: torch.Tensor, pass_id: int, stage: int = 0, sub: int = 0, events=None):
        """One row (0) / column (1, stage code) / 1D step (2) / 3D (3..7) pass over sub-batch `sub`,
        for kernel timing.  events: optional (start, stop) torch.cuda.Event pair (enable_timing) that
        the pass's kernel records at its own start and end (tfsm_launch_pass_timed)."""
        self._check_state(uh)
        ws = self.workspace()
        with torch.cuda.device(self.device):
            if events is None:
                check(lib.tfsm_launch_pass(self._plan, _ptr(uh), _ptr(ws), int(pass_id), int(stage), int(sub),
                                           _stream(self.device)), "launch_pass")
            else:
                e0, e1 = events
                check(lib.tfsm_launch_pass_timed(self._plan, _ptr(uh), _ptr(ws), int(pass_id), int(stage), int(sub),
                                                 ctypes.c_void_p(e0.cuda_event), ctypes.c_void_p(e1.cuda_event),
                                                 _stream(self.device)), "launch_pass_timed")

    def graph(self, uh: torch.Tensor, nsteps: int, counts: Optional[torch.Tensor] = None,
              keep_plan: bool = True) -> "StepGraph":
        """Capture nsteps into a HIP graph (StepGraph).  keep_plan=False for a graph the plan itself
        holds (and closes with it), which must not keep the plan alive."""
        return StepGraph(self, uh, nsteps, counts, keep_plan)


class StepGraph:
    """nsteps of tfsm_step captured into a HIP graph on fixed state/workspace buffers.  With
    `counts` (device int32 [B]), nsteps of tfsm_step_varying: every replay reads the per-sample
    step counts from that tensor, which the caller updates before launching."""

    def __init__(self, plan: SpectralPlan, uh: torch.Tensor, nsteps: int, counts: Optional[torch.Tensor] = None,
                 keep_plan: bool = True):
        # the graph's kernels use the plan's tables: a graph the caller holds keeps the plan alive; a
        # graph the plan caches for itself (keep_plan False) is closed by the plan's close() instead,
        # and holding the plan there would make a reference cycle that only the cycle collector (or
        # interpreter shutdown) frees
        self.plan = plan if keep_plan else None
        self.device = plan.device
        self.uh = uh
        self.nsteps = nsteps
        self.counts = counts
        self._g = ctypes.c_void_p()
        plan._check_state(uh)
        ws = plan.workspace()
        with torch.cuda.device(plan.device):
            if counts is None:
                check(lib.tfsm_graph_create(plan.handle, _ptr(uh), int(nsteps), _ptr(ws), _stream(plan.device),
                                            ctypes.byref(self._g)), "graph")
            else:
                plan._check(counts, (), torch.int32, "steps per sample", plan.batch)
                check(lib.tfsm_graph_create_varying(plan.handle, _ptr(uh), _ptr(counts), int(nsteps), _ptr(ws),
                                                    _stream(plan.device), ctypes.byref(self._g)), "graph")

    def launch(self):
        if self._g is None:
            raise RuntimeError("graph launch: the graph is closed")
        with torch.cuda.device(self.device):
            check(lib.tfsm_graph_launch(self._g, _stream(self.device)), "graph launch")

    def close(self):
        """Destroy the native graph (idempotent; the handle is cleared first, see SpectralPlan.close)."""
        g = self.__dict__.get("_g")
        self._g = None
        if g is not None and g.value and lib is not None:
            lib.tfsm_graph_destroy(g)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_TRANSFORM_PLANS = {}


def transform_plan(mesh_info, device) -> SpectralPlan:
    """Geometry-only plan (no operator) used by FourierMesh.fft / .ifft; cached per mesh and device."""
    key = (tuple(tuple(float(x) if i < 2 else int(x) for i, x in enumerate(m)) for m in mesh_info), str(device))
    p = _TRANSFORM_PLANS.get(key)
    if p is None:
        p = SpectralPlan(mesh_info, 1, 1, device=device)
        _TRANSFORM_PLANS[key] = p
    return p
