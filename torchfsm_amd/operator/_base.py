"""Operator algebra and the integrate loop with the reference's interface
(operator/_base.py:16-1028), compiled onto the MI355X engine.

register_mesh turns the generator list into an engine term descriptor: linear symbols
{Laplacian, Biharmonic, ImplicitSource()} with their coefficients, at most one dealiased product
term {Convection, VorticityConvection} and any number of ExplicitSource spectra.  integrate and
__call__ run entirely in libtfsm; a term outside that set raises ValueError (there is no torch
fallback).
"""

import os
from abc import ABC, abstractmethod
from typing import Callable, List, Literal, Optional, Sequence, Tuple, Union

import torch
from torch import Tensor
from tqdm.auto import tqdm

from ..engine import SpectralPlan, require_gpu_tensor
from ..integrator import ETDRKIntegrator, RKIntegrator, SETDRKIntegrator
from ..mesh import FourierMesh, MeshGrid
from ..traj_recorder import AutoRecorder, CPURecorder, _TrajRecorder
from ..utils import default


class LinearCoef(ABC):
    """A linear symbol generator core (_base.py:16-55)."""

    _engine_kind: Optional[str] = None

    @abstractmethod
    def __call__(self, f_mesh: FourierMesh, n_channel: int) -> Tensor:
        raise NotImplementedError

    def nonlinear_like(self, u_fft, f_mesh, u=None):
        return self(f_mesh, u_fft.shape[1]) * u_fft


class NonlinearFunc(ABC):
    """A nonlinear term core (_base.py:58-109)."""

    _engine_kind: Optional[str] = None

    def __init__(self, dealiasing_swtich: bool = True) -> None:
        self._dealiasing_swtich = dealiasing_swtich

    @abstractmethod
    def __call__(self, u_fft, f_mesh: FourierMesh, u=None):
        raise NotImplementedError

    def spatial_value(self, u_fft, f_mesh: FourierMesh, u=None):
        return f_mesh.ifft(self(u_fft, f_mesh, u)).real


class CoreGenerator(ABC):
    """Builds a core for a mesh and channel count (_base.py:112-131)."""

    @abstractmethod
    def __call__(self, f_mesh: FourierMesh, n_channel: int) -> Union[LinearCoef, NonlinearFunc]:
        raise NotImplementedError


GeneratorLike = Union[CoreGenerator, Callable[[FourierMesh, int], Union[LinearCoef, NonlinearFunc]]]


def check_value_with_mesh(u, mesh):
    """_base.py:144-167."""
    if isinstance(mesh, (FourierMesh, MeshGrid)):
        mesh = mesh.mesh_info
    n_dim = len(mesh)
    value_shape = u.shape
    if len(value_shape) - 2 != n_dim:
        raise ValueError(f"the value shape {value_shape} is not compatible with mesh dim {n_dim}")
    for i in range(n_dim):
        if value_shape[i + 2] != mesh[i][2]:
            raise ValueError(f"Expect to have {mesh[i][2]} points in dim {i} but got {value_shape[i+2]}")


class _MutableMixIn:
    """+, -, scalar * and / on operators (_base.py:170-207)."""

    def __radd__(self, other):
        return self + other

    def __iadd__(self, other):
        return self + other

    def __sub__(self, other):
        try:
            return self + (-1 * other)
        except Exception:
            return NotImplemented

    def __rsub__(self, other):
        try:
            return other + (-1 * self)
        except Exception:
            return NotImplemented

    def __isub__(self, other):
        return self - other

    def __rmul__(self, other):
        return self * other

    def __imul__(self, other):
        return self * other

    def __truediv__(self, other):
        try:
            return self * (1 / other)
        except Exception:
            return NotImplemented


def _coef_value(coef, batch):
    """A python number, or a tensor broadcast over the batch (per-sample coefficients)."""
    if isinstance(coef, (int, float)):
        return float(coef)
    if isinstance(coef, Tensor):
        if coef.is_complex():
            raise ValueError("term not supported by engine: complex coefficient")
        flat = coef.detach().reshape(-1).to("cpu", torch.float64)
        if flat.numel() == 1:
            return float(flat[0])
        if coef.dim() >= 1 and coef.shape[0] == batch and flat.numel() == batch:
            return [float(v) for v in flat]
        raise ValueError(f"term not supported by engine: coefficient of shape {tuple(coef.shape)}")
    if isinstance(coef, complex):
        raise ValueError("term not supported by engine: complex coefficient")
    return float(coef)


GRAPH_STEPS = 16  # steps per captured graph replayed by integrate() on small multi-launch meshes


def _graph_worthwhile(plan) -> bool:
    """Plans whose steps are several short launches (2D without the persistent stepper up to 128^2,
    3D up to 64^3, the unfused path up to 2^16 points): launch overhead is a large part of a step
    there, and a replayed HIP graph removes most of it."""
    if plan.persistent:
        return False
    pts = 1
    for n in plan.n:
        pts *= int(n)
    if plan.general:
        return pts <= 1 << 16
    if plan.ndim == 1:
        return False
    return pts <= (64 ** 3 if plan.ndim == 3 else 128 * 128)


class OperatorLike(_MutableMixIn):
    """Base of all operators (_base.py:284-805)."""

    def __init__(self, operator_generators=None, coefs: Optional[List] = None) -> None:
        super().__init__()
        self.operator_generators = default(operator_generators, [])
        if not isinstance(self.operator_generators, list):
            self.operator_generators = [self.operator_generators]
        self.coefs = default(coefs, [1] * len(self.operator_generators))
        self._state_dict = {
            "f_mesh": None,
            "n_channel": None,
            "linear_coef": None,
            "nonlinear_func": None,
            "operator": None,
            "integrator": None,
            "invert_linear_coef": None,
            "engine": None,
        }
        self._nonlinear_funcs = []
        self._linear_cores = []
        self._de_aliasing_rate = 2 / 3
        self._value_mesh_check_func = lambda dim_value, dim_mesh: True
        self._integrator = "auto"
        self._integrator_config = {}
        self._is_etdrk_integrator = True

    # ------------------------------------------------------------------ composition
    @property
    def is_linear(self) -> bool:
        assert self._state_dict["f_mesh"] is not None, "Mesh should be registered before checking if the operator is linear"
        return len(self._nonlinear_funcs) == 0 and len(self._linear_cores) > 0

    def register_mesh(self, mesh, n_channel: int, device=None, dtype=None):
        """Split generators into linear symbols and nonlinear terms (_base.py:581-624)."""
        if isinstance(mesh, FourierMesh):
            f_mesh = mesh
            if device is not None or dtype is not None:
                f_mesh.to(device=device, dtype=dtype)
        else:
            f_mesh = FourierMesh(mesh, device=device, dtype=dtype)
        for key in self._state_dict:
            self._state_dict[key] = None
        self._state_dict.update({"f_mesh": f_mesh, "n_channel": n_channel, "engine": {}})
        linear, nonlinear = [], []
        for coef, generator in zip(self.coefs, self.operator_generators):
            op = generator(f_mesh, n_channel)
            if isinstance(op, LinearCoef):
                linear.append((coef, op))
            elif isinstance(op, NonlinearFunc):
                nonlinear.append((coef, op))
            else:
                raise ValueError(f"Operator {op} is not supported")
        self._linear_cores = linear
        self._nonlinear_funcs = nonlinear
        if nonlinear:
            f_mesh.set_default_rel_freq_threshold(self._de_aliasing_rate)

    def register_additional_check(self, func: Callable[[int, int], bool]):
        self._value_mesh_check_func = func

    def add_generator(self, generator, coef=1):
        self.operator_generators.append(generator)
        self.coefs.append(coef)

    def set_integrator(self, integrator, **integrator_config):
        """_base.py:646-674."""
        if isinstance(integrator, str):
            assert integrator == "auto", (
                "The integrator should be 'auto' or an instance of ETDRKIntegrator, SETDRKIntegrator or RKIntegrator"
            )
        else:
            assert isinstance(integrator, (ETDRKIntegrator, SETDRKIntegrator, RKIntegrator)), (
                "The integrator should be 'auto' or an instance of ETDRKIntegrator, SETDRKIntegrator or RKIntegrator"
            )
        self._integrator = integrator
        self._integrator_config = integrator_config
        self._state_dict["integrator"] = None

    # ------------------------------------------------------------------ engine compilation
    def _engine_descriptor(self, batch):
        lin = []
        for coef, core in self._linear_cores:
            kind = getattr(core, "_engine_kind", None)
            if kind == "deriv":
                kind = ("deriv", core.dim_index, core.order)
            elif kind not in ("lap", "biharm", "unit"):
                raise ValueError(f"term not supported by engine: linear core {type(core).__name__}")
            lin.append((kind, _coef_value(coef, batch)))
        prod, prod_coef, sources = None, 0.0, []
        for coef, core in self._nonlinear_funcs:
            kind = getattr(core, "_engine_kind", None)
            if kind == "src":
                c = _coef_value(coef, batch)
                if isinstance(c, list):
                    raise ValueError("term not supported by engine: per-sample coefficient on an explicit source")
                sources.append((c, core._source_spatial))
            elif kind in ("conv", "vc", "nsp", "ks", "ks_keep_mean", "cons"):
                if prod is not None:
                    raise ValueError("term not supported by engine: more than one product term")
                if not core._dealiasing_swtich:
                    raise ValueError("term not supported by engine: product term without dealiasing")
                c = _coef_value(coef, batch)
                if isinstance(c, list):
                    raise ValueError("term not supported by engine: per-sample coefficient on a product term")
                prod, prod_coef = kind, c
            else:
                raise ValueError(f"term not supported by engine: nonlinear core {type(core).__name__}")
        if sources and prod is None:
            raise ValueError("term not supported by engine: explicit sources need a product term")
        return lin, prod, prod_coef, sources

    def _engine_forces(self):
        """Explicit sources of NSPressureConvection's external force, as (coefficient, field)."""
        out = []
        for _, core in self._nonlinear_funcs:
            out += list(getattr(core, "_engine_force", ()))
        return out

    def _engine(self, batch: int, device) -> SpectralPlan:
        cache = self._state_dict.get("engine")
        if cache is None:
            cache = {}
            self._state_dict["engine"] = cache
        key = (int(batch), str(device))
        plan = cache.get(key)
        if plan is None:
            lin, prod, prod_coef, sources = self._engine_descriptor(batch)
            f_mesh = self._state_dict["f_mesh"]
            try:
                plan = SpectralPlan(f_mesh.mesh_info, batch, self._state_dict["n_channel"], nonlinear=prod,
                                    nonlinear_coef=prod_coef, linear=lin, device=device,
                                    dealias_rate=self._de_aliasing_rate)
                for c, src in sources:
                    plan.add_source(src, c)
                for c, f in self._engine_forces():
                    plan.add_force(f, c)
            except torch.cuda.OutOfMemoryError as e:
                raise RuntimeError(os.linesep.join([
                    "Cuda out of memory when building the integrator.",
                    "Original error message: {}".format(str(e)),
                    "Please try to use a smaller mesh or a low-order integrator.",
                ]))
            cache[key] = plan
        return plan

    def _scheme(self):
        """_build_integrator (_base.py:441-526): 'auto' = ETDRK0 if linear else SETDRK4."""
        solver = self._integrator
        if solver == "auto":
            solver = ETDRKIntegrator.ETDRK0 if self.is_linear else SETDRKIntegrator.SETDRK4
        if isinstance(solver, RKIntegrator):
            raise ValueError("term not supported by engine: explicit Runge-Kutta integrators are not ETD schemes")
        self._is_etdrk_integrator = True
        if solver == ETDRKIntegrator.ETDRK0:
            assert self.is_linear, "The ETDRK0 integrator only supports linear term"
        cfg = dict(self._integrator_config)
        n_points = int(cfg.pop("n_integration_points", 16))
        radius = float(cfg.pop("integration_radius", 1.0))
        cfg.pop("cpu_cached", None)
        if cfg:
            raise ValueError(f"unknown integrator options {sorted(cfg)}")
        return solver.value, n_points, radius

    def _prepare(self, batch, device, dt):
        plan = self._engine(batch, device)
        name, n_points, radius = self._scheme()
        want = (name, float(dt), n_points, radius)
        if getattr(plan, "_scheme_key", None) != want:
            plan.set_scheme(name, dt, n_points, radius)
            plan._scheme_key = want
        self._state_dict["integrator"] = plan
        return plan

    def _pre_check(self, u=None, u_fft=None, mesh=None) -> Tuple[FourierMesh, int]:
        """_base.py:528-579."""
        if u_fft is None and u is None:
            raise ValueError("Either u or u_fft should be given")
        if u_fft is not None and u is not None:
            assert u.shape == u_fft.shape, "The shape of u and u_fft should be the same"
        assert mesh is not None, "Mesh should be given"
        value_device = u.device if u is not None else u_fft.device
        value_dtype = u.dtype if u is not None else u_fft.dtype
        if not isinstance(mesh, FourierMesh):
            if not isinstance(mesh, MeshGrid):
                mesh = FourierMesh(mesh, device=value_device, dtype=value_dtype)
            else:
                mesh = FourierMesh(mesh)
        n_channel = u.shape[1] if u is not None else u_fft.shape[1]
        value_shape = u.shape if u is not None else u_fft.shape
        assert len(value_shape) == mesh.n_dim + 2, (
            f"the value shape {value_shape} is not compatible with mesh dim {mesh.n_dim}"
        )
        for i in range(mesh.n_dim):
            assert value_shape[i + 2] == mesh.mesh_info[i][2], (
                f"Expect to have {mesh.mesh_info[i][2]} points in dim {i} but got {value_shape[i+2]}"
            )
        assert value_device == mesh.device, (
            "The device of mesh {} and the device of value {} are not the same".format(mesh.device, value_device)
        )
        assert self._value_mesh_check_func(len(value_shape) - 2, mesh.n_dim), "Value and mesh do not match the requirement"
        return mesh, n_channel

    def _to_half(self, plan, u_0, u_0_fft):
        if u_0_fft is None:
            require_gpu_tensor(u_0, "u_0")
            if u_0.dtype != torch.float32:
                raise TypeError(f"the MI355X engine computes in float32; got u_0 of dtype {u_0.dtype}")
            return plan.forward(u_0)
        require_gpu_tensor(u_0_fft, "u_0_fft")
        if u_0_fft.dtype not in (torch.complex64,):
            raise TypeError(f"the MI355X engine computes in complex64; got u_0_fft of dtype {u_0_fft.dtype}")
        return plan.full_to_half(u_0_fft)

    # ------------------------------------------------------------------ integrate
    def integrate(self, u_0=None, u_0_fft=None, dt: float = 1, step: int = 1, mesh=None, progressive: bool = False,
                  trajectory_recorder: Optional[_TrajRecorder] = None, return_in_fourier: bool = False):
        """OperatorLike.integrate (_base.py:676-751) on the engine."""
        if self._state_dict["f_mesh"] is None or mesh is not None:
            mesh, n_channel = self._pre_check(u=u_0, u_fft=u_0_fft, mesh=mesh)
            self.register_mesh(mesh, n_channel)
        else:
            self._pre_check(u=u_0, u_fft=u_0_fft, mesh=self._state_dict["f_mesh"])
        x = u_0 if u_0 is not None else u_0_fft
        require_gpu_tensor(x, "u_0" if u_0 is not None else "u_0_fft")
        step = int(step)
        if trajectory_recorder is not None and step <= 0:
            raise UnboundLocalError("cannot access local variable 'i' where it is not associated with a value")
        if isinstance(self._integrator, RKIntegrator):
            if torch.is_grad_enabled() and self._grad_requested(u_0, u_0_fft):
                raise ValueError("term not supported by engine: gradients through Runge-Kutta integrators")
            return self._integrate_rk(u_0, u_0_fft, dt, step, progressive, trajectory_recorder, return_in_fourier)
        if torch.is_grad_enabled() and self._grad_requested(u_0, u_0_fft):
            return self._integrate_with_adjoint(u_0, u_0_fft, dt, step, trajectory_recorder, return_in_fourier)
        plan = self._prepare(x.shape[0], x.device, dt)
        try:
            uh = self._to_half(plan, u_0, u_0_fft)
            if trajectory_recorder is None:
                self._run_steps(plan, uh, step, progressive)
            elif isinstance(trajectory_recorder, CPURecorder):
                sched = [bool(trajectory_recorder.control_func(i)) for i in range(step + 1)]
                self._integrate_streamed(plan, uh, step, sched, trajectory_recorder, return_in_fourier)
            elif isinstance(trajectory_recorder, AutoRecorder):
                sched = [bool(trajectory_recorder.control_func(i)) for i in range(step + 1)]
                slots, t = [], 0
                for s in sched:
                    slots.append(t if s else -1)
                    t += int(s)
                if t == 0:
                    self._run_steps(plan, uh, step, progressive)
                else:
                    rec = torch.tensor(slots, dtype=torch.int32, device=plan.device)
                    traj = torch.zeros((plan.batch, t) + tuple(plan.spec_shape[1:]), dtype=torch.complex64,
                                       device=plan.device)
                    plan.step(uh, step, rec, traj, rec_slot_host=slots)
                    trajectory_recorder._take_engine_frames(plan, traj)
            else:
                # generic recorder protocol: full complex frames handed to record() at recorded steps
                sched = [bool(trajectory_recorder.control_func(i)) for i in range(step + 1)]
                done = 0
                for i in range(step + 1):
                    if sched[i]:
                        self._run_steps(plan, uh, i - done, False)
                        done = i
                        trajectory_recorder.record(i, plan.half_to_full(uh))
                self._run_steps(plan, uh, step - done, False)
        except torch.cuda.OutOfMemoryError as e:
            raise RuntimeError(os.linesep.join([
                "Cuda out of memory when integrating the operator.",
                "Original error message: {}".format(str(e)),
                "Please try to use a smaller mesh or a low-order integrator.",
            ]))
        if trajectory_recorder is not None:
            trajectory_recorder.return_in_fourier = return_in_fourier
            return trajectory_recorder.trajectory
        if return_in_fourier:
            return plan.half_to_full(uh)
        return plan.inverse(uh)

    # device bytes of one half of the frame ring used when frames stream to host memory
    _stream_ring_bytes = 256 << 20

    def _integrate_streamed(self, plan, uh, step, sched, recorder, return_in_fourier):
        """CPURecorder (traj_recorder.py:127-148) with the frames streamed to pinned host memory while
        the stepping continues: the steps are cut into chunks of F recorded frames; each chunk records
        into one half of a device ring, is transformed on the GPU (inverse, or the full spectrum with
        return_in_fourier) and copied to host on a second stream, overlapping the next chunk's steps.
        The ring half is reused only after its copy has finished (event wait on the compute stream)."""
        frame_steps = [i for i, s in enumerate(sched) if s]
        T = len(frame_steps)
        if T == 0:
            self._run_steps(plan, uh, step, False)
            recorder._take_host_frames(None, return_in_fourier)
            return
        B, dev = plan.batch, plan.device
        half_shape = tuple(plan.spec_shape[1:])
        fe = 1
        for d in half_shape:
            fe *= int(d)
        F = max(1, min(T, self._stream_ring_bytes // max(1, B * fe * 8)))
        ring = [torch.empty(B * F * fe, dtype=torch.complex64, device=dev) for _ in range(2)]
        # every chunk's record schedule, uploaded once: chunk k covers frames [t0, t1), steps [s0, s_end]
        chunks, slots = [], []
        s0, t0 = 0, 0
        while t0 < T:
            t1 = min(T, t0 + F)
            s_end = frame_steps[t1 - 1]
            sl = [-1] * (s_end - s0 + 1)
            for t in range(t0, t1):
                sl[frame_steps[t] - s0] = t - t0
            chunks.append((s0, s_end, t0, t1, len(slots)))
            slots += sl
            s0, t0 = s_end, t1
        rec_all = torch.tensor(slots, dtype=torch.int32).to(dev)
        compute = torch.cuda.current_stream(dev)
        copy_stream = torch.cuda.Stream(device=dev)
        out_shape = (plan.channels,) + tuple(plan.n)
        host = torch.empty((B, T) + out_shape, dtype=torch.complex64 if return_in_fourier else torch.float32,
                           pin_memory=True)
        busy = [None, None]
        for k, (s0, s_end, t0, t1, off) in enumerate(chunks):
            h, nf = k % 2, t1 - t0
            if busy[h] is not None:
                compute.wait_event(busy[h][0])
            frames = ring[h][: B * nf * fe].view((B, nf) + half_shape)
            plan.step(uh, s_end - s0, rec_all[off:off + s_end - s0 + 1], frames,
                      rec_slot_host=slots[off:off + s_end - s0 + 1])
            flat = frames.reshape((B * nf,) + half_shape)
            out = plan.half_to_full(flat) if return_in_fourier else plan.inverse(flat)
            out = out.view((B, nf) + tuple(out.shape[1:]))
            ready = torch.cuda.Event()
            ready.record(compute)
            with torch.cuda.stream(copy_stream):
                copy_stream.wait_event(ready)
                for b in range(B):  # host[b, t0:t1] is contiguous
                    host[b, t0:t1].copy_(out[b], non_blocking=True)
                done = torch.cuda.Event()
                done.record(copy_stream)
            busy[h] = (done, out)  # keeps `out` alive until its copy has finished
        if step > frame_steps[-1]:
            self._run_steps(plan, uh, step - frame_steps[-1], False)
        copy_stream.synchronize()
        recorder._take_host_frames(host, return_in_fourier)

    def _integrate_rk(self, u_0, u_0_fft, dt, step, progressive, trajectory_recorder, return_in_fourier):
        """Explicit Runge-Kutta stepping (_rk.py:43-58 with the operator L u + N(u) of
        _base.py:408-439): each stage is one tfsm_apply on the engine, the stage and update sums
        are tfsm_rk_combine."""
        from ..integrator import rk_tableau

        cfg = dict(self._integrator_config)
        name = self._integrator.name
        if cfg and name in ("Euler", "Midpoint", "RK4", "RK4_38Rule"):  # constructors without options
            raise TypeError(f"{name}.__init__() got an unexpected keyword argument '{next(iter(cfg))}'")
        if cfg.pop("adaptive", False):
            raise ValueError("term not supported by engine: adaptive Runge-Kutta steps")
        cfg.pop("atol", None)
        cfg.pop("rtol", None)
        if cfg:
            raise TypeError(f"{name}.__init__() got an unexpected keyword argument '{next(iter(cfg))}'")
        ca, b = rk_tableau(self._integrator)
        x = u_0 if u_0 is not None else u_0_fft
        plan = self._engine(x.shape[0], x.device)
        self._state_dict["integrator"] = plan
        uh = self._to_half(plan, u_0, u_0_fft)
        sched = [bool(trajectory_recorder.control_func(i)) for i in range(step + 1)] if trajectory_recorder else None
        it = range(step)
        if progressive:
            it = tqdm(it, desc="Integrating")
        for i in it:
            if sched and sched[i]:
                trajectory_recorder.record(i, plan.half_to_full(uh))
            ks = [plan.apply(uh)]
            for s in range(1, len(b)):
                a = ca[s - 1][1:]
                ks.append(plan.apply(plan.rk_combine(uh, ks[:len(a)], a, dt)))
            uh = plan.rk_combine(uh, ks, b, dt)
        if sched and sched[step]:
            trajectory_recorder.record(step, plan.half_to_full(uh))
        if trajectory_recorder is not None:
            trajectory_recorder.return_in_fourier = return_in_fourier
            return trajectory_recorder.trajectory
        if return_in_fourier:
            return plan.half_to_full(uh)
        return plan.inverse(uh)

    def _grad_requested(self, u_0, u_0_fft) -> bool:
        for t in (u_0, u_0_fft):
            if isinstance(t, Tensor) and t.requires_grad:
                return True
        for coef, _ in list(self._linear_cores) + list(self._nonlinear_funcs):
            if isinstance(coef, Tensor) and coef.requires_grad:
                return True
        return False

    def _integrate_with_adjoint(self, u_0, u_0_fft, dt, step, trajectory_recorder, return_in_fourier):
        """integrate() under autograd: the engine's hand-written adjoint with checkpoints every
        `adjoint_checkpoint_every` steps (default 10) replaces autograd through the step loop."""
        from ..adjoint import ADJOINT_SCHEMES, IntegrateAdjoint

        if trajectory_recorder is not None or return_in_fourier or u_0_fft is not None or u_0 is None:
            raise ValueError("term not supported by engine: gradients are available for integrate(u_0, ...) returning "
                             "the final field (no trajectory recorder, u_0_fft or return_in_fourier)")
        name = self._scheme()[0]
        if name not in ADJOINT_SCHEMES:
            raise ValueError(f"term not supported by engine: gradients need one of {ADJOINT_SCHEMES} "
                             f"(the operator uses {name}); call set_integrator")
        for coef, _ in self._nonlinear_funcs:
            if isinstance(coef, Tensor) and coef.requires_grad:
                raise ValueError("term not supported by engine: gradient with respect to a nonlinear coefficient")
        grad_terms = [i for i, (coef, _) in enumerate(self._linear_cores)
                      if isinstance(coef, Tensor) and coef.requires_grad]
        if len(grad_terms) > 1:
            raise ValueError("term not supported by engine: gradients with respect to more than one linear coefficient")
        gt = grad_terms[0] if grad_terms else -1
        theta = self._linear_cores[gt][0] if gt >= 0 else torch.zeros((), device=u_0.device)
        plan = self._prepare(u_0.shape[0], u_0.device, dt)
        every = int(getattr(self, "adjoint_checkpoint_every", 10))
        return IntegrateAdjoint.apply(u_0.to(torch.float32), theta, plan, int(step), max(1, every), gt)

    @staticmethod
    def _run_steps(plan, uh, step, progressive):
        if step <= 0:
            return
        if not progressive:
            if step >= 2 * GRAPH_STEPS and _graph_worthwhile(plan):
                # several launches per step on a small mesh: replay a captured graph of GRAPH_STEPS
                # steps (kept on the plan, over its own state buffer) instead of launching each kernel
                cache = getattr(plan, "_step_graph", None)
                if cache is None:
                    buf = plan.empty_spectrum()
                    cache = plan._step_graph = (buf, plan.graph(buf, GRAPH_STEPS, keep_plan=False))
                buf, graph = cache
                buf.copy_(uh)
                for _ in range(step // GRAPH_STEPS):
                    graph.launch()
                uh.copy_(buf)
                plan.step(uh, step % GRAPH_STEPS)
                return
            plan.step(uh, step)
            return
        chunk = max(1, step // 100)
        with tqdm(total=step, desc="Integrating") as bar:
            done = 0
            while done < step:
                k = min(chunk, step - done)
                plan.step(uh, k)
                torch.cuda.current_stream(plan.device).synchronize()
                done += k
                bar.update(k)

    def __call__(self, u=None, u_fft=None, mesh=None, return_in_fourier=False):
        """L u^ + N(u^) (_base.py:753-790) on the engine."""
        if self._state_dict["f_mesh"] is None or mesh is not None:
            mesh, n_channel = self._pre_check(u, u_fft, mesh)
            self.register_mesh(mesh, n_channel)
        else:
            self._pre_check(u=u, u_fft=u_fft, mesh=self._state_dict["f_mesh"])
        if not self._linear_cores and not self._nonlinear_funcs:
            raise ValueError("Both linear coefficient and nonlinear function are None. Cannot build operator.")
        x = u if u is not None else u_fft
        require_gpu_tensor(x, "u" if u is not None else "u_fft")
        cores = list(self._linear_cores) + list(self._nonlinear_funcs)
        if len(cores) == 1 and getattr(cores[0][1], "_engine_map", None):
            from ._diag import map_call

            coef, core = cores[0]
            return map_call(core, coef, u, u_fft, self._state_dict["f_mesh"], return_in_fourier)
        if any(getattr(core, "_engine_map", None) and not getattr(core, "_engine_kind", None) for _, core in cores):
            raise ValueError("term not supported by engine: a channel-changing operator (Div, Grad, Curl, "
                             "Vorticity2Velocity, pressure) combined with other terms")
        plan = self._engine(x.shape[0], x.device)
        uh = self._to_half(plan, u, u_fft)
        out = plan.apply(uh)
        if return_in_fourier:
            return plan.half_to_full(out)
        return plan.inverse(out)

    def to(self, device=None, dtype=None):
        if self._state_dict is not None and self._state_dict["f_mesh"] is not None:
            self._state_dict["f_mesh"].to(device=device, dtype=dtype)
            self.register_mesh(self._state_dict["f_mesh"], self._state_dict["n_channel"])


class _InverseSolveMixin:
    """LinearOperator.solve (_base.py:209-262): x^ = b^ / L with 1/0 := 1."""

    def solve(self, b=None, b_fft=None, mesh=None, n_channel=None, return_in_fourier=False):
        if not (mesh is not None and n_channel is not None):
            assert self._state_dict["f_mesh"] is not None, "Mesh and n_channel should be given when calling solve"
        if not (mesh is None and n_channel is None):
            mesh = self._state_dict["f_mesh"] if mesh is None else mesh
            n_channel = self._state_dict["n_channel"] if n_channel is None else n_channel
            self.register_mesh(mesh, n_channel)
        x = b if b is not None else b_fft
        require_gpu_tensor(x, "b" if b is not None else "b_fft")
        plan = self._engine(x.shape[0], x.device)
        uh = self._to_half(plan, b, b_fft)
        out = plan.solve(uh)
        if return_in_fourier:
            return plan.half_to_full(out)
        return plan.inverse(out)


class _DeAliasMixin:
    def set_de_aliasing_rate(self, de_aliasing_rate: float):
        """_base.py:265-281."""
        self._de_aliasing_rate = de_aliasing_rate
        self._state_dict = {k: None for k in self._state_dict}


class Operator(OperatorLike, _DeAliasMixin):
    """Linear and nonlinear terms (_base.py:807-850)."""

    def __init__(self, operator_generators=None, coefs: Optional[List] = None) -> None:
        super().__init__(operator_generators, coefs)

    def __add__(self, other):
        if isinstance(other, OperatorLike):
            return Operator(self.operator_generators + other.operator_generators, self.coefs + other.coefs)
        if isinstance(other, Tensor):
            return Operator(self.operator_generators + [lambda f_mesh, n_channel: _ExplicitSourceCore(other)],
                            self.coefs + [1])
        return NotImplemented

    def __mul__(self, other):
        if isinstance(other, OperatorLike):
            return NotImplemented
        return Operator(self.operator_generators, [coef * other for coef in self.coefs])

    def __neg__(self):
        return Operator(self.operator_generators, [-1 * coef for coef in self.coefs])


class LinearOperator(OperatorLike, _InverseSolveMixin):
    """Linear terms only (_base.py:853-919)."""

    def __init__(self, linear_coef=None, coefs: Optional[List] = None) -> None:
        if not isinstance(linear_coef, list):
            linear_coef = [linear_coef]
        super().__init__(
            operator_generators=[
                (c if not isinstance(c, LinearCoef) else (lambda f_mesh, n_channel, _c=c: _c)) for c in linear_coef
            ],
            coefs=coefs,
        )

    @property
    def is_linear(self):
        return True

    def __add__(self, other):
        if isinstance(other, LinearOperator):
            return LinearOperator(self.operator_generators + other.operator_generators, self.coefs + other.coefs)
        if isinstance(other, OperatorLike):
            return Operator(self.operator_generators + other.operator_generators, self.coefs + other.coefs)
        if isinstance(other, Tensor):
            return Operator(self.operator_generators + [lambda f_mesh, n_channel: _ExplicitSourceCore(other)],
                            self.coefs + [1])
        return NotImplemented

    def __mul__(self, other):
        if isinstance(other, OperatorLike):
            return NotImplemented
        return LinearOperator(self.operator_generators, [coef * other for coef in self.coefs])

    def __neg__(self):
        return LinearOperator(self.operator_generators, [-1 * coef for coef in self.coefs])


class NonlinearOperator(OperatorLike, _DeAliasMixin):
    """Nonlinear terms only (_base.py:922-988)."""

    def __init__(self, nonlinear_func=None, coefs: Optional[List] = None) -> None:
        if not isinstance(nonlinear_func, list):
            nonlinear_func = [nonlinear_func]
        super().__init__(
            operator_generators=[
                (f if not isinstance(f, NonlinearFunc) else (lambda f_mesh, n_channel, _f=f: _f)) for f in nonlinear_func
            ],
            coefs=coefs,
        )

    @property
    def is_linear(self):
        return False

    def __add__(self, other):
        if isinstance(other, NonlinearOperator):
            return NonlinearOperator(self.operator_generators + other.operator_generators, self.coefs + other.coefs)
        if isinstance(other, OperatorLike):
            return Operator(self.operator_generators + other.operator_generators, self.coefs + other.coefs)
        if isinstance(other, Tensor):
            return Operator(self.operator_generators + [lambda f_mesh, n_channel: _ExplicitSourceCore(other)],
                            self.coefs + [1])
        return NotImplemented

    def __mul__(self, other):
        if isinstance(other, OperatorLike):
            return NotImplemented
        return NonlinearOperator(self.operator_generators, [coef * other for coef in self.coefs])

    def __neg__(self):
        return NonlinearOperator(self.operator_generators, [-1 * coef for coef in self.coefs])


class _ExplicitSourceCore(NonlinearFunc):
    """Constant source spectrum (_base.py:994-1015); the engine transforms the spatial source on
    the device when the operator is compiled."""

    _engine_kind = "src"

    def __init__(self, source: Tensor) -> None:
        super().__init__(dealiasing_swtich=False)
        self._source_spatial = source

    def __call__(self, u_fft, f_mesh: FourierMesh, u=None):
        src = self._source_spatial.to(f_mesh.device)
        return f_mesh.fft(src)


class ExplicitSource(NonlinearOperator):
    """Explicit source term (_base.py:1018-1028)."""

    def __init__(self, source: Tensor) -> None:
        super().__init__(_ExplicitSourceCore(source))
