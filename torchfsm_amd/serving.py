"""Serving loop for mixed ETD integration requests: continuous batching over shape classes.

A request asks for `steps` steps of one PDE from its own initial field and parameter:

  kind "ks1d"       KuramotoSivashinsky() on [0, L), N in {512..4096}; param = L      (pde.py:27-39)
  kind "ns2d"       NavierStokesVorticity(Re, kolm_force(y)) on [0, 2 pi)^2; param = Re (pde.py:66-83)
  kind "burgers3d"  Burgers(nu), 3 channels on [0, 1)^3; param = nu                   (pde.py:13-25)

Requests of one (kind, N) form a shape class.  A class owns one engine plan with a fixed number
of sample slots and per-sample coefficients: admitting a request re-targets one slot
(tfsm_plan_set_sample: its L and ETD coefficients, and its axis table for 1D lengths) and writes
the forward transform of its field into that slot.  Each class steps on its own HIP stream in
chunks, and every slot advances by its own count within a chunk (tfsm_step_varying): a slot whose
request finishes inside the chunk stops there, and an empty slot is not stepped at all.  1D classes
run one persistent launch per chunk; 2D and 3D classes replay a captured HIP graph of 2^k steps
whose kernels read the per-slot counts from device memory.  A chunk is at most MAX_CHUNK steps
(and at most the longest remaining request), so queued requests wait at most one chunk for a free
slot.  Between chunks, finished slots are harvested asynchronously: the inverse transform and the
copy into a pinned host buffer are queued on the class stream behind an event, and the result is
handed out when the host sees that event complete, so the scheduler thread never blocks on a
copy.  One host thread drives all classes.

Results do not depend on the mix: samples never interact, launch shapes are fixed per shape class
(never per occupancy), and every slot's coefficients come from the same kernels.  A request
served alone and the same request served among others return bit-identical fields.

Memory: each class reserves its plan, state and workspace at creation, charged against an HBM
budget; a request whose class cannot be created within the budget waits until an idle class is
released.
"""

import math
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import torch

from .engine import SpectralPlan, StepGraph

SPECS = {
    # dims, channels, dt, scheme, domain length (placeholder for ks1d)
    "ks1d": (1, 1, 0.1, "SETDRK4", 32 * math.pi),
    "ns2d": (2, 1, 1e-3, "SETDRK4", 2 * math.pi),
    "burgers3d": (3, 3, 2e-3, "SETDRK2", 1.0),
}

DEFAULT_SLOTS = {
    ("ks1d", 512): 64, ("ks1d", 1024): 64, ("ks1d", 2048): 32, ("ks1d", 4096): 32,
    ("ns2d", 64): 32, ("ns2d", 128): 32, ("ns2d", 256): 16, ("ns2d", 512): 16,
    ("burgers3d", 32): 16, ("burgers3d", 64): 8, ("burgers3d", 128): 4,
}

MAX_GRAPH_LOG2 = 6  # graphs of 1, 2, 4, ..., 64 steps
MAX_CHUNK = 1 << MAX_GRAPH_LOG2  # steps per chunk at most: bounds the wait of a queued request


@dataclass
class Request:
    rid: int
    kind: str
    n: int
    steps: int
    param: float
    u0: torch.Tensor          # (1, C, N...) float32, host or device
    arrival: float = 0.0      # seconds after the start of the run


@dataclass
class Result:
    rid: int
    u: torch.Tensor           # (1, C, N...) float32 on the host
    arrival: float
    admitted: float
    completed: float
    key: Tuple[str, int] = ("", 0)  # shape class (kind, N)

    @property
    def latency(self) -> float:
        return self.completed - self.arrival

    @property
    def queue(self) -> float:
        """Time from arrival to admission into a slot."""
        return self.admitted - self.arrival

    @property
    def service(self) -> float:
        """Time from admission until the result is on the host."""
        return self.completed - self.admitted


def linear_coefs(kind: str, param: float) -> List[float]:
    """Linear coefficients in generator order, as the PDE factories build them."""
    if kind == "ks1d":
        return [-1.0, -1.0]               # -Laplacian() - Biharmonic()
    if kind == "ns2d":
        return [1.0 / param, -0.1]        # 1/Re * Laplacian(), kolm_force drag * ImplicitSource()
    if kind == "burgers3d":
        return [float(param)]             # nu * Laplacian()
    raise ValueError(f"unknown request kind {kind}")


def _class_operator(kind: str, n: int, device):
    from . import field, pde
    from .mesh import FourierMesh, MeshGrid

    dims, ch, _, _, length = SPECS[kind]
    mesh_info = [(0, length, n)] * dims
    if kind == "ks1d":
        op = pde.KuramotoSivashinsky()
    elif kind == "ns2d":
        y = MeshGrid(mesh_info, device=device).bc_mesh_grid()[1]
        op = pde.NavierStokesVorticity(1000.0, field.kolm_force(y))
    else:
        op = pde.Burgers(0.01)
    op.register_mesh(FourierMesh(mesh_info, device=device), ch)
    return op, mesh_info


class ShapeClass:
    """One plan with `slots` sample slots for requests of one (kind, N)."""

    def __init__(self, kind: str, n: int, slots: int, device):
        self.kind, self.n, self.slots = kind, int(n), int(slots)
        self.device = torch.device(device)
        dims, ch, dt, scheme, length = SPECS[kind]
        self.dims, self.channels, self.dt, self.scheme = dims, ch, dt, scheme
        op, mesh_info = _class_operator(kind, n, self.device)
        lin, prod, pc, sources = op._engine_descriptor(1)
        lin = [(k, [float(c)] * self.slots) for k, c in lin]
        self.plan = SpectralPlan(mesh_info, self.slots, ch, nonlinear=prod, nonlinear_coef=pc, linear=lin,
                                 device=self.device,
                                 lengths_per_sample=[length] * self.slots if kind == "ks1d" else None)
        for c, src in sources:
            self.plan.add_source(src, c)
        self.plan.set_scheme(scheme, dt)
        self.stream = torch.cuda.Stream(self.device)
        self.uh = self.plan.empty_spectrum()
        self.plan.workspace()
        self.dof = ch * n ** dims
        self.req: List[Optional[Request]] = [None] * self.slots
        self.admitted = [0.0] * self.slots
        self.left = [0] * self.slots
        self.busy = False
        self.event = None
        self.graphs: Dict[int, StepGraph] = {}
        self.counts = torch.zeros(self.slots, dtype=torch.int32, device=self.device)
        # per-slot step counts of the next chunk: staged in pinned host memory, copied on the stream
        self.counts_host = torch.zeros(self.slots, dtype=torch.int32, pin_memory=True)
        # harvest buffers: inverse transforms on the device, their copies in pinned host memory
        self.out_dev = torch.empty((self.slots, ch) + (self.n,) * dims, dtype=torch.float32, device=self.device)
        self.out_host = torch.empty((self.slots, ch) + (self.n,) * dims, dtype=torch.float32, pin_memory=True)
        self.pending: List[Tuple[torch.cuda.Event, List[Tuple[int, Request, float]]]] = []
        self.last_used = time.perf_counter()
        self.reserved_bytes = self.bytes_estimate(kind, n, slots)
        # per-class counters: chunks, steps, slot-steps of active requests, device time (ms)
        self.stats = {"chunks": 0, "steps": 0, "slot_steps": 0, "gpu_ms": 0.0}
        self.ev0 = None
        torch.cuda.synchronize(self.device)

    @staticmethod
    def bytes_estimate(kind: str, n: int, slots: int) -> int:
        dims, ch = SPECS[kind][0], SPECS[kind][1]
        half = (n // 2 + 1 + 15) // 16 * 16 * n ** (dims - 1) * 8
        factor = {1: 4, 2: 7, 3: 12}[dims]  # state + stage arrays + products + pruned fields
        outputs = 2 * slots * ch * n ** dims * 4  # harvest buffer on the device (and pinned on the host)
        return slots * ch * half * factor + slots * 8 * 4 * half // 8 + outputs + (64 << 20)

    @property
    def nbytes(self) -> int:
        """Device bytes actually held (plan workspace and state; tables excluded)."""
        return self.plan.ws_bytes + self.uh.numel() * 8

    # ------------------------------------------------------------------ slots
    def free_slot(self) -> Optional[int]:
        for b in range(self.slots):
            if self.req[b] is None:
                return b
        return None

    def active(self) -> List[int]:
        return [b for b in range(self.slots) if self.req[b] is not None and self.left[b] > 0]

    def admit(self, b: int, r: Request, now: float):
        length = float(r.param) if self.kind == "ks1d" else 0.0
        with torch.cuda.stream(self.stream):
            self.plan.set_sample(b, linear_coefs(self.kind, r.param), length)
            u0 = r.u0.to(self.device, dtype=torch.float32)
            self.plan.forward(u0, out=self.uh[b:b + 1])
        self.req[b] = r
        self.left[b] = int(r.steps)
        self.admitted[b] = now

    def prewarm_graphs(self):
        if self.dims == 1:
            return
        for k in range(MAX_GRAPH_LOG2 + 1):
            self._graph(1 << k)
        torch.cuda.synchronize(self.device)

    def _graph(self, nsteps: int) -> StepGraph:
        g = self.graphs.get(nsteps)
        if g is None:
            with torch.cuda.stream(self.stream):
                g = StepGraph(self.plan, self.uh, nsteps, counts=self.counts)
            self.graphs[nsteps] = g
        return g

    def launch(self) -> int:
        """Start the next chunk on the class stream; returns its step count.  The chunk is the
        largest power of two not above min(MAX_CHUNK, longest remaining request); each slot
        advances by min(its remaining steps, chunk), empty slots by 0."""
        act = self.active()
        if not act:
            return 0
        longest = min(max(self.left[b] for b in act), MAX_CHUNK)
        k = 1 << (longest.bit_length() - 1)
        counts = [min(self.left[b], k) if self.req[b] is not None else 0 for b in range(self.slots)]
        with torch.cuda.stream(self.stream):
            self.ev0 = torch.cuda.Event(enable_timing=True)
            self.ev0.record(self.stream)
            # the previous chunk's copy of the counts has completed (the host saw its event)
            self.counts_host.copy_(torch.tensor(counts, dtype=torch.int32))
            self.counts.copy_(self.counts_host, non_blocking=True)
            if self.dims == 1:
                self.plan.step_varying(self.uh, self.counts, k)
            else:
                self._graph(k).launch()
            self.event = torch.cuda.Event(enable_timing=True)
            self.event.record(self.stream)
        for b in act:
            self.left[b] -= counts[b]
        self.stats["slot_steps"] += sum(counts)
        self.stats["chunks"] += 1
        self.stats["steps"] += k
        self.busy = True
        self.last_used = time.perf_counter()
        return k

    def harvest(self, now: float):
        """Queue the inverse transform and host copy of every finished slot behind an event, and
        free those slots (later work on the class stream is ordered after the copies)."""
        done = [b for b in range(self.slots) if self.req[b] is not None and self.left[b] == 0]
        if not done:
            return
        with torch.cuda.stream(self.stream):
            for b in done:
                self.out_dev[b:b + 1].copy_(self.plan.inverse(self.uh[b:b + 1]))
                self.out_host[b].copy_(self.out_dev[b], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(self.stream)
        self.pending.append((ev, [(b, self.req[b], self.admitted[b]) for b in done]))
        for b in done:
            self.req[b] = None

    def collect(self, t0: float) -> List[Result]:
        """Results whose host copies have completed (never waits)."""
        out = []
        while self.pending and self.pending[0][0].query():
            _, items = self.pending.pop(0)
            now = time.perf_counter() - t0
            for b, r, adm in items:
                out.append(Result(r.rid, self.out_host[b:b + 1].clone(), r.arrival, adm, now, (self.kind, self.n)))
        return out

    def close(self):
        for g in self.graphs.values():
            g.close()
        self.graphs.clear()
        self.plan.close()


class Server:
    """Continuous-batching scheduler over shape classes, one host thread."""

    def __init__(self, device="cuda", hbm_budget: float = 200e9, slots: Optional[Dict[Tuple[str, int], int]] = None):
        self.device = torch.device(device)
        self.budget = float(hbm_budget)
        self.slots = dict(DEFAULT_SLOTS)
        if slots:
            self.slots.update(slots)
        self.classes: Dict[Tuple[str, int], ShapeClass] = {}

    def _slots_for(self, key):
        return self.slots.get(key, 8)

    def reserved(self) -> int:
        return sum(c.reserved_bytes for c in self.classes.values())

    def get_class(self, key) -> Optional[ShapeClass]:
        c = self.classes.get(key)
        if c is not None:
            return c
        need = ShapeClass.bytes_estimate(key[0], key[1], self._slots_for(key))
        if need > self.budget:
            raise RuntimeError(f"shape class {key} needs {need / 2**30:.2f} GiB, above the HBM budget "
                               f"({self.budget / 2**30:.2f} GiB)")
        while self.reserved() + need > self.budget:
            idle = [k for k, v in self.classes.items()
                    if not v.busy and not v.pending and all(r is None for r in v.req)]
            if not idle:
                return None
            k = min(idle, key=lambda x: self.classes[x].last_used)
            self.classes.pop(k).close()
        c = ShapeClass(key[0], key[1], self._slots_for(key), self.device)
        self.classes[key] = c
        return c

    def prewarm(self, keys: Sequence[Tuple[str, int]]):
        """Create the classes (plans, coefficients, graphs) before a timed run."""
        for k in keys:
            c = self.get_class(tuple(k))
            if c is not None:
                c.prewarm_graphs()

    def run(self, requests: Sequence[Request], poll_sleep: float = 2e-5) -> List[Result]:
        """Serve `requests` (arrival times in seconds from the start); returns results in
        completion order."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        waiting: Dict[Tuple[str, int], deque] = {}
        results: List[Result] = []
        total = len(pending)
        t0 = time.perf_counter()
        while len(results) < total:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                r = pending.popleft()
                if r.kind not in SPECS:
                    raise ValueError(f"unknown request kind {r.kind}")
                waiting.setdefault((r.kind, int(r.n)), deque()).append(r)
            launched = False
            for key in list(set(waiting) | set(self.classes)):
                c = self.classes.get(key)
                if c is not None:
                    results.extend(c.collect(t0))
                if c is not None and c.busy:
                    if not c.event.query():
                        continue
                    c.busy = False
                    c.stats["gpu_ms"] += c.ev0.elapsed_time(c.event)
                    c.harvest(time.perf_counter() - t0)
                q = waiting.get(key)
                if q:
                    if c is None:
                        c = self.get_class(key)
                        if c is None:
                            continue
                    now = time.perf_counter() - t0
                    b = c.free_slot()
                    while q and b is not None:
                        c.admit(b, q.popleft(), now)
                        b = c.free_slot()
                if c is not None and not c.busy and c.launch():
                    launched = True
            if not launched and poll_sleep:
                time.sleep(poll_sleep)
        torch.cuda.synchronize(self.device)
        return results

    def close(self):
        for c in self.classes.values():
            c.close()
        self.classes.clear()


def _pct(vals: Sequence[float], p: float) -> float:
    v = sorted(vals)
    i = min(len(v) - 1, max(0, int(math.ceil(p / 100.0 * len(v))) - 1))
    return v[i]


def latency_summary(results: Sequence[Result]) -> dict:
    lat = [r.latency for r in results]
    if not lat:
        return {}
    return {"p50_ms": _pct(lat, 50) * 1e3, "p99_ms": _pct(lat, 99) * 1e3, "max_ms": max(lat) * 1e3, "n": len(lat)}


def latency_by_class(results: Sequence[Result]) -> dict:
    """Per shape class: request count, p50/p99 of latency, and its split into queueing (arrival to
    admission) and service (admission to the result on the host), in milliseconds."""
    out = {}
    for key in sorted({r.key for r in results}):
        rs = [r for r in results if r.key == key]
        row = {"n": len(rs)}
        for name, get in (("latency", lambda r: r.latency), ("queue", lambda r: r.queue),
                          ("service", lambda r: r.service)):
            vals = [get(r) for r in rs]
            row[f"{name}_p50_ms"] = round(_pct(vals, 50) * 1e3, 2)
            row[f"{name}_p99_ms"] = round(_pct(vals, 99) * 1e3, 2)
        out[f"{key[0]}_{key[1]}"] = row
    return out


def synthetic_requests(n_requests: int, rate: float, seed: int = 0, kinds=("ks1d", "ns2d", "burgers3d"),
                       sizes=None, steps=(100, 1000)) -> List[Request]:
    """Poisson arrivals at `rate` per second, kinds mixed evenly; per-request parameters and
    initial fields are seeded by the request id (see DEFAULT_MIX for the parameter ranges)."""
    from .workloads import fourier_modes_field, ns_initial_vorticity

    sizes = sizes or {"ks1d": (512, 1024, 2048, 4096), "ns2d": (64, 128, 256, 512), "burgers3d": (64,)}
    g = torch.Generator().manual_seed(seed)
    t = 0.0
    out = []
    for rid in range(n_requests):
        t += float(-math.log(1.0 - float(torch.rand(1, generator=g, dtype=torch.float64))) / rate)
        kind = kinds[int(torch.randint(len(kinds), (1,), generator=g))]
        ns = sizes[kind]
        n = int(ns[int(torch.randint(len(ns), (1,), generator=g))])
        st = int(torch.randint(steps[0], steps[1] + 1, (1,), generator=g))
        u = float(torch.rand(1, generator=g, dtype=torch.float64))
        if kind == "ks1d":
            param = 32 * math.pi + u * 96 * math.pi                    # L ~ U[32 pi, 128 pi]
            x = torch.arange(n, dtype=torch.float64) * param / n
            gg = torch.Generator().manual_seed(1000 + rid)
            a = torch.randn(4, generator=gg, dtype=torch.float64)
            u0 = sum(a[i] * torch.cos(2 * math.pi * (i + 1) * x / param * 4) for i in range(4))
            u0 = u0.reshape(1, 1, n).float()
        elif kind == "ns2d":
            param = 1.0 / math.exp(math.log(1e-3) + u * (math.log(5e-2) - math.log(1e-3)))  # nu ~ logU
            u0 = ns_initial_vorticity(n, 1, seed=1000 + rid)
        else:
            param = math.exp(math.log(1e-2) + u * (math.log(5e-2) - math.log(1e-2)))        # nu ~ logU
            u0 = fourier_modes_field(n, 3, 3, 1, seed=1000 + rid, nmodes=4)
        out.append(Request(rid, kind, n, st, param, u0.contiguous(), t))
    return out
