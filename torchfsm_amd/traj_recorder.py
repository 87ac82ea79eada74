"""Trajectory recorders with the reference's protocol (traj_recorder.py:7-148).

A recorder decides per step (control_func) whether the state before that step is kept.  The
engine evaluates the schedule once per integrate call, records half-spectrum frames on the device
inside the step loop, and the recorder turns them into the reference's output on demand:
``trajectory`` = ifftn(stack(frames, dim=1)).real of shape (B, T, C, N...) or, with
``return_in_fourier``, the stacked full complex spectra.
"""

import copy
from typing import Callable, Optional

import torch

from .utils import default


class IntervalController:
    """Record when step >= start and (step - start) % interval == 0 (traj_recorder.py:7-22)."""

    def __init__(self, interval: int = 1, start: int = 0) -> None:
        self.start = start
        self.interval = interval

    def __call__(self, step: int) -> bool:
        return step >= self.start and (step - self.start) % self.interval == 0


class _TrajRecorder:
    """Base recorder (traj_recorder.py:24-93)."""

    def __init__(self, control_func: Optional[Callable[[int], bool]] = None, include_initial_state: bool = True):
        control_func = default(control_func, lambda step: True)
        if include_initial_state:
            self.control_func = control_func
        else:
            self.control_func = lambda step: False if step == 0 else control_func(step)
        self.return_in_fourier = False

    def record(self, step: int, frame: torch.Tensor):
        if self.control_func(step):
            self._record(step, frame)

    def _record(self, step: int, frame: torch.Tensor):
        raise NotImplementedError

    @property
    def trajectory(self):
        raise NotImplementedError


class AutoRecorder(_TrajRecorder):
    """Keeps frames on the simulation device (traj_recorder.py:95-125)."""

    def __init__(self, control_func: Optional[Callable[[int], bool]] = None, include_initial_state: bool = True):
        super().__init__(control_func, include_initial_state)
        self._trajectory = []
        self._engine_frames = None  # (plan, half-spectrum frames [B][T][C][...][P])

    # engine fast path: frames arrive in bulk as half spectra
    def _take_engine_frames(self, plan, frames: torch.Tensor):
        if self._trajectory:
            raise RuntimeError("The recorder already holds frames from another integration.")
        self._engine_frames = (plan, frames)

    def _record(self, step: int, frame: torch.Tensor):
        if isinstance(self._trajectory, torch.Tensor):
            raise RuntimeError("The trajectory has been finalized.")
        self._trajectory.append(copy.deepcopy(frame))

    def _frames_out(self, plan, frames):
        B, T = frames.shape[0], frames.shape[1]
        flat = frames.reshape((B * T,) + tuple(frames.shape[2:]))
        if self.return_in_fourier:
            out = plan.half_to_full(flat)
        else:
            out = plan.inverse(flat)
        return out.reshape((B, T) + tuple(out.shape[1:]))

    @property
    def trajectory(self):
        if self._engine_frames is not None:
            return self._frames_out(*self._engine_frames)
        if len(self._trajectory) == 0:
            return None
        stacked = torch.stack(self._trajectory, dim=1)
        if self.return_in_fourier:
            return stacked
        # frames handed over through the generic record() protocol: full complex spectra
        from .mesh import FourierMesh

        n_dim = stacked.dim() - 3
        shape = tuple(stacked.shape[-n_dim:])
        fm = FourierMesh([(0.0, 1.0, n) for n in shape], device=stacked.device)
        return fm.ifft(stacked).real


class CPURecorder(AutoRecorder):
    """Keeps frames in host memory (traj_recorder.py:127-148).  With the engine the frames are
    transformed on the GPU and streamed to pinned host memory while the stepping continues
    (OperatorLike._integrate_streamed); ``trajectory`` returns that CPU tensor."""

    def __init__(self, control_func: Optional[Callable[[int], bool]] = None, include_initial_state: bool = True):
        super().__init__(control_func, include_initial_state)
        self._host = None  # (frames (B, T, ...) on the host, already in the requested space)

    def _take_host_frames(self, frames: Optional[torch.Tensor], in_fourier: bool):
        if self._trajectory or self._engine_frames is not None:
            raise RuntimeError("The recorder already holds frames from another integration.")
        self._host = (frames, in_fourier)

    @property
    def trajectory(self):
        if self._host is not None:
            frames, in_fourier = self._host
            if frames is None or in_fourier == self.return_in_fourier:
                return frames
            raise RuntimeError("return_in_fourier differs from the integration that recorded the frames")
        return super().trajectory

    def _take_engine_frames(self, plan, frames: torch.Tensor):
        host = torch.empty(frames.shape, dtype=frames.dtype, pin_memory=True)
        host.copy_(frames, non_blocking=False)
        super()._take_engine_frames(plan, host)

    def _record(self, step: int, frame: torch.Tensor):
        self._trajectory.append(copy.deepcopy(frame) if frame.is_cpu else frame.cpu())

    def _frames_out(self, plan, frames):
        B, T = frames.shape[0], frames.shape[1]
        outs = []
        chunk = max(1, T // 8)
        for t0 in range(0, T, chunk):
            dev = frames[:, t0:t0 + chunk].to(plan.device)
            outs.append(super()._frames_out(plan, dev).cpu())
        return torch.cat(outs, dim=1)
