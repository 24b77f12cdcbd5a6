"""Device-resident simulations and the batched engine (C ABI tf_sim_* /
tf_batch_*).

A ``Simulation`` owns two state buffers and a 128-byte control block in HBM;
its t, dt, step count and loop flag live on the device, so stepping never
waits for the host. A ``Batch`` advances any set of simulations, of any
sizes, together: one fused stencil launch per tiling class covers the tiles
of every simulation, then one launch updates every simulation's dt.
``step_graph`` replays captured HIP graphs of 16 such steps, and ``reset``
resets every simulation in three launches.

Each cell's arithmetic is the same whatever else is in the batch, so a
simulation's result is bit-identical alone and in any mix.
"""

from __future__ import annotations

import ctypes
import math
from typing import List, Sequence

import numpy as np
import torch

from . import _lib
from .fluid import _stream


class Simulation:
    def __init__(self, state, grid_spacing: float, gamma: float, *, max_steps: int = -1, t_final: float = math.inf,
                 handle_boundaries: bool = False, c_cfl: float = 0.4, device=None, buffers=None, ctrl=None):
        """``state``: (5, Nx, Ny, Nz) float32 primitive state (numpy or torch).
        ``buffers``: optional pair of preallocated device tensors to use."""
        if device is None:
            device = state.device if isinstance(state, torch.Tensor) and state.is_cuda else torch.device("cuda")
        shape = tuple(state.shape)
        if len(shape) != 4 or shape[0] != 5:
            raise TypeError(f"state must have shape (5, Nx, Ny, Nz), got {shape}")
        if buffers is None:
            buf0 = torch.empty(shape, dtype=torch.float32, device=device)
            buf1 = torch.empty_like(buf0)
        else:
            buf0, buf1 = buffers
        src = torch.from_numpy(np.ascontiguousarray(state, dtype=np.float32)) if isinstance(state, np.ndarray) else state
        if src.dtype != torch.float32:
            raise TypeError(f"state must be float32, got {src.dtype}")
        if not (isinstance(src, torch.Tensor) and src.data_ptr() == buf0.data_ptr()):
            buf0.copy_(src, non_blocking=False)
        self.buf = (buf0, buf1)
        self.ctrl = ctrl if ctrl is not None else torch.zeros(_lib.TF_CTRL_BYTES, dtype=torch.uint8, device=device)
        self.shape = shape
        self.grid_spacing = float(grid_spacing)
        self.gamma = float(gamma)
        self.max_steps = int(max_steps)
        self.t_final = float(t_final)
        self.handle_boundaries = bool(handle_boundaries)
        self.c_cfl = float(c_cfl)
        self.desc = _lib.SimDesc(
            buf0.data_ptr(), buf1.data_ptr(), self.ctrl.data_ptr(), shape[1], shape[2], shape[3],
            int(self.handle_boundaries), self.grid_spacing, self.gamma, self.c_cfl, self.t_final, self.max_steps)

    @property
    def cells(self) -> int:
        return self.shape[1] * self.shape[2] * self.shape[3]

    def reset(self):
        """t = 0, n = 0, first dt from buf0 (asynchronous)."""
        _lib.check(_lib.load().tf_sim_reset(ctypes.byref(self.desc), _stream(self.buf[0].device)))

    def status(self) -> _lib.SimStatus:
        """t, dt, step count and flags. Reads the control block on the current
        stream: work that advances the simulation on another stream must be
        waited for first (Batch.status reads on the batch's stream)."""
        raw = self.ctrl[:48].cpu().numpy().tobytes()
        return _lib.SimStatus.from_buffer_copy(raw)

    def state(self, status: _lib.SimStatus = None) -> torch.Tensor:
        st = status if status is not None else self.status()
        return self.buf[int(st.n_iter) & 1]


class StatusCopy:
    """Statuses of a batch being copied to page-locked host memory on the
    batch's stream (Batch.status_async). Keep it, and its batch open, until
    result(): the copy writes into its buffer until the stream passes it."""

    def __init__(self, batch: "Batch", raw: torch.Tensor, event):
        self.batch, self.raw, self.event = batch, raw, event

    def ready(self) -> bool:
        return self.event.query()

    def result(self) -> List[_lib.SimStatus]:
        self.event.synchronize()
        out = (_lib.SimStatus * len(self.batch.sims))()
        _lib.check(_lib.load().tf_batch_status_unpack(self.batch._h, self.raw.data_ptr(), out))
        return list(out)


class Batch:
    def __init__(self, sims: Sequence[Simulation], tile_x: int = 0, tile_y: int = 0, asynchronous: bool = False):
        """``asynchronous``: build the plan without waiting for the current
        stream (tf_batch_create_async); step it on that stream."""
        if not sims:
            raise ValueError("empty batch")
        self.sims: List[Simulation] = list(sims)
        lib = _lib.load()
        n = len(self.sims)
        device = self.sims[0].buf[0].device
        self.workspace = torch.empty(lib.tf_batch_workspace_bytes(n), dtype=torch.uint8, device=device)
        descs = (_lib.SimDesc * n)(*[s.desc for s in self.sims])
        handle = ctypes.c_void_p()
        create = lib.tf_batch_create_async if asynchronous else lib.tf_batch_create
        _lib.check(create(ctypes.byref(handle), descs, n, self.workspace.data_ptr(), self.workspace.numel(),
                          int(tile_x), int(tile_y), _stream(device)))
        self._h = handle
        self.device = device

    def close(self):
        if self._h:
            _lib.load().tf_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tiling(self, index: int = 0):
        """(tile_x, tile_y, waves per z-line) the plan runs simulation `index` with."""
        tx, ty, w = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        _lib.check(_lib.load().tf_batch_sim_tiling(self._h, int(index), ctypes.byref(tx), ctypes.byref(ty),
                                                   ctypes.byref(w)))
        return tx.value, ty.value, w.value

    def reset(self):
        """Simulation.reset of every simulation, in three launches (asynchronous)."""
        _lib.check(_lib.load().tf_batch_reset(self._h, _stream(self.device)))

    def step(self, steps: int = 1):
        _lib.check(_lib.load().tf_batch_step(self._h, int(steps), _stream(self.device)))

    def step_graph(self, steps: int):
        _lib.check(_lib.load().tf_batch_step_graph(self._h, int(steps), _stream(self.device)))

    def run_graph(self, steps: int):
        """``steps`` (1..1024) steps as one captured graph, launched once."""
        _lib.check(_lib.load().tf_batch_run_graph(self._h, int(steps), _stream(self.device)))

    def step_timed(self, steps: int) -> float:
        """Runs ``steps`` steps, waits, returns the summed stencil-kernel ms."""
        ms = ctypes.c_float(0.0)
        _lib.check(_lib.load().tf_batch_step_timed(self._h, int(steps), ctypes.byref(ms), _stream(self.device)))
        return float(ms.value)

    def status_async(self) -> StatusCopy:
        """Enqueues a copy of every simulation's status on the current stream
        and returns at once; StatusCopy.result() waits for it."""
        lib = _lib.load()
        raw = torch.empty(lib.tf_batch_status_bytes(self._h), dtype=torch.uint8, pin_memory=True)
        _lib.check(lib.tf_batch_status_enqueue(self._h, raw.data_ptr(), _stream(self.device)))
        event = torch.cuda.Event()
        event.record(torch.cuda.current_stream(self.device))
        return StatusCopy(self, raw, event)

    def status(self) -> List[_lib.SimStatus]:
        n = len(self.sims)
        out = (_lib.SimStatus * n)()
        _lib.check(_lib.load().tf_batch_status(self._h, out, _stream(self.device)))
        return list(out)
