"""Serving: continuous batching of whole registrations on one device, over `fw_server_*`.

The reference runs several registrations on one GPU as independent device instances, up to four at
once on worker threads (python/fastwarp/_api.py:1104-1130, sized by `registrations_at_once`,
crates/fastwarp-cli/src/engine.rs:1015-1068). `Server` keeps such instances as slots inside the
library (one HIP stream and buffer pool each, on its own native thread) and admits queued requests
to them by device-memory footprint as slots and memory free up. A request's result is the same bits
alone and in any mix.
"""
from __future__ import annotations

import ctypes as C
import threading
from dataclasses import dataclass

import numpy as np

from ._abi import Affine, Grid, Plan, Result, ServedTiming, dptr
from .backend import FW_ERR_INVALID, FW_OK, RegError, lib


@dataclass
class Ticket:
    """A queued request: its id and the output arrays its fields are written into."""

    id: int
    forward: np.ndarray | None
    inverse: np.ndarray | None


@dataclass
class Served:
    affine: Affine
    forward: np.ndarray | None  # (n, 3) on the fixed grid
    inverse: np.ndarray | None
    result: Result
    timing: ServedTiming


def _f32_channels(arrays):
    arrs = [np.ascontiguousarray(a, dtype=np.float32).reshape(-1) for a in arrays]
    ptrs = (C.POINTER(C.c_float) * len(arrs))(*[a.ctypes.data_as(C.POINTER(C.c_float)) for a in arrs])
    return arrs, ptrs


def _error(status: int) -> RegError:
    return RegError(status, lib().fw_server_last_error().decode())


class Server:
    """`slots` 0 takes 4; `budget_bytes` 0 takes the device's free memory less 1 GiB."""

    def __init__(self, device: int = 0, slots: int = 0, budget_bytes: int = 0):
        self.handle = C.c_void_p()
        # ticket id -> the arrays and structs the native slots read and write until the request
        # has ended: held here, not by the Ticket, so that a ticket dropped without a wait cannot
        # free memory a slot still uses
        self._keep: dict[int, tuple] = {}
        self._lock = threading.Lock()
        st = lib().fw_server_create(device, slots, budget_bytes, C.byref(self.handle))
        if st != FW_OK:
            raise _error(st)

    def submit(self, plan: Plan, fixed_grid: Grid, fixed, moving_grid: Grid, moving, fields: bool = True) -> Ticket:
        fa, fp = _f32_channels(fixed)
        ma, mp = _f32_channels(moving)
        n = fixed_grid.n
        fwd = np.empty((n, 3)) if fields and plan.run_syn else None
        inv = np.empty((n, 3)) if fields and plan.run_syn else None
        ticket = C.c_uint64()
        pl, fg, mg = Plan.from_buffer_copy(plan), Grid.from_buffer_copy(fixed_grid), Grid.from_buffer_copy(moving_grid)
        st = lib().fw_server_submit(self.handle, C.byref(pl), C.byref(fg), fp, C.byref(mg), mp,
                                    dptr(fwd) if fwd is not None else None, dptr(inv) if inv is not None else None,
                                    C.byref(ticket))
        if st != FW_OK:
            raise _error(st)
        self._keep[ticket.value] = (fa, fp, ma, mp, pl, fg, mg, fwd, inv)
        return Ticket(ticket.value, fwd, inv)

    def wait(self, ticket: Ticket) -> Served:
        """Blocks until the request has ended; raises RegError when it failed or was cancelled.
        A ticket is waited for once: a second wait, also a concurrent one, is refused."""
        with self._lock:
            keep = self._keep.pop(ticket.id, None)
        if keep is None:
            raise RegError(FW_ERR_INVALID, f"server: unknown ticket {ticket.id}")
        res, timing = Result(), ServedTiming()
        # `keep` holds the request's arrays until the native wait has returned, i.e. until the
        # request has ended and its slot no longer touches them
        st = lib().fw_server_wait(self.handle, ticket.id, C.byref(res), C.byref(timing))
        if st != FW_OK:
            raise _error(st)
        del keep
        return Served(res.affine, ticket.forward, ticket.inverse, res, timing)

    def cancel(self, ticket: Ticket) -> None:
        st = lib().fw_server_cancel(self.handle, ticket.id)
        if st != FW_OK:
            raise _error(st)

    def counts(self) -> tuple[int, int, int]:
        """(waiting, running, device bytes admitted to the running requests)."""
        w, r, b = C.c_uint32(), C.c_uint32(), C.c_uint64()
        st = lib().fw_server_counts(self.handle, C.byref(w), C.byref(r), C.byref(b))
        if st != FW_OK:
            raise _error(st)
        return w.value, r.value, b.value

    def close(self) -> None:
        """Cancels what has not ended and stops the slots, then lets go of the requests' arrays."""
        if self.handle:
            lib().fw_server_destroy(self.handle)
            self.handle = C.c_void_p()
        self._keep.clear()

    def __enter__(self) -> "Server":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
