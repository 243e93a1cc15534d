"""Whole registrations on the HIP engine: the front end of fastwarp's `register_syn`
(python/fastwarp/_api.py:98 -> stages.rs:286 register_stages_async) over `fw_register`.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._abi import Affine, Grid, Plan, Result, TraceStep, dptr, read_trace
from .backend import FW_OK, DeviceField, DeviceImage, HipBackend, RegError, lib, write_composite_pair


def cohort_plan(channels: int = 1) -> Plan:
    """The cohort recipe (Rigid/Affine MI, SyN CC radius 4, 8x4x2x1, 3x2x1x0 vox)."""
    p = Plan()
    lib().fw_plan_cohort(C.byref(p), channels)
    return p


def plan_2mm(channels: int = 1) -> Plan:
    """The 2 mm plain registration: Rigid/Affine MI at shrink 4/2/1 (100/50/25 iterations,
    sigmas 2/1/0 vox), SyN CC radius 4, step 0.1, update variance 3 at shrink 4/2/1 (40/20/10)."""
    p = cohort_plan(channels)
    p.set_linear([100, 50, 25], [4, 2, 1], [2.0, 1.0, 0.0])
    p.set_syn([40, 20, 10], [4, 2, 1], [2.0, 1.0, 0.0])
    return p


@dataclass
class Registration:
    affine: Affine
    forward: np.ndarray | None  # (n, 3) on the fixed grid
    inverse: np.ndarray | None
    result: Result


def write_composite(reg: Registration, fixed_grid: Grid, prefix) -> None:
    """A registration's transforms as antsRegistration writes them with --write-composite-transform
    (ants_io.rs:100-141): <prefix>Composite.h5 = [affine, forward field] and
    <prefix>InverseComposite.h5 = [inverse field, affine^-1]; [affine] / [affine^-1] when the plan
    had no SyN stage or the fields were not downloaded."""
    if reg.forward is None or reg.inverse is None:
        write_composite_pair(prefix, reg.affine)
    else:
        write_composite_pair(prefix, reg.affine, fixed_grid, reg.forward, reg.inverse)


def _sig():
    return lib()


def channel_array(arrays):
    arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in arrays]
    ptrs = (C.POINTER(C.c_double) * len(arrs))(*[dptr(a) for a in arrs])
    return arrs, ptrs


def register(backend: HipBackend, plan: Plan, fixed_grid: Grid, fixed, moving_grid: Grid, moving,
             fields: bool = True, cancel: C.c_int32 | None = None) -> Registration:
    """One full normalisation on the device of `backend`."""
    h = _sig()
    fa, fp = channel_array(fixed)
    ma, mp = channel_array(moving)
    n = fixed_grid.n
    fwd = np.empty((n, 3)) if fields and plan.run_syn else None
    inv = np.empty((n, 3)) if fields and plan.run_syn else None
    res = Result()
    st = h.fw_register(backend.ctx, C.byref(plan), C.byref(fixed_grid), fp, C.byref(moving_grid), mp,
                       dptr(fwd) if fwd is not None else None, dptr(inv) if inv is not None else None, C.byref(res),
                       C.byref(cancel) if cancel is not None else None)
    if st != FW_OK:
        raise RegError(st, h.fw_last_error(backend.ctx).decode())
    return Registration(res.affine, fwd, inv, res)


@dataclass
class ResidentRegistration:
    affine: Affine
    forward: DeviceField | None  # on the fixed grid, in HBM
    inverse: DeviceField | None
    result: Result


def register_resident(backend: HipBackend, plan: Plan, fixed: list[DeviceImage], moving: list[DeviceImage],
                      fields: bool = True, cancel: C.c_int32 | None = None) -> ResidentRegistration:
    """One full normalisation from raw images already on the device (upload_image handles, kept);
    the fields stay on the device."""
    h = _sig()
    fp = (C.c_void_p * len(fixed))(*[im.handle for im in fixed])
    mp = (C.c_void_p * len(moving))(*[im.handle for im in moving])
    fwd, inv = C.c_void_p(), C.c_void_p()
    want = fields and plan.run_syn
    res = Result()
    st = h.fw_register_resident(backend.ctx, C.byref(plan), fp, mp, C.byref(fwd) if want else None,
                                C.byref(inv) if want else None, C.byref(res), C.byref(cancel) if cancel is not None else None)
    if st != FW_OK:
        raise RegError(st, h.fw_last_error(backend.ctx).decode())
    wrap = lambda p: DeviceField(backend, p) if p.value else None
    return ResidentRegistration(res.affine, wrap(fwd) if want else None, wrap(inv) if want else None, res)


def register_batch(backend: HipBackend, plan: Plan, fixed: list[DeviceImage], moving: list[list[DeviceImage]],
                   fields: bool = True, workers: int = 0, cancel: C.c_int32 | None = None) -> list[ResidentRegistration]:
    """Several subjects (moving[s] holds subject s's channels) registered to one template (the
    fixed images): the template's preprocessing, centre of mass and smoothed levels are shared, the
    linear chains run side by side on `workers` worker streams (0: 8), and the SyN stages run
    batched on the device (fw_register_batch). Each subject's result is the bits of its own
    register_resident run."""
    h = _sig()
    ns = len(moving)
    fp = (C.c_void_p * len(fixed))(*[im.handle for im in fixed])
    mp = (C.c_void_p * max(ns * len(fixed), 1))(*[im.handle for subj in moving for im in subj])
    want = fields and plan.run_syn
    fwd, inv = (C.c_void_p * max(ns, 1))(), (C.c_void_p * max(ns, 1))()
    res = (Result * max(ns, 1))()
    st = h.fw_register_batch(backend.ctx, C.byref(plan), fp, ns, mp, workers, fwd if want else None, inv if want else None,
                             res, C.byref(cancel) if cancel is not None else None)
    if st != FW_OK:
        raise RegError(st, h.fw_last_error(backend.ctx).decode())
    wrap = lambda p: DeviceField(backend, C.c_void_p(p)) if p else None
    return [ResidentRegistration(res[s].affine, wrap(fwd[s]) if want else None, wrap(inv[s]) if want else None, res[s])
            for s in range(ns)]


def syn_stage(backend: HipBackend, plan: Plan, fixed: list[DeviceImage], moving: list[DeviceImage],
              initial: Affine | None = None, fixed_mask: DeviceImage | None = None, moving_mask: DeviceImage | None = None,
              prior: DeviceField | None = None, prior_inverse: DeviceField | None = None,
              cancel: C.c_int32 | None = None) -> ResidentRegistration:
    """One SyN stage (the plan's SyN recipe) on images already on the device, taken as they are,
    from `initial`, with optional masks and the fields of earlier SyN stages, which are collapsed
    with this stage's (fw_syn_stage). The fields stay on the device."""
    h = _sig()
    fp = (C.c_void_p * len(fixed))(*[im.handle for im in fixed])
    mp = (C.c_void_p * len(moving))(*[im.handle for im in moving])
    fwd, inv = C.c_void_p(), C.c_void_p()
    res = Result()
    handle = lambda x: x.handle if x is not None else None
    st = h.fw_syn_stage(backend.ctx, C.byref(plan), fp, mp, C.byref(initial) if initial is not None else None,
                        handle(fixed_mask), handle(moving_mask), handle(prior), handle(prior_inverse), C.byref(fwd),
                        C.byref(inv), C.byref(res), C.byref(cancel) if cancel is not None else None)
    if st != FW_OK:
        raise RegError(st, h.fw_last_error(backend.ctx).decode())
    wrap = lambda p: DeviceField(backend, p) if p.value else None
    return ResidentRegistration(res.affine, wrap(fwd), wrap(inv), res)


def linear_batch(backend: HipBackend, plan: Plan, fixed: list[DeviceImage], moving: list[DeviceImage], workers: int = 0,
                 cancel: C.c_int32 | None = None) -> list[Result]:
    """Independent linear fits (centre of mass, Rigid, Affine) of fixed[i] / moving[i] (raw images
    on the device, each pair its own grids), side by side on `workers` worker streams (0: 8)
    (fw_linear_batch). Each fit's result is the bits of its own register_resident run without SyN."""
    h = _sig()
    n = len(fixed)
    if len(moving) != n:
        raise ValueError("one moving image per fixed image")
    fp = (C.c_void_p * max(n, 1))(*[im.handle for im in fixed])
    mp = (C.c_void_p * max(n, 1))(*[im.handle for im in moving])
    res = (Result * max(n, 1))()
    st = h.fw_linear_batch(backend.ctx, C.byref(plan), n, fp, mp, workers, res, C.byref(cancel) if cancel is not None else None)
    if st != FW_OK:
        raise RegError(st, h.fw_last_error(backend.ctx).decode())
    return [res[i] for i in range(n)]


def last_batch_trace(subject: int) -> list[TraceStep]:
    """last_trace for one subject (or fit) of this thread's last register_batch or linear_batch call."""
    return read_trace(lambda out, cap: lib().fw_last_batch_trace(subject, out, cap))


def last_trace() -> list[TraceStep]:
    """Per-iteration records (CG values, SyN energies and inversions) of this thread's last
    register / register_resident call."""
    return read_trace(lib().fw_last_trace)
