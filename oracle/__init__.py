"""ORACLE — test infrastructure only.

ctypes bindings of the f64 CPU restatement of the reference's CPU engine (oracle/oracle.hpp). Only
the tests, the smoke check and the CPU-baseline leg of the benchmark import this package, and only
as the checker; the product never does.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from fastwarp_amd._abi import Affine, Grid, InvertParams, InvertStats, LinearTransform, MiBinning, aos, dptr, fptr, u32ptr

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libfastwarp_oracle.so"
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            from . import build as _b

            _b.build()
        h = C.CDLL(str(LIB_PATH))
        h.fwo_last_error.restype = C.c_char_p
        h.fwo_bessel.restype = C.c_double
        h.fwo_bessel.argtypes = [C.c_int, C.c_double]
        h.fwo_cubic_bspline.restype = C.c_double
        h.fwo_cubic_bspline.argtypes = [C.c_double, C.c_int]
        h.fwo_last_trace.restype = C.c_int32
        _lib = h
    return _lib


class OracleError(RuntimeError):
    pass


def _check(st: int):
    if st != 0:
        raise OracleError(lib().fwo_last_error().decode())


def smooth_image(grid: Grid, data, sigma_mm) -> np.ndarray:
    """smooth_recursive_gaussian (pyramid.rs:429-437), f64."""
    d = np.ascontiguousarray(data, dtype=np.float64)
    s = np.ascontiguousarray(sigma_mm, dtype=np.float64).reshape(3)
    out = np.empty(grid.n)
    _check(lib().fwo_smooth_image(C.byref(grid), dptr(d), dptr(s), dptr(out)))
    return out


def preprocess(grid: Grid, data, lower: float = 0.005, upper: float = 0.995) -> np.ndarray:
    """Winsorise by the 256-bin histogram quantiles and window to [0, 1] (preprocess.rs:159-232), f64."""
    d = np.ascontiguousarray(data, dtype=np.float64)
    out = np.empty(grid.n)
    _check(lib().fwo_preprocess(C.byref(grid), dptr(d), C.c_float(lower), C.c_float(upper), dptr(out)))
    return out


def resample_isotropic(grid: Grid, data, mm: float) -> tuple[Grid, np.ndarray]:
    """The pipeline's isotropic resample (register.rs:82-142), f64 rounded to f32: (grid, values)."""
    d = np.ascontiguousarray(data, dtype=np.float64)
    og = Grid()
    _check(lib().fwo_resample_isotropic(C.byref(grid), dptr(d), C.c_double(mm), C.byref(og), None, C.c_int64(0)))
    out = np.empty(og.n)
    _check(lib().fwo_resample_isotropic(C.byref(grid), dptr(d), C.c_double(mm), C.byref(og), dptr(out), C.c_int64(og.n)))
    return og, out


def centre_of_gravity(grid: Grid, data) -> np.ndarray:
    """Intensity-weighted centre in mm (init.rs:47-86), f64."""
    d = np.ascontiguousarray(data, dtype=np.float64)
    c = np.zeros(3)
    _check(lib().fwo_centre_of_gravity(C.byref(grid), dptr(d), dptr(c)))
    return c


def sample_values(grid: Grid, data, points) -> tuple[np.ndarray, np.ndarray]:
    """(nearest-voxel, linear) values at physical points, NaN outside (mi.rs:161-169, interp.rs:43-87)."""
    d = np.ascontiguousarray(data, dtype=np.float64)
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    near, lin = np.empty(len(pts)), np.empty(len(pts))
    _check(lib().fwo_sample_values(C.byref(grid), dptr(d), dptr(pts), C.c_int64(len(pts)), dptr(near), dptr(lin)))
    return near, lin


def warp(src_grid: Grid, src, field_grid: Grid, field, pre_grid: Grid | None = None, pre=None, affine: Affine | None = None,
         nearest: bool = False, product=None) -> np.ndarray:
    s = np.ascontiguousarray(src, dtype=np.float64)
    f = aos(field)
    p = aos(pre) if pre is not None else None
    pr = np.ascontiguousarray(product, dtype=np.float64) if product is not None else None
    out = np.empty(field_grid.n)
    _check(
        lib().fwo_warp(
            C.byref(src_grid), dptr(s), C.byref(field_grid), dptr(f), C.byref(pre_grid) if pre_grid else None,
            dptr(p) if p is not None else None, C.byref(affine) if affine else None, int(nearest),
            dptr(pr) if pr is not None else None, dptr(out),
        )
    )
    return out


def cc(grid: Grid, fixed, moving, radius: int, *, moving_grid: Grid | None = None, moving_affine: Affine | None = None,
       mask=None, derivative: bool = True) -> dict:
    """CcMetric::evaluate_masked (metric/cc.rs:170-321)."""
    mg = moving_grid or grid
    f = np.ascontiguousarray(fixed, dtype=np.float64)
    m = np.ascontiguousarray(moving, dtype=np.float64)
    mk = np.ascontiguousarray(mask, dtype=np.float64) if mask is not None else None
    n = grid.n
    value = C.c_double()
    nv = C.c_int64()
    local = np.zeros(n)
    coeff = np.zeros(n)
    deriv = np.zeros((n, 3))
    valid = np.zeros(n, dtype=np.uint8)
    _check(
        lib().fwo_cc(
            C.byref(grid), dptr(f), C.byref(mg), dptr(m), C.byref(moving_affine) if moving_affine else None,
            dptr(mk) if mk is not None else None, int(radius), int(derivative), C.byref(value), C.byref(nv),
            dptr(local), dptr(coeff), dptr(deriv), valid.ctypes.data_as(C.POINTER(C.c_uint8)),
        )
    )
    return {"value": value.value, "num_valid": nv.value, "local_cc": local, "coefficient": coeff, "derivative": deriv,
            "valid": valid.astype(bool)}


def mi_dense(fixed_grid: Grid, fixed, moving_grid: Grid, moving, bins: int, *, field_grid: Grid | None = None, field=None,
             fixed_mask=None, moving_mask=None, derivative: bool = True) -> dict:
    """MattesMi::evaluate_dense (metric/mi.rs:578-676)."""
    f = np.ascontiguousarray(fixed, dtype=np.float64)
    m = np.ascontiguousarray(moving, dtype=np.float64)
    fl = np.ascontiguousarray(field, dtype=np.float32).reshape(-1) if field is not None else None
    fm = np.ascontiguousarray(fixed_mask, dtype=np.float64) if fixed_mask is not None else None
    mm = np.ascontiguousarray(moving_mask, dtype=np.float64) if moving_mask is not None else None
    value = C.c_double()
    nv = C.c_int64()
    deriv = np.zeros((fixed_grid.n, 3))
    _check(
        lib().fwo_mi_dense(
            C.byref(fixed_grid), dptr(f), C.byref(moving_grid), dptr(m), C.byref(field_grid) if field_grid else None,
            fptr(fl) if fl is not None else None, dptr(fm) if fm is not None else None, dptr(mm) if mm is not None else None,
            int(bins), int(derivative), C.byref(value), C.byref(nv), dptr(deriv),
        )
    )
    return {"value": value.value, "num_valid": nv.value, "derivative": deriv}


def mi_sampled(binning: MiBinning, moving_grid: Grid, moving, points, fixed_bin, t: LinearTransform, derivative=True) -> dict:
    """MattesMi::evaluate with prepared fixed bins (metric/mi.rs:443-511)."""
    m = np.ascontiguousarray(moving, dtype=np.float64)
    p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    b = np.ascontiguousarray(fixed_bin, dtype=np.uint32)
    value = C.c_double()
    nv = C.c_int64()
    n = 6 if t.kind == 0 else 12
    d = np.zeros(n)
    pdf = np.zeros(binning.bins * binning.bins)
    _check(
        lib().fwo_mi_sampled(
            C.byref(binning), C.byref(moving_grid), dptr(m), dptr(p), u32ptr(b), p.shape[0], C.byref(t), int(derivative),
            C.byref(value), dptr(d), C.byref(nv), dptr(pdf),
        )
    )
    return {"value": value.value, "derivative": d, "num_valid": nv.value, "joint_pdf": pdf}


def mi_finish(bins: int, moving_bin_size: float, num_valid: int, joint, fixed_marginal) -> tuple[float, np.ndarray]:
    """finish_histogram (metric/mi.rs:702-769) on given cells: (value, ratio weights)."""
    j = np.ascontiguousarray(joint, dtype=np.float64).reshape(-1)
    f = np.ascontiguousarray(fixed_marginal, dtype=np.float64).reshape(-1)
    value = C.c_double()
    ratio = np.zeros(bins * bins)
    _check(lib().fwo_mi_finish(C.c_int64(bins), C.c_double(moving_bin_size), C.c_int64(num_valid), dptr(j), dptr(f),
                               C.byref(value), dptr(ratio)))
    return value.value, ratio


def gaussian_taps(variance: float, max_error: float = 0.001, max_width: int = 1000) -> np.ndarray:
    out = np.zeros(4096)
    n = C.c_int64()
    _check(lib().fwo_gaussian_taps(C.c_double(variance), C.c_double(max_error), C.c_int64(max_width), dptr(out), 4096, C.byref(n)))
    return out[: n.value].copy()


def bessel(n: int, y: float) -> float:
    return lib().fwo_bessel(n, y)


def cubic_bspline(u: float, derivative: bool = False) -> float:
    return lib().fwo_cubic_bspline(u, int(derivative))


def smooth(grid: Grid, field, variance: float) -> np.ndarray:
    f = aos(field)
    out = np.empty_like(f)
    _check(lib().fwo_smooth(C.byref(grid), dptr(f), C.c_double(variance), dptr(out)))
    return out


def compose(grid: Grid, displacement, warping) -> np.ndarray:
    u, w = aos(displacement), aos(warping)
    out = np.empty_like(u)
    _check(lib().fwo_compose(C.byref(grid), dptr(u), dptr(w), dptr(out)))
    return out


def invert(grid: Grid, field, initial=None, params: InvertParams | None = None):
    params = params or InvertParams.syn()
    f = aos(field)
    init = aos(initial) if initial is not None else None
    out = np.empty_like(f)
    st = InvertStats()
    _check(lib().fwo_invert(C.byref(grid), dptr(f), dptr(init) if init is not None else None, C.byref(params), dptr(out), C.byref(st)))
    return out, (int(st.iterations), st.mean_error_norm, st.max_error_norm)


def resample(grid: Grid, field, target: Grid) -> np.ndarray:
    f = aos(field)
    out = np.empty((target.n, 3))
    _check(lib().fwo_resample(C.byref(grid), dptr(f), C.byref(target), dptr(out)))
    return out


def norms(grid: Grid, field) -> tuple[float, float]:
    f = aos(field)
    a, b = C.c_double(), C.c_double()
    _check(lib().fwo_norms(C.byref(grid), dptr(f), C.byref(a), C.byref(b)))
    return a.value, b.value


def register(plan, fixed_grid: Grid, fixed, moving_grid: Grid, moving, fields: bool = True):
    """The whole f64 CPU engine (engine="cpu") on the same plan as the HIP path: the generic
    drivers over the oracle's ops. Returns (Result, forward, inverse)."""
    from fastwarp_amd._abi import Plan, Result
    from fastwarp_amd.register import channel_array

    h = lib()
    fa, fp = channel_array(fixed)
    ma, mp = channel_array(moving)
    n = fixed_grid.n
    fwd = np.empty((n, 3)) if fields and plan.run_syn else None
    inv = np.empty((n, 3)) if fields and plan.run_syn else None
    res = Result()
    _check(
        h.fwo_register(C.byref(plan), C.byref(fixed_grid), fp, C.byref(moving_grid), mp,
                       dptr(fwd) if fwd is not None else None, dptr(inv) if inv is not None else None, C.byref(res))
    )
    return res, fwd, inv


def syn_stage(plan, fixed_grid: Grid, fixed, moving_grid: Grid, moving, initial: Affine | None = None,
              fixed_mask=None, moving_mask=None, prior=None):
    """One SyN stage on the f64 CPU engine (fw_syn_stage's counterpart): masks as (grid, data)
    pairs, the earlier stages' fields as (grid, forward (n, 3), inverse (n, 3)). Returns (Result,
    forward, inverse) on the fixed grid."""
    from fastwarp_amd._abi import Result
    from fastwarp_amd.register import channel_array

    h = lib()
    fa, fp = channel_array(fixed)
    ma, mp = channel_array(moving)
    n = fixed_grid.n
    fwd, inv = np.empty((n, 3)), np.empty((n, 3))
    res = Result()
    keep = []

    def arr(x):
        a = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        keep.append(a)
        return dptr(a)

    fm_grid, fm = (C.byref(fixed_mask[0]), arr(fixed_mask[1])) if fixed_mask is not None else (None, None)
    mm_grid, mm = (C.byref(moving_mask[0]), arr(moving_mask[1])) if moving_mask is not None else (None, None)
    if prior is not None:
        p_grid, pw, pi = C.byref(prior[0]), arr(prior[1]), arr(prior[2])
    else:
        p_grid, pw, pi = None, None, None
    _check(h.fwo_syn_stage(C.byref(plan), C.byref(fixed_grid), fp, C.byref(moving_grid), mp,
                           C.byref(initial) if initial is not None else None, fm_grid, fm, mm_grid, mm, p_grid, pw, pi,
                           dptr(fwd), dptr(inv), C.byref(res)))
    return res, fwd, inv


def generic_syn(on: bool):
    """Run the CPU engine's SyN stages through the product's generic loop (True) or the
    oracle's own restatement of syn.rs (False, the default)."""
    lib().fwo_generic_syn(int(on))


def generic_linear(on: bool):
    """Run the CPU engine's linear chains through the product's generic driver (True) or the
    oracle's own restatement of linear.rs and optimizer.rs (False, the default)."""
    lib().fwo_generic_linear(int(on))


def last_trace():
    """Per-iteration records of this thread's last `register` call (fastwarp_amd._abi.TraceStep)."""
    from fastwarp_amd._abi import read_trace

    return read_trace(lib().fwo_last_trace)


def register_lockstep(plan, fixed_grid: Grid, fixed, moving_grid: Grid, moving, trace, fields: bool = True,
                      start: Affine | None = None):
    """The CPU engine's run replaying another engine's discrete SyN decisions, taken from that
    run's per-iteration trace (fastwarp_amd._abi.TraceStep records): every SyN level runs the
    traced iteration count (convergence test off) and every inversion the traced step count.
    With `start` (that run's collapsed linear result), the linear stages are skipped and the SyN
    stage starts from it, so that the two SyN stages see the same inputs.
    Returns (Result, forward, inverse, disagreements), where each disagreement is (inversion call,
    body, the oracle's own decision (1 continue), max error, mean error) at a check where the
    oracle's tolerance rule would have decided otherwise."""
    from fastwarp_amd._abi import Plan

    syn = [t for t in trace if t.stage == 2]
    steps = [int(s) for t in syn for s in t.inversion_steps]
    p = Plan()
    C.pointer(p)[0] = plan
    if start is not None:
        p.run_linear = 0
        p.initial_kind = 2
        p.initial_affine = start
    for level in range(p.syn_levels):
        p.syn_iterations[level] = sum(1 for t in syn if t.level == level)
    p.syn_convergence_threshold = -1e300
    arr = (C.c_int32 * max(len(steps), 1))(*steps)
    h = lib()
    h.fwo_lockstep.argtypes = [C.POINTER(C.c_int32), C.c_int64]
    h.fwo_lockstep(arr, len(steps))
    try:
        res, fwd, inv = register(p, fixed_grid, fixed, moving_grid, moving, fields=fields)
    finally:
        h.fwo_lockstep(arr, -1)
    cap = 4096
    call = (C.c_int64 * cap)()
    body = (C.c_int32 * cap)()
    nat = (C.c_int32 * cap)()
    mx = (C.c_double * cap)()
    mn = (C.c_double * cap)()
    h.fwo_lockstep_disagreements.restype = C.c_int64
    n = h.fwo_lockstep_disagreements(call, body, nat, mx, mn, cap)
    dis = [(call[i], body[i], nat[i], mx[i], mn[i]) for i in range(min(n, cap))]
    return res, fwd, inv, dis
