"""Python mirror of the reference's `Backend` / `LinearEngine` stage interface over the C ABI.

`HipBackend` exposes the trait methods of crates/fastwarp-reg/src/backend.rs:39-305 with the same
names and argument meaning; images and fields are device handles (`DeviceImage`, `DeviceField`)
that free themselves. Host arrays cross as float64 numpy arrays: images shaped (n,), fields
(n, 3), both x-fastest. There is no CPU fallback: without the built gfx950 library every call
raises.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from ._abi import (FW_CHAIN_AFFINE, FW_CHAIN_FIELD, Affine, Capacity, ChainElement, Grid, InvertParams, InvertStats,
                   LinearTransform, MiBinning, Plan, Result, ServedTiming, TraceStep, aos, dptr, read_trace, u32ptr)

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libfastwarp_hip.so"

FW_OK = 0
FW_ERR_INVALID = -1
FW_ERR_DEVICE = -2
FW_ERR_CANCELLED = -3


class RegError(RuntimeError):
    """A failed operation (RegError::Invalid / device errors / Cancelled), message as the library gives it."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


_lib = None

# (name, restype, argtypes) for every entry point of include/fastwarp_hip.h
_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
_D = C.POINTER(C.c_double)
_SIGNATURES = [
    ("fw_ctx_create", C.c_int, [C.c_int, _PP]),
    ("fw_ctx_destroy", None, [_P]),
    ("fw_last_error", C.c_char_p, [_P]),
    ("fw_release_idle", C.c_int, [_P]),
    ("fw_memory_stats", C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("fw_image_upload", C.c_int, [_P, C.POINTER(Grid), _D, _PP]),
    ("fw_image_upload_f32", C.c_int, [_P, C.POINTER(Grid), C.POINTER(C.c_float), _PP]),
    ("fw_image_download", C.c_int, [_P, _P, _D]),
    ("fw_image_grid", C.c_int, [_P, C.POINTER(Grid)]),
    ("fw_image_free", None, [_P, _P]),
    ("fw_field_zero", C.c_int, [_P, C.POINTER(Grid), _PP]),
    ("fw_field_upload", C.c_int, [_P, C.POINTER(Grid), _D, _PP]),
    ("fw_field_download", C.c_int, [_P, _P, _D]),
    ("fw_field_grid", C.c_int, [_P, C.POINTER(Grid)]),
    ("fw_field_copy", C.c_int, [_P, _P, _PP]),
    ("fw_field_free", None, [_P, _P]),
    ("fw_warp", C.c_int, [_P, _P, _P, _P, C.POINTER(Affine), _PP]),
    ("fw_set_tuning", C.c_int, [_P, C.c_char_p, C.c_int64]),
    ("fw_warp_mask", C.c_int, [_P, _P, _P, _P, C.POINTER(Affine), _P, _PP]),
    ("fw_cc_pair", C.c_int, [_P, _P, _P, _P, C.c_uint32, _D, _D, _PP, _PP]),
    ("fw_cc_pair_noalloc", C.c_int, [C.POINTER(Grid), _P, _P, C.c_uint32, _P, _P, _P, _P]),
    ("fw_mi_dense_pair", C.c_int, [_P, _P, _P, _P, _P, C.c_uint32, _D, _D, _PP, _PP]),
    ("fw_smooth_field", C.c_int, [_P, _P, C.c_double, _PP]),
    ("fw_smooth_image", C.c_int, [_P, _P, _D, _PP]),
    ("fw_sample_values", C.c_int, [_P, _P, _D, C.c_int64, _D, _D]),
    ("fw_preprocess_image", C.c_int, [_P, _P, C.c_float, C.c_float, _PP]),
    ("fw_centre_of_gravity", C.c_int, [_P, _P, _D]),
    ("fw_max_scaled_norm_pair", C.c_int, [_P, _P, _P, _D]),
    ("fw_l2_norm", C.c_int, [_P, _P, _D]),
    ("fw_scale", C.c_int, [_P, _P, C.c_double]),
    ("fw_axpy", C.c_int, [_P, _P, C.c_double, _P]),
    ("fw_compose", C.c_int, [_P, _P, _P, _PP]),
    ("fw_invert", C.c_int, [_P, _P, _P, C.POINTER(InvertParams), _PP, C.POINTER(InvertStats)]),
    (
        "fw_invert_pair",
        C.c_int,
        [_P, _P, _P, C.POINTER(InvertParams), _PP, C.POINTER(InvertStats), _PP, C.POINTER(InvertStats)],
    ),
    ("fw_resample_field", C.c_int, [_P, _P, C.POINTER(Grid), _PP]),
    ("fw_mi_level_create", C.c_int, [_P, _P, _D, C.POINTER(C.c_uint32), C.c_int64, C.POINTER(MiBinning), _PP]),
    ("fw_mi_level_value", C.c_int, [_P, C.POINTER(LinearTransform), _D]),
    ("fw_mi_level_histogram", C.c_int, [_P, C.POINTER(LinearTransform), _D, _D, C.POINTER(C.c_int64)]),
    ("fw_mi_level_values", C.c_int, [_P, C.POINTER(LinearTransform), C.c_int64, _D]),
    ("fw_mi_level_value_and_derivative", C.c_int, [_P, C.POINTER(LinearTransform), _D, _D]),
    ("fw_mi_level_free", None, [_P]),
    ("fw_set_profiling", C.c_int, [_P, C.c_int]),
    ("fw_set_profiling_focus", C.c_int, [_P, C.c_char_p, C.c_int]),
    ("fw_take_profile", C.c_int, [_P, C.c_char_p, C.c_size_t]),
    ("fw_plan_capacity", C.c_int, [C.POINTER(Plan), C.POINTER(Grid), C.POINTER(Grid), C.POINTER(Capacity)]),
    ("fw_memory_budget", C.c_int, [_P, C.POINTER(C.c_uint64)]),
    ("fw_fits", C.c_int, [_P, C.POINTER(Plan), C.POINTER(Grid), C.POINTER(Grid), C.POINTER(C.c_int32)]),
    (
        "fw_registrations_at_once",
        C.c_int,
        [_P, C.POINTER(Plan), C.POINTER(Grid), C.POINTER(Grid), C.c_uint32, C.POINTER(C.c_uint32)],
    ),
    ("fw_plan_cohort", None, [C.POINTER(Plan), C.c_int32]),
    (
        "fw_register",
        C.c_int,
        [_P, C.POINTER(Plan), C.POINTER(Grid), C.POINTER(C.POINTER(C.c_double)), C.POINTER(Grid),
         C.POINTER(C.POINTER(C.c_double)), _D, _D, C.POINTER(Result), C.POINTER(C.c_int32)],
    ),
    ("fw_register_resident", C.c_int, [_P, C.POINTER(Plan), C.POINTER(_P), C.POINTER(_P), _PP, _PP, C.POINTER(Result),
                                       C.POINTER(C.c_int32)]),
    ("fw_last_trace", C.c_int32, [C.POINTER(TraceStep), C.c_int32]),
    ("fw_register_batch", C.c_int, [_P, C.POINTER(Plan), C.POINTER(_P), C.c_uint32, C.POINTER(_P), C.c_uint32,
                                    C.POINTER(_P), C.POINTER(_P), C.POINTER(Result), C.POINTER(C.c_int32)]),
    ("fw_linear_batch", C.c_int, [_P, C.POINTER(Plan), C.c_uint32, C.POINTER(_P), C.POINTER(_P), C.c_uint32,
                                  C.POINTER(Result), C.POINTER(C.c_int32)]),
    ("fw_syn_stage", C.c_int, [_P, C.POINTER(Plan), C.POINTER(_P), C.POINTER(_P), C.POINTER(Affine), _P, _P, _P, _P, _PP,
                               _PP, C.POINTER(Result), C.POINTER(C.c_int32)]),
    ("fw_last_batch_trace", C.c_int32, [C.c_uint32, C.POINTER(TraceStep), C.c_int32]),
    ("fw_resample_isotropic", C.c_int, [_P, _P, C.c_double, _PP]),
    ("fw_write_affine_mat", C.c_int, [C.c_char_p, C.POINTER(Affine)]),
    ("fw_read_affine_mat", C.c_int, [C.c_char_p, C.POINTER(Affine)]),
    ("fw_write_composite_h5", C.c_int, [C.c_char_p, C.POINTER(ChainElement), C.c_uint32]),
    ("fw_read_composite_h5", C.c_int, [C.c_char_p, C.POINTER(ChainElement), C.c_uint32, C.POINTER(C.c_uint32)]),
    ("fw_write_composite_pair", C.c_int, [C.c_char_p, C.POINTER(Affine), C.POINTER(Grid), _D, _D]),
    ("fw_write_composite_fields", C.c_int, [_P, C.c_char_p, C.POINTER(Affine), _P, _P]),
    ("fw_server_create", C.c_int, [C.c_int32, C.c_uint32, C.c_uint64, _PP]),
    ("fw_server_destroy", None, [_P]),
    ("fw_server_last_error", C.c_char_p, []),
    (
        "fw_server_submit",
        C.c_int,
        [_P, C.POINTER(Plan), C.POINTER(Grid), C.POINTER(C.POINTER(C.c_float)), C.POINTER(Grid),
         C.POINTER(C.POINTER(C.c_float)), _D, _D, C.POINTER(C.c_uint64)],
    ),
    ("fw_server_wait", C.c_int, [_P, C.c_uint64, C.POINTER(Result), C.POINTER(ServedTiming)]),
    ("fw_server_cancel", C.c_int, [_P, C.c_uint64]),
    ("fw_server_counts", C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)]),
]


def write_affine_mat(path, affine: Affine) -> None:
    """The affine as ANTs writes it (0GenericAffine.mat, itk_mat.rs:150-190)."""
    if lib().fw_write_affine_mat(str(path).encode(), C.byref(affine)) != FW_OK:
        raise RegError(FW_ERR_INVALID, lib().fw_last_error(None).decode())


def read_affine_mat(path) -> Affine:
    """The first transform of an ITK MATLAB transform file as an affine (itk_mat.rs:102-133)."""
    out = Affine()
    if lib().fw_read_affine_mat(str(path).encode(), C.byref(out)) != FW_OK:
        raise RegError(FW_ERR_INVALID, lib().fw_last_error(None).decode())
    return out


def _field_array(grid: Grid, values) -> np.ndarray:
    n = int(grid.size[0]) * int(grid.size[1]) * int(grid.size[2])
    a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
    if a.size != 3 * n:
        raise ValueError(f"a field on this grid needs {3 * n} values, got {a.size}")
    return a


def write_composite_h5(path, chain) -> None:
    """An ITK HDF5 transform file (itk_h5.rs:114-199). `chain` lists, in queue order, Affine
    objects and (Grid, values) pairs, values being 3 per voxel ((x, y, z) displacements, x fastest)."""
    elems = (ChainElement * max(1, len(chain)))()
    keep = []
    for i, item in enumerate(chain):
        if isinstance(item, Affine):
            elems[i].kind = FW_CHAIN_AFFINE
            elems[i].affine = item
        else:
            grid, values = item
            a = _field_array(grid, values)
            keep.append(a)
            elems[i].kind = FW_CHAIN_FIELD
            elems[i].grid = grid
            elems[i].values = a.ctypes.data_as(C.POINTER(C.c_double))
    if lib().fw_write_composite_h5(str(path).encode(), elems, len(chain)) != FW_OK:
        raise RegError(FW_ERR_INVALID, lib().fw_last_error(None).decode())


def read_composite_h5(path) -> list:
    """The chain of an ITK HDF5 transform file in queue order (itk_h5.rs:54-110): Affine objects
    and (Grid, values) pairs, values of shape (voxels, 3)."""
    count = C.c_uint32()
    name = str(path).encode()

    def call(elems, capacity):
        if lib().fw_read_composite_h5(name, elems, capacity, C.byref(count)) != FW_OK:
            raise RegError(FW_ERR_INVALID, lib().fw_last_error(None).decode())

    call(None, 0)
    n = count.value
    elems = (ChainElement * n)()
    call(elems, n)
    buffers = {}
    for i in range(n):
        if elems[i].kind == FW_CHAIN_FIELD:
            g = elems[i].grid
            buffers[i] = np.empty((int(g.size[0]) * int(g.size[1]) * int(g.size[2]), 3), dtype=np.float64)
            elems[i].values = buffers[i].ctypes.data_as(C.POINTER(C.c_double))
    if buffers:
        call(elems, n)
    out = []
    for i in range(n):
        if elems[i].kind == FW_CHAIN_FIELD:
            g = Grid()
            C.memmove(C.byref(g), C.byref(elems[i].grid), C.sizeof(Grid))
            out.append((g, buffers[i]))
        else:
            a = Affine()
            C.memmove(C.byref(a), C.byref(elems[i].affine), C.sizeof(Affine))
            out.append(a)
    return out


def write_composite_pair(prefix, affine: Affine, grid: Grid | None = None, field=None, inverse_field=None) -> None:
    """<prefix>Composite.h5 = [affine, field] and <prefix>InverseComposite.h5 = [inverse_field,
    affine^-1], or [affine] / [affine^-1] without a grid (ants_io.rs:100-141)."""
    if grid is None:
        rc = lib().fw_write_composite_pair(str(prefix).encode(), C.byref(affine), None, None, None)
    else:
        f = _field_array(grid, field)
        b = _field_array(grid, inverse_field)
        rc = lib().fw_write_composite_pair(str(prefix).encode(), C.byref(affine), C.byref(grid),
                                           f.ctypes.data_as(_D), b.ctypes.data_as(_D))
    if rc != FW_OK:
        raise RegError(FW_ERR_INVALID, lib().fw_last_error(None).decode())


def plan_capacity(plan: Plan, fixed_grid: Grid, moving_grid: Grid | None = None) -> Capacity:
    """Device bytes one registration of `plan` needs on this engine (no device is touched)."""
    out = Capacity()
    if lib().fw_plan_capacity(C.byref(plan), C.byref(fixed_grid), C.byref(moving_grid or fixed_grid), C.byref(out)) != FW_OK:
        raise RegError(-1, "capacity: invalid plan or grid")
    return out


def exported_symbols() -> list[str]:
    return [name for name, _, _ in _SIGNATURES]


def lib() -> C.CDLL:
    """The loaded gfx950 library; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"the HIP library is not built ({LIB_PATH}); run __graft_entry__.build()")
        handle = C.CDLL(str(LIB_PATH))
        for name, res, args in _SIGNATURES:
            fn = getattr(handle, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


class DeviceImage:
    def __init__(self, backend: "HipBackend", handle: C.c_void_p):
        self._b = backend
        self.handle = handle

    @property
    def grid(self) -> Grid:
        g = Grid()
        lib().fw_image_grid(self.handle, C.byref(g))
        return g

    def free(self):
        # a handle outliving its closed context is dropped: the context took the device memory
        if self.handle and self._b.ctx:
            lib().fw_image_free(self._b.ctx, self.handle)
        self.handle = None

    def take(self) -> C.c_void_p:
        h, self.handle = self.handle, None
        return h

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceField(DeviceImage):
    @property
    def grid(self) -> Grid:
        g = Grid()
        lib().fw_field_grid(self.handle, C.byref(g))
        return g

    def free(self):
        if self.handle and self._b.ctx:
            lib().fw_field_free(self._b.ctx, self.handle)
        self.handle = None


@dataclass
class Stats:
    iterations: int
    mean_error_norm: float
    max_error_norm: float


def _stats(s: InvertStats) -> Stats:
    return Stats(int(s.iterations), float(s.mean_error_norm), float(s.max_error_norm))


class HipBackend:
    """The HIP implementation of `Backend` (backend.rs:39-305) on one device."""

    def __init__(self, device: int = 0):
        self.ctx = C.c_void_p()
        st = lib().fw_ctx_create(device, C.byref(self.ctx))
        if st != FW_OK:
            raise RegError(st, lib().fw_last_error(None).decode())

    def close(self):
        if self.ctx:
            lib().fw_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def name(self) -> str:
        return "hip-gfx950"

    def set_profiling(self, on: bool):
        self._check(lib().fw_set_profiling(self.ctx, int(on)))

    def set_profiling_focus(self, family: str | None = None, every: int = 1):
        """Time only every `every`-th launch of `family` (None: every launch of every family)."""
        self._check(lib().fw_set_profiling_focus(self.ctx, (family or "").encode(), int(every)))

    def take_profile(self) -> dict[str, tuple[int, float, float, int]]:
        """Per kernel family: (launches, milliseconds of all launches, algorithmic bytes, timed
        launches). With a focus, a family's milliseconds are its timed launches' mean times its
        launches, and NaN for a family with no timed launch."""
        buf = C.create_string_buffer(1 << 16)
        self._check(lib().fw_take_profile(self.ctx, buf, len(buf)))
        out = {}
        for line in buf.value.decode().splitlines():
            name, launches, ms, nbytes, timed = line.split()
            n, t = int(launches), int(timed)
            out[name] = (n, float(ms) * n / t if t else float("nan"), float(nbytes), t)
        return out

    def _check(self, status: int):
        if status != FW_OK:
            raise RegError(status, lib().fw_last_error(self.ctx).decode())

    def _img(self, fn, *args) -> DeviceImage:
        out = C.c_void_p()
        self._check(fn(self.ctx, *args, C.byref(out)))
        return DeviceImage(self, out)

    def _fld(self, fn, *args) -> DeviceField:
        out = C.c_void_p()
        self._check(fn(self.ctx, *args, C.byref(out)))
        return DeviceField(self, out)

    # --- admission (capacity.rs, fits_gpu / registrations_at_once) ---
    def set_tuning(self, key: str, value: int):
        """One launch-shape choice of this context (fw_set_tuning); value < 0 removes it."""
        self._check(lib().fw_set_tuning(self.ctx, key.encode(), int(value)))

    def memory_budget(self) -> int:
        b = C.c_uint64()
        self._check(lib().fw_memory_budget(self.ctx, C.byref(b)))
        return b.value

    def fits(self, plan: Plan, fixed_grid: Grid, moving_grid: Grid | None = None) -> bool:
        f = C.c_int32()
        self._check(lib().fw_fits(self.ctx, C.byref(plan), C.byref(fixed_grid), C.byref(moving_grid or fixed_grid), C.byref(f)))
        return bool(f.value)

    def registrations_at_once(self, plan: Plan, fixed_grid: Grid, moving_grid: Grid | None = None, max: int = 1 << 16) -> int:
        n = C.c_uint32()
        self._check(lib().fw_registrations_at_once(self.ctx, C.byref(plan), C.byref(fixed_grid), C.byref(moving_grid or fixed_grid),
                                                   max, C.byref(n)))
        return n.value

    # --- transfers ---
    def release_idle_buffers(self):
        self._check(lib().fw_release_idle(self.ctx))

    def memory_stats(self) -> tuple[int, int]:
        held, peak = C.c_uint64(), C.c_uint64()
        self._check(lib().fw_memory_stats(self.ctx, C.byref(held), C.byref(peak)))
        return held.value, peak.value

    def upload_image(self, grid: Grid, data: np.ndarray) -> DeviceImage:
        d = np.ascontiguousarray(data, dtype=np.float64).reshape(-1)
        assert d.size == grid.n
        return self._img(lib().fw_image_upload, C.byref(grid), dptr(d))

    def download_image(self, image: DeviceImage) -> np.ndarray:
        out = np.empty(image.grid.n, dtype=np.float64)
        self._check(lib().fw_image_download(self.ctx, image.handle, dptr(out)))
        return out

    def zero_field(self, grid: Grid) -> DeviceField:
        return self._fld(lib().fw_field_zero, C.byref(grid))

    def upload_field(self, grid: Grid, data: np.ndarray) -> DeviceField:
        d = aos(data)
        assert d.shape[0] == grid.n
        return self._fld(lib().fw_field_upload, C.byref(grid), dptr(d))

    def download_field(self, field: DeviceField) -> np.ndarray:
        out = np.empty((field.grid.n, 3), dtype=np.float64)
        self._check(lib().fw_field_download(self.ctx, field.handle, dptr(out)))
        return out

    def field_grid(self, field: DeviceField) -> Grid:
        return field.grid

    def write_composite_fields(self, prefix, affine: Affine, forward: DeviceField, inverse: DeviceField) -> None:
        """<prefix>Composite.h5 / <prefix>InverseComposite.h5 from fields resident on this device
        (fw_write_composite_fields; ants_io.rs:100-141)."""
        self._check(lib().fw_write_composite_fields(self.ctx, str(prefix).encode(), C.byref(affine), forward.handle,
                                                    inverse.handle))

    def copy_field(self, field: DeviceField) -> DeviceField:
        return self._fld(lib().fw_field_copy, field.handle)

    # --- warps ---
    def warp_through_field(self, source: DeviceImage, field: DeviceField, pre=None, affine: Affine | None = None):
        return self._img(
            lib().fw_warp, source.handle, field.handle, pre.handle if pre else None, C.byref(affine) if affine else None
        )

    def warp_mask_through_field(self, mask, field, pre=None, affine=None, product: DeviceImage | None = None):
        return self._img(
            lib().fw_warp_mask,
            mask.handle,
            field.handle,
            pre.handle if pre else None,
            C.byref(affine) if affine else None,
            product.take() if product is not None else None,
        )

    # --- metrics ---
    def _cc(self, fixed, moving, valid, radius, norms):
        vals = (C.c_double * 2)()
        nrm = (C.c_double * 2)() if norms else None
        d0, d1 = C.c_void_p(), C.c_void_p()
        self._check(
            lib().fw_cc_pair(
                self.ctx,
                fixed.handle,
                moving.handle,
                valid.handle if valid is not None else None,
                radius,
                vals,
                nrm,
                C.byref(d0),
                C.byref(d1),
            )
        )
        pair = ((vals[0], DeviceField(self, d0)), (vals[1], DeviceField(self, d1)))
        return (pair, [nrm[0], nrm[1]]) if norms else pair

    def cc_value_derivative_pair(self, fixed, moving, radius: int):
        return self._cc(fixed, moving, None, radius, False)

    def cc_value_derivative_pair_with_norms(self, fixed, moving, radius: int):
        return self._cc(fixed, moving, None, radius, True)

    def cc_value_derivative_pair_masked(self, fixed, moving, valid, radius: int):
        return self._cc(fixed, moving, valid, radius, False)

    def cc_value_derivative_pair_masked_with_norms(self, fixed, moving, valid, radius: int):
        return self._cc(fixed, moving, valid, radius, True)

    def cc_value_derivative(self, fixed, moving, radius: int):
        return self._cc(fixed, moving, None, radius, False)[0]

    def _mi(self, fixed, moving, fm, mm, bins, norms):
        vals = (C.c_double * 2)()
        nrm = (C.c_double * 2)() if norms else None
        d0, d1 = C.c_void_p(), C.c_void_p()
        self._check(
            lib().fw_mi_dense_pair(
                self.ctx,
                fixed.handle,
                moving.handle,
                fm.handle if fm is not None else None,
                mm.handle if mm is not None else None,
                bins,
                vals,
                nrm,
                C.byref(d0),
                C.byref(d1),
            )
        )
        pair = ((vals[0], DeviceField(self, d0)), (vals[1], DeviceField(self, d1)))
        return (pair, [nrm[0], nrm[1]]) if norms else pair

    def mi_value_derivative_pair(self, fixed, moving, bins: int):
        return self._mi(fixed, moving, None, None, bins, False)

    def mi_value_derivative_pair_masked(self, fixed, moving, fixed_mask, moving_mask, bins: int):
        return self._mi(fixed, moving, fixed_mask, moving_mask, bins, False)

    def mi_value_derivative_pair_with_norms(self, fixed, moving, bins: int):
        return self._mi(fixed, moving, None, None, bins, True)

    # --- field algebra ---
    def smooth_image(self, image: DeviceImage, sigma_mm) -> DeviceImage:
        """Level smoothing on the device (pyramid.rs:429-437), per-axis sigma in mm."""
        s = np.ascontiguousarray(sigma_mm, dtype=np.float64).reshape(3)
        return self._img(lib().fw_smooth_image, image.handle, dptr(s))

    def resample_isotropic(self, image: DeviceImage, mm: float) -> DeviceImage:
        """The image on the pipeline's isotropic `mm` grid (register.rs:112-142), on the device."""
        return self._img(lib().fw_resample_isotropic, image.handle, C.c_double(mm))

    def preprocess_image(self, raw: DeviceImage, lower: float = 0.005, upper: float = 0.995) -> DeviceImage:
        """Winsorise and window a raw image on the device (preprocess.rs:159-232)."""
        return self._img(lib().fw_preprocess_image, raw.handle, C.c_float(lower), C.c_float(upper))

    def centre_of_gravity(self, image: DeviceImage) -> np.ndarray:
        """Intensity-weighted centre (mm) of an image on the device (init.rs:47-86)."""
        c = np.zeros(3)
        self._check(lib().fw_centre_of_gravity(self.ctx, image.handle, dptr(c)))
        return c

    def sample_values(self, image: DeviceImage, points: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """(nearest-voxel, linear) values at physical points, NaN outside."""
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        m = pts.shape[0]
        near, lin = np.empty(m), np.empty(m)
        self._check(lib().fw_sample_values(self.ctx, image.handle, dptr(pts), m, dptr(near), dptr(lin)))
        return near, lin

    def smooth_field(self, field: DeviceField, variance: float) -> DeviceField:
        return self._fld(lib().fw_smooth_field, field.take(), float(variance))

    def max_scaled_norm(self, field) -> float:
        out = (C.c_double * 2)()
        self._check(lib().fw_max_scaled_norm_pair(self.ctx, field.handle, None, out))
        return out[0]

    def max_scaled_norm_pair(self, a, b) -> tuple[float, float]:
        out = (C.c_double * 2)()
        self._check(lib().fw_max_scaled_norm_pair(self.ctx, a.handle, b.handle, out))
        return out[0], out[1]

    def l2_norm(self, field) -> float:
        out = C.c_double()
        self._check(lib().fw_l2_norm(self.ctx, field.handle, C.byref(out)))
        return out.value

    def scale(self, field, s: float):
        self._check(lib().fw_scale(self.ctx, field.handle, float(s)))

    def axpy(self, acc, a: float, x):
        self._check(lib().fw_axpy(self.ctx, acc.handle, float(a), x.handle))

    def compose(self, displacement, warping) -> DeviceField:
        return self._fld(lib().fw_compose, displacement.handle, warping.handle)

    def invert(self, field, initial: DeviceField, params: InvertParams | None = None):
        params = params or InvertParams.syn()
        out = C.c_void_p()
        st = InvertStats()
        self._check(lib().fw_invert(self.ctx, field.handle, initial.take(), C.byref(params), C.byref(out), C.byref(st)))
        return DeviceField(self, out), _stats(st)

    def invert_pair(self, field: DeviceField, inverse: DeviceField, params: InvertParams | None = None):
        params = params or InvertParams.syn()
        a, b = C.c_void_p(), C.c_void_p()
        sa, sb = InvertStats(), InvertStats()
        self._check(
            lib().fw_invert_pair(
                self.ctx,
                field.take(),
                inverse.take(),
                C.byref(params),
                C.byref(a),
                C.byref(sa),
                C.byref(b),
                C.byref(sb),
            )
        )
        return (DeviceField(self, a), _stats(sa)), (DeviceField(self, b), _stats(sb))

    def resample_field(self, field, grid: Grid) -> DeviceField:
        return self._fld(lib().fw_resample_field, field.handle, C.byref(grid))


class MiLevel:
    """`LinearMetric` of one prepared linear-stage level on the device (GpuMiLevel, mi.rs:144-524)."""

    def __init__(self, backend: HipBackend, moving: DeviceImage, points: np.ndarray, fixed_bin: np.ndarray, binning: MiBinning):
        self._b = backend
        self._moving = moving
        self._pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        self._bins = np.ascontiguousarray(fixed_bin, dtype=np.uint32)
        self.handle = C.c_void_p()
        backend._check(
            lib().fw_mi_level_create(
                backend.ctx, moving.handle, dptr(self._pts), u32ptr(self._bins), self._pts.shape[0], C.byref(binning), C.byref(self.handle)
            )
        )

    def value(self, t: LinearTransform) -> float:
        v = C.c_double()
        self._b._check(lib().fw_mi_level_value(self.handle, C.byref(t), C.byref(v)))
        return v.value

    def values(self, ts: list) -> np.ndarray:
        """The values of several transforms in batched passes (each equals value(t))."""
        arr = (LinearTransform * len(ts))(*ts)
        out = np.zeros(len(ts))
        self._b._check(lib().fw_mi_level_values(self.handle, arr, len(ts), dptr(out)))
        return out

    def value_and_derivative(self, t: LinearTransform):
        v = C.c_double()
        n = 6 if t.kind == 0 else 12
        d = np.zeros(n)
        self._b._check(lib().fw_mi_level_value_and_derivative(self.handle, C.byref(t), C.byref(v), dptr(d)))
        return v.value, d

    def histogram(self, t: LinearTransform, bins: int):
        """The merged histogram before the finish: (joint (bins, bins), fixed marginal, valid count)."""
        joint = np.zeros(bins * bins)
        marginal = np.zeros(bins)
        valid = C.c_int64()
        self._b._check(lib().fw_mi_level_histogram(self.handle, C.byref(t), dptr(joint), dptr(marginal), C.byref(valid)))
        return joint.reshape(bins, bins), marginal, valid.value

    def __del__(self):
        try:
            if self.handle and self._b.ctx:
                lib().fw_mi_level_free(self.handle)
            self.handle = None
        except Exception:
            pass
