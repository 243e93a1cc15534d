"""ctypes mirror of the structs in include/fastwarp_hip.h (shared by the library and test bindings)."""
from __future__ import annotations

import ctypes as C

import numpy as np


class Grid(C.Structure):
    """fw_grid: an ITK grid in LPS mm (crates/fastwarp-reg/src/grid.rs:26-44)."""

    _fields_ = [
        ("size", C.c_int64 * 3),
        ("spacing", C.c_double * 3),
        ("origin", C.c_double * 3),
        ("direction", C.c_double * 9),
    ]

    @classmethod
    def make(cls, size, spacing=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0), direction=None) -> "Grid":
        g = cls()
        for a in range(3):
            g.size[a] = int(size[a])
            g.spacing[a] = float(spacing[a])
            g.origin[a] = float(origin[a])
        d = np.eye(3) if direction is None else np.asarray(direction, dtype=np.float64).reshape(3, 3)
        for i, v in enumerate(d.reshape(-1)):
            g.direction[i] = float(v)
        return g

    @property
    def shape_xyz(self) -> tuple[int, int, int]:
        return (self.size[0], self.size[1], self.size[2])

    @property
    def n(self) -> int:
        return self.size[0] * self.size[1] * self.size[2]

    def dir_matrix(self) -> np.ndarray:
        return np.array(list(self.direction), dtype=np.float64).reshape(3, 3)

    def index_to_point(self, idx) -> np.ndarray:
        idx = np.asarray(idx, dtype=np.float64)
        return np.array(list(self.origin)) + self.dir_matrix() @ (np.array(list(self.spacing)) * idx)

    def same(self, other: "Grid") -> bool:
        return (
            tuple(self.size) == tuple(other.size)
            and tuple(self.spacing) == tuple(other.spacing)
            and tuple(self.origin) == tuple(other.origin)
            and tuple(self.direction) == tuple(other.direction)
        )


class Affine(C.Structure):
    """fw_affine: y = M (x - c) + c + t (transform/affine.rs:34-45)."""

    _fields_ = [("matrix", C.c_double * 9), ("translation", C.c_double * 3), ("center", C.c_double * 3)]

    @classmethod
    def make(cls, matrix=None, translation=(0, 0, 0), center=(0, 0, 0)) -> "Affine":
        a = cls()
        m = np.eye(3) if matrix is None else np.asarray(matrix, dtype=np.float64).reshape(3, 3)
        for i, v in enumerate(m.reshape(-1)):
            a.matrix[i] = float(v)
        for i in range(3):
            a.translation[i] = float(translation[i])
            a.center[i] = float(center[i])
        return a

    def apply(self, p) -> np.ndarray:
        m = np.array(list(self.matrix)).reshape(3, 3)
        c = np.array(list(self.center))
        return m @ (np.asarray(p, dtype=np.float64) - c) + c + np.array(list(self.translation))


FW_CHAIN_AFFINE = 0
FW_CHAIN_FIELD = 1


class ChainElement(C.Structure):
    """fw_chain_element: one element of an ITK transform chain (an affine or a displacement field)."""

    _fields_ = [
        ("kind", C.c_int32),
        ("affine", Affine),
        ("grid", Grid),
        ("values", C.POINTER(C.c_double)),
    ]


class InvertParams(C.Structure):
    _fields_ = [
        ("max_iterations", C.c_uint32),
        ("mean_error_tolerance", C.c_double),
        ("max_error_tolerance", C.c_double),
        ("enforce_boundary_condition", C.c_int32),
    ]

    @classmethod
    def syn(cls) -> "InvertParams":
        """InvertParams::SYN (field_ops.rs:290-295)."""
        return cls(20, 0.001, 0.1, 1)


class InvertStats(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("mean_error_norm", C.c_double), ("max_error_norm", C.c_double)]


class MiBinning(C.Structure):
    _fields_ = [
        ("bins", C.c_int64),
        ("fixed_min", C.c_double),
        ("fixed_max", C.c_double),
        ("moving_min", C.c_double),
        ("moving_max", C.c_double),
        ("fixed_bin_size", C.c_double),
        ("moving_bin_size", C.c_double),
        ("fixed_normalized_min", C.c_double),
        ("moving_normalized_min", C.c_double),
    ]

    @classmethod
    def from_ranges(cls, bins, fixed, moving) -> "MiBinning":
        """MiBinning::from_ranges (metric/mi.rs:238-276)."""
        b = cls()
        b.bins = bins
        b.fixed_min, b.fixed_max = fixed
        b.moving_min, b.moving_max = moving
        span = float(bins - 4)
        b.fixed_bin_size = (b.fixed_max - b.fixed_min) / span
        b.moving_bin_size = (b.moving_max - b.moving_min) / span
        b.fixed_normalized_min = b.fixed_min / b.fixed_bin_size - 2.0
        b.moving_normalized_min = b.moving_min / b.moving_bin_size - 2.0
        return b


class LinearTransform(C.Structure):
    """fw_linear_transform: x -> outer(optimised(x)) (transform/param.rs:306-315)."""

    _fields_ = [
        ("outer", Affine),
        ("kind", C.c_int32),
        ("parameters", C.c_double * 12),
        ("center", C.c_double * 3),
    ]

    @classmethod
    def make(cls, kind: int, params, center=(0, 0, 0), outer: Affine | None = None) -> "LinearTransform":
        t = cls()
        t.outer = outer if outer is not None else Affine.make()
        t.kind = kind
        for i, v in enumerate(params):
            t.parameters[i] = float(v)
        for i in range(3):
            t.center[i] = float(center[i])
        return t


def aos(field: np.ndarray) -> np.ndarray:
    """A field as a contiguous f64 (n, 3) array."""
    return np.ascontiguousarray(field, dtype=np.float64).reshape(-1, 3)


def dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def u32ptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


class Plan(C.Structure):
    """fw_plan: the stage plan of one full normalisation (include/fastwarp_hip.h)."""

    _fields_ = [
        ("linear_levels", C.c_int32),
        ("linear_iterations", C.c_uint32 * 8),
        ("linear_shrink", C.c_uint32 * 8),
        ("linear_sigmas", C.c_double * 8),
        ("linear_gradient_step", C.c_double),
        ("mi_bins", C.c_int32),
        ("sampling_percentage", C.c_double),
        ("seed", C.c_uint32),
        ("linear_convergence_threshold", C.c_double),
        ("linear_convergence_window", C.c_int32),
        ("syn_levels", C.c_int32),
        ("syn_iterations", C.c_uint32 * 8),
        ("syn_shrink", C.c_uint32 * 8),
        ("syn_sigmas", C.c_double * 8),
        ("syn_gradient_step", C.c_double),
        ("update_variance", C.c_double),
        ("total_variance", C.c_double),
        ("cc_radius", C.c_int32),
        ("channels", C.c_int32),
        ("channel_weights", C.c_double * 8),
        ("syn_convergence_threshold", C.c_double),
        ("syn_convergence_window", C.c_int32),
        ("winsorize_lower", C.c_double),
        ("winsorize_upper", C.c_double),
        ("run_linear", C.c_int32),
        ("run_syn", C.c_int32),
        ("initial_kind", C.c_int32),
        ("initial_affine", Affine),
        ("syn_channel_metrics", C.c_int32 * 8),
    ]

    def set_linear(self, iterations, shrink, sigmas):
        self.linear_levels = len(iterations)
        for i, (it, sh, sg) in enumerate(zip(iterations, shrink, sigmas)):
            self.linear_iterations[i], self.linear_shrink[i], self.linear_sigmas[i] = it, sh, sg

    def set_syn(self, iterations, shrink, sigmas):
        self.syn_levels = len(iterations)
        for i, (it, sh, sg) in enumerate(zip(iterations, shrink, sigmas)):
            self.syn_iterations[i], self.syn_shrink[i], self.syn_sigmas[i] = it, sh, sg


class Capacity(C.Structure):
    """fw_capacity: device bytes of one registration on this engine (include/fastwarp_hip.h)."""

    _fields_ = [
        ("fixed_voxels", C.c_uint64),
        ("moving_voxels", C.c_uint64),
        ("smoothing_cache_bytes", C.c_uint64),
        ("linear_bytes", C.c_uint64),
        ("syn_bytes", C.c_uint64),
        ("peak_bytes", C.c_uint64),
        ("largest_buffer_bytes", C.c_uint64),
    ]


class Result(C.Structure):
    """fw_result: the collapsed affine, phase timings and per-level counts of one registration."""

    _fields_ = [
        ("affine", Affine),
        ("seconds_total", C.c_double),
        ("seconds_preprocess", C.c_double),
        ("seconds_linear", C.c_double),
        ("seconds_syn", C.c_double),
        ("syn_levels", C.c_int32),
        ("syn_iterations", C.c_int32 * 8),
        ("syn_last_energy", C.c_double * 8),
        ("linear_iterations", (C.c_int32 * 8) * 2),
        ("linear_final_value", C.c_double * 2),
        ("rigid_parameters", C.c_double * 6),
        ("affine_parameters", C.c_double * 12),
        ("syn_level_seconds", C.c_double * 8),
        ("syn_iteration_seconds", C.c_double * 8),
        ("linear_level_seconds", (C.c_double * 8) * 2),
        ("linear_iteration_seconds", (C.c_double * 8) * 2),
        ("linear_evaluations", (C.c_int64 * 8) * 2),
        ("syn_inversion_steps", C.c_int64 * 8),
    ]


class TraceStep(C.Structure):
    """fw_trace_step: one CG iteration (stage 0 rigid, 1 affine) or SyN iteration (stage 2)."""

    _fields_ = [
        ("stage", C.c_int32),
        ("level", C.c_int32),
        ("iteration", C.c_int32),
        ("pad", C.c_int32),
        ("value", C.c_double),
        ("convergence_value", C.c_double),
        ("update_scale", C.c_double * 2),
        ("inversion_steps", C.c_int32 * 4),
        ("inversion_mean_error", C.c_double * 4),
        ("inversion_max_error", C.c_double * 4),
    ]


def read_trace(fn) -> list[TraceStep]:
    """The records of the last registration on this thread from `fn` (fw_last_trace or the
    oracle's counterpart)."""
    n = fn(None, 0)
    buf = (TraceStep * max(n, 1))()
    fn(buf, n)
    return list(buf[:n])


class ServedTiming(C.Structure):
    """fw_served_timing: where a served request's time went."""

    _fields_ = [("queued_s", C.c_double), ("run_s", C.c_double), ("total_s", C.c_double),
                ("footprint_bytes", C.c_uint64), ("slot", C.c_int32), ("pad", C.c_int32)]
