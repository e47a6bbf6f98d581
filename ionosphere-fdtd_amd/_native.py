"""ctypes binding of the C ABI declared in include/geofdtd.h.

The shared library ``libgeofdtd.so`` is built in-tree by ``_build.build()``.
There is no fallback: if it is missing or no gfx950 device is present, the
caller gets ``BackendUnavailableError``.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

from .backends.base import BackendUnavailableError

LIBRARY_PATH = Path(__file__).resolve().parent / "libgeofdtd.so"

GEO_OK = 0
GEO_ERR_INVALID = -1
GEO_ERR_UNAVAILABLE = -2
GEO_ERR_HIP = -3

STORAGE_CODES = {"float32": 0, "float64": 1, "float16": 2, "bfloat16": 3}

_p_i64 = C.POINTER(C.c_int64)
_p_f64 = C.POINTER(C.c_double)


class GeoMeshDesc(C.Structure):
    _fields_ = [
        ("n_vertices", C.c_int64), ("n_edges", C.c_int64), ("n_faces", C.c_int64),
        ("edges", _p_i64), ("faces", _p_i64), ("face_edges", _p_i64),
        ("face_edge_signs", _p_i64), ("edge_left_faces", _p_i64),
        ("edge_right_faces", _p_i64), ("vertex_edges", _p_i64), ("vertex_edge_signs", _p_i64),
        ("max_degree", C.c_int64),
        ("primal_edge_angles", _p_f64), ("dual_edge_angles", _p_f64),
        ("face_solid_angles", _p_f64), ("dual_cell_solid_angles", _p_f64),
    ]


class GeoRadialDesc(C.Structure):
    _fields_ = [("radial_cells", C.c_int64), ("radii_m", _p_f64),
                ("full_spherical", C.c_int32), ("reserved", C.c_int32)]


class GeoPhysicsDesc(C.Structure):
    _fields_ = [
        ("storage", C.c_int32), ("scaling", C.c_int32), ("time_step_s", C.c_double),
        ("ca_er", _p_f64), ("cb_er", _p_f64), ("er_coef_rows", C.c_int64),
        ("ca_et", _p_f64), ("cb_et", _p_f64), ("et_coef_rows", C.c_int64),
        ("eps_er", _p_f64), ("eps_et", _p_f64),
        ("scale_er", C.c_double), ("scale_et", C.c_double), ("scale_hr", C.c_double),
        ("scale_ht", C.c_double),
        ("si_terms", C.c_int32), ("reserved1", C.c_int32),
        ("si_decay", _p_f64), ("si_drive", _p_f64), ("si_history_weights", _p_f64),
        ("si_scale", _p_f64), ("si_scale_rows", C.c_int64),
        ("n_source", C.c_int32), ("reserved2", C.c_int32),
        ("source_vertices", _p_i64), ("source_layers", _p_i64), ("source_weights", _p_f64),
        ("source_length_m", C.c_double),
        ("n_tsource", C.c_int32), ("reserved3", C.c_int32),
        ("tsource_edges", _p_i64), ("tsource_layers", _p_i64), ("tsource_weights", _p_f64),
        ("n_species", C.c_int32), ("magnetized", C.c_int32),
        ("band_lo", C.c_int32), ("band_hi", C.c_int32),
        ("plasma_reconstruction", _p_f64), ("face_tangents", _p_f64), ("face_centers", _p_f64),
        ("left_tangents", _p_f64), ("right_tangents", _p_f64),
        ("vertex_faces", _p_i64), ("vertex_face_weights", _p_f64), ("vf_width", C.c_int64),
        ("plasma_coef", _p_f64), ("plasma_coef_rows", C.c_int64),
        ("plasma_b", _p_f64), ("plasma_b_rows", C.c_int64),
        ("sigma_er", _p_f64), ("sigma_et", _p_f64),
        ("storage_rounding", C.c_int32), ("thread_mapping", C.c_int32),
    ]


class GeoState(C.Structure):
    _fields_ = [
        ("er", C.c_void_p), ("er_pitch", C.c_int64), ("ht", C.c_void_p), ("ht_pitch", C.c_int64),
        ("et", C.c_void_p), ("et_pitch", C.c_int64), ("hr", C.c_void_p), ("hr_pitch", C.c_int64),
        ("si_memory", C.c_void_p), ("plasma_current", C.c_void_p), ("plasma_pitch", C.c_int64),
    ]


class GeoLayout(C.Structure):
    _fields_ = [("node_pitch", C.c_int64), ("mid_pitch", C.c_int64), ("band_lo", C.c_int64),
                ("band_hi", C.c_int64), ("plasma_pitch", C.c_int64),
                ("plasma_components", C.c_int64),
                ("vector_elems", C.c_int64), ("compute_bytes", C.c_int64),
                ("si_pitch", C.c_int64), ("plasma_basis", C.c_int64),
                ("radial_closed_form", C.c_int64), ("h_thread_mapping", C.c_int64),
                ("e_thread_mapping", C.c_int64)]


class GeoDiagOut(C.Structure):
    _fields_ = [("max_abs_er", C.c_double), ("max_abs_et", C.c_double),
                ("max_abs_hr", C.c_double), ("max_abs_ht", C.c_double),
                ("energy_er", C.c_double), ("energy_et", C.c_double),
                ("energy_hr", C.c_double), ("energy_ht", C.c_double),
                ("nonfinite", C.c_int64),
                ("divb_max_abs", C.c_double), ("divb_flux_scale", C.c_double),
                ("storage_overflow", C.c_int64),
                ("loss_er", C.c_double), ("loss_et", C.c_double),
                ("sumsq_er", C.c_double), ("sumsq_et", C.c_double),
                ("sumsq_hr", C.c_double), ("sumsq_ht", C.c_double),
                ("nonfinite_er", C.c_int64), ("nonfinite_et", C.c_int64),
                ("nonfinite_hr", C.c_int64), ("nonfinite_ht", C.c_int64)]


class GeoProfiles(C.Structure):
    _fields_ = [("energy_er", _p_f64), ("energy_et", _p_f64), ("energy_hr", _p_f64),
                ("energy_ht", _p_f64), ("loss_er", _p_f64), ("loss_et", _p_f64)]


EXPORTED_SYMBOLS = (
    "geo_plan_create", "geo_plan_layout", "geo_plan_bind", "geo_step", "geo_step_prepare",
    "geo_plan_set_step_index", "geo_set_receivers", "geo_set_h_receivers", "geo_step_h",
    "geo_sample_er", "geo_diagnostics", "geo_radial_profiles", "geo_time_kernels",
    "geo_plan_destroy",
    "geo_last_error", "geo_version", "geo_stream_probe", "geo_get_storage_scales",
    "geo_set_storage_scales", "geo_ensemble_create", "geo_ensemble_configure",
    "geo_ensemble_tune", "geo_ensemble_step", "geo_ensemble_destroy",
    "geo_mesh_ops_create", "geo_mesh_op", "geo_array_diff", "geo_mesh_ops_destroy",
    # include/geosurrogate.h
    "geo_gemm_bf16", "geo_gemm_bf16_noalloc", "geo_f32_to_bf16", "geo_silu_forward",
    "geo_silu_backward", "geo_column_sums", "geo_im2col_up", "geo_col2im_up", "geo_adam_step",
    "geo_trace_loss",
)

_lib = None


def load_library(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load libgeofdtd.so (no device call happens here)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    target = Path(path) if path is not None else LIBRARY_PATH
    if not target.exists():
        raise BackendUnavailableError(
            f"HIP extension {target} is not built; run __graft_entry__.build()")
    lib = C.CDLL(str(target))
    vp = C.c_void_p
    lib.geo_plan_create.argtypes = [C.POINTER(GeoMeshDesc), C.POINTER(GeoRadialDesc),
                                    C.POINTER(GeoPhysicsDesc), C.POINTER(vp)]
    lib.geo_plan_layout.argtypes = [vp, C.POINTER(GeoLayout)]
    lib.geo_plan_bind.argtypes = [vp, C.POINTER(GeoState)]
    lib.geo_step.argtypes = [vp, vp, C.c_int64, C.c_int32, vp, vp]
    lib.geo_step_prepare.argtypes = [vp, C.c_int32]
    lib.geo_plan_set_step_index.argtypes = [vp, C.c_int64, vp]
    lib.geo_set_receivers.argtypes = [vp, _p_i64, _p_i64, _p_f64, C.c_int32]
    lib.geo_set_h_receivers.argtypes = [vp, _p_i64, _p_i64, _p_f64, C.c_int32, C.c_int32,
                                        _p_i64, _p_i64, _p_f64, C.c_int32, C.c_int32]
    lib.geo_step_h.argtypes = [vp, vp, C.c_int64, C.c_int32, C.c_int32, vp, vp, vp]
    lib.geo_sample_er.argtypes = [vp, vp, vp]
    lib.geo_diagnostics.argtypes = [vp, C.POINTER(GeoDiagOut), vp]
    lib.geo_radial_profiles.argtypes = [vp, C.POINTER(GeoProfiles), vp]
    lib.geo_time_kernels.argtypes = [vp, C.c_int64, vp, _p_f64]
    lib.geo_stream_probe.argtypes = [vp, vp, C.c_int64, C.c_int32, C.c_int32, vp, _p_f64]
    lib.geo_get_storage_scales.argtypes = [vp, _p_f64, vp]
    lib.geo_set_storage_scales.argtypes = [vp, _p_f64, vp]
    lib.geo_plan_destroy.argtypes = [vp]
    lib.geo_plan_destroy.restype = None
    lib.geo_ensemble_create.argtypes = [C.POINTER(vp), C.c_int32, C.POINTER(vp)]
    lib.geo_ensemble_configure.argtypes = [vp, C.c_int32, C.c_int32]
    lib.geo_ensemble_tune.argtypes = [vp, C.c_int32, vp, C.POINTER(C.c_int32),
                                      C.POINTER(C.c_int32), _p_f64]
    lib.geo_ensemble_step.argtypes = [vp, vp, C.c_int64, C.c_int32, vp, vp]
    lib.geo_ensemble_destroy.argtypes = [vp]
    lib.geo_ensemble_destroy.restype = None
    lib.geo_mesh_ops_create.argtypes = [C.POINTER(GeoMeshDesc), C.POINTER(vp)]
    lib.geo_mesh_op.argtypes = [vp, C.c_int32, C.c_int32, vp, C.c_int64, vp, vp]
    lib.geo_array_diff.argtypes = [C.c_int32, vp, C.c_int64, C.c_int64, C.c_int32, vp, vp]
    lib.geo_mesh_ops_destroy.argtypes = [vp]
    lib.geo_mesh_ops_destroy.restype = None
    i64, i32, f32 = C.c_int64, C.c_int32, C.c_float
    lib.geo_gemm_bf16.argtypes = [i64, i64, i64, vp, i64, i64, vp, i64, i64, vp, vp, vp, i64,
                                  i32, vp]
    lib.geo_gemm_bf16_noalloc.argtypes = list(lib.geo_gemm_bf16.argtypes)
    lib.geo_f32_to_bf16.argtypes = [i64, vp, vp, vp]
    lib.geo_silu_forward.argtypes = [i64, vp, vp, vp]
    lib.geo_silu_backward.argtypes = [i64, vp, vp, vp, vp, vp]
    lib.geo_column_sums.argtypes = [i64, i64, vp, i64, vp, vp]
    lib.geo_im2col_up.argtypes = [i64, i64, i64, i32, i32, vp, vp, vp]
    lib.geo_col2im_up.argtypes = [i64, i64, i64, i32, i32, vp, vp, vp]
    lib.geo_adam_step.argtypes = [i64, vp, vp, vp, vp, vp, f32, f32, f32, f32, i64, vp]
    lib.geo_trace_loss.argtypes = [i64, i64, vp, vp, vp, i64, vp, f32, vp, vp, vp, vp, vp]
    lib.geo_last_error.restype = C.c_char_p
    lib.geo_last_error.argtypes = []
    lib.geo_version.restype = C.c_int
    for name in EXPORTED_SYMBOLS:
        if name not in ("geo_plan_destroy", "geo_last_error", "geo_version",
                        "geo_ensemble_destroy"):
            getattr(lib, name).restype = C.c_int
    if path is None:
        _lib = lib
    return lib


def stream_probe(torch, bytes_: int, mode: int, repeats: int = 10) -> float:
    """Achieved GB/s of the library's 16-byte-strip probe kernel (see the header)."""
    lib = load_library()
    n = bytes_ // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.zeros_like(src)
    out = C.c_double()
    check(lib.geo_stream_probe(dst.data_ptr(), src.data_ptr(), n * 4, mode, repeats,
                               C.c_void_p(torch.cuda.current_stream().cuda_stream),
                               C.byref(out)), lib)
    if mode in (8, 9):  # rows of 272 B read and written, a third as many gathered rows read
        rows = (n * 4) // 272
        moved = 2 * rows * 272 + (rows // 3) * 272 + (8 * rows if mode == 9 else 0)
    else:
        moved = (2 if (mode & 3) == 0 else 1) * n * 4
    return moved / (out.value * 1e-3) / 1e9


def check(code: int, lib: C.CDLL | None = None) -> None:
    """Raise the Python exception matching a C ABI error code."""
    if code == GEO_OK:
        return
    lib = lib or load_library()
    message = (lib.geo_last_error() or b"").decode(errors="replace")
    if code == GEO_ERR_INVALID:
        raise ValueError(message)
    if code == GEO_ERR_UNAVAILABLE:
        raise BackendUnavailableError(message)
    raise RuntimeError(message)


def f64_ptr(array):
    """ctypes double* into a C-contiguous float64 numpy array (None -> NULL)."""
    if array is None:
        return None
    return array.ctypes.data_as(_p_f64)


def i64_ptr(array):
    if array is None:
        return None
    return array.ctypes.data_as(_p_i64)
