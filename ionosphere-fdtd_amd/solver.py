"""GeodesicFDTD on MI355X: the reference's solver API over the HIP step library.

``SimulationConfig`` and ``GeodesicFDTD`` keep the reference's constructor,
validation, attributes and methods (solver.py:30-1294).  Setup stays on the
host in FP64 (mesh, radial grid, material sampling, CFL, exponential loss
coefficients, source / receiver distributions, surface-impedance and plasma
tables); ``libgeofdtd`` turns those tables into device tables once, and every
``step`` runs the two fused HIP kernels per leapfrog step -- as a captured HIP
graph of ``compile_chunk_size`` steps when ``compile_step=True``, otherwise
as individual launches of the same kernels.

Fields live in padded radial-innermost device buffers; ``er``, ``et``, ``hr``
and ``ht`` are (entity, radial) views of them for FP32/FP64 storage, so
in-place edits such as ``sim.er[:, k] = mode`` reach the kernels.  With
FP16/BF16 storage the properties return unscaled FP32 copies and assignment
goes through the setters.
"""

from __future__ import annotations

import ctypes as C
import dataclasses
from dataclasses import dataclass
from pathlib import Path
from typing import Any

import numpy as np
from numpy.typing import NDArray

from . import _native
from .backends import ArrayBackend, BackendUnavailableError, create_backend
from .constants import C_0, EARTH_RADIUS_M, EPSILON_0, MU_0
from .materials import (EarthIonosphereMaterial, blend_cell_anomalies, blend_node_anomalies,
                        conservative_anomaly_fractions)
from .mesh import GeodesicMesh, build_geodesic_mesh, vertex_incidence
from .plasma import MeshPlasmaModel, plasma_tables
from .radial_grid import validate_radial_grid
from .sources import GaussianCurrent, TangentialGaussianCurrent
from .surface_impedance import ConductiveHalfSpaceSurface, surface_ade_tables

DEFAULT_NARROW_SCALES = {"er": 1.0, "et": 1.0, "hr": 256.0, "ht": 256.0}
_FIELD_ORDER = ("er", "et", "hr", "ht")


# Rounding of narrow (FP16/BF16) stored values: the C ABI's storage_rounding codes.
STORAGE_ROUNDING = {"nearest": 0, "stochastic": 1, "rotation": 2}
# Thread mapping of the step kernels: the C ABI's GEO_MAPPING_* codes.
THREAD_MAPPING = {"auto": 0, "rows": 1, "flat": 2}
# A source envelope longer than this many steps drives quasi-static fields.
STAGNATION_STEPS = 1.0e4


def _is_scale(value) -> bool:
    """A power of two in [2^-60, 2^60] (exact to multiply by in every storage type)."""
    try:
        v = float(value)
    except (TypeError, ValueError):
        return False
    if not (v > 0.0 and np.isfinite(v)):
        return False
    mant, exp = np.frexp(v)
    return mant == 0.5 and -60 <= exp - 1 <= 60


@dataclass(frozen=True, slots=True)
class SimulationConfig:
    """Geometry and integration controls (reference solver.py:30-171)."""

    subdivision: int = 2
    radial_cells: int = 24
    minimum_altitude_m: float = -100_000.0
    maximum_altitude_m: float = 100_000.0
    earth_radius_m: float = EARTH_RADIUS_M
    courant_factor: float = 0.35
    time_step_s: float | None = None
    mesh_relaxations: int = 0
    mesh_optimization_steps: int = 0
    mesh_orientation: str = "polar"
    radial_altitudes_m: tuple[float, ...] | None = None
    radial_material_support: str = "point"
    tangential_material_support: str = "point"
    horizontal_anomaly_mode: str = "point"
    radial_boundary_condition: str = "pec"
    loss_integration: str = "exponential"
    radial_grid_policy: str = "smooth"
    geometry_mode: str = "full-spherical"
    compress_uniform_material_coefficients: bool = False

    def __post_init__(self) -> None:
        ints = (self.subdivision, self.radial_cells, self.mesh_relaxations,
                self.mesh_optimization_steps)
        if any(isinstance(v, bool) or not isinstance(v, (int, np.integer)) for v in ints):
            raise ValueError("mesh and radial cell controls must be integers")
        if self.subdivision < 0:
            raise ValueError("subdivision must be non-negative")
        if self.radial_cells < 2:
            raise ValueError("radial_cells must be at least 2")
        if self.mesh_relaxations < 0:
            raise ValueError("mesh_relaxations must be non-negative")
        if self.mesh_optimization_steps < 0:
            raise ValueError("mesh_optimization_steps must be non-negative")
        if self.mesh_orientation not in {"native", "polar"}:
            raise ValueError("mesh_orientation must be 'native' or 'polar'")
        if self.radial_material_support not in {"point", "dual-cell"}:
            raise ValueError("radial_material_support must be 'point' or 'dual-cell'")
        if self.tangential_material_support not in {"point", "edge-diamond"}:
            raise ValueError("tangential_material_support must be 'point' or 'edge-diamond'")
        if self.horizontal_anomaly_mode not in {"point", "conservative-nearest"}:
            raise ValueError("horizontal_anomaly_mode must be 'point' or 'conservative-nearest'")
        if self.radial_boundary_condition not in {"pec", "surface-impedance"}:
            raise ValueError("radial_boundary_condition must be 'pec' or 'surface-impedance'")
        if self.radial_boundary_condition == "surface-impedance" and \
                self.minimum_altitude_m != 0.0:
            raise ValueError("surface-impedance boundary requires minimum_altitude_m=0")
        if self.loss_integration not in {"exponential", "trapezoidal"}:
            raise ValueError("loss_integration must be 'exponential' or 'trapezoidal'")
        if self.radial_grid_policy not in {"smooth", "balanced-2to1", "allow-abrupt"}:
            raise ValueError("radial_grid_policy must be 'smooth', 'balanced-2to1', "
                             "or 'allow-abrupt'")
        if self.geometry_mode not in {"full-spherical", "thin-shell"}:
            raise ValueError("geometry_mode must be 'full-spherical' or 'thin-shell'")
        finite = (self.minimum_altitude_m, self.maximum_altitude_m, self.earth_radius_m,
                  self.courant_factor)
        if not all(np.isfinite(v) for v in finite):
            raise ValueError("geometry and Courant controls must be finite")
        if self.minimum_altitude_m >= self.maximum_altitude_m:
            raise ValueError("altitude bounds are reversed")
        if self.earth_radius_m + self.minimum_altitude_m <= 0.0:
            raise ValueError("minimum radius must be positive")
        if not 0.0 < self.courant_factor <= 1.0:
            raise ValueError("courant_factor must be in (0, 1]")
        if self.time_step_s is not None and (not np.isfinite(self.time_step_s)
                                             or self.time_step_s <= 0.0):
            raise ValueError("time_step_s must be finite and positive")
        if self.radial_altitudes_m is not None:
            alt = np.asarray(self.radial_altitudes_m, dtype=np.float64)
            if (alt.ndim != 1 or len(alt) < 3 or not np.all(np.isfinite(alt))
                    or not np.all(np.diff(alt) > 0.0)):
                raise ValueError("radial_altitudes_m must contain at least three finite, "
                                 "increasing values")
            if len(alt) != self.radial_cells + 1:
                raise ValueError("radial_cells must equal len(radial_altitudes_m) - 1")
            if alt[0] != self.minimum_altitude_m or alt[-1] != self.maximum_altitude_m:
                raise ValueError(
                    "custom radial altitudes must include the configured altitude bounds")
            steps = np.diff(alt)
            change = self.radial_cells * np.abs(np.diff(steps)) / np.minimum(steps[:-1],
                                                                              steps[1:])
            if self.radial_grid_policy == "smooth" and np.any(change > 8.0 * (1.0 + 1.0e-12)):
                raise ValueError(
                    "custom radial grid is not smoothly graded; use a smooth node mapping or "
                    "explicitly select radial_grid_policy='allow-abrupt'")
            if self.radial_grid_policy == "balanced-2to1":
                validate_radial_grid(alt, maximum_adjacent_step_ratio=2.0)


class _SurfaceImpedanceState:
    """Stand-in for the reference's SurfaceImpedanceADE attribute (memory view)."""

    def __init__(self, owner: "GeodesicFDTD", terms: int):
        self._owner = owner
        self.terms = terms

    @property
    def memory(self):
        return self._owner._si_memory

    @memory.setter
    def memory(self, values) -> None:
        self._owner._si_memory.copy_(self._owner._tensor(values, self._owner._si_memory.dtype))

    @property
    def state_bytes(self) -> int:
        return self._owner.backend.nbytes(self._owner._si_memory)


class GeodesicFDTD:
    """Earth-ionosphere FDTD on staggered geodesic radial planes (HIP backend).

    ``er``/``ht`` live on integer radial planes, ``hr``/``et`` halfway between
    them; magnetic fields lag the electric fields by half a step.

    ``dtype`` selects the field storage: ``"float32"`` (default), ``"float64"``
    or the narrow ``"float16"`` / ``"bfloat16"``, which keep FP32 update
    arithmetic and FP32 plasma and surface-impedance state.  Narrow storage
    carries ~1e-3..1e-2 relative error against FP64 (FP16; BF16 ~5x more).
    ``storage_rounding`` picks how each stored narrow value is rounded:
    ``"rotation"`` (default; stochastic, with a hashed per-element phase
    advanced by the golden ratio each step), ``"stochastic"`` (independent
    hashed thresholds every step) or ``"nearest"``.  Rounding to nearest
    stagnates once a field's per-step change drops below half an ulp, as it
    does under slowly varying sources (the 20 Hz benchmark source: O(1)
    errors after 5000-10000 steps); the constructor warns for that case.
    ``thread_mapping`` picks how the step kernels' threads cover the field
    rows: ``"rows"`` (whole rows per 256-thread block), ``"flat"`` (rows may
    straddle blocks, every lane busy) or ``"auto"`` (default: flat for a
    kernel whose whole-row blocks would leave 16 or more lanes idle, as at
    80 or 96 radial cells).  All three give the same bits.
    """

    def __init__(
        self,
        config: SimulationConfig | None = None,
        material: Any | None = None,
        source: GaussianCurrent | TangentialGaussianCurrent | None = None,
        surface_impedance: ConductiveHalfSpaceSurface | None = None,
        plasma: MeshPlasmaModel | None = None,
        mesh: GeodesicMesh | None = None,
        backend: str = "hip",
        device: str = "auto",
        dtype: str = "auto",
        compile_step: bool = False,
        compile_chunk_size: int = 8,
        torch_threads: int | None = None,
        storage_scales: dict[str, float] | None = None,
        storage_scaling: str = "auto",
        storage_rounding: str = "rotation",
        thread_mapping: str = "auto",
    ) -> None:
        if (isinstance(compile_chunk_size, bool)
                or not isinstance(compile_chunk_size, (int, np.integer))
                or compile_chunk_size < 1):
            raise ValueError("compile_chunk_size must be a positive integer")
        if storage_rounding not in STORAGE_ROUNDING:
            raise ValueError("storage_rounding must be one of " + ", ".join(
                repr(k) for k in STORAGE_ROUNDING))
        if thread_mapping not in THREAD_MAPPING:
            raise ValueError("thread_mapping must be one of " + ", ".join(
                repr(k) for k in THREAD_MAPPING))
        self.compile_chunk_size = int(compile_chunk_size)
        self.config = config or SimulationConfig()
        cfg = self.config
        if mesh is None:
            self.mesh = build_geodesic_mesh(cfg.subdivision, cfg.mesh_relaxations,
                                            cfg.mesh_orientation, cfg.mesh_optimization_steps)
        else:
            if cfg.mesh_relaxations or cfg.mesh_optimization_steps:
                raise ValueError("mesh relaxation and optimization controls cannot accompany "
                                 "a provided mesh")
            if mesh.topology_kind == "uniform":
                polar = np.count_nonzero((mesh.vertex_degree == 5) & np.isclose(
                    np.abs(mesh.vertices[:, 2]), 1.0, rtol=0.0, atol=1.0e-13))
                if polar != (2 if cfg.mesh_orientation == "polar" else 0):
                    raise ValueError("provided mesh orientation does not match config")
            self.mesh = mesh
        if self.mesh.subdivision is not None and self.mesh.subdivision != cfg.subdivision:
            raise ValueError("provided mesh subdivision does not match config")
        self.backend: ArrayBackend = create_backend(backend, self.mesh, device=device,
                                                    dtype=dtype, torch_threads=torch_threads)
        self._lib = _native.load_library()
        self.material = material or EarthIonosphereMaterial()
        self.source = source
        self.surface_impedance = surface_impedance
        self.plasma = plasma
        if cfg.radial_boundary_condition == "surface-impedance":
            if surface_impedance is None:
                raise ValueError("surface-impedance boundary requires a surface model")
        elif surface_impedance is not None:
            raise ValueError("surface impedance model requires radial_boundary_condition="
                             "'surface-impedance'")
        # radial grid (solver.py:260-283)
        if cfg.radial_altitudes_m is None:
            self.altitudes_m = np.linspace(cfg.minimum_altitude_m, cfg.maximum_altitude_m,
                                           cfg.radial_cells + 1)
        else:
            self.altitudes_m = np.asarray(cfg.radial_altitudes_m, dtype=np.float64)
        self.radii_m = cfg.earth_radius_m + self.altitudes_m
        self.radial_midpoints_m = 0.5 * (self.radii_m[:-1] + self.radii_m[1:])
        self.radial_midpoint_altitudes_m = self.radial_midpoints_m - cfg.earth_radius_m
        self.radial_steps_m = np.diff(self.radii_m)
        ctrl = np.empty(len(self.radii_m))
        ctrl[0] = 0.5 * self.radial_steps_m[0]
        ctrl[-1] = 0.5 * self.radial_steps_m[-1]
        ctrl[1:-1] = np.diff(self.radial_midpoints_m)
        self.radial_node_control_lengths_m = ctrl
        if plasma is not None:
            plasma.validate_grid(self.mesh, self.radial_midpoint_altitudes_m)
        sampled = sample_material(self.material, self.mesh, self.altitudes_m,
                                  self.radial_midpoint_altitudes_m, cfg.earth_radius_m,
                                  radial_support=cfg.radial_material_support,
                                  tangential_support=cfg.tangential_material_support,
                                  anomaly_mode=cfg.horizontal_anomaly_mode)
        self.sigma_er, self.epsilon_r_er = sampled.sigma_er, sampled.epsilon_r_er
        self.sigma_et, self.epsilon_r_et = sampled.sigma_et, sampled.epsilon_r_et
        if sampled.fractions_er is not None:
            self.anomaly_horizontal_fractions_er = sampled.fractions_er
            self.anomaly_horizontal_fractions_et = sampled.fractions_et
        self.cfl_time_step_limit_s = cfl_time_step_limit(
            self.mesh, self.radii_m, self.radial_steps_m, self.epsilon_r_er, self.epsilon_r_et,
            plasma)
        self.maximum_stable_time_step_s = cfg.courant_factor * self.cfl_time_step_limit_s
        self.time_step_s = (cfg.time_step_s if cfg.time_step_s is not None
                            else self.maximum_stable_time_step_s)
        if self.time_step_s > self.maximum_stable_time_step_s * (1.0 + 1.0e-12):
            raise ValueError(f"time step {self.time_step_s:.6e} s exceeds conservative limit "
                             f"{self.maximum_stable_time_step_s:.6e} s")
        if self.source is not None and \
                self.source.carrier_frequency_hz >= 0.5 / self.time_step_s:
            raise ValueError("source carrier frequency must be below the time-step Nyquist "
                             "limit")
        self._prepare_material_coefficients()
        self._source_distribution = None
        self._tangential_source_distribution = None
        if isinstance(self.source, TangentialGaussianCurrent):
            self._tangential_source_distribution = self.source.edge_distribution(self)
        elif self.source is not None:
            self._source_distribution = self.source.staggered_distribution(self)
        scales = dict(DEFAULT_NARROW_SCALES)
        unknown = set(storage_scales or {}) - set(scales)
        if unknown:
            raise ValueError(f"unknown storage scale fields: {sorted(unknown)}")
        scales.update(storage_scales or {})
        for name, value in scales.items():
            if not _is_scale(value):
                raise ValueError(f"storage scale for {name} must be a power of two in "
                                 "[2^-60, 2^60]")
        if storage_scaling not in ("auto", "dynamic", "static"):
            raise ValueError("storage_scaling must be 'auto', 'dynamic' or 'static'")
        if storage_scaling == "auto":
            # FP16 needs its range managed; BF16 has the exponent range of FP32
            storage_scaling = "dynamic" if self.backend.dtype_name == "float16" else "static"
        self.storage_scaling = storage_scaling
        # Narrow storage rounds each stored value stochastically by default:
        # to nearest, every update smaller than half an ulp of the stored
        # value is lost, which freezes slowly varying (quasi-static) fields.
        self.storage_rounding = storage_rounding
        self.thread_mapping = thread_mapping
        self._initial_scales = scales
        self.time_s = 0.0
        self.steps = 0
        self.compiled = bool(compile_step)
        self._plan = None
        self._keepalive: list[Any] = []
        self._create_plan()
        self._warn_if_narrow_storage_stagnates()

    # ------------------------------------------------------------------ setup
    def _prepare_material_coefficients(self) -> None:
        """FP64 exponential (or trapezoidal) loss coefficients (solver.py:706-761)."""
        dt = self.time_step_s
        # Horizontally uniform materials (every benchmark profile) are computed
        # on one radial row: the same per-element arithmetic on fewer elements.
        sig_er, rel_er = _shared_rows(self.sigma_er, self.epsilon_r_er)
        sig_et, rel_et = _shared_rows(self.sigma_et, self.epsilon_r_et)
        eps_er = EPSILON_0 * rel_er
        eps_et = EPSILON_0 * rel_et
        if self.config.loss_integration == "trapezoidal":
            l_er = sig_er * dt / (2.0 * eps_er)
            l_et = sig_et * dt / (2.0 * eps_et)
            ca_er, cb_er = (1.0 - l_er) / (1.0 + l_er), dt / (eps_er * (1.0 + l_er))
            ca_et, cb_et = (1.0 - l_et) / (1.0 + l_et), dt / (eps_et * (1.0 + l_et))
        else:
            ca_er, cb_er = exponential_loss_coefficients(sig_er, eps_er, dt)
            ca_et, cb_et = exponential_loss_coefficients(sig_et, eps_et, dt)
        if self.config.compress_uniform_material_coefficients:
            ca_er = self._uniform_radial_profile(ca_er, "radial electric")
            cb_er = self._uniform_radial_profile(cb_er, "radial electric")
            ca_et = self._uniform_radial_profile(ca_et, "tangential electric")
            cb_et = self._uniform_radial_profile(cb_et, "tangential electric")
        # Host FP64 copies (the reference keeps backend arrays of the same values).
        self._ca_er, self._cb_er, self._ca_et, self._cb_et = ca_er, cb_er, ca_et, cb_et
        self._eps_er, self._eps_et = eps_er, eps_et

    @staticmethod
    def _uniform_radial_profile(values, label):
        if values.shape[0] == 1:
            return values
        profile = values[:1]
        if not np.array_equal(values, np.broadcast_to(profile, values.shape)):
            raise ValueError(f"cannot compress horizontally varying {label} coefficients")
        return profile

    @staticmethod
    def _rows(*arrays):
        """Collapse exactly horizontally-uniform coefficient arrays to one row."""
        rows = max(a.shape[0] for a in arrays)
        full = [np.broadcast_to(a, (rows, a.shape[1])) for a in arrays]
        if all(np.array_equal(a, np.broadcast_to(a[:1], a.shape)) for a in full):
            return tuple(np.ascontiguousarray(a[:1]) for a in full)
        return tuple(np.ascontiguousarray(a) for a in full)

    def _tensor(self, values, dtype):
        torch = self.backend.torch
        if torch.is_tensor(values):
            return values.to(device=self.backend.torch_device, dtype=dtype)
        array = np.asarray(values)
        if not array.flags.writeable:
            array = array.copy()
        return torch.as_tensor(array, dtype=dtype, device=self.backend.torch_device)

    def _create_plan(self) -> None:
        """Build the C-ABI descriptors, create the plan and bind the state."""
        torch = self.backend.torch
        mesh = self.mesh
        R = len(self.radial_midpoints_m)
        keep = self._keepalive = []

        def f64(a):
            arr = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
            keep.append(arr)
            return _native.f64_ptr(arr)

        def i64(a):
            arr = np.ascontiguousarray(np.asarray(a, dtype=np.int64))
            keep.append(arr)
            return _native.i64_ptr(arr)

        vedges, vsigns = vertex_incidence(mesh)
        self._vertex_edges, self._vertex_edge_signs = vedges, vsigns
        md = _native.GeoMeshDesc(
            n_vertices=mesh.n_vertices, n_edges=mesh.n_edges, n_faces=mesh.n_faces,
            edges=i64(mesh.edges), faces=i64(mesh.faces), face_edges=i64(mesh.face_edges),
            face_edge_signs=i64(mesh.face_edge_signs), edge_left_faces=i64(mesh.edge_left_faces),
            edge_right_faces=i64(mesh.edge_right_faces), vertex_edges=i64(vedges),
            vertex_edge_signs=i64(vsigns), max_degree=vedges.shape[1],
            primal_edge_angles=f64(mesh.primal_edge_angles),
            dual_edge_angles=f64(mesh.dual_edge_angles),
            face_solid_angles=f64(mesh.face_solid_angles),
            dual_cell_solid_angles=f64(mesh.dual_cell_solid_angles))
        rd = _native.GeoRadialDesc(radial_cells=R, radii_m=f64(self.radii_m),
                                   full_spherical=int(self.config.geometry_mode ==
                                                      "full-spherical"))
        ca_er, cb_er, eps_er, sig_er = self._rows(
            self._ca_er, self._cb_er, np.broadcast_to(self._eps_er, self.sigma_er.shape),
            self.sigma_er)
        ca_et, cb_et, eps_et, sig_et = self._rows(
            self._ca_et, self._cb_et, np.broadcast_to(self._eps_et, self.sigma_et.shape),
            self.sigma_et)
        storage = self.backend.dtype_name
        narrow = storage in ("float16", "bfloat16")
        ph = _native.GeoPhysicsDesc(
            storage=_native.STORAGE_CODES[storage], time_step_s=self.time_step_s,
            ca_er=f64(ca_er), cb_er=f64(cb_er), er_coef_rows=ca_er.shape[0],
            ca_et=f64(ca_et), cb_et=f64(cb_et), et_coef_rows=ca_et.shape[0],
            eps_er=f64(eps_er), eps_et=f64(eps_et),
            sigma_er=f64(sig_er), sigma_et=f64(sig_et),
            scaling=int(narrow and self.storage_scaling == "dynamic"),
            storage_rounding=STORAGE_ROUNDING[self.storage_rounding] if narrow else 0,
            thread_mapping=THREAD_MAPPING[self.thread_mapping],
            scale_er=self._initial_scales["er"] if narrow else 1.0,
            scale_et=self._initial_scales["et"] if narrow else 1.0,
            scale_hr=self._initial_scales["hr"] if narrow else 1.0,
            scale_ht=self._initial_scales["ht"] if narrow else 1.0)
        si_terms = 0
        if self.surface_impedance is not None:
            tables = surface_ade_tables(self.surface_impedance, mesh.n_edges, self.time_step_s)
            si_terms = int(self.surface_impedance.terms)
            ph.si_terms = si_terms
            ph.si_decay = f64(tables["decay"])
            ph.si_drive = f64(tables["drive"])
            ph.si_history_weights = f64(tables["history_weights"])
            scale = tables["scale"]
            if scale.size > 1 and np.all(scale == scale[0]):
                scale = scale[:1]
            ph.si_scale = f64(scale)
            ph.si_scale_rows = scale.size
        if self._source_distribution is not None:
            v, l, w = self._source_distribution
            ph.n_source = len(v)
            ph.source_vertices, ph.source_layers, ph.source_weights = i64(v), i64(l), f64(w)
            ph.source_length_m = self.source.vertical_element_length_m
        if self._tangential_source_distribution is not None:
            e, l, w = self._tangential_source_distribution
            ph.n_tsource = len(e)
            ph.tsource_edges, ph.tsource_layers, ph.tsource_weights = i64(e), i64(l), f64(w)
        self._plasma_info = None
        if self.plasma is not None:
            pt = plasma_tables(self.plasma, mesh, self.time_step_s)
            lo, hi = pt["band"]
            if hi > lo:
                ph.n_species = len(self.plasma.species)
                ph.magnetized = int(pt["magnetized"])
                ph.band_lo, ph.band_hi = lo, hi
                ph.plasma_reconstruction = f64(pt["reconstruction"])
                ph.face_tangents = f64(np.ascontiguousarray(pt["face_tangents"]))
                ph.face_centers = f64(mesh.face_centers)
                ph.left_tangents = f64(pt["left_tangents"])
                ph.right_tangents = f64(pt["right_tangents"])
                ph.vertex_faces = i64(pt["vertex_faces"])
                ph.vertex_face_weights = f64(pt["vertex_face_weights"])
                ph.vf_width = pt["vertex_faces"].shape[1]
                ph.plasma_coef = f64(pt["coef"])
                ph.plasma_coef_rows = pt["coef"].shape[2]
                ph.plasma_b = f64(pt["bdir"])
                ph.plasma_b_rows = pt["bdir"].shape[0]
            self._plasma_info = pt
        plan = C.c_void_p()
        _native.check(self._lib.geo_plan_create(C.byref(md), C.byref(rd), C.byref(ph),
                                                C.byref(plan)), self._lib)
        self._plan = plan
        layout = _native.GeoLayout()
        _native.check(self._lib.geo_plan_layout(plan, C.byref(layout)), self._lib)
        self._layout = layout
        dev = self.backend.torch_device
        sdt = self.backend.storage_dtype
        cdt = self.backend.compute_dtype
        pn, pm = layout.node_pitch, layout.mid_pitch
        self._er_buf = torch.zeros((mesh.n_vertices, pn), dtype=sdt, device=dev)
        self._ht_buf = torch.zeros((mesh.n_edges, pn), dtype=sdt, device=dev)
        self._et_buf = torch.zeros((mesh.n_edges, pm), dtype=sdt, device=dev)
        self._hr_buf = torch.zeros((mesh.n_faces, pm), dtype=sdt, device=dev)
        # edge-major (E, si_pitch) rows; the reference's (E, terms) array is a view
        self._si_buf = (torch.zeros((mesh.n_edges, layout.si_pitch), dtype=cdt, device=dev)
                        if si_terms else None)
        self._si_memory = self._si_buf[:, :si_terms] if si_terms else None
        n_species = ph.n_species
        self._plasma_buf = (torch.zeros((n_species, layout.plasma_components, mesh.n_faces,
                                         layout.plasma_pitch), dtype=cdt, device=dev)
                            if n_species else None)
        self._surface_impedance_ade = (_SurfaceImpedanceState(self, si_terms) if si_terms
                                       else None)
        self._bind()

    def _bind(self) -> None:
        st = _native.GeoState(
            er=self._er_buf.data_ptr(), er_pitch=self._er_buf.shape[1],
            ht=self._ht_buf.data_ptr(), ht_pitch=self._ht_buf.shape[1],
            et=self._et_buf.data_ptr(), et_pitch=self._et_buf.shape[1],
            hr=self._hr_buf.data_ptr(), hr_pitch=self._hr_buf.shape[1],
            si_memory=self._si_buf.data_ptr() if self._si_buf is not None else None,
            plasma_current=self._plasma_buf.data_ptr() if self._plasma_buf is not None else None,
            plasma_pitch=self._layout.plasma_pitch)
        _native.check(self._lib.geo_plan_bind(self._plan, C.byref(st)), self._lib)

    def __del__(self) -> None:
        plan = getattr(self, "_plan", None)
        lib = getattr(self, "_lib", None)
        if plan is not None and lib is not None:
            try:
                self.backend.synchronize()
            except Exception:
                pass
            lib.geo_plan_destroy(plan)
            self._plan = None

    # ------------------------------------------------------------ field views
    @property
    def storage_scales(self) -> dict[str, float]:
        """Current storage scales (stored value = field * scale); all 1 for
        FP32/FP64 storage.  With dynamic scaling they change between chunks."""
        out = (C.c_double * 4)()
        _native.check(self._lib.geo_get_storage_scales(self._plan, out, self._stream()),
                      self._lib)
        return dict(zip(_FIELD_ORDER, (float(x) for x in out)))

    def _narrow(self) -> bool:
        return self.backend.dtype_name in ("float16", "bfloat16")

    def _field_view(self, name: str):
        buf = getattr(self, f"_{name}_buf")
        cols = len(self.radii_m) if name in ("er", "ht") else len(self.radial_midpoints_m)
        view = buf[:, :cols]
        if self._narrow():
            return view.float() / self.storage_scales[name]
        return view

    def _set_field(self, name: str, values) -> None:
        buf = getattr(self, f"_{name}_buf")
        cols = len(self.radii_m) if name in ("er", "ht") else len(self.radial_midpoints_m)
        data = self._tensor(values, self.backend.compute_dtype)
        if tuple(data.shape) != (buf.shape[0], cols):
            raise ValueError(f"{name} must have shape {(buf.shape[0], cols)}")
        if self._narrow():
            scales = self.storage_scales
            peak = float(data.abs().max()) if data.numel() else 0.0
            if not np.isfinite(peak):
                raise ValueError(f"{name} values must be finite")
            if self.storage_scaling == "dynamic" and peak > 0.0:
                # place the peak at 2^10, as the device rescale does; the next
                # rescale point couples it to the other fields
                exponent = int(np.clip(np.floor(np.log2(1024.0 / peak)), -60, 60))
                scales[name] = float(2.0 ** exponent)
            limit = 65504.0 if self.backend.dtype_name == "float16" else 3.0e38
            if peak * scales[name] >= limit:
                raise ValueError(f"{name} values overflow {self.backend.dtype_name} storage at "
                                 f"scale {scales[name]:g}")
            data = data * scales[name]
            buf[:, :cols].copy_(data)
            arr = (C.c_double * 4)(*(scales[k] for k in _FIELD_ORDER))
            _native.check(self._lib.geo_set_storage_scales(self._plan, arr, self._stream()),
                          self._lib)
            return
        buf[:, :cols].copy_(data)

    er = property(lambda self: self._field_view("er"),
                  lambda self, v: self._set_field("er", v))
    et = property(lambda self: self._field_view("et"),
                  lambda self, v: self._set_field("et", v))
    hr = property(lambda self: self._field_view("hr"),
                  lambda self, v: self._set_field("hr", v))
    ht = property(lambda self: self._field_view("ht"),
                  lambda self, v: self._set_field("ht", v))

    def _plasma_bases(self):
        """Device (normal, tangents, reconstruction, w) for state <-> Cartesian J,
        with t_2 = w[:, 0] t_0 + w[:, 1] t_1 (a face's edge tangents are
        orthogonal to its centre, hence coplanar)."""
        if getattr(self, "_plasma_basis_cache", None) is None:
            pt = self._plasma_info
            cdt = self.backend.compute_dtype
            t = pt["face_tangents"]
            gram = np.einsum("fcd,fed->fce", t[:, :2], t[:, :2])
            rhs = np.einsum("fcd,fd->fc", t[:, :2], t[:, 2])
            w = np.linalg.solve(gram, rhs[..., None])[..., 0]
            self._plasma_basis_cache = (self._tensor(self.mesh.face_centers, cdt),
                                        self._tensor(t, cdt),
                                        self._tensor(pt["reconstruction"], cdt),
                                        self._tensor(w, cdt))
        return self._plasma_basis_cache

    @property
    def plasma_current_density(self) -> list:
        """Per-species (face, radial, 3) Cartesian current density (copies).

        With B = 0 the device state holds the projections (J.n, J.t_0, J.t_1)
        the scatter needs; with J.t_2 = w1 J.t_0 + w2 J.t_1,
        J = (J.n) n + sum_c (J.t_c) R_c recovers the Cartesian vector exactly
        (R = pinv(T).T spans the tangent plane)."""
        torch = self.backend.torch
        R = len(self.radial_midpoints_m)
        if self._plasma_buf is None:
            if self.plasma is None:
                return []
            return [torch.zeros((self.mesh.n_faces, R, 3), dtype=self.backend.compute_dtype,
                                device=self.backend.torch_device) for _ in self.plasma.species]
        lo, hi = self._layout.band_lo, self._layout.band_hi
        out = []
        for s in range(self._plasma_buf.shape[0]):
            state = self._plasma_buf[s, :, :, : hi - lo]  # (comps, F, band)
            if self._layout.plasma_basis == 0:
                band_j = state.permute(1, 2, 0)
            else:
                normal, _, recon, w = self._plasma_bases()
                t2 = w[:, 0, None] * state[1] + w[:, 1, None] * state[2]
                tang = torch.stack((state[1], state[2], t2))
                band_j = (state[0][..., None] * normal[:, None, :]
                          + torch.einsum("cfb,fcd->fbd", tang, recon))
            full = torch.zeros((self.mesh.n_faces, R, 3), dtype=self._plasma_buf.dtype,
                               device=self._plasma_buf.device)
            full[:, lo:hi, :] = band_j
            out.append(full)
        return out

    def set_plasma_current_density(self, index: int, values) -> None:
        data = self._tensor(values, self.backend.compute_dtype)
        R = len(self.radial_midpoints_m)
        if tuple(data.shape) != (self.mesh.n_faces, R, 3):
            raise ValueError("plasma current must have shape (face, radial, 3)")
        if self._plasma_buf is None:
            if bool((data != 0).any()):
                raise ValueError("this plasma model has no drive; its current must be zero")
            return
        lo, hi = self._layout.band_lo, self._layout.band_hi
        outside = torch_any_nonzero(data[:, :lo]) or torch_any_nonzero(data[:, hi:])
        if outside:
            raise ValueError("plasma current is nonzero outside the layers with plasma drive")
        band = data[:, lo:hi, :]
        if self._layout.plasma_basis == 0:
            self._plasma_buf[index, :, :, : hi - lo] = band.permute(2, 0, 1)
        else:
            normal, tangents, _, _ = self._plasma_bases()
            self._plasma_buf[index, 0, :, : hi - lo] = (band * normal[:, None, :]).sum(-1)
            self._plasma_buf[index, 1:, :, : hi - lo] = self.backend.torch.einsum(
                "fbd,fcd->cfb", band, tangents[:, :2])

    # --------------------------------------------------------------- stepping
    def _stream(self):
        return C.c_void_p(self.backend.current_stream_handle())

    def _source_currents(self, count: int):
        """FP64 currents at (n + 1/2) dt for the next ``count`` steps (solver.py:851-865)."""
        if self.source is None or count == 0:
            return None
        times = (self.steps + np.arange(count, dtype=np.float64) + 0.5) * self.time_step_s
        values = self.source.currents_a(times, self.time_step_s)
        return self._tensor(values, self.backend.compute_dtype)

    def _warn_if_narrow_storage_stagnates(self) -> None:
        """Round-to-nearest FP16/BF16 storage under a slowly varying source:
        once the per-step change of a field falls below half an ulp of its
        stored value it is lost every step, so quasi-static fields stop
        following the source (relative errors of order one measured on the
        20 Hz benchmark source after 5000-10000 steps)."""
        waveform = getattr(self.source, "waveform_parameters", None)
        if not self._narrow() or self.storage_rounding != "nearest" or waveform is None:
            return
        half_width, _ = waveform(self.time_step_s)
        if half_width > STAGNATION_STEPS * self.time_step_s:
            import warnings

            warnings.warn(
                f"{self.backend.dtype_name} storage rounded to nearest with a source envelope "
                f"of {half_width / self.time_step_s:.3g} steps: updates below half an ulp of "
                "the stored field are lost and quasi-static fields stagnate; use "
                "storage_rounding='rotation' (the default)", RuntimeWarning, stacklevel=3)

    def _sync_step_index(self, advancing: int) -> None:
        """Tell the plan the step index of the bound state when it differs from
        the steps the plan has advanced itself (a loaded checkpoint, a caller
        setting ``steps``): the stochastic rounding of narrow storage is keyed
        on it.  ``advancing`` steps are about to be queued."""
        if getattr(self, "_device_steps", 0) != self.steps:
            _native.check(self._lib.geo_plan_set_step_index(self._plan, int(self.steps),
                                                            self._stream()), self._lib)
        self._device_steps = self.steps + advancing

    def step(self, count: int = 1) -> None:
        """Advance the fields by ``count`` complete leapfrog time steps."""
        count = self._validated_count(count, "step count", minimum=0)
        if count == 0:
            return
        self._sync_step_index(count)
        currents = self._source_currents(count)
        chunk = self.compile_chunk_size if self.compiled else 1
        _native.check(self._lib.geo_step(
            self._plan, None if currents is None else currents.data_ptr(), count, chunk, None,
            self._stream()), self._lib)
        self._pending = currents  # keep alive until the queued work has read it
        self.steps += count
        self.time_s = self.steps * self.time_step_s

    def prepare_compiled_step(self) -> None:
        """Capture the HIP graph of ``compile_chunk_size`` steps ahead of the
        first ``step`` that replays it (no-op when ``compile_step=False`` or
        the chunk is a single step).  Setting new receivers or rebinding the
        state (checkpoint load) rebuilds it on the next use."""
        if self.compiled and self.compile_chunk_size > 1:
            _native.check(self._lib.geo_step_prepare(self._plan, self.compile_chunk_size),
                          self._lib)

    def _advance_fields(self, current_a: float = 0.0) -> None:
        """One leapfrog step with an explicit source current (reference signature)."""
        value = float(current_a)
        currents = self._tensor(np.asarray([value]), self.backend.compute_dtype)
        self._sync_step_index(0)
        self._device_steps += 1  # the caller advances ``steps`` itself
        _native.check(self._lib.geo_step(self._plan, currents.data_ptr(), 1, 1, None,
                                         self._stream()), self._lib)
        self._pending = currents

    def record_er_observations(self, vertex_indices, radial_layers, weights, steps: int, *,
                               synchronize_every: int = 128) -> NDArray[np.generic]:
        """Advance and record weighted Er samples (solver.py:1080-1136).

        Row 0 is the initial field and row n the field after step n.  The
        samples are written by the step kernels into a device trace buffer;
        one device-to-host copy happens at the end.
        """
        steps = self._validated_count(steps, "step count", minimum=0)
        self._validated_count(synchronize_every, "synchronize_every", minimum=1)
        n_rx = self._set_receivers(vertex_indices, radial_layers, weights)
        torch = self.backend.torch
        traces = torch.zeros((steps + 1, n_rx), dtype=self.backend.compute_dtype,
                             device=self.backend.torch_device)
        if n_rx == 0:
            self.step(steps)
            return self.to_numpy(traces)
        self._sync_step_index(steps)
        currents = self._source_currents(steps)
        chunk = self.compile_chunk_size if self.compiled else 1
        _native.check(self._lib.geo_step(
            self._plan, None if currents is None else currents.data_ptr(), steps, chunk,
            traces.data_ptr(), self._stream()), self._lib)
        self.steps += steps
        self.time_s = self.steps * self.time_step_s
        self.backend.synchronize()
        return self.to_numpy(traces)

    def _set_receivers(self, vertex_indices, radial_layers, weights) -> int:
        """Validate receivers (solver.py:1097-1120) and upload them to the plan."""
        vertices = self._validated_index_array(vertex_indices, "vertex_indices")
        layers = self._validated_index_array(radial_layers, "radial_layers")
        w = np.asarray(weights, dtype=np.float64)
        if vertices.ndim != 2 or w.shape != vertices.shape:
            raise ValueError("vertex_indices and weights must have matching 2-D shapes")
        if layers.shape != (vertices.shape[0],):
            raise ValueError("radial_layers must contain one layer per observation")
        if np.any(vertices < 0) or np.any(vertices >= self.mesh.n_vertices):
            raise ValueError("observation vertex index is out of range")
        if np.any(layers < 0) or np.any(layers >= len(self.radii_m)):
            raise ValueError("observation radial layer is out of range")
        if not np.all(np.isfinite(w)):
            raise ValueError("observation weights must be finite")
        if not np.allclose(w.sum(axis=1), 1.0):
            raise ValueError("observation weights must sum to one")
        n_rx = vertices.shape[0]
        width = vertices.shape[1]
        if width > 3:
            raise ValueError("each observation may combine at most three vertices")
        if n_rx == 0:
            return 0
        v3 = np.zeros((n_rx, 3), dtype=np.int64)
        w3 = np.zeros((n_rx, 3), dtype=np.float64)
        v3[:, :width] = vertices
        v3[:, width:] = vertices[:, :1]
        w3[:, :width] = w
        v3c, l3c, w3c = (np.ascontiguousarray(v3), np.ascontiguousarray(layers),
                         np.ascontiguousarray(w3))
        key = (v3c.tobytes(), l3c.tobytes(), w3c.tobytes())
        if key == getattr(self, "_receiver_key", None):
            return n_rx  # same receivers: the plan (and its step graph) stay as they are
        _native.check(self._lib.geo_set_receivers(self._plan, _native.i64_ptr(v3c),
                                                  _native.i64_ptr(l3c), _native.f64_ptr(w3c),
                                                  n_rx), self._lib)
        self._receiver_key = key
        return n_rx

    def record_h_observations(self, face_indices, face_radial_layers, face_weights,
                              edge_indices, edge_radial_layers, edge_weights, steps: int, *,
                              synchronize_every: int = 128, sample_every: int = 1):
        """Weighted Hr/Ht samples at half-step times (solver.py:1138-1229)."""
        faces = self._validated_index_array(face_indices, "face_indices")
        flayers = self._validated_index_array(face_radial_layers, "face_radial_layers")
        fw = np.asarray(face_weights, dtype=np.float64)
        edges = self._validated_index_array(edge_indices, "edge_indices")
        elayers = self._validated_index_array(edge_radial_layers, "edge_radial_layers")
        ew = np.asarray(edge_weights, dtype=np.float64)
        steps = self._validated_count(steps, "step count", minimum=0)
        self._validated_count(synchronize_every, "synchronize_every", minimum=1)
        sample_every = self._validated_count(sample_every, "sample_every", minimum=1)
        if faces.ndim != 2 or fw.shape != faces.shape:
            raise ValueError("face indices and weights must have matching 2-D shapes")
        if edges.ndim != 2 or ew.shape != edges.shape:
            raise ValueError("edge indices and weights must have matching 2-D shapes")
        if flayers.shape != faces.shape:
            raise ValueError("face radial layers must match the face indices")
        if elayers.shape != edges.shape:
            raise ValueError("edge radial layers must match the edge indices")
        if np.any(faces < 0) or np.any(faces >= self.mesh.n_faces):
            raise ValueError("observation face index is out of range")
        if np.any(edges < 0) or np.any(edges >= self.mesh.n_edges):
            raise ValueError("observation edge index is out of range")
        if np.any(flayers < 0) or np.any(flayers >= len(self.radial_midpoints_m)):
            raise ValueError("radial H observation layer is out of range")
        if np.any(elayers < 0) or np.any(elayers >= len(self.radii_m)):
            raise ValueError("tangential H observation layer is out of range")
        if not np.all(np.isfinite(fw)) or not np.all(np.isfinite(ew)):
            raise ValueError("observation weights must be finite")
        n_f, n_e = faces.shape[0], edges.shape[0]
        key = tuple(np.ascontiguousarray(a).tobytes() for a in (faces, flayers, fw, edges,
                                                                 elayers, ew)) + (
            faces.shape, edges.shape)
        if key != getattr(self, "_h_receiver_key", None):
            tables = [np.ascontiguousarray(a, dtype=t) for a, t in (
                (faces, np.int64), (flayers, np.int64), (fw, np.float64),
                (edges, np.int64), (elayers, np.int64), (ew, np.float64))]
            _native.check(self._lib.geo_set_h_receivers(
                self._plan, _native.i64_ptr(tables[0]), _native.i64_ptr(tables[1]),
                _native.f64_ptr(tables[2]), n_f, faces.shape[1],
                _native.i64_ptr(tables[3]), _native.i64_ptr(tables[4]),
                _native.f64_ptr(tables[5]), n_e, edges.shape[1]), self._lib)
            self._h_receiver_key = key
        torch = self.backend.torch
        rows = steps // sample_every + 1 + (1 if steps % sample_every else 0)
        rt = torch.zeros((rows, n_f), dtype=self.backend.compute_dtype,
                         device=self.backend.torch_device)
        tt = torch.zeros((rows, n_e), dtype=self.backend.compute_dtype,
                         device=self.backend.torch_device)
        # One device call: row 0 before the first step, then one row from the
        # E launch of every sampled step (graph chunks included).
        self._sync_step_index(steps)
        currents = self._source_currents(steps)
        chunk = self.compile_chunk_size if self.compiled else 1
        _native.check(self._lib.geo_step_h(
            self._plan, None if currents is None else currents.data_ptr(), steps, chunk,
            sample_every, rt.data_ptr() if n_f else None, tt.data_ptr() if n_e else None,
            self._stream()), self._lib)
        self._pending = currents
        self.steps += steps
        self.time_s = self.steps * self.time_step_s
        self.backend.synchronize()
        return self.to_numpy(rt), self.to_numpy(tt)

    # ---------------------------------------------------------- diagnostics
    def field_diagnostics(self) -> dict[str, float]:
        """FP64 device reductions: max |field|, staggered energies, non-finite count."""
        out = _native.GeoDiagOut()
        _native.check(self._lib.geo_diagnostics(self._plan, C.byref(out), self._stream()),
                      self._lib)
        return {name: getattr(out, name) for name, _ in out._fields_}

    def radial_profiles(self) -> dict[str, np.ndarray]:
        """Per-layer FP64 energy (J) and conductive loss (W) of each field:
        the horizontal sums of verification/physics_diagnostics/model.py:242-267
        times the radial weights (r^2 L_k at nodes, rbar^2 dr at midpoints)."""
        R = len(self.radial_midpoints_m)
        out = {"energy_er": np.zeros(R + 1), "energy_et": np.zeros(R),
               "energy_hr": np.zeros(R), "energy_ht": np.zeros(R + 1),
               "loss_er": np.zeros(R + 1), "loss_et": np.zeros(R)}
        desc = _native.GeoProfiles(**{k: _native.f64_ptr(v) for k, v in out.items()})
        _native.check(self._lib.geo_radial_profiles(self._plan, C.byref(desc), self._stream()),
                      self._lib)
        return out

    def energy_j(self) -> float:
        """Total staggered field energy (physics_diagnostics/model.py:242-258 weights)."""
        d = self.field_diagnostics()
        return d["energy_er"] + d["energy_et"] + d["energy_hr"] + d["energy_ht"]

    def diagnostics(self) -> dict[str, float | int | str]:
        """Scalar diagnostics without saving field data (solver.py:991-1031)."""
        d = self.field_diagnostics()
        return {
            "step": self.steps,
            "time_s": self.time_s,
            "electric_time_s": self.electric_time_s,
            "magnetic_time_s": self.magnetic_time_s,
            "backend": self.backend.name,
            "device": self.backend.device,
            "dtype": self.backend.dtype_name,
            "compiled": self.compiled,
            "compile_chunk_size": self.compile_chunk_size,
            "radial_boundary_condition": self.config.radial_boundary_condition,
            "surface_impedance_terms": (self.surface_impedance.terms
                                        if self.surface_impedance is not None else 0),
            "surface_impedance_state_bytes": (self.backend.nbytes(self._si_memory)
                                              if self._si_memory is not None else 0),
            "plasma_species": len(self.plasma.species) if self.plasma else 0,
            "plasma_state_bytes": (self.backend.nbytes(self._plasma_buf)
                                   if self._plasma_buf is not None else 0),
            "loss_integration": self.config.loss_integration,
            "geometry_mode": self.config.geometry_mode,
            "cfl_time_step_limit_s": self.cfl_time_step_limit_s,
            "courant_factor": self.config.courant_factor,
            "field_memory_bytes": self.memory_bytes,
            "persistent_backend_bytes": self.persistent_backend_bytes,
            "max_abs_er_v_m": d["max_abs_er"],
            "max_abs_et_v_m": d["max_abs_et"],
            "max_abs_hr_a_m": d["max_abs_hr"],
            "max_abs_ht_a_m": d["max_abs_ht"],
            "energy_j": d["energy_er"] + d["energy_et"] + d["energy_hr"] + d["energy_ht"],
            "field_finite": d["nonfinite"] == 0 and d["storage_overflow"] == 0,
            "storage_scaling": self.storage_scaling if self._narrow() else "none",
        }

    @property
    def electric_time_s(self) -> float:
        return self.steps * self.time_step_s

    @property
    def magnetic_time_s(self) -> float:
        return (self.steps - 0.5) * self.time_step_s

    def to_numpy(self, values: Any) -> NDArray[np.generic]:
        return self.backend.to_numpy(values)

    def field_value(self, field: str, *indices: int) -> float:
        if field not in {"er", "et", "hr", "ht"}:
            raise ValueError("field must be er, et, hr, or ht")
        return self.backend.scalar(getattr(self, field)[indices])

    @property
    def memory_bytes(self) -> int:
        """Bytes of the four padded field buffers."""
        return sum(self.backend.nbytes(b) for b in (self._er_buf, self._et_buf, self._hr_buf,
                                                    self._ht_buf))

    @property
    def persistent_backend_bytes(self) -> int:
        total = self.memory_bytes
        for b in (self._si_memory, self._plasma_buf):
            if b is not None:
                total += self.backend.nbytes(b)
        return total

    def kernel_times_ms(self, steps: int) -> tuple[float, float]:
        """Average (H, E) kernel durations of ``steps`` event-timed single steps.

        The timed steps advance the state with a zero source current."""
        out = (C.c_double * 2)()
        self._sync_step_index(steps)
        _native.check(self._lib.geo_time_kernels(self._plan, steps, self._stream(), out),
                      self._lib)
        self.steps += steps
        self.time_s = self.steps * self.time_step_s
        return float(out[0]), float(out[1])

    def save_checkpoint(self, path: str | Path) -> Path:
        from .checkpoint import save_checkpoint

        return save_checkpoint(self, path)

    @classmethod
    def load_checkpoint(cls, path: str | Path, *, backend: str = "hip", device: str = "auto",
                        dtype: str | None = None, compile_step: bool = False,
                        compile_chunk_size: int = 8, torch_threads: int | None = None):
        from .checkpoint import load_checkpoint

        return load_checkpoint(path, backend=backend, device=device, dtype=dtype,
                               compile_step=compile_step, compile_chunk_size=compile_chunk_size,
                               torch_threads=torch_threads)

    @staticmethod
    def _validated_count(value: int, label: str, *, minimum: int) -> int:
        if isinstance(value, bool) or not isinstance(value, (int, np.integer)):
            raise ValueError(f"{label} must be an integer")
        result = int(value)
        if result < minimum:
            raise ValueError(f"{label} must be {'non-negative' if minimum == 0 else 'positive'}")
        return result

    @staticmethod
    def _validated_index_array(values: Any, label: str) -> NDArray[np.int64]:
        array = np.asarray(values)
        if np.issubdtype(array.dtype, np.bool_) or not np.issubdtype(array.dtype, np.integer):
            raise ValueError(f"{label} must contain integers")
        return np.asarray(array, dtype=np.int64)


def _validated_material(values, shape, label):
    try:
        sigma, eps = values
    except (TypeError, ValueError) as error:
        raise ValueError(f"{label} material sampler must return conductivity and "
                         "permittivity") from error
    sigma = np.asarray(sigma, dtype=np.float64)
    eps = np.asarray(eps, dtype=np.float64)
    if sigma.shape != shape or eps.shape != shape:
        raise ValueError(f"{label} material arrays must have shape {shape}")
    if not np.all(np.isfinite(sigma)) or not np.all(np.isfinite(eps)):
        raise ValueError(f"{label} material arrays must be finite")
    if np.any(sigma < 0.0):
        raise ValueError(f"{label} conductivity cannot be negative")
    if np.any(eps <= 0.0):
        raise ValueError(f"{label} relative permittivity must be positive")
    return sigma, eps


@dataclass
class MaterialSample:
    """Material on the staggered grid: Er on vertices x radial nodes, Et on
    edges x radial midpoints, and the horizontal anomaly fractions of the
    conservative-nearest mode (None otherwise)."""

    sigma_er: NDArray[np.float64]
    epsilon_r_er: NDArray[np.float64]
    sigma_et: NDArray[np.float64]
    epsilon_r_et: NDArray[np.float64]
    fractions_er: tuple | None = None
    fractions_et: tuple | None = None


# wedges sampled per call of a material's sampler in the dual-cell average
_WEDGES_PER_SAMPLE = 65_536


def _dual_cell_average(material, mesh, altitudes_m, earth_radius_m):
    """Er material averaged over each polygonal dual cell (solver.py:631-705):
    one quadrature point per (vertex, incident edge) wedge, weighted by the
    wedge's share of the cell area, summed per vertex in edge order."""
    ends = mesh.edges.ravel()
    order = np.argsort(ends, kind="stable")
    wv, we = ends[order], order // 2
    directions, area = mesh.dual_cell_wedge_quadrature(wv, we)
    weight = area / mesh.dual_cell_solid_angles[wv]
    if not np.allclose(np.bincount(wv, weights=weight, minlength=mesh.n_vertices), 1.0,
                       rtol=0.0, atol=2.0e-12):
        raise RuntimeError("dual-cell material weights do not close")
    shape = (mesh.n_vertices, len(altitudes_m))
    sigma, eps = np.zeros(shape), np.zeros(shape)
    for lo in range(0, len(wv), _WEDGES_PER_SAMPLE):
        part = slice(lo, lo + _WEDGES_PER_SAMPLE)
        s, e = _validated_material(material.sample(directions[part], altitudes_m, earth_radius_m),
                                   (len(directions[part]), len(altitudes_m)),
                                   "radial dual-cell support")
        np.add.at(sigma, wv[part], weight[part, None] * s)
        np.add.at(eps, wv[part], weight[part, None] * e)
    return sigma, eps


def _sample_tangential(material, directions, altitudes_m, midpoint_altitudes_m, earth_radius_m):
    """Et material at unit directions: a material's radial-cell sampler when
    it has one (sample_tangential_cells over [alt_j, alt_j+1]), else point
    samples at the radial midpoints."""
    cells = getattr(material, "sample_tangential_cells", None)
    if cells is None:
        values = material.sample(directions, midpoint_altitudes_m, earth_radius_m)
    else:
        values = cells(directions, altitudes_m[:-1], altitudes_m[1:], earth_radius_m)
    return _validated_material(values, (len(directions), len(midpoint_altitudes_m)),
                               "tangential")


def _edge_diamond_average(material, mesh, altitudes_m, midpoint_altitudes_m, earth_radius_m):
    """Et material averaged over each edge diamond (solver.py:579-606): one
    quadrature point per quarter-diamond triangle, weighted by its area."""
    mid = mesh.edge_midpoints()
    tail, head = mesh.vertices[mesh.edges[:, 0]], mesh.vertices[mesh.edges[:, 1]]
    left = mesh.face_centers[mesh.edge_left_faces]
    right = mesh.face_centers[mesh.edge_right_faces]
    quadrants = mesh.edge_diamond_quadrant_solid_angles()
    share = quadrants / np.sum(quadrants, axis=1, keepdims=True)
    shape = (mesh.n_edges, len(midpoint_altitudes_m))
    sigma, eps = np.zeros(shape), np.zeros(shape)
    for k, (a, b) in enumerate(((tail, left), (left, head), (head, right), (right, tail))):
        d = mid + a + b
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        s, e = _sample_tangential(material, d, altitudes_m, midpoint_altitudes_m, earth_radius_m)
        sigma += share[:, k, None] * s
        eps += share[:, k, None] * e
    return sigma, eps


def sample_material(material, mesh, altitudes_m, midpoint_altitudes_m, earth_radius_m, *,
                    radial_support="point", tangential_support="point",
                    anomaly_mode="point") -> MaterialSample:
    """Material on the staggered grid (solver.py:467-629).

    * Er: point samples at vertices x radial nodes, or (radial_support
      "dual-cell") the dual-cell wedge average;
    * Et: samples at edge midpoints x radial midpoints, or
      (tangential_support "edge-diamond") the quarter-diamond average;
    * anomaly_mode "conservative-nearest": the material's spherical anomalies
      are taken out of point sampling and rasterised onto the dual cells and
      edge diamonds with their exact areas instead.
    A material with a ``sample_mesh`` hook provides all four arrays itself.
    """
    altitudes_m = np.asarray(altitudes_m, dtype=np.float64)
    midpoint_altitudes_m = np.asarray(midpoint_altitudes_m, dtype=np.float64)
    shape_er = (mesh.n_vertices, len(altitudes_m))
    shape_et = (mesh.n_edges, len(midpoint_altitudes_m))
    sampler = getattr(material, "sample_mesh", None)
    if sampler is not None:
        values = sampler(mesh, altitudes_m, earth_radius_m, radial_material_support=radial_support,
                         tangential_material_support=tangential_support,
                         horizontal_anomaly_mode=anomaly_mode)
        try:
            s_er, e_er, s_et, e_et = values
        except (TypeError, ValueError) as error:
            raise ValueError("mesh material sampler must return four property arrays") \
                from error
        return MaterialSample(*_validated_material((s_er, e_er), shape_er, "mesh-native radial"),
                              *_validated_material((s_et, e_et), shape_et,
                                                   "mesh-native tangential"))
    anomalies = getattr(material, "anomalies", None)
    if anomaly_mode == "conservative-nearest" and anomalies is None:
        raise ValueError("conservative-nearest anomalies require a material with an "
                         "anomalies collection")
    conservative = anomaly_mode == "conservative-nearest" and bool(anomalies)
    base = material
    if conservative:
        try:
            base = dataclasses.replace(material, anomalies=())
        except TypeError as error:
            raise ValueError("conservative-nearest anomalies require a dataclass material") \
                from error
    if radial_support == "point":
        s_er, e_er = _validated_material(base.sample(mesh.vertices, altitudes_m, earth_radius_m),
                                         shape_er, "radial")
    else:
        s_er, e_er = _dual_cell_average(base, mesh, altitudes_m, earth_radius_m)
    if tangential_support == "point":
        s_et, e_et = _sample_tangential(base, mesh.edge_midpoints(), altitudes_m,
                                        midpoint_altitudes_m, earth_radius_m)
    else:
        s_et, e_et = _edge_diamond_average(base, mesh, altitudes_m, midpoint_altitudes_m,
                                           earth_radius_m)
    out = MaterialSample(s_er, e_er, s_et, e_et)
    if conservative:
        anomalies = tuple(anomalies)
        out.fractions_er = tuple(conservative_anomaly_fractions(
            mesh.vertices, mesh.dual_cell_solid_angles, a, earth_radius_m) for a in anomalies)
        blend_node_anomalies(s_er, e_er, altitudes_m, anomalies, out.fractions_er)
        out.fractions_et = tuple(conservative_anomaly_fractions(
            mesh.edge_midpoints(), mesh.edge_diamond_solid_angles(), a, earth_radius_m)
            for a in anomalies)
        blend_cell_anomalies(s_et, e_et, altitudes_m[:-1], altitudes_m[1:], anomalies,
                             out.fractions_et)
    return out


def cfl_time_step_limit(mesh, radii_m, radial_steps_m, epsilon_r_er, epsilon_r_et,
                        plasma=None) -> float:
    """Conservative lossless CFL limit, plasma-limited (solver.py:377-406)."""
    smallest = float(np.min(radii_m))
    primal = smallest * float(mesh.primal_edge_angles.min())
    dual = smallest * float(mesh.dual_edge_angles.min())
    radial = float(np.min(radial_steps_m))
    inv_len2 = primal**-2 + dual**-2 + (2.0 / radial) ** 2
    min_eps = min(float(np.min(epsilon_r_er)), float(np.min(epsilon_r_et)))
    geometric = np.sqrt(min_eps) / (C_0 * np.sqrt(inv_len2))
    if plasma is None:
        return geometric
    wp = np.sqrt(plasma.maximum_plasma_drive() / (EPSILON_0 * min_eps))
    if wp == 0.0:
        return geometric
    return min(geometric, 2.0 / wp)


def stable_time_step(config: SimulationConfig, mesh, material=None, plasma=None) -> float:
    """The time step GeodesicFDTD would choose, without creating any device state."""
    material = material or EarthIonosphereMaterial()
    if config.radial_altitudes_m is None:
        altitudes = np.linspace(config.minimum_altitude_m, config.maximum_altitude_m,
                                config.radial_cells + 1)
    else:
        altitudes = np.asarray(config.radial_altitudes_m, dtype=np.float64)
    radii = config.earth_radius_m + altitudes
    mids = 0.5 * (radii[:-1] + radii[1:])
    m = sample_material(material, mesh, altitudes, mids - config.earth_radius_m,
                        config.earth_radius_m, radial_support=config.radial_material_support,
                        tangential_support=config.tangential_material_support,
                        anomaly_mode=config.horizontal_anomaly_mode)
    limit = cfl_time_step_limit(mesh, radii, np.diff(radii), m.epsilon_r_er, m.epsilon_r_et,
                                plasma)
    return config.time_step_s if config.time_step_s is not None else config.courant_factor * limit


def _shared_rows(*arrays):
    """The arrays' first rows if every array repeats its first row, else the arrays."""
    if all(a.shape[0] > 0 and np.array_equal(a, np.broadcast_to(a[:1], a.shape)) for a in arrays):
        return tuple(a[:1] for a in arrays)
    return arrays


def torch_any_nonzero(t) -> bool:
    return bool(t.numel()) and bool((t != 0).any())


def exponential_loss_coefficients(sigma, epsilon, dt):
    """Exact conductive decay with midpoint forcing: Ca = exp(-q),
    Cb = dt/eps * (-expm1(-q)/q) (= dt/eps at q = 0), q = sigma dt / eps."""
    rate = sigma * dt / epsilon
    decay = np.exp(-rate)
    phi1 = np.ones_like(rate)
    nz = rate != 0.0
    phi1[nz] = -np.expm1(-rate[nz]) / rate[nz]
    return decay, dt / epsilon * phi1
