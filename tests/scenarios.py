"""Problem definitions shared by the tests, smoke() and bench.py.

A scenario is a plain dict; ``build_sim`` turns it into the product
``GeodesicFDTD`` (HIP) and ``build_oracle`` into the CPU oracle's
``OracleSim`` with the same mesh, grid, material, boundary, plasma and source.

Workloads (sizes as in the benchmark definitions):
  * ``c0``   -- s2 x 24 radial cells, 0-100 km, r_E = 6370 km, ground SI
                sigma = 1e-2 S/m, sigma(z) = 1e-6 exp((z-70 km)/5 km) S/m,
                20 Hz Gaussian vertical current at 35.15 N 126.85 E, 4 receivers
                1000/5000/10000/20000 km due east, FP64.
  * ``c1``   -- s7 x 64 cells, FP32, the same profile and ground plus an
                electron Drude plasma above 60 km:
                n_e = 1.43e13 exp(-0.15 h') exp((beta-0.15)(z-h')) m^-3,
                h' = 74 km, beta = 0.3 /km; nu = 1.8e11 exp(-0.15 z[km]) /s;
                32 receivers drawn area-uniformly with seed 20260901.
  * ``c3``   -- s8 x 96 cells with a 2:1-refined band over 60-90 km
                (44 cells of 1.364 km below, 44 of 0.682 km in the band, 8 of
                1.25 km above), SI ground, the c1 plasma and receivers; the
                precision sweep runs it with FP64/FP32/FP16/BF16 storage.
                ``c3_small`` is the same grid at s3 for oracle comparisons.
"""

from __future__ import annotations

import os
import sys
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parents[1]
if str(REPO) not in sys.path:
    sys.path.insert(0, str(REPO))
ORACLE_DIR = REPO / "oracle"

EPSILON_0 = 8.854_187_812_8e-12
RECEIVER_SEED = 20260901
SOURCE_LAT, SOURCE_LON = 35.15, 126.85


def oracle_module():
    if str(ORACLE_DIR) not in sys.path:
        sys.path.insert(0, str(ORACLE_DIR))
    import geofdtd_oracle

    return geofdtd_oracle


def drude_profile(midpoint_altitudes_m, h_prime_km=74.0, beta_per_km=0.3):
    """(n_e, nu) on radial midpoints; n_e = 0 below 60 km."""
    z = np.asarray(midpoint_altitudes_m, dtype=np.float64) / 1000.0
    density = 1.43e13 * np.exp(-0.15 * h_prime_km) * np.exp((beta_per_km - 0.15) * (z - h_prime_km))
    density = np.where(z >= 60.0, density, 0.0)
    collision = 1.8e11 * np.exp(-0.15 * z)
    return density, collision


def refined_band_altitudes() -> tuple[float, ...]:
    """96 radial cells, 0-100 km, 2:1 refinement over 60-90 km (balanced 2:1)."""
    nodes = np.concatenate((np.linspace(0.0, 60e3, 45), np.linspace(60e3, 90e3, 45)[1:],
                            np.linspace(90e3, 100e3, 9)[1:]))
    return tuple(float(x) for x in nodes)


def scenario(name: str, **overrides) -> dict:
    base = {
        "subdivision": 2, "radial_cells": 24, "minimum_altitude_m": 0.0,
        "maximum_altitude_m": 100e3, "earth_radius_m": 6.370e6, "courant_factor": 0.35,
        "geometry_mode": "full-spherical", "loss_integration": "exponential",
        "material": {"ionosphere_prefactor_hz": 1e-6 / EPSILON_0,
                     "ionosphere_reference_height_m": 70e3,
                     "ionosphere_scale_height_m": 5e3},
        "vacuum": False, "surface_sigma": 1e-2, "plasma": None,
        "source": {"latitude_deg": SOURCE_LAT, "longitude_deg": SOURCE_LON, "altitude_m": 0.0,
                   "carrier_frequency_hz": 20.0},
        "tangential": None, "dtype": "float64", "receivers": "c0", "time_step_s": None,
        "seed_fields": False, "radial_altitudes_m": None,
    }
    if name == "c0":
        pass
    elif name == "c0_pulse":
        # source active inside a short run: pulse half width 30 steps, centre 90 steps
        base["source"] = dict(base["source"], one_over_e_half_width_s="30dt",
                              center_time_s="90dt")
    elif name == "c0_plasma":
        base["plasma"] = {"h_prime_km": 74.0, "beta_per_km": 0.3}
        base["source"] = dict(base["source"], one_over_e_half_width_s="30dt",
                              center_time_s="90dt")
    elif name == "plasma_random":
        # seeded fields everywhere so the plasma band is driven from the first step
        base["plasma"] = {"h_prime_km": 74.0, "beta_per_km": 0.3}
        base["seed_fields"] = True
    elif name == "drude_two_species":
        # B = 0 with electrons and O+ ions: the general species loop of the
        # Drude update and scatter, driven by seeded fields from the first step
        base["plasma"] = {"h_prime_km": 74.0, "beta_per_km": 0.3, "ions": True}
        base["seed_fields"] = True
    elif name == "magnetized_random":
        base.update(subdivision=1, radial_cells=12, seed_fields=True)
        base["plasma"] = {"h_prime_km": 74.0, "beta_per_km": 0.3, "magnetized": True}
    elif name == "magnetized":
        base.update(subdivision=1, radial_cells=12)
        base["plasma"] = {"h_prime_km": 74.0, "beta_per_km": 0.3, "magnetized": True}
        base["source"] = dict(base["source"], one_over_e_half_width_s="20dt",
                              center_time_s="60dt")
    elif name == "pec_vacuum":
        base.update(vacuum=True, surface_sigma=None, source=None, minimum_altitude_m=0.0,
                    subdivision=2, radial_cells=16, seed_fields=True)
    elif name == "thin_trapezoid":
        base.update(geometry_mode="thin-shell", loss_integration="trapezoidal",
                    surface_sigma=None, subdivision=1, radial_cells=10)
        base["source"] = dict(base["source"], altitude_m=2_500.0,
                              one_over_e_half_width_s="20dt", center_time_s="60dt")
    elif name == "tangential":
        # midpoints must straddle the 0 m line source (reference sources.py:138-139)
        base.update(subdivision=2, radial_cells=11, surface_sigma=None,
                    minimum_altitude_m=-10e3)
        base["source"] = None
        base["tangential"] = {"latitude_deg": 46.5, "longitude_deg": -90.9, "altitude_m": 0.0,
                              "azimuths_deg": (0.0, 90.0), "line_lengths_m": (22_500.0, 22_500.0),
                              "one_over_e_half_width_s": "20dt", "center_time_s": "60dt"}
    elif name == "c1":
        base.update(subdivision=7, radial_cells=64, dtype="float32", receivers="c1")
        base["plasma"] = {"h_prime_km": 74.0, "beta_per_km": 0.3}
    elif name == "c1_noplasma":
        base.update(subdivision=7, radial_cells=64, dtype="float32", receivers="c1")
    elif name == "c1_pec":
        base.update(subdivision=7, radial_cells=64, dtype="float32", receivers="c1",
                    surface_sigma=None)
        base["plasma"] = {"h_prime_km": 74.0, "beta_per_km": 0.3}
    elif name in ("c3", "c3_small"):
        alt = refined_band_altitudes()
        base.update(subdivision=8 if name == "c3" else 3, radial_cells=len(alt) - 1,
                    dtype="float32", receivers="c1", radial_altitudes_m=alt)
        base["plasma"] = {"h_prime_km": 74.0, "beta_per_km": 0.3}
        if name == "c3_small":
            # pulsed so that a short oracle run sees the source
            base["source"] = dict(base["source"], one_over_e_half_width_s="30dt",
                                  center_time_s="90dt")
    elif name == "c3_seeded":
        # The C3 radial grid (2:1 band over 60-90 km, 96 cells) at s2 with SI
        # ground and the Drude plasma.  Seeded fields make every layer active
        # from the first step, so the graded stencils, the band junctions and
        # the plasma band all carry weight in a relative-L2 comparison; a 1 A
        # pulsed source at 59.5 km (between the 58.6 km node and the 60 km
        # junction node, whose control lengths differ) keeps its fields at the
        # seeded fields' 1e-6 scale.
        alt = refined_band_altitudes()
        base.update(subdivision=2, radial_cells=len(alt) - 1, radial_altitudes_m=alt,
                    seed_fields=True, receivers="c1")
        base["plasma"] = {"h_prime_km": 74.0, "beta_per_km": 0.3}
        base["source"] = dict(base["source"], altitude_m=59_500.0, peak_current_a=1.0,
                              one_over_e_half_width_s="20dt", center_time_s="60dt")
    elif name in ("vacuum_s7_r64", "vacuum_s7_r80"):
        # the reference's benchmark workload: vacuum, PEC, seeded fields, no source
        base.update(subdivision=7, radial_cells=int(name[-2:]), dtype="float32",
                    receivers="c1", vacuum=True, surface_sigma=None, source=None,
                    seed_fields=True, earth_radius_m=6.371e6)
    else:
        raise KeyError(name)
    base.update(overrides)
    base["name"] = name
    return base


C2_SEED = 20260902


def c2_runs(n_runs: int = 64, seed: int = C2_SEED, subdivision: int = 5,
            radial_choices=(32, 40, 48), n_receivers: int = 16) -> list[dict]:
    """Seeded scenarios of the batched ensemble: per run R in radial_choices,
    h' ~ U(68, 80) km, beta ~ U(0.25, 0.55) /km (electron Drude profile as c1),
    ground sigma ~ logU(1e-4, 1e-1) S/m (SI boundary), source lat/lon uniform
    on the sphere, carrier ~ U(10, 40) Hz, FP32, receivers uniform on the
    sphere; the c0 conductivity profile and r_E = 6370 km throughout."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_runs):
        R = int(rng.choice(radial_choices))
        h_prime = float(rng.uniform(68.0, 80.0))
        beta = float(rng.uniform(0.25, 0.55))
        sigma_g = float(10.0 ** rng.uniform(-4.0, -1.0))
        lat = float(np.rad2deg(np.arcsin(rng.uniform(-1.0, 1.0))))
        lon = float(rng.uniform(-180.0, 180.0))
        carrier = float(rng.uniform(10.0, 40.0))
        rx_lat = np.rad2deg(np.arcsin(rng.uniform(-1.0, 1.0, n_receivers)))
        rx_lon = rng.uniform(-180.0, 180.0, n_receivers)
        sc = scenario("c0", subdivision=subdivision, radial_cells=R, dtype="float32",
                      surface_sigma=sigma_g, receivers=str(n_receivers),
                      plasma={"h_prime_km": h_prime, "beta_per_km": beta},
                      source={"latitude_deg": lat, "longitude_deg": lon, "altitude_m": 0.0,
                              "carrier_frequency_hz": carrier})
        sc["name"] = f"c2_run{i}"
        sc["receiver_points"] = list(zip(rx_lat.tolist(), rx_lon.tolist()))
        out.append(sc)
    return out


def _resolve_time(value, dt):
    if isinstance(value, str) and value.endswith("dt"):
        return float(value[:-2]) * dt
    return value


def receiver_points(sc: dict) -> list[tuple[float, float]]:
    from ionosphere_fdtd_amd.sources import destination_point

    if sc.get("receiver_points") is not None:
        return list(sc["receiver_points"])
    if sc["receivers"] == "c0":
        return [destination_point(SOURCE_LAT, SOURCE_LON, 90.0, d * 1e3, sc["earth_radius_m"])
                for d in (1000.0, 5000.0, 10000.0, 20000.0)]
    rng = np.random.default_rng(RECEIVER_SEED)
    n = 32 if sc["receivers"] == "c1" else int(sc["receivers"])
    lat = np.rad2deg(np.arcsin(rng.uniform(-1.0, 1.0, n)))
    lon = rng.uniform(-180.0, 180.0, n)
    return list(zip(lat.tolist(), lon.tolist()))


def receiver_tables(mesh, altitudes_m, points):
    from ionosphere_fdtd_amd.sources import barycentric_distribution

    verts, layers, weights = [], [], []
    for lat, lon in points:
        v, layer, w = barycentric_distribution(mesh, altitudes_m, lat, lon, 0.0)
        verts.append(v)
        layers.append(layer)
        weights.append(w)
    return np.asarray(verts), np.asarray(layers), np.asarray(weights)


class VacuumMaterial:
    def sample(self, directions, altitudes_m, earth_radius_m):
        shape = (len(directions), len(altitudes_m))
        return np.zeros(shape), np.ones(shape)


def _plasma_arrays(sc, mesh, midpoint_altitudes):
    p = sc["plasma"]
    density, collision = drude_profile(midpoint_altitudes, p["h_prime_km"], p["beta_per_km"])
    F, R = mesh.n_faces, len(midpoint_altitudes)
    dens = np.broadcast_to(density, (F, R)).copy()
    coll = np.broadcast_to(collision, (F, R)).copy()
    b = np.zeros((F, R, 3))
    species = [("electron", -1.602_176_634e-19, 9.109_383_713_9e-31, dens, coll)]
    if p.get("magnetized"):
        # dipole-like field along the face normal tilted by latitude + an ion species
        n = mesh.face_centers
        tilt = np.stack((n[:, 1], -n[:, 0], 0.3 * np.ones(F)), axis=1)
        b = np.broadcast_to((5e-5 * tilt / np.linalg.norm(tilt, axis=1, keepdims=True))[:, None, :],
                            (F, R, 3)).copy()
    if p.get("magnetized") or p.get("ions"):
        species.append(("o_plus", 1.602_176_634e-19, 16 * 1.672_621_923_69e-27, 0.5 * dens,
                        0.01 * coll))
    return b, species


def build_sim(sc: dict, *, compile_step=False, compile_chunk_size=8, dtype=None,
              storage_scales=None, storage_scaling="auto", mesh=None,
              storage_rounding="rotation", thread_mapping="auto"):
    import ionosphere_fdtd_amd as g

    cfg = g.SimulationConfig(
        subdivision=sc["subdivision"], radial_cells=sc["radial_cells"],
        minimum_altitude_m=sc["minimum_altitude_m"], maximum_altitude_m=sc["maximum_altitude_m"],
        earth_radius_m=sc["earth_radius_m"], courant_factor=sc["courant_factor"],
        radial_boundary_condition="surface-impedance" if sc["surface_sigma"] else "pec",
        geometry_mode=sc["geometry_mode"], loss_integration=sc["loss_integration"],
        time_step_s=sc["time_step_s"], radial_altitudes_m=sc["radial_altitudes_m"],
        radial_grid_policy="balanced-2to1" if sc["radial_altitudes_m"] else "smooth")
    material = VacuumMaterial() if sc["vacuum"] else g.EarthIonosphereMaterial(**sc["material"])
    surface = g.ConductiveHalfSpaceSurface(sc["surface_sigma"]) if sc["surface_sigma"] else None
    if mesh is None:
        mesh = g.build_geodesic_mesh(sc["subdivision"])
    plasma = None
    if sc["plasma"]:
        alt = _midpoint_altitudes(sc)
        b, species = _plasma_arrays(sc, mesh, alt)
        plasma = g.MeshPlasmaModel.from_mesh(
            mesh, alt, b, tuple(g.ColdPlasmaSpecies(*s) for s in species))
    kwargs = dict(material=material, surface_impedance=surface, plasma=plasma, mesh=mesh,
                  backend="hip", dtype=dtype or sc["dtype"], compile_step=compile_step,
                  compile_chunk_size=compile_chunk_size, storage_scales=storage_scales,
                  storage_scaling=storage_scaling, storage_rounding=storage_rounding,
                  thread_mapping=thread_mapping)
    from ionosphere_fdtd_amd.solver import stable_time_step

    timed = [v for part in (sc["source"] or {}, sc["tangential"] or {}) for v in part.values()
             if isinstance(v, str) and v.endswith("dt")]
    dt = stable_time_step(cfg, mesh, material, plasma) if timed else None
    source = None
    if sc["source"]:
        s = {k: _resolve_time(v, dt) for k, v in sc["source"].items()}
        source = g.GaussianCurrent(**s)
    elif sc["tangential"]:
        s = {k: _resolve_time(v, dt) for k, v in sc["tangential"].items()}
        source = g.TangentialGaussianCurrent(**s)
    return g.GeodesicFDTD(cfg, source=source, **kwargs)


def _midpoint_altitudes(sc):
    if sc["radial_altitudes_m"] is not None:
        alt = np.asarray(sc["radial_altitudes_m"], dtype=np.float64)
    else:
        alt = np.linspace(sc["minimum_altitude_m"], sc["maximum_altitude_m"],
                          sc["radial_cells"] + 1)
    radii = sc["earth_radius_m"] + alt
    return 0.5 * (radii[:-1] + radii[1:]) - sc["earth_radius_m"]


def build_oracle(sc: dict, *, dtype=None, _dt_only=False, mesh=None):
    """The oracle's OracleSim for a scenario.  ``mesh`` may be the product's
    mesh (bit-identical to the reference's, so to the oracle's own builder):
    at s7 and above it saves the oracle's loop-based mesh construction."""
    o = oracle_module()
    if mesh is None:
        mesh = o.build_mesh(sc["subdivision"])
    plasma = None
    if sc["plasma"]:
        alt = _midpoint_altitudes(sc)
        b, species = _plasma_arrays(sc, mesh, alt)
        plasma = o.OraclePlasma(b, [o.OracleSpecies(q, m, n, nu) for _, q, m, n, nu in species])
    material_arrays = None
    if sc["vacuum"]:
        R = sc["radial_cells"]
        material_arrays = (np.zeros((mesh.n_vertices, R + 1)), np.ones((mesh.n_vertices, R + 1)),
                           np.zeros((mesh.n_edges, R)), np.ones((mesh.n_edges, R)))
    common = dict(subdivision=sc["subdivision"], radial_cells=sc["radial_cells"],
                  minimum_altitude_m=sc["minimum_altitude_m"],
                  maximum_altitude_m=sc["maximum_altitude_m"],
                  earth_radius_m=sc["earth_radius_m"], courant_factor=sc["courant_factor"],
                  time_step_s=sc["time_step_s"], loss_integration=sc["loss_integration"],
                  radial_altitudes_m=sc["radial_altitudes_m"],
                  geometry_mode=sc["geometry_mode"], material=None if sc["vacuum"] else sc["material"],
                  material_arrays=material_arrays,
                  surface=o.OracleSurface(sc["surface_sigma"]) if sc["surface_sigma"] else None,
                  plasma=plasma, mesh=mesh, dtype=dtype or "float64")
    if _dt_only:
        return o.OracleSim(**common)
    dt = o.OracleSim(**common).time_step_s
    if sc["source"]:
        s = {k: _resolve_time(v, dt) for k, v in sc["source"].items()}
        common["source"] = o.OracleSource(**s)
    sim = o.OracleSim(**common)
    if sc["tangential"]:
        s = {k: _resolve_time(v, dt) for k, v in sc["tangential"].items()}
        attach_tangential_source(sim, s)
    return sim


def attach_tangential_source(sim, params):
    """Give the oracle a TangentialGaussianCurrent (sources.py:257-335, solver.py:973-987)."""
    o = oracle_module()
    from ionosphere_fdtd_amd.sources import geographic_tangent_basis

    mesh = sim.mesh
    face = o.geographic_face_index(mesh, params["latitude_deg"], params["longitude_deg"])
    edges = mesh.face_edges[face].copy()
    ends = mesh.vertices[mesh.edges[edges]]
    dirs = ends[:, 1] - ends[:, 0]
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    east, north = geographic_tangent_basis(params["latitude_deg"], params["longitude_deg"])
    lengths = mesh.primal_edge_angles[edges] * (sim.earth_radius_m + params["altitude_m"])
    weights = np.zeros(len(edges))
    for az, line in zip(params["azimuths_deg"], params["line_lengths_m"]):
        a = np.deg2rad(az)
        weights += line * (dirs @ (np.cos(a) * north + np.sin(a) * east)) / lengths
    layers, rw = o.radial_linear_distribution(sim.radial_midpoint_altitudes_m, params["altitude_m"])
    n = len(layers)
    t_edges = np.repeat(edges, n)
    t_layers = np.tile(layers, len(edges))
    t_weights = np.repeat(weights, n) * np.tile(rw, len(edges))
    src = o.OracleSource(latitude_deg=params["latitude_deg"],
                         longitude_deg=params["longitude_deg"], altitude_m=params["altitude_m"],
                         center_time_s=params["center_time_s"],
                         one_over_e_half_width_s=params["one_over_e_half_width_s"])
    sim.source = src
    sim.source_distribution = None
    orig = sim.update_electric_fields

    def update(current_a=0.0):
        orig(current_a)
        cd = (t_weights * current_a * sim._inv_dual[t_edges, 0] / sim._mid[t_layers]
              / sim._steps[t_layers])
        sim.et[t_edges, t_layers] -= sim._cb_et[t_edges if sim._cb_et.shape[0] > 1 else 0,
                                                t_layers] * cd

    sim.update_electric_fields = update
    return sim


def apply_seed_fields(sim, set_field):
    """Seed the four fields of a scenario with seed_fields=True (same values on
    every implementation): set_field(name, values)."""
    R = sim.radial_midpoints_m.shape[0]
    shapes = {"er": (sim.mesh.n_vertices, R + 1), "et": (sim.mesh.n_edges, R),
              "hr": (sim.mesh.n_faces, R), "ht": (sim.mesh.n_edges, R + 1)}
    for name, values in seeded_fields(shapes).items():
        set_field(name, values)


def seeded_fields(shape_dict, seed=20260814, scale=1e-6):
    """Random initial fields like the reference benchmark (backend_matrix.py:179-183)."""
    rng = np.random.default_rng(seed)
    return {k: rng.standard_normal(v) * scale for k, v in shape_dict.items()}


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))


def family_errors(got: dict, want: dict) -> dict:
    """Per-field error relative to the norm of its field family.

    A field that is only roundoff in an excitation (Hr under a vertical
    source: Et stays a discrete gradient, so its face circulation vanishes)
    has no meaningful relative error of its own; measuring it against the
    (Er, Et) or (Hr, Ht) family norm keeps the bar meaningful.
    """
    norms = {}
    for fam in (("er", "et"), ("hr", "ht")):
        n = np.sqrt(sum(float(np.sum(np.asarray(want[k], dtype=np.float64) ** 2)) for k in fam))
        for k in fam:
            norms[k] = n
    out = {}
    for k in ("er", "et", "hr", "ht"):
        diff = np.linalg.norm(np.asarray(got[k], dtype=np.float64) - np.asarray(want[k],
                                                                                dtype=np.float64))
        out[k] = float(diff / norms[k]) if norms[k] > 0 else float(diff)
    return out


def gpu_available() -> bool:
    if os.environ.get("GEOFDTD_FORCE_NO_GPU"):
        return False
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False
