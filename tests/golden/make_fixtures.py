"""Write golden vectors from the reference implementation (run in the build
container only; the reference is not available where the tests run).

    PYTHONPATH=/root/reference/src python tests/golden/make_fixtures.py

Outputs (data only, no reference code):
  * mesh_s0_s3.npz       -- integer topology arrays of build_geodesic_mesh(s),
                            s = 0..3, polar and native orientation;
  * mesh_hashes.json     -- sha256 of every mesh array (canonical <i8 / <f8
                            bytes) for s = 0..7;
  * trajectory_<name>.npz -- GeodesicFDTD (NumPy backend, float64) runs of the
                            scenarios in tests/scenarios.py: time step, source
                            and receiver distributions, Er traces, final fields,
                            surface-impedance memory and plasma currents;
  * full_field.csv       -- artifacts/analytic-solutions/full-field.csv of the
                            reference (measured cavity frequencies and time
                            steps), copied as data;
  * checkpoint_<name>.npz -- a reference checkpoint (format version 4) after 60
                            steps, and checkpoint_<name>_continued.npz the
                            reference's fields after reloading it and running
                            40 more steps;
  * radial_grids.json    -- build_refined_radial_grid inputs and the reference's
                            nodes (float.hex) or error message, for the
                            reference's own test cases and seeded random ones.

    PYTHONPATH=/root/reference/src python tests/golden/make_fixtures.py c3_seeded
    PYTHONPATH=/root/reference/src python tests/golden/make_fixtures.py radial_grids
    PYTHONPATH=/root/reference/src python tests/golden/make_fixtures.py mesh_hash 8
write one trajectory, the radial-grid vectors, or the hashes of one more
subdivision (merged into mesh_hashes.json).
These pin the parts of the oracle that no closed form pins: integer mesh
ordering, and the end-to-end conventions of the surface-impedance ADE, the
plasma coupler and the source/receiver distributions.
"""

from __future__ import annotations

import hashlib
import json
import shutil
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parents[1]
REFERENCE = Path("/root/reference")
sys.path.insert(0, str(REFERENCE / "src"))
sys.path.insert(0, str(REPO / "tests"))

from ionosphere_fdtd.data_artifacts import DatasetProvenance, VariableProvenance  # noqa: E402
from ionosphere_fdtd.materials import EarthIonosphereMaterial  # noqa: E402
from ionosphere_fdtd.mesh import build_geodesic_mesh  # noqa: E402
from ionosphere_fdtd.plasma import ColdPlasmaSpecies, MeshPlasmaModel  # noqa: E402
from ionosphere_fdtd.radial_grid import (RadialRefinementRegion,  # noqa: E402
                                         build_refined_radial_grid)
from ionosphere_fdtd.solver import GeodesicFDTD, SimulationConfig  # noqa: E402
from ionosphere_fdtd.sources import (GaussianCurrent, TangentialGaussianCurrent,  # noqa: E402
                                     geographic_distribution)
from ionosphere_fdtd.surface_impedance import ConductiveHalfSpaceSurface  # noqa: E402

import scenarios as S  # noqa: E402  (parameter dicts only)

INT_FIELDS = ("faces", "edges", "face_edges", "face_edge_signs", "edge_left_faces",
              "edge_right_faces", "vertex_degree")
FLOAT_FIELDS = ("vertices", "face_centers", "primal_edge_angles", "dual_edge_angles",
                "face_solid_angles", "dual_cell_solid_angles")
STEPS = {"c0": 1000, "c0_pulse": 200, "c0_plasma": 200, "magnetized": 150,
         "thin_trapezoid": 150, "tangential": 150, "pec_vacuum": 100, "plasma_random": 120,
         "magnetized_random": 120, "c3_seeded": 150}


def canonical_sha(values) -> str:
    a = np.asarray(values)
    dt = "<f8" if np.issubdtype(a.dtype, np.floating) else "<i8"
    return hashlib.sha256(np.asarray(a, dtype=dt, order="C").tobytes()).hexdigest()


def provenance():
    return DatasetProvenance(
        dataset_id="golden.synthetic.drude.v1", title="Synthetic Drude profile",
        version="1", source_url="https://example.invalid/drude", citation="Synthetic.",
        license="CC0-1.0", retrieved_at="2026-09-01T00:00:00Z", source_sha256="0" * 64,
        coordinate_reference_system="geocentric Cartesian",
        variables=(VariableProvenance("B", "T", "T", "identity"),
                   VariableProvenance("number_density", "m^-3", "m^-3", "identity"),
                   VariableProvenance("collision_frequency", "Hz", "Hz", "identity")))


class Vacuum:
    def sample(self, directions, altitudes_m, earth_radius_m):
        shape = (len(directions), len(altitudes_m))
        return np.zeros(shape), np.ones(shape)


def reference_sim(sc):
    cfg = SimulationConfig(
        subdivision=sc["subdivision"], radial_cells=sc["radial_cells"],
        minimum_altitude_m=sc["minimum_altitude_m"], maximum_altitude_m=sc["maximum_altitude_m"],
        earth_radius_m=sc["earth_radius_m"], courant_factor=sc["courant_factor"],
        radial_boundary_condition="surface-impedance" if sc["surface_sigma"] else "pec",
        geometry_mode=sc["geometry_mode"], loss_integration=sc["loss_integration"],
        radial_altitudes_m=sc["radial_altitudes_m"],
        radial_grid_policy="balanced-2to1" if sc["radial_altitudes_m"] else "smooth")
    material = Vacuum() if sc["vacuum"] else EarthIonosphereMaterial(**sc["material"])
    surface = ConductiveHalfSpaceSurface(sc["surface_sigma"]) if sc["surface_sigma"] else None
    mesh = build_geodesic_mesh(sc["subdivision"])
    plasma = None
    if sc["plasma"]:
        alt = S._midpoint_altitudes(sc)
        b, species = S._plasma_arrays(sc, mesh, alt)
        plasma = MeshPlasmaModel.from_mesh(
            mesh, alt, b, tuple(ColdPlasmaSpecies(*s) for s in species),
            provenance=(provenance(),), interpolation="synthetic")
    probe = GeodesicFDTD(cfg, material=material, surface_impedance=surface, plasma=plasma,
                         mesh=mesh, dtype="float64")
    dt = probe.time_step_s
    source = None
    if sc["source"]:
        source = GaussianCurrent(**{k: S._resolve_time(v, dt) for k, v in sc["source"].items()})
    elif sc["tangential"]:
        source = TangentialGaussianCurrent(
            **{k: S._resolve_time(v, dt) for k, v in sc["tangential"].items()})
    return GeodesicFDTD(cfg, material=material, source=source, surface_impedance=surface,
                        plasma=plasma, mesh=mesh, dtype="float64")


def receiver_points(sc):
    # same rule as tests/scenarios.receiver_points, restated without the product
    if sc["receivers"] == "c0":
        out = []
        lat0, lon0 = np.deg2rad(S.SOURCE_LAT), np.deg2rad(S.SOURCE_LON)
        start = np.array([np.cos(lat0) * np.cos(lon0), np.cos(lat0) * np.sin(lon0), np.sin(lat0)])
        east = np.array([-np.sin(lon0), np.cos(lon0), 0.0])
        for d in (1000.0, 5000.0, 10000.0, 20000.0):
            delta = d * 1e3 / sc["earth_radius_m"]
            p = np.cos(delta) * start + np.sin(delta) * east
            out.append((float(np.rad2deg(np.arcsin(np.clip(p[2], -1, 1)))),
                        float(np.rad2deg(np.arctan2(p[1], p[0])))))
        return out
    rng = np.random.default_rng(S.RECEIVER_SEED)
    lat = np.rad2deg(np.arcsin(rng.uniform(-1.0, 1.0, 32)))
    lon = rng.uniform(-180.0, 180.0, 32)
    return list(zip(lat.tolist(), lon.tolist()))


def write_trajectory(name):
    sc = S.scenario(name)
    sim = reference_sim(sc)
    out = {"time_step_s": np.float64(sim.time_step_s),
           "cfl_time_step_limit_s": np.float64(sim.cfl_time_step_limit_s)}
    if sc["seed_fields"]:
        S.apply_seed_fields(sim, lambda k, v: getattr(sim, k).__setitem__(slice(None), v))
    if sim._source_distribution is not None:
        v, l, w = sim.source.staggered_distribution(sim)
        out.update(source_vertices=v, source_layers=l, source_weights=w)
    if sim._tangential_source_distribution is not None:
        e, l, w = sim.source.edge_distribution(sim)
        out.update(tsource_edges=e, tsource_layers=l, tsource_weights=w)
    pts = receiver_points(sc)
    rv, rl, rw = [], [], []
    for lat, lon in pts:
        v, layer, w = geographic_distribution(sim, lat, lon, 0.0)
        rv.append(v)
        rl.append(layer)
        rw.append(w)
    rv, rl, rw = np.asarray(rv), np.asarray(rl), np.asarray(rw)
    out.update(receiver_points=np.asarray(pts), receiver_vertices=rv, receiver_layers=rl,
               receiver_weights=rw)
    traces = sim.record_er_observations(rv, rl, rw, STEPS[name])
    out.update(traces=traces, steps=np.int64(STEPS[name]))
    for k in ("er", "et", "hr", "ht"):
        out[k] = np.asarray(getattr(sim, k))
    if sim._surface_impedance_ade is not None:
        out["si_memory"] = np.asarray(sim._surface_impedance_ade.memory)
    if sim._plasma_coupler is not None:
        for i, j in enumerate(sim._plasma_coupler.ade.current_density):
            out[f"plasma_current_{i}"] = np.asarray(j)
    np.savez_compressed(HERE / f"trajectory_{name}.npz", **out)
    print(name, "dt", sim.time_step_s, "max|er|", float(np.max(np.abs(out["er"]))), flush=True)


def write_checkpoint_case(name, first=60, second=40):
    """A reference checkpoint after `first` steps and the reference's own
    continuation for `second` more steps after reloading it."""
    sc = S.scenario(name)
    sim = reference_sim(sc)
    if sc["seed_fields"]:
        S.apply_seed_fields(sim, lambda k, v: getattr(sim, k).__setitem__(slice(None), v))
    sim.step(first)
    path = HERE / f"checkpoint_{name}.npz"
    sim.save_checkpoint(path)
    resumed = GeodesicFDTD.load_checkpoint(path, dtype="float64")
    resumed.step(second)
    out = {k: np.asarray(getattr(resumed, k)) for k in ("er", "et", "hr", "ht")}
    out["steps"] = np.int64(resumed.steps)
    if resumed._surface_impedance_ade is not None:
        out["si_memory"] = np.asarray(resumed._surface_impedance_ade.memory)
    if resumed._plasma_coupler is not None:
        for i, j in enumerate(resumed._plasma_coupler.ade.current_density):
            out[f"plasma_current_{i}"] = np.asarray(j)
    np.savez_compressed(HERE / f"checkpoint_{name}_continued.npz", **out)
    print("checkpoint", name, resumed.steps, flush=True)


def radial_grid_cases():
    """The reference's own cases (tests/test_radial_grid.py) plus seeded random
    domains, backgrounds and up to three overlapping regions."""
    cases = [((0.0, 100e3, 10e3), [(60e3, 90e3, 1e3)]),
             ((0.0, 100e3, 10e3), [(40e3, 80e3, 2e3), (60e3, 70e3, 500.0)]),
             ((0.0, 100e3, 10e3), [(60e3, 90e3, 2e3)]),
             ((0.0, 100e3, 2e3), [(60e3, 90e3, 500.0)]),
             ((0.0, 100e3, 10e3), [(-1.0, 1e3, 100.0)]),
             ((0.0, 100e3, 60e3), [])]
    rng = np.random.default_rng(20261204)
    for _ in range(60):
        lo = float(rng.uniform(-50e3, 0.0))
        hi = lo + float(rng.uniform(20e3, 200e3))
        bg = float(rng.uniform(1e3, 30e3))
        regions = []
        for _ in range(int(rng.integers(0, 4))):
            a = float(rng.uniform(lo, hi))
            b = float(rng.uniform(a, hi))
            if b > a:
                regions.append((a, b, float(bg * 2.0 ** -rng.uniform(0.0, 6.0))))
        cases.append(((lo, hi, bg), regions))
    return cases


def write_radial_grids():
    out = []
    for (lo, hi, bg), regions in radial_grid_cases():
        entry = {"minimum_altitude_m": lo.hex(), "maximum_altitude_m": hi.hex(),
                 "background_step_m": bg.hex(),
                 "regions": [[a.hex(), b.hex(), c.hex()] for a, b, c in regions]}
        try:
            nodes = build_refined_radial_grid(
                lo, hi, bg, tuple(RadialRefinementRegion(*r) for r in regions))
            entry["nodes"] = [float(x).hex() for x in nodes]
        except ValueError as exc:
            entry["error"] = str(exc)
        out.append(entry)
    (HERE / "radial_grids.json").write_text(json.dumps(out, indent=0))
    print("radial grids", len(out), flush=True)


def write_mesh_hash(s):
    path = HERE / "mesh_hashes.json"
    hashes = json.loads(path.read_text())
    m = build_geodesic_mesh(s, orientation="polar")
    hashes[f"s{s}_polar"] = {f: canonical_sha(getattr(m, f)) for f in INT_FIELDS + FLOAT_FIELDS}
    path.write_text(json.dumps(hashes, indent=1, sort_keys=True))
    print("mesh hash", s, flush=True)


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "radial_grids":
        write_radial_grids()
        return
    if len(sys.argv) > 2 and sys.argv[1] == "mesh_hash":
        write_mesh_hash(int(sys.argv[2]))
        return
    if len(sys.argv) > 1 and sys.argv[1] == "checkpoints":
        for name in ("c0_pulse", "plasma_random"):
            write_checkpoint_case(name)
        return
    if len(sys.argv) > 1:
        for name in sys.argv[1:]:
            write_trajectory(name)
        return
    meshes = {}
    hashes = {}
    for s in range(8):
        for orient in (("polar", "native") if s <= 3 else ("polar",)):
            m = build_geodesic_mesh(s, orientation=orient)
            key = f"s{s}_{orient}"
            hashes[key] = {f: canonical_sha(getattr(m, f)) for f in INT_FIELDS + FLOAT_FIELDS}
            if s <= 3:
                for f in INT_FIELDS:
                    meshes[f"{key}_{f}"] = np.asarray(getattr(m, f))
            print("mesh", key, flush=True)
    np.savez_compressed(HERE / "mesh_s0_s3.npz", **meshes)
    (HERE / "mesh_hashes.json").write_text(json.dumps(hashes, indent=1, sort_keys=True))
    for name in STEPS:
        write_trajectory(name)
    shutil.copyfile(REFERENCE / "artifacts" / "analytic-solutions" / "full-field.csv",
                    HERE / "full_field.csv")
    for name in ("c0_pulse", "plasma_random"):
        write_checkpoint_case(name)


if __name__ == "__main__":
    main()
