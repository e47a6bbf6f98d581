"""Host-side setup of the product package and the C-ABI boundary (CPU only).

The product builds its tables with vectorised code of its own; these tests
pin them against the reference fixtures (bit-exact mesh arrays at every
level used by the benchmark configurations, source / receiver
distributions, time steps) and against the oracle (surface-impedance and
plasma tables), and check that the shared library loads and exports every
function declared in include/geofdtd.h without touching a GPU.
"""

from __future__ import annotations

import hashlib
import json
import re
from pathlib import Path

import numpy as np
import pytest

import scenarios as S

import ionosphere_fdtd_amd as g
from ionosphere_fdtd_amd import _native
from ionosphere_fdtd_amd.plasma import plasma_tables
from ionosphere_fdtd_amd.solver import stable_time_step
from ionosphere_fdtd_amd.surface_impedance import surface_ade_tables

GOLDEN = Path(__file__).resolve().parent / "golden"
REPO = Path(__file__).resolve().parents[1]
o = S.oracle_module()
INT_FIELDS = ("faces", "edges", "face_edges", "face_edge_signs", "edge_left_faces",
              "edge_right_faces", "vertex_degree")
FLOAT_FIELDS = ("vertices", "face_centers", "primal_edge_angles", "dual_edge_angles",
                "face_solid_angles", "dual_cell_solid_angles")


def _sha(values):
    a = np.asarray(values)
    dt = "<f8" if np.issubdtype(a.dtype, np.floating) else "<i8"
    return hashlib.sha256(np.asarray(a, dtype=dt, order="C").tobytes()).hexdigest()


@pytest.mark.parametrize("s", range(9))
def test_product_mesh_is_bit_exact_with_reference(s):
    """Integer ordering and FP64 geometry identical to the reference, s0..s8
    (s8 is the precision-sweep mesh, 655 362 dual cells)."""
    hashes = json.loads((GOLDEN / "mesh_hashes.json").read_text())[f"s{s}_polar"]
    mesh = g.build_geodesic_mesh(s)
    for f in INT_FIELDS + FLOAT_FIELDS:
        assert _sha(getattr(mesh, f)) == hashes[f], f


@pytest.mark.parametrize("s", range(4))
def test_product_native_orientation_matches_reference(s):
    fx = np.load(GOLDEN / "mesh_s0_s3.npz")
    mesh = g.build_geodesic_mesh(s, orientation="native")
    for f in INT_FIELDS:
        assert np.array_equal(getattr(mesh, f), fx[f"s{s}_native_{f}"]), f


def test_vertex_incidence_matches_oracle():
    mesh = g.build_geodesic_mesh(3)
    from ionosphere_fdtd_amd.mesh import vertex_incidence

    ve, vs = vertex_incidence(mesh)
    oe, os_ = o.vertex_incidence(o.build_mesh(3))
    assert np.array_equal(ve, oe)
    assert np.array_equal(vs, os_.astype(np.int64))


@pytest.mark.parametrize("name", ("c0", "c0_plasma", "magnetized", "tangential",
                                  "thin_trapezoid", "pec_vacuum"))
def test_time_step_and_distributions_match_reference(name):
    sc = S.scenario(name)
    fx = np.load(GOLDEN / f"trajectory_{name}.npz")
    mesh = g.build_geodesic_mesh(sc["subdivision"])
    cfg = g.SimulationConfig(subdivision=sc["subdivision"], radial_cells=sc["radial_cells"],
                             minimum_altitude_m=sc["minimum_altitude_m"],
                             maximum_altitude_m=sc["maximum_altitude_m"],
                             earth_radius_m=sc["earth_radius_m"],
                             courant_factor=sc["courant_factor"])
    material = S.VacuumMaterial() if sc["vacuum"] else g.EarthIonosphereMaterial(**sc["material"])
    plasma = None
    if sc["plasma"]:
        alt = S._midpoint_altitudes(sc)
        b, species = S._plasma_arrays(sc, mesh, alt)
        plasma = g.MeshPlasmaModel.from_mesh(mesh, alt, b,
                                             tuple(g.ColdPlasmaSpecies(*s) for s in species))
    assert stable_time_step(cfg, mesh, material, plasma) == float(fx["time_step_s"])
    altitudes = np.linspace(sc["minimum_altitude_m"], sc["maximum_altitude_m"],
                            sc["radial_cells"] + 1)
    verts, layers, weights = S.receiver_tables(mesh, altitudes,
                                               [tuple(p) for p in fx["receiver_points"]])
    assert np.array_equal(verts, fx["receiver_vertices"])
    assert np.array_equal(layers, fx["receiver_layers"])
    np.testing.assert_allclose(weights, fx["receiver_weights"], rtol=0, atol=1e-15)
    if "source_vertices" in fx:
        from ionosphere_fdtd_amd.sources import barycentric_distribution, radial_linear_distribution

        src = sc["source"]
        v, _, hw = barycentric_distribution(mesh, altitudes, src["latitude_deg"],
                                            src["longitude_deg"], src["altitude_m"])
        lay, rw = radial_linear_distribution(altitudes, src["altitude_m"])
        n = len(lay)
        assert np.array_equal(np.repeat(v, n), fx["source_vertices"])
        assert np.array_equal(np.tile(lay, len(v)), fx["source_layers"])
        np.testing.assert_allclose(np.repeat(hw, n) * np.tile(rw, len(v)), fx["source_weights"],
                                   rtol=0, atol=1e-15)


def test_source_waveform_vectorised_equals_scalar_form():
    src = g.GaussianCurrent(carrier_frequency_hz=20.0)
    ref = o.OracleSource(carrier_frequency_hz=20.0)
    dt = 2.4e-6
    times = (np.arange(5000) + 0.5) * dt * 37
    got = src.currents_a(times, dt)
    want = np.array([ref.current_a(t, dt) for t in times])
    # near carrier zero crossings the cosine's relative error grows; bound by the peak
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-14 * src.peak_current_a)
    assert src.current_a(0.1, 1e-6) == 1.0e6  # peak at the default centre (4 half widths)


def test_surface_impedance_tables_match_oracle():
    model = g.ConductiveHalfSpaceSurface(1e-2)
    t = surface_ade_tables(model, 480, 2.4e-6)
    alpha = 0.5 * 2.4e-6 * o.si_pole_rates()
    np.testing.assert_array_equal(model.pole_rates_s_inv, o.si_pole_rates())
    np.testing.assert_array_equal(model.diffusive_weights_sqrt_s_inv, o.si_weights())
    np.testing.assert_array_equal(t["decay"], (1 - alpha) / (1 + alpha))
    np.testing.assert_array_equal(t["history_weights"], o.si_weights() / (1 + alpha))
    # positive-real fit within the reference's 6e-3 of sqrt(mu s / sigma) over the band
    f = np.linspace(5.0, 45.0, 41)
    fit = model.impedance_ohm(f)[0]
    exact = model.exact_impedance_ohm(f)[0]
    assert np.max(np.abs(fit - exact) / np.abs(exact)) < 6e-3


def test_plasma_tables_match_oracle_coupler():
    sc = S.scenario("magnetized")
    mesh = g.build_geodesic_mesh(sc["subdivision"])
    alt = S._midpoint_altitudes(sc)
    b, species = S._plasma_arrays(sc, mesh, alt)
    model = g.MeshPlasmaModel.from_mesh(mesh, alt, b, tuple(g.ColdPlasmaSpecies(*s)
                                                            for s in species))
    dt = 1.6e-7
    t = plasma_tables(model, mesh, dt)
    om = o.build_mesh(sc["subdivision"])
    coupler = o.OraclePlasmaCoupler(o.OraclePlasma(b, [o.OracleSpecies(q, m, n, nu)
                                                       for _, q, m, n, nu in species]), om, dt)
    np.testing.assert_array_equal(t["reconstruction"], coupler.reconstruction)
    np.testing.assert_array_equal(t["left_tangents"], coupler.left_tangents)
    np.testing.assert_array_equal(t["right_tangents"], coupler.right_tangents)
    assert np.array_equal(t["vertex_faces"], coupler.vertex_faces)
    np.testing.assert_array_equal(t["vertex_face_weights"], coupler.vertex_face_weights)
    rows = t["coef"].shape[2]
    for s, coefs in enumerate(coupler.coefficients):
        for q, c in enumerate(coefs):
            want = c if rows > 1 else c[:1]
            np.testing.assert_array_equal(t["coef"][s, q], want)
    assert t["magnetized"]
    assert t["band"] == (int(np.flatnonzero(alt >= 60e3)[0]), len(alt))


def test_pinv_reconstruction_recovers_tangential_vectors():
    """plasma.py reconstruction: any tangent vector is recovered from its three
    edge-tangent projections (test_plasma.py:98-108 tolerance 8e-16 scale)."""
    mesh = g.build_geodesic_mesh(2)
    model = g.electron_drude_model(mesh, np.array([80e3]), 1e9, 1e5)
    t = plasma_tables(model, mesh, 1e-7)
    rng = np.random.default_rng(3)
    v = rng.standard_normal((mesh.n_faces, 3))
    v -= np.einsum("ij,ij->i", v, mesh.face_centers)[:, None] * mesh.face_centers
    proj = np.einsum("fcd,fd->fc", t["face_tangents"], v)
    back = np.einsum("fc,fcd->fd", proj, t["reconstruction"])
    np.testing.assert_allclose(back, v, atol=8e-15)


def test_config_validation_mirrors_reference():
    with pytest.raises(ValueError, match="radial_cells"):
        g.SimulationConfig(radial_cells=1)
    with pytest.raises(ValueError, match="surface-impedance"):
        g.SimulationConfig(radial_boundary_condition="surface-impedance")
    with pytest.raises(ValueError, match="courant"):
        g.SimulationConfig(courant_factor=1.5)
    with pytest.raises(ValueError, match="integers"):
        g.SimulationConfig(subdivision=2.0)
    with pytest.raises(ValueError, match="reversed"):
        g.SimulationConfig(minimum_altitude_m=10.0, maximum_altitude_m=0.0)
    with pytest.raises(ValueError, match="geometry_mode"):
        g.SimulationConfig(geometry_mode="flat")
    g.SimulationConfig(radial_cells=3, minimum_altitude_m=0.0, maximum_altitude_m=3.0,
                       radial_altitudes_m=(0.0, 1.0, 2.0, 3.0))


def test_library_exports_every_declared_symbol():
    header = "".join(p.read_text() for p in sorted((REPO / "include").glob("*.h")))
    declared = set(re.findall(r"\b(geo_[a-z_0-9]+)\s*\(", header))
    assert declared == set(_native.EXPORTED_SYMBOLS)
    lib = _native.load_library()
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.geo_version() == 1


def test_product_refuses_cpu_backends_and_missing_device(monkeypatch):
    with pytest.raises(g.BackendUnavailableError):
        g.create_backend("numpy", g.build_geodesic_mesh(0))
    with pytest.raises(ValueError):
        g.create_backend("jax", g.build_geodesic_mesh(0))
    import torch

    if not torch.cuda.is_available():
        with pytest.raises(g.BackendUnavailableError):
            g.GeodesicFDTD(g.SimulationConfig(subdivision=0, radial_cells=2))


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(g.BackendUnavailableError, match="not built"):
        _native.load_library(tmp_path / "libgeofdtd.so")


def test_radial_refinement_matches_reference_rules():
    from ionosphere_fdtd_amd.radial_grid import RadialRefinementRegion

    nodes = g.build_refined_radial_grid(0.0, 100e3, 2e3,
                                        (RadialRefinementRegion(60e3, 90e3, 500.0),))
    steps = np.diff(nodes)
    assert np.isclose(steps.min(), 500.0) and np.isclose(steps.max(), 2000.0)
    assert np.max(np.maximum(steps[:-1] / steps[1:], steps[1:] / steps[:-1])) <= 2.0


def test_radial_refinement_equals_reference_vectors():
    """Nodes bit-equal to the reference's build_refined_radial_grid (or the
    same error) on its own test cases and 60 seeded random domains with up to
    three overlapping regions (tests/golden/radial_grids.json)."""
    from ionosphere_fdtd_amd.radial_grid import RadialRefinementRegion

    cases = json.loads((GOLDEN / "radial_grids.json").read_text())
    assert len(cases) == 66
    built = 0
    for case in cases:
        lo, hi, bg = (float.fromhex(case[k]) for k in
                      ("minimum_altitude_m", "maximum_altitude_m", "background_step_m"))
        args = (lo, hi, bg)
        if "error" in case:
            with pytest.raises(ValueError) as info:
                regions = tuple(RadialRefinementRegion(*map(float.fromhex, r))
                                for r in case["regions"])
                g.build_refined_radial_grid(*args, regions)
            assert str(info.value) == case["error"]
            continue
        regions = tuple(RadialRefinementRegion(*map(float.fromhex, r)) for r in case["regions"])
        nodes = g.build_refined_radial_grid(*args, regions)
        assert [x.hex() for x in nodes] == case["nodes"]
        built += 1
    assert built >= 60


def test_radial_refinement_reference_cases():
    """tests/test_radial_grid.py:12-105 of the reference, restated."""
    from ionosphere_fdtd_amd.radial_grid import RadialRefinementRegion, validate_radial_grid

    region = RadialRefinementRegion(60_000.0, 90_000.0, 1_000.0)
    alt = g.build_refined_radial_grid(0.0, 100_000.0, 10_000.0, (region,))
    v = validate_radial_grid(alt)
    steps, lower, upper = np.diff(alt), np.asarray(alt[:-1]), np.asarray(alt[1:])
    inside = (upper > region.minimum_altitude_m) & (lower < region.maximum_altitude_m)
    assert alt[0] == 0.0 and alt[-1] == 100_000.0
    assert v.cells == 61 and v.maximum_adjacent_step_ratio <= 2.0
    assert np.max(steps[inside]) <= region.maximum_step_m

    regions = (RadialRefinementRegion(40_000.0, 80_000.0, 2_000.0),
               RadialRefinementRegion(60_000.0, 70_000.0, 500.0))
    alt = g.build_refined_radial_grid(0.0, 100_000.0, 10_000.0, regions)
    steps, lower, upper = np.diff(alt), np.asarray(alt[:-1]), np.asarray(alt[1:])
    for r in regions:
        inside = (upper > r.minimum_altitude_m) & (lower < r.maximum_altitude_m)
        assert np.max(steps[inside]) <= r.maximum_step_m
    assert validate_radial_grid(alt).maximum_adjacent_step_ratio <= 2.0

    with pytest.raises(ValueError, match="adjacent-step ratio"):
        g.SimulationConfig(subdivision=0, radial_cells=3, minimum_altitude_m=0.0,
                           maximum_altitude_m=7_000.0,
                           radial_altitudes_m=(0.0, 1_000.0, 2_000.0, 7_000.0),
                           radial_grid_policy="balanced-2to1")
    with pytest.raises(ValueError, match="outside"):
        g.build_refined_radial_grid(0.0, 100_000.0, 10_000.0,
                                    (RadialRefinementRegion(-1.0, 1_000.0, 100.0),))
    with pytest.raises(ValueError, match="reversed"):
        RadialRefinementRegion(2.0, 1.0, 1.0)
    with pytest.raises(TypeError):
        g.build_refined_radial_grid(0.0, 1.0, 0.1, ((0.0, 1.0, 0.1),))


def test_nearest_narrow_storage_with_slow_source_warns(monkeypatch):
    """Round-to-nearest narrow storage under a source whose envelope spans
    more than 1e4 steps warns (the stagnation check runs on the host only)."""
    from ionosphere_fdtd_amd import solver

    class Probe:
        _narrow = solver.GeodesicFDTD._narrow
        _warn_if_narrow_storage_stagnates = solver.GeodesicFDTD._warn_if_narrow_storage_stagnates

    for rounding, source, expect in (("nearest", g.GaussianCurrent(carrier_frequency_hz=20.0), True),
                                     ("rotation", g.GaussianCurrent(carrier_frequency_hz=20.0), False),
                                     ("nearest", g.GaussianCurrent(one_over_e_half_width_s=1e-6), False)):
        probe = Probe()
        probe.storage_rounding = rounding
        probe.source = source
        probe.time_step_s = 1.27e-7
        probe.backend = type("B", (), {"dtype_name": "float16"})()
        import warnings

        with warnings.catch_warnings(record=True) as caught:
            warnings.simplefilter("always")
            probe._warn_if_narrow_storage_stagnates()
        assert any("stagnate" in str(w.message) for w in caught) == expect, (rounding, source)
    with pytest.raises(ValueError, match="storage_rounding"):
        g.GeodesicFDTD(g.SimulationConfig(subdivision=0, radial_cells=2), dtype="float16",
                       storage_rounding="up")
    with pytest.raises(ValueError, match="thread_mapping"):
        g.GeodesicFDTD(g.SimulationConfig(subdivision=0, radial_cells=2), thread_mapping="cols")


def test_trace_query_rings_and_pair_weights():
    """Vertex rings equal breadth-first graph distances along primal edges,
    and a query's vertex-pair weights s_j w_i sum to one."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import shortest_path

    from ionosphere_fdtd_amd import trace_queries as tq

    mesh = g.build_geodesic_mesh(3)
    rings = tq.MeshRings(mesh)
    e = np.asarray(mesh.edges)
    adj = coo_matrix((np.ones(len(e)), (e[:, 0], e[:, 1])), shape=(mesh.n_vertices,) * 2)
    dist = shortest_path(adj, directed=False, unweighted=True)
    for u in (0, 11, 137, mesh.n_vertices - 1):
        verts, ring = rings.neighbourhood(u)
        assert np.array_equal(np.sort(verts), np.flatnonzero(dist[u] <= tq.RINGS))
        assert np.array_equal(ring, dist[u, verts].astype(int))
        assert verts[0] == u and np.all(np.diff(ring) >= 0)
    s5 = g.build_geodesic_mesh(tq.SUBDIVISION)
    q = tq.random_queries(6, 3, seed=5)
    (qi, u, v, w), covered = tq.decompose(q, s5)
    assert covered.all()
    np.testing.assert_allclose(np.bincount(qi, weights=w, minlength=len(q)), 1.0, atol=1e-12)
    far = q[:1].copy()
    far[0, 5], far[0, 6] = tq.destination_point(far[0, 3], far[0, 4], 0.0, 300e3,
                                                tq.EARTH_RADIUS_M)
    assert not tq.in_range(far, s5)[0]
    assert tq.in_range(q, s5).all()


def test_batched_surface_tables_equal_the_per_point_distribution_bitwise():
    """The trace-query service's batched face search and barycentric weights
    equal sources.barycentric_distribution point by point, bit for bit: at
    random points and at the ties of the face search (mesh vertices, edge
    midpoints, face centres, the poles and the antimeridian)."""
    from ionosphere_fdtd_amd import trace_queries as tq
    from ionosphere_fdtd_amd.sources import barycentric_distribution

    mesh = g.build_geodesic_mesh(5)
    rng = np.random.default_rng(17)
    lat = np.degrees(np.arcsin(rng.uniform(-1.0, 1.0, 2000)))
    lon = rng.uniform(-180.0, 180.0, 2000)
    v = np.asarray(mesh.vertices)
    e = np.asarray(mesh.edges)
    mids = v[e[:200, 0]] + v[e[:200, 1]]
    special = np.concatenate([v[:200], mids / np.linalg.norm(mids, axis=1, keepdims=True),
                              np.asarray(mesh.face_centers)[:200]])
    lat = np.concatenate([lat, np.degrees(np.arcsin(np.clip(special[:, 2], -1.0, 1.0))),
                          [90.0, -90.0, 0.0]])
    lon = np.concatenate([lon, np.degrees(np.arctan2(special[:, 1], special[:, 0])),
                          [0.0, 0.0, 180.0]])
    tables = tq.surface_tables(mesh, lat, lon)
    for i, (la, lo) in enumerate(zip(lat, lon)):
        verts, _, weights = barycentric_distribution(mesh, [0.0, 1.0], la, lo, 0.0)
        assert np.array_equal(tables[i][0], verts), i
        assert tables[i][2].tobytes() == weights.tobytes(), i


def test_committee_sums_pair_traces_per_query_in_pair_order():
    """The committee's trace of a query is the sum of its pairs' weighted
    traces, added in pair order (the np.add.at restatement, bit for bit); the
    disagreement is the largest member distance from the mean."""
    from ionosphere_fdtd_amd import trace_queries as tq

    class Member:  # a deterministic stand-in for a trained network
        amp_coef = None  # no host amplitude baseline

        def __init__(self, phase):
            self.phase = phase

        def predict_pairs(self, feats, basis=None):
            k = np.arange(tq.SAMPLES)
            return np.sin(np.outer(feats @ np.linspace(0.3, 1.1, feats.shape[1]), k / 97.0)
                          + self.phase) * np.exp(feats[:, :1])

    s5 = g.build_geodesic_mesh(tq.SUBDIVISION)
    q = tq.random_queries(7, 2, seed=9)
    members = [Member(0.0), Member(0.2)]
    mean, spread = tq.SurrogateCommittee(members).predict(q, s5)
    (qi, u, v, w), _ = tq.decompose(q, s5)
    dts = tq.time_steps(q[:, 0], q[:, 1], s5)
    feats = tq.pair_features(q[qi][:, [0, 1, 2, 7]], dts[qi], u, v, tq.mesh_rings(s5))
    traces = []
    for m in members:
        t = np.zeros((len(q), tq.SAMPLES))
        np.add.at(t, qi, m.predict_pairs(feats) * w[:, None])
        traces.append(t)
    want = np.mean(traces, axis=0)
    assert mean.tobytes() == want.tobytes()
    np.testing.assert_array_equal(
        spread, np.max([tq.trace_rel_l2(t, want) for t in traces], axis=0))
    assert np.all(spread > 0)


def test_range_test_hands_over_the_pairs_of_the_in_range_queries():
    """in_range(..., with_pairs=True) returns decompose() of exactly the
    in-range queries (indexed among them), so the service decomposes once;
    queries outside the physics box or too far from their source drop out."""
    from ionosphere_fdtd_amd import trace_queries as tq

    s5 = g.build_geodesic_mesh(tq.SUBDIVISION)
    q = tq.random_queries(9, 2, seed=21)
    q[3, 0] = tq.BOX["h_prime_km"][1] + 5.0  # physics outside the box
    q[10, 5], q[10, 6] = tq.destination_point(q[10, 3], q[10, 4], 0.0, 400e3, tq.EARTH_RADIUS_M)
    ok, pairs = tq.in_range(q, s5, with_pairs=True)
    assert np.array_equal(ok, tq.in_range(q, s5))
    assert not ok[3] and not ok[10] and ok.sum() == len(q) - 2
    want, covered = tq.decompose(q[ok], s5)
    assert covered.all()
    for got, exp in zip(pairs, want):
        assert got.dtype == exp.dtype and got.tobytes() == exp.tobytes()
    none_ok, none_pairs = tq.in_range(q[[3]], s5, with_pairs=True)
    assert not none_ok.any() and all(len(a) == 0 for a in none_pairs)
