"""Material sampling modes of the host setup (reference solver.py:467-705,
materials.py:22-326) against vectors the reference wrote
(tests/golden/make_material_fixtures.py), plus the invariants that fix them:
wedge and quarter-diamond weights tile their cells, and a conservatively
rasterised anomaly covers exactly its area."""

from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import numpy as np
import pytest

import ionosphere_fdtd_amd as g
from ionosphere_fdtd_amd.solver import sample_material, stable_time_step

GOLDEN = Path(__file__).resolve().parent / "golden" / "materials_s2.npz"
CASES = {
    "point": ("point", "point", "point"),
    "dual_cell": ("dual-cell", "point", "point"),
    "edge_diamond": ("point", "edge-diamond", "point"),
    "conservative": ("point", "point", "conservative-nearest"),
    "all": ("dual-cell", "edge-diamond", "conservative-nearest"),
}
GRID = dict(subdivision=2, radial_cells=12, minimum_altitude_m=-20e3, maximum_altitude_m=100e3)


def _anomalies():
    return (
        g.SphericalAnomaly(latitude_deg=35.15, longitude_deg=126.85, radius_m=900e3,
                           altitude_min_m=-15e3, altitude_max_m=-2e3, conductivity_factor=40.0,
                           relative_permittivity=25.0,
                           maximum_background_conductivity_s_m=5e-3),
        g.SphericalAnomaly(latitude_deg=-20.0, longitude_deg=-60.0, radius_m=1500e3,
                           altitude_min_m=-20e3, altitude_max_m=5e3, conductivity_factor=0.1,
                           relative_permittivity=4.0, target_area_m2=5.0e12),
    )


def _grid():
    cfg = g.SimulationConfig(**GRID)
    alt = np.linspace(cfg.minimum_altitude_m, cfg.maximum_altitude_m, cfg.radial_cells + 1)
    return cfg, alt, 0.5 * (alt[1:] + alt[:-1])


@pytest.mark.parametrize("case", sorted(CASES))
def test_material_modes_match_reference_vectors(case):
    radial, tangential, mode = CASES[case]
    fx = np.load(GOLDEN)
    cfg, alt, mid = _grid()
    mesh = g.build_geodesic_mesh(cfg.subdivision)
    mat = g.EarthIonosphereMaterial(anomalies=_anomalies())
    got = sample_material(mat, mesh, alt, mid, cfg.earth_radius_m, radial_support=radial,
                          tangential_support=tangential, anomaly_mode=mode)
    for name in ("sigma_er", "epsilon_r_er", "sigma_et", "epsilon_r_et"):
        want = fx[f"{case}/{name}"]
        np.testing.assert_allclose(getattr(got, name), want, rtol=1e-14, atol=0, err_msg=name)
    for fam in ("er", "et"):
        key = f"{case}/fractions_{fam}"
        frac = getattr(got, f"fractions_{fam}")
        if key in fx.files:
            np.testing.assert_allclose(np.stack(frac), fx[key], rtol=1e-14, atol=1e-15)
        else:
            assert frac is None
    cfg_case = g.SimulationConfig(**GRID, radial_material_support=radial,
                                  tangential_material_support=tangential,
                                  horizontal_anomaly_mode=mode)
    dt = stable_time_step(cfg_case, mesh, mat)
    assert dt == pytest.approx(float(fx[f"{case}/time_step_s"]), rel=1e-14)


def test_quadrature_weights_tile_their_cells():
    mesh = g.build_geodesic_mesh(3)
    q = mesh.edge_diamond_quadrant_solid_angles()
    assert np.all(q > 0)
    np.testing.assert_allclose(q.sum(axis=1), mesh.edge_diamond_solid_angles(), rtol=1e-13)
    # the quarter diamonds also tile the sphere, like the dual cells
    assert q.sum() == pytest.approx(4 * np.pi, rel=1e-13)
    ends = mesh.edges.ravel()
    directions, areas = mesh.dual_cell_wedge_quadrature(ends, np.repeat(np.arange(mesh.n_edges), 2))
    assert np.all(areas > 0)
    np.testing.assert_allclose(np.bincount(ends, areas, mesh.n_vertices),
                               mesh.dual_cell_solid_angles, rtol=1e-13)
    np.testing.assert_allclose(np.linalg.norm(directions, axis=1), 1.0, rtol=1e-15)
    with pytest.raises(ValueError, match="incident"):
        mesh.dual_cell_wedge_quadrature(np.array([0]), np.array([int(np.nonzero(
            (mesh.edges != 0).all(axis=1))[0][0])]))


def test_uniform_material_is_unchanged_by_cell_averages():
    """A horizontally uniform material: the dual-cell and quarter-diamond
    averages return the point samples (the weights sum to one)."""
    cfg, alt, mid = _grid()
    mesh = g.build_geodesic_mesh(2)
    mat = g.EarthIonosphereMaterial()
    a = sample_material(mat, mesh, alt, mid, cfg.earth_radius_m)
    b = sample_material(mat, mesh, alt, mid, cfg.earth_radius_m, radial_support="dual-cell",
                        tangential_support="edge-diamond")
    for name in ("sigma_er", "epsilon_r_er", "sigma_et", "epsilon_r_et"):
        np.testing.assert_allclose(getattr(b, name), getattr(a, name), rtol=1e-14)


def test_conservative_rasterisation_preserves_the_area():
    mesh = g.build_geodesic_mesh(3)
    r_e = g.EARTH_RADIUS_M
    for anomaly in _anomalies():
        for points, areas in ((mesh.vertices, mesh.dual_cell_solid_angles),
                              (mesh.edge_midpoints(), mesh.edge_diamond_solid_angles())):
            f = g.materials.conservative_anomaly_fractions(points, areas, anomaly, r_e)
            assert np.all((f >= 0) & (f <= 1))
            assert np.sum(f * areas) == pytest.approx(anomaly.cap_solid_angle(r_e), rel=1e-13)
            # filled nearest first: every support closer than a partly filled one is full
            d = anomaly.arc_distance(points)
            partial = np.nonzero((f > 0) & (f < 1))[0]
            assert len(partial) <= 1
            if len(partial):
                assert np.all(f[d < d[partial[0]]] == 1.0)
    huge = g.SphericalAnomaly(0.0, 0.0, 1e3, -1e3, 0.0, 2.0, target_area_m2=1e16)
    with pytest.raises(ValueError, match="exceeds"):
        g.materials.conservative_anomaly_fractions(mesh.vertices, mesh.dual_cell_solid_angles,
                                                   huge, r_e)


def test_conservative_mode_refusals():
    cfg, alt, mid = _grid()
    mesh = g.build_geodesic_mesh(2)

    class NoAnomalies:
        def sample(self, directions, altitudes_m, earth_radius_m):
            shape = (len(directions), len(altitudes_m))
            return np.full(shape, 1e-3), np.ones(shape)

    with pytest.raises(ValueError, match="anomalies collection"):
        sample_material(NoAnomalies(), mesh, alt, mid, cfg.earth_radius_m,
                        anomaly_mode="conservative-nearest")

    class Plain(NoAnomalies):
        anomalies = _anomalies()

    with pytest.raises(ValueError, match="dataclass material"):
        sample_material(Plain(), mesh, alt, mid, cfg.earth_radius_m,
                        anomaly_mode="conservative-nearest")

    @dataclass(frozen=True)
    class Empty(NoAnomalies):
        anomalies: tuple = ()

    # no anomalies: the mode changes nothing
    a = sample_material(Empty(), mesh, alt, mid, cfg.earth_radius_m,
                        anomaly_mode="conservative-nearest")
    assert a.fractions_er is None
    with pytest.raises(ValueError, match="conductivity_factor"):
        g.SphericalAnomaly(0.0, 0.0, 1e5, -1e3, 0.0, 0.0)
    with pytest.raises(ValueError, match="reversed"):
        g.SphericalAnomaly(0.0, 0.0, 1e5, 1e3, 0.0, 2.0)


def test_solver_takes_the_modes():
    """GeodesicFDTD accepts every mode (no device work happens before the
    backend is created, so the check stops at the backend on a CPU host)."""
    cfg = g.SimulationConfig(**GRID, radial_material_support="dual-cell",
                             tangential_material_support="edge-diamond",
                             horizontal_anomaly_mode="conservative-nearest")
    mat = g.EarthIonosphereMaterial(anomalies=_anomalies())
    try:
        sim = g.GeodesicFDTD(cfg, material=mat, backend="hip")
    except g.BackendUnavailableError:
        return
    fx = np.load(GOLDEN)
    np.testing.assert_allclose(sim.sigma_er, fx["all/sigma_er"], rtol=1e-14)
    assert len(sim.anomaly_horizontal_fractions_er) == 2
