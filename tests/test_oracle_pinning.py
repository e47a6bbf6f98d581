"""Pin the CPU oracle before trusting it (CPU only).

Every check here can fail for a plausible mistake in the oracle:
  * reference fixtures (tests/golden/, written by make_fixtures.py from the
    reference itself): integer mesh arrays bit-exact, full FP64 trajectories
    of the scenarios (SI ground, Drude and magnetised plasma, vertical and
    tangential sources, thin-shell trapezoid, random-field PEC vacuum);
  * the reference's published golden data (artifacts/analytic-solutions/
    full-field.csv): CFL time steps and measured cavity frequencies;
  * mathematics: discrete adjoint identities, the exact leapfrog recurrence of
    TM/TE surface eigenmodes, exact stiff exponential decay, second-order
    convergence of exponential forcing, passivity, and discrete div-B;
  * planted mistakes in the oracle (sign flips, dropped terms, swapped
    operands) must make the fixture comparison fail.
"""

from __future__ import annotations

import csv
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

import cavity_modes as CM
import scenarios as S

GOLDEN = Path(__file__).resolve().parent / "golden"
o = S.oracle_module()
INT_FIELDS = ("faces", "edges", "face_edges", "face_edge_signs", "edge_left_faces",
              "edge_right_faces", "vertex_degree")
FLOAT_FIELDS = ("vertices", "face_centers", "primal_edge_angles", "dual_edge_angles",
                "face_solid_angles", "dual_cell_solid_angles")
TRAJECTORIES = ("c0", "c0_pulse", "c0_plasma", "magnetized", "thin_trapezoid", "tangential",
                "pec_vacuum", "plasma_random", "magnetized_random", "c3_seeded")


def _sha(values):
    a = np.asarray(values)
    dt = "<f8" if np.issubdtype(a.dtype, np.floating) else "<i8"
    return hashlib.sha256(np.asarray(a, dtype=dt, order="C").tobytes()).hexdigest()


# ----------------------------------------------------------------------- mesh
@pytest.mark.parametrize("s", range(4))
@pytest.mark.parametrize("orientation", ("polar", "native"))
def test_oracle_mesh_integer_arrays_equal_reference(s, orientation):
    fx = np.load(GOLDEN / "mesh_s0_s3.npz")
    mesh = o.build_mesh(s, orientation)
    for f in INT_FIELDS:
        assert np.array_equal(getattr(mesh, f), fx[f"s{s}_{orientation}_{f}"]), f


@pytest.mark.parametrize("s", range(5))
def test_oracle_mesh_hashes_equal_reference(s):
    hashes = json.loads((GOLDEN / "mesh_hashes.json").read_text())[f"s{s}_polar"]
    mesh = o.build_mesh(s, "polar")
    for f in INT_FIELDS + FLOAT_FIELDS:
        assert _sha(getattr(mesh, f)) == hashes[f], f


def test_oracle_mesh_counts_and_closure():
    for s in range(4):
        m = o.build_mesh(s)
        assert (m.n_vertices, m.n_edges, m.n_faces) == (10 * 4**s + 2, 30 * 4**s, 20 * 4**s)
        assert np.count_nonzero(m.vertex_degree == 5) == 12
        assert m.face_solid_angles.sum() == pytest.approx(4 * np.pi, rel=1e-12)
        assert m.dual_cell_solid_angles.sum() == pytest.approx(4 * np.pi, rel=1e-11)
        grad = m.edge_difference(np.arange(m.n_vertices, dtype=float))
        assert np.allclose(m.face_circulation(grad), 0.0)  # boundary of boundary


# ------------------------------------------------------------ CFL golden dts
def _full_field_rows():
    with open(GOLDEN / "full_field.csv") as fh:
        return list(csv.DictReader(fh))


@pytest.mark.parametrize("row", _full_field_rows(), ids=lambda r: f"{r['case']}-{r['polarization']}-s{r['subdivision']}-r{r['radial_cells']}-{r['refinement']}")
def test_cfl_time_step_matches_reference_artifact(row):
    s, r = int(row["subdivision"]), int(row["radial_cells"])
    sim = o.OracleSim(subdivision=s, radial_cells=r, minimum_altitude_m=0.0,
                      maximum_altitude_m=100e3, earth_radius_m=o.EARTH_RADIUS_M,
                      courant_factor=0.4, material_arrays=_vacuum(s, r))
    assert sim.time_step_s == pytest.approx(float(row["time_step_s"]), rel=1e-14)


def _vacuum(s, r):
    V, E = 10 * 4**s + 2, 30 * 4**s
    return (np.zeros((V, r + 1)), np.ones((V, r + 1)), np.zeros((E, r)), np.ones((E, r)))


def _cavity_run(s, r, polarization, radial_index, steps=None, periods=None):
    sim = o.OracleSim(subdivision=s, radial_cells=r, minimum_altitude_m=0.0,
                      maximum_altitude_m=100e3, earth_radius_m=o.EARTH_RADIUS_M,
                      courant_factor=0.4, material_arrays=_vacuum(s, r))
    mode = CM.electric_mode(sim.mesh, sim.radii_m, sim.radial_midpoints_m,
                            sim.radial_node_control_lengths_m, sim.radial_steps_m, 1,
                            polarization, radial_index)
    sim.er[:] = mode["er"]
    sim.et[:] = mode["et"]
    analytic = mode["k"] * o.C_0 / (2 * np.pi)
    if periods is not None:
        steps = int(np.ceil(periods / (analytic * sim.time_step_s)))
    return CM.measure_frequency(sim, mode, steps), analytic


def test_a2_tm_cavity_frequency_matches_reference_artifact():
    """A2 horizontal TM, s1 x 8, 2000 steps: measured 10.305426732600361 Hz."""
    row = next(r for r in _full_field_rows() if r["case"] == "A2" and r["subdivision"] == "1")
    measured, analytic = _cavity_run(1, 8, "TM", 0, steps=2000)
    assert analytic == pytest.approx(float(row["analytic_frequency_hz"]), rel=1e-12)
    # The artifact was written by another reference revision / NumPy build; the
    # recurrence fit reproduces it to ~4e-10.  An operator mistake moves it ~1e-2.
    assert measured == pytest.approx(float(row["measured_frequency_hz"]), rel=1e-8)


@pytest.mark.parametrize("polarization,index", (("TE", 0), ("TM", 1)))
def test_a4_radial_cavity_frequency_matches_reference_artifact(polarization, index):
    row = next(r for r in _full_field_rows() if r["case"] == "A4" and r["refinement"] ==
               "radial" and r["polarization"] == polarization and r["radial_cells"] == "8")
    measured, analytic = _cavity_run(2, 8, polarization, index, periods=5.0)
    assert analytic == pytest.approx(float(row["analytic_frequency_hz"]), rel=1e-12)
    assert measured == pytest.approx(float(row["measured_frequency_hz"]), rel=1e-8)


# ---------------------------------------------------------------- invariants
def test_discrete_adjoint_identities():
    """test_maxwell_invariants.py:76-89 restated on the oracle operators."""
    m = o.build_mesh(2)
    g = np.random.default_rng(20260805)
    ev, fv, vv = g.standard_normal(m.n_edges), g.standard_normal(m.n_faces), g.standard_normal(
        m.n_vertices)
    assert fv @ m.face_circulation(ev) == pytest.approx(m.dual_edge_difference(fv) @ ev,
                                                         abs=2e-13)
    assert vv @ m.dual_cell_circulation(ev) == pytest.approx(-m.edge_difference(vv) @ ev,
                                                              abs=2e-13)


def _surface_mode_sim():
    return o.OracleSim(subdivision=1, radial_cells=2, minimum_altitude_m=-1.0e9,
                       maximum_altitude_m=1.0e9, earth_radius_m=1.0e10, courant_factor=0.8,
                       material_arrays=_vacuum(1, 2))


def _operator(n, apply):
    eye = np.eye(n)
    return np.column_stack([apply(eye[:, i]) for i in range(n)])


def test_tm_surface_eigenmode_follows_exact_leapfrog_recurrence():
    sim = _surface_mode_sim()
    m = sim.mesh

    def lap(v):
        flux = m.dual_edge_angles / m.primal_edge_angles * m.edge_difference(v)
        return m.dual_cell_circulation(flux) / m.dual_cell_solid_angles

    w, vecs = np.linalg.eig(_operator(m.n_vertices, lap))
    i = int(np.argmin(np.abs(w.real + 2.0)))
    lam, mode = -float(w[i].real), vecs[:, i].real
    sim.er[:, 1] = mode
    amps = [1.0]
    for _ in range(100):
        sim.update_magnetic_fields()
        sim.update_electric_fields()
        sim.et[:] = 0.0
        sim.hr[:] = 0.0
        amps.append(float(sim.er[:, 1] @ mode / (mode @ mode)))
    q = (o.C_0 * sim.time_step_s) ** 2 * lam / sim.radii_m[1] ** 2
    ph = np.arccos(1.0 - 0.5 * q)
    n = np.arange(len(amps))
    np.testing.assert_allclose(amps, np.cos((n + 0.5) * ph) / np.cos(0.5 * ph), atol=8e-14)


def test_te_surface_eigenmode_follows_exact_leapfrog_recurrence():
    sim = _surface_mode_sim()
    m = sim.mesh

    def te(v):
        grad = m.primal_edge_angles / m.dual_edge_angles * m.dual_edge_difference(v)
        return m.face_circulation(grad) / m.face_solid_angles

    w, vecs = np.linalg.eig(_operator(m.n_faces, te))
    i = int(np.argmin(np.abs(w.real - 2.0)))
    lam, mode = float(w[i].real), vecs[:, i].real
    sim.hr[:, 1] = mode
    amps = [1.0]
    for _ in range(100):
        sim.update_magnetic_fields()
        sim.ht[:] = 0.0
        sim.update_electric_fields()
        sim.er[:] = 0.0
        amps.append(float(sim.hr[:, 1] @ mode / (mode @ mode)))
    q = (o.C_0 * sim.time_step_s) ** 2 * lam / sim.radial_midpoints_m[1] ** 2
    ph = np.arccos(1.0 - 0.5 * q)
    n = np.arange(len(amps))
    np.testing.assert_allclose(amps, np.cos((n - 0.5) * ph) / np.cos(0.5 * ph), atol=8e-14)


def _uniform_sigma(s, r, sigma):
    V, E = 10 * 4**s + 2, 30 * 4**s
    return (np.full((V, r + 1), sigma), np.ones((V, r + 1)), np.full((E, r), sigma),
            np.ones((E, r)))


def test_exponential_loss_is_exact_stiff_decay():
    sim = o.OracleSim(subdivision=0, radial_cells=2, courant_factor=0.2,
                      material_arrays=_uniform_sigma(0, 2, 1.0))
    sim.er.fill(1.0)
    sim.step()
    np.testing.assert_allclose(sim.er, np.exp(-1.0 * sim.time_step_s / o.EPSILON_0), atol=1e-300)
    assert np.all(sim.er >= 0.0)


def test_exponential_forcing_converges_at_second_order():
    """test_maxwell_invariants.py:237-275: error ratio 4 within 0.3 %."""
    base = o.OracleSim(subdivision=0, radial_cells=2, courant_factor=1.0,
                       material_arrays=_uniform_sigma(0, 2, 0.0))
    coarse = 0.4 * base.cfl_time_step_limit_s
    duration = 20 * coarse
    sigma = 0.5 * o.EPSILON_0 / duration
    errors = []
    for refine in (1, 2):
        dt = coarse / refine
        sim = o.OracleSim(subdivision=0, radial_cells=2, courant_factor=1.0, time_step_s=dt,
                          material_arrays=_uniform_sigma(0, 2, sigma))
        a, b = float(sim._ca_er[0, 0]), float(sim._cb_er[0, 0])
        f = 0.0
        for n in range(20 * refine):
            f = a * f + b * (n + 0.5) * dt
        exact = duration / sigma - o.EPSILON_0 / sigma**2 * (1 - np.exp(-sigma * duration
                                                                        / o.EPSILON_0))
        errors.append(abs(f - exact))
    assert errors[0] / errors[1] == pytest.approx(4.0, rel=3e-3)


def test_conductive_update_is_passive():
    sim = o.OracleSim(subdivision=0, radial_cells=2, courant_factor=0.2,
                      material_arrays=_uniform_sigma(0, 2, 1.0))
    sim.er.fill(1.0)
    norms = [np.linalg.norm(sim.er)]
    for _ in range(20):
        sim.step()
        norms.append(np.linalg.norm(sim.er))
    assert np.all(np.diff(norms) <= 0.0)


def test_discrete_div_b_stays_at_roundoff():
    """Derived invariant: the H update conserves the prism-cell divergence of B
    exactly (telescoping radial stencil, boundary of boundary = 0)."""
    sc = S.scenario("c0_pulse", surface_sigma=None)
    sim = S.build_oracle(sc)
    g = np.random.default_rng(7)
    sim.er[:] = g.standard_normal(sim.er.shape)
    sim.et[:] = g.standard_normal(sim.et.shape)
    sim.step(200)
    D, scale = o.divergence_b(sim.mesh, sim.radii_m, sim.radial_midpoints_m,
                              sim.radial_node_control_lengths_m, sim.hr, sim.ht)
    assert np.max(np.abs(D)) <= 1e-13 * np.max(scale)
    assert np.max(scale) > 0


# ------------------------------------------------------- graded radial grids
GRADED = (-10_000.0, -6_000.0, -2_000.0, -1_000.0, 0.0, 4_000.0)


def _graded_sim(geometry_mode, altitudes=GRADED, **kw):
    return o.OracleSim(subdivision=0, radial_cells=len(altitudes) - 1,
                       minimum_altitude_m=altitudes[0], maximum_altitude_m=altitudes[-1],
                       radial_altitudes_m=altitudes, geometry_mode=geometry_mode,
                       material_arrays=_vacuum_grid(0, len(altitudes) - 1), **kw)


def _vacuum_grid(s, r):
    return _vacuum(s, r)


def _swap_center_distance(sim):
    """Planted mistake: the node-centred derivative divides by the cell step
    above the node instead of the midpoint distance D_k = rbar_k - rbar_{k-1}
    (the two agree on a uniform grid only)."""
    sim._center_dist = sim._steps[1:].copy()


def _radial_adjoint(sim, weights_ht, weights_et):
    g = np.random.default_rng(20260805)
    sim.et[:] = g.standard_normal(sim.et.shape)
    sim.ht[:] = g.standard_normal(sim.ht.shape)
    left = np.sum(sim.ht * sim.radial_derivative_et() * weights_ht[None, :])
    right = np.sum(sim.et * sim.radial_derivative_ht() * weights_et[None, :])
    return left, right


@pytest.mark.parametrize("plant", [None, _swap_center_distance])
def test_graded_radial_stencils_weighted_adjoint_thin_shell(plant):
    """test_maxwell_invariants.py:344-374 restated: on a nonuniform grid the
    thin-shell node and midpoint derivatives are negative adjoints under the
    control-length weights L_k and the steps dr_j; the planted D_k <-> dr_k
    swap breaks the identity by O(1)."""
    sim = _graded_sim("thin-shell")
    if plant:
        plant(sim)
    left, right = _radial_adjoint(sim, sim.radial_node_control_lengths_m, sim.radial_steps_m)
    if plant:
        assert abs(left + right) > 1e-2 * max(abs(left), abs(right))
    else:
        assert left + right == pytest.approx(0.0, abs=2.0e-11)


@pytest.mark.parametrize("plant", [None, _swap_center_distance])
def test_graded_radial_stencils_weighted_adjoint_full_spherical(plant):
    """test_maxwell_invariants.py:399-429 restated: full-spherical weights
    r_k^2 L_k at nodes and rbar_j^2 dr_j at midpoints."""
    sim = _graded_sim("full-spherical")
    if plant:
        plant(sim)
    left, right = _radial_adjoint(sim, sim.radial_node_control_lengths_m * sim.radii_m ** 2,
                                  sim.radial_steps_m * sim.radial_midpoints_m ** 2)
    if plant:
        assert abs(left + right) > 1e-2 * max(abs(left), abs(right))
    else:
        assert abs(left + right) <= 2.0e-15 * max(abs(left), abs(right))


def test_graded_radial_derivative_annihilates_constants():
    """test_maxwell_invariants.py:377-396 restated (thin shell, constant Ht)."""
    sim = _graded_sim("thin-shell")
    sim.ht.fill(1.0)
    assert np.array_equal(sim.radial_derivative_ht(), np.zeros_like(sim.et))


@pytest.mark.parametrize("altitudes", [None, (-100e3, -60e3, -40e3, -20e3, -10e3, 0.0, 10e3,
                                              20e3, 40e3, 100e3)])
def test_full_spherical_metric_annihilates_inverse_radius(altitudes):
    """test_maxwell_invariants.py:432-460 restated, on the uniform grid of the
    reference's test and on a graded one: r*Ht = 1 and rbar*Et = 1 have zero
    interior radial derivative; the thin-shell metric does not annihilate 1/r."""
    kw = dict(subdivision=0, radial_cells=8, minimum_altitude_m=-100e3,
              maximum_altitude_m=100e3, material_arrays=_vacuum(0, 8))
    if altitudes is not None:
        kw.update(radial_cells=len(altitudes) - 1, radial_altitudes_m=altitudes,
                  material_arrays=_vacuum(0, len(altitudes) - 1))
    sim = o.OracleSim(geometry_mode="full-spherical", **kw)
    sim.et[:] = 1.0 / sim.radial_midpoints_m[None, :]
    sim.ht[:] = 1.0 / sim.radii_m[None, :]
    # rbar * (1 / rbar) is 1 to an ulp, divided by the smallest spacing and r:
    # 1e-27 at the reference's 25 km spacing, 1e-26 with 10 km cells; a wrong
    # metric leaves ~1/r^2 = 2.5e-14
    atol = 1.0e-27 if altitudes is None else 1.0e-26
    np.testing.assert_allclose(sim.radial_derivative_et()[:, 1:-1], 0.0, atol=atol)
    np.testing.assert_allclose(sim.radial_derivative_ht(), 0.0, atol=atol)
    thin = o.OracleSim(geometry_mode="thin-shell", **kw)
    thin.et[:] = 1.0 / thin.radial_midpoints_m[None, :]
    assert np.max(np.abs(thin.radial_derivative_et()[:, 1:-1])) > 2.0e-14


def test_graded_grid_stays_bounded_at_the_cfl_limit():
    """test_maxwell_invariants.py:463-496 restated: courant factor 1 on the
    abrupt graded grid, 1000 steps from random E, max |E| below twice its start."""
    sim = _graded_sim("full-spherical", courant_factor=1.0)
    g = np.random.default_rng(20260805)
    sim.er[:] = g.standard_normal(sim.er.shape)
    sim.et[:] = g.standard_normal(sim.et.shape)
    start = max(np.max(np.abs(sim.er)), np.max(np.abs(sim.et)))
    worst = start
    for _ in range(1000):
        sim.step()
        worst = max(worst, np.max(np.abs(sim.er)), np.max(np.abs(sim.et)))
    assert np.isfinite(worst) and worst < 2.0 * start


# ----------------------------------------------------- reference trajectories
def _compare_to_fixture(sim, fx, bar):
    errs = {}
    for k in ("er", "et", "hr", "ht"):
        errs[k] = S.rel_l2(getattr(sim, k), fx[k])
    return errs


def _oracle_trajectory(name, plant=None):
    sc = S.scenario(name)
    fx = np.load(GOLDEN / f"trajectory_{name}.npz")
    sim = S.build_oracle(sc)
    if plant:
        plant(sim)
    if sc["seed_fields"]:
        S.apply_seed_fields(sim, lambda k, v: getattr(sim, k).__setitem__(slice(None), v))
    traces = sim.record_er_observations(fx["receiver_vertices"], fx["receiver_layers"],
                                        fx["receiver_weights"], int(fx["steps"]))
    return sim, traces, fx


@pytest.mark.parametrize("name", TRAJECTORIES)
def test_oracle_trajectory_equals_reference(name):
    sim, traces, fx = _oracle_trajectory(name)
    assert sim.time_step_s == float(fx["time_step_s"])
    errs = _compare_to_fixture(sim, fx, 1e-12)
    errs["traces"] = S.rel_l2(traces, fx["traces"])
    if "si_memory" in fx:
        errs["si"] = S.rel_l2(sim.si_memory, fx["si_memory"])
    if "plasma_current_0" in fx:
        for i, j in enumerate(sim.coupler.current_density):
            errs[f"j{i}"] = S.rel_l2(j, fx[f"plasma_current_{i}"])
    bad = {k: v for k, v in errs.items() if not v <= 1e-12}
    assert not bad, bad


def _flip_radial_sign(sim):
    orig = sim.radial_derivative_et
    sim.radial_derivative_et = lambda: -orig()


def _drop_si_history(sim):
    sim.si_history_weights[:] = 0.0


def _swap_plasma_sides(sim):
    c = sim.coupler
    c.left_tangents, c.right_tangents = c.right_tangents, -c.left_tangents


def _drop_vertex_average(sim):
    sim.coupler.vertex_face_weights[:, -1] = 0.0


def _transpose_source_layers(sim):
    v, l, w = sim.source_distribution
    sim.source_distribution = (v, l[::-1].copy(), w)


@pytest.mark.parametrize("name,plant", [
    ("c0_pulse", _flip_radial_sign), ("c0_pulse", _drop_si_history),
    ("plasma_random", _swap_plasma_sides), ("magnetized_random", _swap_plasma_sides),
    ("thin_trapezoid", _transpose_source_layers), ("plasma_random", _drop_vertex_average),
    ("c3_seeded", _swap_center_distance),
])
def test_planted_oracle_mistakes_are_caught(name, plant):
    sim, traces, fx = _oracle_trajectory(name, plant)
    errs = _compare_to_fixture(sim, fx, 1e-12)
    errs["traces"] = S.rel_l2(traces, fx["traces"])
    assert max(errs.values()) > 1e-6, errs


# ------------------------------------------------------ torch-CPU restatement
@pytest.mark.parametrize("name,steps", [("c0_plasma", 60), ("magnetized", 40),
                                        ("thin_trapezoid", 40), ("c3_seeded", 30)])
def test_torch_cpu_leg_equals_numpy_oracle(name, steps):
    """The CPU baseline's torch leg (oracle/torch_cpu.py) is the oracle's step
    on torch CPU tensors: FP64 fields, SI memory and plasma current within
    1e-12 of the NumPy oracle (SI ground, Drude and magnetised plasma,
    vertical source, thin shell, graded radii)."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("torch_cpu", S.ORACLE_DIR / "torch_cpu.py")
    tc = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(tc)
    sc = S.scenario(name)
    ref = S.build_oracle(sc)
    if sc["seed_fields"]:
        S.apply_seed_fields(ref, lambda k, v: getattr(ref, k).__setitem__(slice(None), v))
    sim = tc.TorchCpuSim(ref, threads=2)
    sim.step(steps)
    ref.step(steps)
    errs = {k: S.rel_l2(getattr(sim, k).numpy(), getattr(ref, k)) for k in ("er", "et", "hr",
                                                                             "ht")}
    if ref.surface is not None:
        errs["si"] = S.rel_l2(sim.si_memory.numpy(), ref.si_memory)
    if ref.coupler is not None:
        errs["j"] = S.rel_l2(sim.current_density[0].numpy(), ref.coupler.current_density[0])
    bad = {k: v for k, v in errs.items() if not v <= 1e-12}
    assert not bad, bad


# ---------------------------------------------------------------------------
# Trace surrogate oracle (no reference counterpart): pinned by finite
# differences and by the adjoint / convolution identities of its layout ops.
# ---------------------------------------------------------------------------
def _surrogate_oracle():
    import importlib.util

    spec = importlib.util.spec_from_file_location(
        "surrogate_oracle", S.ORACLE_DIR / "surrogate_oracle.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_surrogate_im2col_is_upsampled_convolution_and_col2im_its_adjoint():
    so = _surrogate_oracle()
    rng = np.random.default_rng(5)
    x = rng.standard_normal((2, 6, 3))
    cols = so.im2col_up(x, 2, 5)
    assert cols.shape == (2 * 12, 15)
    # one output channel with weights w equals np.convolve on the upsampled signal
    w = rng.standard_normal((5, 3))
    out = (cols @ w.reshape(15, 1)).reshape(2, 12)
    up = np.repeat(x, 2, axis=1)
    for b in range(2):
        want = sum(np.convolve(up[b, :, c], w[::-1, c], mode="same") for c in range(3))
        np.testing.assert_allclose(out[b], want, rtol=1e-12, atol=1e-12)
    y = rng.standard_normal(cols.shape)
    lhs = float(np.sum(cols * y))
    rhs = float(np.sum(x * so.col2im_up(y, 2, 6, 3, 2, 5)))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_surrogate_oracle_gradients_match_finite_differences():
    so = _surrogate_oracle()
    from ionosphere_fdtd_amd.surrogate import SurrogateConfig, initial_parameters

    cfg = SurrogateConfig(in_features=5, width=7, depth=2, seq0=4, channels=(3, 2, 2), ksize=3,
                          amp_weight=0.7)
    rng = np.random.default_rng(11)
    params = {k: v.astype(np.float64) for k, v in initial_parameters(cfg, seed=3).items()}
    for k in params:
        if not k.startswith("W"):
            params[k] = 0.1 * rng.standard_normal(params[k].shape)
    x = rng.standard_normal((3, 5))
    st = rng.standard_normal((3, cfg.length))
    at = rng.standard_normal(3)

    def loss_of(p):
        shape, amp, _ = so.forward(p, x, cfg, False)
        return so.trace_loss(shape, amp, st, at, cfg.amp_weight)[0]

    shape, amp, cache = so.forward(params, x, cfg, False)
    _, ds, da = so.trace_loss(shape, amp, st, at, cfg.amp_weight)
    grads = so.backward(cache, ds, da, cfg, False)
    assert set(grads) == set(params)
    eps = 1e-6
    for name, value in params.items():
        flat = value.reshape(-1)
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            keep = flat[idx]
            flat[idx] = keep + eps
            up = loss_of(params)
            flat[idx] = keep - eps
            down = loss_of(params)
            flat[idx] = keep
            fd = (up - down) / (2 * eps)
            g = grads[name].reshape(-1)[idx]
            assert abs(fd - g) <= 1e-6 * max(1.0, abs(fd)), (name, idx, fd, g)
