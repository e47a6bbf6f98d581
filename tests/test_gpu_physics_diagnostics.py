"""Device physics diagnostics (energy, conductive loss, profiles, RMS) against
independent volume integrals and an FP64 NumPy evaluation of the reference
sampler's formulas (verification/physics_diagnostics/model.py:242-275; test
model: verification/tests/test_physics_diagnostics.py:43-116)."""

from __future__ import annotations

import math

import numpy as np
import pytest

import scenarios as S

pytestmark = pytest.mark.gpu

EPSILON_0 = 8.8541878128e-12
MU_0 = 1.25663706212e-6


def _volumes(sim):
    mesh = sim.mesh
    node_radial = sim.radii_m ** 2 * sim.radial_node_control_lengths_m
    cell_radial = sim.radial_midpoints_m ** 2 * sim.radial_steps_m
    edge_h = mesh.primal_edge_angles * mesh.dual_edge_angles
    return {"er": mesh.dual_cell_solid_angles[:, None] * node_radial[None, :],
            "et": edge_h[:, None] * cell_radial[None, :],
            "hr": mesh.face_solid_angles[:, None] * cell_radial[None, :],
            "ht": edge_h[:, None] * node_radial[None, :]}


def test_sampler_matches_independent_volume_integrals():
    """Constant fields (as the reference's test does): every energy and loss
    total is a closed-form volume integral; the fields are left unchanged."""
    import ionosphere_fdtd_amd as g

    sc = S.scenario("c0")  # Earth-ionosphere conductivity, SI ground
    sim = S.build_sim(sc)
    values = {"er": 2.0, "et": 3.0, "hr": 4.0, "ht": 5.0}
    for name, v in values.items():
        setattr(sim, name, np.full(tuple(getattr(sim, name).shape), v))
    before = {name: sim.to_numpy(getattr(sim, name)).copy() for name in values}
    snap = g.PhysicsDiagnosticSampler(sim).sample({"A": 1.25})
    vol = _volumes(sim)
    want = {"er": 0.5 * np.sum(EPSILON_0 * sim.epsilon_r_er * 4.0 * vol["er"]),
            "et": 0.5 * np.sum(EPSILON_0 * sim.epsilon_r_et * 9.0 * vol["et"]),
            "hr": 0.5 * MU_0 * 16.0 * np.sum(vol["hr"]),
            "ht": 0.5 * MU_0 * 25.0 * np.sum(vol["ht"])}
    loss = (np.sum(sim.sigma_er * 4.0 * vol["er"]) + np.sum(sim.sigma_et * 9.0 * vol["et"]))
    for name in values:
        assert snap.scalars[f"energy/{name}_j"] == pytest.approx(want[name], rel=1e-12)
        assert snap.scalars[f"field_rms/{name}"] == pytest.approx(values[name], rel=1e-12)
        assert snap.scalars[f"field_max_abs/{name}"] == values[name]
        assert snap.scalars[f"field_finite/{name}"] == 1.0
    assert loss > 0
    assert snap.scalars["conductive_loss/total_w"] == pytest.approx(loss, rel=1e-12)
    assert snap.receiver_values == {"A": 1.25}
    for name in values:
        np.testing.assert_array_equal(sim.to_numpy(getattr(sim, name)), before[name])


@pytest.mark.parametrize("dtype,rtol", [("float64", 1e-12), ("float32", 1e-6)])
def test_profiles_match_fp64_evaluation_of_the_reference_formulas(dtype, rtol):
    """After driven steps with plasma and a surface-impedance ground, every
    per-layer energy and loss profile equals the reference sampler's formula
    evaluated in FP64 on the same (device) fields."""
    import ionosphere_fdtd_amd as g

    sc = S.scenario("plasma_random")
    sim = S.build_sim(sc, dtype=dtype, compile_step=True)
    S.apply_seed_fields(sim, lambda k, v: setattr(sim, k, v))
    sim.step(40)
    snap = g.PhysicsDiagnosticSampler(sim).sample()
    mesh = sim.mesh
    f = {name: sim.to_numpy(getattr(sim, name)).astype(np.float64)
         for name in ("er", "et", "hr", "ht")}
    node_radial = sim.radii_m ** 2 * sim.radial_node_control_lengths_m
    cell_radial = sim.radial_midpoints_m ** 2 * sim.radial_steps_m
    edge_h = (mesh.primal_edge_angles * mesh.dual_edge_angles)[:, None]
    want = {
        "er": 0.5 * (EPSILON_0 * sim.epsilon_r_er * f["er"] ** 2
                     * mesh.dual_cell_solid_angles[:, None]).sum(axis=0) * node_radial,
        "et": 0.5 * (EPSILON_0 * sim.epsilon_r_et * f["et"] ** 2 * edge_h).sum(axis=0)
        * cell_radial,
        "hr": 0.5 * MU_0 * (f["hr"] ** 2 * mesh.face_solid_angles[:, None]).sum(axis=0)
        * cell_radial,
        "ht": 0.5 * MU_0 * (f["ht"] ** 2 * edge_h).sum(axis=0) * node_radial,
    }
    want_loss = {
        "er": (sim.sigma_er * f["er"] ** 2 * mesh.dual_cell_solid_angles[:, None]).sum(axis=0)
        * node_radial,
        "et": (sim.sigma_et * f["et"] ** 2 * edge_h).sum(axis=0) * cell_radial,
    }
    for name in want:
        got = getattr(snap, f"radial_energy_{name}_j")
        scale = float(np.max(np.abs(want[name])))
        assert scale > 0
        np.testing.assert_allclose(got, want[name], rtol=0, atol=rtol * scale)
        rms = math.sqrt(float(np.mean(f[name] ** 2)))
        assert snap.scalars[f"field_rms/{name}"] == pytest.approx(rms, rel=rtol)
        assert snap.scalars[f"field_max_abs/{name}"] == pytest.approx(
            float(np.max(np.abs(f[name]))), rel=1e-15)
    for name in want_loss:
        got = getattr(snap, f"radial_loss_{name}_w")
        scale = float(np.max(np.abs(want_loss[name])))
        assert scale > 0
        np.testing.assert_allclose(got, want_loss[name], rtol=0, atol=rtol * scale)
    # the region split partitions the totals
    total = sum(snap.scalars[f"energy_region/{r}_j"] for r in ("earth", "atmosphere",
                                                              "ionosphere"))
    assert total == pytest.approx(snap.scalars["energy/total_staggered_j"], rel=1e-12)
    # the device totals agree with the profile sums
    d = sim.field_diagnostics()
    assert d["loss_er"] + d["loss_et"] == pytest.approx(
        snap.scalars["conductive_loss/total_w"], rel=1e-10)


def test_injected_nan_clears_the_finite_flag_of_that_field_only():
    """A NaN written into one Er value (the reference's failure-detection
    check): that field reports field_finite 0 and a NaN RMS, the device count
    sees one non-finite value, and every other field stays finite."""
    import ionosphere_fdtd_amd as g

    sc = S.scenario("c0")
    sim = S.build_sim(sc)
    rng = np.random.default_rng(3)
    for name in ("er", "et", "hr", "ht"):
        setattr(sim, name, rng.standard_normal(tuple(getattr(sim, name).shape)))
    sim.er[5, 3] = float("nan")
    snap = g.PhysicsDiagnosticSampler(sim).sample()
    assert snap.scalars["field_finite/er"] == 0.0
    assert math.isnan(snap.scalars["field_rms/er"])
    for name in ("et", "hr", "ht"):
        assert snap.scalars[f"field_finite/{name}"] == 1.0
        assert math.isfinite(snap.scalars[f"field_rms/{name}"])
    d = sim.field_diagnostics()
    assert d["nonfinite"] == 1 and d["nonfinite_er"] == 1
    assert not sim.diagnostics()["field_finite"]
