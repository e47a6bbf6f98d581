"""Ensembles of independent runs (uneven radial grids, own time steps,
coefficients, plasma profiles, ground conductivity, sources and receivers)
advanced by shared launches.

Bars: every run of an ensemble is bitwise identical to the same run stepped
alone (same strip arithmetic, gather-only kernels), for every column-block
shape; the FP64 ensemble matches the FP64 oracle within 1e-10 (family
relative L2, as the single-run parity tests); the autotuner leaves every
run's state bitwise unchanged.
"""

from __future__ import annotations

import numpy as np
import pytest

import scenarios as S

pytestmark = pytest.mark.gpu

FIELDS = ("er", "et", "hr", "ht")


def _scenarios(n=5, radial=(12, 16, 24)):
    scs = S.c2_runs(n_runs=n, subdivision=2, radial_choices=radial, n_receivers=4)
    for sc in scs:  # pulsed so that a short run sees the source
        sc["source"] = dict(sc["source"], one_over_e_half_width_s="30dt", center_time_s="90dt")
    return scs


def _build(scs, dtype, **kw):
    return [S.build_sim(sc, dtype=dtype, **kw) for sc in scs]


def _receivers(sims, scs):
    return [S.receiver_tables(sim.mesh, sim.altitudes_m, S.receiver_points(sc))
            for sim, sc in zip(sims, scs)]


def _state(sim):
    out = {k: sim.to_numpy(getattr(sim, f"_{k}_buf")).copy() for k in FIELDS}
    if sim._si_buf is not None:
        out["si"] = sim.to_numpy(sim._si_buf).copy()
    if sim._plasma_buf is not None:
        out["plasma"] = sim.to_numpy(sim._plasma_buf).copy()
    return out


@pytest.mark.parametrize("strips,order", [(1, 0), (2, 0), (4, 0), (1, 1), (2, 1), (4, 1),
                                          (1, 2), (1, 3)])
def test_ensemble_runs_equal_runs_alone_bitwise(strips, order):
    import ionosphere_fdtd_amd as g

    scs = _scenarios()
    solo = _build(scs, "float32", compile_step=True, compile_chunk_size=50)
    rx = _receivers(solo, scs)
    solo_traces = [s.record_er_observations(*r, 130) for s, r in zip(solo, rx)]
    runs = _build(scs, "float32")
    ens = g.EnsembleFDTD(runs, compile_chunk_size=50, autotune=False)
    ens.configure(strips, order)
    traces = ens.record_er_observations(rx, 130)
    print("radial", [len(r.radial_midpoints_m) for r in runs],
          "dt", [f"{r.time_step_s:.3e}" for r in runs], flush=True)
    assert traces.shape == (len(scs), 131, 4)
    for i, (a, b) in enumerate(zip(runs, solo)):
        np.testing.assert_array_equal(traces[i], solo_traces[i])
        sa, sb = _state(a), _state(b)
        for k in sa:
            np.testing.assert_array_equal(sa[k], sb[k], err_msg=f"run {i} {k}")
        assert a.steps == b.steps == 130
    assert np.any(traces != 0)


@pytest.mark.parametrize("strips,radial", [(1, (80, 96)), (2, (80, 96)), (1, (24, 96))])
def test_ensemble_of_flat_mapped_runs_equals_runs_alone_bitwise(strips, radial):
    """Runs of 80 and 96 radial cells take the flat thread mapping, runs of 24
    whole rows; with one strip per thread the ensemble keeps each run's
    mapping (mixed in one launch for 24 and 96), with two it runs whole rows."""
    import ionosphere_fdtd_amd as g

    scs = _scenarios(n=4, radial=radial)
    solo = _build(scs, "float32", compile_step=True, compile_chunk_size=20)
    for sc, s in zip(scs, solo):
        assert s._layout.e_thread_mapping == (2 if sc["radial_cells"] >= 80 else 1)
    rx = _receivers(solo, scs)
    solo_traces = [s.record_er_observations(*r, 60) for s, r in zip(solo, rx)]
    runs = _build(scs, "float32")
    ens = g.EnsembleFDTD(runs, compile_chunk_size=20, autotune=False)
    ens.configure(strips, 0)
    traces = ens.record_er_observations(rx, 60)
    for i, (a, b) in enumerate(zip(runs, solo)):
        np.testing.assert_array_equal(traces[i], solo_traces[i])
        sa, sb = _state(a), _state(b)
        for k in sa:
            np.testing.assert_array_equal(sa[k], sb[k], err_msg=f"run {i} {k}")
    assert np.any(traces != 0)


def test_ensemble_fp64_matches_oracle():
    import ionosphere_fdtd_amd as g

    scs = _scenarios(n=3, radial=(12, 16, 20))
    runs = _build(scs, "float64")
    rx = _receivers(runs, scs)
    ens = g.EnsembleFDTD(runs, compile_chunk_size=40, autotune=False)
    ens.configure(2, 0)
    traces = ens.record_er_observations(rx, 120)
    for i, (sc, run) in enumerate(zip(scs, runs)):
        ref = S.build_oracle(sc)
        assert ref.time_step_s == run.time_step_s
        want = ref.record_er_observations(*rx[i], 120)
        err = S.rel_l2(traces[i], want)
        got = {k: run.to_numpy(getattr(run, k)) for k in FIELDS}
        fam = S.family_errors(got, {k: getattr(ref, k) for k in FIELDS})
        print(sc["name"], f"trace {err:.2e}", {k: f"{v:.1e}" for k, v in fam.items()},
              flush=True)
        assert err <= 1e-10
        assert max(fam.values()) <= 1e-10


def test_autotuner_restores_state_and_picks_a_shape():
    import ionosphere_fdtd_amd as g

    scs = _scenarios(n=4)
    runs = _build(scs, "float32")
    ens = g.EnsembleFDTD(runs, compile_chunk_size=25, autotune=False)
    ens.step(60)
    before = [_state(r) for r in runs]
    shape = ens.tune(5)
    print("tuned", shape, flush=True)
    assert shape["strips_per_thread"] in (1, 2, 4) and shape["block_order"] in (0, 1, 2, 3)
    assert shape["strips_per_thread"] == 1 or shape["block_order"] < 2
    assert shape["ms_per_step"] > 0
    for b, r in zip(before, runs):
        a = _state(r)
        for k in b:
            np.testing.assert_array_equal(a[k], b[k])
    # stepping after tuning continues the same trajectory as without tuning
    ref = _build(scs, "float32")
    g.EnsembleFDTD(ref, autotune=False).step(80)
    ens.step(20)
    for a, b in zip(runs, ref):
        np.testing.assert_array_equal(_state(a)["er"], _state(b)["er"])


def test_ensemble_rejects_mixed_runs():
    import ionosphere_fdtd_amd as g

    scs = _scenarios(n=2)
    with pytest.raises(ValueError, match="dtype"):
        g.EnsembleFDTD([S.build_sim(scs[0], dtype="float32"),
                        S.build_sim(scs[1], dtype="float64")], autotune=False)
    no_plasma = dict(scs[1], plasma=None)
    with pytest.raises(ValueError, match="plasma"):
        g.EnsembleFDTD([S.build_sim(scs[0], dtype="float32"),
                        S.build_sim(no_plasma, dtype="float32")], autotune=False)
    runs = _build(scs, "float32")
    ens = g.EnsembleFDTD(runs, autotune=False)
    rx = _receivers(runs, scs)
    short = (rx[1][0][:2], rx[1][1][:2], rx[1][2][:2])
    with pytest.raises(ValueError, match="receivers"):
        ens.record_er_observations([rx[0], short], 5)
    with pytest.raises(ValueError, match="one strip per thread"):
        ens.configure(2, 2)
    with pytest.raises(ValueError, match="policy"):
        ens.configure(1, 4)
    # the same run twice: refused by the Python layer and by the C ABI itself
    with pytest.raises(ValueError, match="only once"):
        g.EnsembleFDTD([runs[0], runs[1], runs[0]], autotune=False)
    import ctypes as C

    from ionosphere_fdtd_amd import _native

    lib = _native.load_library()
    plans = (C.c_void_p * 2)(runs[0]._plan, runs[0]._plan)
    handle = C.c_void_p()
    assert lib.geo_ensemble_create(plans, 2, C.byref(handle)) == -1
    assert b"only once" in lib.geo_last_error()
