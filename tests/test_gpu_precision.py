"""FP16 / BF16 field storage (FP32 update arithmetic) on the MI355X.

Narrow storage rounds every stored field value once per step, to 11 (FP16)
or 8 (BF16) significant bits.  Rounding to nearest loses every update
smaller than half an ulp of the stored value, which freezes quasi-static
fields (the 20 Hz source with a ~0.1 us step: per-step changes of ~1e-5 of
the field).  The default rounding is stochastic with rotation thresholds: a
hashed per-element phase advanced by the golden ratio each step, so the
rounding error of a steady sub-ulp update stays O(log N) ulp instead of
growing with N.  The bars below bound the relative L2 error against the
FP64 oracle after the stated number of steps, each field measured against
the norm of its family (Er, Et) or (Hr, Ht), traces against their own norm:

  * wave-dominated runs (150-300 steps): FP16 with dynamic power-of-two
    scaling 2e-2, BF16 1.5e-1;
  * quasi-static runs (the 20 Hz source, 2000-10000 steps): FP16 1e-2, BF16
    5e-2, with round-to-nearest shown to be at least 4x worse (FP16) and
    O(1) (BF16) on the same runs.

Both exceed the FP32 bar (1e-5) by design; the storage mode is opt-in.

Dynamic scaling is exact (powers of two) and the rounding thresholds depend
only on (step, field, element), so the scaled problem is the same problem: a
run whose source is 2^-20 times smaller gives exactly 2^-20 times the
fields, bit for bit.  That invariant, graph/single-launch identity,
repeatability, resume equivalence and the host-side scale selection are
tested here as well.
"""

from __future__ import annotations

import numpy as np
import pytest

import scenarios as S

pytestmark = pytest.mark.gpu

FIELDS = ("er", "et", "hr", "ht")
BARS = {"float16": 2e-2, "bfloat16": 1.5e-1}
CASES = [("c0_pulse", 300), ("c0_plasma", 300), ("plasma_random", 200), ("pec_vacuum", 200),
         ("tangential", 150), ("c3_small", 200)]


def _fields(sim):
    return {name: sim.to_numpy(getattr(sim, name)) for name in FIELDS}


def _run_pair(name, steps, dtype):
    sc = S.scenario(name)
    sim = S.build_sim(sc, dtype=dtype, compile_step=True, compile_chunk_size=50)
    ref = S.build_oracle(sc)
    if sc["seed_fields"]:
        S.apply_seed_fields(ref, lambda k, v: getattr(ref, k).__setitem__(slice(None), v))
        S.apply_seed_fields(sim, lambda k, v: setattr(sim, k, v))
    pts = S.receiver_points(sc)
    v, l, w = S.receiver_tables(sim.mesh, sim.altitudes_m, pts)
    traces = sim.record_er_observations(v, l, w, steps)
    want = ref.record_er_observations(v, l, w, steps)
    return sim, ref, traces, want


@pytest.mark.parametrize("dtype", ["float16", "bfloat16"])
@pytest.mark.parametrize("name,steps", CASES)
def test_narrow_storage_error_vs_fp64_oracle(name, steps, dtype):
    sim, ref, traces, want = _run_pair(name, steps, dtype)
    got = _fields(sim)
    errs = S.family_errors(got, {k: getattr(ref, k) for k in FIELDS})
    # Traces far below the largest field (Er at the ground under a horizontal
    # source: 1e-13 of Et) lie outside FP16's 2^-24..2^16 range at any scale.
    peak = max(float(np.max(np.abs(getattr(ref, k)))) for k in ("er", "et"))
    trace_checked = float(np.max(np.abs(want))) >= 1e-6 * peak
    trace_err = S.rel_l2(traces, want) if trace_checked else 0.0
    diag = sim.field_diagnostics()
    print(f"{name} {dtype} steps={steps} trace={trace_err:.2e}{'' if trace_checked else '(n/a)'} "
          + " ".join(f"{k}={v:.2e}" for k, v in errs.items())
          + f" divb={diag['divb_max_abs'] / max(diag['divb_flux_scale'], 1e-300):.2e}"
          + f" scales={sim.storage_scales}", flush=True)
    assert diag["nonfinite"] == 0 and diag["storage_overflow"] == 0
    bar = BARS[dtype]
    assert trace_err <= bar, f"trace rel-L2 {trace_err:.3e} > {bar}"
    for k, e in errs.items():
        assert e <= bar, f"{k} rel-L2 {e:.3e} > {bar}"


def test_fp16_dynamic_scaling_is_amplitude_invariant():
    """Fields of a 2^-20 weaker source are exactly 2^-20 times the fields."""
    base = S.scenario("c0_pulse")
    weak = S.scenario("c0_pulse")
    weak["source"] = dict(weak["source"], peak_current_a=1.0e6 * 2.0 ** -20)
    out = []
    for sc in (base, weak):
        sim = S.build_sim(sc, dtype="float16", compile_step=True, compile_chunk_size=40)
        sim.step(240)
        out.append((_fields(sim), sim.storage_scales))
    (fa, sa), (fb, sb) = out
    print("scales", sa, sb, flush=True)
    for k in FIELDS:
        assert sa[k] == sb[k] * 2.0 ** -20 or not np.any(fa[k])
        np.testing.assert_array_equal(fa[k] * 2.0 ** -20, fb[k])
    assert np.any(fa["er"])


def test_fp16_graph_chunks_equal_single_launches():
    """Rescale points fall every 32 steps on both paths, so results are bitwise equal."""
    sc = S.scenario("c0_plasma")
    a = S.build_sim(sc, dtype="float16", compile_step=True, compile_chunk_size=32)
    b = S.build_sim(sc, dtype="float16", compile_step=False)
    a.step(96)
    a.step(64)
    b.step(160)
    for k in FIELDS:
        np.testing.assert_array_equal(a.to_numpy(getattr(a, f"_{k}_buf")),
                                      b.to_numpy(getattr(b, f"_{k}_buf")))
    assert a.storage_scales == b.storage_scales


def test_fp16_setter_places_peak_and_round_trips():
    sc = S.scenario("pec_vacuum")
    sim = S.build_sim(sc, dtype="float16")
    rng = np.random.default_rng(7)
    values = rng.standard_normal((sim.mesh.n_vertices, len(sim.radii_m))) * 3.0e-9
    sim.er = values
    scale = sim.storage_scales["er"]
    peak = np.max(np.abs(values))
    assert scale == 2.0 ** np.floor(np.log2(1024.0 / peak))
    got = sim.to_numpy(sim.er)
    # round to nearest in FP16: relative error <= 2^-11 for normal values
    normal = np.abs(values) * scale >= 2.0 ** -14
    assert np.all(np.abs(got - values)[normal] <= 2.0 ** -11 * np.abs(values)[normal])


def test_static_fp16_scale_overflow_is_rejected():
    sc = S.scenario("pec_vacuum")
    sim = S.build_sim(sc, dtype="float16", storage_scaling="static",
                      storage_scales={"er": 2.0 ** 40})
    with pytest.raises(ValueError, match="overflow"):
        sim.er = np.full((sim.mesh.n_vertices, len(sim.radii_m)), 1.0e-6)


QUASI_STATIC = {"float16": 1e-2, "bfloat16": 5e-2}


def _quasi_static_case(name):
    """The scenario with the benchmark's 20 Hz carrier source (no pulse
    override) and receivers where the near field is: at the source and
    200 / 400 km from it."""
    from ionosphere_fdtd_amd.sources import destination_point

    sc = S.scenario(name)
    sc["source"] = {"latitude_deg": S.SOURCE_LAT, "longitude_deg": S.SOURCE_LON,
                    "altitude_m": 0.0, "carrier_frequency_hz": 20.0}
    sc["receiver_points"] = [(S.SOURCE_LAT, S.SOURCE_LON)] + [
        destination_point(S.SOURCE_LAT, S.SOURCE_LON, az, d, sc["earth_radius_m"])
        for az, d in ((90.0, 200e3), (0.0, 400e3))]
    return sc


@pytest.fixture(scope="module")
def quasi_static_oracles():
    """FP64 oracle runs: c0 (s2 x 24, SI ground, 10 000 steps) and c3_small
    (the graded 96-cell C3 grid at s3 with the Drude plasma, 2000 steps)."""
    out = {}
    for name, steps in (("c0", 10_000), ("c3_small", 2000)):
        sc = _quasi_static_case(name)
        ref = S.build_oracle(sc)
        import ionosphere_fdtd_amd as g

        mesh = g.build_geodesic_mesh(sc["subdivision"])
        v, l, w = S.receiver_tables(mesh, ref.altitudes_m, S.receiver_points(sc))
        traces = ref.record_er_observations(v, l, w, steps)
        out[name] = (sc, steps, (v, l, w), {k: getattr(ref, k).copy() for k in FIELDS}, traces)
    return out


@pytest.mark.parametrize("name", ["c0", "c3_small"])
def test_narrow_storage_follows_a_quasi_static_source(quasi_static_oracles, name):
    sc, steps, (v, l, w), want, want_tr = quasi_static_oracles[name]
    assert float(np.max(np.abs(want_tr))) > 0
    errs = {}
    for dtype, rounding in (("float16", "rotation"), ("float16", "nearest"),
                            ("bfloat16", "rotation"), ("bfloat16", "nearest")):
        sim = S.build_sim(sc, dtype=dtype, compile_step=True, compile_chunk_size=100,
                          storage_rounding=rounding)
        traces = sim.record_er_observations(v, l, w, steps)
        e = S.family_errors(_fields(sim), want)
        e["traces"] = S.rel_l2(traces, want_tr)
        diag = sim.field_diagnostics()
        assert diag["nonfinite"] == 0 and diag["storage_overflow"] == 0
        errs[dtype, rounding] = e
        print(f"{name} {steps} steps {dtype} {rounding}: "
              + " ".join(f"{k}={x:.2e}" for k, x in e.items()), flush=True)
    for dtype, bar in QUASI_STATIC.items():
        worst = max(errs[dtype, "rotation"].values())
        assert worst <= bar, f"{dtype}: {errs[dtype, 'rotation']}"
        # what the rounding buys: nearest on the same run is far worse
        assert max(errs[dtype, "nearest"].values()) >= 4.0 * worst


def test_stochastic_rounding_is_repeatable_and_keyed_on_the_step_index(tmp_path):
    """Rounding thresholds are hashed from (step, field, element): two runs
    give the same bits, a run split over several calls equals one call, and
    a run resumed from a checkpoint at step 60 continues exactly as the run
    that was never interrupted, because the plan's step index is restored; a
    resumed run left at step index 0 does not.  BF16 (static scales) on c0_pulse
    with seeded fields: SI ground and a live source, and every piece of state
    (fields, SI memory) survives the checkpoint exactly."""
    import ionosphere_fdtd_amd as g

    sc = S.scenario("c0_pulse")

    def build(**kw):
        sim = S.build_sim(sc, dtype="bfloat16", compile_step=True, compile_chunk_size=20, **kw)
        S.apply_seed_fields(sim, lambda k, v: setattr(sim, k, v))
        return sim

    def state(sim):
        out = {k: sim.to_numpy(getattr(sim, f"_{k}_buf")) for k in FIELDS}
        out["si"] = sim.to_numpy(sim._si_memory)
        return out

    runs = []
    for split in ((100,), (60, 40), (100,)):
        sim = build()
        for n in split:
            sim.step(n)
        runs.append(state(sim))
    for k in runs[0]:
        np.testing.assert_array_equal(runs[0][k], runs[1][k])
        np.testing.assert_array_equal(runs[0][k], runs[2][k])
    sim = build()
    sim.step(60)
    path = sim.save_checkpoint(tmp_path / "bf16.npz")
    for restore_index in (True, False):
        resumed = g.GeodesicFDTD.load_checkpoint(path, dtype="bfloat16", compile_step=True,
                                                 compile_chunk_size=20)
        assert resumed.steps == 60
        if not restore_index:
            resumed._device_steps = 60  # planted: the plan keeps step index 0
        resumed.step(40)
        got = state(resumed)
        same = all(np.array_equal(got[k], runs[0][k]) for k in runs[0])
        assert same == restore_index, restore_index
    # nearest rounding differs from the stochastic thresholds
    nearest = build(storage_rounding="nearest")
    nearest.step(100)
    assert not np.array_equal(nearest.to_numpy(nearest._ht_buf), runs[0]["ht"])
