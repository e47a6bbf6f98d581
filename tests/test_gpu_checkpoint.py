"""Checkpoints of the device state (GPU): resume from reference files,
resume equivalence, dtype conversion (reference tests/test_checkpoint.py:79-232)."""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import scenarios as S

import ionosphere_fdtd_amd as g

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"
FIELDS = ("er", "et", "hr", "ht")


@pytest.mark.parametrize("name", ("c0_pulse", "plasma_random"))
def test_resume_reference_checkpoint_matches_reference_continuation(name):
    sim = g.GeodesicFDTD.load_checkpoint(GOLDEN / f"checkpoint_{name}.npz", dtype="float64",
                                         compile_step=True, compile_chunk_size=16)
    assert sim.steps == 60
    sim.step(40)
    fx = np.load(GOLDEN / f"checkpoint_{name}_continued.npz")
    errs = S.family_errors({k: sim.to_numpy(getattr(sim, k)) for k in FIELDS},
                           {k: fx[k] for k in FIELDS})
    if "si_memory" in fx:
        errs["si"] = S.rel_l2(sim.to_numpy(sim._si_memory), fx["si_memory"])
    for i, j in enumerate(sim.plasma_current_density):
        errs[f"j{i}"] = S.rel_l2(sim.to_numpy(j), fx[f"plasma_current_{i}"])
    print(name, {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    assert max(errs.values()) <= 1e-10, errs


@pytest.mark.parametrize("dtype", ("float64", "float32"))
@pytest.mark.parametrize("name", ("c0_pulse", "plasma_random"))
def test_interrupted_run_resumes_bitwise(tmp_path, name, dtype):
    sc = S.scenario(name)
    straight = S.build_sim(sc, dtype=dtype, compile_step=True, compile_chunk_size=8)
    first = S.build_sim(sc, dtype=dtype, compile_step=True, compile_chunk_size=8)
    for sim in (straight, first):
        if sc["seed_fields"]:
            S.apply_seed_fields(sim, lambda k, v, s=sim: setattr(s, k, v))
    straight.step(50)
    first.step(30)
    path = first.save_checkpoint(tmp_path / "mid.npz")
    resumed = g.GeodesicFDTD.load_checkpoint(path, compile_step=True, compile_chunk_size=8)
    assert resumed.backend.dtype_name == dtype and resumed.steps == 30
    resumed.step(20)
    for k in FIELDS:
        a = straight.to_numpy(getattr(straight, k))
        b = resumed.to_numpy(getattr(resumed, k))
        if name != "plasma_random":
            assert np.array_equal(a, b), k
        else:
            # the projected plasma state round-trips through Cartesian J, which
            # rounds once; the continuation then differs at roundoff only
            bar = 1e-12 if dtype == "float64" else 1e-6
            assert S.rel_l2(b, a) <= bar, k


def test_checkpoint_dtype_conversion_on_load(tmp_path):
    sc = S.scenario("c0_pulse")
    sim = S.build_sim(sc)
    sim.step(25)
    path = sim.save_checkpoint(tmp_path / "f64.npz")
    as32 = g.GeodesicFDTD.load_checkpoint(path, dtype="float32")
    assert as32.backend.dtype_name == "float32"
    for k in FIELDS:
        np.testing.assert_allclose(as32.to_numpy(getattr(as32, k)),
                                   sim.to_numpy(getattr(sim, k)).astype(np.float32), rtol=0,
                                   atol=0)
