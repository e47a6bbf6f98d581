"""Whole-row and flat thread mappings of the step kernels give the same bits.

The two mappings assign the same strips to different threads: a strip's
neighbour values come from an adjacent lane or from a load of the same
address, and the surface-impedance history from four lanes or from one in
the same order, so every stored value, plasma state, surface-impedance
memory and trace must agree bit for bit.  The automatic choice takes the
flat mapping where whole-row blocks would leave 16 or more of their 256
lanes idle (96 radial cells: rows of 25 and 24 strips) and whole rows
otherwise (24 radial cells: 7 and 6 strips, 4 idle lanes each).
"""

from __future__ import annotations

import numpy as np
import pytest

import scenarios as S

pytestmark = pytest.mark.gpu

ROWS, FLAT = 1, 2


def _state(sim):
    out = {name: sim.to_numpy(getattr(sim, name)).copy() for name in ("er", "et", "hr", "ht")}
    if sim._plasma_buf is not None:
        out["plasma"] = sim._plasma_buf.cpu().numpy().copy()
    if sim._si_memory is not None:
        out["si"] = sim._si_memory.cpu().numpy().copy()
    return out


def _run(sc, dtype, mapping, steps=40):
    sim = S.build_sim(sc, dtype=dtype, compile_step=True, compile_chunk_size=16,
                      thread_mapping=mapping)
    if sc["seed_fields"]:
        S.apply_seed_fields(sim, lambda k, v: setattr(sim, k, v))
    v, l, w = S.receiver_tables(sim.mesh, sim.altitudes_m, S.receiver_points(sc))
    traces = sim.record_er_observations(v, l, w, steps)
    rng = np.random.default_rng(11)
    R = len(sim.radial_midpoints_m)
    faces = rng.integers(0, sim.mesh.n_faces, (5, 3))
    edges = rng.integers(0, sim.mesh.n_edges, (4, 2))
    h_r, h_t = sim.record_h_observations(
        faces, rng.integers(0, R, (5, 3)), rng.random((5, 3)),
        edges, rng.integers(0, R + 1, (4, 2)), rng.random((4, 2)), 9, sample_every=2)
    return sim, {"traces": traces, "h_r": h_r, "h_t": h_t, **_state(sim)}


@pytest.mark.parametrize("name,dtype", [
    ("c3_seeded", "float32"), ("c3_seeded", "float64"), ("c0_plasma", "float32"),
    ("drude_two_species", "float32"), ("magnetized_random", "float32"),
    ("tangential", "float32"), ("pec_vacuum", "float32"), ("c0_pulse", "float16"),
])
def test_thread_mappings_give_the_same_bits(name, dtype):
    sc = S.scenario(name)
    rows_sim, rows = _run(sc, dtype, "rows")
    flat_sim, flat = _run(sc, dtype, "flat")
    assert (rows_sim._layout.h_thread_mapping, rows_sim._layout.e_thread_mapping) == (ROWS, ROWS)
    assert (flat_sim._layout.h_thread_mapping, flat_sim._layout.e_thread_mapping) == (FLAT, FLAT)
    assert rows.keys() == flat.keys()
    for key in rows:
        assert rows[key].shape == flat[key].shape, key
        assert rows[key].tobytes() == flat[key].tobytes(), f"{name} {dtype}: {key} differs"
    assert float(np.max(np.abs(rows["traces"]))) > 0.0


def test_automatic_mapping_follows_idle_lanes():
    c3 = S.build_sim(S.scenario("c3_seeded"), dtype="float32")
    assert (c3._layout.h_thread_mapping, c3._layout.e_thread_mapping) == (FLAT, FLAT)
    c0 = S.build_sim(S.scenario("c0_plasma"), dtype="float32")
    assert (c0._layout.h_thread_mapping, c0._layout.e_thread_mapping) == (ROWS, ROWS)
    with pytest.raises(ValueError, match="thread_mapping"):
        S.build_sim(S.scenario("c0_plasma"), thread_mapping="diagonal")
