"""Checkpoint files (CPU): reference-written version-4 files are read and
validated, corrupt or unsupported files raise CheckpointError
(reference tests/test_checkpoint.py:245-257)."""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

import ionosphere_fdtd_amd as g
from ionosphere_fdtd_amd.checkpoint import (CheckpointError, _plasma_from_payload,
                                            read_checkpoint)

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_reads_reference_checkpoint_with_surface_impedance():
    meta, arrays = read_checkpoint(GOLDEN / "checkpoint_c0_pulse.npz")
    assert meta["version"] == 4 and meta["state"]["steps"] == 60
    assert meta["surface_impedance"]["terms"] == 16
    assert arrays["surface_impedance_memory"].shape == (480, 16)
    mesh = g.build_geodesic_mesh_from_topology(arrays["mesh_vertices"], arrays["mesh_faces"],
                                               subdivision=2, normalize_vertices=False)
    ref = g.build_geodesic_mesh(2)
    assert np.array_equal(mesh.edges, ref.edges) and np.array_equal(mesh.vertices, ref.vertices)


def test_reads_reference_plasma_payload_and_validates_mesh():
    meta, arrays = read_checkpoint(GOLDEN / "checkpoint_plasma_random.npz")
    model = _plasma_from_payload(meta["plasma"], {k: arrays[f"plasma_model_{k}"]
                                                  for k in meta["plasma"]["arrays"]})
    mesh = g.build_geodesic_mesh(2)
    model.validate_grid(mesh, model.radial_midpoint_altitudes_m)
    assert arrays["plasma_current_0"].shape == (mesh.n_faces, 24, 3)


def _rewrite(src, dst, mutate):
    with np.load(src, allow_pickle=False) as archive:
        arrays = {k: archive[k] for k in archive.files}
    mutate(arrays)
    np.savez(dst, **arrays)
    return dst


def test_rejects_unsupported_version(tmp_path):
    def bump(arrays):
        meta = json.loads(str(arrays["metadata"].item()))
        meta["version"] = 99
        arrays["metadata"] = np.asarray(json.dumps(meta))

    path = _rewrite(GOLDEN / "checkpoint_c0_pulse.npz", tmp_path / "v99.npz", bump)
    with pytest.raises(CheckpointError, match="unsupported checkpoint version"):
        read_checkpoint(path)


def test_rejects_missing_field_and_foreign_file(tmp_path):
    path = _rewrite(GOLDEN / "checkpoint_c0_pulse.npz", tmp_path / "no_er.npz",
                    lambda a: a.pop("er"))
    with pytest.raises(CheckpointError, match="missing arrays: er"):
        read_checkpoint(path)
    np.savez(tmp_path / "other.npz", metadata=np.asarray(json.dumps({"format": "x"})))
    with pytest.raises(CheckpointError, match="not an ionosphere-fdtd checkpoint"):
        read_checkpoint(tmp_path / "other.npz")


def test_rejects_tampered_plasma_array(tmp_path):
    def tamper(arrays):
        arrays["plasma_model_magnetic_field_t"] = arrays["plasma_model_magnetic_field_t"] + 1e-9

    path = _rewrite(GOLDEN / "checkpoint_plasma_random.npz", tmp_path / "t.npz", tamper)
    meta, arrays = read_checkpoint(path)
    with pytest.raises(CheckpointError, match="checksum"):
        _plasma_from_payload(meta["plasma"], {k: arrays[f"plasma_model_{k}"]
                                              for k in meta["plasma"]["arrays"]})
