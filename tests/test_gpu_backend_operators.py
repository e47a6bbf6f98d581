"""Surface operators of the HIP backend (backends/base.py:38-56) against the
reference's NumPy formulas (numpy_backend.py:90-114), evaluated here in the
same order of additions: bitwise equal in FP64 and FP32."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _numpy_ops(mesh):
    from ionosphere_fdtd_amd.mesh import vertex_incidence

    ve, vs = vertex_incidence(mesh)

    def dual_cell(e):
        shape = (mesh.n_vertices,) + (1,) * (e.ndim - 1)
        out = e[ve[:, 0]] * vs[:, 0].reshape(shape).astype(e.dtype)
        for s in range(1, ve.shape[1]):
            out = out + e[ve[:, s]] * vs[:, s].reshape(shape).astype(e.dtype)
        return out

    def face_circ(e):
        sel = e[mesh.face_edges]
        sg = mesh.face_edge_signs.astype(e.dtype)
        while sg.ndim < sel.ndim:
            sg = sg[..., None]
        prod = sel * sg
        return (prod[:, 0] + prod[:, 1]) + prod[:, 2]

    return {
        "edge_difference": lambda v: v[mesh.edges[:, 1]] - v[mesh.edges[:, 0]],
        "dual_edge_difference": lambda f: f[mesh.edge_left_faces] - f[mesh.edge_right_faces],
        "face_circulation": face_circ,
        "dual_cell_circulation": dual_cell,
    }


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("trailing", [(), (7,), (3, 5)])
def test_surface_operators_equal_reference_formulas(dtype, trailing):
    import ionosphere_fdtd_amd as g

    mesh = g.build_geodesic_mesh(3)
    backend = g.create_backend("hip", mesh, dtype=dtype)
    ref = _numpy_ops(mesh)
    rng = np.random.default_rng(11)
    rows_in = {"edge_difference": mesh.n_vertices, "dual_edge_difference": mesh.n_faces,
               "face_circulation": mesh.n_edges, "dual_cell_circulation": mesh.n_edges}
    for name, rows in rows_in.items():
        x = rng.standard_normal((rows,) + trailing).astype(dtype)
        got = getattr(backend, name)(backend.torch.as_tensor(x, device=backend.torch_device))
        want = ref[name](x)
        assert got.dtype == (backend.torch.float32 if dtype == "float32" else backend.torch.float64)
        np.testing.assert_array_equal(got.cpu().numpy(), want, err_msg=name)
    # the host operators of the mesh agree with the backend (np.add.at order)
    e = rng.standard_normal((mesh.n_edges,) + trailing)
    got = backend.dual_cell_circulation(backend.torch.as_tensor(e.astype(dtype),
                                                                device=backend.torch_device))
    np.testing.assert_allclose(got.cpu().numpy(), mesh.dual_cell_circulation(e.astype(dtype)),
                               rtol=0, atol=1e-12 if dtype == "float64" else 1e-5)


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_diff_equals_numpy(dtype):
    import ionosphere_fdtd_amd as g

    mesh = g.build_geodesic_mesh(1)
    backend = g.create_backend("hip", mesh, dtype=dtype)
    rng = np.random.default_rng(4)
    for shape, axis in [((40,), 0), ((40,), -1), ((9, 13), 0), ((9, 13), 1), ((9, 13), -1)]:
        x = rng.standard_normal(shape).astype(dtype)
        got = backend.diff(backend.torch.as_tensor(x, device=backend.torch_device), axis)
        np.testing.assert_array_equal(got.cpu().numpy(), np.diff(x, axis=axis))
    with pytest.raises(ValueError):
        backend.diff(backend.zeros((2, 2, 2)), 0)
    with pytest.raises(ValueError):
        backend.edge_difference(backend.zeros((mesh.n_vertices + 1,)))
