"""Torch-CPU leg of the CPU baseline (test and benchmark infrastructure only).

The reference runs the same array algorithm on torch tensors when it is given
``backend="torch", device="cpu"`` (backends/torch_backend.py:151-200: fancy
indexing for the surface operators, in-place ``mul_`` / ``add_``, one slot at
a time for the circulations; torch.set_num_threads at :55).  This module
restates the oracle's step (oracle/geofdtd_oracle.py, which cites the solver
lines it follows) on torch CPU tensors so that ``bench.py``'s CPU baseline
can time that path on every host core.  It is only imported by ``tests/`` and
``bench.py``'s ``cpu_baseline`` leg, never by the product, and
tests/test_oracle_pinning.py holds it to the NumPy oracle at 1e-12 (FP64).

Only what the benchmark workloads use is covered: full-spherical or
thin-shell geometry, PEC or surface-impedance ground, vertical source and
the B = 0 (Drude) or magnetised cold-plasma ADE.
"""

from __future__ import annotations

import numpy as np

MU_0 = 1.256_637_062_12e-6


class TorchCpuSim:
    """The state and tables of an ``OracleSim`` as torch CPU tensors, and its
    leapfrog step (solver.py:789-989) on them."""

    def __init__(self, oracle, threads: int | None = None):
        import torch

        self.torch = torch
        if threads is not None:
            torch.set_num_threads(int(threads))
        o = oracle
        self.o = o
        dt = {np.dtype(np.float64): torch.float64, np.dtype(np.float32): torch.float32}[o.dtype]
        self.dtype = dt

        def t(a):  # copies: the mesh arrays are read-only views
            return torch.from_numpy(np.array(a, copy=True, order="C")).to(dt)

        def ix(a):
            return torch.from_numpy(np.array(a, dtype=np.int64, copy=True, order="C"))

        m = o.mesh
        self.edges = ix(m.edges)
        self.face_edges = ix(m.face_edges)
        self.face_edge_signs = t(o.face_edge_signs)
        self.left, self.right = ix(m.edge_left_faces), ix(m.edge_right_faces)
        self.vertex_edges = ix(o.vertex_edges)
        self.vertex_edge_signs = t(o.vertex_edge_signs)
        self.n_vertices = m.n_vertices
        self.full = o.geometry_mode == "full-spherical"
        for name in ("_primal", "_inv_primal", "_dual", "_inv_dual", "_inv_dual_area",
                     "_inv_face_area", "_radii", "_inv_radii", "_mid", "_inv_mid", "_steps",
                     "_control", "_center_dist", "_ca_er", "_cb_er", "_ca_et", "_cb_et"):
            setattr(self, name, t(getattr(o, name)))
        self.er, self.et, self.hr, self.ht = (t(getattr(o, k)) for k in ("er", "et", "hr", "ht"))
        self.dt = o.time_step_s
        self.si = o.surface is not None
        if self.si:
            self.si_memory = t(o.si_memory)
            self.si_decay, self.si_drive = t(o.si_decay), t(o.si_drive)
            self.si_weights, self.si_scale = t(o.si_history_weights), t(o.si_scale)
            self.si_metric = (o.radial_midpoints_m[0] / o.radii_m[0] if self.full else 1.0)
        self.source = o.source
        if o.source_distribution is not None:
            v, l, w = o.source_distribution
            self.src = (ix(v), ix(l), t(w))
        else:
            self.src = None
        c = o.coupler
        self.plasma = c is not None
        if self.plasma:
            self.faces = ix(m.faces)
            self.reconstruction = t(c.reconstruction)
            self.face_centers = t(c.face_centers)
            self.left_tangents, self.right_tangents = t(c.left_tangents), t(c.right_tangents)
            self.vertex_faces = ix(c.vertex_faces)
            self.vertex_face_weights = t(c.vertex_face_weights)
            self.b = t(c.b)
            self.magnetized = bool(np.any(np.asarray(c.b) != 0.0))
            self.coefficients = [tuple(t(x) for x in k) for k in c.coefficients]
            self.current_density = [t(j) for j in c.current_density]
        self.steps = o.steps
        self.time_s = o.time_s

    # -- surface operators (torch_backend.py:170-200) ---------------------------
    def _edge_difference(self, v):
        return v[self.edges[:, 1]] - v[self.edges[:, 0]]

    def _dual_edge_difference(self, f):
        return f[self.left] - f[self.right]

    def _face_circulation(self, ev):
        result = ev[self.face_edges[:, 0]]
        result.mul_(self.face_edge_signs[:, 0, None])
        for corner in (1, 2):
            term = ev[self.face_edges[:, corner]]
            term.mul_(self.face_edge_signs[:, corner, None])
            result.add_(term)
        return result

    def _dual_cell_circulation(self, ev):
        result = ev[self.vertex_edges[:, 0]]
        result.mul_(self.vertex_edge_signs[:, 0, None])
        for slot in range(1, self.vertex_edges.shape[1]):
            term = ev[self.vertex_edges[:, slot]]
            term.mul_(self.vertex_edge_signs[:, slot, None])
            result.add_(term)
        return result

    # -- radial derivatives (solver.py:443-465) ---------------------------------
    def _radial_derivative_et(self):
        values = self.et * self._mid[None, :] if self.full else self.et
        result = self.torch.empty_like(self.ht)
        result[:, 0] = 2.0 * values[:, 0] / self._steps[0]
        result[:, -1] = -2.0 * values[:, -1] / self._steps[-1]
        if self.ht.shape[1] > 2:
            result[:, 1:-1] = self.torch.diff(values, dim=1) / self._center_dist
        if self.full:
            result.mul_(self._inv_radii)
        return result

    def _radial_derivative_ht(self):
        values = self.ht * self._radii[None, :] if self.full else self.ht
        result = self.torch.diff(values, dim=1) / self._steps[None, :]
        if self.full:
            result.mul_(self._inv_mid)
        return result

    # -- surface impedance (surface_impedance.py:225-270) -----------------------
    def _si_lower_boundary(self, h_old, electric_first_cell, surface_gradient):
        dt = self.dt
        history = self.si_scale * (self.si_memory * self.si_weights[None, :]).sum(dim=1)
        gain = self.si_scale * self.si_weights.sum()
        curl_scale = 2.0 * dt / (MU_0 * self._steps[0])
        forcing = dt / MU_0 * surface_gradient
        coupling = curl_scale * gain
        h_new = ((1.0 - 0.5 * coupling) * h_old + forcing - curl_scale * electric_first_cell
                 + curl_scale * history) / (1.0 + 0.5 * coupling)
        self.si_memory.mul_(self.si_decay[None, :]).add_(
            (h_new + h_old)[:, None] * self.si_drive[None, :])
        return h_new

    # -- plasma coupler and ADE (plasma.py:413-442, 489-527) ----------------------
    @staticmethod
    def _cross_with_b(values, b):
        out = values.new_empty(values.shape)
        out[..., 0] = values[..., 1] * b[..., 2] - values[..., 2] * b[..., 1]
        out[..., 1] = values[..., 2] * b[..., 0] - values[..., 0] * b[..., 2]
        out[..., 2] = values[..., 0] * b[..., 1] - values[..., 1] * b[..., 0]
        return out

    def _ade_advance(self, electric):
        total = self.torch.zeros_like(electric)
        b = self.b
        e_par = (electric * b).sum(dim=2, keepdim=True) * b
        e_perp = electric - e_par
        e_cross = self._cross_with_b(e_perp, b)
        for index, (decay, cosine, sine, d_par, d_re, d_im) in enumerate(self.coefficients):
            current = self.current_density[index]
            c_par = (current * b).sum(dim=2, keepdim=True) * b
            c_perp = current - c_par
            c_cross = self._cross_with_b(c_perp, b)
            updated = (decay[..., None] * c_par + d_par[..., None] * e_par
                       + decay[..., None] * (cosine[..., None] * c_perp + sine[..., None] * c_cross)
                       + d_re[..., None] * e_perp + d_im[..., None] * e_cross)
            current.copy_(updated)
            total.add_(updated)
        return total

    def _plasma_advance(self):
        er, et = self.er, self.et
        nodes_at_faces = er[self.faces].sum(dim=1) / 3.0
        radial = 0.5 * (nodes_at_faces[:, :-1] + nodes_at_faces[:, 1:])
        edge_values = et[self.face_edges]
        tangential = (edge_values[..., None] * self.reconstruction[:, :, None, :]).sum(dim=1)
        electric = tangential + radial[..., None] * self.face_centers[:, None, :]
        current = self._ade_advance(electric)
        left = (current[self.left] * self.left_tangents[:, None, :]).sum(dim=2)
        right = (current[self.right] * self.right_tangents[:, None, :]).sum(dim=2)
        tangential_current = 0.5 * (left + right)
        radial_at_faces = (current * self.face_centers[:, None, :]).sum(dim=2)
        radial_at_vertices = (radial_at_faces[self.vertex_faces]
                              * self.vertex_face_weights[:, :, None]).sum(dim=1)
        radial_current = radial_at_vertices.new_zeros(
            (radial_at_vertices.shape[0], radial_at_vertices.shape[1] + 1))
        radial_current[:, 0] = radial_at_vertices[:, 0]
        radial_current[:, -1] = radial_at_vertices[:, -1]
        if radial_at_vertices.shape[1] > 1:
            radial_current[:, 1:-1] = 0.5 * (radial_at_vertices[:, :-1]
                                             + radial_at_vertices[:, 1:])
        return radial_current, tangential_current

    # -- half steps (solver.py:877-989) ------------------------------------------
    def update_magnetic_fields(self):
        dt = self.dt
        g = self._edge_difference(self.er)
        g.mul_(self._inv_primal).mul_(self._inv_radii)
        lower_g = g[:, 0].clone() if self.si else None
        lower_h_old = self.ht[:, 0].clone() if self.si else None
        g.sub_(self._radial_derivative_et())
        g.mul_(dt / MU_0)
        self.ht.add_(g)
        if self.si:
            self.ht[:, 0] = self._si_lower_boundary(lower_h_old, self.si_metric * self.et[:, 0],
                                                    lower_g)
        circ = self._face_circulation(self.et * self._primal)
        circ.mul_(self._inv_face_area).mul_(self._inv_mid).mul_(dt / MU_0)
        self.hr.sub_(circ)

    def update_electric_fields(self, current_a=0.0):
        plasma_current = self._plasma_advance() if self.plasma else None
        circ = self._dual_cell_circulation(self.ht * self._dual)
        circ.mul_(self._inv_dual_area).mul_(self._inv_radii)
        cd = None
        if self.src is not None:
            verts, layers, weights = self.src
            cd = (weights * current_a * self.source.vertical_element_length_m
                  * self._inv_dual_area[verts, 0] / self._radii[layers] ** 2
                  / self._control[layers])
        self.er.mul_(self._ca_er)
        circ.mul_(self._cb_er)
        self.er.add_(circ)
        if cd is not None:
            verts, layers, _ = self.src
            cv = verts if self._cb_er.shape[0] > 1 else self.torch.zeros_like(verts)
            self.er[verts, layers] -= self._cb_er[cv, layers] * cd
        if plasma_current is not None:
            self.er.sub_(self._cb_er * plasma_current[0])
        g = self._dual_edge_difference(self.hr)
        g.mul_(self._inv_dual).mul_(self._inv_mid)
        g.sub_(self._radial_derivative_ht())
        self.et.mul_(self._ca_et)
        g.mul_(self._cb_et)
        self.et.add_(g)
        if plasma_current is not None:
            self.et.sub_(self._cb_et * plasma_current[1])

    def step(self, count: int = 1) -> None:
        for _ in range(count):
            current = (self.source.current_a(self.time_s + 0.5 * self.dt, self.dt)
                       if self.source is not None else 0.0)
            self.update_magnetic_fields()
            self.update_electric_fields(current)
            self.steps += 1
            self.time_s = self.steps * self.dt
