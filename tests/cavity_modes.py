"""Standing-mode initialisation and frequency measurement for a PEC shell.

Test helper restating the reference's verification procedure
(verification/analytic_solutions/cavity.py:52-199, model.py:98-144) so the
measured frequencies and time steps in its artifacts/analytic-solutions/
full-field.csv can be reproduced against any step implementation that exposes
er / et / hr / ht arrays and a ``step()`` method.
"""

from __future__ import annotations

import numpy as np
from numpy.polynomial.legendre import Legendre
from scipy.optimize import brentq
from scipy.special import spherical_jn, spherical_yn

EPSILON_0 = 8.854_187_812_8e-12
MU_0 = 1.256_637_062_12e-6
C_0 = 1.0 / (EPSILON_0 * MU_0) ** 0.5


def shell_wavenumbers(degree, inner, outer, polarization, count=1, max_frequency_hz=10_000.0):
    """Roots of the PEC-shell determinant (model.py:98-144)."""

    def radial(kind, x):
        f = spherical_jn if kind == "j" else spherical_yn
        if polarization == "TE":
            return f(degree, x)
        return f(degree, x) + x * f(degree, x, derivative=True)

    def det(k):
        return float(radial("j", k * inner) * radial("y", k * outer)
                     - radial("y", k * inner) * radial("j", k * outer))

    kmax = 2.0 * np.pi * max_frequency_hz / C_0
    spacing = np.pi / (outer - inner)
    samples = max(20_000, int(80.0 * kmax / spacing))
    grid = np.linspace(kmax / samples, kmax, samples)
    values = np.asarray([det(v) for v in grid])
    roots = []
    for a, b, fa, fb in zip(grid[:-1], grid[1:], values[:-1], values[1:]):
        if np.isfinite(fa) and np.isfinite(fb) and fa * fb < 0.0:
            root = brentq(det, a, b, xtol=1e-15, rtol=1e-13)
            if not roots or abs(root - roots[-1]) > 1e-10 * root:
                roots.append(root)
                if len(roots) == count:
                    break
    return np.asarray(roots)


def _zonal(points, degree):
    z = np.clip(points[:, 2], -1.0, 1.0)
    poly = Legendre.basis(degree)
    values = poly(z)
    dtheta = -np.sqrt(np.maximum(0.0, 1.0 - z**2)) * poly.deriv()(z)
    lon = np.arctan2(points[:, 1], points[:, 0])
    sin_t = np.sqrt(np.maximum(0.0, 1.0 - z**2))
    e_theta = np.column_stack((z * np.cos(lon), z * np.sin(lon), -sin_t))
    return values, dtheta[:, None] * e_theta


def _radial_combination(degree, k, radii, inner, polarization):
    x = k * np.asarray(radii)
    xa = k * inner

    def val(kind, arg):
        return (spherical_jn if kind == "j" else spherical_yn)(degree, arg)

    def dcombo(kind, arg):
        f = spherical_jn if kind == "j" else spherical_yn
        return f(degree, arg) + arg * f(degree, arg, derivative=True)

    if polarization == "TE":
        radial = val("j", x) * val("y", xa) - val("y", x) * val("j", xa)
        return radial, radial
    cj = dcombo("y", xa)
    cy = dcombo("j", xa)
    radial = val("j", x) * cj - val("y", x) * cy
    deriv = dcombo("j", x) * cj - dcombo("y", x) * cy
    return degree * (degree + 1.0) * radial / x, deriv / x


def electric_mode(mesh, radii, midpoints, control, steps_m, degree, polarization, radial_index):
    """Sampled axisymmetric standing mode on the E degrees of freedom (cavity.py:52-135)."""
    k = float(shell_wavenumbers(degree, radii[0], radii[-1], polarization,
                                count=radial_index + 1)[radial_index])
    vertex_y, _ = _zonal(mesh.vertices, degree)
    mids = mesh.vertices[mesh.edges].sum(axis=1)
    mids /= np.linalg.norm(mids, axis=1, keepdims=True)
    _, grad = _zonal(mids, degree)
    ends = mesh.vertices[mesh.edges]
    tangent = ends[:, 1] - ends[:, 0]
    tangent /= np.linalg.norm(tangent, axis=1, keepdims=True)
    if polarization == "TE":
        ang = np.einsum("ij,ij->i", np.cross(mids, grad), tangent)
        er = np.zeros((mesh.n_vertices, len(radii)))
        rad = _radial_combination(degree, k, midpoints, radii[0], "TE")[0]
        et = ang[:, None] * rad[None, :]
    else:
        ang = np.einsum("ij,ij->i", grad, tangent)
        rer, _ = _radial_combination(degree, k, radii, radii[0], "TM")
        _, ret = _radial_combination(degree, k, midpoints, radii[0], "TM")
        er = vertex_y[:, None] * rer[None, :]
        et = ang[:, None] * ret[None, :]
    scale = max(float(np.max(np.abs(er))), float(np.max(np.abs(et))))
    er, et = er / scale, et / scale
    tail = mesh.vertices[mesh.edges[:, 0]]
    head = mesh.vertices[mesh.edges[:, 1]]
    lc = mesh.face_centers[mesh.edge_left_faces]
    rc = mesh.face_centers[mesh.edge_right_faces]

    def area(a, b, c):
        num = np.einsum("ij,ij->i", a, np.cross(b, c))
        den = 1.0 + np.einsum("ij,ij->i", a, b) + np.einsum("ij,ij->i", b, c) + np.einsum(
            "ij,ij->i", c, a)
        return 2.0 * np.arctan2(num, den)

    diamond = area(tail, head, lc) - area(tail, head, rc)
    er_w = EPSILON_0 * mesh.dual_cell_solid_angles[:, None] * (radii**2 * control)[None, :]
    et_w = EPSILON_0 * diamond[:, None] * (midpoints**2 * steps_m)[None, :]
    return {"k": k, "er": er, "et": et, "er_w": er_w, "et_w": et_w}


def measure_frequency(sim, mode, steps, get=lambda sim, n: getattr(sim, n)):
    """Fit the exact leapfrog recurrence of the mode amplitude (cavity.py:154-199)."""
    amp = np.empty(steps + 1)
    den = np.sum(mode["er_w"] * mode["er"] ** 2) + np.sum(mode["et_w"] * mode["et"] ** 2)
    for i in range(steps + 1):
        er = np.asarray(get(sim, "er"), dtype=np.float64)
        et = np.asarray(get(sim, "et"), dtype=np.float64)
        num = np.sum(mode["er_w"] * er * mode["er"]) + np.sum(mode["et_w"] * et * mode["et"])
        amp[i] = num / den
        if i < steps:
            sim.step()
    c = amp[1:-1]
    cosine = float(np.clip(np.dot(c, amp[2:] + amp[:-2]) / (2.0 * np.dot(c, c)), -1.0, 1.0))
    return np.arccos(cosine) / (2.0 * np.pi * sim.time_step_s)
