"""Receiver-trace queries: FDTD ensembles, the learned surrogate and the router.

No counterpart in the reference.  A query asks for the E_r trace at a
receiver for one configuration of the batched-ensemble family:

    (h' [km], beta [1/km], log10 ground sigma [S/m], source lat, source lon,
     receiver lat, receiver lon, carrier [Hz])

answered by a C2-style FDTD run: s5 mesh, 32 radial cells 0-100 km,
r_E = 6370 km, sigma(z) = 1e-6 exp((z - 70 km) / 5 km) S/m, electron Drude
plasma n_e = 1.43e13 exp(-0.15 h') exp((beta - 0.15)(z - h')) above 60 km with
nu = 1.8e11 exp(-0.15 z[km]), surface-impedance ground, a vertical Gaussian
current at the source (default timing for its carrier), FP32, courant 0.35.
The trace is E_r at the receiver after every 8th step of 4096 steps (512
samples; step numbers, not seconds, since the time step follows the plasma
CFL limit of each configuration: 0.9 ns to 0.5 us across the box).

**Decomposition.**  The run is linear in its source, the source is injected
at the three vertices of the mesh face containing it with barycentric weights
s_j (sources.py staggered_distribution at the ground node), and the receiver
reads the three vertices of its face with weights w_i.  So the FDTD trace of
a query is exactly

    trace = sum_j sum_i s_j w_i G(u_j, v_i),

with G(u, v) the E_r trace at vertex v of the same configuration driven by a
unit-weight source at vertex u.  The surrogate learns G for vertex pairs up to
three mesh rings apart from the configuration and the pair's local geometry
(distance, ring, dual-cell areas, edge lengths, degrees), and answers a query
with the same weighted sum of predicted G.  Sources may sit anywhere on the
sphere; a query is *in range* when its physics lies in the ensemble box
(h' 68-80 km, beta 0.25-0.55 /km, ground sigma 1e-4..1e-1 S/m, carrier
10-40 Hz), its receiver lies within 250 km of the source and every weighted
(u, v) pair is within three rings (at s5 an edge is ~240 km; three rings hold
every pair of every receiver within 300 km).  The router answers other
queries with FDTD.

**Router.**  A committee of networks, each trained on its own seeded half of
the training runs, answers in-range queries with the mean of its members'
traces; a query whose members disagree by more than `disagreement` (relative
L2) goes to FDTD, and one configuration per batch is spot-checked by FDTD
(any error above the tolerance sends the whole batch to FDTD).
"""

from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass

import numpy as np

from .constants import C_0, EPSILON_0
from .ensemble import EnsembleFDTD
from .materials import EarthIonosphereMaterial
from .mesh import build_geodesic_mesh
from .plasma import electron_drude_model
from .solver import GeodesicFDTD, SimulationConfig
from .sources import GaussianCurrent, barycentric_distribution, destination_point
from .surface_impedance import ConductiveHalfSpaceSurface

SUBDIVISION = 5
RADIAL_CELLS = 32
EARTH_RADIUS_M = 6.370e6
STEPS = 4096
STRIDE = 8
SAMPLES = STEPS // STRIDE
BOX = {"h_prime_km": (68.0, 80.0), "beta_per_km": (0.25, 0.55),
       "log10_ground_sigma": (-4.0, -1.0), "carrier_hz": (10.0, 40.0)}
MAX_DISTANCE_M = 250e3   # receivers in range: within this great-circle distance of the source
RINGS = 3                # vertex pairs the surrogate covers
COLUMNS = ("h_prime_km", "beta_per_km", "log10_ground_sigma", "src_lat", "src_lon", "rx_lat",
           "rx_lon", "carrier_hz")
ELECTRON_CHARGE_C = 1.602_176_634e-19
ELECTRON_MASS_KG = 9.109_383_713_9e-31


def _unit(lat, lon):
    la, lo = np.deg2rad(lat), np.deg2rad(lon)
    return np.stack((np.cos(la) * np.cos(lo), np.cos(la) * np.sin(lo), np.sin(la)), axis=-1)


def _latlon(unit):
    u = np.asarray(unit, dtype=np.float64)
    return (np.rad2deg(np.arcsin(np.clip(u[..., 2], -1.0, 1.0))),
            np.rad2deg(np.arctan2(u[..., 1], u[..., 0])))


def _altitudes():
    return np.linspace(0.0, 100e3, RADIAL_CELLS + 1)


# --------------------------------------------------------------- sampling
def _sample_physics(rng, n) -> np.ndarray:
    out = np.empty((n, 4))
    out[:, 0] = rng.uniform(*BOX["h_prime_km"], n)
    out[:, 1] = rng.uniform(*BOX["beta_per_km"], n)
    out[:, 2] = rng.uniform(*BOX["log10_ground_sigma"], n)
    out[:, 3] = rng.uniform(*BOX["carrier_hz"], n)
    return out


def random_queries(n_configs: int, per_config: int, seed: int) -> np.ndarray:
    """(n_configs * per_config, 8) queries: physics uniform in the box, the
    source uniform on the sphere, each receiver uniform over the disc of
    MAX_DISTANCE_M around its source (area-uniform distance, uniform azimuth)."""
    rng = np.random.default_rng(seed)
    phys = _sample_physics(rng, n_configs)
    src_lat = np.rad2deg(np.arcsin(rng.uniform(-1.0, 1.0, n_configs)))
    src_lon = rng.uniform(-180.0, 180.0, n_configs)
    rows = []
    for k in range(n_configs):
        for _ in range(per_config):
            d = MAX_DISTANCE_M * np.sqrt(rng.uniform())
            az = rng.uniform(0.0, 360.0)
            rlat, rlon = destination_point(src_lat[k], src_lon[k], az, d, EARTH_RADIUS_M)
            h, b, ls, f = phys[k]
            rows.append((h, b, ls, src_lat[k], src_lon[k], rlat, rlon, f))
    return np.asarray(rows, dtype=np.float64)


def training_runs(n: int, seed: int, mesh) -> np.ndarray:
    """(n, 5) training runs: physics uniform in the box and a source vertex
    drawn uniformly from the mesh vertices."""
    rng = np.random.default_rng(seed)
    out = np.empty((n, 5))
    out[:, :4] = _sample_physics(rng, n)
    out[:, 4] = rng.integers(0, mesh.n_vertices, n)
    return out


# ------------------------------------------------------- mesh neighbourhoods
class MeshRings:
    """Vertex rings (graph distance along primal edges) and the per-vertex
    geometry the pair features use."""

    def __init__(self, mesh, rings: int = RINGS):
        self.mesh = mesh
        self.rings = rings
        V = mesh.n_vertices
        e = np.asarray(mesh.edges, dtype=np.int64)
        order = np.argsort(np.concatenate((e[:, 0], e[:, 1])), kind="stable")
        ends = np.concatenate((e[:, 1], e[:, 0]))[order]
        counts = np.bincount(np.concatenate((e[:, 0], e[:, 1])), minlength=V)
        starts = np.concatenate(([0], np.cumsum(counts)))
        self.adjacent = [ends[starts[v]:starts[v + 1]] for v in range(V)]
        self.degree = counts
        theta = np.asarray(mesh.primal_edge_angles, dtype=np.float64)
        phi = np.asarray(mesh.dual_edge_angles, dtype=np.float64)
        incident = np.concatenate((theta, theta))[order]
        self.edge_length = np.add.reduceat(incident, starts[:-1]) / counts * EARTH_RADIUS_M
        # phi / theta: the weight of an edge in the surface Laplacian (the
        # coupling of its two vertices); per vertex, the mean over its edges
        self.edge_coupling = phi / theta
        incident_w = np.concatenate((self.edge_coupling, self.edge_coupling))[order]
        self.coupling = np.add.reduceat(incident_w, starts[:-1]) / counts
        self.mean_coupling = float(self.edge_coupling.mean())
        self.edge_of = {(int(a), int(b)): k for k, (a, b) in enumerate(e)}
        self.area = np.asarray(mesh.dual_cell_solid_angles, dtype=np.float64)
        self.mean_area = float(self.area.mean())
        self.mean_edge = float(self.edge_length.mean())
        self.unit = np.asarray(mesh.vertices, dtype=np.float64)
        self._hoods: dict[int, tuple[np.ndarray, np.ndarray]] = {}
        # padded adjacency (-1 past a vertex's degree) and sorted edge keys
        # min * V + max, for the vectorised pair lookups
        self.adjacent6 = np.full((V, 6), -1, dtype=np.int64)
        for v in range(V):
            self.adjacent6[v, :counts[v]] = self.adjacent[v]
        ekey = np.minimum(e[:, 0], e[:, 1]) * V + np.maximum(e[:, 0], e[:, 1])
        self._edge_order = np.argsort(ekey, kind="stable")
        self._edge_keys = ekey[self._edge_order]
        self._pair_keys = None
        self._pair_rings = None

    def _ring_table(self):
        """Sorted keys u * V + v of every ordered pair within `rings` rings and
        their ring numbers: breadth-first layers of the sparse adjacency."""
        if self._pair_keys is None:
            from scipy import sparse

            V = self.mesh.n_vertices
            e = np.asarray(self.mesh.edges, dtype=np.int64)
            adj = sparse.coo_matrix((np.ones(2 * len(e), dtype=np.int8),
                                     (np.concatenate((e[:, 0], e[:, 1])),
                                      np.concatenate((e[:, 1], e[:, 0])))), shape=(V, V)).tocsr()
            seen = sparse.identity(V, dtype=np.int8, format="csr")
            keys = [np.arange(V, dtype=np.int64) * (V + 1)]
            rings = [np.zeros(V, dtype=np.int64)]
            frontier = seen
            for r in range(1, self.rings + 1):
                reach = ((frontier @ adj) > 0).astype(np.int8)
                new = (reach - reach.multiply(seen)).tocoo()
                new.eliminate_zeros()
                keys.append(new.row.astype(np.int64) * V + new.col.astype(np.int64))
                rings.append(np.full(new.nnz, r, dtype=np.int64))
                seen = ((seen + new.tocsr()) > 0).astype(np.int8)
                frontier = new.tocsr()
            k = np.concatenate(keys)
            order = np.argsort(k)
            self._pair_keys, self._pair_rings = k[order], np.concatenate(rings)[order]
        return self._pair_keys, self._pair_rings

    def rings_of(self, u, v) -> np.ndarray:
        """Vectorised ring(u, v): ring numbers of the pairs, rings + 1 beyond."""
        keys, rings = self._ring_table()
        q = np.asarray(u, dtype=np.int64) * self.mesh.n_vertices + np.asarray(v, dtype=np.int64)
        at = np.minimum(np.searchsorted(keys, q), len(keys) - 1)
        return np.where(keys[at] == q, rings[at], self.rings + 1)

    def common_neighbours_of(self, u, v) -> np.ndarray:
        """Vectorised common_neighbours(u, v)."""
        a = self.adjacent6[np.asarray(u, dtype=np.int64)]
        b = self.adjacent6[np.asarray(v, dtype=np.int64)]
        return np.sum((a[:, :, None] == b[:, None, :]) & (a[:, :, None] >= 0), axis=(1, 2))

    def pair_couplings(self, u, v) -> np.ndarray:
        """Vectorised pair_coupling(u, v) for adjacent pairs (NaN otherwise)."""
        u = np.asarray(u, dtype=np.int64)
        v = np.asarray(v, dtype=np.int64)
        q = np.minimum(u, v) * self.mesh.n_vertices + np.maximum(u, v)
        at = np.minimum(np.searchsorted(self._edge_keys, q), len(self._edge_keys) - 1)
        hit = self._edge_keys[at] == q
        return np.where(hit, self.edge_coupling[self._edge_order[at]], np.nan)

    def neighbourhood(self, u: int) -> tuple[np.ndarray, np.ndarray]:
        """Vertices within `rings` rings of u (u first) and their ring numbers."""
        hit = self._hoods.get(u)
        if hit is not None:
            return hit
        ring = {u: 0}
        frontier = [u]
        for r in range(1, self.rings + 1):
            nxt = []
            for w in frontier:
                for x in self.adjacent[w]:
                    x = int(x)
                    if x not in ring:
                        ring[x] = r
                        nxt.append(x)
            frontier = nxt
        verts = np.asarray(sorted(ring, key=lambda x: (ring[x], x)), dtype=np.int64)
        out = (verts, np.asarray([ring[int(x)] for x in verts], dtype=np.int64))
        self._hoods[u] = out
        return out

    def ring(self, u: int, v: int) -> int:
        """Ring number of v around u, or rings + 1 beyond the neighbourhood."""
        verts, rings = self.neighbourhood(int(u))
        hit = np.flatnonzero(verts == int(v))
        return int(rings[hit[0]]) if hit.size else self.rings + 1

    def common_neighbours(self, u: int, v: int) -> int:
        return len(np.intersect1d(self.adjacent[int(u)], self.adjacent[int(v)]))

    def pair_coupling(self, u: int, v: int) -> float:
        """phi / theta of the edge joining u and v (adjacent vertices)."""
        k = self.edge_of.get((int(u), int(v)), self.edge_of.get((int(v), int(u))))
        return float(self.edge_coupling[k])


_RINGS: dict = {}


def mesh_rings(mesh) -> MeshRings:
    hit = _RINGS.get(id(mesh))
    if hit is None or hit.mesh is not mesh:
        hit = MeshRings(mesh)
        _RINGS[id(mesh)] = hit
    return hit


# ---------------------------------------------------------------- physics
def drude_profile(h_prime_km, beta_per_km, midpoint_altitudes_m):
    z = np.asarray(midpoint_altitudes_m) / 1000.0
    density = 1.43e13 * np.exp(-0.15 * h_prime_km) * np.exp((beta_per_km - 0.15) * (z - h_prime_km))
    return np.where(z >= 60.0, density, 0.0), 1.8e11 * np.exp(-0.15 * z)


def time_steps(h_prime_km, beta_per_km, mesh) -> np.ndarray:
    """time_step for arrays of (h', beta), one value per run (the same
    elementwise operations, evaluated for all runs at once)."""
    h = np.atleast_1d(np.asarray(h_prime_km, dtype=np.float64))[:, None]
    b = np.atleast_1d(np.asarray(beta_per_km, dtype=np.float64))[:, None]
    alt = _altitudes()
    radii = EARTH_RADIUS_M + alt
    mids = 0.5 * (radii[1:] + radii[:-1]) - EARTH_RADIUS_M
    smallest = float(radii.min())
    inv = (smallest * mesh.primal_edge_angles.min()) ** -2 + \
        (smallest * mesh.dual_edge_angles.min()) ** -2 + (2.0 / np.diff(radii).min()) ** 2
    geometric = 1.0 / (C_0 * np.sqrt(inv))
    density, _ = drude_profile(h, b, mids[None, :])
    wp = np.sqrt(density.max(axis=1) * ELECTRON_CHARGE_C ** 2 / ELECTRON_MASS_KG / EPSILON_0)
    with np.errstate(divide="ignore"):
        return 0.35 * np.where(wp > 0, np.minimum(geometric, 2.0 / wp), geometric)


def time_step(h_prime_km, beta_per_km, mesh) -> float:
    """The run's time step: courant 0.35 x min(geometric CFL, 2 / omega_p max)
    (the solver's rule, solver.py:377-406, for a vacuum-permittivity grid)."""
    alt = _altitudes()
    radii = EARTH_RADIUS_M + alt
    mids = 0.5 * (radii[1:] + radii[:-1]) - EARTH_RADIUS_M
    smallest = float(radii.min())
    inv = (smallest * mesh.primal_edge_angles.min()) ** -2 + \
        (smallest * mesh.dual_edge_angles.min()) ** -2 + (2.0 / np.diff(radii).min()) ** 2
    geometric = 1.0 / (C_0 * np.sqrt(inv))
    density, _ = drude_profile(h_prime_km, beta_per_km, mids)
    wp = np.sqrt(density.max() * ELECTRON_CHARGE_C ** 2 / ELECTRON_MASS_KG / EPSILON_0)
    return 0.35 * (min(geometric, 2.0 / wp) if wp > 0 else geometric)


def pair_features(physics: np.ndarray, dts: np.ndarray, u: np.ndarray, v: np.ndarray,
                  rings: MeshRings) -> np.ndarray:
    """(n, 16) FP32 features of G(u, v): the configuration (normalised h',
    beta, log sigma_g, carrier, log dt), the pair's great-circle distance and
    ring, the local geometry of both vertices (dual-cell area, mean edge
    length, mean Laplacian coupling phi/theta at u, pentagon count), the
    coupling of the edge joining a ring-1 pair or the common neighbours of a
    ring-2 pair, and the steps a wave needs to cover the distance."""
    p = np.asarray(physics, dtype=np.float64)
    u = np.asarray(u, dtype=np.int64)
    v = np.asarray(v, dtype=np.int64)
    x = np.zeros((p.shape[0], 16))
    x[:, 0] = (p[:, 0] - 74.0) / 6.0
    x[:, 1] = (p[:, 1] - 0.40) / 0.15
    x[:, 2] = (p[:, 2] + 2.5) / 1.5
    x[:, 3] = (p[:, 3] - 25.0) / 15.0
    x[:, 4] = np.log(np.asarray(dts) / 1e-7)
    dist = np.arccos(np.clip(np.sum(rings.unit[u] * rings.unit[v], axis=1), -1.0, 1.0)) \
        * EARTH_RADIUS_M
    ring = rings.rings_of(u, v)
    x[:, 5] = dist / 100e3
    x[:, 6] = ring == 1
    x[:, 7] = ring == 2
    x[:, 8] = (rings.coupling[u] / rings.mean_coupling - 1.0) * 100.0
    x[:, 9] = (rings.area[u] / rings.mean_area - 1.0) * 10.0
    x[:, 10] = (rings.area[v] / rings.mean_area - 1.0) * 10.0
    x[:, 11] = (rings.edge_length[u] / rings.mean_edge - 1.0) * 10.0
    x[:, 12] = (rings.edge_length[v] / rings.mean_edge - 1.0) * 10.0
    x[:, 13] = (rings.degree[u] == 5).astype(float) + (rings.degree[v] == 5)
    with np.errstate(invalid="ignore"):
        x[:, 14] = np.where(ring == 1, (rings.pair_couplings(u, v) / rings.mean_coupling - 1.0)
                            * 10.0,
                            np.where(ring == 2, rings.common_neighbours_of(u, v) / 2.0, 0.0))
    x[:, 15] = np.log1p(dist / (C_0 * np.asarray(dts)) / STEPS)
    return x.astype(np.float32)


# ------------------------------------------------------------------ FDTD
def build_run(mesh, h_prime_km, beta_per_km, log10_sigma, src_lat, src_lon, carrier_hz,
              dtype="float32") -> GeodesicFDTD:
    cfg = SimulationConfig(subdivision=SUBDIVISION, radial_cells=RADIAL_CELLS,
                           minimum_altitude_m=0.0, maximum_altitude_m=100e3,
                           earth_radius_m=EARTH_RADIUS_M, courant_factor=0.35,
                           radial_boundary_condition="surface-impedance")
    alt = _altitudes()
    radii = EARTH_RADIUS_M + alt
    mids = 0.5 * (radii[1:] + radii[:-1]) - EARTH_RADIUS_M
    density, collision = drude_profile(h_prime_km, beta_per_km, mids)
    plasma = electron_drude_model(mesh, mids, density, collision)
    material = EarthIonosphereMaterial(ionosphere_prefactor_hz=1e-6 / EPSILON_0,
                                       ionosphere_reference_height_m=70e3,
                                       ionosphere_scale_height_m=5e3)
    source = GaussianCurrent(latitude_deg=float(src_lat), longitude_deg=float(src_lon),
                             altitude_m=0.0, carrier_frequency_hz=float(carrier_hz))
    return GeodesicFDTD(cfg, material=material, source=source,
                        surface_impedance=ConductiveHalfSpaceSurface(10.0 ** float(log10_sigma)),
                        plasma=plasma, mesh=mesh, backend="hip", dtype=dtype,
                        compile_step=True, compile_chunk_size=128)


_RECEIVER_TABLES: dict = {}
_FACE_TREES: dict = {}


def _face_tree(mesh):
    hit = _FACE_TREES.get(id(mesh))
    if hit is None or hit[0] is not mesh:
        from scipy.spatial import cKDTree

        hit = (mesh, cKDTree(np.asarray(mesh.face_centers, dtype=np.float64)))
        _FACE_TREES[id(mesh)] = hit
    return hit[1]


def surface_tables(mesh, lats, lons) -> list:
    """Ground-level barycentric distributions (vertices, layer 0, weights) of
    many points at once, equal to sources.barycentric_distribution point by
    point: the containing face is searched among the 24 nearest face centres
    (a k-d tree) with the same edge-plane test, and points without a passing
    face there take the full per-point search."""
    from .sources import _face_planes, face_index, geographic_direction

    lats = np.asarray(lats, dtype=np.float64).ravel()
    lons = np.asarray(lons, dtype=np.float64).ravel()
    if lats.size == 0:
        return []
    dirs = np.stack([geographic_direction(a, b) for a, b in zip(lats, lons)])
    k = min(24, mesh.n_faces)
    _, near = _face_tree(mesh).query(dirs, k=k)
    near = np.sort(np.asarray(near).reshape(len(dirs), k), axis=1)
    planes = _face_planes(mesh)
    signs = np.stack([np.einsum("nkj,nj->nk", p[near], dirs) for p in planes], axis=2)
    inside = np.all(signs >= -1.0e-12, axis=2) | np.all(signs <= 1.0e-12, axis=2)
    # the face: the one passing candidate; among several (a point on a shared
    # edge or vertex) the one whose centre is nearest, as the per-point search
    # picks it; without a candidate the full per-point search
    counts = inside.sum(axis=1)
    face = np.empty(len(dirs), dtype=np.int64)
    one = counts == 1
    face[one] = near[one, np.argmax(inside[one], axis=1)]
    for i in np.flatnonzero(counts > 1):
        faces = near[i, np.flatnonzero(inside[i])]
        face[i] = int(faces[np.argmax(mesh.face_centers[faces] @ dirs[i])])
    for i in np.flatnonzero(counts == 0):
        face[i] = face_index(mesh, lats[i], lons[i])
    # barycentric weights of every point at once; the stacked 3-vector dot
    # products round as the per-point ones do, so the weights are bitwise
    # those of barycentric_distribution
    vertices = np.asarray(mesh.faces)[face]
    corners = np.asarray(mesh.vertices)[vertices]
    a, b, c = corners[:, 0], corners[:, 1], corners[:, 2]

    def dot(x, y):
        return np.matmul(x[:, None, :], y[:, :, None])[:, 0, 0]

    ab, ac = b - a, c - a
    normal = np.cross(ab, ac)
    point = dirs * (dot(normal, a) / dot(normal, dirs))[:, None]
    off = point - a
    d11, d12, d22 = dot(ab, ab), dot(ab, ac), dot(ac, ac)
    o1, o2 = dot(off, ab), dot(off, ac)
    # Python's float power (C pow), as the per-point formula squares d12;
    # an array ** 2 multiplies instead and can differ in the last bit
    den = d11 * d22 - np.asarray([x ** 2 for x in d12.tolist()])
    wb = (d22 * o1 - d12 * o2) / den
    wc = (d11 * o2 - d12 * o1) / den
    weights = np.clip(np.stack((1.0 - wb - wc, wb, wc), axis=1), 0.0, None)
    weights = weights / weights.sum(axis=1, keepdims=True)
    return [(vertices[i].copy(), 0, weights[i]) for i in range(len(dirs))]


def _receiver_table(mesh, altitudes, lat: float, lon: float):
    """Surface receiver distribution at (lat, lon), kept per mesh."""
    key = (id(mesh), float(lat), float(lon))
    hit = _RECEIVER_TABLES.get(key)
    if hit is not None and hit[0] is mesh:
        return hit[1]
    table = barycentric_distribution(mesh, altitudes, lat, lon, 0.0)
    _remember_table(mesh, key, table)
    return table


def _remember_table(mesh, key, table) -> None:
    _RECEIVER_TABLES[key] = (mesh, table)
    if len(_RECEIVER_TABLES) > 16384:
        _RECEIVER_TABLES.pop(next(iter(_RECEIVER_TABLES)))


def _prefetch_tables(mesh, points) -> None:
    """Fill the table cache for many (lat, lon) points with one batched search."""
    missing = []
    for lat, lon in points:
        key = (id(mesh), float(lat), float(lon))
        hit = _RECEIVER_TABLES.get(key)
        if hit is None or hit[0] is not mesh:
            missing.append(key)
    missing = list(dict.fromkeys(missing))
    if missing:
        tables = surface_tables(mesh, [k[1] for k in missing], [k[2] for k in missing])
        for key, table in zip(missing, tables):
            _remember_table(mesh, key, table)


def _run_ensembles(specs, mesh, batch, workers, progress=None):
    """Advance runs in ensembles of `batch`; specs: (config key, receiver
    table triple) per run, every table with the same receiver count.  The
    next batch is built on host threads while the current one runs; the
    ensembles then step by single launches, because a HIP graph capture on
    this thread fails when other threads allocate device memory meanwhile.
    Returns a list of (n_rx, SAMPLES) traces."""
    import torch

    # Initialise the device on this thread and let torch cache the device
    # count: its uncached query (through amdsmi) reports no devices from the
    # setup worker threads.
    torch.cuda.init()
    if torch.cuda.device_count() < 1:
        raise RuntimeError("no ROCm device is visible")

    def make(spec):
        key, _ = spec
        return build_run(mesh, key[0], key[1], key[2], key[3], key[4], key[5])

    out = [None] * len(specs)
    groups = [list(range(s, min(s + batch, len(specs)))) for s in range(0, len(specs), batch)]
    with ThreadPoolExecutor(max_workers=max(1, workers)) as pool:
        pending = pool.map(make, [specs[i] for i in groups[0]]) if groups else None
        for gi, group in enumerate(groups):
            built = list(pending)
            if gi + 1 < len(groups):
                pending = pool.map(make, [specs[i] for i in groups[gi + 1]])
            overlapped = gi + 1 < len(groups)
            ens = EnsembleFDTD(built, compile_chunk_size=1 if overlapped else 128,
                               autotune=False)
            traces = ens.record_er_observations([specs[i][1] for i in group], STEPS)
            for k, i in enumerate(group):
                out[i] = traces[k, STRIDE::STRIDE, :].T.astype(np.float64)
            del ens, built
            if progress is not None:
                progress(group[-1] + 1, len(specs))
    return out


def fdtd_traces(queries: np.ndarray, mesh=None, *, batch: int = 64, workers: int = 8,
                progress=None) -> np.ndarray:
    """Answer queries with FDTD: one run per distinct configuration, all of its
    receivers recorded, runs advanced in ensembles of `batch`.  Returns
    (n, SAMPLES) FP64 traces (E_r after steps 8, 16, ..., 4096)."""
    mesh = mesh if mesh is not None else build_geodesic_mesh(SUBDIVISION)
    q = np.asarray(queries, dtype=np.float64)
    keys: dict = {}
    order = []
    for i, row in enumerate(q):
        key = (row[0], row[1], row[2], row[3], row[4], row[7])
        if key not in keys:
            keys[key] = []
            order.append(key)
        keys[key].append(i)
    width = max(len(v) for v in keys.values())
    alt = _altitudes()
    _prefetch_tables(mesh, q[:, 5:7])
    specs = []
    for key in order:
        idx = keys[key]
        pts = [(q[i, 5], q[i, 6]) for i in idx]
        pts += [pts[0]] * (width - len(pts))  # equal receiver counts across the ensemble
        tables = [_receiver_table(mesh, alt, lat, lon) for lat, lon in pts]
        specs.append((key, (np.asarray([t[0] for t in tables]),
                            np.asarray([t[1] for t in tables]),
                            np.asarray([t[2] for t in tables]))))
    traces = _run_ensembles(specs, mesh, batch, workers, progress)
    out = np.zeros((q.shape[0], SAMPLES))
    for key, tr in zip(order, traces):
        for j, i in enumerate(keys[key]):
            out[i] = tr[j]
    return out


def pair_training_data(runs: np.ndarray, mesh, *, batch: int = 64, workers: int = 8,
                       progress=None):
    """G(u, v) for every vertex v within RINGS rings of each run's source
    vertex u (the source placed exactly at u, receivers read exactly at v).
    Returns (physics (P, 4), dt (P,), u (P,), v (P,), traces (P, SAMPLES))."""
    rings = mesh_rings(mesh)
    hoods = [rings.neighbourhood(int(r[4]))[0] for r in runs]
    width = max(len(h) for h in hoods)
    lat, lon = _latlon(rings.unit)
    specs = []
    for r, hood in zip(runs, hoods):
        u = int(r[4])
        verts = np.concatenate((hood, np.full(width - len(hood), u)))
        table = (np.repeat(verts[:, None], 3, axis=1), np.zeros(width, dtype=np.int64),
                 np.tile(np.array([1.0, 0.0, 0.0]), (width, 1)))
        specs.append(((r[0], r[1], r[2], lat[u], lon[u], r[3]), table))
    traces = _run_ensembles(specs, mesh, batch, workers, progress)
    phys, dts, us, vs, ts = [], [], [], [], []
    dt_cache: dict = {}
    for r, hood, tr in zip(runs, hoods, traces):
        key = (r[0], r[1])
        if key not in dt_cache:
            dt_cache[key] = time_step(r[0], r[1], mesh)
        n = len(hood)
        phys.append(np.repeat(r[None, :4], n, axis=0))
        dts.append(np.full(n, dt_cache[key]))
        us.append(np.full(n, int(r[4])))
        vs.append(hood)
        ts.append(tr[:n])
    return (np.concatenate(phys), np.concatenate(dts), np.concatenate(us), np.concatenate(vs),
            np.concatenate(ts))


# ------------------------------------------------------- query decomposition
def decompose(queries: np.ndarray, mesh):
    """Split queries into weighted vertex pairs.

    Returns (query index, u, v, weight s_j w_i) per pair with nonzero weight,
    and an (n,) flag: the receiver lies within MAX_DISTANCE_M of the source and
    every pair is within RINGS rings."""
    q = np.asarray(queries, dtype=np.float64)
    rings = mesh_rings(mesh)
    alt = _altitudes()
    qi, us, vs, ws = [], [], [], []
    covered = np.ones(q.shape[0], dtype=bool)
    src = _unit(q[:, 3], q[:, 4])
    rx = _unit(q[:, 5], q[:, 6])
    dist = np.arccos(np.clip(np.sum(src * rx, axis=1), -1.0, 1.0)) * EARTH_RADIUS_M
    covered &= dist <= MAX_DISTANCE_M * (1.0 + 1e-9)
    _prefetch_tables(mesh, np.concatenate((q[:, 3:5], q[:, 5:7])))
    n = q.shape[0]
    src = [_receiver_table(mesh, alt, row[3], row[4]) for row in q]
    rcv = [_receiver_table(mesh, alt, row[5], row[6]) for row in q]
    ks = max((len(t[0]) for t in src), default=1)
    kr = max((len(t[0]) for t in rcv), default=1)
    su, sw = np.zeros((n, ks), dtype=np.int64), np.zeros((n, ks))
    rv, rw = np.zeros((n, kr), dtype=np.int64), np.zeros((n, kr))
    for i in range(n):  # padding entries carry weight 0 and drop out below
        su[i, :len(src[i][0])], sw[i, :len(src[i][2])] = src[i][0], src[i][2]
        rv[i, :len(rcv[i][0])], rw[i, :len(rcv[i][2])] = rcv[i][0], rcv[i][2]
    # pairs in the order (query, source entry, receiver entry), zero weights dropped
    keep = ((sw != 0.0)[:, :, None] & (rw != 0.0)[:, None, :]).ravel()
    qi = np.repeat(np.arange(n, dtype=np.int64), ks * kr)[keep]
    us = np.broadcast_to(su[:, :, None], (n, ks, kr)).ravel()[keep]
    vs = np.broadcast_to(rv[:, None, :], (n, ks, kr)).ravel()[keep]
    ws = (sw[:, :, None] * rw[:, None, :]).ravel()[keep]
    far = rings.rings_of(us, vs) > rings.rings
    covered[np.unique(qi[far])] = False
    return (qi, us, vs, ws), covered


def in_range(queries: np.ndarray, mesh=None, *, with_pairs: bool = False):
    """Physics in the ensemble box, receiver within MAX_DISTANCE_M of the
    source, every weighted vertex pair within RINGS rings.

    With `with_pairs`, also returns decompose(queries[ok]) -- the pairs of
    the in-range queries, indexed among them -- which the range test has
    computed already (a query's pairs depend on that query alone)."""
    mesh = mesh if mesh is not None else build_geodesic_mesh(SUBDIVISION)
    q = np.asarray(queries, dtype=np.float64)
    ok = np.ones(q.shape[0], dtype=bool)
    for j, key in enumerate(("h_prime_km", "beta_per_km", "log10_ground_sigma")):
        lo, hi = BOX[key]
        ok &= (q[:, j] >= lo) & (q[:, j] <= hi)
    lo, hi = BOX["carrier_hz"]
    ok &= (q[:, 7] >= lo) & (q[:, 7] <= hi)
    ok &= np.all(np.isfinite(q), axis=1) & (np.abs(q[:, [3, 5]]) <= 90.0).all(axis=1)
    empty = np.zeros(0, dtype=np.int64)
    pairs = (empty, empty, empty, np.zeros(0))
    if ok.any():
        (qi, u, v, w), covered = decompose(q[ok], mesh)
        ok[np.flatnonzero(ok)] = covered
        keep = covered[qi]
        pairs = ((np.cumsum(covered) - 1)[qi[keep]], u[keep], v[keep], w[keep])
    return (ok, pairs) if with_pairs else ok


def normalise(traces: np.ndarray, floor: float = 1e-30):
    """(shape with unit RMS, natural-log RMS amplitude) of each trace."""
    t = np.asarray(traces, dtype=np.float64)
    rms = np.sqrt(np.mean(t * t, axis=1))
    rms = np.maximum(rms, floor)
    return t / rms[:, None], np.log(rms)


def trace_rel_l2(pred: np.ndarray, true: np.ndarray) -> np.ndarray:
    num = np.linalg.norm(pred - true, axis=1)
    den = np.linalg.norm(true, axis=1)
    return num / np.maximum(den, 1e-300)


# ---------------------------------------------------------------- networks
def amplitude_basis(features: np.ndarray) -> np.ndarray:
    """FP64 quadratic basis of the pair features (1, x_i, x_i x_j for i <= j):
    the host-side baseline of the log amplitude."""
    x = np.asarray(features, dtype=np.float64)
    cols = [np.ones(x.shape[0])] + [x[:, i] for i in range(x.shape[1])]
    for i in range(x.shape[1]):
        for j in range(i, x.shape[1]):
            cols.append(x[:, i] * x[:, j])
    return np.stack(cols, axis=1)


def fit_amplitude_baseline(features: np.ndarray, log_amp: np.ndarray,
                           ridge: float = 1e-6) -> np.ndarray:
    """Ridge least-squares coefficients of the log amplitude on amplitude_basis."""
    b = amplitude_basis(features)
    gram = b.T @ b
    gram += ridge * b.shape[0] * np.eye(b.shape[1])
    return np.linalg.solve(gram, b.T @ np.asarray(log_amp, dtype=np.float64))


@dataclass
class SurrogateModel:
    """One trained network for G(u, v).

    The pair traces' RMS amplitudes span ~17 decades across the box and the
    rings.  The log amplitude is therefore split into an FP64 host baseline
    (a quadratic least-squares fit on the pair features, `amp_coef`) and the
    network's residual, normalised by `amp_mean` / `amp_std`: the network's
    BF16 arithmetic then resolves a residual of ~0.1 natural-log units
    instead of the whole range."""

    net: object
    amp_mean: float
    amp_std: float
    amp_coef: np.ndarray | None = None

    def log_amplitude(self, features: np.ndarray, amp_n: np.ndarray,
                      basis: np.ndarray | None = None) -> np.ndarray:
        """`basis`: amplitude_basis(features) when the caller holds it already."""
        amp = np.asarray(amp_n, dtype=np.float64) * self.amp_std + self.amp_mean
        if self.amp_coef is not None:
            amp = amp + (amplitude_basis(features) if basis is None else basis) @ self.amp_coef
        return amp

    def predict_pairs(self, features: np.ndarray, basis: np.ndarray | None = None) -> np.ndarray:
        shape, amp_n = self.net.predict(features)
        amp = self.log_amplitude(features, amp_n, basis)
        g = shape.astype(np.float64)
        g *= np.exp(amp)[:, None]
        return g

    def save(self, path) -> None:
        params = self.net.host_parameters()
        extra = {} if self.amp_coef is None else {"amp_coef": self.amp_coef}
        np.savez(path, amp_mean=self.amp_mean, amp_std=self.amp_std, **extra,
                 **{f"param_{k}": v for k, v in params.items()})

    @classmethod
    def load(cls, path):
        from .surrogate import SurrogateConfig, TraceSurrogate

        with np.load(path, allow_pickle=False) as data:
            params = {k[len("param_"):]: data[k] for k in data.files if k.startswith("param_")}
            mean, std = float(data["amp_mean"]), float(data["amp_std"])
            coef = np.asarray(data["amp_coef"]) if "amp_coef" in data.files else None
        return cls(TraceSurrogate(SurrogateConfig(), params), mean, std, coef)


class SurrogateCommittee:
    """Independently trained networks answering queries through the vertex-pair
    decomposition; their disagreement is the router's uncertainty."""

    def __init__(self, members: list[SurrogateModel]):
        if not members:
            raise ValueError("a committee needs at least one member")
        self.members = list(members)

    @classmethod
    def load(cls, paths):
        return cls([SurrogateModel.load(p) for p in paths])

    def predict(self, queries: np.ndarray, mesh, pairs=None) -> tuple[np.ndarray, np.ndarray]:
        """(mean trace (n, SAMPLES), member disagreement (n,)) for in-range queries:
        the disagreement is the largest relative L2 distance of a member's
        trace from the committee mean.  `pairs`: decompose(queries)[0] when
        the caller holds it already (in_range(..., with_pairs=True))."""
        q = np.asarray(queries, dtype=np.float64)
        qi, u, v, w = decompose(q, mesh)[0] if pairs is None else pairs
        dts = time_steps(q[:, 0], q[:, 1], mesh)
        feats = pair_features(q[qi][:, [0, 1, 2, 7]], dts[qi], u, v, mesh_rings(mesh))
        # the members share the features, so the amplitude baseline's basis too
        basis = amplitude_basis(feats) if any(m.amp_coef is not None for m in self.members) \
            else None
        # Each query's trace is the sum of its pairs' weighted traces.  The
        # sparse (queries x pairs) product adds every query's pairs in pair
        # order, as np.add.at does (bit for bit), at a third of its cost
        # (0.07 -> 0.024 s per member for 2000 queries).
        from scipy.sparse import csr_matrix

        scatter = csr_matrix((np.ones(len(qi)), (qi, np.arange(len(qi)))),
                             shape=(q.shape[0], len(qi)))
        per_member = []
        for m in self.members:
            g = m.predict_pairs(feats, basis)
            g *= w[:, None]
            per_member.append(np.asarray(scatter @ g))
        mean = np.mean(per_member, axis=0)
        spread = np.max([trace_rel_l2(t, mean) for t in per_member], axis=0)
        return mean, spread


def train(features: np.ndarray, traces: np.ndarray, *, steps: int = 6000, batch: int = 256,
          lr: float = 2e-3, seed: int = 0, log=None, validation=None) -> SurrogateModel:
    """Fit one network to (features, traces) with Adam and a cosine schedule.
    `validation` is (queries, FDTD traces, mesh): query-level errors of the
    network through the pair decomposition are logged during training."""
    import torch

    from .surrogate import TraceSurrogate

    shape, amp = normalise(traces)
    coef = fit_amplitude_baseline(features, amp)
    amp = amp - amplitude_basis(features) @ coef
    mean, std = float(amp.mean()), float(amp.std() + 1e-12)
    net = TraceSurrogate(seed=seed)
    dev = net.device
    x = torch.as_tensor(np.asarray(features, dtype=np.float32), device=dev)
    s = torch.as_tensor(shape.astype(np.float32), device=dev)
    a = torch.as_tensor(((amp - mean) / std).astype(np.float32), device=dev)
    rng = np.random.default_rng(seed)
    n = x.shape[0]
    for step in range(steps):
        idx = torch.as_tensor(rng.integers(0, n, batch), device=dev)
        rate = lr * 0.5 * (1.0 + np.cos(np.pi * step / steps)) + 1e-6
        loss = net.train_step(x.index_select(0, idx), s.index_select(0, idx),
                              a.index_select(0, idx), rate)
        if log is not None and (step % 1000 == 0 or step == steps - 1):
            msg = ""
            if validation is not None and (step % 5000 == 0 or step == steps - 1):
                vq, vt, vmesh = validation
                pred, _ = SurrogateCommittee([SurrogateModel(net, mean, std, coef)]).predict(
                    vq, vmesh)
                err = trace_rel_l2(pred, vt)
                msg = f" validation median {np.median(err):.4f} p90 {np.quantile(err, 0.9):.4f}"
            log(step, loss, msg)
    return SurrogateModel(net, mean, std, coef)


class TraceQueryService:
    """Answer trace queries with the surrogate committee, FDTD staying the arbiter of error.

    In-range queries get the committee's mean trace unless (a) its members
    disagree by more than `disagreement` -- that query goes to FDTD -- or (b)
    the batch's spot check fails: `spot_checks` configurations of the batch,
    chosen at random among the confident queries, are simulated and every
    confident query of those configurations is compared with FDTD; any error
    above `tolerance` sends the whole batch to FDTD.  Out-of-range queries
    always go to FDTD.
    """

    def __init__(self, model: SurrogateCommittee, mesh=None, *, tolerance: float = 0.02,
                 disagreement: float = 0.01, spot_checks: int = 1, seed: int = 0):
        if isinstance(model, SurrogateModel):
            model = SurrogateCommittee([model])
        self.model = model
        self.mesh = mesh if mesh is not None else build_geodesic_mesh(SUBDIVISION)
        self.tolerance = tolerance
        self.disagreement = disagreement
        self.spot_checks = spot_checks
        self.rng = np.random.default_rng(seed)

    def answer(self, queries: np.ndarray):
        q = np.asarray(queries, dtype=np.float64)
        t0 = time.perf_counter()
        out = np.zeros((q.shape[0], SAMPLES))
        served = np.full(q.shape[0], "fdtd", dtype=object)
        ok, pairs = in_range(q, self.mesh, with_pairs=True)
        report = {"queries": int(q.shape[0]), "in_range": int(ok.sum())}
        fdtd_rows = np.flatnonzero(~ok)
        rows = np.zeros(0, dtype=np.int64)
        check = np.zeros(0, dtype=bool)
        pred = None
        if ok.any():
            rows = np.flatnonzero(ok)
            t1 = time.perf_counter()
            pred, spread = self.model.predict(q[rows], self.mesh, pairs)
            report["surrogate_seconds"] = time.perf_counter() - t1
            confident = spread <= self.disagreement
            report["uncertain"] = int(np.sum(~confident))
            fdtd_rows = np.concatenate((fdtd_rows, rows[~confident]))
            rows, pred = rows[confident], pred[confident]
            if rows.size:
                keys = [tuple(q[i, [0, 1, 2, 3, 4, 7]]) for i in rows]
                distinct = sorted(set(keys))
                picks = self.rng.choice(len(distinct), size=min(self.spot_checks, len(distinct)),
                                        replace=False)
                chosen = {distinct[p] for p in picks}
                check = np.asarray([k in chosen for k in keys])
        # One FDTD pass for the out-of-range and uncertain queries and the
        # spot-checked configurations (one set of ensembles instead of two).
        first = np.concatenate((fdtd_rows, rows[check])).astype(np.int64)
        if first.size:
            t3 = time.perf_counter()
            out[first] = fdtd_traces(q[first], self.mesh)
            report["fdtd_seconds"] = time.perf_counter() - t3
        if rows.size:
            errors = trace_rel_l2(pred[check], out[rows[check]])
            report["spot_check_errors"] = errors.tolist()
            if np.all(errors <= self.tolerance):
                out[rows[~check]] = pred[~check]
                served[rows[~check]] = "surrogate"
            else:
                rest = rows[~check]
                report["fallback"] = "spot check failed"
                if rest.size:
                    t4 = time.perf_counter()
                    out[rest] = fdtd_traces(q[rest], self.mesh)
                    report["fdtd_seconds"] = report.get("fdtd_seconds", 0.0) + \
                        time.perf_counter() - t4
        report["seconds"] = time.perf_counter() - t0
        report["served_by_surrogate"] = int(np.sum(served == "surrogate"))
        return out, served, report


def _main(argv=None) -> int:
    """Generate pair training data on the GPU, train one committee member, save it."""
    import argparse
    import json
    import sys

    ap = argparse.ArgumentParser(prog="python -m ionosphere_fdtd_amd.trace_queries")
    ap.add_argument("--runs", type=int, default=10000)
    ap.add_argument("--seed", type=int, default=20261205)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", required=True)
    ap.add_argument("--validation-configs", type=int, default=64)
    ap.add_argument("--max-seconds", type=float, default=None,
                    help="stop generating runs once this much time has passed (the runs "
                         "generated so far are used)")
    args = ap.parse_args(argv)
    mesh = build_geodesic_mesh(SUBDIVISION)
    t0 = time.perf_counter()
    runs = training_runs(args.runs, args.seed, mesh)
    if args.max_seconds is not None:
        # time a first slice to size the rest of the generation
        probe = pair_training_data(runs[:256], mesh)
        rate = 256 / (time.perf_counter() - t0)
        keep = max(256, min(args.runs, int(rate * args.max_seconds)))
        print(f"{rate:.1f} runs/s: generating {keep} runs", file=sys.stderr, flush=True)
        rest = pair_training_data(runs[256:keep], mesh, progress=lambda d, n: (
            d % 1024 == 0 and print(f"fdtd {256 + d}/{keep} runs, "
                                    f"{time.perf_counter() - t0:.0f} s", file=sys.stderr,
                                    flush=True))) if keep > 256 else None
        data = probe if rest is None else tuple(np.concatenate((a, b)) for a, b in zip(probe, rest))
        runs = runs[:keep]
    else:
        data = pair_training_data(runs, mesh, progress=lambda d, n: (
            d % 1024 == 0 and print(f"fdtd {d}/{n} runs, {time.perf_counter() - t0:.0f} s",
                                    file=sys.stderr, flush=True)))
    phys, dts, us, vs, traces = data
    gen = time.perf_counter() - t0
    feats = pair_features(phys, dts, us, vs, mesh_rings(mesh))
    vq = random_queries(args.validation_configs, 8, seed=args.seed + 1)
    vt = fdtd_traces(vq, mesh)
    t1 = time.perf_counter()
    model = train(feats, traces, steps=args.steps, batch=args.batch, lr=args.lr,
                  seed=args.seed % 1000, validation=(vq, vt, mesh),
                  log=lambda s, l, m: print(f"train step {s} loss {l:.5f}{m}", file=sys.stderr,
                                            flush=True))
    fit = time.perf_counter() - t1
    model.save(args.out)
    pred, _ = SurrogateCommittee([model]).predict(vq, mesh)
    err = trace_rel_l2(pred, vt)
    print(json.dumps({"runs": int(len(runs)), "pairs": int(len(traces)),
                      "generate_s": round(gen, 1), "train_s": round(fit, 1),
                      "validation_median": float(np.median(err)),
                      "validation_p90": float(np.quantile(err, 0.9)),
                      "validation_p99": float(np.quantile(err, 0.99)),
                      "validation_within_2pct": float(np.mean(err <= 0.02))}))
    return 0


if __name__ == "__main__":
    raise SystemExit(_main())
