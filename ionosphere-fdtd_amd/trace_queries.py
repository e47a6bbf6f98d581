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
CFL limit of each configuration).

In 4096 steps (0.02-0.8 ms) a signal covers at most ~240 km, so receivers
drawn anywhere on the sphere would see nothing; queries therefore come from a
station network -- eight transmitter sites, each with eight receivers within
250 km.  A query is *in range* when its physical parameters lie in the
ensemble box (h' 68-80 km, beta 0.25-0.55 /km, ground sigma 1e-4..1e-1 S/m,
carrier 10-40 Hz) and its source/receiver pair is a station pair; the router
answers other queries with FDTD.
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
TRANSMITTERS = ((35.15, 126.85), (46.5, -90.9), (-33.9, 18.4), (64.8, -147.7),
                (-12.0, -77.0), (28.6, 77.2), (51.5, -0.1), (-41.3, 174.8))
STATION_OFFSETS = ((0.0, 0.0), (30.0, 45.0), (60.0, 90.0), (90.0, 135.0), (120.0, 180.0),
                   (160.0, 225.0), (200.0, 270.0), (250.0, 315.0))  # (km, azimuth deg)
COLUMNS = ("h_prime_km", "beta_per_km", "log10_ground_sigma", "src_lat", "src_lon", "rx_lat",
           "rx_lon", "carrier_hz")
ELECTRON_CHARGE_C = 1.602_176_634e-19
ELECTRON_MASS_KG = 9.109_383_713_9e-31


def stations() -> np.ndarray:
    """(transmitter, station, lat/lon) receiver positions."""
    if _STATIONS is None:
        out = np.zeros((len(TRANSMITTERS), len(STATION_OFFSETS), 2))
        for t, (lat, lon) in enumerate(TRANSMITTERS):
            for s, (d_km, az) in enumerate(STATION_OFFSETS):
                out[t, s] = destination_point(lat, lon, az, d_km * 1e3, EARTH_RADIUS_M)
        globals()["_STATIONS"] = out
    return _STATIONS.copy()


_STATIONS = None


def sample_runs(n: int, seed: int) -> np.ndarray:
    """(n, 5) runs: h', beta, log10 sigma_g, carrier, transmitter index."""
    rng = np.random.default_rng(seed)
    out = np.empty((n, 5))
    out[:, 0] = rng.uniform(*BOX["h_prime_km"], n)
    out[:, 1] = rng.uniform(*BOX["beta_per_km"], n)
    out[:, 2] = rng.uniform(*BOX["log10_ground_sigma"], n)
    out[:, 3] = rng.uniform(*BOX["carrier_hz"], n)
    out[:, 4] = rng.integers(0, len(TRANSMITTERS), n)
    return out


def station_queries(runs: np.ndarray) -> np.ndarray:
    """(n_runs * 8, 8) queries: every station of each run's transmitter."""
    st = stations()
    rows = []
    for h, b, ls, f, t in runs:
        tx = TRANSMITTERS[int(t)]
        for s in range(len(STATION_OFFSETS)):
            rows.append((h, b, ls, tx[0], tx[1], st[int(t), s, 0], st[int(t), s, 1], f))
    return np.asarray(rows, dtype=np.float64)


def _unit(lat, lon):
    la, lo = np.deg2rad(lat), np.deg2rad(lon)
    return np.stack((np.cos(la) * np.cos(lo), np.cos(la) * np.sin(lo), np.sin(la)), axis=-1)


def _altitudes():
    return np.linspace(0.0, 100e3, RADIAL_CELLS + 1)


def drude_profile(h_prime_km, beta_per_km, midpoint_altitudes_m):
    z = np.asarray(midpoint_altitudes_m) / 1000.0
    density = 1.43e13 * np.exp(-0.15 * h_prime_km) * np.exp((beta_per_km - 0.15) * (z - h_prime_km))
    return np.where(z >= 60.0, density, 0.0), 1.8e11 * np.exp(-0.15 * z)


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


def encode(queries: np.ndarray, mesh) -> np.ndarray:
    """(n, 8) queries -> (n, 16) FP32 features."""
    q = np.asarray(queries, dtype=np.float64)
    x = np.zeros((q.shape[0], 16))
    x[:, 0] = (q[:, 0] - 74.0) / 6.0
    x[:, 1] = (q[:, 1] - 0.40) / 0.15
    x[:, 2] = (q[:, 2] + 2.5) / 1.5
    x[:, 3] = (q[:, 7] - 25.0) / 15.0
    src, rx = _unit(q[:, 3], q[:, 4]), _unit(q[:, 5], q[:, 6])
    x[:, 4:7] = src
    x[:, 7:10] = rx
    x[:, 10:13] = (rx - src) * (EARTH_RADIUS_M / 100e3)
    x[:, 13] = np.arccos(np.clip(np.sum(src * rx, axis=1), -1.0, 1.0)) * EARTH_RADIUS_M / 100e3
    dts = {}
    for i, (h, b) in enumerate(q[:, :2]):
        key = (h, b)
        if key not in dts:
            dts[key] = time_step(h, b, mesh)
        x[i, 14] = np.log(dts[key] / 1e-7)
    return x.astype(np.float32)


def in_range(queries: np.ndarray, tolerance_deg: float = 1e-6) -> np.ndarray:
    q = np.asarray(queries, dtype=np.float64)
    ok = np.ones(q.shape[0], dtype=bool)
    for j, key in enumerate(("h_prime_km", "beta_per_km", "log10_ground_sigma")):
        lo, hi = BOX[key]
        ok &= (q[:, j] >= lo) & (q[:, j] <= hi)
    lo, hi = BOX["carrier_hz"]
    ok &= (q[:, 7] >= lo) & (q[:, 7] <= hi)
    st = stations()
    pair = np.zeros(q.shape[0], dtype=bool)
    for t, (lat, lon) in enumerate(TRANSMITTERS):
        at_tx = (np.abs(q[:, 3] - lat) <= tolerance_deg) & (np.abs(q[:, 4] - lon) <= tolerance_deg)
        for s in range(st.shape[1]):
            pair |= at_tx & (np.abs(q[:, 5] - st[t, s, 0]) <= tolerance_deg) & \
                (np.abs(q[:, 6] - st[t, s, 1]) <= tolerance_deg)
    return ok & pair


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


def _receiver_table(mesh, altitudes, lat: float, lon: float):
    """Surface receiver distribution at (lat, lon), kept per mesh: the station
    network repeats the same few points in every query batch."""
    key = (id(mesh), float(lat), float(lon))
    hit = _RECEIVER_TABLES.get(key)
    if hit is not None and hit[0] is mesh:
        return hit[1]
    table = barycentric_distribution(mesh, altitudes, lat, lon, 0.0)
    _RECEIVER_TABLES[key] = (mesh, table)
    if len(_RECEIVER_TABLES) > 4096:
        _RECEIVER_TABLES.pop(next(iter(_RECEIVER_TABLES)))
    return table


def fdtd_traces(queries: np.ndarray, mesh=None, *, batch: int = 64, workers: int = 8,
                progress=None) -> np.ndarray:
    """Answer queries with FDTD: one run per distinct configuration, all of its
    receivers recorded, runs advanced in ensembles of `batch`.  Returns
    (n, SAMPLES) FP64 traces (E_r after steps 8, 16, ..., 4096)."""
    mesh = mesh if mesh is not None else build_geodesic_mesh(SUBDIVISION)
    q = np.asarray(queries, dtype=np.float64)
    keys = {}
    order = []
    for i, row in enumerate(q):
        key = (row[0], row[1], row[2], row[3], row[4], row[7])
        if key not in keys:
            keys[key] = []
            order.append(key)
        keys[key].append(i)
    width = max(len(v) for v in keys.values())
    out = np.zeros((q.shape[0], SAMPLES))
    alt = _altitudes()

    def make(key):
        sim = build_run(mesh, key[0], key[1], key[2], key[3], key[4], key[5])
        idx = keys[key]
        pts = [(q[i, 5], q[i, 6]) for i in idx]
        pts += [pts[0]] * (width - len(pts))  # equal receiver counts across the ensemble
        tables = [_receiver_table(mesh, alt, lat, lon) for lat, lon in pts]
        rx = (np.asarray([t[0] for t in tables]), np.asarray([t[1] for t in tables]),
              np.asarray([t[2] for t in tables]))
        return sim, rx

    import torch

    # Initialise the device on this thread and let torch cache the device
    # count: its uncached query (through amdsmi) reports no devices from the
    # setup worker threads.
    torch.cuda.init()
    if torch.cuda.device_count() < 1:
        raise RuntimeError("no ROCm device is visible")
    with ThreadPoolExecutor(max_workers=max(1, min(workers, len(order)))) as pool:
        for start in range(0, len(order), batch):
            group = order[start:start + batch]
            built = list(pool.map(make, group))
            ens = EnsembleFDTD([b[0] for b in built], compile_chunk_size=128, autotune=False)
            traces = ens.record_er_observations([b[1] for b in built], STEPS)
            for k, key in enumerate(group):
                for j, i in enumerate(keys[key]):
                    out[i] = traces[k, STRIDE::STRIDE, j]
            del ens, built
            if progress is not None:
                progress(min(start + batch, len(order)), len(order))
    return out


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


@dataclass
class SurrogateModel:
    """A trained network with its amplitude normalisation."""

    net: object
    amp_mean: float
    amp_std: float

    def predict(self, queries: np.ndarray, mesh) -> np.ndarray:
        shape, amp_n = self.net.predict(encode(queries, mesh))
        amp = amp_n.astype(np.float64) * self.amp_std + self.amp_mean
        return shape.astype(np.float64) * np.exp(amp)[:, None]

    def save(self, path) -> None:
        params = self.net.host_parameters()
        np.savez(path, amp_mean=self.amp_mean, amp_std=self.amp_std,
                 **{f"param_{k}": v for k, v in params.items()})

    @classmethod
    def load(cls, path):
        from .surrogate import SurrogateConfig, TraceSurrogate

        with np.load(path, allow_pickle=False) as data:
            params = {k[len("param_"):]: data[k] for k in data.files if k.startswith("param_")}
            mean, std = float(data["amp_mean"]), float(data["amp_std"])
        return cls(TraceSurrogate(SurrogateConfig(), params), mean, std)


def train(features: np.ndarray, traces: np.ndarray, *, steps: int = 6000, batch: int = 256,
          lr: float = 2e-3, seed: int = 0, log=None, validation=None) -> SurrogateModel:
    """Fit the surrogate to (features, traces) with Adam and a cosine schedule."""
    import torch

    from .surrogate import TraceSurrogate

    shape, amp = normalise(traces)
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
        if log is not None and (step % 500 == 0 or step == steps - 1):
            msg = ""
            if validation is not None and (step % 2000 == 0 or step == steps - 1):
                vq, vt, vmesh = validation
                err = trace_rel_l2(SurrogateModel(net, mean, std).predict(vq, vmesh), vt)
                msg = f" validation median {np.median(err):.4f} p90 {np.quantile(err, 0.9):.4f}"
            log(step, loss, msg)
    return SurrogateModel(net, mean, std)


class TraceQueryService:
    """Answer trace queries with the surrogate, FDTD staying the judge.

    In-range queries get the surrogate's trace unless the batch's spot check
    fails: one configuration of the batch, chosen at random, is simulated and
    every in-range query of that configuration is compared with FDTD; any
    error above `tolerance` sends the whole batch to FDTD.  Out-of-range
    queries always go to FDTD.
    """

    def __init__(self, model: SurrogateModel, mesh=None, *, tolerance: float = 0.02,
                 spot_checks: int = 1, seed: int = 0):
        self.model = model
        self.mesh = mesh if mesh is not None else build_geodesic_mesh(SUBDIVISION)
        self.tolerance = tolerance
        self.spot_checks = spot_checks
        self.rng = np.random.default_rng(seed)

    def answer(self, queries: np.ndarray):
        q = np.asarray(queries, dtype=np.float64)
        t0 = time.perf_counter()
        out = np.zeros((q.shape[0], SAMPLES))
        served = np.full(q.shape[0], "fdtd", dtype=object)
        ok = in_range(q)
        report = {"queries": int(q.shape[0]), "in_range": int(ok.sum())}
        fdtd_rows = np.flatnonzero(~ok)
        if ok.any():
            rows = np.flatnonzero(ok)
            t1 = time.perf_counter()
            pred = self.model.predict(q[rows], self.mesh)
            report["surrogate_seconds"] = time.perf_counter() - t1
            keys = [tuple(q[i, [0, 1, 2, 3, 4, 7]]) for i in rows]
            distinct = sorted(set(keys))
            picks = self.rng.choice(len(distinct), size=min(self.spot_checks, len(distinct)),
                                    replace=False)
            chosen = {distinct[p] for p in picks}
            check = np.asarray([k in chosen for k in keys])
            t2 = time.perf_counter()
            truth = fdtd_traces(q[rows[check]], self.mesh, batch=64)
            report["spot_check_seconds"] = time.perf_counter() - t2
            errors = trace_rel_l2(pred[check], truth)
            report["spot_check_errors"] = errors.tolist()
            if np.all(errors <= self.tolerance):
                out[rows] = pred
                out[rows[check]] = truth
                served[rows] = "surrogate"
                served[rows[check]] = "fdtd"
            else:
                fdtd_rows = np.concatenate((fdtd_rows, rows))
                report["fallback"] = "spot check failed"
        if fdtd_rows.size:
            out[fdtd_rows] = fdtd_traces(q[fdtd_rows], self.mesh)
        report["seconds"] = time.perf_counter() - t0
        report["served_by_surrogate"] = int(np.sum(served == "surrogate"))
        return out, served, report


def _main(argv=None) -> int:
    """Generate C2-style training data on the GPU, train, save the model."""
    import argparse
    import json
    import sys

    ap = argparse.ArgumentParser(prog="python -m ionosphere_fdtd_amd.trace_queries")
    ap.add_argument("--runs", type=int, default=8192)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", required=True)
    ap.add_argument("--validation-runs", type=int, default=128)
    args = ap.parse_args(argv)
    mesh = build_geodesic_mesh(SUBDIVISION)
    t0 = time.perf_counter()
    runs = sample_runs(args.runs + args.validation_runs, args.seed)
    queries = station_queries(runs)

    def progress(done, total):
        print(f"fdtd {done}/{total} runs, {time.perf_counter() - t0:.0f} s", file=sys.stderr,
              flush=True)

    traces = fdtd_traces(queries, mesh, progress=progress)
    gen = time.perf_counter() - t0
    feats = encode(queries, mesh)
    n_train = args.runs * len(STATION_OFFSETS)
    t1 = time.perf_counter()
    model = train(feats[:n_train], traces[:n_train], steps=args.steps, batch=args.batch,
                  lr=args.lr, seed=1, validation=(queries[n_train:], traces[n_train:], mesh),
                  log=lambda s, l, m: print(f"train step {s} loss {l:.5f}{m}", file=sys.stderr,
                                            flush=True))
    fit = time.perf_counter() - t1
    model.save(args.out)
    err = trace_rel_l2(model.predict(queries[n_train:], mesh), traces[n_train:])
    print(json.dumps({"runs": args.runs, "traces": n_train, "generate_s": round(gen, 1),
                      "train_s": round(fit, 1), "validation_median": float(np.median(err)),
                      "validation_p90": float(np.quantile(err, 0.9)),
                      "validation_within_2pct": float(np.mean(err <= 0.02))}))
    return 0


if __name__ == "__main__":
    raise SystemExit(_main())
