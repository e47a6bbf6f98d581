"""Golden vectors of the reference's material sampling modes (run in the build
container only; the reference is not available where the tests run).

    PYTHONPATH=/root/reference/src python tests/golden/make_material_fixtures.py

Writes materials_s2.npz: for every case below, the reference GeodesicFDTD's
(NumPy backend) sigma / epsilon_r arrays on the Er (vertex x node) and Et
(edge x midpoint) grids, its time step, and, for the conservative-nearest
anomaly mode, the horizontal anomaly fractions.  Data only; these pin the
dual-cell, edge-diamond and anomaly sampling of the product's host setup,
which no closed form or trajectory fixture covers.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, "/root/reference/src")

from ionosphere_fdtd.materials import EarthIonosphereMaterial, SphericalAnomaly  # noqa: E402
from ionosphere_fdtd.solver import GeodesicFDTD, SimulationConfig  # noqa: E402

# (name, radial support, tangential support, anomaly mode)
CASES = (
    ("point", "point", "point", "point"),
    ("dual_cell", "dual-cell", "point", "point"),
    ("edge_diamond", "point", "edge-diamond", "point"),
    ("conservative", "point", "point", "conservative-nearest"),
    ("all", "dual-cell", "edge-diamond", "conservative-nearest"),
)
GRID = dict(subdivision=2, radial_cells=12, minimum_altitude_m=-20e3, maximum_altitude_m=100e3)


def anomalies():
    """Two crust anomalies: a conductive, higher-permittivity cap limited to a
    resistive background, and a resistive cap with its own target area that
    reaches into the air layers."""
    return (
        SphericalAnomaly(latitude_deg=35.15, longitude_deg=126.85, radius_m=900e3,
                         altitude_min_m=-15e3, altitude_max_m=-2e3, conductivity_factor=40.0,
                         relative_permittivity=25.0,
                         maximum_background_conductivity_s_m=5e-3),
        SphericalAnomaly(latitude_deg=-20.0, longitude_deg=-60.0, radius_m=1500e3,
                         altitude_min_m=-20e3, altitude_max_m=5e3, conductivity_factor=0.1,
                         relative_permittivity=4.0, target_area_m2=5.0e12),
    )


def main() -> None:
    out = {}
    material = EarthIonosphereMaterial(anomalies=anomalies())
    for name, radial, tangential, mode in CASES:
        cfg = SimulationConfig(**GRID, radial_material_support=radial,
                               tangential_material_support=tangential,
                               horizontal_anomaly_mode=mode)
        sim = GeodesicFDTD(cfg, material=material, backend="numpy", dtype="float64")
        out[f"{name}/sigma_er"] = sim.sigma_er
        out[f"{name}/epsilon_r_er"] = sim.epsilon_r_er
        out[f"{name}/sigma_et"] = sim.sigma_et
        out[f"{name}/epsilon_r_et"] = sim.epsilon_r_et
        out[f"{name}/time_step_s"] = np.float64(sim.time_step_s)
        for fam in ("er", "et"):
            frac = getattr(sim, f"anomaly_horizontal_fractions_{fam}", None)
            if frac is not None:
                out[f"{name}/fractions_{fam}"] = np.stack(frac)
        print(name, sim.time_step_s, float(sim.sigma_er.sum()), float(sim.sigma_et.sum()))
    np.savez_compressed(HERE / "materials_s2.npz", **out)


if __name__ == "__main__":
    main()
