"""MI355X-native GeodesicFDTD time-step loop (HIP/gfx950).

Drop-in for the time-step path of ruddyscent/ionosphere-fdtd: the same
``SimulationConfig`` / ``GeodesicFDTD`` / source / material / surface
impedance / plasma / checkpoint API, with the H and E updates, exponential
conductive integration, plasma ADE current, surface-impedance boundary,
Gaussian current injection and receiver sampling executed by hand-written
gfx950 kernels in ``libgeofdtd.so`` (C ABI in include/geofdtd.h).
"""

from .backends import ArrayBackend, BackendUnavailableError, create_backend
from .constants import C_0, EARTH_RADIUS_M, EPSILON_0, MU_0
from .ensemble import EnsembleFDTD
from .materials import EarthIonosphereMaterial, SphericalAnomaly
from .mesh import GeodesicMesh, build_geodesic_mesh, build_geodesic_mesh_from_topology
from .physics_diagnostics import PhysicsDiagnosticSampler, PhysicsSnapshot
from .plasma import ColdPlasmaSpecies, MeshPlasmaModel, electron_drude_model
from .radial_grid import (RadialGridValidation, RadialRefinementRegion,
                          build_refined_radial_grid, validate_radial_grid)
from .solver import GeodesicFDTD, SimulationConfig
from .sources import GaussianCurrent, TangentialGaussianCurrent
from .surface_impedance import ConductiveHalfSpaceSurface

__all__ = [
    "ArrayBackend", "BackendUnavailableError", "C_0", "ColdPlasmaSpecies",
    "ConductiveHalfSpaceSurface", "EARTH_RADIUS_M", "EPSILON_0", "EarthIonosphereMaterial",
    "EnsembleFDTD",
    "GaussianCurrent", "GeodesicFDTD", "GeodesicMesh", "MU_0", "MeshPlasmaModel",
    "PhysicsDiagnosticSampler", "PhysicsSnapshot",
    "RadialGridValidation", "RadialRefinementRegion", "SimulationConfig", "SphericalAnomaly",
    "TangentialGaussianCurrent", "build_geodesic_mesh", "build_geodesic_mesh_from_topology",
    "build_refined_radial_grid", "create_backend", "electron_drude_model",
    "validate_radial_grid",
]
