"""endobeams_amd — MI355X-native hot path of EndoBeams.jl behind a C ABI.

Python mirror of the reference's user interface (NodesBeams, Beams, BeamsConfiguration,
SimulationParams, RigidInteraction, run_simulation!) over libendobeams_hip.so.
"""
from . import abi
from .model import (Beams, BeamElementSurface, BeamsConfiguration, BoundaryConditions, ConcentratedForce,
                    DiscreteSignedDistanceField, Encastre, ImposedDisplacement, InteractionProperties, Loads,
                    NodesBeams, RigidInteraction, SimulationParams, SphereSurface, read_vtk_sdf)
from .engine import BatchEngine, EBError, Engine, HipBatchEngine, HipEngine, LIB_PATH, load_hip_library
from .simulation import RunStats, SimulationAborted, Stepper, run_simulation
from . import configs

__all__ = [
    "abi", "Beams", "BeamElementSurface", "BeamsConfiguration", "BoundaryConditions", "ConcentratedForce",
    "DiscreteSignedDistanceField", "Encastre", "ImposedDisplacement", "InteractionProperties", "Loads",
    "NodesBeams", "RigidInteraction", "SimulationParams", "SphereSurface", "read_vtk_sdf",
    "BatchEngine", "EBError", "Engine", "HipBatchEngine", "HipEngine", "LIB_PATH", "load_hip_library",
    "RunStats", "SimulationAborted", "Stepper", "run_simulation", "configs",
]
