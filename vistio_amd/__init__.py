"""vistio_amd — Vistio's Projective Dynamics step on MI355X (gfx950 HIP).

The numerical path lives in libvistio_hip.so (vistio_amd/csrc, C ABI in
include/vistio_hip.h); this package is the host-side mirror of the
reference's solver interface.
"""
from .solver import (FabricProperties, GpuError, HipPdSolver, InvalidConfig, InvalidMesh, IpcConfig,
                     MaterialDatabase, SimulationState, SolverConfig, StepResult, Topology, TriangleMesh,
                     VistioError, estimate_initial_kappa, quad_grid)
from .backend import ComputeBuffer, HipBackend
from .narrow import CandidatePair, HipVertexTriangleTest

__all__ = ["CandidatePair", "ComputeBuffer", "FabricProperties", "GpuError", "HipBackend", "HipPdSolver", "HipVertexTriangleTest", "InvalidConfig",
           "InvalidMesh", "IpcConfig", "MaterialDatabase", "SimulationState", "SolverConfig", "StepResult", "Topology", "TriangleMesh",
           "VistioError", "estimate_initial_kappa", "quad_grid"]
