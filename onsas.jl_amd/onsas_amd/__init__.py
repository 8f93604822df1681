"""onsas_amd -- MI355X engine for the nonlinear static Newton path of ONSAS.jl.

Host-side mirror of the reference API (Structure / Mesh / NonLinearStaticAnalysis
/ NewtonRaphson / solve!) over the C ABI of libonsas_hip.so (include/onsas_hip.h).
Julia's `f!` names are spelled `f_` here.
"""
from . import _capi, meshes, workloads
from .analysis import (AbstractConvergenceCriterion, ConvergenceSettings, MaxIterCriterion,
                       NewtonRaphson, NonLinearStaticAnalysis, NotConvergedYet,
                       ResidualForceCriterion, ResidualsIterationStep, Solution, ΔUCriterion,
                       DeltaUCriterion, ΔU_and_ResidualForce_Criteria, assemble_, criterion,
                       current_iteration, current_state, displacement_tol, isconverged_,
                       iterations, load_factors, max_iter_tol, reset_, residual_forces_tol,
                       solve, solve_, step_, tolerances, update_)
from .backend import GpuStaticState, MI355XBackend, internal_forces
from .boundary_conditions import FixedField, GlobalLoad, Pressure, StructuralBoundaryCondition
from .entities import (Circle, GenericCrossSection, GreenStrain, Mesh, NodeSet,
                       RotatedEngineeringStrain, Square, Tetrahedra, TriangularFaces, Trusses)
from .linear_analysis import LinearStaticAnalysis
from .materials import (IsotropicLinearElastic, NeoHookean, SVK, bulk_modulus, elasticity_modulus,
                        lame_parameters, poisson_ratio, shear_modulus)
from .structure import StructuralMaterial, Structure

__all__ = [n for n in dir() if not n.startswith("_")]
