"""MI355X-native (gfx950) 2D staggered-grid viscoelastic velocity-stress stepper.

The compute path is the HIP library ``libvi_mi355x.so`` behind the C ABI of
``include/vi_mi355x.h``; this package is its host-side mirror of the reference solver API.
"""
from . import segy
from .api import (CPMLConfig, DirichletBC, Fields, FreeSurfaceBC, IsotropicModel, K,
                  Parameters, Problem, QModel, RickerSource, Solution, StepParameters, VTIModel,
                  _ricker, apply_parameters, dK_dw,
                  find_optimal_taun, get_bc, get_Q, get_tau, gradient_to_thomsen_params,
                  gradient_to_velocity_params,
                  leapfrog_step, make_problem, misfit_and_pullback, solve_problem,
                  stress_eq, thomsen_to_stiffness, velocity_eq)
from .batch import Batch, Model, cached_bytes, mem_info, release_cached
from .stepper import COMPONENTS, FIELDS, SOURCE_KINDS, Simulation

__all__ = [
    "CPMLConfig", "DirichletBC", "Fields", "FreeSurfaceBC", "IsotropicModel", "K",
    "Parameters", "Problem", "QModel", "RickerSource", "Solution", "_ricker", "dK_dw",
    "find_optimal_taun", "get_bc", "get_Q", "get_tau", "leapfrog_step", "make_problem",
    "solve_problem", "Simulation", "FIELDS", "COMPONENTS", "SOURCE_KINDS", "Batch", "Model",
    "mem_info", "cached_bytes", "release_cached", "gradient_to_velocity_params",
    "misfit_and_pullback", "VTIModel", "thomsen_to_stiffness", "segy",
    "gradient_to_thomsen_params", "StepParameters", "apply_parameters", "stress_eq",
    "velocity_eq",
]
