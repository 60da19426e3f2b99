"""MI355X-native DNS step of GBGHomogeneousTurbulence.jl: a HIP library (libhdns.so, C ABI in
include/hdns.h) and the reference's run API over it.

The folder name carries a dot, so it is imported by path; ``load_package()`` in
``__graft_entry__.py`` registers it as the module ``gbg_amd``.
"""
from . import _lib
from .api import (AB, ABM, ConstantPower, Ensemble, EnsembleBox, EnsembleMember, FixedDomain, FlowConditions, Heun,
                  IFHeun, IFLowStorageRK3, IFLowStorageRK4, IFSSPRK3, Linear, NoForcing, VarianceController,
                  VarianceRescaling, ResolvedFields, Simulation, Tracers, TwoThirds, HouLi, PhaseShift,
                  StationarityCriterion, cfl_bound, integral_length, predicted_viscosity,
                  recommended_cfl, truncate_to_band, warmup, reynolds_box, reynolds_taylor, taylor_length,
                  kolmogorov_length, kolmogorov_time, kolmogorov_velocity, unwrapped_positions, revolve_schedule,
                  SCHEDULE_FAMILIES)
from .seeding import random_phase_scalar, random_phase_velocity, uniform_positions
from .stochastic import StochasticOU

__all__ = [
    "AB", "ABM", "ConstantPower", "Ensemble", "EnsembleBox", "EnsembleMember", "FixedDomain", "FlowConditions", "Heun",
    "IFHeun", "IFLowStorageRK3", "IFLowStorageRK4", "IFSSPRK3", "Linear", "NoForcing", "VarianceController",
    "VarianceRescaling", "ResolvedFields", "Simulation", "Tracers", "TwoThirds", "HouLi", "PhaseShift",
    "predicted_viscosity",
    "recommended_cfl", "cfl_bound", "integral_length", "StationarityCriterion", "warmup", "truncate_to_band", "random_phase_scalar", "random_phase_velocity",
    "uniform_positions", "StochasticOU", "reynolds_box", "reynolds_taylor", "taylor_length", "kolmogorov_length",
    "kolmogorov_time", "kolmogorov_velocity", "unwrapped_positions", "revolve_schedule", "SCHEDULE_FAMILIES",
]
