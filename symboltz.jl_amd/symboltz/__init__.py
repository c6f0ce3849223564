"""symboltz: MI355X-native perturbation-stage solver behind SymBoltz's problem/solve API.

Public names mirror the reference's exports (src/SymBoltz.jl:50-58) for the perturbation path.
"""
from .models import LCDM, w0waCDM, Model, parameters_Planck18, make_params
from .problem import (CosmologyProblem, CosmologySolution, issuccess, parameter_updater, remake, solve, solve_sets,
                      solvebg, spectrum_matter, spectrum_primordial)
from .constants import k_dimensionless

ΛCDM = LCDM

__all__ = ["LCDM", "ΛCDM", "w0waCDM", "Model", "parameters_Planck18", "make_params", "CosmologyProblem",
           "CosmologySolution", "issuccess", "parameter_updater", "remake", "solve", "solve_sets", "solvebg",
           "spectrum_matter", "spectrum_primordial", "k_dimensionless"]
