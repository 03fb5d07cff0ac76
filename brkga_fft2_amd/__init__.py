"""MI355X-native BRKGA population decoder (gfx950 HIP kernels behind a C ABI)."""
from .problem import (PartData, MachinePartData, MachineData, ProblemData, INFEASIBLE_PENALTY,
                      prepare_part, build_problem, thresholds_for, load_reference_instance,
                      synthetic_problem, synthetic_population, pack_problem, SYNTHETIC_CONFIGS)

__all__ = ["PartData", "MachinePartData", "MachineData", "ProblemData", "INFEASIBLE_PENALTY",
           "prepare_part", "build_problem", "thresholds_for", "load_reference_instance",
           "synthetic_problem", "synthetic_population", "pack_problem", "SYNTHETIC_CONFIGS"]
