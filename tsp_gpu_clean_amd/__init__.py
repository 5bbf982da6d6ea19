"""tsp_gpu_clean_amd — MI355X-native memetic-TSP hot path behind the reference's strategy API.

The HIP kernels live in ``csrc/`` and are reached only through the C ABI of ``libtspmi.so``
(``include/tspmi.h``) via ``_native``.  Importing the package needs no GPU; calling a GPU
strategy without a device raises ``GPUUnavailableError``.
"""

from .context import ProblemContext
from .errors import GPUUnavailableError, VRAMInsufficientError
from .ga import STAGE_SPECS, GenerationState, GeneticAlgorithm
from .pipeline_steps import (
  PIPELINE_STEP_REGISTRY,
  EliteOnlyLocalSearchStep,
  PeriodicLocalSearchStep,
  PostInitLocalSearchStep,
  register_pipeline_step,
)
from .problem import Problem, load_tsplib_file
from .strategies import (
  STRATEGY_REGISTRY,
  CrossoverCPUPermOrder,
  CrossoverCPUPermPartiallyMapped,
  CrossoverGPUPermOrder,
  CrossoverGPUPermPartiallyMapped,
  EvaluationCPUCost,
  EvaluationGPUCost,
  InitializationCPUPermRandom,
  InitializationCPUTourNearestNeighbor,
  InitializationGPUPermRandom,
  InitializationGPUTourNearestNeighbor,
  LocalSearchCPUTourTwoOpt,
  LocalSearchGPUTourTwoOpt,
  LocalSearchNoOp,
  MutationCPUPermInversion,
  MutationCPUPermSwap,
  MutationGPUPermInversion,
  MutationGPUPermSwap,
  SelectionCPUTournament,
  SelectionGPUTournament,
  StrategyBase,
  register_strategy,
)

__all__ = [
  "PIPELINE_STEP_REGISTRY",
  "STAGE_SPECS",
  "STRATEGY_REGISTRY",
  "CrossoverCPUPermOrder",
  "CrossoverCPUPermPartiallyMapped",
  "EvaluationCPUCost",
  "InitializationCPUPermRandom",
  "InitializationCPUTourNearestNeighbor",
  "LocalSearchCPUTourTwoOpt",
  "MutationCPUPermInversion",
  "MutationCPUPermSwap",
  "SelectionCPUTournament",
  "CrossoverGPUPermPartiallyMapped",
  "CrossoverGPUPermOrder",
  "EliteOnlyLocalSearchStep",
  "EvaluationGPUCost",
  "GPUUnavailableError",
  "GenerationState",
  "GeneticAlgorithm",
  "InitializationGPUPermRandom",
  "InitializationGPUTourNearestNeighbor",
  "LocalSearchGPUTourTwoOpt",
  "LocalSearchNoOp",
  "MutationGPUPermInversion",
  "MutationGPUPermSwap",
  "PeriodicLocalSearchStep",
  "PostInitLocalSearchStep",
  "Problem",
  "ProblemContext",
  "SelectionGPUTournament",
  "StrategyBase",
  "VRAMInsufficientError",
  "load_tsplib_file",
  "register_pipeline_step",
  "register_strategy",
]
