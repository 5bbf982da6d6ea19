"""Domain type aliases (reference src/protocols/types.py:19-55).

Host arrays are NumPy; device arrays are torch tensors on the HIP device.
"""

from __future__ import annotations

from typing import Any, Literal

import numpy as np
from numpy.typing import NDArray

Route = NDArray[np.int32]  # (n,) tour
Routes = NDArray[np.int32]  # (N, n) batch
RouteOrRoutes = Any  # Route | Routes | device tensor of either shape
Fitness = NDArray[np.float64]  # (N,) costs
DistanceMatrix = NDArray[np.int32]
SeedLike = Any  # int | Sequence[int] | Generator | Callable[[], Generator] | None

LocalSearchStage = Literal["post_crossover", "post_mutation", "post_evaluation", "post_evolution", "none"]
