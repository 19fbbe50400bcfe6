"""gpusteer_amd — MI355X (gfx950) implementation of GPUSteer's per-step
simulation path behind a C-ABI (include/gpusteer.h).

Importing the package loads lib/libgpusteer.so; there is no CPU fallback.
"""
from .build import LIB  # noqa: F401
from . import _capi  # noqa: F401  (raises ImportError when the library is not built)
from .api import (  # noqa: F401
    KNN_BINNED,
    KNN_DEFAULT,
    KNN_NO_SEED,
    MODE_GPUREF,
    MODE_OPENSTEER,
    AgentGroup,
    Context,
    Frame,
    KNNBinData,
    KNNData,
    ObstacleGroup,
    PolylinePathway,
    WallGroup,
    antiPenetrationAgents,
    clearAppliedKernels,
    findKNearestNeighbors,
    findKNearestNeighborsBruteForce,
    findKNearestNeighborsWithin,
    steerForAlignment,
    steerForCohesion,
    steerForFlocking,
    steerForSeek,
    steerForSeparation,
    steerToAvoidNeighbors,
    steerToAvoidObstacles,
    steerToAvoidWalls,
    steerToFollowPath,
    stepBoids,
    stepBoidsTimed,
    updateGroup,
    updateGroupWithAntiPenetration,
    updateKNNDatabase,
    wrapWorld,
)
from .ensemble import Calibration, Ensemble, Sequential  # noqa: F401
from .state import AgentState, WallState  # noqa: F401
