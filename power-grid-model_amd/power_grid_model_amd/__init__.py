"""power_grid_model_amd -- batched Newton-Raphson power flow on AMD MI355X (gfx950).

Drop-in for the power-flow path of power-grid-model: the same input / update / output datasets and the
same PowerGridModel.calculate_power_flow(update_data=...) call, computed by hand-written HIP kernels.
"""

from ._lib import LIB_PATH, lib
from .errors import (
    AutomaticTapInputError,
    ConflictID,
    ConflictVoltage,
    DatasetError,
    IDNotFound,
    IDWrongType,
    InvalidCalculationMethod,
    InvalidMeasuredObject,
    InvalidTransformerClock,
    IterationDiverge,
    MaxIterationReached,
    NotObservableError,
    PowerGridBatchError,
    PowerGridError,
    SparseMatrixError,
)
from .meta import initialize_array, power_grid_meta_data
from .model import (
    BranchSide,
    CalculationMethod,
    CalculationType,
    ComponentAttributeFilterOptions,
    DeviceBatch,
    FactorPrecision,
    LoadGenType,
    PowerGridModel,
    TapChangingStrategy,
    WindingType,
    block_lu_solve,
    inspect_topology,
    minimum_degree_ordering,
    plan_stats,
    tap_regulator_ranks,
)

__all__ = [
    "LIB_PATH", "lib", "ConflictID", "ConflictVoltage", "DatasetError", "IDNotFound", "IDWrongType",
    "InvalidCalculationMethod", "InvalidTransformerClock", "IterationDiverge", "PowerGridBatchError",
    "PowerGridError", "SparseMatrixError", "NotObservableError", "InvalidMeasuredObject", "initialize_array", "power_grid_meta_data", "BranchSide",
    "CalculationMethod", "CalculationType", "ComponentAttributeFilterOptions", "DeviceBatch", "FactorPrecision", "LoadGenType", "PowerGridModel",
    "WindingType", "block_lu_solve", "inspect_topology", "minimum_degree_ordering", "plan_stats",
    "AutomaticTapInputError", "MaxIterationReached", "TapChangingStrategy", "tap_regulator_ranks",
]
