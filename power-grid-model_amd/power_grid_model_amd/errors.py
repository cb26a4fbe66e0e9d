"""Exceptions, named and raised like the reference's src/power_grid_model/_core/error_handling.py
(the message text decides the class)."""

from __future__ import annotations

import re

import numpy as np


class PowerGridError(RuntimeError):
    """Generic power grid error."""


class PowerGridBatchError(PowerGridError):
    """Error of a batch calculation; lists the failed scenarios."""

    failed_scenarios: np.ndarray
    succeeded_scenarios: np.ndarray
    error_messages: list[str]
    errors: list[PowerGridError]


class IterationDiverge(PowerGridError):
    """Newton-Raphson did not converge within the maximum number of iterations."""


class SparseMatrixError(PowerGridError):
    """The linear system is singular."""


class IDNotFound(PowerGridError):
    """An id in the input or update data is unknown."""


class IDWrongType(PowerGridError):
    """An id refers to a component of another type."""


class ConflictID(PowerGridError):
    """Two components share an id."""


class ConflictVoltage(PowerGridError):
    """A line connects nodes of different rated voltage."""


class InvalidTransformerClock(PowerGridError):
    """Transformer clock inconsistent with its windings."""


class InvalidCalculationMethod(PowerGridError):
    """Calculation method not valid for the calculation type."""


class DatasetError(PowerGridError):
    """Malformed dataset."""


class MaxIterationReached(IterationDiverge):
    """The tap changer did not settle within its iteration limit."""


class AutomaticTapInputError(PowerGridError):
    """A regulated transformer is controlled from its non-source side towards the source side."""


class InvalidRegulatedObject(PowerGridError):
    """A transformer_tap_regulator regulates an object that is not a transformer."""


class DuplicativelyRegulatedObject(PowerGridError):
    """An object has more than one regulator."""


class NotObservableError(PowerGridError):
    pass


class InvalidMeasuredObject(PowerGridError):
    pass


_PATTERNS = [
    (r"Not enough measurements available for state estimation", NotObservableError),
    (r"measurement is not supported for object of type", InvalidMeasuredObject),
    (r"Maximum number of iterations reached!", MaxIterationReached),
    (r"Automatic tap changer has invalid configuration", AutomaticTapInputError),
    (r"regulator is not supported for object", InvalidRegulatedObject),
    (r"regulated by more than one regulator", DuplicativelyRegulatedObject),
    (r"Iteration failed to converge", IterationDiverge),
    (r"Sparse matrix error", SparseMatrixError),
    (r"The id cannot be found:", IDNotFound),
    (r"Wrong type for object with id", IDWrongType),
    (r"Conflicting id detected:", ConflictID),
    (r"Conflicting voltage for line", ConflictVoltage),
    (r"Invalid clock for transformer", InvalidTransformerClock),
    (r"The calculation method is invalid", InvalidCalculationMethod),
    (r"Dataset error", DatasetError),
]


def error_class(message: str) -> type[PowerGridError]:
    for pattern, cls in _PATTERNS:
        if re.search(pattern, message):
            return cls
    return PowerGridError


def raise_regular_error(message: str):
    raise error_class(message)(message)
