"""Exceptions of the Python front-end (reference ``src/monoprop/exceptions.py``)."""

from __future__ import annotations


class NumberOfModesInvalidError(ValueError):
    """Raised when the number of Fermionic modes is invalid."""
