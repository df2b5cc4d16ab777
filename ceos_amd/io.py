"""Binary dataset files: raw little-endian FP32, one row of n_features values after
another (the format of the reference's loadbinData, src/Utilities.cpp:142-214).

load_bin maps the file (numpy memmap, nothing copied until used) and returns the first
n_points rows as an (n_points, n_features) array. The reference allocates its matrix
as n_features x n_points but fills it as n_points x n_features (Utilities.cpp:180-194);
here the shape is the intended one.
"""
from __future__ import annotations

import os

import numpy as np


def rows_in(path: str, n_features: int) -> int:
    """Whole rows of n_features FP32 values in the file (Utilities.cpp:161)."""
    return os.path.getsize(path) // (4 * int(n_features))


def load_bin(path: str, n_points: int, n_features: int) -> np.ndarray:
    total = rows_in(path, n_features)
    if n_points > total:  # Utilities.cpp:172-176
        raise ValueError(f"{path}: n_points {n_points} exceeds the {total} rows in the file")
    return np.memmap(path, dtype="<f4", mode="r", shape=(int(n_points), int(n_features)))


def save_bin(path: str, X: np.ndarray) -> None:
    np.ascontiguousarray(X, dtype="<f4").tofile(path)
