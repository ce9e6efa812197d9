"""Build recipe of the oracle's native pieces (test infrastructure).

The reference's kernels are CUDA sources compiled by NVRTC at run time and its host chain needs
cupy and tidy3d, so no part of the reference's path builds here with gcc; ``oracle/_ref`` stays
empty and the oracle is the float64 NumPy restatement in ``oracle/yee64.py``.
"""

from __future__ import annotations


def build_oracle() -> None:
    """Nothing to compile: the oracle is pure NumPy."""
    return None
