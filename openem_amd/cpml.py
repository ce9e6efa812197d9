"""CPML profiles and recursion coefficients, plus SI constants.

The grading follows tidy3d's PML definition as the reference restates it
(``openem/cpml.py:1-16``, ``:55-128``)::

    sigma(x) = sigma_max * v / (eta0 * dl) * step**sigma_order
    kappa(x) = kappa_min + (kappa_max - kappa_min) * step**kappa_order
    alpha(x) = alpha_min + (alpha_max - alpha_min) * step**alpha_order

``step`` is the normalised depth at the E or H position of each layer (E positions are the
backward-difference points, H positions are half a cell further out). The psi recursion is the
Roden-Gedney form ``psi <- b*psi + a*d`` with

    b = exp(-dt/eps0 * (sigma/kappa + alpha)),   a = (b - 1) * sigma / (sigma + alpha*kappa) / kappa

and ``a = 0`` where ``sigma = 0``.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

EPSILON_0 = 8.8541878128e-12       # F/m
MU_0 = 1.25663706212e-6            # H/m
ETA_0 = 376.7303136668535          # ohm
C_0 = 299792458.0                  # m/s


@dataclass(frozen=True)
class PMLParams:
    """``td.PML()`` defaults (reference ``openem/cpml.py:37-52``)."""

    num_layers: int
    sigma_max: float = 1.5
    sigma_order: int = 3
    kappa_min: float = 1.0
    kappa_max: float = 3.0
    kappa_order: int = 3
    alpha_min: float = 0.0
    alpha_max: float = 0.0
    alpha_order: int = 1


def steps(n_pml: int, side: str, at: str) -> np.ndarray:
    """Normalised depth of each layer, ordered by increasing global index
    (reference ``openem/cpml.py:55-76``). The E side is deliberately asymmetric: at the low end
    the outermost E layer has depth 1, at the high end the innermost E layer has depth 0."""
    if n_pml <= 0:
        return np.zeros(0)
    i = np.arange(n_pml, dtype=np.float64)
    if side == "lo":
        return (n_pml - i) / n_pml if at == "E" else (n_pml - i - 0.5) / n_pml
    return i / n_pml if at == "E" else (i + 0.5) / n_pml


def profile(p: PMLParams, dl: float, side: str, at: str, avg_speed: float = 1.0):
    """``(sigma, kappa, alpha)`` per layer (reference ``openem/cpml.py:79-112``)."""
    s = steps(p.num_layers, side, at)
    sigma = p.sigma_max * avg_speed / (ETA_0 * dl) * s ** p.sigma_order
    kappa = p.kappa_min + (p.kappa_max - p.kappa_min) * s ** p.kappa_order
    alpha = p.alpha_min + (p.alpha_max - p.alpha_min) * s ** p.alpha_order
    return sigma, kappa, alpha


def recursion_coeffs(sigma, kappa, alpha, dt: float):
    """``(a, b)`` of ``psi <- b*psi + a*d`` (reference ``openem/cpml.py:115-128``)."""
    b = np.exp(-dt / EPSILON_0 * (sigma / kappa + alpha))
    denom = sigma + alpha * kappa
    with np.errstate(divide="ignore", invalid="ignore"):
        a = (b - 1.0) * sigma / denom / kappa
    return np.nan_to_num(a, nan=0.0, posinf=0.0, neginf=0.0), b


@dataclass(frozen=True)
class PMLCoeffs:
    """Coefficients of one PML face, each ``(num_layers,)``."""

    a_E: np.ndarray
    b_E: np.ndarray
    inv_kappa_E: np.ndarray
    a_H: np.ndarray
    b_H: np.ndarray
    inv_kappa_H: np.ndarray

    @property
    def num_layers(self) -> int:
        return int(self.a_E.size)


def build(p: PMLParams, dl: float, side: str, dt: float, avg_speed: float = 1.0) -> PMLCoeffs:
    """Parameters -> profile -> recursion coefficients (reference ``openem/cpml.py:149-164``)."""
    out = {}
    for at in ("E", "H"):
        sigma, kappa, alpha = profile(p, dl, side, at, avg_speed)
        a, b = recursion_coeffs(sigma, kappa, alpha, dt)
        out[f"a_{at}"] = a
        out[f"b_{at}"] = b
        out[f"inv_kappa_{at}"] = 1.0 / kappa
    return PMLCoeffs(**out)


def courant_dt(dls: tuple[float, float, float], courant: float = 0.99) -> float:
    """Time step of a uniform grid: ``courant / (c * sqrt(sum 1/dl^2))`` (tidy3d's rule for
    ``dt``; the reference takes ``sim.dt`` from tidy3d, ``openem/scene/build.py:271``)."""
    inv = np.sqrt(sum(1.0 / (d * d) for d in dls))
    return float(courant / (C_0 * inv))
