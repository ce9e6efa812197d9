"""CPML coefficients of one PML face, and the SI constants.

Grading: tidy3d's published PML definition (``PML`` / ``PMLParams`` in tidy3d's documentation;
the reference states it in ``openem/cpml.py:1-16``).  At normalised depth ``d`` in [0, 1]
(0 at the interface with the domain, 1 at the outer wall)::

    sigma(d) = sigma_max * v / (eta0 * dl) * d**sigma_order      dl: the grid step at that face
    kappa(d) = kappa_min + (kappa_max - kappa_min) * d**kappa_order
    alpha(d) = alpha_min + (alpha_max - alpha_min) * d**alpha_order

Depth of a sample: a face of ``n`` layers spans ``n`` cells; E samples sit on primal edges and H
samples at cell centres, so a sample ``p`` cells from the interface has depth ``p / n``.  On a
low face the interface is the stack's high edge (the E sample on the outer wall has depth 1,
the last H sample depth 1/(2n)); on a high face it is the stack's low edge (the first E sample
lies on the interface itself, depth 0).

Recursion: the convolutional PML of Roden & Gedney (2000), ``psi <- b psi + a d`` with::

    b = exp(-(sigma / kappa + alpha) dt / eps0)
    a = sigma (b - 1) / (kappa (sigma + kappa alpha))      (a = 0 where sigma = 0)

and the stretched derivative ``d / kappa + psi``.
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
    """Grading parameters of one face; the defaults are ``td.PML()``'s (field names as tidy3d's
    ``PMLParams``)."""

    num_layers: int
    sigma_max: float = 1.5
    sigma_order: int = 3
    kappa_min: float = 1.0
    kappa_max: float = 3.0
    kappa_order: int = 3
    alpha_min: float = 0.0
    alpha_max: float = 0.0
    alpha_order: int = 1

    def grade(self, depth: np.ndarray, dl: float, avg_speed: float = 1.0):
        """``(sigma, kappa, alpha)`` at the given normalised depths."""
        sigma_peak = self.sigma_max * avg_speed / (ETA_0 * dl)
        sigma = sigma_peak * depth ** self.sigma_order
        kappa = self.kappa_min + (self.kappa_max - self.kappa_min) * depth ** self.kappa_order
        alpha = self.alpha_min + (self.alpha_max - self.alpha_min) * depth ** self.alpha_order
        return sigma, kappa, alpha


def sample_depths(n_layers: int, side: str, at: str) -> np.ndarray:
    """Normalised depth of each layer's E (``at="E"``) or H sample, by increasing global
    index (module docstring)."""
    if n_layers <= 0:
        return np.zeros(0)
    pos = np.arange(n_layers, dtype=np.float64) + (0.5 if at == "H" else 0.0)
    from_interface = n_layers - pos if side == "lo" else pos
    return from_interface / n_layers


def recursion(sigma: np.ndarray, kappa: np.ndarray, alpha: np.ndarray, dt: float):
    """``(a, b)`` of the psi recursion at each sample."""
    b = np.exp(-(dt / EPSILON_0) * (sigma / kappa + alpha))
    a = np.zeros_like(b)
    live = sigma > 0
    s, k, al = sigma[live], kappa[live], alpha[live]
    a[live] = s * (b[live] - 1.0) / (k * (s + k * al))
    return a, b


@dataclass(frozen=True)
class PMLCoeffs:
    """Coefficients of one PML face, each ``(num_layers,)`` (the Scene contract of the
    reference's ``cpml.PMLCoeffs``)."""

    a_E: np.ndarray
    b_E: np.ndarray
    inv_kappa_E: np.ndarray
    a_H: np.ndarray
    b_H: np.ndarray
    inv_kappa_H: np.ndarray

    @property
    def num_layers(self) -> int:
        return int(self.a_E.size)


def face_coeffs(p: PMLParams, dl: float, side: str, dt: float,
                avg_speed: float = 1.0) -> PMLCoeffs:
    """The face's coefficients at its E and H samples."""
    def at(where):
        sigma, kappa, alpha = p.grade(sample_depths(p.num_layers, side, where), dl, avg_speed)
        a, b = recursion(sigma, kappa, alpha, dt)
        return a, b, 1.0 / kappa

    (ae, be, ke), (ah, bh, kh) = at("E"), at("H")
    return PMLCoeffs(a_E=ae, b_E=be, inv_kappa_E=ke, a_H=ah, b_H=bh, inv_kappa_H=kh)


def courant_dt(dls: tuple[float, float, float], courant: float = 0.99) -> float:
    """Time step of a uniform grid: ``courant / (c * sqrt(sum 1/dl^2))`` (tidy3d's rule for
    ``dt``; the reference takes ``sim.dt`` from tidy3d, ``openem/scene/build.py:271``)."""
    inv = np.sqrt(sum(1.0 / (d * d) for d in dls))
    return float(courant / (C_0 * inv))
