"""Early termination (reference ``openem/shutoff.py:1-253``; checks driven from
``openem/solver.py:103-203``).

Every ``interval`` steps (400, Tidy3D's check interval) three criteria are evaluated; the run
may stop when any fires, but never before the sources have finished injecting:

* field decay: instantaneous ``sum V |E|^2`` (Yee dual volumes, PML included, no eps, no H)
  over its running maximum, below ``threshold``;
* energy decay: ``1/2 (eps0 sum V eps |E|^2 + mu0 sum V |H|^2)`` over its running maximum,
  below ``threshold``; its history also yields the decay time tau (a straight-line fit of
  ``ln(ratio)`` against the step over the second half of the checks);
* phasor convergence: the largest change of the monitored fluxes since the previous check,
  relative to their largest magnitude, scaled by ``tau / interval`` (the change still to come),
  below ``tol`` on ``patience`` consecutive checks.

The sums come from the device (:meth:`openem_amd.solver.Stepper.energy`, float64
accumulation of float32 terms) and the fluxes from the device flux reduction.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

CHECK_INTERVAL = 400
#: injection counts as over once every source envelope stays below this fraction of its peak
SOURCE_TAIL = 1e-3


@dataclass
class RatioToPeak:
    """``value / max(value so far)`` against a threshold (the field and energy criteria)."""

    threshold: float
    peak: float = 0.0
    ratio: float = float("nan")
    history: list = field(default_factory=list)        # (step, ratio)
    values: list = field(default_factory=list)         # (step, value)

    def update(self, step: int, value: float) -> bool:
        value = float(value)
        self.values.append((step, value))
        if value > self.peak:
            self.peak = value
        self.ratio = value / self.peak if self.peak > 0.0 else 0.0
        self.history.append((step, self.ratio))
        return self.peak > 0.0 and self.ratio < self.threshold

    @property
    def deepest(self) -> tuple:
        if not self.history:
            return (0, float("nan"))
        return min(self.history, key=lambda sr: sr[1])

    @property
    def tau(self) -> float:
        """Decay time in steps from the second half of the history; ``inf`` while the quantity
        is not decaying (or too few checks exist to tell)."""
        if len(self.history) < 8:
            return float("inf")
        tail = np.asarray(self.history[len(self.history) // 2:], dtype=np.float64)
        ok = tail[:, 1] > 0.0
        if np.count_nonzero(ok) < 4:
            return float("inf")
        slope = np.polyfit(tail[ok, 0], np.log(tail[ok, 1]), 1)[0]
        return float(-1.0 / slope) if np.isfinite(slope) and slope < 0.0 else float("inf")


@dataclass
class FluxConvergence:
    """The extrapolated remaining relative change of the fluxes (see the module docstring)."""

    tol: float = 1e-3
    patience: int = 3
    window_steps: int = CHECK_INTERVAL
    last: np.ndarray | None = None
    streak: int = 0
    rel: float = float("inf")
    per_window: float = float("inf")
    history: list = field(default_factory=list)

    def update(self, step: int, fluxes, tau: float | None = None) -> bool:
        cur = np.asarray(fluxes, dtype=np.float64)
        prev, self.last = self.last, cur
        if prev is None:
            return False
        top = float(np.abs(cur).max()) if cur.size else 0.0
        self.per_window = float(np.abs(cur - prev).max()) / top if top > 0.0 else float("inf")
        scale = tau / max(self.window_steps, 1) if tau is not None and 0.0 < tau < np.inf else 1.0
        self.rel = self.per_window * scale
        self.history.append((step, self.rel))
        self.streak = self.streak + 1 if self.rel < self.tol else 0
        return self.streak >= self.patience


@dataclass
class Policy:
    """Any enabled criterion firing stops the run; ``reason`` names the ones that fired."""

    field_decay: RatioToPeak | None = None
    energy_decay: RatioToPeak | None = None
    phasor: FluxConvergence | None = None
    interval: int = CHECK_INTERVAL
    reason: str = ""

    def check(self, step: int, e2: float, energy: float, fluxes) -> bool:
        fired = []
        if self.field_decay is not None and self.field_decay.update(step, e2):
            fired.append(f"field_decay<{self.field_decay.threshold:g}")
        if self.energy_decay is not None and self.energy_decay.update(step, energy):
            fired.append(f"energy_decay<{self.energy_decay.threshold:g}")
        tau = self.energy_decay.tau if self.energy_decay is not None else None
        if self.phasor is not None and np.size(fluxes) and self.phasor.update(step, fluxes, tau):
            fired.append(f"phasor_converged(remaining~{self.phasor.rel:.2e}<{self.phasor.tol:g}"
                         f" x{self.phasor.patience})")
        if fired:
            self.reason = " + ".join(fired)
        return bool(fired)


def default_policy(threshold: float, interval: int = CHECK_INTERVAL) -> Policy:
    """All three criteria: the scene's shutoff threshold for the two decays, flux tolerance
    1e-3 held for 3 checks."""
    return Policy(field_decay=RatioToPeak(threshold), energy_decay=RatioToPeak(threshold),
                  phasor=FluxConvergence(tol=1e-3, patience=3, window_steps=interval),
                  interval=interval)


def source_end_step(sc) -> int:
    """Last step at which any source's sampled envelope exceeds SOURCE_TAIL of its peak (the
    tables that are injected, so the test cannot disagree with the injection)."""
    end = 0
    waves = ([s.waveform for s in sc.sources] + [d.waveform for d in sc.dipoles]
             + [t.waveform for t in sc.tfsf_sources])
    for w in waves:
        a = np.abs(np.asarray(w.amp_int, dtype=np.float64))
        top = float(a.max()) if a.size else 0.0
        if top <= 0.0:
            continue
        live = np.flatnonzero(a > SOURCE_TAIL * top)
        if live.size:
            end = max(end, int(live[-1]))
    return end
