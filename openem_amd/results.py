"""Run results (reference ``openem/results.py:18-84``)."""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Phasors:
    """Plane phasors ``(4, nf, n1, n2)`` complex128 in the order E_t1, E_t2, H_t1, H_t2."""

    data: np.ndarray
    freqs: np.ndarray


@dataclass
class Result:
    phasors: dict = field(default_factory=dict)
    field_phasors: dict = field(default_factory=dict)
    mode_amps: dict = field(default_factory=dict)      # name -> a_plus, a_minus, freqs, n_eff
    steps_run: int = 0
    wall_seconds: float = 0.0
    fields: np.ndarray | None = None
    shutoff_triggered: bool = False
    stop_reason: str = "pinned step count"
    decay_trace: list = field(default_factory=list)     # (step, field-decay ratio) per check
    energy_trace: list = field(default_factory=list)    # (step, total energy) per check
