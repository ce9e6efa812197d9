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
    """The reference's ``Result`` fields (``openem/results.py:50-84``) plus ``mode_amps``."""

    phasors: dict = field(default_factory=dict)
    field_phasors: dict = field(default_factory=dict)    # name -> (6, nf, ni, nj, nk), Yee positions
    mode_amps: dict = field(default_factory=dict)        # name -> a_plus, a_minus, freqs, n_eff
    steps_run: int = 0
    wall_seconds: float = 0.0
    fields: dict = field(default_factory=dict)           # "Ex".."Hz" -> (nx, ny, nz), with return_fields
    shutoff_triggered: bool = False
    stop_reason: str = ""                                # the criterion that stopped the run, or ""
    shutoff_summary: str = ""
    source_end: int = -1                                 # step at which every source is done
    decay_trace: list = field(default_factory=list)      # (step, field-decay ratio) per check
    energy_trace: list = field(default_factory=list)     # (step, total energy) per check
    time_samples: dict = field(default_factory=dict)     # FieldTimeMonitor name -> samples
    time_slots: dict = field(default_factory=dict)       # FieldTimeMonitor name -> slots filled
    flux_time: dict = field(default_factory=dict)        # FluxTimeMonitor name -> flux series


def downsample_indices(n: int, k: int) -> np.ndarray:
    """Points kept along an axis of ``n`` samples by tidy3d's ``interval_space`` stride ``k``:
    every ``k``-th from the first, the last always kept (reference
    ``openem/nb/backend.py:1942-1955``)."""
    n, k = int(n), max(int(k), 1)
    idx = np.arange(0, n, k)
    if n > 0 and idx[-1] != n - 1:
        idx = np.append(idx, n - 1)
    return idx


def strided_field_phasors(result: Result, mon) -> np.ndarray:
    """A field monitor's ``(6, nf, ni, nj, nk)`` phasors thinned by its ``interval_space``: the
    box is accumulated at full resolution and the stride applies here (reference
    ``openem/nb/backend.py:1958-1966``)."""
    data = result.field_phasors[mon.name]
    for ax, k in enumerate(getattr(mon, "interval_space", (1, 1, 1))):
        data = np.take(data, downsample_indices(data.shape[2 + ax], k), axis=2 + ax)
    return data
