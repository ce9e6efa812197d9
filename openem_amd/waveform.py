"""Source time waveforms.

The reference samples tidy3d's ``source_time.amp_time`` into step tables
(``openem/waveform.py:48-65``). tidy3d is not available here, so :class:`GaussianPulse`
restates the published tidy3d ``GaussianPulse.amp_time`` (tidy3d 2.x,
``components/source/time.py``)::

    twidth  = 1 / (2 pi fwidth)
    ts      = t - offset * twidth
    pulse   = amplitude * exp(i phase) * exp(-i 2 pi f0 t) * exp(-ts^2 / (2 twidth^2))
    pulse  *= (1j + ts / (twidth^2 * 2 pi f0))      if remove_dc_component   (default)
    pulse  *= 1j                                    otherwise

The DC-removal factor makes the real part the time derivative of a Gaussian envelope carrier,
so its time integral vanishes (pinned by a test, since tidy3d itself cannot be imported).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class GaussianPulse:
    freq0: float
    fwidth: float
    amplitude: float = 1.0
    phase: float = 0.0
    offset: float = 5.0
    remove_dc_component: bool = True

    @property
    def twidth(self) -> float:
        return 1.0 / (2.0 * np.pi * self.fwidth)

    def amp_time(self, t) -> np.ndarray:
        t = np.asarray(t, dtype=np.float64)
        omega0 = 2.0 * np.pi * self.freq0
        tw = self.twidth
        ts = t - self.offset * tw
        amp = (self.amplitude * np.exp(1j * self.phase) * np.exp(-1j * omega0 * t)
               * np.exp(-(ts ** 2) / (2.0 * tw ** 2)))
        if self.remove_dc_component:
            return amp * (1j + ts / (tw ** 2) / omega0)
        return amp * 1j

    def end_time(self, margin: float = 2.0) -> float:
        """Time after which the envelope is below ~exp(-(offset+margin)^2/2)."""
        return (2.0 * self.offset + margin) * self.twidth


@dataclass(frozen=True)
class Waveform:
    """A pre-sampled waveform (reference ``openem/waveform.py:19-45``).

    ``amp_int[n] = Re a(n dt)``, ``amp_half[n] = Re a((n+1/2) dt)``; the complex versions feed
    the Bloch path.
    """

    amp_int: np.ndarray
    amp_half: np.ndarray
    amp_int_complex: np.ndarray
    amp_half_complex: np.ndarray
    dt: float

    @property
    def num_steps(self) -> int:
        return self.amp_int.size - 1


def sample(source_time, dt: float, num_steps: int) -> Waveform:
    n = np.arange(num_steps + 1, dtype=np.float64)
    a_int = np.asarray(source_time.amp_time(n * dt), dtype=np.complex128)
    a_half = np.asarray(source_time.amp_time((n + 0.5) * dt), dtype=np.complex128)
    return Waveform(amp_int=a_int.real.copy(), amp_half=a_half.real.copy(),
                    amp_int_complex=a_int, amp_half_complex=a_half, dt=dt)
