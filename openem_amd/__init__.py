"""MI355X-native FDTD time-stepper behind OpenEM's ``install()`` / ``run()`` API.

``run(scene)`` advances a :class:`~openem_amd.model.Scene` on the GPU through the HIP library
``libopenem_hip.so`` (C ABI in ``include/openem_hip.h``) and returns a
:class:`~openem_amd.results.Result` (reference API: ``openem/__init__.py:31-63``,
``openem/solver.py:403``).
"""

from __future__ import annotations

from . import cpml, flux, grid, model, scenes, tables, waveform  # noqa: F401
from .model import Scene
from .results import Result

__version__ = "0.1.0"


def run(sim, task_name: str | None = None, verbose: bool = True, **kw) -> Result:
    """Solve a ``Scene`` (reference ``openem.run``, ``openem/__init__.py:56-63``). A tidy3d
    ``Simulation`` needs the tidy3d scene builder, which is not part of this package."""
    from . import solver
    if not isinstance(sim, Scene):
        raise NotImplementedError(
            "openem_amd.run takes a Scene; building one from a tidy3d Simulation needs tidy3d")
    return solver.run(sim, verbose=verbose, **kw)


def install(steps: int | None = None, max_cells: int = 80_000_000, **_kw) -> bool:
    """Route tidy3d's ``web.run`` to this stepper when tidy3d is importable (reference
    ``openem/__init__.py:31-47``). Returns False (and patches nothing) without tidy3d."""
    try:
        import tidy3d  # noqa: F401
    except ImportError:
        return False
    raise NotImplementedError("the tidy3d scene builder is not part of this package")
