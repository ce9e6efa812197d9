"""Launch autotuner: picks the step form (the fused H+E march or the two sweeps) and the x chunk
each work-group marches, by timing a few candidates on the scene itself, once per scene shape.

The reference fixes its launch shapes by hand: a global block-size knob measured once on one
GPU (``openem/knobs.py:64``, ``openem/device.py:203-230``) and fusion gates written as size
rules (``openem/fusion.py:49-216``).  Here a candidate is a throwaway context of the same tables
(:class:`openem_amd.solver.Stepper` with ``form`` and ``chunk``) stepped a few warm-up steps and
then a timed block on its own stream; the fastest plan is cached under the scene's shape key and
handed to the contexts that step the scene for real.

Results never depend on the choice.  Every chunk gives the same bits (each chunk recomputes what
it carries across its boundary plane).  The two forms give the same bits unless a point list
corrects H (magnetic dipoles, TF/SF box faces, mode sources: the march adds those before the
curl, the sweeps after it), so on such scenes only the chunk of the library's default form is
tuned and a run's numbers never hinge on a timing.
"""

from __future__ import annotations

import json
import os
import threading
from pathlib import Path

import numpy as np

from . import tables as tables_mod

#: chunk candidates per form (x planes per work-group); the library's defaults are 12 (fused,
#: real fields), 16 (fused, complex) and a size rule for the sweeps
CHUNKS = {"fused": (8, 12, 16, 24), "sweeps": (16, 32, 64)}

_cache: dict = {}
_lock = threading.Lock()


def shape_key(tabs: tables_mod.StepTables) -> str:
    """What the launch plan depends on: the grid, real or complex fields, the CPML widths."""
    pml = tuple((a.pml_lo, a.pml_hi) for a in tabs.axes)
    return f"{tuple(tabs.shape)}|{'c' if tabs.bloch else 'r'}|{pml}"


def has_h_points(tabs: tables_mod.StepTables) -> bool:
    """Whether a point list corrects H (then the two forms differ in rounding)."""
    return (any(d.magnetic for d in tabs.dipoles) or bool(tabs.tfsf) or bool(tabs.modes))


def _cache_file() -> Path | None:
    p = os.environ.get("OEM_TUNE_CACHE")
    return Path(p) if p else None


def _load_file() -> dict:
    p = _cache_file()
    if p is None or not p.exists():
        return {}
    try:
        return json.loads(p.read_text())
    except (OSError, ValueError):
        return {}


def _store_file(key: str, plan: dict) -> None:
    p = _cache_file()
    if p is None:
        return
    data = _load_file()
    data[key] = plan
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        p.write_text(json.dumps(data, indent=1, sort_keys=True))
    except OSError:
        pass


def candidates(tabs: tables_mod.StepTables, default_form: str) -> list:
    """``(form, chunk)`` pairs to time.  Both forms when they give the same bits, else only the
    library's default form; the fused march only while a field set fits 32-bit offsets."""
    n = int(np.prod(tabs.shape))
    esz = 2 if tabs.bloch else 1
    fused_ok = 6 * n * 4 * esz < (1 << 32)
    forms = ["fused", "sweeps"] if fused_ok else ["sweeps"]
    if has_h_points(tabs):
        forms = [default_form]
    nx = tabs.shape[0]
    out = []
    for f in forms:
        seen = set()
        for c in CHUNKS[f]:
            c = min(c, nx)
            if c not in seen:
                seen.add(c)
                out.append((f, c))
    return out


def tune(tabs: tables_mod.StepTables, device: int = 0, dft_batch: int = 16, steps: int = 16,
         warmup: int = 4, force: bool = False) -> dict:
    """The fastest launch plan for ``tabs``' shape: ``{"form", "chunk", "ms_per_step",
    "timings": {"form/chunk": ms per step}, "key"}``.  Cached per shape key (in this process,
    and in the JSON file ``OEM_TUNE_CACHE`` names, if set); ``force`` times again."""
    from .solver import Stepper
    key = shape_key(tabs)
    with _lock:
        if not force:
            if key in _cache:
                return _cache[key]
            stored = _load_file().get(key)
            if stored is not None:
                _cache[key] = stored
                return stored
    # the library's default form for this scene, from a context that only prepares
    probe = Stepper(tabs, device=device, dft_batch=dft_batch)
    try:
        default = probe.launch
    finally:
        probe.close()
    need = warmup + steps
    if tabs.n_total < need:
        raise ValueError(f"the tables hold {tabs.n_total} steps; tuning steps {need}")
    timings = {}
    for form, chunk in candidates(tabs, default["form"]):
        st = Stepper(tabs, device=device, dft_batch=dft_batch, form=form, chunk=chunk)
        try:
            st.run(warmup)
            timings[f"{form}/{chunk}"] = st.timed_run(steps) / steps
        finally:
            st.close()
    best = min(timings, key=timings.get)
    form, chunk = best.split("/")
    plan = {"form": form, "chunk": int(chunk), "ms_per_step": round(timings[best], 5),
            "timings": {k: round(v, 5) for k, v in timings.items()}, "key": key,
            "default": default}
    with _lock:
        _cache[key] = plan
    _store_file(key, plan)
    return plan


def stepper(tabs: tables_mod.StepTables, device: int = 0, dft_batch: int = 16, **kw):
    """A :class:`openem_amd.solver.Stepper` of ``tabs`` on the tuned launch plan."""
    from .solver import Stepper
    plan = tune(tabs, device=device, dft_batch=dft_batch, **kw)
    return Stepper(tabs, device=device, dft_batch=dft_batch, form=plan["form"],
                   chunk=plan["chunk"])
