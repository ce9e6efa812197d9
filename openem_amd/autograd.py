"""PyTorch autograd over the device adjoint: an inverse-design objective as a differentiable
function of the design permittivities.

The reference reaches gradients by patching tidy3d's autograd engine so that its forward and
adjoint solves run locally (``openem/nb/autograd_hook.py:342-411``, ``_run_tidy3d`` /
``_run_async_tidy3d``) and tidy3d assembles the vector-Jacobian product.  tidy3d is absent
here, and the adjoint is this library's own discrete transpose of the time step
(:mod:`openem_amd.adjoint`), so the wrapper is a plain :class:`torch.autograd.Function`:
forward runs the stepper with the adjoint (one forward run and one backward sweep per design on
the MI355X) and keeps ``dJ/d eps``; backward scales it by the incoming gradient.

Only this module imports torch; the stepper itself never does.
"""

from __future__ import annotations

from typing import Callable

import numpy as np
import torch

from . import adjoint

#: ``build(design) -> (scene, cells)``: a Scene whose permittivity at ``cells[p, ...]`` (flat
#: cell indices, all three eps components) is the value of design pixel ``p``
SceneBuilder = Callable[[np.ndarray], tuple]


def _pixel_gradients(designs: np.ndarray, build: SceneBuilder, monitor: str, concurrency: int,
                     kw: dict, weights=None):
    """J and dJ/d pixel per design: the mode power of ``monitor`` (``weights`` None) or the power
    of its projection on ``weights`` (:func:`openem_amd.adjoint.overlap_power_gradient`)."""
    jobs, shapes = [], []
    for d in designs:
        sc, cells = build(d)
        cells = np.asarray(cells)
        if cells.shape[:d.ndim] != d.shape:
            raise ValueError(f"design cells {cells.shape} do not start with the design shape "
                             f"{d.shape}")
        comp, cell = np.repeat(np.arange(3), cells.size), np.tile(cells.ravel(), 3)
        jobs.append((sc, monitor, comp, cell) if weights is None
                    else (sc, monitor, weights, comp, cell))
        shapes.append(cells.shape)
    if weights is None:
        res = adjoint.mode_power_gradients(jobs, concurrency=concurrency, **kw)
    else:
        res = adjoint.overlap_power_gradients(jobs, concurrency=concurrency, **kw)
    J = np.array([g.J for g in res])
    grads = []
    for g, shp, d in zip(res, shapes, designs):
        # sum over the three eps components and every cell a pixel drives
        per = g.grad.reshape(3, *shp).sum(axis=0)
        grads.append(per.reshape(*d.shape, -1).sum(axis=-1))
    return J, np.stack(grads)


class ModePower(torch.autograd.Function):
    """``J = |a+|^2`` of mode monitor ``monitor`` (or, with ``weights``, ``J = |sum weights .
    phasors|^2`` of plane monitor ``monitor``) for each design of a batch ``(B, *pixels)``; the
    vector-Jacobian product is the device adjoint's ``dJ/d eps`` per pixel."""

    @staticmethod
    def forward(ctx, designs, build, monitor, concurrency, kw, weights=None):
        if designs.dim() < 2:
            raise ValueError("designs must be a batch (B, *pixels)")
        host = designs.detach().to("cpu", torch.float64).numpy()
        J, g = _pixel_gradients(host, build, monitor, concurrency, kw, weights)
        ctx.save_for_backward(torch.from_numpy(g).to(designs.device, designs.dtype))
        return torch.from_numpy(J).to(designs.device, designs.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        shape = (-1,) + (1,) * (g.dim() - 1)
        return grad_out.reshape(shape) * g, None, None, None, None, None


def mode_power(designs: torch.Tensor, build: SceneBuilder, monitor: str = "out",
               batched: bool = True, concurrency: int = 2, **kw) -> torch.Tensor:
    """Differentiable ``J`` per design: ``designs`` is ``(B, *pixels)`` (one design without the
    batch axis when ``batched`` is False), ``build(design)`` returns ``(scene, cells)`` as
    :func:`openem_amd.scenes.bend_scene` does; ``kw`` goes to
    :func:`openem_amd.adjoint.mode_power_gradient` (``checkpoint_every``, ``record_bytes``,
    ``f_index`` ...).  ``concurrency`` designs run at once, each on its own context and stream."""
    if not batched:
        return ModePower.apply(designs.unsqueeze(0), build, monitor, concurrency, kw)[0]
    return ModePower.apply(designs, build, monitor, concurrency, kw)


def overlap_power(designs: torch.Tensor, build: SceneBuilder, monitor: str, weights,
                  batched: bool = True, concurrency: int = 2, **kw) -> torch.Tensor:
    """Differentiable ``J = |sum weights . phasors|^2`` of plane monitor ``monitor`` per design
    (``weights`` (4, n1, n2) complex over its E t1, E t2, H t1, H t2 slots: a diffraction
    order's projection on a Bloch cell, for one; real or complex fields), otherwise as
    :func:`mode_power`; ``kw`` goes to :func:`openem_amd.adjoint.overlap_power_gradient`."""
    w = np.asarray(weights, dtype=np.complex128)
    if not batched:
        return ModePower.apply(designs.unsqueeze(0), build, monitor, concurrency, kw, w)[0]
    return ModePower.apply(designs, build, monitor, concurrency, kw, w)
