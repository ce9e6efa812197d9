"""CPU tests of the torch autograd wrapper's plumbing (the device adjoint replaced by known
per-entry gradients): J per design, and backward = incoming gradient x dJ/d eps summed over
the three eps components and every cell a pixel drives."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from openem_amd import adjoint, autograd  # noqa: E402


def test_backward_scales_the_per_pixel_sums(monkeypatch):
    npx, layers = 3, 2
    cells = np.arange(npx * npx * layers).reshape(npx, npx, layers) + 100
    rng = np.random.default_rng(0)
    fake = {}

    def build(d):
        return ("scene", float(d.sum())), cells

    def grads(jobs, concurrency, **kw):
        out = []
        for (sc, mon, comp, cell) in jobs:
            assert mon == "out" and comp.size == 3 * cells.size
            assert np.array_equal(cell, np.tile(cells.ravel(), 3))
            g = rng.standard_normal(comp.size)
            fake[sc[1]] = g
            out.append(adjoint.Gradient(J=sc[1] ** 2, a_plus=0j, grad=g, steps=1, segments=1))
        return out

    monkeypatch.setattr(adjoint, "mode_power_gradients", grads)
    designs = torch.tensor(rng.uniform(2, 12, (2, npx, npx)), requires_grad=True)
    J = autograd.mode_power(designs, build)
    sums = designs.detach().sum(dim=(1, 2)).numpy()
    assert np.allclose(J.detach().numpy(), sums ** 2)
    w = torch.tensor([0.5, -3.0], dtype=torch.float64)
    (w * J).sum().backward()
    for q in range(2):
        per = fake[float(sums[q])].reshape(3, npx, npx, layers).sum(axis=(0, 3))
        assert np.allclose(designs.grad[q].numpy(), w[q].item() * per, rtol=1e-12, atol=0)


def test_single_design_without_batch_axis(monkeypatch):
    cells = np.arange(8).reshape(2, 2, 2)

    def grads(jobs, concurrency, **kw):
        return [adjoint.Gradient(J=7.0, a_plus=0j, grad=np.ones(24), steps=1, segments=1)
                for _ in jobs]

    monkeypatch.setattr(adjoint, "mode_power_gradients", grads)
    d = torch.ones(2, 2, dtype=torch.float64, requires_grad=True)
    J = autograd.mode_power(d, lambda x: (None, cells), batched=False)
    assert J.shape == () and J.item() == 7.0
    J.backward()
    assert np.array_equal(d.grad.numpy(), np.full((2, 2), 6.0))


def test_cells_must_follow_the_design_shape(monkeypatch):
    monkeypatch.setattr(adjoint, "mode_power_gradients", lambda *a, **k: [])
    d = torch.ones(1, 2, 2, dtype=torch.float64)
    with pytest.raises(ValueError):
        autograd.mode_power(d, lambda x: (None, np.arange(6).reshape(3, 2)))


def test_overlap_power_passes_its_weights_and_scales_backward(monkeypatch):
    """The plane-projection objective: each job carries the weights, J per design comes back,
    backward is the incoming gradient times dJ/d eps summed per pixel."""
    cells = np.arange(2 * 2 * 3).reshape(2, 2, 3) + 40
    K = np.full((4, 5, 6), 0.5 + 0.25j)
    seen = []

    def grads(jobs, concurrency, **kw):
        out = []
        for (sc, mon, w, comp, cell) in jobs:
            assert mon == "T" and w is not None and np.array_equal(w, K)
            assert np.array_equal(cell, np.tile(cells.ravel(), 3))
            seen.append(kw)
            out.append(adjoint.Gradient(J=2.0 * sc, a_plus=0j, grad=np.arange(comp.size) * 1.0,
                                        steps=1, segments=1))
        return out

    monkeypatch.setattr(adjoint, "overlap_power_gradients", grads)
    d = torch.ones(3, 2, 2, dtype=torch.float64, requires_grad=True)
    J = autograd.overlap_power(d, lambda x: (float(x.sum()), cells), "T", K, f_index=1)
    assert np.allclose(J.detach().numpy(), 8.0)
    assert all(k == {"f_index": 1} for k in seen)
    J.sum().backward()
    per = np.arange(36.0).reshape(3, 2, 2, 3).sum(axis=(0, 3))
    assert np.allclose(d.grad.numpy(), np.broadcast_to(per, (3, 2, 2)))
