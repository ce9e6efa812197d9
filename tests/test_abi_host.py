"""CPU tests of the C ABI surface and of the host-side table builders (no GPU compute calls)."""

import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from openem_amd import _lib, cpml, scenes, tables, waveform
from openem_amd.model import Dispersion, lorentz_poles, eps_poles, trap_coeffs

ROOT = Path(__file__).resolve().parent.parent


def _header_functions():
    text = "".join(h.read_text() for h in sorted((ROOT / "include").glob("*.h")))
    return sorted(set(re.findall(r"^int (oem_\w+)\(", text, flags=re.M)))


def test_header_declarations_match_binding():
    assert _header_functions() == sorted(_lib.EXPORTS)


def test_library_loads_and_exports_every_declared_symbol():
    lib = _lib.load()
    assert lib.oem_version() == 1
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    syms = set(re.findall(r"\bT (oem_\w+)", out))
    for name in _header_functions():
        assert name in syms, name
        assert hasattr(lib, name)


def test_error_paths_without_a_device():
    """No GPU in this container: device count is 0 and context creation fails with a message
    (never a crash); argument errors are reported before touching the device."""
    lib = _lib.load()
    n = C.c_int32(-1)
    assert lib.oem_device_count(C.byref(n)) == _lib.OEM_OK
    d = _lib.Dims(nx=4, ny=4, nz=4, n_steps=10, dft_batch=8, ch=1.0, cw=1.0)
    ctx = C.c_void_p()
    if n.value == 0:
        rc = lib.oem_ctx_create(C.byref(d), 0, C.byref(ctx))
        assert rc == _lib.OEM_EHIP and not ctx.value
        assert "no HIP device" in _lib.last_error(lib)
    bad = _lib.Dims(nx=4, ny=4, nz=4, n_steps=10, dft_batch=100, ch=1.0, cw=1.0)
    assert lib.oem_ctx_create(C.byref(bad), 0, C.byref(ctx)) == _lib.OEM_EINVAL
    assert "dft_batch" in _lib.last_error(lib)
    bad = _lib.Dims(nx=4, ny=4, nz=4, n_steps=10, dft_batch=8, ch=1.0, cw=1.0)
    bad.pml_lo[2] = 3
    bad.pml_hi[2] = 3
    assert lib.oem_ctx_create(C.byref(bad), 0, C.byref(ctx)) == _lib.OEM_EINVAL
    assert lib.oem_run(None, 1) == _lib.OEM_EINVAL


def test_material_classes_roundtrip():
    sc = scenes.mini_scene(sigma_on=True, pec_on=True)
    t = tables.build(sc)
    cls, lut = t.material_classes()
    assert cls.shape == sc.shape and lut.shape[1] == tables.LUT_STRIDE
    rows = lut[cls]
    for c in range(3):
        assert np.array_equal(rows[..., 2 * c], t.ca[c].astype(np.float32))
        assert np.array_equal(rows[..., 2 * c + 1], t.cb[c].astype(np.float32))
    flags = rows[..., 6].view(np.uint32)
    for c in range(3):
        assert np.array_equal((flags >> c) & 1, t.disp_flag[c].astype(np.uint32))


def test_e_coefficients_closed_form():
    sc = scenes.mini_scene(sigma_on=True, lorentz_on=False, pec_on=True)
    ca, cb = tables.e_coeffs(sc)
    s = sc.sigma_ex * sc.dt / (2 * cpml.EPSILON_0)
    want_ca = (sc.eps_ex - s) / (sc.eps_ex + s)
    want_ca.reshape(-1)[sc.pec_ex] = 0.0
    assert np.allclose(ca[0], want_ca, rtol=1e-15, atol=0)
    assert np.all(cb[0].reshape(-1)[sc.pec_ex] == 0)


def test_cpml_profile_closed_form():
    """tidy3d's grading: sigma = smax/(eta0 dl) * s^3 and kappa = 1 + 2 s^3 at the E/H depths
    (reference openem/cpml.py:55-112)."""
    p = cpml.PMLParams(num_layers=12)
    dl, dt = 20e-9, 3.8e-17
    c = cpml.face_coeffs(p, dl, "lo", dt)
    s_e = (12 - np.arange(12)) / 12
    sig = 1.5 / (cpml.ETA_0 * dl) * s_e ** 3
    kap = 1 + 2 * s_e ** 3
    b = np.exp(-dt / cpml.EPSILON_0 * sig / kap)
    assert np.allclose(c.b_E, b, rtol=1e-14)
    assert np.allclose(c.a_E, (b - 1) * sig / sig / kap, rtol=1e-14)
    assert np.allclose(c.inv_kappa_E, 1 / kap)
    hi = cpml.face_coeffs(p, dl, "hi", dt)
    assert hi.a_E[0] == 0.0 and hi.b_E[0] == 1.0          # innermost E layer at depth 0


def test_gaussian_pulse_invariants():
    """Restated tidy3d GaussianPulse: the DC-removed real waveform integrates to ~0, the
    envelope peaks at offset*twidth, and its spectrum peaks near freq0."""
    g = waveform.GaussianPulse(freq0=3e14, fwidth=6e13)
    dt = 1e-17
    t = np.arange(0, 2 * g.end_time(), dt)
    a = g.amp_time(t)
    assert abs(np.sum(a.real) * dt) < 1e-6 * np.abs(a.real).max() * g.twidth
    env = np.abs(a)
    assert abs(t[np.argmax(env)] - g.offset * g.twidth) < 0.02 * g.twidth
    f = np.linspace(1e14, 5e14, 401)
    spec = np.abs(np.exp(2j * np.pi * f[:, None] * t[None, :]) @ a.real)
    assert abs(f[np.argmax(spec)] - 3e14) < 0.05 * 3e14


def test_lorentz_pole_conversion_matches_continuous_model():
    de, w0, delta = 1.7, 2.1e15, 3e13
    poles = lorentz_poles(de, w0, delta)
    w = np.linspace(1e14, 4e15, 50)
    want = 2.0 + de * w0 ** 2 / (w0 ** 2 - 2j * w * delta - w ** 2)
    assert np.allclose(eps_poles(2.0, poles, w), want, rtol=1e-12)


def test_trapezoid_coefficients_closed_form():
    q, r, dt = complex(-1e13, -2e15), complex(0, 3e30), 4e-17
    am1, b = trap_coeffs([(q, r)], dt)
    A = (1 + q * dt / 2) / (1 - q * dt / 2)
    assert abs(am1[0] - (A - 1)) < 1e-15 and abs(b[0] - r * dt / 2 / (1 - q * dt / 2)) < 1e-3


def test_dispersion_validation_rejects_duplicates():
    d = Dispersion(comp=np.array([0, 0], np.int32), cell=np.array([5, 5], np.int32),
                   pole_ofs=np.array([0, 1, 2], np.int32), am1=np.zeros(2, complex),
                   b=np.zeros(2, complex), g=np.zeros(2))
    with pytest.raises(ValueError, match="duplicate"):
        d.validate((4, 4, 4))


def test_plane_source_coefficients():
    sc = scenes.slab_scene(size=(0.04e-6, 0.04e-6, 3e-6))
    t = tables.build(sc)
    ps = t.plane_srcs[0]
    assert (ps.ks, ps.kh) == (37, 36)
    assert np.isclose(ps.coef_h, sc.dt / cpml.MU_0 / 20e-9, rtol=1e-12)
    assert np.isclose(ps.coef_e, 1 / 20e-9, rtol=1e-12)


def test_box_soup_structure():
    sc = scenes.box_soup_scene(n=64, n_boxes=6, num_steps=20)
    assert sc.shape == (64, 64, 64)
    sc.dispersion.validate(sc.shape)
    t = tables.build(sc, 20)
    assert t.axes[0].pml_lo == 12 and t.axes[2].pml_hi == 12
    assert len(t.dipoles) == 1 and t.dipoles[0].comp.size == 2     # Ez between two z centres
    cls, lut = t.material_classes()
    assert lut.shape[0] < 100
