"""Writes ``scene_v10_small.npz`` with the reference's own ``openem.serialize.save``.

The reference's scene file layout is pinned by no closed form, so this script imports the
reference's numpy-only modules (``openem.model``, ``openem.grid``, ``openem.cpml``,
``openem.waveform``, ``openem.serialize``) from the read-only reference checkout, builds a small
Scene from fixed seeded arrays with the reference's own classes, and saves it.  The resulting
npz (data only: the arrays below and the keys ``save`` chose for them) is committed as the
fixture that ``tests/test_scene_io_host.py`` reads; nothing of the reference is committed or
needed at test time.  Run from the repository root:  python tests/golden/make_scene_npz.py
"""

import sys
from pathlib import Path

import numpy as np

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent / "scene_v10_small.npz"


def main() -> None:
    sys.path.insert(0, str(REF))
    from openem import cpml, model, serialize, waveform
    from openem.grid import Axis, Grid

    rng = np.random.default_rng(20260)
    nx, ny, nz = 6, 5, 9
    dl = 2e-8
    ex_ = np.concatenate([[0.0], np.cumsum(rng.uniform(0.8, 1.2, nx) * dl)]) - 5e-8
    ey_ = np.arange(ny + 1) * dl - 4e-8
    ez_ = np.arange(nz + 1) * dl - 9e-8
    grid = Grid(Axis(ex_, "Periodic", "Periodic"), Axis(ey_, "PECBoundary", "PECBoundary"),
                Axis(ez_, "PML", "PML"))
    shp = (nx, ny, nz)
    eps = [1.0 + rng.uniform(0, 3, shp) for _ in range(3)]
    sig = [rng.uniform(0, 1e3, shp) for _ in range(3)]
    pec = [np.array([3, 17], dtype=np.int64), np.array([], dtype=np.int64),
           np.array([40], dtype=np.int64)]
    layers = 3

    def face(o):
        return cpml.PMLCoeffs(*(rng.uniform(0.1, 0.9, layers) + o for _ in range(6)))
    pml = {(2, "lo"): face(0.0), (2, "hi"): face(1.0)}
    n = 12
    dt = 3.5e-17

    def wf(complex_half=False):
        ai = rng.normal(size=n + 1)
        ah = rng.normal(size=n + 1)
        ic = ai + 1j * rng.normal(size=n + 1)
        hc = ah + 1j * rng.normal(size=n + 1) if complex_half else None
        return waveform.Waveform(amp_int=ai, amp_half=ah, amp_int_complex=ic, dt=dt,
                                 amp_half_complex=hc)
    src = model.PlaneWaveSource(axis=2, plane_index=4, direction=1, pol_axis=0, waveform=wf(),
                                ex_inc=rng.normal(size=n + 1), hy_inc=rng.normal(size=n + 1),
                                angle_theta=0.0, num_freqs=1, center=-1.3e-8)
    dips = [model.PointDipole(component=2, indices=np.array([[1, 2, 3], [1, 2, 4]]),
                              coef=np.array([0.25, 0.75]) * 1e21, waveform=wf(True),
                              magnetic=False),
            model.PointDipole(component=0, indices=np.array([[2, 2, 2]]),
                              coef=np.array([1e21]), waveform=wf(), magnetic=True)]
    m1 = rng.normal(size=(ny, nz)) + 1j * rng.normal(size=(ny, nz))
    msrc = model.ModeSource(plane_index=3, direction=-1, ey_inc=m1, ez_inc=2 * m1,
                            hy_inc=3 * m1, hz_inc=4 * m1,
                            amp_e=rng.normal(size=(n + 1,)) + 0j,
                            amp_h=rng.normal(size=(n + 1,)) + 0j, n_eff=2.1 + 0.01j, axis=0)
    disp = model.Dispersion(comp=np.array([0, 2, 1]), cell=np.array([5, 77, 120]),
                            pole_ofs=np.array([0, 2, 3, 4]),
                            am1=rng.normal(size=4) + 1j * rng.normal(size=4),
                            b=rng.normal(size=4) + 1j * rng.normal(size=4),
                            g=rng.normal(size=3))
    freqs = np.array([1.8e14, 2.0e14, 2.2e14])
    fmons = [model.FluxMonitor(name="T", axis=2, plane_index=6, normal_dir=1, freqs=freqs,
                               source_spectrum=np.ones(3, complex),
                               transverse=((1, 4), (0, 3)), apodization=(1e-15, None, 2e-15),
                               t_bounds=((-3e-8, 4e-8), (-2e-8, 3e-8))),
             model.FluxMonitor(name="R", axis=1, plane_index=2, normal_dir=-1, freqs=freqs[:2],
                               source_spectrum=np.ones(2, complex))]
    fldm = [model.FieldMonitor(name="box", freqs=freqs, origin=(1, 0, 2), box=(3, 4, 5),
                               apodization=None, interval_space=(1, 2, 1)),
            model.FieldMonitor(name="mplane", freqs=freqs[1:], origin=(0, 0, 5), box=(6, 5, 2))]
    sc = model.Scene(grid=grid, dt=dt, num_time_steps=n, shutoff=1e-5, eps_ex=eps[0],
                     eps_ey=eps[1], eps_ez=eps[2], pml=pml, sigma_ex=sig[0], sigma_ey=sig[1],
                     sigma_ez=sig[2], pec_ex=pec[0], pec_ey=pec[1], pec_ez=pec[2],
                     dispersion=disp, bloch_k=(0.0, 0.0, 0.0), sources=[src], dipoles=dips,
                     mode_sources=[msrc], flux_monitors=fmons, field_monitors=fldm,
                     mode_monitors=[model.ModeMonitorSpec(name="mm", plane_name="mplane")],
                     permittivity_monitors=[model.PermittivityMonitor(
                         name="perm", freqs=freqs[:1], origin=(2, 1, 3), box=(3, 2, 4))],
                     projection_monitors=[model.ProjectionMonitor(
                         name="far", surface_names=("box", "mplane"), normal_dirs=("+", "-"))],
                     box_flux_monitors=[model.BoxFluxMonitor(name="net", face_names=("T", "R"))])
    serialize.save(sc, OUT)
    print("wrote", OUT, OUT.stat().st_size, "bytes")


if __name__ == "__main__":
    main()
