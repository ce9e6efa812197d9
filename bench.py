"""Benchmark: Yee Mcell-steps/s and HBM-roofline fraction on one MI355X.

Workload (the roofline configuration): 512x512x512 cells at dl = 20 nm, fp32 fields, 12-layer
CPML on all six faces, 32 seeded boxes (plain eps or 2-pole Lorentz ADE), a centred Ez point
dipole and a field monitor on the central xy plane (8 frequencies).  One "step" is one full
Yee time step of that scene (H sweep, ADE, E sweep, dipole, DFT), inputs resident in HBM.

Prints one JSON line.  ``value`` = Mcell-steps/s over the timed steps (HIP events on the
stepper's stream).  ``roofline`` is for the dominant kernel: algorithmic bytes per launch /
its average device time from event pairs around each launch.  ``cpu_baseline`` is the float64
NumPy oracle on a bounded sample of the same workload.  ``serving`` is the batched-serving
half of the metric: the first ``--serve-requests`` requests of the seeded serving stream (Poisson
arrivals at 20 / s, cubic sides log-uniform in [64, 320], 2000-8000 steps, 1-3 flux monitors of
5-21 frequencies, slab / box-soup / Lorentz-sphere scenes) served by continuous batching and by
the serial plain loop on the same GPU: sims/s and p50/p99 time-to-result of each.  ``bloch``
times the complex field path on the full-size Bloch / oblique unit cell (256 x 256 x 600).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Yee Mcell-steps/s & %HBM peak; batched sims/s + p50/p99 latency vs serial"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(tabs) -> dict:
    """Algorithmic bytes per launch in the reference's stepping accounting: 78 B per cell-step
    (the H sweep reads E 12 B and reads+writes H 24 B = 36 B; the E sweep reads H 12 B,
    reads+writes E 24 B and reads three uint16 material indices 6 B = 42 B), + 32 B per cell per
    CPML slab axis (2 psi read+write on each side, 16 B per sweep), + 4 B per dispersive
    component (the history term the E update reads).  Complex (Bloch) fields double the field and
    psi terms.  The fused H+E step does both sweeps' work in one launch, so its figure is the
    sum.  ``minimal_*`` is what each kernel of this build must move at the least: the fused step
    reads E and H once and writes them once (48 B) plus one 2-byte class index per cell.
    """
    nx, ny, nz = tabs.shape
    n = nx * ny * nz
    slab_cells = 0
    for a, ax in enumerate(tabs.axes):
        w = ax.pml_lo + ax.pml_hi
        slab_cells += n // tabs.shape[a] * w
    disp_comps = 0 if tabs.disp is None else tabs.disp.comp.size
    c = 2 if tabs.bloch else 1          # complex (Bloch) fields: every field value is 8 B
    h = c * 36 * n + c * 16 * slab_cells
    e = c * 36 * n + 6 * n + c * 16 * slab_cells + c * 4 * disp_comps
    he_min = c * 48 * n + 2 * n + c * 32 * slab_cells + c * 4 * disp_comps
    return {"update_h": h, "update_e": e, "update_he": h + e,
            "minimal_update_h": h, "minimal_update_e": e - 4 * n, "minimal_update_he": he_min,
            "cells": n, "slab_cells": slab_cells}


def measured_traffic(kernel: str):
    """HBM bytes per launch of ``kernel`` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (profiles/yee_box_soup_512.md), or None when absent."""
    p = ROOT / "profiles" / "yee_box_soup_512.md"
    if not p.exists():
        return None
    text = p.read_text()
    try:
        blob = text.split("```json", 1)[1].split("```", 1)[0]
        return json.loads(blob)["traffic_bytes_per_launch"].get("k_" + kernel)
    except (IndexError, KeyError, ValueError):
        return None


def _oracle_sample_time(n: int, steps: int, nx_sample: int, barrier=None) -> float:
    """Seconds for ``steps`` float64 oracle steps on the (nx_sample, n, n) box-soup sample (one
    untimed step first); with ``barrier``, the timed steps start together with the other
    workers'."""
    from openem_amd import scenes, tables
    from oracle.yee64 import Oracle64
    sc = scenes.box_soup_scene(shape=(nx_sample, n, n), num_steps=steps + 1)
    tabs = tables.build(sc, steps + 1)
    o = Oracle64(tabs)
    o.step()                       # first step touches every array once
    if barrier is not None:
        barrier.wait()
    t0 = time.perf_counter()
    o.run(steps)
    return time.perf_counter() - t0


def _oracle_worker(n, steps, nx_sample, barrier, out):
    out.put(_oracle_sample_time(n, steps, nx_sample, barrier))


def cpu_baseline(n: int, steps: int = 4, nx_sample: int = 64, workers: int = 8) -> dict:
    """The float64 oracle on the host: one process on one x-slab sample, and ``workers`` processes
    stepping one x-slab each at the same time (an x-slab decomposition without its halo
    exchange, which would add one plane per step and neighbour).  ``steps`` timed steps after
    one untimed step: fewer than the 10-20 of the baseline plan, to keep the whole benchmark
    inside its time budget (the per-step time of the oracle does not change from step to step,
    its arrays being fixed in size)."""
    import multiprocessing as mp
    dt = _oracle_sample_time(n, steps, nx_sample)
    cells = nx_sample * n * n
    out = {"value": round(cells * steps / dt / 1e6, 3), "unit": "Mcell-steps/s", "cores": 1,
           "kind": "port",
           "sample": f"float64 NumPy oracle, {steps} timed steps (after one untimed step) on a "
                     f"{nx_sample}x{n}x{n} x-slab of the same box-soup workload (same seed, PML, "
                     f"boxes, dipole, monitor), single process; the baseline plan's 10-20 steps "
                     f"cut to {steps} for the benchmark's time budget"}
    if workers > 1:
        ctx = mp.get_context("spawn")          # fresh interpreters: no fork of this process
        bar, q = ctx.Barrier(workers), ctx.Queue()
        procs = [ctx.Process(target=_oracle_worker, args=(n, steps, nx_sample, bar, q))
                 for _ in range(workers)]
        for pr in procs:
            pr.start()
        times = [q.get() for _ in procs]
        for pr in procs:
            pr.join()
        out["parallel"] = {"value": round(workers * cells * steps / max(times) / 1e6, 3),
                           "unit": "Mcell-steps/s", "cores": workers,
                           "sample": f"{workers} processes, each the single-process sample "
                                     f"(one {nx_sample}-plane x-slab), timed together: an "
                                     f"x-slab decomposition without its halo exchange"}
    return out


def serving_bench(n_req: int, arena_gib: int, workers: int, solo_lanes: int = 2,
                  best_rate: float | None = None, solo_running: int = 2) -> dict:
    """The seeded serving stream, served by continuous batching and by the serial plain loop.
    ``best_rate`` (Mcell-steps/s, the single-scene headline rate) gives the stream's lower
    bound: its cell-steps at that rate, the time no schedule of these scenes can beat on this
    stepper; the serial loop's summed stepping time is the device time the same scenes take
    one at a time."""
    from openem_amd import scenes, serving
    specs = scenes.serving_requests(256, rate=20.0, seed=0)[:n_req]
    cs = sum(sp["side"] ** 3 * sp["n_steps"] for sp in specs)
    print(f"[bench] serving {n_req} requests ({cs / 1e12:.2f} Tcell-steps): batched",
          file=sys.stderr, flush=True)
    b = serving.serve(specs, arena_bytes=arena_gib << 30, workers=workers,
                      solo_lanes=solo_lanes, solo_running=solo_running, report_every=16)
    sb = serving.summary(b)
    print(f"[bench] batched {sb}, peak running {b['peak_running']}; serial", file=sys.stderr,
          flush=True)
    s = serving.serve_serial(specs, report_every=16)
    ss = serving.summary(s)
    print(f"[bench] serial {ss}", file=sys.stderr, flush=True)
    errors = [f"batched #{r.spec['id']} ({r.spec['kind']} {r.spec['side']}^3): {r.error}"
              for r in b["requests"] if r.error] + \
             [f"serial #{r.spec['id']} ({r.spec['kind']} {r.spec['side']}^3): {r.error}"
              for r in s["requests"] if r.error]
    for e in errors:
        print(f"[bench] failed request {e}", file=sys.stderr, flush=True)
    worst = 0.0
    for rb, rs in zip(b["requests"], s["requests"]):
        for x, y in zip(rb.flux, rs.flux):
            worst = max(worst, float(np.abs(x - y).max() / max(np.abs(y).max(), 1e-300)))
    out = {"requests": n_req, "stream": "seed 0, Poisson 20/s, side log-U[64,320], "
                                        "2000-8000 steps, 1-3 flux monitors x 5-21 freqs",
           "cell_steps": cs, "sims_per_s": round(sb["sims_per_s"], 4),
           "p50_s": round(sb["p50_s"], 3), "p99_s": round(sb["p99_s"], 3),
           "span_s": round(sb["span_s"], 2),
           "serial_sims_per_s": round(ss["sims_per_s"], 4),
           "serial_p50_s": round(ss["p50_s"], 3), "serial_p99_s": round(ss["p99_s"], 3),
           "serial_span_s": round(ss["span_s"], 2),
           "serial_stepping_s": round(sum(r.run_s for r in s["requests"]), 2),
           "speedup": round(sb["sims_per_s"] / ss["sims_per_s"], 3),
           "peak_batch": b["peak_running"], "solo_lanes": solo_lanes,
           "solo_running": solo_running,
           "failed": sb["failed"] + ss["failed"], "errors": errors[:8],
           "max_rel_flux_diff_vs_serial": worst, "arena_gib": arena_gib}
    if best_rate:
        bound = cs / (best_rate * 1e6)
        out["bound"] = {"span_s": round(bound, 2), "rate_mcell_steps_per_s": best_rate,
                        "what": "stream cell-steps at the single-scene headline rate",
                        "batched_frac": round(bound / sb["span_s"], 4),
                        "speedup_ceiling": round(ss["span_s"] / bound, 3)}
    return out


def adjoint_bench(n_designs: int, steps: int, check_pixels: int, n: int = 400,
                  concurrency: int = 3, record_gib: float = 24.0) -> dict:
    """Inverse-design gradients of the 90-degree bend: 400 x 400 x 48 cells, a 160 x 160-pixel
    design region, the fundamental mode in and the output mode power as the objective, 6000
    steps.  Each design runs forward + device adjoint (the forward run records E at the design
    cells when the record fits in HBM, else checkpoints every 500 steps and recomputes);
    ``concurrency`` designs are in flight at once, each on its own context and stream.  The
    gradient of the first design is checked by central differences of the same stepper
    (h = 1e-3 in eps) on ``check_pixels`` seeded pixels, the perturbed runs ``concurrency`` at
    a time."""
    from concurrent.futures import ThreadPoolExecutor
    from openem_amd import adjoint, scenes
    t0 = time.perf_counter()
    jobs, shapes = [], []
    for d in range(n_designs):
        sc, cells = scenes.bend_scene(design=scenes.bend_design(seed=d), n=n, num_steps=steps)
        jobs.append((sc, "out", np.repeat(np.arange(3), cells.size), np.tile(cells.ravel(), 3)))
        shapes.append(cells.shape)
    t1 = time.perf_counter()
    print(f"[bench] adjoint: {n_designs} designs built in {t1 - t0:.1f} s", file=sys.stderr,
          flush=True)

    def done(d, g):
        print(f"[bench] design {d}: J = {g.J:.6e} at {time.perf_counter() - t1:.1f} s",
              file=sys.stderr, flush=True)

    res = adjoint.mode_power_gradients(jobs, concurrency=concurrency, checkpoint_every=500,
                                       on_done=done, record_bytes=int(record_gib * (1 << 30)))
    wall = time.perf_counter() - t1
    jobs.clear()
    grads = [(g.J, g.grad.reshape(3, *shp).sum(axis=(0, 3))) for g, shp in zip(res, shapes)]
    check = None
    if check_pixels:
        rng = np.random.default_rng(1)
        npx = shapes[0][0]
        pix = [(int(rng.integers(npx)), int(rng.integers(npx))) for _ in range(check_pixels)]

        def power(args):
            (a, b), sgn = args
            des = scenes.bend_design(seed=0)
            des[a, b] += sgn * 1e-3
            sc, _ = scenes.bend_scene(design=des, n=n, num_steps=steps)
            return _mode_power(sc)

        t2 = time.perf_counter()
        with ThreadPoolExecutor(max(1, concurrency)) as ex:
            vals = list(ex.map(power, [(p, sgn) for p in pix for sgn in (+1, -1)]))
        scale = float(np.abs(grads[0][1]).max())
        errs = []
        for q, (a, b) in enumerate(pix):
            fd = (vals[2 * q] - vals[2 * q + 1]) / 2e-3
            errs.append(abs(fd - grads[0][1][a, b]) / scale)
        print(f"[bench] finite differences at {check_pixels} pixels in "
              f"{time.perf_counter() - t2:.1f} s: max error {max(errs):.3e} of max |grad|",
              file=sys.stderr, flush=True)
        check = {"pixels": check_pixels, "h": 1e-3, "seed": 1,
                 "max_abs_err_over_max_grad": max(errs),
                 "median_abs_err_over_max_grad": float(np.median(errs)),
                 "max_abs_grad": scale}
    return {"designs": n_designs, "concurrency": concurrency, "steps": steps,
            "checkpoint_every": 500, "segments": res[0].segments,
            "grid": [n, n, 48], "design_pixels": list(shapes[0][:2]),
            "s_per_design": round(wall / n_designs, 3),
            "designs_per_s": round(n_designs / wall, 4), "wall_s": round(wall, 2),
            "J": [float(j) for j, _ in grads], "fd_check": check,
            "setup_s": round(t1 - t0, 1)}


def bloch_bench(steps: int, warmup: int = 10, profile_steps: int = 10) -> dict:
    """The Bloch / oblique unit cell at its full size (256 x 256 x 600 complex cells, Bloch x/y,
    CPML z, oblique TF/SF box, Drude + Lorentz metal cylinders, four 21-frequency flux planes):
    the complex field path's throughput and the roofline of its dominant sweep."""
    from openem_amd import scenes, tables
    from openem_amd.solver import Stepper
    n_total = warmup + steps + profile_steps + 8
    t0 = time.time()
    sc = scenes.bloch_cell_scene(num_steps=max(n_total, 400))
    tabs = tables.build(sc, n_total)
    st = Stepper(tabs)
    setup = time.time() - t0
    try:
        st.run(warmup)
        ms = st.timed_run(steps)
        prof = st.profile_run(profile_steps)
    finally:
        st.close()
    cells = int(np.prod(sc.shape))
    ab = algorithmic_bytes(tabs)
    k_dom = max(("update_h", "update_e", "update_he"),
                key=lambda k: prof[k][0] / max(prof[k][1], 1))
    avg = prof[k_dom][0] / prof[k_dom][1]
    achieved = ab[k_dom] / (avg * 1e-3) / 1e9
    kern = {k: round(v[0] / v[1], 5) for k, v in prof.items() if v[1]}
    print(f"[bench] bloch cell {sc.shape}: setup {setup:.1f}s, per-kernel avg ms {kern}",
          file=sys.stderr, flush=True)
    return {"workload": "Bloch unit cell 256x256x600 complex (Bloch x/y, CPML z, oblique TF/SF, "
                        "Drude+Lorentz metal, 4 flux planes x 21 freqs)",
            "cells": cells, "steps": steps, "ms_per_step": round(ms / steps, 4),
            "mcell_steps_per_s": round(cells / (ms / steps * 1e-3) / 1e6, 1),
            "dtype": "complex f32",
            "roofline": {"bound": "hbm", "kernel": k_dom, "achieved": round(achieved, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "fused": k_dom == "update_he",
                         "algorithmic_bytes_per_launch": ab[k_dom],
                         "minimal_bytes_per_launch": ab["minimal_" + k_dom],
                         "avg_launch_ms": round(avg, 5)}}


def bloch_gradient_bench(steps: int, n_design: int = 4096, seed: int = 0) -> dict:
    """One gradient on the full-size Bloch cell (complex fields, the cell of
    :func:`bloch_bench` over ``steps`` steps): J = |sum K . phasors|^2 of the transmission plane
    with seeded weights K, dJ/d eps at ``n_design`` seeded (component, cell) entries in the
    cylinders' layer, the forward run recording the design (no checkpoints); beside it the
    forward run alone.  The gradient's wall time includes its table build."""
    from openem_amd import adjoint, scenes, tables
    from openem_amd.solver import Stepper
    sc = scenes.bloch_cell_scene(num_steps=steps)
    tabs = tables.build(sc)
    m = tabs.monitors[[q.name for q in tabs.monitors].index("T")]
    rng = np.random.default_rng(seed)
    K = rng.standard_normal((4, m.n1, m.n2)) + 1j * rng.standard_normal((4, m.n1, m.n2))
    nx, ny, nz = sc.shape
    d = sc.dispersion
    kd = d.cell % nz
    k_lo, k_hi = int(kd.min()), int(kd.max()) + 1                     # the cylinders' layer
    comp = rng.integers(0, 3, n_design)
    cell = (rng.integers(0, nx, n_design) * ny + rng.integers(0, ny, n_design)) * nz \
        + rng.integers(k_lo, k_hi, n_design)
    t0 = time.perf_counter()
    g = adjoint.overlap_power_gradient(sc, "T", K, comp, cell)
    wall = time.perf_counter() - t0
    st = Stepper(tabs)
    try:
        st.run(16)
        fwd = st.timed_run(steps - 16) / (steps - 16)
    finally:
        st.close()
    print(f"[bench] bloch gradient: {wall:.2f} s for {steps} steps ({fwd:.4f} ms per forward "
          f"step), J {g.J:.6e}", file=sys.stderr, flush=True)
    return {"workload": "Bloch unit cell 256x256x600 complex, J = |K . T-plane phasors|^2",
            "cells": int(np.prod(sc.shape)), "steps": steps, "design_entries": n_design,
            "s_per_gradient": round(wall, 3), "forward_ms_per_step": round(fwd, 4),
            "max_abs_grad": float(np.abs(g.grad).max())}


def _mode_power(sc) -> float:
    from openem_amd import solver
    r = solver.run(sc, use_shutoff=False, verbose=False)
    return float(abs(r.mode_amps["out"]["a_plus"][0]) ** 2)


def pool_bench(k: int, n: int, blocks: int = 8, dft_batch: int = 16) -> dict:
    """The serving pool's kernel throughput: ``k`` box soups of side ``n`` resident and stepped
    together (one graph of ``dft_batch`` steps per block over every scene), ``blocks`` blocks
    timed after one warm-up block, against one of the scenes stepped alone."""
    from openem_amd import scenes, serving, tables
    from openem_amd.solver import Stepper
    steps = (blocks + 1) * dft_batch
    tabs = tables.build(scenes.box_soup_scene(n=n, num_steps=steps + 32, n_boxes=8, seed=3),
                        steps + 32)
    pool = serving.Pool(int(serving.device_bytes(tabs, dft_batch) * k * 1.3) + (64 << 20),
                        dft_batch)
    sts = [Stepper(tabs, dft_batch=dft_batch, pool=pool) for _ in range(k)]
    try:
        for st in sts:
            pool.attach(st)
        pool.launch(1)
        pool.wait(0)
        t0 = time.perf_counter()
        pool.launch(blocks)
        pool.wait(0)
        dt = time.perf_counter() - t0
    finally:
        for st in sts:
            st.close()
        pool.close()
    alone = Stepper(tabs, dft_batch=dft_batch)
    try:
        alone.run(32)
        ms = alone.timed_run(blocks * dft_batch)
    finally:
        alone.close()
    cs = k * n ** 3 * blocks * dft_batch
    return {"scenes": k, "side": n, "steps": blocks * dft_batch,
            "pool_mcell_steps_per_s": round(cs / dt / 1e6, 1),
            "alone_mcell_steps_per_s": round(n ** 3 * blocks * dft_batch / ms / 1e3, 1)}


def _serving_estimate_s(specs, best_rate: float) -> float:
    """Wall seconds the serving leg (batched, then serial) is expected to take: the stream's
    cell-steps at 0.60 and 0.51 of the single-scene rate, host work included (the whole stream
    measured 0.60-0.66 and 0.51-0.55)."""
    cs = sum(sp["side"] ** 3 * sp["n_steps"] for sp in specs)
    return cs / (best_rate * 1e6) * (1 / 0.60 + 1 / 0.51)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--dft-batch", type=int, default=16,
                    help="steps per DFT snapshot batch and per captured step graph")
    ap.add_argument("--form", choices=["fused", "sweeps"], default=None,
                    help="step form of the headline run (skips the tuning)")
    ap.add_argument("--chunk", type=int, default=0,
                    help="x planes per work-group of the headline run (skips the tuning)")
    ap.add_argument("--tune", type=int, default=1,
                    help="time the launch candidates first and step on the fastest (0: the "
                         "library's default plan)")
    ap.add_argument("--serve-requests", type=int, default=256,
                    help="requests of the serving stream to run batched and serially (0: skip)")
    ap.add_argument("--serve-only", action="store_true")
    ap.add_argument("--arena-gib", type=int, default=96)
    ap.add_argument("--serve-workers", type=int, default=6)
    ap.add_argument("--solo-lanes", type=int, default=2)
    ap.add_argument("--adjoint-only", action="store_true")
    ap.add_argument("--adjoint-designs", type=int, default=32,
                    help="designs of the gradient leg (0: skip)")
    ap.add_argument("--adjoint-concurrency", type=int, default=3,
                    help="designs in flight at once, each on its own context and stream")
    ap.add_argument("--adjoint-steps", type=int, default=6000)
    ap.add_argument("--adjoint-check", type=int, default=16)
    ap.add_argument("--adjoint-record-gib", type=float, default=24.0,
                    help="device memory for E at the design cells over the whole run (0: "
                         "checkpoints every 500 steps and recomputation)")
    ap.add_argument("--bloch-steps", type=int, default=60,
                    help="steps of the full-size Bloch unit cell to time (0: skip)")
    ap.add_argument("--bloch-only", action="store_true",
                    help="only the Bloch unit cell leg")
    ap.add_argument("--bloch-gradient-only", type=int, default=0, metavar="STEPS",
                    help="only one gradient on the full-size Bloch cell over STEPS steps")
    ap.add_argument("--pool-only", type=int, nargs=2, metavar=("K", "N"),
                    help="only the serving pool's throughput: K box soups of side N stepped "
                         "together")
    ap.add_argument("--budget-s", type=float, default=540.0,
                    help="wall-clock budget of the whole run: a leg that would overrun it is "
                         "shortened (and says so in its record)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(a scene is stepped on a single MI355X)", file=sys.stderr)
        return 2
    t_start = time.perf_counter()

    from openem_amd import scenes, tables
    from openem_amd.solver import Stepper

    if args.adjoint_only:
        print(json.dumps({"metric": "inverse-design gradients (forward + device adjoint)",
                          "adjoint": adjoint_bench(args.adjoint_designs, args.adjoint_steps,
                                                   args.adjoint_check,
                                                   concurrency=args.adjoint_concurrency,
                                                   record_gib=args.adjoint_record_gib)}),
              flush=True)
        return 0
    if args.bloch_only:
        print(json.dumps({"metric": METRIC, "bloch": bloch_bench(args.bloch_steps)}), flush=True)
        return 0
    if args.bloch_gradient_only:
        print(json.dumps({"metric": "Bloch-cell gradient (forward + device adjoint)",
                          "bloch_gradient": bloch_gradient_bench(args.bloch_gradient_only)}),
              flush=True)
        return 0
    if args.pool_only:
        print(json.dumps({"metric": "serving pool throughput (Mcell-steps/s)",
                          "pool": pool_bench(*args.pool_only)}), flush=True)
        return 0
    if args.serve_only:
        print(json.dumps({"metric": METRIC, "serving": serving_bench(
            args.serve_requests, args.arena_gib, args.serve_workers, args.solo_lanes)}),
            flush=True)
        return 0

    n_total = max(args.warmup + args.dft_batch + args.steps + args.profile_steps + 8, 32)
    t0 = time.time()
    sc = scenes.box_soup_scene(n=args.n, num_steps=max(n_total, 4000))
    tabs = tables.build(sc, n_total)
    plan = None
    if args.form or args.chunk:
        st = Stepper(tabs, dft_batch=args.dft_batch, form=args.form or None, chunk=args.chunk)
    elif args.tune:
        from openem_amd import autotune
        plan = autotune.tune(tabs, dft_batch=args.dft_batch, steps=12, warmup=2)
        print(f"[bench] launch plan {plan['form']}/{plan['chunk']} of {plan['timings']} "
              f"(ms per step)", file=sys.stderr, flush=True)
        st = Stepper(tabs, dft_batch=args.dft_batch, form=plan["form"], chunk=plan["chunk"])
    else:
        st = Stepper(tabs, dft_batch=args.dft_batch)
    launch = st.launch
    setup_s = time.time() - t0
    print(f"[bench] setup {setup_s:.1f}s, {sc.shape}, classes {st.n_classes}, "
          f"ADE entries {tabs.disp.comp.size if tabs.disp else 0}, launch {launch}",
          file=sys.stderr, flush=True)
    # the warmup runs on to a whole graph block, so the timed steps replay captured blocks
    # (they start at a block boundary) rather than capture one
    warm = args.warmup + (-args.warmup) % args.dft_batch
    st.run(warm)
    t1 = time.perf_counter()
    ms = st.timed_run(args.steps)
    wall = time.perf_counter() - t1
    prof = st.profile_run(args.profile_steps)
    st.close()

    cells = int(np.prod(sc.shape))
    ms_step = ms / args.steps
    value = cells / (ms_step * 1e-3) / 1e6
    ab = algorithmic_bytes(tabs)
    k_dom = max(("update_h", "update_e", "update_he"),
                key=lambda k: prof[k][0] / max(prof[k][1], 1))
    avg_ms = prof[k_dom][0] / prof[k_dom][1]
    achieved = ab[k_dom] / (avg_ms * 1e-3) / 1e9
    traffic = measured_traffic(k_dom) if args.n == 512 else None
    step_bytes = ab["update_h"] + ab["update_e"]       # the reference's accounting of a step
    kern = {k: round(v[0] / v[1], 5) for k, v in prof.items() if v[1]}
    print(f"[bench] per-kernel avg ms: {kern}", file=sys.stderr, flush=True)
    print(f"[bench] timed {args.steps} steps: {ms:.2f} ms (events), {wall * 1e3:.2f} ms (wall)",
          file=sys.stderr, flush=True)
    del sc, tabs
    blo = bloch_bench(args.bloch_steps) if args.bloch_steps > 0 else None
    cpu = None if args.no_cpu else cpu_baseline(args.n)

    # serving, then gradients, each shortened when it would overrun the budget
    srv = None
    if args.serve_requests > 0:
        specs = scenes.serving_requests(256, rate=20.0, seed=0)[:args.serve_requests]
        # keep room for the smallest gradient leg (4 designs and the central differences)
        left = (args.budget_s - (time.perf_counter() - t_start)
                - (4 * 2.2 + 2.0 * args.adjoint_check + 4 if args.adjoint_designs > 0 else 0))
        n_req = len(specs)
        while n_req > 16 and _serving_estimate_s(specs[:n_req], value) > left:
            n_req -= 16
        if n_req < len(specs):
            print(f"[bench] serving shortened to {n_req} requests by the {args.budget_s:.0f} s "
                  f"budget", file=sys.stderr, flush=True)
        srv = serving_bench(n_req, args.arena_gib, args.serve_workers, args.solo_lanes,
                            best_rate=value)
        if n_req < args.serve_requests:
            srv["shortened_from"] = args.serve_requests
    adj = None
    if args.adjoint_designs > 0:
        left = args.budget_s - (time.perf_counter() - t_start)
        designs = args.adjoint_designs
        # seconds per design with three in flight (2.1-2.2 measured) / per pixel's pair of runs
        per, fd = 2.2, 2.0 * args.adjoint_check
        while designs > 4 and designs * per + fd + 4 > left:
            designs -= 4
        if designs < args.adjoint_designs:
            print(f"[bench] gradients shortened to {designs} designs by the "
                  f"{args.budget_s:.0f} s budget", file=sys.stderr, flush=True)
        adj = adjoint_bench(designs, args.adjoint_steps, args.adjoint_check,
                            concurrency=args.adjoint_concurrency)
        if designs < args.adjoint_designs:
            adj["shortened_from"] = args.adjoint_designs
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "warmup_steps_run": warm,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"box-soup {args.n}^3 (CPML x6, 32 boxes eps/2-pole Lorentz, "
                               f"Ez dipole, 8-freq xy field monitor), one Yee step",
                   "cells": cells, "dft_batch": args.dft_batch, "launch": launch,
                   "launch_tuning_ms_per_step": plan["timings"] if plan else None},
        "roofline": {"bound": "hbm", "kernel": k_dom, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "traffic_frac": (round(traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                                      if traffic else None),
                     "fused": k_dom == "update_he",
                     "algorithmic_bytes_per_launch": ab[k_dom],
                     "minimal_bytes_per_launch": ab["minimal_" + k_dom],
                     "avg_launch_ms": round(avg_ms, 5)},
        "step_hbm_frac": round(step_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "cpu_baseline": cpu,
        "serving": srv,
        "adjoint": adj,
        "bloch": blo,
        "wall_s": round(time.perf_counter() - t_start, 1),
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
