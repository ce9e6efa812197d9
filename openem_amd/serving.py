"""Serving: continuous batching of many scenes on one GPU.

The reference runs one scene per ``openem.solver.run`` call (``openem/solver.py:403-511``).
Here a :class:`Pool` keeps many scenes resident in one device arena and advances all of them
with the same launches (batched kernels over every resident scene in a hipGraph of ``dft_batch`` steps; see
``csrc/oem_pool.hip``).  :func:`serve` is the request loop:

* a client thread releases requests at their arrival times (building each Scene just ahead);
* setup workers build the step tables on the host and, when the arena can hold the scene
  (admission against the HBM budget), create it in the arena and upload it;
* the stepping thread attaches ready scenes at the next block boundary, keeps a bounded number
  of blocks in flight and hands finished scenes to the readout, which reads the phasors,
  computes the fluxes and releases the scene's memory;
* scenes of at least ``solo_cells`` cells nearly fill the GPU by themselves: they run as
  standalone contexts (hipGraph blocks; the fused H+E step from 10 M cells) in ``solo_lanes``
  lanes beside the batch, each lane one scene at a time on its own stream, so one large scene's
  setup, readout and release overlap another's stepping, and two large scenes together fill the
  GPU better than one.  A free lane takes the waiting large scene with the most cell-steps.

:func:`serve_serial` is the plain loop the batched form is measured against: one
:func:`openem_amd.solver.run` per request, in arrival order.
"""

from __future__ import annotations

import ctypes as C
import queue
import sys
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib, flux, scenes, solver, tables as tables_mod


class Pool:
    """A device arena of ``arena_bytes`` plus the batched stepping graph."""

    def __init__(self, arena_bytes: int, dft_batch: int = 16, device: int = 0):
        self.lib = _lib.load()
        self.dft_batch = int(dft_batch)
        self.handle = C.c_void_p()
        _lib.check(self.lib, self.lib.oem_pool_create(int(device), int(arena_bytes),
                                                      self.dft_batch, C.byref(self.handle)))

    def attach(self, st: solver.Stepper) -> None:
        _lib.check(self.lib, self.lib.oem_pool_attach(self.handle, st.ctx))

    def launch(self, n_blocks: int = 1) -> None:
        _lib.check(self.lib, self.lib.oem_pool_launch(self.handle, int(n_blocks)))

    def wait(self, max_pending: int = 0) -> None:
        _lib.check(self.lib, self.lib.oem_pool_wait(self.handle, int(max_pending)))

    def done(self, st: solver.Stepper, wait: bool = False) -> bool:
        d = C.c_int32(0)
        _lib.check(self.lib, self.lib.oem_pool_ctx_done(self.handle, st.ctx, int(wait),
                                                        C.byref(d)))
        return bool(d.value)

    def stats(self) -> dict:
        used, cap = C.c_uint64(0), C.c_uint64(0)
        run, blocks = C.c_int32(0), C.c_int64(0)
        _lib.check(self.lib, self.lib.oem_pool_stats(self.handle, C.byref(used), C.byref(cap),
                                                     C.byref(run), C.byref(blocks)))
        return dict(used=used.value, capacity=cap.value, running=run.value, blocks=blocks.value)

    def close(self) -> None:
        if self.handle.value:
            _lib.check(self.lib, self.lib.oem_pool_destroy(self.handle))
            self.handle = C.c_void_p()

    def run_to_completion(self, steppers: list) -> None:
        """Attach ``steppers`` and launch blocks until all of them are done."""
        for st in steppers:
            self.attach(st)
        while self.stats()["running"]:
            self.launch(1)
        self.wait(0)


def device_bytes(tabs: tables_mod.StepTables, dft_batch: int = 16) -> int:
    """Arena bytes a scene needs (every buffer the library allocates, each rounded up to the
    arena's 256-byte granule).  Used for admission; the arena itself is the hard check."""
    nx, ny, nz = tabs.shape
    n = nx * ny * nz
    r = lambda b: (int(b) + 255) // 256 * 256  # noqa: E731
    total = 2 * r(12 * n)                                          # E, H
    w = [a.pml_lo + a.pml_hi for a in tabs.axes]
    psi = 2 * (w[0] * ny * nz + nx * w[1] * nz + nx * ny * w[2])
    total += 2 * r(4 * max(psi, 1)) + r(8)
    total += sum(r(4 * 16 * m) + 2 * r(4 * m) for m in (nx, ny, nz))  # axis tables
    mc = tabs.material_classes()
    if mc is not None:
        total += r(2 * n) + r(32 * mc[1].shape[0])
    else:
        total += 2 * r(12 * n)
    if tabs.disp is not None:
        d = tabs.disp
        hb = d.hist_box
        total += 3 * r(4 * d.comp.size) + r(4 * (d.comp.size + 1)) + r(4 * d.am1.size)
        total += 4 * r(4 * d.lut_am1.size) + 2 * r(4 * d.am1.size)
        total += r(12 * (hb[1] - hb[0]) * (hb[3] - hb[2]) * (hb[5] - hb[4]))
    total += 2 * r(8 * (tabs.n_total + 1)) * len(tabs.plane_srcs)
    for dp in tabs.dipoles:
        total += 6 * r(4 * dp.comp.size) + r(8 * dp.amp.size * (2 if np.iscomplexobj(dp.amp) else 1))
    for m in tabs.monitors:
        npts = m.n1 * m.n2 if m.kind == "plane" else int(np.prod(m.box))
        nc = 4 if m.kind == "plane" else 6
        total += r(4 * dft_batch * nc * npts) + 2 * r(8 * nc * m.nf * npts)
        total += r(8 * tabs.n_total * m.nf * 4)
    return int(total)


@dataclass
class Request:
    spec: dict
    t_arrival: float = 0.0          # when the request entered the system (seconds, run clock)
    t_admit: float = 0.0            # attached to the batch
    t_done: float = 0.0             # results read back
    run_s: float = 0.0              # wall seconds of its own stepping call (standalone runs)
    flux: list = field(default_factory=list)   # per monitor, (nf,) float64
    error: str | None = None


def _flux_of(st: solver.Stepper) -> list:
    out = []
    for i, m in enumerate(st.tabs.monitors):
        data = st.read_monitor(i)
        mon = m.spec if m.kind == "plane" and m.modes is None else None
        out.append(flux.monitor_flux(data, st.tabs.scene.grid, mon) if mon is not None else data)
    return out


def _report(results: list, total: int, t0: float, label: str, every: int) -> None:
    """One line on stderr each time another ``every`` requests have finished."""
    if every and (len(results) % every == 0 or len(results) == total):
        print(f"[{label}] {len(results)}/{total} requests done at "
              f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)


def _client(specs, t0, outq, build_ahead: int, stop):
    """Release each request at its arrival time; scenes are built up to ``build_ahead``
    requests early so a request's clock starts at its arrival, not after its construction."""
    pending = queue.Queue(maxsize=build_ahead)

    def builder():
        for sp in specs:
            if stop.is_set():
                break
            pending.put((sp, scenes.serving_scene(sp)))
        pending.put(None)

    th = threading.Thread(target=builder, daemon=True)
    th.start()
    while True:
        item = pending.get()
        if item is None:
            break
        sp, sc = item
        delay = t0 + sp["arrival"] - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        outq.put((Request(spec=sp, t_arrival=time.perf_counter() - t0), sc))
    outq.put(None)
    th.join()


#: cells stepped together before a new request waits for room (FIFO): enough to fill the GPU
#: with small scenes, few enough that large ones are not time-shared many ways
BATCH_CELLS = 48_000_000


#: scenes at least this large run standalone in the solo lane
SOLO_CELLS = 8_000_000


def serve(specs: list, arena_bytes: int = 64 << 30, dft_batch: int = 16, workers: int = 6,
          max_in_flight: int = 2, build_ahead: int = 8, device: int = 0,
          batch_cells: int = BATCH_CELLS, solo_cells: int = SOLO_CELLS, solo_lanes: int = 2,
          solo_running: int = 2, on_admit=None, report_every: int = 0) -> dict:
    """Continuous batching of ``specs`` (from :func:`openem_amd.scenes.serving_requests`).
    Ready scenes join the running batch in arrival order while the batch's cells stay within
    ``batch_cells`` (a scene that alone exceeds it runs when the batch is empty).  Large scenes
    go to ``solo_lanes`` lanes, of which at most ``solo_running`` step at once (the others set
    up, read out or release meanwhile).  With
    ``report_every`` a progress line goes to stderr each time that many more requests finish.
    Returns ``dict(requests=[Request...], wall=seconds, peak_running=int)``."""
    pool = Pool(arena_bytes, dft_batch, device)
    t0 = time.perf_counter()
    arrivals: queue.Queue = queue.Queue()
    ready: queue.Queue = queue.Queue()
    # the solo lanes take the waiting large scene with the most cell-steps first: those scenes
    # set the end of the stream, and the longest-first order shortens its tail
    solo: queue.PriorityQueue = queue.PriorityQueue()
    seq = iter(range(1 << 62))
    stepping = threading.Semaphore(max(1, int(solo_running)))
    done_q: queue.Queue = queue.Queue()
    stop = threading.Event()
    mem = threading.Condition()
    budget = {"free": int(arena_bytes * 0.97)}
    results: list = []

    def setup_worker():
        while True:
            item = arrivals.get()
            if item is None:
                arrivals.put(None)
                ready.put(None)
                return
            req, sc = item
            try:
                tabs = tables_mod.build(sc)
                if int(np.prod(tabs.shape)) >= solo_cells:
                    work = int(np.prod(tabs.shape)) * int(tabs.n_total)
                    solo.put((-work, next(seq), (req, tabs)))
                    continue
                need = device_bytes(tabs, dft_batch)
                with mem:
                    while budget["free"] < need and not stop.is_set():
                        mem.wait(0.05)
                    budget["free"] -= need
                st = None
                while st is None:
                    try:
                        st = solver.Stepper(tabs, device=device, dft_batch=dft_batch, pool=pool)
                    except _lib.OpenEMError as e:       # arena fragmentation: wait for a release
                        if e.code != _lib.OEM_ENOMEM or stop.is_set():
                            raise
                        with mem:
                            mem.wait(0.05)
                ready.put((req, st, need))
            except Exception as e:  # noqa: BLE001 - reported per request
                req.error = repr(e)
                req.t_done = time.perf_counter() - t0
                done_q.put((req, None, 0))

    def readout_worker():
        while True:
            item = done_q.get()
            if item is None:
                return
            req, st, need = item
            if st is not None:
                try:
                    req.flux = _flux_of(st)
                except Exception as e:  # noqa: BLE001
                    req.error = repr(e)
                st.close()
                with mem:
                    budget["free"] += need
                    mem.notify_all()
                req.t_done = time.perf_counter() - t0
            with mem:
                results.append(req)
                _report(results, len(specs), t0, "serve", report_every)

    def solo_worker():
        while True:
            item = solo.get()[2]
            if item is None:
                return
            req, tabs = item
            req.t_admit = time.perf_counter() - t0
            try:
                st = solver.Stepper(tabs, device=device, dft_batch=dft_batch)
                try:
                    with stepping:
                        tr = time.perf_counter()
                        st.run(tabs.n_total)
                        req.run_s = time.perf_counter() - tr
                    req.flux = _flux_of(st)
                finally:
                    st.close()
            except Exception as e:  # noqa: BLE001
                req.error = repr(e)
            req.t_done = time.perf_counter() - t0
            with mem:
                results.append(req)
                _report(results, len(specs), t0, "serve", report_every)

    client = threading.Thread(target=_client, args=(specs, t0, arrivals, build_ahead, stop),
                              daemon=True)
    setups = [threading.Thread(target=setup_worker, daemon=True) for _ in range(workers)]
    solo_ths = [threading.Thread(target=solo_worker, daemon=True)
                for _ in range(max(1, int(solo_lanes)))]
    for th in solo_ths:
        th.start()
    reader = threading.Thread(target=readout_worker, daemon=True)
    client.start()
    for th in setups:
        th.start()
    reader.start()

    running: list = []
    waiting: list = []              # ready, not yet admitted (arrival order)
    finished_setups = 0
    peak = 0
    try:
        while True:
            # collect what is ready (wait for something when idle and more may come)
            while finished_setups < workers:
                idle = not running and not waiting
                try:
                    item = ready.get(block=idle, timeout=0.5 if idle else None)
                except queue.Empty:
                    if idle:
                        continue
                    break
                if item is None:
                    finished_setups += 1
                    continue
                waiting.append(item)
            waiting.sort(key=lambda it: it[0].spec["id"])
            cells = sum(int(np.prod(st.tabs.shape)) for _, st, _ in running)
            while waiting:
                req, st, need = waiting[0]
                c = int(np.prod(st.tabs.shape))
                if running and cells + c > batch_cells:
                    break
                waiting.pop(0)
                pool.attach(st)
                if on_admit is not None:
                    on_admit(req, st)
                req.t_admit = time.perf_counter() - t0
                running.append((req, st, need))
                cells += c
            if not running and not waiting and finished_setups == workers:
                break
            peak = max(peak, len(running))
            if running:
                pool.launch(1)
                pool.wait(max_in_flight - 1)
            still = []
            for req, st, need in running:
                if pool.done(st):
                    done_q.put((req, st, need))
                else:
                    still.append((req, st, need))
            running = still
    finally:
        stop.set()
        for _ in solo_ths:
            solo.put((float("inf"), next(seq), None))
        for th in solo_ths:
            th.join()
        done_q.put(None)
        reader.join()
        client.join(timeout=5)
        pool.wait(0)
        pool.close()
    wall = max((r.t_done for r in results), default=0.0)
    return dict(requests=sorted(results, key=lambda r: r.spec["id"]), wall=wall,
                peak_running=peak)


def serve_serial(specs: list, dft_batch: int = 16, device: int = 0, build_ahead: int = 8,
                 report_every: int = 0) -> dict:
    """The plain loop: requests in arrival order, one :func:`openem_amd.solver.run`-equivalent
    each (tables, context, stepping, readout, release)."""
    t0 = time.perf_counter()
    arrivals: queue.Queue = queue.Queue()
    stop = threading.Event()
    client = threading.Thread(target=_client, args=(specs, t0, arrivals, build_ahead, stop),
                              daemon=True)
    client.start()
    results = []
    while True:
        item = arrivals.get()
        if item is None:
            break
        req, sc = item
        try:
            tabs = tables_mod.build(sc)
            st = solver.Stepper(tabs, device=device, dft_batch=dft_batch)
            try:
                tr = time.perf_counter()
                st.run(tabs.n_total)
                req.run_s = time.perf_counter() - tr
                req.flux = _flux_of(st)
            finally:
                st.close()
        except Exception as e:  # noqa: BLE001
            req.error = repr(e)
        req.t_done = time.perf_counter() - t0
        results.append(req)
        _report(results, len(specs), t0, "serial", report_every)
    client.join(timeout=5)
    wall = max((r.t_done for r in results), default=0.0)
    return dict(requests=results, wall=wall)


def summary(res: dict) -> dict:
    """sims/s over the run (first arrival to last result) and latency percentiles."""
    reqs = res["requests"]
    ok = [r for r in reqs if r.error is None]
    lat = np.array([r.t_done - r.t_arrival for r in ok]) if ok else np.zeros(0)
    t_first = min((r.t_arrival for r in reqs), default=0.0)
    span = max(res["wall"] - t_first, 1e-9)
    return dict(n=len(reqs), failed=len(reqs) - len(ok), sims_per_s=len(ok) / span,
                p50_s=float(np.percentile(lat, 50)) if lat.size else None,
                p99_s=float(np.percentile(lat, 99)) if lat.size else None,
                span_s=span)
