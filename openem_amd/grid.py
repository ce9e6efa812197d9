"""Non-uniform Yee grid: spacings, boundary topology tables, Yee dual volumes.

Follows the reference's grid contract (``openem/grid.py:139-304``): edges are the primal cell
boundaries in metres, including any PML layers (which sit *outside* the physical domain, the
tidy3d convention). H updates divide by the primal spacing ``dl[i] = edges[i+1]-edges[i]``, E
updates by the dual spacing ``(dl[i]+dl[i-1])/2``.

Yee layout (component -> on-edge flags per axis), the same table the reference uses
(``openem/grid.py:50-57``)::

    Ex(ctr, edge, edge)  Ey(edge, ctr, edge)  Ez(edge, edge, ctr)
    Hx(edge, ctr, ctr)   Hy(ctr, edge, ctr)   Hz(ctr, ctr, edge)

Memory layout everywhere is C order ``(nx, ny, nz)``, z contiguous, flat index
``(i*ny + j)*nz + k``.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

YEE_ON_EDGE: dict[str, tuple[bool, bool, bool]] = {
    "Ex": (False, True, True),
    "Ey": (True, False, True),
    "Ez": (True, True, False),
    "Hx": (True, False, False),
    "Hy": (False, True, False),
    "Hz": (False, False, True),
}

PERIODIC = ("Periodic", "BlochBoundary")
ABSORBING = ("PML", "StablePML", "Absorber")


def transverse_axes(axis: int) -> tuple[int, int]:
    """The two axes other than ``axis``, ascending (monitor (t1, t2) convention,
    reference ``openem/grid.py:65-68``)."""
    return tuple(i for i in range(3) if i != axis)  # type: ignore[return-value]


def cyclic_axes(axis: int) -> tuple[int, int]:
    """``((axis+1)%3, (axis+2)%3)``: the right-handed (u, v) pair of mode sources / TFSF."""
    return (axis + 1) % 3, (axis + 2) % 3


def outer3(a, b, c):
    """``a[i]*b[j]*c[k]`` with the multiplication order fixed at (a*b)*c."""
    return a[:, None, None] * b[None, :, None] * c[None, None, :]


@dataclass(frozen=True)
class Axis:
    """Grid of one axis.

    Attributes:
        edges: ``(n+1,)`` primal boundaries [m], PML layers included.
        boundary_lo / boundary_hi: "PML", "Periodic", "BlochBoundary", "PECBoundary", ...
    """

    edges: np.ndarray
    boundary_lo: str
    boundary_hi: str

    def __post_init__(self) -> None:
        e = np.asarray(self.edges, dtype=np.float64)
        if e.ndim != 1 or e.size < 2:
            raise ValueError(f"edges must be 1D with >= 2 entries, got shape {e.shape}")
        if not np.all(np.diff(e) > 0):
            raise ValueError("edges must be strictly increasing")
        object.__setattr__(self, "edges", e)

    @property
    def n(self) -> int:
        return self.edges.size - 1

    @property
    def dl(self) -> np.ndarray:
        """Primal spacing ``(n,)`` (H update)."""
        return np.diff(self.edges)

    @property
    def is_periodic(self) -> bool:
        return self.boundary_lo in PERIODIC

    @property
    def dl_dual(self) -> np.ndarray:
        """Dual spacing ``(n,)`` (E update): ``(dl[i] + dl[i-1])/2``; at i=0 a periodic axis
        wraps to ``dl[n-1]``, an absorbing end extends with ``dl[0]``
        (reference ``openem/grid.py:198-215``)."""
        dl = self.dl
        prev = np.empty_like(dl)
        prev[1:] = dl[:-1]
        prev[0] = dl[-1] if self.is_periodic else dl[0]
        return 0.5 * (dl + prev)

    @property
    def centers(self) -> np.ndarray:
        return 0.5 * (self.edges[:-1] + self.edges[1:])

    def index_tables(self, bloch_k: float = 0.0) -> dict[str, np.ndarray]:
        """Neighbour indices and masks that encode the boundary topology
        (reference ``openem/grid.py:219-304``).

        - ``nxt``/``prv``: forward/backward neighbour (wrap on a periodic axis, clamp otherwise)
        - ``mnx``/``mpv``: 1 where that neighbour exists, 0 outside an absorbing end; on a Bloch
          axis the wrap entries carry ``exp(+-i 2 pi k)`` and the arrays are complex128
        - ``pec``: 0 at i=0 of an absorbing axis (the PML is backed by PEC)
        """
        n = self.n
        i = np.arange(n)
        if self.is_periodic:
            mnx = np.ones(n)
            mpv = np.ones(n)
            if bloch_k != 0.0:
                mnx = mnx.astype(np.complex128)
                mpv = mpv.astype(np.complex128)
                mnx[n - 1] = np.exp(+2j * np.pi * bloch_k)
                mpv[0] = np.exp(-2j * np.pi * bloch_k)
            return {"nxt": ((i + 1) % n).astype(np.int32),
                    "prv": ((i - 1) % n).astype(np.int32),
                    "mnx": mnx, "mpv": mpv, "pec": np.ones(n)}
        if bloch_k != 0.0:
            raise ValueError(f"bloch_k={bloch_k} needs a periodic axis, this one is "
                             f"{self.boundary_lo}/{self.boundary_hi}")
        pec = np.ones(n)
        mpv = (i > 0).astype(np.float64)
        nxt = np.minimum(i + 1, n - 1).astype(np.int32)
        mnx = (i + 1 < n).astype(np.float64)
        if self.boundary_lo == "PMCBoundary":
            mpv[0] = -1.0
        else:
            pec[0] = 0.0
        if self.boundary_hi == "PMCBoundary":
            mnx[n - 1] = 1.0
        return {"nxt": nxt, "prv": np.maximum(i - 1, 0).astype(np.int32),
                "mnx": mnx, "mpv": mpv, "pec": pec}


@dataclass(frozen=True)
class Grid:
    x: Axis
    y: Axis
    z: Axis

    @property
    def axes(self) -> tuple[Axis, Axis, Axis]:
        return (self.x, self.y, self.z)

    @property
    def shape(self) -> tuple[int, int, int]:
        return (self.x.n, self.y.n, self.z.n)

    @property
    def num_cells(self) -> int:
        nx, ny, nz = self.shape
        return nx * ny * nz

    def yee_dual_volumes(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """Dual volume of Ex/Ey/Ez at their own Yee positions
        (reference ``openem/grid.py:350-368``)."""
        xp, yp, zp = (a.dl for a in self.axes)
        xd, yd, zd = (a.dl_dual for a in self.axes)
        return outer3(xp, yd, zd), outer3(xd, yp, zd), outer3(xd, yd, zp)

    def dual_volume_at(self, comp: int, magnetic: bool, i: int, j: int, k: int) -> float:
        """One entry of :meth:`yee_dual_volumes` (``magnetic``: :meth:`yee_h_dual_volumes`),
        with the same multiplication order."""
        primal = (comp != 0, comp != 1, comp != 2) if magnetic else \
            (comp == 0, comp == 1, comp == 2)
        f = [a.dl if p else a.dl_dual for a, p in zip(self.axes, primal)]
        return float((f[0][i] * f[1][j]) * f[2][k])

    def yee_h_dual_volumes(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """Dual volume of Hx/Hy/Hz (reference ``openem/grid.py:370-386``)."""
        xp, yp, zp = (a.dl for a in self.axes)
        xd, yd, zd = (a.dl_dual for a in self.axes)
        return outer3(xd, yp, zp), outer3(xp, yd, zp), outer3(xp, yp, zd)

    def yee_coords(self, comp: int, magnetic: bool) -> tuple[np.ndarray, ...]:
        """Per-axis coordinates of a component's Yee points."""
        name = ("H" if magnetic else "E") + "xyz"[comp]
        return tuple(a.edges[:-1] if on else a.centers
                     for a, on in zip(self.axes, YEE_ON_EDGE[name]))


def uniform_axis(n: int, dl: float, lo: float, boundary_lo: str,
                 boundary_hi: str | None = None) -> Axis:
    """Uniform axis of ``n`` cells starting at coordinate ``lo``."""
    edges = lo + dl * np.arange(n + 1, dtype=np.float64)
    return Axis(edges, boundary_lo, boundary_hi or boundary_lo)


def nearest_edge_index(edges: np.ndarray, coord: float) -> int:
    """Snap a plane coordinate to the nearest primal boundary index (within half a local cell),
    as the reference does for source and flux planes (``openem/grid.py:397-425``)."""
    i = int(np.argmin(np.abs(edges - coord)))
    resid = abs(float(edges[i]) - coord)
    dl = np.diff(edges)
    local = max(float(dl[i]) if i < dl.size else 0.0, float(dl[i - 1]) if i > 0 else 0.0)
    if resid > 0.5 * local:
        raise ValueError(f"coordinate {coord!r} is {resid!r} off the nearest edge {edges[i]!r}")
    return i
