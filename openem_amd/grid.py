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
        """Dual spacing ``(n,)`` (E update): the mean of the two primal cells that meet at each
        primal edge.  Before cell 0 a periodic axis sees its last cell, an absorbing end a copy
        of cell 0 (the grid continued one cell outwards)."""
        half = 0.5 * self.dl
        out = half + np.roll(half, 1)
        if not self.is_periodic:
            out[0] = self.dl[0]
        return out

    @property
    def centers(self) -> np.ndarray:
        return 0.5 * (self.edges[:-1] + self.edges[1:])

    def _end_image(self, side: str, bloch_k: float):
        """What a difference reads past one end of the axis: ``(index, weight, pec)``.

        The forward difference at the last cell reads E one edge past the high end, the backward
        difference at cell 0 reads H one centre before the low end.  ``weight`` multiplies that
        read (0 where the field outside is zero), ``index`` is the in-range cell it is read
        from, and ``pec`` (low end only) is what tangential E on edge 0 is multiplied by.
        """
        n = self.n
        if self.is_periodic:
            # Bloch condition u(x + L) = u(x) exp(+i 2 pi k): the wrap picks up the phase
            ph = np.exp((2j if side == "hi" else -2j) * np.pi * bloch_k) if bloch_k else 1.0
            return (0 if side == "hi" else n - 1), ph, 1.0
        if bloch_k != 0.0:
            raise ValueError(f"bloch_k={bloch_k} needs a periodic axis, this one is "
                             f"{self.boundary_lo}/{self.boundary_hi}")
        kind = self.boundary_lo if side == "lo" else self.boundary_hi
        if kind.startswith("Symmetry"):
            raise NotImplementedError(f"{kind}: symmetry folding is not supported")
        if kind == "PMCBoundary":
            # magnetic wall: tangential H is odd about it, tangential E even
            return (0, -1.0, 1.0) if side == "lo" else (n - 1, 1.0, 1.0)
        # PEC wall, or the PEC that backs a PML: nothing is read from outside, and tangential
        # E on the low wall is held at zero
        return (0 if side == "lo" else n - 1), 0.0, 0.0

    def index_tables(self, bloch_k: float = 0.0) -> dict[str, np.ndarray]:
        """Neighbour indices and weights encoding the boundary topology (the reference's
        contract, ``openem/grid.py:219-304``): ``nxt``/``prv`` the forward/backward neighbour of
        every cell (int32), ``mnx``/``mpv`` the weight of that neighbour (1 inside; at the ends
        the image of :meth:`_end_image`, complex128 on a Bloch axis), ``pec`` 1 except on a PEC
        low wall."""
        n = self.n
        cell = np.arange(n, dtype=np.int32)
        nxt, prv = np.roll(cell, -1), np.roll(cell, 1)
        dtype = np.complex128 if (self.is_periodic and bloch_k != 0.0) else np.float64
        mnx, mpv, pec = np.ones(n, dtype), np.ones(n, dtype), np.ones(n)
        nxt[n - 1], mnx[n - 1], _ = self._end_image("hi", bloch_k)
        prv[0], mpv[0], pec[0] = self._end_image("lo", bloch_k)
        return {"nxt": nxt, "prv": prv, "mnx": mnx, "mpv": mpv, "pec": pec}


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

    def _yee_lengths(self, name: str) -> list[np.ndarray]:
        """Per-axis lengths of the dual cell around component ``name``'s Yee points: the dual
        spacing along an axis where the point sits on a primal edge, the primal one where it
        sits at a cell centre."""
        return [a.dl_dual if on else a.dl for a, on in zip(self.axes, YEE_ON_EDGE[name])]

    def yee_dual_volumes(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """Volume ``(nx, ny, nz)`` of the dual cell of Ex, Ey, Ez at their Yee points
        (reference ``openem/grid.py:350-368``: Ex -> dx dy_dual dz_dual, ...)."""
        return tuple(outer3(*self._yee_lengths(c)) for c in ("Ex", "Ey", "Ez"))

    def yee_h_dual_volumes(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """The same for Hx, Hy, Hz (reference ``openem/grid.py:370-386``)."""
        return tuple(outer3(*self._yee_lengths(c)) for c in ("Hx", "Hy", "Hz"))

    def dual_volume_at(self, comp: int, magnetic: bool, i: int, j: int, k: int) -> float:
        """One entry of :meth:`yee_dual_volumes` (``magnetic``: :meth:`yee_h_dual_volumes`),
        with the same multiplication order."""
        lx, ly, lz = self._yee_lengths(("H" if magnetic else "E") + "xyz"[comp])
        return float((lx[i] * ly[j]) * lz[k])

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
    """Index of the primal edge nearest to a plane coordinate (sources and flux planes sit on
    edges, where tangential E lives; the reference snaps the same way, ``openem/grid.py:397``).
    A coordinate inside the grid always has an edge within half its cell; one outside must lie
    within half of the end cell, else it is refused (ties go to the lower edge)."""
    e = np.asarray(edges, dtype=np.float64)
    k = int(np.clip(np.searchsorted(e, coord), 1, e.size - 1))
    i = k - 1 if coord - e[k - 1] <= e[k] - coord else k
    if coord < e[0]:
        reach = 0.5 * (e[1] - e[0])
    elif coord > e[-1]:
        reach = 0.5 * (e[-1] - e[-2])
    else:
        reach = np.inf
    if abs(coord - e[i]) > reach:
        raise ValueError(f"coordinate {coord!r} is {abs(coord - e[i])!r} off the grid's end "
                         f"edge {e[i]!r}")
    return i
