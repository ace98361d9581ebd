"""One-dimensional node grid (the reference's jsolver/grid_1d.py:4-114).

Same constructor, attributes (mx, nx, centered, rim, xb, xe, grid_type,
length, del_, idel, is_linear, centers, widths, edges) and value semantics:
hash and equality on (mx, centered, grid_type, rim, xb, xe), so a grid can key
plan caches the way the reference's grids key jax.jit caches.  Host-side
numpy; no device work.
"""
import numpy as np


class grid_1D:
    """Linear (grid_type 0) or sinusoidal (grid_type 1) grid; the solvers
    accept linear grids only."""

    __slots__ = ("mx", "nx", "centered", "rim", "xb", "xe", "grid_type", "length",
                 "del_", "idel", "is_linear", "centers", "widths", "edges")

    def __init__(self, mx, xb=0.0, xe=1.0, rim=1, grid_type=0, centered=True):
        self.centered = centered
        self.mx = mx
        self.rim = rim
        self.xb = xb
        self.xe = xe
        self.grid_type = grid_type
        self.compute_dependent_vars()
        self.build_grid()

    def _key(self):
        return (self.mx, self.centered, self.grid_type, self.rim, self.xb, self.xe)

    def __hash__(self):
        return hash(self._key())

    def __eq__(self, other):
        return isinstance(other, grid_1D) and self._key() == other._key()

    def __repr__(self):
        return (f"grid_1D(mx={self.mx}, xb={self.xb}, xe={self.xe}, rim={self.rim}, "
                f"grid_type={self.grid_type}, centered={self.centered})")

    def compute_dependent_vars(self):
        """nx = mx + 1, cell width del_ = length / mx (centered) or / nx."""
        self.nx = self.mx + 1
        self.length = self.xe - self.xb
        self.del_ = self.length / (self.mx if self.centered else self.nx)
        self.idel = 1 / self.del_

    def build_grid(self):
        """edges at integer positions -rim .. mx+rim+1 (shifted by half a cell
        when centered), centers and widths between consecutive edges."""
        pos = np.arange(-self.rim, self.mx + self.rim + 2)
        if self.grid_type == 0:
            edges = (pos - 0.5 if self.centered else pos) * self.del_ + self.xb
        elif self.grid_type == 1:
            edges = (pos + 2 * np.sin(2 * np.pi * pos / self.nx)) * self.del_ + self.xb
        else:
            raise ValueError("No such grid grid_type: {}".format(self.grid_type))
        self.edges = edges
        lo, hi = edges[:-1], edges[1:]
        self.centers = 0.5 * (lo + hi)
        self.widths = hi - lo
        self.is_linear = self.grid_type == 0
