"""GUDHI-style Vietoris-Rips complex object over rips_persistence.

Same public surface as the reference's RipsComplex (complex.py:10-140):
constructor, compute_persistence, persistence,
persistence_intervals_in_dimension, betti_numbers, persistent_betti_numbers,
num_vertices, with the same defaults, exceptions and messages.  The
diagrams are computed once on the device and brought to the host once; every
accessor then works on those FP32 host copies.  Comparisons against the
caller's bounds are made in FP32, as the reference's comparisons of FP32
tensors with Python scalars are.
"""
from __future__ import annotations

import numpy as np
import torch

from flash_ph.persistence import rips_persistence

_NOT_COMPUTED = "Call compute_persistence() or persistence() first"


class RipsComplex:
    """Vietoris-Rips complex of a point cloud, filtered by edge length.

    Parameters
    ----------
    points : Tensor or ndarray
        (n, d) point coordinates.
    max_edge_length : float
        Edges of length >= max_edge_length are not in the complex.

    Examples
    --------
    >>> rc = RipsComplex(points, max_edge_length=2.0)
    >>> rc.compute_persistence(max_dim=1)
    >>> rc.persistence_intervals_in_dimension(1)
    >>> rc.betti_numbers()
    """

    def __init__(self, points, max_edge_length: float):
        if max_edge_length < 0:
            raise ValueError(
                f"max_edge_length must be non-negative, got {max_edge_length}")
        if isinstance(points, np.ndarray):
            points = torch.from_numpy(points).float()
            if torch.cuda.is_available():
                points = points.cuda()
        self._points = points
        self._max_edge_length = max_edge_length
        self._max_dim = None
        self._host = None     # per dimension: (K, 2) float32 host array

    # -- computation ------------------------------------------------------

    def compute_persistence(self, max_dim: int = 1, homology_coeff_field: int = 2,
                            backend: str = 'auto'):
        """Compute the diagrams up to ``max_dim`` and keep them (returns None).
        Only Z/2 coefficients exist (``homology_coeff_field=2``)."""
        if homology_coeff_field != 2:
            raise ValueError("Only Z/2Z coefficients are supported (homology_coeff_field=2)")
        diagrams = rips_persistence(self._points, self._max_edge_length, max_dim=max_dim,
                                    backend=backend)
        self._host = [np.ascontiguousarray(d.detach().cpu().numpy(), dtype=np.float32)
                      for d in diagrams]
        self._max_dim = max_dim

    def persistence(self, max_dim: int = 1, homology_coeff_field: int = 2):
        """Compute, then return every interval as ``(dim, (birth, death))``,
        sorted by dimension, birth, death."""
        self.compute_persistence(max_dim, homology_coeff_field)
        out = [(dim, (float(b), float(d)))
               for dim, rows in enumerate(self._host) for b, d in rows.tolist()]
        out.sort(key=lambda r: (r[0], r[1][0], r[1][1]))
        return out

    # -- accessors --------------------------------------------------------

    def _diagrams(self):
        if self._host is None:
            raise RuntimeError(_NOT_COMPUTED)
        return self._host

    def persistence_intervals_in_dimension(self, dim: int):
        """(K, 2) float64 array of the ``(birth, death)`` pairs of one
        dimension; empty outside 0..max_dim."""
        host = self._diagrams()
        if not 0 <= dim <= self._max_dim or host[dim].shape[0] == 0:
            return np.empty((0, 2), dtype=np.float64)
        return host[dim].astype(np.float64)

    def betti_numbers(self):
        """Per dimension, the number of intervals that never die."""
        return [int(np.count_nonzero(rows[:, 1] == np.float32(np.inf)))
                for rows in self._diagrams()]

    def persistent_betti_numbers(self, from_value: float, to_value: float):
        """Per dimension, the intervals born at or before ``from_value`` and
        still alive after ``to_value``."""
        lo, hi = np.float32(from_value), np.float32(to_value)
        return [int(np.count_nonzero((rows[:, 0] <= lo) & (rows[:, 1] > hi)))
                for rows in self._diagrams()]

    def num_vertices(self) -> int:
        """Number of input points."""
        return self._points.shape[0]
