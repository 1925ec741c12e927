"""flash-ph for MI355X: exact Vietoris-Rips persistent homology (H0, H1, H2
over Z/2) on one AMD Instinct GPU through hand-written gfx950 HIP kernels.

Drop-in for the reference package's public API (reference __init__.py:3-9).
"""

__version__ = "0.1.0"

from flash_ph.persistence import rips_persistence
from flash_ph.threshold import auto_threshold, enclosing_radius
from flash_ph.complex import RipsComplex
from flash_ph.batch import auto_threshold_batch, rips_persistence_batch

__all__ = ["rips_persistence", "auto_threshold", "enclosing_radius", "RipsComplex",
           "rips_persistence_batch", "auto_threshold_batch"]
