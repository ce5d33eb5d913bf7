"""Helpers of the unresolved couplings (src/unlbdem/utils.py)."""
import math


def threedelta(r: float) -> float:
    """Roma-Peskin 3-point kernel applied to a radial distance (eqlattice.py:451-462)."""
    if r < 0.5:
        return (1.0 + math.sqrt(1.0 - 3.0 * r * r)) / 3.0
    if r <= 1.5:
        return (5.0 - 3.0 * r - math.sqrt(1.0 - 3.0 * (1.0 - r) ** 2)) / 6.0
    return 0.0
