"""Broad-phase collision detection (spatial hashing with Morton codes) and prefix sums."""
from .bpcd import BPCD
from .prefixsum import PrefixSumExecutor
from .utils import next_pow2, round32, cal_neighbor_search_radius

__all__ = ["BPCD", "PrefixSumExecutor", "next_pow2", "round32", "cal_neighbor_search_radius"]
