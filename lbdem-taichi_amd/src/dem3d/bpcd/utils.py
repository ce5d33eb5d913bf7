"""Helpers of the broad phase (src/dem3d/bpcd/utils.py)."""


def next_pow2(x):
    """Smallest power of two >= x (x >= 1); 0 maps to 0 as in the reference's bit trick."""
    x = int(x)
    return 0 if x <= 0 else 1 << (x - 1).bit_length()


def round32(n):
    return n if n % 32 == 0 else ((n >> 5) + 1) << 5


def cal_neighbor_search_radius(max_radius):
    return max_radius * 1.1
