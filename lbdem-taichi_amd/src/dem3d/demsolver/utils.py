"""Host-side helpers of the DEM driver (src/dem3d/demsolver/utils.py)."""
import numpy as np


def quat2RotMatrix(quat):
    """Rotation matrix of a unit quaternion (w, x, y, z) (demsolver/utils.py:34-61)."""
    w, i, j, k = (float(v) for v in quat)
    return np.array([[w * w + i * i - j * j - k * k, 2 * (i * j - k * w), 2 * (j * w + i * k)],
                     [2 * (i * j + k * w), w * w - i * i + j * j - k * k, 2 * (j * k - i * w)],
                     [2 * (i * k - j * w), 2 * (j * k + i * w), w * w - i * i - j * j + k * k]])
