"""DEM package (API of src/dem3d in the reference)."""
from . import bpcd, contactmanager, demconfig, dateclass  # noqa: F401

__all__ = ["bpcd", "contactmanager", "demsolver", "dateclass"]
