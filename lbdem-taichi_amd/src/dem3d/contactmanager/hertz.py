"""Hertz-Mindlin contact model (src/dem3d/contactmanager/hertz.py)."""
import math

from .contactmodel import ContactModel


class HertzMindlinContactModel(ContactModel):
    model_id = 0

    @staticmethod
    def min_time(r, rho, Gm, v):
        """Rayleigh critical time step (hertz.py:156-162)."""
        return math.pi * r * math.sqrt(rho / Gm) / (0.1631 * v + 0.8766)
