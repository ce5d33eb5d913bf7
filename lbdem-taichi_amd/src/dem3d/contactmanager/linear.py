"""Linear spring-dashpot contact model (src/dem3d/contactmanager/linear.py)."""
import math

from .contactmodel import ContactModel


class LinearContactModel(ContactModel):
    model_id = 1

    @staticmethod
    def min_time(m, K):
        """sqrt(m / k) critical time step (linear.py:155-160)."""
        return math.sqrt(m / K)
