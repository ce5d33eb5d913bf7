"""Contact model registry: linear spring-dashpot and Hertz-Mindlin."""
from .contactmodel import ContactModel
from .hertz import HertzMindlinContactModel
from .linear import LinearContactModel

__all__ = ["ContactModel", "HertzMindlinContactModel", "LinearContactModel"]
