"""Core configuration and data types for DEM simulations."""
from .contact_model import ContactModelConfig, HertzContactConfig, LinearContactConfig
from .types import DomainBounds, ParticleProperties, Vector3, WallProperties
from .demconfig import DEMSolverConfig

__all__ = ["ContactModelConfig", "HertzContactConfig", "LinearContactConfig", "DomainBounds",
           "ParticleProperties", "WallProperties", "DEMSolverConfig", "Vector3"]
