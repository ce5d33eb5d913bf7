"""DEM solver configuration (src/dem3d/demconfig/demconfig.py)."""
from typing import Optional

from .contact_model import ContactModelConfig, HertzContactConfig, LinearContactConfig
from .types import DomainBounds, ParticleProperties, Vector3, WallProperties


class DEMSolverConfig:
    """Configuration for the DEM solver."""
    BOUNDARY_TYPE_WALL = 0
    BOUNDARY_TYPE_PERIODIC = 1

    def __init__(self, domain: DomainBounds, dt: float, gravity, contact_model: ContactModelConfig,
                 init_particles: Optional[object] = None):
        if dt <= 0:
            raise ValueError(f"Time step dt must be positive, got {dt}")
        contact_model.validate()
        self.domain = domain
        self.dt = dt
        self.gravity = Vector3(*[float(v) for v in gravity])
        self.init_particles = init_particles
        self.contact_model = contact_model
        self.particle_props = ParticleProperties()
        self.wall_props = WallProperties()
        self.domain_min, self.domain_max = domain.get_extended_bounds()
        self._setup_walls()

    def _setup_walls(self):
        d = self.domain
        # six axis-aligned planes n.x = d with outward normals (+x, -x, +y, -y, +z, -z)
        self.wall_normals = [Vector3(s * (a == 0), s * (a == 1), s * (a == 2))
                             for a in range(3) for s in (1.0, -1.0)]
        self.wall_distances = [d.xmax, -d.xmin, d.ymax, -d.ymin, d.zmax, -d.zmin]
        self.boundaryType = [self.BOUNDARY_TYPE_WALL] * 6

    def set_particle_properties(self, **kwargs) -> "DEMSolverConfig":
        for key, value in kwargs.items():
            if not hasattr(self.particle_props, key):
                raise ValueError(f"Unknown particle property: {key}")
            setattr(self.particle_props, key, value)
        return self

    def set_wall_properties(self, **kwargs) -> "DEMSolverConfig":
        for key, value in kwargs.items():
            if not hasattr(self.wall_props, key):
                raise ValueError(f"Unknown wall property: {key}")
            setattr(self.wall_props, key, value)
        return self

    def set_periodic_boundaries(self, x_periodic=False, y_periodic=False, z_periodic=False):
        for axis, on in enumerate((x_periodic, y_periodic, z_periodic)):
            if on:
                self.boundaryType[2 * axis] = self.BOUNDARY_TYPE_PERIODIC
                self.boundaryType[2 * axis + 1] = self.BOUNDARY_TYPE_PERIODIC
        return self

    def update_contact_model(self, **kwargs) -> "DEMSolverConfig":
        for key, value in kwargs.items():
            if not hasattr(self.contact_model, key):
                raise ValueError(f"Unknown parameter '{key}' for {self.contact_model.get_model_name()} contact model")
            setattr(self.contact_model, key, value)
        self.contact_model.validate()
        return self

    def periodic_axes(self):
        return [self.boundaryType[2 * a] == self.BOUNDARY_TYPE_PERIODIC for a in range(3)]

    def summary(self) -> str:
        m = self.contact_model
        d = self.domain
        g = self.gravity
        lines = [
            "", "DEM Solver Configuration:", "========================",
            f"Domain: [{d.xmin}, {d.xmax}] x [{d.ymin}, {d.ymax}] x [{d.zmin}, {d.zmax}]",
            f"Time step: {self.dt} s",
            f"Gravity: ({g.x}, {g.y}, {g.z}) m/s^2", "",
            "Particle Properties:",
            f"- Elastic modulus: {self.particle_props.elastic_modulus} Pa",
            f"- Poisson ratio: {self.particle_props.poisson_ratio}",
            f"- Density: {self.particle_props.density} kg/m^3", "",
            "Wall Properties:",
            f"- Density: {self.wall_props.density} kg/m^3",
            f"- Elastic modulus: {self.wall_props.elastic_modulus} Pa",
            f"- Poisson ratio: {self.wall_props.poisson_ratio}", "",
            f"Contact Model: {m.get_model_name().upper()}",
        ]
        if isinstance(m, LinearContactConfig):
            lines += [f"- Normal stiffness: {m.stiffness_normal} N/m",
                      f"- Tangential stiffness: {m.stiffness_tangential} N/m",
                      f"- Normal damping: {m.damping_normal}",
                      f"- Tangential damping: {m.damping_tangential}"]
        elif isinstance(m, HertzContactConfig):
            lines += [f"- P-P restitution: {m.pp_restitution}", f"- P-W restitution: {m.pw_restitution}"]
        lines += [f"- P-P friction: {m.pp_friction}", f"- P-W friction: {m.pw_friction}",
                  f"- Max coordination number: {self.particle_props.max_coordinate_number}", ""]
        return "\n".join(lines)
