"""Record layouts of the DEM entities (src/dem3d/dateclass).

The reference declares them as Taichi structs; here each entity is stored structure-of-arrays
in device tensors and these classes only name the members and their widths.
"""


class _Layout:
    members = ()

    @classmethod
    def names(cls):
        return [m for m, _ in cls.members]


class Grain(_Layout):
    """Spherical particle with 6-DOF state (grain.py:23-52)."""
    members = (("ID", 1), ("groupID", 1), ("materialType", 1), ("mass", 1), ("radius", 1), ("position", 3),
               ("velocity", 3), ("acceleration", 3), ("force", 3), ("quaternion", 4), ("omega", 3),
               ("omega_dot", 3), ("inertia", 9), ("moment", 3), ("freeze", 1), ("fixvel", 1),
               ("force_fluid", 3), ("moment_fluid", 3))


class Contact(_Layout):
    """Contact record (contact.py:20-33); stored: partner, activity, point, force, shear."""
    members = (("i", 1), ("j", 1), ("isActive", 1), ("materialType_i", 1), ("materialType_j", 1),
               ("position", 3), ("force_a", 3), ("moment_a", 3), ("moment_b", 3), ("shear_displacement", 3))


class Wall(_Layout):
    """Planar wall n.x = d (wall.py:18-24)."""
    members = (("normal", 3), ("distance", 1), ("materialType", 1), ("boundaryType", 1))


class Material(_Layout):
    """Material properties (material.py:13-30)."""
    members = (("density", 1), ("elasticModulus", 1), ("poissonRatio", 1), ("coefficientFriction", 1),
               ("coefficientRestitution", 1), ("stiffness_normal", 1), ("stiffness_tangent", 1),
               ("dp_nratio", 1), ("dp_sratio", 1))


__all__ = ["Contact", "Material", "Wall", "Grain"]
