"""Contact model interface (src/dem3d/contactmanager/contactmodel.py).

The force laws run inside the DEM kernels; these classes select the law and expose the
critical time step estimates on the host.
"""


class ContactModel:
    model_id = -1

    def __init__(self, material_field):
        self.mf = material_field

    def particle_particle_force(self, *args, **kwargs):
        raise NotImplementedError("evaluated on the device by lbdem_dem_run")

    def particle_wall_force(self, *args, **kwargs):
        raise NotImplementedError("evaluated on the device by lbdem_dem_run")
