"""Cell-type flags and unit conversion (API of src/lbm3d/lbmutils.py in the reference).

CellType keeps the reference's bit values (lbmutils.py:69-97) because drivers build their
masks from them; UnitConverter keeps its method names (lbmutils.py:14-63).
"""


class UnitConverter:
    """Physical <-> lattice units for spacing dx [m], step dt [s] and density rho [kg/m^3]."""

    def __init__(self, dx: float, dt: float, rho: float):
        self.dx = dx
        self.dt = dt
        self.rho = rho

    # scale factors physical -> lattice
    def _v(self):
        return self.dt / self.dx

    def _a(self):
        return self.dt * self.dt / self.dx

    def _F(self):
        return self.dt * self.dt / (self.rho * self.dx ** 4)

    def _sig(self):
        return self.dt * self.dt / (self.rho * self.dx * self.dx)

    def _nu(self):
        return self.dt / (self.dx * self.dx)

    # physical -> lattice
    def getLbLength(self, physLength): return physLength / self.dx
    def getLbTime(self, physTime): return physTime / self.dt
    def getLbVel(self, physVel): return physVel * self._v()
    def getLbAccel(self, physAccel): return physAccel * self._a()
    def getLbStrainRate(self, physStrainRate): return physStrainRate * self.dt
    def getLbRho(self, physRho): return physRho / self.rho
    def getLbNu(self, physNu): return physNu * self._nu()
    def getLbSigma(self, physSigma): return physSigma * self._sig()
    def getLbForce(self, physForce): return physForce * self._F()

    # lattice -> physical
    def getPhysLength(self, lbLength): return lbLength * self.dx
    def getPhysTime(self, lbTime): return lbTime * self.dt
    def getPhysVel(self, lbVel): return lbVel / self._v()
    def getPhysAccell(self, lbAccel): return lbAccel / self._a()
    def getPhysStrainRate(self, lbStrainRate): return lbStrainRate / self.dt
    def getPhysRho(self, lbRho): return lbRho * self.rho
    def getPhysNu(self, lbNu): return lbNu / self._nu()
    def getPhysSigma(self, lbSigma): return lbSigma / self._sig()
    def getPhysForce(self, lbForce): return lbForce / self._F()


class CellType:
    """Cell-type bit flags (values fixed by the reference, lbmutils.py:69-97)."""
    FLUID = 1 << 0
    INTERFACE = 1 << 1
    GAS = 1 << 2
    OBSTACLE = 1 << 3
    FREE_SLIP = 1 << 4
    VEL_LADD = 1 << 5
    VEL_ZOUHE = 1 << 6
    VEL_EXIT = 1 << 7
    Pre_ZOUHE = 1 << 8
    NO_FLUID_NEIGH = 1 << 9
    TO_FLUID = 1 << 10
    TO_GAS = 1 << 11
    RIGHT = 1 << 12
    LEFT = 1 << 13
    TOP = 1 << 14
    BOTTOM = 1 << 15
    FRONT = 1 << 16
    BACK = 1 << 17
    VEL_INLET_LADD = 1 << 18
