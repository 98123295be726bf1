"""MagMaterial and the predefined uniaxial anisotropy.

Mirrors interface/nmag/main.py:340-471 (coefficients in simulation units) and
interface/nmag/anisotropy.py:92-136 (uniaxial_anisotropy).
"""
import numpy as np

from .units import SI, mu0, to_su


class NmagUserError(Exception):
    """User error, as raised by the reference front end."""


class PredefinedAnisotropy:
    def __init__(self, anis_type, function, order, axis1=None, K1=None, K2=None):
        self.anis_type = anis_type
        self.function = function
        self.order = order
        self.axis1 = axis1
        self.K1 = K1
        self.K2 = K2


def uniaxial_anisotropy(axis, K1, K2=0):
    """E_anis = -K1 <axis,m>^2 - K2 <axis,m>^4 (interface/nmag/anisotropy.py:92-136)."""
    axis = np.asarray(axis, dtype=np.float64)
    axis = axis / np.linalg.norm(axis)

    def f(m):
        a = float(np.dot(axis, m))
        return -K1 * a ** 2 - K2 * a ** 4

    return PredefinedAnisotropy("uniaxial", f, 4 if K2 else 2, axis1=axis, K1=K1, K2=K2)


class MagMaterial:
    def __init__(self, name, Ms=SI(0.86e6, "A/m"), llg_damping=0.5,
                 llg_gamma_G=SI(2.210173e5, "m/A s"), llg_normalisationfactor=SI(0.1e12, "1/s"),
                 do_precession=True, exchange_coupling=SI(1.3e-11, "J/m"), anisotropy=None,
                 anisotropy_order=None, properties=("magnetic", "material")):
        self.name = name
        self.Ms = Ms
        self.llg_damping = llg_damping
        self.llg_gamma_G = llg_gamma_G
        self.llg_normalisationfactor = llg_normalisationfactor
        self.do_precession = do_precession
        self.exchange_coupling = exchange_coupling
        self.properties = list(properties)
        if isinstance(anisotropy, PredefinedAnisotropy):
            if anisotropy_order:
                raise NmagUserError("Cannot specify custom anisotropy_order when using a "
                                    "predefined anisotropy.")
            if anisotropy.anis_type != "uniaxial":
                raise NmagUserError("only uniaxial anisotropy is implemented on this path")
            self.anisotropy = anisotropy
            self.anisotropy_order = anisotropy.order
        elif anisotropy is not None:
            raise NmagUserError("custom anisotropy functions are not supported on this path; "
                                "use uniaxial_anisotropy(axis, K1, K2)")
        else:
            self.anisotropy = None
            self.anisotropy_order = None

        # simulation-unit coefficients (interface/nmag/main.py:423-471)
        self.su_Ms = float(to_su(Ms, "A/m"))
        self.su_llg_gamma_G = float(to_su(llg_gamma_G, "m/A s"))
        self.su_llg_damping = float(to_su(llg_damping, ""))
        self.su_llg_normalisationfactor = float(to_su(llg_normalisationfactor, "1/s"))
        self.su_exchange_coupling = float(to_su(exchange_coupling, "J/m"))
        gilbert_to_ll = 1.0 / (1.0 + self.su_llg_damping ** 2)
        self.su_llg_coeff1 = -self.su_llg_gamma_G * gilbert_to_ll
        self.su_llg_coeff2 = self.su_llg_coeff1 * self.su_llg_damping
        if not do_precession:
            self.su_llg_coeff1 = 0.0
        if self.su_exchange_coupling < 0.0:
            raise NmagUserError("The exchange coupling constant must be positive. For "
                                "material '%s', you specified: %s." % (name, exchange_coupling))
        self.su_mu0 = float(to_su(mu0))
        self.su_exch_prefactor = 2.0 * self.su_exchange_coupling / (self.su_mu0 * self.su_Ms)

    def su_coeffs(self):
        """The nsim_material struct of the C ABI."""
        return {"Ms": self.su_Ms, "negJ": -self.su_exch_prefactor, "c1": self.su_llg_coeff1,
                "c2": self.su_llg_coeff2, "c3": self.su_llg_normalisationfactor,
                "mu0": self.su_mu0}

    def su_anisotropy(self):
        """(K1_su, K2_su, axis) of a uniaxial anisotropy, or None."""
        if self.anisotropy is None:
            return None
        a = self.anisotropy
        return (float(to_su(a.K1, "J/m^3")), float(to_su(a.K2, "J/m^3")) if a.K2 else 0.0,
                np.asarray(a.axis1, dtype=np.float64))
