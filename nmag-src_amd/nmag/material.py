"""MagMaterial: a material's coefficients in simulation units.

Mirrors interface/nmag/material.py:20-230 and interface/nmag/main.py:423-471.
"""
import numpy as np

from .anisotropy import (NmagUserError, PredefinedAnisotropy, fit_polynomial, su_energy,
                         uniaxial_anisotropy)
from .units import SI, mu0, to_su

__all__ = ["MagMaterial", "NmagUserError", "uniaxial_anisotropy"]


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
        # interface/nmag/material.py:175-196: a PredefinedAnisotropy carries
        # its order; a custom energy function a(m) needs anisotropy_order
        if isinstance(anisotropy, PredefinedAnisotropy):
            if anisotropy_order:
                raise NmagUserError("Cannot specify custom anisotropy_order when using a "
                                    "predefined anisotropy.")
            if not anisotropy.has_function():
                raise NmagUserError("the anisotropy has no energy function")
            self.anisotropy = anisotropy
            self.anisotropy_order = anisotropy.order
        elif anisotropy is not None or anisotropy_order is not None:
            if anisotropy is None or not callable(anisotropy):
                raise NmagUserError("anisotropy must be a PredefinedAnisotropy or a function "
                                    "of the magnetisation direction")
            if not anisotropy_order:
                raise NmagUserError("You need to specify the anisotropy_order when using a "
                                    "custom anisotropy function.")
            self.anisotropy = PredefinedAnisotropy(anisotropy, anisotropy_order)
            self.anisotropy_order = anisotropy_order
        else:
            self.anisotropy = None
            self.anisotropy_order = None
        self._poly = None

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
        """(K1_su, K2_su, axis) of a uniaxial anisotropy (the device's closed
        form, per-node K1 and K2), or None for none or another kind."""
        a = self.anisotropy
        if a is None or a.anis_type != "uniaxial":
            return None
        return (float(to_su(a.K1, "J/m^3")), float(to_su(a.K2, "J/m^3")) if a.K2 else 0.0,
                np.asarray(a.axis1, dtype=np.float64))

    def su_anisotropy_poly(self):
        """(powers, coeffs) of any other anisotropy: the homogeneous polynomial
        of its order equal to its energy density (simulation units) on the unit
        sphere, or None."""
        a = self.anisotropy
        if a is None or a.anis_type == "uniaxial":
            return None
        if self._poly is None:
            self._poly = fit_polynomial(su_energy(a.function), a.order)
        return self._poly
