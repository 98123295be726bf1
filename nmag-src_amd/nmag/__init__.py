"""nmag-compatible front end of the MI355X-native LLG right-hand-side path.

Keeps the surface of the reference's interface/nmag package for this path:
Simulation, MagMaterial, SI, uniaxial_anisotropy, load_mesh, set_m, set_H_ext,
advance_time, relax, get_subfield.  All compute runs in libnsim_hip.so.
"""
from .units import SI, mu0, to_su, from_su
from .material import MagMaterial, NmagUserError, uniaxial_anisotropy
from .anisotropy import PredefinedAnisotropy, cubic_anisotropy
from .mesh import kuhn_cuboid, cuboid_sizes, read_nmesh, write_nmesh
from .simulation import (Simulation, every, at, get_subfield_from_h5file,
                         get_subfield_positions_from_h5file)
from .when import never
from .datafile import read_ndt
from .hlib import HMatrixSetup, default_hmatrix_setup
from . import _native

__all__ = ["SI", "mu0", "to_su", "from_su", "MagMaterial", "NmagUserError",
           "uniaxial_anisotropy", "cubic_anisotropy", "PredefinedAnisotropy", "kuhn_cuboid", "cuboid_sizes", "read_nmesh", "write_nmesh",
           "Simulation", "every", "at", "never", "get_subfield_from_h5file",
           "get_subfield_positions_from_h5file", "read_ndt", "HMatrixSetup",
           "default_hmatrix_setup"]
