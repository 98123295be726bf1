"""nsim_psolve_noalloc: the preconditioner solve of nsim_psolve with its
GMRES workspace in the library's static device scratch instead of memory the
library allocates.  Its output must equal nsim_psolve's on the same handle,
on two calls in a row, the second on a smaller magnet, and it must refuse a
magnet whose workspace exceeds the scratch."""
import numpy as np
import pytest

from helpers import material, random_m, rel_err
from nmag import kuhn_cuboid
from nmag._native import NativeMagnet, NsimError

pytestmark = pytest.mark.gpu


def _noalloc(g, r):
    return g.psolve_noalloc(r)


def _magnet(dims, seed, gamma=0.05):
    print("[case] no-alloc psolve on", dims, flush=True)
    coords, tets = kuhn_cuboid(*dims)
    g = NativeMagnet(coords, tets, material(alpha=0.2).su_coeffs(), nbc_rtol=1e-10,
                     dbc_rtol=1e-10)
    g.set_field("H_ext", np.tile([0.03, -0.01, 0.02], (g.n, 1)))
    g.psetup(random_m(g.n, seed).reshape(-1), gamma)
    return g


@pytest.mark.parametrize("tols", [(1e-2, 1e-7), (1e-11, 0.0)])
def test_noalloc_psolve_equals_psolve_on_two_calls_in_a_row(tols):
    big = _magnet((12, 10, 4, 2.5), 51)      # 715 nodes
    small = _magnet((4, 3, 2, 2.5), 52)      # 60 nodes
    rng = np.random.default_rng(53)
    for g in (big, small):
        g.set_tols(ksp_rtol=tols[0], ksp_atol=tols[1])
        r = rng.standard_normal(3 * g.n)
        z_ref = g.psolve(r)
        z = _noalloc(g, r)
        assert rel_err(z, z_ref) <= 1e-12
        assert np.max(np.abs(z)) > 0
    big.close()
    small.close()


def test_noalloc_psolve_refuses_a_magnet_beyond_the_scratch():
    # 31 x 31 x 23 nodes = 22,103 > the 20,400-node limit of the 16 MiB scratch
    coords, tets = kuhn_cuboid(30, 30, 22, 2.5)
    print("[case] no-alloc psolve size limit, %d nodes" % len(coords), flush=True)
    g = NativeMagnet(coords, tets, material().su_coeffs(), do_demag=False)
    g.psetup(random_m(g.n, 54).reshape(-1), 0.05)
    with pytest.raises(NsimError, match="static scratch"):
        _noalloc(g, np.ones(3 * g.n))
    g.close()
