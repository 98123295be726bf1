"""Differentiable LLG trajectories: torch.autograd.Function over the library's
discrete adjoint (nsim_traj_forward / nsim_traj_adjoint).

Everything is in simulation units.  PyTorch is only the autograd front: the
forward and backward sweeps run in libnsim_hip.so on the GPU, and the tensors
exchanged here are small (parameters, the loss and its cotangents).

Per-element anisotropy has no counterpart in the reference, whose anisotropy
is per material and evaluated at the nodes (interface/nmag/nmag_lam.py:485-501).
It is mapped to nodal K1 by the lumped projection
    K1_i = sum_{e containing i} (vol_e / 4) K_e / V_i,   V_i = sum_e vol_e / 4,
which reproduces the reference exactly when K is uniform.
"""
import numpy as np
import scipy.sparse as sps
import torch


def element_to_node(coords, tets):
    """Sparse P (n_nodes x n_tets) with K1_nodes = P @ K_elements."""
    x = np.asarray(coords, dtype=np.float64).reshape(-1, 3)
    t = np.asarray(tets, dtype=np.int64).reshape(-1, 4)
    e1, e2, e3 = (x[t[:, k]] - x[t[:, 0]] for k in (1, 2, 3))
    vol = np.abs(np.einsum("ij,ij->i", e1, np.cross(e2, e3))) / 6.0
    rows = t.reshape(-1)
    cols = np.repeat(np.arange(len(t)), 4)
    w = np.repeat(vol / 4.0, 4)
    V = np.bincount(rows, weights=w, minlength=len(x))
    return sps.csr_matrix((w / V[rows], (rows, cols)), shape=(len(x), len(t)))


def mean_mx_weights(volumes):
    """d<m_x>/dm for the volume-weighted average <m_x> (3N)."""
    V = np.asarray(volumes, dtype=np.float64)
    w = np.zeros((len(V), 3))
    w[:, 0] = V / V.sum()
    return w.reshape(-1)


class TrajectoryProblem:
    """The fixed pieces of a differentiable run: the magnet (a NativeMagnet
    with one magnet), its base material coefficients (simulation units), the
    initial m, the anisotropy axis, the static field, the time grid and the
    pulse layout, and the terminal loss weights (loss = w . m(T))."""

    def __init__(self, magnet, coeffs, coords, tets, m0, axis, dt, n_steps, n_pulses,
                 pulse_len, pulse_dir=(1.0, 0.0, 0.0), checkpoint_every=50, weights=None,
                 H_static=None):
        self.magnet = magnet
        self.coeffs = dict(coeffs)
        # exchange constant in simulation units, fixed while Ms varies
        self.A = -self.coeffs["negJ"] * self.coeffs["mu0"] * self.coeffs["Ms"] / 2.0
        self.P = element_to_node(coords, tets)
        self.m0 = np.asarray(m0, dtype=np.float64).reshape(-1, 3)
        self.axis = np.asarray(axis, dtype=np.float64).reshape(-1, 3)
        self.dt, self.n_steps = float(dt), int(n_steps)
        self.n_pulses, self.pulse_len = int(n_pulses), float(pulse_len)
        self.pulse_dir = tuple(float(v) for v in pulse_dir)
        self.checkpoint_every = int(checkpoint_every)
        self.weights = (mean_mx_weights(magnet.get_field("volumes")) if weights is None
                        else np.asarray(weights, dtype=np.float64).reshape(-1))
        self.H_static = (np.zeros_like(self.m0) if H_static is None
                         else np.asarray(H_static, dtype=np.float64).reshape(-1, 3))

    def material(self, Ms):
        c = dict(self.coeffs)
        c["Ms"] = float(Ms)
        c["negJ"] = -2.0 * self.A / (c["mu0"] * float(Ms))
        return c

    def kwargs(self):
        return dict(dt=self.dt, n_steps=self.n_steps, checkpoint_every=self.checkpoint_every,
                    pulse_len=self.pulse_len, pulse_dir=self.pulse_dir)

    def run(self, K_elem, Ms, amps):
        """Forward only: the loss w . m(T)."""
        g = self.magnet
        g.set_material(self.material(Ms))
        g.set_field("axis", self.axis)
        g.set_field("K1", self.P @ np.asarray(K_elem, dtype=np.float64))
        g.set_field("H_ext", self.H_static)
        g.set_field("m", self.m0)
        g.traj_forward(amps=np.asarray(amps, dtype=np.float64), **self.kwargs())
        return float(self.weights @ g.get_field("m").reshape(-1))

    def gradient(self, amps, scale=1.0):
        """Backward of the last run: cotangents of (K_elem, Ms, amps)."""
        r = self.magnet.traj_adjoint(scale * self.weights, amps=np.asarray(amps, np.float64),
                                     **self.kwargs())
        return self.P.T @ r["K1"], r["Ms"], r["amps"], r["m0"]


class LLGTrajectoryLoss(torch.autograd.Function):
    """loss = w . m(T) as a function of per-element K (n_tets), Ms (scalar)
    and the pulse amplitudes (n_pulses), all float64 tensors."""

    @staticmethod
    def forward(ctx, K_elem, Ms, amps, problem):
        loss = problem.run(K_elem.detach().cpu().numpy(), float(Ms.detach().cpu()),
                           amps.detach().cpu().numpy())
        ctx.problem = problem
        ctx.amps = amps.detach().cpu().numpy()
        ctx.dev = K_elem.device
        return torch.tensor(loss, dtype=torch.float64, device=K_elem.device)

    @staticmethod
    def backward(ctx, grad_out):
        gK, gMs, gA, _ = ctx.problem.gradient(ctx.amps, scale=float(grad_out))
        t = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.float64, device=ctx.dev)
        return t(gK), t(gMs), t(gA), None


def trajectory_loss(K_elem, Ms, amps, problem):
    return LLGTrajectoryLoss.apply(K_elem, Ms, amps, problem)
