"""Differentiation form: gradient of a CIC density loss with respect to the
initial Zel'dovich displacement field, through the solve-and-move steps of
the reference's time loop (controller.cpp:106-208) with the octree,
computational particles, neighbour lists and particle-to-leaf map held fixed
after step 0.

The forward pass runs the library's step (solve + kick + drift) with the
topology frozen and saves the particle state on the device every
`ckpt_every` steps; the backward pass walks the steps in reverse, restoring
the nearest checkpoint and replaying forward to regenerate each step's node
fields, and applies the library's adjoint kernels (drift, then kick with the
interpolation).  `CICLoss` wraps it as a torch.autograd.Function.

With the topology frozen the node densities come from fixed per-node
particle counts (solver.cpp:116-172: rhs1 = c_dens n_p / h^3), so the node
fields do not depend on the particle positions and the displacement gradient
needs only the drift and kick adjoints.  The fields do depend on the particle
mass c_dens; `backward(mass=True)` also returns dL/dc_dens through the whole
field chain of every step: the kick adjoint accumulates dL/dg and dL/dH per
node, the transpose of the gradient fit gives dL/dphi, a transposed GMRES on
the Poisson matrix gives dL/drhs2, the transpose of the mean removal
(solver.cpp:244-252) and c_pot give dL/drho, and a transposed GMRES on the
density matrix gives dL/drhs1 = dL/dc_dens / (n_p / h^3).

Einstein-de Sitter in normalised comoving units (mass 1/N per particle,
C_pot = 1.5/a, C_field = -1, Cx = sqrt(a), Cv = Cx/a^2) as the bench uses.
"""
import numpy as np

from . import Context, make_params, ksp


def eds(a):
    """C_pot, C_field, Cx(a), Cv(a) for Omega = 1 in normalised units."""
    cx = np.sqrt(a)
    return 1.5 / a, -1.0, cx, cx / a / a


class FrozenSim:
    """x0 = q + psi (wrapped), v0 = vfac * psi; initial half drift; `steps`
    leapfrog steps from a_start to a_end; loss = sum over a periodic grid^3
    mesh of (CIC density contrast)^2 of the final positions."""

    def __init__(self, q, psi0, vfac, a_start=0.02, a_end=0.1, steps=10, grid=32, ckpt_every=2,
                 lm=27, leaf_capacity=8, rtol=1e-12, restart=60, max_it=4000, pc="jacobi"):
        self.q = np.asarray(q, np.float64).reshape(3, -1)
        self.N = self.q.shape[1]
        self.vfac = float(vfac)
        self.a_start, self.a_end, self.steps = a_start, a_end, steps
        self.da = (a_end - a_start) / steps
        self.grid, self.ckpt_every = grid, ckpt_every
        self.o1 = ksp(rtol=rtol, restart=restart, max_it=max_it, pc="jacobi")
        self.o2 = ksp(rtol=rtol, restart=restart, max_it=max_it, pc=pc)
        self.oadj = ksp(rtol=rtol, restart=restart, max_it=max_it, pc="jacobi")
        self.mass = 1.0 / self.N  # c_dens of every step
        self.ctx = Context(make_params(bbox=(0, 1, 0, 1, 0, 1), depth=20, bc=1, c=0.0, lm=lm,
                                       leaf_capacity=leaf_capacity))
        self.forwards = 0
        # topology from the base field, after the initial half drift
        x0, v0 = self._initial(psi0)
        self.ctx.set_particles(*x0, *v0)
        _, _, cx0, _ = eds(a_start)
        self.ctx.drift(cx0, 0.5 * self.da)
        self.ctx.octree_build()
        self.ctx.select()
        self.ctx.freeze_topology(True)
        comp = self.ctx.comp()
        M = comp["nc"] + comp["nv"]
        h3 = comp["h"][:M] ** 3
        self._b1 = comp["np"][:M] / h3  # rhs1 per unit c_dens
        self._h3 = h3                   # weights of the periodic mean of rho

    def loss_weights(self, w=None):
        """Per-particle weights of the CIC loss (lattice order); None: equal."""
        self.ctx.adjoint_cic_weights(w)

    def _initial(self, psi):
        psi = np.asarray(psi, np.float64).reshape(3, self.N)
        return np.mod(self.q + psi, 1.0), self.vfac * psi

    def _schedule(self, k):
        """(t_k, dt_k) of step k after the half drift (controller.cpp:119-208)."""
        t = self.a_start + 0.5 * self.da + k * self.da
        return t, (min(self.da, self.a_end - t) if t + self.da > self.a_end else self.da)

    def _coeffs(self, k):
        t, dt = self._schedule(k)
        cpot, cf, _, cv = eds(t)
        _, _, cxh, _ = eds(t + dt / 2)
        return self.mass, cpot, cf, cv, cxh, dt

    def _step(self, k):
        cd, cpot, cf, cv, cxh, dt = self._coeffs(k)
        self.ctx.step(cd, cpot, cf, cv, cxh, dt, self.o1, self.o2)

    def forward(self, psi):
        """Loss for displacement field psi (3 x N or N x 3 in lattice order)."""
        psi = np.asarray(psi, np.float64)
        if psi.shape == (self.N, 3):
            psi = psi.T
        x0, v0 = self._initial(psi)
        self.ctx.set_state(*x0, *v0)
        _, _, cx0, _ = eds(self.a_start)
        self.ctx.drift(cx0, 0.5 * self.da)
        for k in range(self.steps):
            if k % self.ckpt_every == 0:
                self.ctx.checkpoint(k // self.ckpt_every)
            self._step(k)
        self.forwards += 1
        return self.ctx.adjoint_seed_cic(self.grid)

    def backward(self, mass=False):
        """dL/dpsi (3 x N) for the last forward pass (its adjoint seed); with
        mass=True, (dL/dpsi, dL/dc_dens)."""
        lam = self.ctx.adjoint_get()
        dmass = 0.0
        for k in reversed(range(self.steps)):
            seg = k // self.ckpt_every
            self.ctx.checkpoint(seg, restore=True)
            for j in range(seg * self.ckpt_every, k):  # replay to the state of step k
                self._step(j)
            cd, cpot, cf, cv, cxh, dt = self._coeffs(k)
            self.ctx.set_constants(cd, cpot, cf)
            self.ctx.solve(self.o1, self.o2)  # node fields of step k
            self.ctx.adjoint_set(*lam)
            self.ctx.adjoint_drift(cxh, dt)
            self.ctx.adjoint_kick(cv, dt, accumulate=mass, clear=mass)
            lam = self.ctx.adjoint_get()
            if mass:
                dmass += self._mass_adjoint(cpot)
        _, _, cx0, _ = eds(self.a_start)
        self.ctx.adjoint_set(*lam)
        self.ctx.adjoint_drift(cx0, 0.5 * self.da)
        lx, ly, lz, lvx, lvy, lvz = self.ctx.adjoint_get()
        g = np.stack([lx + self.vfac * lvx, ly + self.vfac * lvy, lz + self.vfac * lvz])
        return (g, dmass) if mass else g

    def _mass_adjoint(self, cpot):
        """dL/dc_dens of one step from the node-field adjoint its kick adjoint
        accumulated (the system and fields of that step are current)."""
        dg, dH = self.ctx.adjoint_fields()
        dphi = self.ctx.gradient_adjoint(dg, dH)
        drhs2, its, reason = self.ctx.solve_adjoint(2, self.oadj, dphi)
        if reason <= 0:
            raise RuntimeError("transposed Poisson solve did not converge (%d, %d)" % (its, reason))
        drho = cpot * (drhs2 - self._h3 * (drhs2.sum() / self._h3.sum()))
        drhs1, its, reason = self.ctx.solve_adjoint(1, self.oadj, drho)
        if reason <= 0:
            raise RuntimeError("transposed density solve did not converge (%d, %d)" % (its, reason))
        return float(np.dot(drhs1, self._b1))


try:
    import torch

    class CICLoss(torch.autograd.Function):
        """loss = CICLoss.apply(psi, sim): psi an (N, 3) float64 tensor of
        displacements in lattice order, sim a FrozenSim."""

        @staticmethod
        def forward(fctx, psi, sim):
            loss = sim.forward(psi.detach().cpu().numpy())
            fctx.sim = sim
            fctx.fwd = sim.forwards
            fctx.save_for_backward(psi)
            return psi.new_tensor(loss)

        @staticmethod
        def backward(fctx, gout):
            sim = fctx.sim
            (psi,) = fctx.saved_tensors
            if sim.forwards != fctx.fwd:  # the adjoint seed belongs to a later pass
                sim.forward(psi.detach().cpu().numpy())
            g = sim.backward().T
            return torch.as_tensor(np.ascontiguousarray(g), dtype=psi.dtype) * gout, None
except ImportError:  # torch is optional for the library itself
    CICLoss = None
