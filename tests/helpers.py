"""Shared builders: the same seeded inputs go to the oracle and to the GPU path."""
import math

import numpy as np

from oracle import spinor_oracle as O
import becgpu as G


def grids(n, L):
    return O.make_grid(n, L), G.make_grid(G.GridConfig(tuple(n), tuple(L)))


def random_psi(shape, D, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    psi = rng.normal(size=(D,) + shape) + 1j * rng.normal(size=(D,) + shape)
    return psi * scale


def pair(n, L, F, psi, dt, itp, *, omega=None, p=0.0, q=0.0, c0=0.0, c1=0.0, c_lhy=0.0,
         ddi=False, cdd=0.0, secular=False, n_steps=10, save_every=None, precision="fp64",
         c_extra=(), loss=None, padding=False, raman=None):
    """Oracle workspace and GPU workspace built from identical inputs."""
    og, gg = grids(n, L)
    ows = O.make_workspace(og, F, psi, dt, itp, omega=omega, p=p, q=q, c0=c0, c1=c1, c_lhy=c_lhy,
                           ddi=ddi, cdd=cdd, secular=secular, n_steps=n_steps,
                           save_every=save_every, c_extra=c_extra, loss=loss, padding=padding,
                           raman=raman)
    sp = G.SimParams(dt=dt, n_steps=n_steps, imaginary_time=itp, save_every=save_every)
    pot = G.HarmonicTrap(*omega) if omega is not None else G.NoPotential()
    atom = G.AtomSpecies("test", 1.0, F)
    gws = G.make_workspace(grid=gg, atom=atom,
                           interactions=G.InteractionParams(c0, c1, c_lhy, tuple(c_extra)),
                           zeeman=G.ZeemanParams(p, q), potential=pot, sim_params=sp,
                           psi_init=psi, enable_ddi=ddi, c_dd=cdd if ddi else float("nan"),
                           secular_ddi=secular, ddi_padding=padding, precision=precision,
                           loss=None if loss is None else G.LossParams(*loss),
                           raman=None if raman is None else G.RamanCoupling(*raman))
    return ows, gws


def rel_l2(a, b):
    return O.relative_l2(a, b)
