"""torch.autograd binding of the device solver's adjoint.

``AmunEvolve`` maps an initial conserved state U0 (a float64 tensor
U[nv][nz][ny][nx]) and the adiabatic index gamma to the state after
``nsteps`` steps.  Its backward is the hand-written adjoint of the device
step (amun_dev_tape_backward: reverse sweep over the stages with local
Jacobians, steps recomputed from checkpoints every ``interval`` steps), with
dt and the GLM cmax of every step held at their forward values.  PyTorch only
carries the tensors in and out; the solve and its adjoint run in the HIP
library, which must be built (no CPU fallback).

``c5_objective`` builds the Orszag-Tang differentiation case: the initial
density field and gamma are the inputs, velocity, field and pressure are the
fixed Orszag-Tang ones, and J is the volume-weighted kinetic energy of the
final state.
"""
from __future__ import annotations

import copy
import math

import numpy as np
import torch

from .config import Params
from .problems import cell_centres
from .solver import Solver


class AmunEvolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, gamma, params: Params, nsteps: int, interval: int):
        p = copy.deepcopy(params)
        p.gamma = float(gamma)
        s = Solver(p)
        s.upload(u0.detach().cpu().numpy())
        s.tape_forward(nsteps, interval)
        un = s.download()
        ctx.solver = s
        return torch.from_numpy(un).to(device=u0.device, dtype=u0.dtype)

    @staticmethod
    def backward(ctx, gu):
        lam0, dg = ctx.solver.tape_backward(gu.detach().cpu().numpy())
        ctx.solver.close()
        g0 = torch.from_numpy(lam0).to(device=gu.device, dtype=gu.dtype)
        return g0, torch.tensor(dg, dtype=gu.dtype, device=gu.device), None, None, None


def c5_params(n: int = 256, precision: str = "fp64") -> Params:
    """2D ideal MHD Orszag-Tang vortex on [0,1]^2, n^2 as (n/16)^2 blocks of
    16^2 (+2 ghosts), TVD-MC + HLLD + SSPRK(3,3).  The CFL number is not
    given for this case; 0.4 is used."""
    return Params(ndims=2, equation_system="mhd", reconstruction="tvd", limiter="mc",
                  riemann_solver="hlld", time_advance="rk3", precision=precision,
                  gamma=5.0 / 3.0, cfl=0.4, ncells=16, nghosts=2, xblocks=n // 16,
                  yblocks=n // 16, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0)


def ot_fields(p: Params):
    """The fixed Orszag-Tang velocity, field and pressure (problems.orszag_tang_2d)."""
    nz, ny, nx = p.shape
    z, y, x = cell_centres(p)
    Y, X = np.meshgrid(y, x, indexing="ij")
    v = np.zeros((3, nz, ny, nx))
    b = np.zeros((3, nz, ny, nx))
    v[0] = -np.sin(2.0 * np.pi * Y)
    v[1] = np.sin(2.0 * np.pi * X)
    b0 = 1.0 / math.sqrt(4.0 * math.pi)
    b[0] = -b0 * np.sin(2.0 * np.pi * Y)
    b[1] = b0 * np.sin(4.0 * np.pi * X)
    pr = np.full((nz, ny, nx), 5.0 / (12.0 * math.pi))
    return v, b, pr


def mhd_prim_to_cons(rho, v, b, pr, gamma):
    """prim2cons_mhd_adi (equations.F90:2815-2868) on tensors; differentiable
    in rho and gamma."""
    m = rho.unsqueeze(0) * v
    ek = 0.5 * (m * v).sum(0)
    em = 0.5 * (b * b).sum(0)
    e = pr / (gamma - 1.0) + ek + em
    psi = torch.zeros_like(rho)
    return torch.cat([rho.unsqueeze(0), m, b, psi.unsqueeze(0), e.unsqueeze(0)], 0)


def kinetic_energy(u, dvol: float):
    """Volume-weighted kinetic energy sum 0.5 |m|^2 / rho dV."""
    return 0.5 * ((u[1:4] ** 2).sum(0) / u[0]).sum() * dvol


def c5_objective(rho0, gamma, p: Params, nsteps: int = 300, interval: int = 20):
    """J(rho0, gamma) of the differentiation case; rho0 is (nz, ny, nx)."""
    v, b, pr = (torch.from_numpy(a).to(rho0) for a in ot_fields(p))
    u0 = mhd_prim_to_cons(rho0, v, b, pr, gamma)
    un = AmunEvolve.apply(u0, gamma, p, nsteps, interval)
    dx = p.dx()
    return kinetic_energy(un, dx[0] * dx[1])
