"""Gradients on the GPU: the hand-written adjoint (amun_dev_tape_backward)
against the tangent-linear run (amun_dev_tape_tangent) by the dot-product
identity, the tangent against central finite differences of the nonlinear
solver with the step sizes held (amun_dev_run_fixed), and the Orszag-Tang
differentiation case through torch.autograd against finite differences
(eps = 1e-6, FP64) and forward-mode derivatives along 32 seeded directions."""
import numpy as np
import pytest
import torch

from amun_amd.autograd import c5_objective, c5_params, kinetic_energy, ot_fields
from amun_amd.config import Params, c1_params
from amun_amd.problems import blast_hd_2d, orszag_tang_2d, prim_to_cons
from amun_amd.solver import Solver

pytestmark = pytest.mark.gpu


def _ot(n=32, ta="rk3", rs="hlld", lim="mc"):
    p = c5_params(n)
    p.time_advance, p.riemann_solver, p.limiter = ta, rs, lim
    return p, orszag_tang_2d(p)


CASES = [("ot", "rk3", "hlld", "mc"), ("ot", "rk2", "hll", "mm"), ("ot", "euler", "hlld", "vl"),
         ("hd", "rk2", "hllc", "mc"), ("mhd3d", "rk3", "hlld", "mc")]


def _case(name, ta, rs, lim):
    if name == "ot":
        return _ot(32, ta, rs, lim)
    if name == "hd":
        p = c1_params()
        p.xblocks = p.yblocks = 2
        p.time_advance, p.riemann_solver, p.limiter = ta, rs, lim
        return p, blast_hd_2d(p)
    p = Params(ndims=3, equation_system="mhd", reconstruction="tvd", limiter=lim,
               riemann_solver=rs, time_advance=ta, ncells=8, nghosts=2, xblocks=2,
               yblocks=2, zblocks=2, cfl=0.3)
    from test_gpu_parity import _mhd_state
    return p, _mhd_state(p)


@pytest.mark.parametrize("case", CASES)
def test_dot_product_identity(case):
    """<lam, T(du0, dg)> = <A(lam), du0> + A_gamma(lam) dg to round-off, over
    7 steps recomputed from checkpoints every 3."""
    p, u0 = _case(*case)
    print(f"[case] {case}", flush=True)
    rng = np.random.default_rng(5)
    lam = rng.normal(size=u0.shape)
    du0 = rng.normal(size=u0.shape) * np.abs(u0).mean(axis=(1, 2, 3), keepdims=True) * 1e-3
    dg = 0.01
    with Solver(p) as s:
        s.upload(u0)
        s.tape_forward(7, 3)
        du = s.tape_tangent(du0, dg)
        lam0, gg = s.tape_backward(lam)
        # the live state is left at the end of the taped run
        assert s.step == 7
    lhs = float((lam * du).sum())
    rhs = float((lam0 * du0).sum()) + gg * dg
    print(f"  <lam, T du> = {lhs:.15e}  <A lam, du> = {rhs:.15e}")
    assert abs(lhs - rhs) <= 1e-11 * (abs(lhs) + abs(rhs))


def test_tangent_vs_finite_differences():
    """The tangent of the final state against central differences of the
    nonlinear map with dt and cmax held at the taped values."""
    p, u0 = _ot(32)
    rng = np.random.default_rng(6)
    drho = rng.normal(size=u0.shape[1:]) * 1e-2
    v, b, pr = ot_fields(p)
    q0 = np.zeros_like(u0)
    q0[0] = u0[0]
    q0[1:4] = v
    q0[4:7] = b
    q0[8] = pr
    def state(rho, gamma):
        q = q0.copy()
        q[0] = rho
        pp = Params(**{k: x for k, x in p.__dict__.items() if k != "extra"})
        pp.gamma = gamma
        return prim_to_cons(pp, q)
    dg = 0.01
    eps = 1e-6
    with Solver(p) as s:
        s.upload(u0)
        s.tape_forward(10, 4)
        dt, cm = s.tape_steps()
        du0 = (state(u0[0] + eps * drho, p.gamma) - state(u0[0] - eps * drho, p.gamma)) / (2 * eps)
        du0 += (state(u0[0], p.gamma + eps * dg) - state(u0[0], p.gamma - eps * dg)) / (2 * eps)
        du = s.tape_tangent(du0, dg)
    outs = []
    for sgn in (1, -1):
        pp = Params(**{k: x for k, x in p.__dict__.items() if k != "extra"})
        pp.gamma = p.gamma + sgn * eps * dg
        with Solver(pp) as s:
            s.upload(state(u0[0] + sgn * eps * drho, pp.gamma))
            s.run_fixed(dt, cm)
            outs.append(s.download())
    fd = (outs[0] - outs[1]) / (2 * eps)
    err = np.linalg.norm(fd - du) / np.linalg.norm(du)
    print(f"tangent vs central differences: relative norm error {err:.3e}")
    assert err <= 1e-5


def test_c5_gradient_vs_finite_differences_and_forward_mode():
    """The Orszag-Tang case, 256^2 x 300 steps, checkpoints every 20 steps:
    dJ/drho0 and dJ/dgamma from torch.autograd (the device adjoint) along 32
    seeded random directions against central differences (eps = 1e-6, FP64,
    step sizes held) and against forward-mode derivatives (tangent run)."""
    p = c5_params(256)
    u0 = orszag_tang_2d(p)
    rho0 = torch.tensor(u0[0], dtype=torch.float64, requires_grad=True)
    gamma = torch.tensor(p.gamma, dtype=torch.float64, requires_grad=True)
    J = c5_objective(rho0, gamma, p, nsteps=300, interval=20)
    J.backward()
    grho, ggam = rho0.grad.numpy(), float(gamma.grad)
    print(f"J = {float(J):.12e}, |dJ/drho0| = {np.linalg.norm(grho):.6e}, dJ/dgamma = {ggam:.6e}")
    # the taped step sizes of the base run
    v, b, pr = ot_fields(p)
    dx = p.dx()
    dvol = dx[0] * dx[1]

    def u_of(rho, g):
        q = np.zeros_like(u0)
        q[0], q[1:4], q[4:7], q[8] = rho, v, b, pr
        pp = c5_params(256)
        pp.gamma = g
        return prim_to_cons(pp, q), pp

    with Solver(p) as s:
        s.upload(u0)
        s.tape_forward(300, 20)
        dt, cm = s.tape_steps()
        uN = s.download()
        lamN = np.zeros_like(uN)
        vv = uN[1:4] / uN[0]
        lamN[1:4] = vv * dvol
        lamN[0] = -0.5 * (vv * vv).sum(0) * dvol
        rng = np.random.default_rng(20261202)
        eps = 1e-6
        worst_fd, worst_fw = 0.0, 0.0
        for n in range(32):
            drho = rng.normal(size=u0.shape[1:]) * 0.01
            dg = float(rng.normal()) * 0.01
            adj = float((grho * drho).sum()) + ggam * dg
            # forward mode: tangent of U0 for (drho, dg), then dJ = lamN . dU
            up, _ = u_of(u0[0] + eps * drho, p.gamma + eps * dg)
            um, _ = u_of(u0[0] - eps * drho, p.gamma - eps * dg)
            du0 = (up - um) / (2 * eps)      # prim2cons is linear in rho and 1/(g-1)
            fwd = float((lamN * s.tape_tangent(du0, dg)).sum())
            js = []
            for sgn in (1, -1):
                uu, pp = u_of(u0[0] + sgn * eps * drho, p.gamma + sgn * eps * dg)
                with Solver(pp) as f:
                    f.upload(uu)
                    f.run_fixed(dt, cm)
                    un = f.download()
                js.append(0.5 * ((un[1:4] ** 2).sum(0) / un[0]).sum() * dvol)
            fd = (js[0] - js[1]) / (2 * eps)
            e_fd = abs(adj - fd) / abs(fd)
            e_fw = abs(adj - fwd) / abs(fwd)
            worst_fd, worst_fw = max(worst_fd, e_fd), max(worst_fw, e_fw)
            if n < 4:
                print(f"  direction {n}: adjoint {adj:.12e} forward {fwd:.12e} fd {fd:.12e}")
    print(f"32 directions: max relative difference adjoint vs FD {worst_fd:.3e}, "
          f"adjoint vs forward mode {worst_fw:.3e}")
    assert worst_fd <= 1e-5
    assert worst_fw <= 1e-10
