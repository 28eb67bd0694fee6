"""The discrete adjoint behind torch.autograd (dgpr_amd.torch_adjoint):
  * torch.autograd.gradcheck of the residual f(U) = -Minv R(U) (every entry
    of the Jacobian by central differences against the adjoint's
    vector-Jacobian products) and of the final kinetic energy after RK4
    steps with respect to the initial state and to mu;
  * twenty seeded random directions of central differences (eps = 1e-6)
    against the autograd gradient of the final kinetic energy;
  * the tangent-linear dot-product test w . (df/dU v) = v . (df/dU)^T w,
    with df/dU v by central differences of the forward and the adjoint
    through torch.autograd.grad;
  * device (float64 HIP tensors, handed over by address) and host tensors
    give the same results.
Central differences carry O(eps^2) truncation and O(1e-16 |f| / eps)
rounding; the tolerances sit above both and far below a missing or
mis-signed term of the adjoint."""
import numpy as np
import pytest
import torch

from dgpr_amd import cases, dgx
from dgpr_amd import torch_adjoint as ta

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    dgx.lib()
    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device visible to torch")
    return torch.device("cuda:0")


def tiny():
    return cases.tgv_case(n=2, p=1, seed=61, M=0.3, Re=400.0, amp=0.1, dt=3e-3)


def test_residual_gradcheck(dev):
    c = tiny()
    with dgx.Solver(c.mesh, c.params) as s:
        U = torch.tensor(c.U0, dtype=torch.float64, device=dev, requires_grad=True)
        print("residual gradcheck: %d inputs" % U.numel(), flush=True)
        assert torch.autograd.gradcheck(lambda x: ta.residual(x, s), (U,), eps=1e-6, atol=1e-7,
                                        rtol=1e-5, nondet_tol=0.0)


def test_kinetic_energy_gradcheck(dev):
    c = tiny()
    with dgx.Solver(c.mesh, c.params) as s:
        U = torch.tensor(c.U0, dtype=torch.float64, device=dev, requires_grad=True)
        mu = torch.tensor(c.params.mu, dtype=torch.float64, device=dev, requires_grad=True)
        print("kinetic-energy gradcheck: %d + 1 inputs, 3 steps" % U.numel(), flush=True)
        assert torch.autograd.gradcheck(
            lambda x, m: ta.kinetic_energy_after_rk4(x, m, s, c.dt, 3, 2), (U, mu), eps=1e-6,
            atol=1e-8, rtol=1e-6, nondet_tol=0.0)


def test_kinetic_energy_fd_twenty_directions(dev):
    c = cases.tgv_case(n=3, p=2, seed=62, M=0.2, Re=800.0, amp=0.1, dt=2e-3)
    nsteps = 8
    with dgx.Solver(c.mesh, c.params) as s:
        U = torch.tensor(c.U0, dtype=torch.float64, device=dev, requires_grad=True)
        J = ta.kinetic_energy_after_rk4(U, c.params.mu, s, c.dt, nsteps)
        (g,) = torch.autograd.grad(J, U)
        g = g.cpu().numpy()

        def J_at(x):
            return float(ta.kinetic_energy_after_rk4(
                torch.tensor(x, dtype=torch.float64, device=dev), c.params.mu, s, c.dt, nsteps))

        rng = np.random.default_rng(620)
        eps = 1e-6
        worst = 0.0
        for k in range(20):
            d = c.U0 * rng.uniform(-1.0, 1.0, c.U0.shape)
            fd = (J_at(c.U0 + eps * d) - J_at(c.U0 - eps * d)) / (2 * eps)
            ad = float(g @ d)
            worst = max(worst, abs(fd - ad) / abs(ad))
        print("  20 directions: worst rel err %.2e" % worst, flush=True)
        assert worst < 1e-6


def test_residual_tangent_linear_dot_product(dev):
    c = cases.tgv_case(n=3, p=3, seed=63, M=0.25, Re=600.0, amp=0.12, dt=1e-3)
    with dgx.Solver(c.mesh, c.params) as s:
        rng = np.random.default_rng(630)
        v = c.U0 * rng.uniform(-1.0, 1.0, c.U0.shape)
        w = rng.standard_normal(c.U0.shape)
        U = torch.tensor(c.U0, dtype=torch.float64, device=dev, requires_grad=True)
        f = ta.residual(U, s)
        (adj,) = torch.autograd.grad(f, U, grad_outputs=torch.tensor(w, device=dev))
        eps = 1e-6

        def f_at(x):
            return ta.residual(torch.tensor(x, dtype=torch.float64, device=dev), s).cpu().numpy()

        tl = (f_at(c.U0 + eps * v) - f_at(c.U0 - eps * v)) / (2 * eps)
        lhs, rhs = float(w @ tl), float(adj.cpu().numpy() @ v)
        print("  <w, J v> %.15e  <J^T w, v> %.15e" % (lhs, rhs), flush=True)
        assert abs(lhs - rhs) <= 1e-7 * abs(rhs)


def test_host_and_device_tensors_agree(dev):
    c = tiny()
    with dgx.Solver(c.mesh, c.params) as s:
        w = np.random.default_rng(64).standard_normal(c.U0.shape)
        out = {}
        for name, d in (("host", torch.device("cpu")), ("device", dev)):
            U = torch.tensor(c.U0, dtype=torch.float64, device=d, requires_grad=True)
            f = ta.residual(U, s)
            (g,) = torch.autograd.grad(f, U, grad_outputs=torch.tensor(w, device=d))
            mu = torch.tensor(c.params.mu, dtype=torch.float64, device=d, requires_grad=True)
            J = ta.kinetic_energy_after_rk4(U.detach().requires_grad_(), mu, s, c.dt, 2)
            (gmu,) = torch.autograd.grad(J, mu)
            out[name] = (f.detach().cpu().numpy(), g.cpu().numpy(), float(J.detach()), float(gmu))
        for a, b in zip(out["host"], out["device"]):
            assert np.allclose(a, b, rtol=1e-13, atol=0.0)
