"""torch.autograd front end of the discrete adjoint of libdgx.

``ResidualFunction``  f = -Minv R(U), the residual of the reference's CPU
    solver (solve.cpp:150-807); backward gives (df/dU)^T g through
    ``dgx_residual_vjp``.
``KineticEnergyRK4``  J(U0, mu) = volume-integrated kinetic energy after n
    RK4 steps (stage algebra of time_distributed.cpp:46-102); backward gives
    dJ/dU0 and dJ/dmu through ``dgx_adjoint_rk4`` (discrete adjoint of the
    stages: every stage state kept from the forward sweep when it fits in
    device memory, else checkpoints every ceil(sqrt(n)) steps, or every
    ``interval`` steps when given).  The stored stages stay with the solver
    for later backward passes until ``solver.adjoint_release()`` or
    ``solver.close()``.

Both act on meshes with any boundary types.  States are float64 tensors in the simstate layout U[e][r][bf]:
device tensors are handed to the library by address after the caller's
stream is drained, host tensors are copied.  The library runs on its own
stream and every call returns after completion, so results are ready when
these functions return.  There is no CPU fallback: without libdgx or a HIP
device the calls raise.
"""
import torch

from . import dgx


def _check_state(U, solver):
    if U.dtype != torch.float64:
        raise TypeError("states are float64 tensors")
    if U.numel() != solver.nstate:
        raise ValueError("state has %d entries, expected %d" % (U.numel(), solver.nstate))


def _load_state(solver, U):
    """Make U (any device) the solver's current state."""
    if U.is_cuda:
        Uc = U.detach().contiguous()
        torch.cuda.current_stream(Uc.device).synchronize()
        solver.set_state_device(Uc.data_ptr())
    else:
        solver.set_state(U.detach().contiguous().numpy())


class ResidualFunction(torch.autograd.Function):
    """f(U) = -Minv R(U) on ``solver``'s mesh and parameters."""

    @staticmethod
    def forward(ctx, U, solver):
        _check_state(U, solver)
        ctx.solver = solver
        ctx.save_for_backward(U)
        if U.is_cuda:
            Uc = U.detach().contiguous()
            f = torch.empty_like(Uc)
            torch.cuda.current_stream(Uc.device).synchronize()
            solver.residual_device(Uc.data_ptr(), 0, f.data_ptr())
            return f.view_as(U)
        solver.set_state(U.detach().contiguous().numpy())
        _, f = solver.residual()
        return torch.from_numpy(f).view_as(U)

    @staticmethod
    def backward(ctx, g):
        (U,) = ctx.saved_tensors
        solver = ctx.solver
        _load_state(solver, U)
        gU = solver.residual_vjp(g.detach().reshape(-1).cpu().numpy())
        return torch.from_numpy(gU).to(U.device).view_as(U), None


class KineticEnergyRK4(torch.autograd.Function):
    """J(U0, mu) after ``nsteps`` RK4 steps of size ``dt``; ``interval`` is
    the checkpoint spacing of the backward pass: 0 keeps every stage state
    from the forward sweep when they fit in three quarters of the free
    device memory (nothing recomputed) and checkpoints every
    ceil(sqrt(nsteps)) steps otherwise; k >= 1 checkpoints every k steps."""

    @staticmethod
    def forward(ctx, U0, mu, solver, dt, nsteps, interval=0):
        _check_state(U0, solver)
        prm = solver.params.copy()
        prm.mu = float(mu)
        solver.set_params(prm)
        _load_state(solver, U0)
        solver.step(dt, "RK4", int(nsteps), norms=False)
        J = solver.kinetic_energy()
        ctx.solver, ctx.dt, ctx.nsteps, ctx.interval = solver, dt, int(nsteps), int(interval)
        ctx.save_for_backward(U0, mu)
        return torch.tensor(J, dtype=torch.float64, device=U0.device)

    @staticmethod
    def backward(ctx, gJ):
        U0, mu = ctx.saved_tensors
        solver = ctx.solver
        prm = solver.params.copy()
        prm.mu = float(mu)
        solver.set_params(prm)
        _load_state(solver, U0)
        want_mu = bool(ctx.needs_input_grad[1])
        _, g, dmu = solver.adjoint_rk4(ctx.dt, ctx.nsteps, ctx.interval, want_mu=want_mu)
        s = gJ.detach().to(torch.float64)
        gU = torch.from_numpy(g).to(U0.device).view_as(U0) * s.to(U0.device)
        gmu = (torch.tensor(dmu, dtype=torch.float64, device=mu.device) * s.to(mu.device)
               if want_mu else None)
        return gU, gmu, None, None, None, None


def residual(U, solver):
    """f(U) with autograd support (see ResidualFunction)."""
    return ResidualFunction.apply(U, solver)


def kinetic_energy_after_rk4(U0, mu, solver, dt, nsteps, interval=0):
    """J(U0, mu) with autograd support (see KineticEnergyRK4)."""
    if not torch.is_tensor(mu):
        mu = torch.tensor(float(mu), dtype=torch.float64)
    return KineticEnergyRK4.apply(U0, mu, solver, dt, nsteps, interval)


__all__ = ["ResidualFunction", "KineticEnergyRK4", "residual", "kinetic_energy_after_rk4"]
