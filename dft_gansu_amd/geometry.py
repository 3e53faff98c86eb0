"""Coordinates -> RHF energy as a differentiable torch function, and the
geometry optimiser driven by its analytic gradient.

The reference has no gradient; this is the Python test binding of the
engine's backward kernels (gansu_rhf_gradient in include/gansu_fock.h):
forward runs the SCF on the GPU, backward returns the analytic nuclear
gradient computed alongside it.
"""
import numpy as np
import torch

from .engine import Parameters, rhf_gradient


class RHFEnergy(torch.autograd.Function):
    """E_total(coords) for fixed atoms, basis and SCF parameters.

    coords: (n_atoms, 3) tensor in bohr.  backward gives dE/dcoords from the
    analytic gradient (hartree/bohr), which assumes the SCF converged tightly.
    """

    @staticmethod
    def forward(ctx, coords, mol, params):
        m = mol.with_coordinates(coords.detach().cpu().numpy().astype(np.float64))
        r = rhf_gradient(m, params)
        if not r["converged"]:
            raise RuntimeError("SCF did not converge; the analytic gradient needs the SCF solution")
        ctx.save_for_backward(torch.as_tensor(r["grad"], dtype=coords.dtype, device=coords.device))
        return coords.new_tensor(r["e_total"])

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g, None, None


def rhf_energy(coords, mol, params):
    return RHFEnergy.apply(coords, mol, params)


def gradient_params(**kv):
    """SCF parameters tight enough for an analytic gradient."""
    kv.setdefault("convergence_energy_threshold", 1e-11)
    kv.setdefault("maxiter", 200)
    return Parameters(**kv)


def optimize_geometry(mol, params=None, gtol=3e-4, max_steps=100, max_step=0.3, callback=None, warm_start=True):
    """Minimise the RHF energy over Cartesian coordinates by BFGS with a
    backtracking line search.

    gtol: largest gradient component at convergence (hartree/bohr).
    max_step: largest displacement of any coordinate in one step (bohr).
    warm_start: start each SCF from the last converged density (D/2, D/2),
    as the reference's optimisers hand densities to HF::solve.
    Returns dict(mol, e_total, grad, steps, converged, history, scf_iterations).
    """
    params = params or gradient_params()
    x = np.array(mol.xyz, dtype=np.float64).reshape(-1)
    last = {"D": None, "iterations": 0}

    def evaluate(xv):
        m = mol.with_coordinates(xv.reshape(-1, 3))
        D = last["D"] if warm_start else None
        r = rhf_gradient(m, params, density=None if D is None else (0.5 * D, 0.5 * D), force_density=True)
        if not r["converged"]:
            raise RuntimeError("SCF did not converge during the optimisation")
        last["D"] = r["D"]
        last["iterations"] += r["iterations"]
        return m, r["e_total"], r["grad"].reshape(-1)

    m, e, g = evaluate(x)
    n = x.size
    Hinv = np.eye(n) * 0.5  # inverse-Hessian guess (bohr^2/hartree)
    history = [(e, float(np.max(np.abs(g))))]
    converged = False
    steps = 0
    for steps in range(1, max_steps + 1):
        if np.max(np.abs(g)) < gtol:
            converged = True
            steps -= 1
            break
        p = -Hinv @ g
        if p @ g >= 0:  # not a descent direction: restart from steepest descent
            Hinv = np.eye(n) * 0.5
            p = -0.5 * g
        scale = min(1.0, max_step / max(np.max(np.abs(p)), 1e-30))
        p *= scale
        t = 1.0
        while True:
            xn = x + t * p
            mn, en, gn = evaluate(xn)
            if en <= e + 1e-4 * t * (g @ p) or t < 1e-3:
                break
            t *= 0.5
        s, y = xn - x, gn - g
        sy = s @ y
        if sy > 1e-12:
            rho = 1.0 / sy
            I = np.eye(n)
            Hinv = (I - rho * np.outer(s, y)) @ Hinv @ (I - rho * np.outer(y, s)) + rho * np.outer(s, s)
        x, m, e, g = xn, mn, en, gn
        history.append((e, float(np.max(np.abs(g)))))
        if callback:
            callback(steps, e, g)
    else:
        converged = np.max(np.abs(g)) < gtol
    return dict(mol=m, e_total=e, grad=g.reshape(-1, 3), steps=steps, converged=bool(converged), history=history,
                scf_iterations=last["iterations"])


__all__ = ["RHFEnergy", "rhf_energy", "gradient_params", "optimize_geometry"]
