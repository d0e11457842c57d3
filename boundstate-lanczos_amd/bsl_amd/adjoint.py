"""Differentiable bound-state energies: a torch.autograd.Function over the
MI355X calculator.

Forward: build the cage operator on the GPU from the potential parameters,
run the Lanczos recurrence, pick the lowest converged levels, rebuild their
Ritz vectors with a second Lanczos pass (recomputing the Krylov vectors
rather than storing them) and refine each level by an FP64 Rayleigh
quotient.  Backward: Hellmann-Feynman, dE_k/dtheta = <psi_k|dV/dtheta|psi_k>,
evaluated by the library's adjoint kernels for the incoming weights dL/dE_k
in one pass, on the Ritz vectors the forward pass left on the device
(bsl_hf_gradient_ritz).

Near-degenerate levels (m-multiplets) are individually differentiable only
while they stay resolved; their sum is smooth, so losses should use sums over
`multiplets(...)` groups where levels approach each other.

The parameters are plain torch tensors (float64, CPU): site positions (nm),
site charges (e), the two Alavi H2 charges (e) and the LJ coefficients A, B of
the H2 centre - water O pair.  `water_model_sites` maps water-model
parameters (two charges, LJ sigma/epsilon of O and of the H2 centre) onto
those with torch operations, so autograd carries gradients to them too.
"""
import numpy as np
import torch

from . import Calculator

__all__ = ["CageLevels", "cage_levels", "water_model_sites"]


class CageLevels(torch.autograd.Function):
    """energies = CageLevels.apply(site_pos, site_q, q_h2, lj_ab, config)

    site_pos (n, 3), site_q (n), q_h2 (2: H, centre), lj_ab (2: A, B).
    config: dict with "calc" (Calculator keyword arguments other than the
    potential parameters, e.g. nx, x_max, lmax, n_theta, n_phi, mass,
    rot_const, ceiling, precision), "lj_sites" (n flags: 1 = LJ site),
    "niter", "count" and optionally "seed"."""

    @staticmethod
    def forward(ctx, site_pos, site_q, q_h2, lj_ab, config):
        pos = site_pos.detach().cpu().to(torch.float64).numpy().reshape(-1, 3)
        q = site_q.detach().cpu().to(torch.float64).numpy()
        qh = q_h2.detach().cpu().to(torch.float64).numpy()
        ab = lj_ab.detach().cpu().to(torch.float64).numpy()
        lj = np.asarray(config["lj_sites"], dtype=np.int32)
        calc = Calculator(potential="cage", sites=(pos, q, lj), q_h2_h=float(qh[0]), q_h2_cm=float(qh[1]),
                          lj_a=float(ab[0]), lj_b=float(ab[1]), **config["calc"])
        # the Ritz vectors stay on the device for the backward pass
        e, _, _ = calc.levels(config["niter"], config["count"], seed=config.get("seed", 0), vectors=False)
        ctx.calc = calc
        return torch.as_tensor(e, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_e):
        g = grad_e.detach().cpu().to(torch.float64).numpy()
        gr = ctx.calc.hf_gradient_ritz([(np.arange(len(g)), g)])[0]
        ctx.calc.close()
        ctx.calc = None
        return (torch.as_tensor(gr["site_pos"]), torch.as_tensor(gr["site_q"]),
                torch.tensor([gr["q_h2_h"], gr["q_h2_cm"]], dtype=torch.float64),
                torch.tensor([gr["lj_a"], gr["lj_b"]], dtype=torch.float64), None)


def cage_levels(site_pos, site_q, q_h2, lj_ab, config):
    """Functional form of CageLevels.apply."""
    return CageLevels.apply(site_pos, site_q, q_h2, lj_ab, config)


def water_model_sites(charges, lj, positions, kinds, mixing="geometric"):
    """Site charges and LJ coefficients from water-model parameters.

    charges: tensor (2,) = (q_O, q_H) of the water model; lj: tensor (4,) =
    (sigma_O nm, eps_O kJ/mol, sigma_H2 nm, eps_H2 kJ/mol); positions: (n, 3);
    kinds: sequence of "O"/"H" per site.  Returns (site_pos, site_q, lj_ab)
    as differentiable tensors: A = 4 eps sigma^12, B = 4 eps sigma^6 with
    eps = sqrt(eps_O eps_H2) and sigma = sqrt(sigma_O sigma_H2) ("geometric",
    the reference's rule, SPCE_AH2_EngRoutines.h:35-38) or the arithmetic
    sigma ("lorentz-berthelot")."""
    is_o = torch.tensor([k == "O" for k in kinds], dtype=torch.bool)
    site_q = torch.where(is_o, charges[0], charges[1])
    eps = torch.sqrt(lj[1] * lj[3])
    sig = torch.sqrt(lj[0] * lj[2]) if mixing == "geometric" else 0.5 * (lj[0] + lj[2])
    lj_ab = torch.stack([4.0 * eps * sig ** 12, 4.0 * eps * sig ** 6])
    return positions, site_q, lj_ab
