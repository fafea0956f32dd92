"""Differentiable power spectrum: P(k) of a CIC-deposited particle set with
gradients w.r.t. positions, masses and the normalising total mass.

Forward: deposit -> r2c -> deconvolved binning (the reference's fieldize +
FFTW + powerspectrum, gen-pk.cpp:208-234), all on the GPU through the C ABI.

Backward (hand-written adjoint, no counterpart in the reference): with
lambda_b = dL/dP_b,
  A_k = 2 lambda_b(k) g_k F_k / (M^2 C_b(k)),   dL/drho = c2r(A),
  dL/dx_j = m_j (d/box) sum_c (dw_jc/du) dL/drho(c),  dL/dm_j = sum_c w_jc dL/drho(c),
  dL/dM = -2 sum_b lambda_b P_b / M.
dL/drho = c2r(A) is formed without writing A: the adjoint's x pass runs
forward FFT -> weight -> inverse FFT per column block in LDS, starting from
the field after the forward z and y passes.  The gather reads dL/drho by
deposit tile, in the particles' tile order.  Where the forward state comes
from is the recompute policy of PowerSpectrumFunction.
"""
import torch

from . import Plan

__all__ = ["power_spectrum", "PowerSpectrumFunction", "chi2_loss"]


_POLICIES = ("auto", "always", "never")


class PowerSpectrumFunction(torch.autograd.Function):
    """Forward/backward of P(k) through one plan.

    recompute policies (the forward spectrum is what the backward needs):
      "auto":   the forward leaves its state in the plan (the field after the
                z and y passes, plus the particles' tile order); the backward
                starts from it if the plan has not been used since, and
                otherwise recomputes deposit + transform.
      "always": the backward recomputes deposit + transform; the deposit
                runs from the particles' tile order the forward kept (their
                sorted copy, particle data), so the partition is not redone
                while the plan still holds it.
      "never":  the forward keeps a copy of the spectrum (d*d*(d/2+1)
                complex values) for the backward.
    """

    @staticmethod
    def forward(ctx, pos, masses, total_mass, plan, boxsize, nrbins, mass, recompute):
        M = float(total_mass)
        mdet = None if masses is None else masses.detach()
        plan.clear()
        plan.deposit(pos.detach(), boxsize, mdet, mass, retain=True)
        spec = None
        if recompute == "never":
            spec = plan.r2c().clone()
            P, _, _ = plan.bin(nrbins, mass=M)
        else:
            P, _, _ = plan.power(nrbins, mass=M)
        ctx.plan, ctx.boxsize, ctx.nrbins, ctx.mass, ctx.M = plan, boxsize, nrbins, mass, M
        ctx.spec, ctx.recompute, ctx.generation = spec, recompute, plan.generation
        ctx.has_masses = masses is not None
        ctx.m_device = total_mass.device
        ctx.save_for_backward(pos, masses if masses is not None else torch.empty(0, device=pos.device), P)
        return P

    @staticmethod
    def backward(ctx, gP):
        pos, masses, P = ctx.saved_tensors
        mdet = masses.detach() if ctx.has_masses else None
        plan = ctx.plan
        if ctx.spec is not None:
            plan.import_spectrum(ctx.spec)
        elif ctx.recompute == "always" or plan.generation != ctx.generation:
            # recompute the forward density (the adjoint then redoes its
            # transform): from the tile order the forward kept when it is
            # still this call's, else by a full deposit
            plan.clear()
            if plan.keeps_order_of(pos.detach(), ctx.boxsize, mdet, ctx.mass):
                plan.redeposit()
            else:
                plan.deposit(pos.detach(), ctx.boxsize, mdet, ctx.mass, retain=True)
        lam = gP.detach().to(torch.float64).contiguous()
        plan.adjoint_density(lam, ctx.M, ctx.nrbins)
        want_m = ctx.needs_input_grad[1]
        gpos, gm = plan.adjoint_gather(pos.detach(), ctx.boxsize, mdet, ctx.mass, want_mass_grad=want_m)
        gM = None
        if ctx.needs_input_grad[2]:
            gM = (-2.0 * (lam * P).sum() / ctx.M).to(device=ctx.m_device, dtype=torch.float64)
        return gpos, (gm if want_m else None), gM, None, None, None, None, None


def power_spectrum(pos, boxsize, dims, nrbins=None, masses=None, mass=1.0, total_mass=None, precision="f64",
                   plan=None, recompute="auto"):
    """Binned, CIC-deconvolved P(k) of particles (auto spectrum), differentiable.

    pos: float32 [n, 3] device tensor (may require grad); masses: float32 [n]
    or None (every particle then has ``mass``).  total_mass: the M of the
    normalisation P /= M^2 (powerspectrum.c:104); a tensor keeps it in the
    autograd graph (e.g. ``masses.sum()``); default: sum of the masses.
    Returns P (float64 [nrbins]); counts and keffs are plan.bin() outputs.
    """
    nrbins = dims if nrbins is None else nrbins
    if plan is None:
        plan = Plan(dims, precision)
    if total_mass is None:
        total_mass = masses.double().sum() if masses is not None else torch.tensor(float(mass) * pos.shape[0],
                                                                                   dtype=torch.float64)
    if not torch.is_tensor(total_mass):
        total_mass = torch.tensor(float(total_mass), dtype=torch.float64)
    if recompute is True:
        recompute = "always"
    elif recompute is False:
        recompute = "never"
    assert recompute in _POLICIES, f"recompute must be one of {_POLICIES}"
    needs = torch.is_grad_enabled() and (pos.requires_grad or total_mass.requires_grad or
                                         (masses is not None and masses.requires_grad))
    if not needs:
        plan.clear()
        plan.deposit(pos.detach(), boxsize, None if masses is None else masses.detach(), mass)
        return plan.power(int(nrbins), mass=float(total_mass))[0]
    return PowerSpectrumFunction.apply(pos, masses, total_mass, plan, float(boxsize), int(nrbins), float(mass),
                                       recompute)


def chi2_loss(P, target, count):
    """Gaussian band-power loss sum_b (P_b - T_b)^2 / sigma_b^2 over the
    occupied bins, sigma_b = T_b sqrt(2 / count_b) (cosmic variance of a
    band power estimated from count_b modes)."""
    cnt = count.to(torch.float64)
    occ = (cnt > 0) & (target > 0)
    sig2 = torch.where(occ, target * target * 2.0 / torch.clamp(cnt, min=1), torch.ones_like(target))
    return torch.where(occ, (P - target) ** 2 / sig2, torch.zeros_like(P)).sum()
