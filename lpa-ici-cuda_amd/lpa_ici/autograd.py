"""torch.autograd.Function over the C ABI — the gradient harness.

LpaIciFunction.apply(x, taps, template, sigma, gamma, keep_estimates)
denoises one device image x (H x W float32, rows may be pitched) with the
kernel set whose run weights are `taps` (a 1-D float32 tensor in the
run-weight order of template.runs(), e.g. template_taps(template)), and
returns y.  The forward saves the uint8 scale map (16 B per pixel for 16
directions) and the adjoint recomputes each direction's estimate from it;
with keep_estimates=True the forward also keeps the D x H x W selected
estimates and the adjoint reads them instead.  The selection stays frozen, so
y is linear in x and rational in the taps through sigma^2 * sum(w^2).

The kernel set built from `taps` is kept per template and reused while the
taps hold the same values (a device comparison per forward), so a loop that
does not change the taps uploads the set once.
"""
import weakref

import numpy as np
import torch

from . import KernelSet, backward, denoise


def template_taps(template, device="cuda"):
    """The template set's run weights as a tensor (interior zeros included)."""
    return torch.from_numpy(template.runs()["weights"].copy()).to(device)


def _kernel_set_with_taps(template, taps):
    """A KernelSet with the template's geometry and these run weights."""
    r = template.runs()
    dense = template.weights.copy()
    t = taps.detach().cpu().numpy().astype(np.float32)
    for k in range(template.K):
        R, C, o = template.rows[k], template.cols[k], template.offsets[k]
        m = dense[o:o + R * C].reshape(R, C)
        for i in range(r["kernel_first_run"][k], r["kernel_first_run"][k] + r["kernel_num_runs"][k]):
            row, col = r["run_row"][i] + R // 2, r["run_col"][i] + C // 2
            ln, wo = r["run_len"][i], r["run_woff"][i]
            m[row, col:col + ln] = t[wo:wo + ln]
    return KernelSet(template.D, template.S, dense, template.offsets, template.rows, template.cols)


# template -> (copy of the taps it was built with, KernelSet)
_sets = weakref.WeakKeyDictionary()


def kernel_set_for(template, taps):
    """The KernelSet for these taps: the cached one while the values match."""
    t = taps.detach()
    hit = _sets.get(template)
    if hit is not None and hit[0].shape == t.shape and hit[0].device == t.device and torch.equal(hit[0], t):
        return hit[1]
    ks = _kernel_set_with_taps(template, t)
    _sets[template] = (t.clone(), ks)
    return ks


class LpaIciFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, taps, template, sigma, gamma, keep_estimates=False):
        ks = kernel_set_for(template, taps)
        # rows may be pitched (unit stride along a row is all the C ABI needs)
        xin = x if x.dim() == 2 and x.stride(1) == 1 else x.contiguous()
        if keep_estimates:
            y, maps, est = denoise(ks, xin, sigma, gamma, scale_maps=True, keep_estimates=True)
            ctx.save_for_backward(xin, y, maps, est)
        else:
            y, maps = denoise(ks, xin, sigma, gamma, scale_maps=True)
            ctx.save_for_backward(xin, y, maps)
        ctx.ks, ctx.sigma, ctx.gamma = ks, sigma, gamma
        return y

    @staticmethod
    def backward(ctx, gy):
        saved = ctx.saved_tensors
        x, y, maps = saved[:3]
        est = [saved[3]] if len(saved) > 3 else None
        g = gy if gy.stride() == y.stride() else _like(y, gy)
        dxs, dk = backward(ctx.ks, [x], [y], [maps], [g], ctx.sigma, ctx.gamma, estimates=est)
        return dxs[0], dk, None, None, None, None


def _like(y, g):
    """g copied into a tensor with y's layout (the adjoint reads dy at y's pitch)."""
    out = torch.empty_strided(tuple(y.shape), y.stride(), dtype=y.dtype, device=y.device)
    out.copy_(g)
    return out


def lpa_ici(x, taps, template, sigma=20.0 / 255.0, gamma=2.0, keep_estimates=False):
    return LpaIciFunction.apply(x, taps, template, sigma, gamma, keep_estimates)
