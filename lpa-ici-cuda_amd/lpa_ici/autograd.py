"""torch.autograd.Function over the C ABI — the gradient harness.

LpaIciFunction.apply(x, taps, template, sigma, gamma) denoises one device
image x (H x W float32) with the kernel set whose run weights are `taps` (a
1-D float32 tensor in the run-weight order of template.runs(), e.g.
template_taps(template)), and returns y.  The forward keeps y, the uint8
scale map and the per-direction estimates at the selected scales; the
backward calls lpa_ici_backward with them (the selection stays frozen, so y
is linear in x and rational in the taps through sigma^2 * sum(w^2)).
"""
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


class LpaIciFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, taps, template, sigma, gamma):
        ks = _kernel_set_with_taps(template, taps)
        y, maps, est = denoise(ks, x.contiguous(), sigma, gamma, scale_maps=True, keep_estimates=True)
        ctx.save_for_backward(x, y, maps, est)
        ctx.ks, ctx.sigma, ctx.gamma = ks, sigma, gamma
        return y

    @staticmethod
    def backward(ctx, gy):
        x, y, maps, est = ctx.saved_tensors
        dxs, dk = backward(ctx.ks, [x], [y], [maps], [gy.contiguous()], ctx.sigma, ctx.gamma, estimates=[est])
        return dxs[0], dk, None, None, None


def lpa_ici(x, taps, template, sigma=20.0 / 255.0, gamma=2.0):
    return LpaIciFunction.apply(x, taps, template, sigma, gamma)
