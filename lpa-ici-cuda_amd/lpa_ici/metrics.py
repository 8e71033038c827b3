"""Work counts for throughput and roofline figures.

useful taps: the walk evaluates, for every pixel and direction, the nonzero
taps of scales 0..min(s*+1, S-1), where s* is the selected scale (the first
rejected scale is evaluated to see the interval break).  FLOP = 2 per tap.
"""
import numpy as np

# MI355X vector FP32: 256 CUs x 2.4 GHz x 256 FLOP/clk (packed FP32 FMA).
MI355X_FP32_PEAK_TFLOPS = 157.3
MI355X_HBM_PEAK_GBS = 8000.0


def useful_taps(scale_maps, nnz, num_scales):
    """scale_maps: uint8 [D, H, W] (numpy or torch); nnz: [D*S] nonzero taps."""
    try:
        import torch
        if isinstance(scale_maps, torch.Tensor):
            return _useful_taps_torch(scale_maps, nnz, num_scales)
    except ImportError:  # pragma: no cover
        pass
    m = np.asarray(scale_maps)
    D = m.shape[0]
    cum = np.cumsum(np.asarray(nnz, np.int64).reshape(D, num_scales), axis=1)
    total = 0
    for d in range(D):
        hist = np.bincount(m[d].ravel(), minlength=num_scales)[:num_scales]
        depth = np.minimum(np.arange(num_scales) + 1, num_scales - 1)
        total += int((hist * cum[d, depth]).sum())
    return total


def _useful_taps_torch(maps, nnz, S):
    import torch
    D = maps.shape[0]
    cum = np.cumsum(np.asarray(nnz, np.int64).reshape(D, S), axis=1)
    depth = np.minimum(np.arange(S) + 1, S - 1)
    total = 0
    for d in range(D):
        hist = torch.bincount(maps[d].reshape(-1).to(torch.int64), minlength=S)[:S].cpu().numpy()
        total += int((hist * cum[d, depth]).sum())
    return total
