"""Shared helpers for the parity tests."""
import numpy as np


def interior(a, cfg):
    if cfg.ndim == 2:
        return a[..., 2:-2, 2:-2]
    return a[..., :, 2:-2]


def rel_l1(a, b, cfg):
    """Relative L1 error per variable over interior cells: sum|a-b|/sum|b|
    (variables that are identically zero use the absolute L1 per cell)."""
    ai, bi = interior(a, cfg), interior(b, cfg)
    out = []
    for n in range(ai.shape[0]):
        den = np.abs(bi[n]).sum()
        num = np.abs(ai[n] - bi[n]).sum()
        out.append(num / den if den > 0 else num / bi[n].size)
    return np.array(out)
