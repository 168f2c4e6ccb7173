"""Shared helpers for the parity tests."""
import json
import os
import re
import shutil
import subprocess

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


# A variable is "order one" in a run when its mean magnitude over the
# interior is at least ORDER_ONE times its fluid's mean density (densities,
# energies, the background field, a shear flow's momentum).  The others
# (perturbation-born momenta and fields, B_z, psi) are set by the flow's
# perturbation or by round-off, and ulp-level differences are large relative
# to them.  The classification uses the oracle's state.
ORDER_ONE = 0.05


def order_one(U, cfg):
    Ui = interior(U, cfg)
    rho = np.abs(Ui[0]).mean()
    return [k for k in range(Ui.shape[0]) if np.abs(Ui[k]).mean() >= ORDER_ONE * rho]


def relaxed_errors(gm, gh, om, oh, cfg):
    """(max relative L1 over the order-one variables, max over the others)."""
    em, eh = rel_l1(gm, om, cfg), rel_l1(gh, oh, cfg)
    bm, bh = order_one(om, cfg), order_one(oh, cfg)
    big = [em[i] for i in bm] + [eh[i] for i in bh]
    small = [em[i] for i in range(len(em)) if i not in bm] + [eh[i] for i in range(5) if i not in bh]
    return max(big), max(small)


# the datasets the reference's output path creates per flag setting
SNAPSHOT_TABLE = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "snapshot_datasets.json")))


def snapshot_case(name):
    return next(c for c in SNAPSHOT_TABLE["cases"] if c["name"] == name)


def h5_listing(path):
    """{dataset name: (dims tuple or 'SCALAR', type)} from h5dump -H, or None
    when h5dump is not installed."""
    from pip_amd import build
    h5dump = shutil.which("h5dump") or os.path.join(build.HDF5_PREFIX, "bin", "h5dump")
    if not os.path.exists(h5dump):
        return None
    txt = subprocess.run([h5dump, "-H", path], capture_output=True, text=True, check=True).stdout
    out = {}
    for m in re.finditer(r'DATASET "([^"]+)" \{\s*DATATYPE\s+(\S+)\s*DATASPACE\s+(SCALAR|SIMPLE \{ \( ([^)]*) \))',
                         txt):
        dims = "SCALAR" if m.group(3) == "SCALAR" else tuple(int(v) for v in m.group(4).split(","))
        out[m.group(1)] = (dims, m.group(2))
    return out
