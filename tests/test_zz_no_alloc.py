"""pip_adj_step_noalloc, the allocation-free variant of pip_adj_step, through
the C-ABI: on two calls in a row -- a 96^2 grid, then a smaller 40^2 grid,
each pulled back through two steps -- its adjoint of the state and of the
collision field equal pip_adj_step's bit for bit (same kernels; the work
buffers live in the library's fixed device array instead of memory the
library allocates).  A grid larger than that array is rejected with an
error before any device work, and the saved primitives are required."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from pip_amd import problems  # noqa: E402
from pip_amd._lib import check, lib  # noqa: E402

pytestmark = pytest.mark.gpu


def _adjoint_inputs(n, nsteps=3):
    """A recorded Alfven-wave run of n^2 cells, replayed with the stage
    inputs and their primitives saved."""
    from pip_amd.adjoint import TwoFluidRun, _ptr
    cfg, uh, um, ac = problems.alfven_damping(n)
    run = TwoFluidRun(cfg, nsteps, ckpt_every=nsteps)
    dev = torch.device("cuda:0")
    U0 = torch.tensor(np.concatenate([um, uh]), device=dev)
    acT = torch.tensor(ac, device=dev)
    _, _, hist = run.forward_record(U0, acT)
    bufs = [run._empty_state(nsteps) for _ in range(4)]
    run._start(U0, acT)
    check(lib().pip_replay_steps(run.h, nsteps, _ptr(hist), *[_ptr(b) for b in bufs]))
    run.solver.sync()
    return cfg, run, hist, bufs


def _pull_back(fn, run, hist, bufs, lam, lam_ac, steps):
    from pip_amd.adjoint import _ptr
    sb = run.nvt * run.cfg.ix * run.cfg.jx * 8
    Un, U1, Vn, V1 = bufs
    torch.cuda.synchronize()
    for k in steps:
        check(fn(run.h, _ptr(Un) + sb * k, _ptr(U1) + sb * k, _ptr(Vn) + sb * k, _ptr(V1) + sb * k,
                 _ptr(hist) + 16 * k, _ptr(lam), _ptr(lam_ac)))
    run.solver.sync()


@pytest.mark.parametrize("n", [96, 40])
def test_noalloc_adjoint_step_equals_pip_adj_step(n):
    print(f"case pip_adj_step_noalloc vs pip_adj_step n={n}, steps 2 and 1", flush=True)
    cfg, run, hist, bufs = _adjoint_inputs(n)
    g = torch.Generator(device="cuda:0").manual_seed(n)
    lam0 = torch.randn((run.nvt, cfg.jx, cfg.ix), dtype=torch.float64, device="cuda:0", generator=g)
    lam0[:, :2] = 0.0
    lam0[:, -2:] = 0.0
    lam0[:, :, :2] = 0.0
    lam0[:, :, -2:] = 0.0
    out = {}
    for name in ("pip_adj_step", "pip_adj_step_noalloc"):
        lam = lam0.clone()
        lam_ac = torch.zeros((cfg.jx, cfg.ix), dtype=torch.float64, device="cuda:0")
        _pull_back(getattr(lib(), name), run, hist, bufs, lam, lam_ac, steps=(2, 1))
        out[name] = (lam, lam_ac)
    a, b = out["pip_adj_step"], out["pip_adj_step_noalloc"]
    assert torch.isfinite(a[0]).all() and float(a[1].abs().max()) > 0.0
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])
    run.close()


def test_noalloc_adjoint_step_rejects_large_grids_and_missing_primitives():
    print("case pip_adj_step_noalloc argument checks (n=340)", flush=True)
    from pip_amd.adjoint import TwoFluidRun
    cfg = problems.alfven_damping(340)[0]
    assert cfg.ix * cfg.jx * (5 * 14 + 4) * 8 > 64 << 20
    run = TwoFluidRun(cfg, 1)
    d = torch.zeros(16, dtype=torch.float64, device="cuda:0")
    p = ctypes.c_void_p(d.data_ptr())
    rc = lib().pip_adj_step_noalloc(run.h, p, p, p, p, p, p, p)
    assert rc == -1 and b"work memory" in lib().pip_last_error()
    run.close()
    cfg, run, hist, bufs = _adjoint_inputs(24, nsteps=1)
    rc = lib().pip_adj_step_noalloc(run.h, p, p, None, None, p, p, p)
    assert rc == -1 and b"Vn and V1" in lib().pip_last_error()
    run.close()
