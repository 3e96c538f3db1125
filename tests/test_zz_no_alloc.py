"""mg_propagate_noalloc (no allocation: the split's partial sums in the
library's fixed device array) against mg_propagate_ws with a caller
workspace, through the C ABI: two calls in a row, the second smaller, then a
size over the array's maximum, which is rejected before any launch."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from magreete_amd import _lib as L
from magreete_amd import workloads as W

DEV = "cuda:0"


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _case(n_scat, n_meas, T, seed):
    rng = np.random.default_rng(seed)
    scat = torch.from_numpy(W.unit_ball(n_scat, rng)).to(DEV)
    meas = torch.from_numpy(2.0 * W.fibonacci_sphere(n_meas)).to(DEV)
    Ek = torch.from_numpy(rng.normal(size=(3 * n_scat, T)) + 1j * rng.normal(size=(3 * n_scat, T))).to(DEV)
    E0 = torch.from_numpy(rng.normal(size=(3 * n_meas, T)) + 1j * rng.normal(size=(3 * n_meas, T))).to(DEV)
    return scat, meas, Ek, E0


def _run(lib, fn, scat, meas, Ek, E0, k0, alpha, work=None):
    out = E0.clone()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    args = [L.MG_3D_VECTOR, _p(meas), meas.shape[0], _p(scat), scat.shape[0], k0, alpha.real, alpha.imag, 0, 0.0,
            _p(Ek), Ek.shape[1], Ek.shape[1], 1.0, 0.0, _p(out), out.shape[1]]
    if fn == "ws":
        st = lib.mg_propagate_ws(*args, _p(work) if work is not None else ctypes.c_void_p(0),
                                 work.numel() if work is not None else 0, s)
    else:
        st = lib.mg_propagate_noalloc(*args, s)
    L.check(st, fn)
    return out


def test_noalloc_matches_workspace_entry():
    lib = L.load()
    cap = lib.mg_propagate_noalloc_max_bytes()
    assert cap == 64 << 20
    k0 = 9.0
    alpha = 0.002 - 0.0007j
    for n_scat, n_meas, T, seed in ((1500, 700, 96, 1), (700, 130, 17, 2)):
        print("noalloc", n_scat, n_meas, T, flush=True)
        scat, meas, Ek, E0 = _case(n_scat, n_meas, T, seed)
        need = lib.mg_propagate_work_bytes(L.MG_3D_VECTOR, n_meas, n_scat, T)
        assert 0 < need <= cap, need          # both calls take the split path
        work = torch.empty(need, dtype=torch.uint8, device=DEV)
        a = _run(lib, "ws", scat, meas, Ek, E0, k0, alpha, work)
        b = _run(lib, "noalloc", scat, meas, Ek, E0, k0, alpha)
        torch.cuda.synchronize()
        assert torch.equal(a, b)
        assert torch.isfinite(b).all()
    # over the maximum: rejected before anything is launched (pointers unused)
    # (3200 points x 256 columns: 400 output tiles of 96 x 64, split in two)
    big = lib.mg_propagate_work_bytes(L.MG_3D_VECTOR, 3200, 600, 256)
    assert big > cap, big
    null = ctypes.c_void_p(0)
    st = lib.mg_propagate_noalloc(L.MG_3D_VECTOR, null, 3200, null, 600, 1.0, 1.0, 0.0, 0, 0.0, null, 256, 256,
                                  1.0, 0.0, null, 256, null)
    assert st == -1 and b"scratch" in lib.mg_last_error()
