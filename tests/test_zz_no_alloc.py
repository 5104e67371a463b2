"""emx_dynsmag_noalloc: the dynamic model with its scratch in a static
device array (no allocation by the call), through the C ABI.

  * against emx_dynsmag (the same kernels on context scratch) on two calls
    in a row, the second on a smaller grid, so the second call reuses the
    array the first one filled: nu_t and vis bit for bit;
  * sizes beyond the array and null pointers are rejected before any
    launch, with the library's status codes.
"""
import ctypes

import numpy as np
import pytest
import torch

from emix3d_amd import Operator, fields
from emix3d_amd import _lib
from emix3d_amd.api import _ptr

pytestmark = pytest.mark.gpu

MAX_POINTS = 1 << 21


def noalloc():
    return _lib.lib().emx_dynsmag_noalloc


def state(shape, seed):
    nx, ny, nz = shape
    op = Operator(nx, ny, nz, "c6c")
    x = fields.grid_tanh(nx, ny, nz, device="cuda")
    u = fields.velocity_sines(x, seed).contiguous()
    jbn = op.jacobeano(x)
    sgd, _, _ = op.inisgdm(x)
    dvel, _ = op.deriv_vel(u)
    return op, u, dvel, jbn, sgd[0].contiguous()


def run_noalloc(u, dvel, jbn, delta, c4):
    nz, ny, nx = delta.shape
    nut = torch.full_like(delta, float("nan"))
    vis = torch.full_like(delta, float("nan"))
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = noalloc()(nx, ny, nz, _ptr(u), _ptr(dvel), _ptr(jbn), _ptr(delta), _ptr(nut), c4,
                   _ptr(vis), stream)
    _lib.check("emx_dynsmag_noalloc", rc)
    return nut, vis


def test_noalloc_matches_context_entry_two_calls():
    c4 = 1e-3
    for shape, seed in (((40, 36, 34), 31), ((21, 18, 16), 32)):
        op, u, dvel, jbn, delta = state(shape, seed)
        nut, vis = run_noalloc(u, dvel, jbn, delta, c4)
        rvis = torch.empty_like(delta)
        rnut = op.dynsmag(u, dvel, jbn, delta, c4=c4, vis=rvis)
        torch.cuda.synchronize()
        a, b = nut.cpu().numpy(), rnut.cpu().numpy()
        print(f"  {shape}: nu_t max {b.max():.3e}, positive {np.mean(b > 0):.2f}, "
              f"max |diff| {np.abs(a - b).max():.1e}", flush=True)
        assert (b > 0).mean() > 0.1
        assert np.array_equal(a, b), shape
        assert np.array_equal(vis.cpu().numpy(), rvis.cpu().numpy()), shape


def test_noalloc_rejects_before_launch():
    fn = noalloc()
    d = torch.zeros((4, 4, 4), dtype=torch.float64, device="cuda")
    big = (256, 128, 65)  # one plane more than the array holds
    assert big[0] * big[1] * big[2] > MAX_POINTS
    p = _ptr(d)
    rc = fn(*big, p, p, p, p, p, 0.0, None, None)
    assert rc == -3  # EMX_ERR_SIZE
    assert "exceed" in _lib.lib().emx_last_error().decode()
    rc = fn(4, 4, 4, p, None, p, p, p, 0.0, None, None)
    assert rc == -1  # EMX_ERR_ARG
    rc = fn(0, 4, 4, p, p, p, p, p, 0.0, None, None)
    assert rc == -3
