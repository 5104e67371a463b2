"""Level-set reinitialisation (cuReinitLsJB) on the GPU against the oracle,
which test_oracle_reinit.py pins to the mathematics; 1e-12 normwise
relative.  At 128^3: the plane distance is a fixed point 20 cells from the
faces."""
import numpy as np
import pytest
import torch

from conftest import normwise_rel
from emix3d_amd import Operator, fields

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


@pytest.mark.parametrize("shape", [(12, 10, 9), (30, 20, 25), (48, 40, 36), (64, 17, 33)])
def test_reinit_vs_oracle(oracle, shape):
    nx, ny, nz = shape
    f = fields.stage_fields(nx, ny, nz, 61, curvilinear=True, turbulent=False)
    x = np.ascontiguousarray(f["x"].numpy())
    jbn = oracle.jacobeano(2, x)
    r = np.sqrt(((x - 0.5) ** 2).sum(0))
    d = r - 0.3
    phi = np.ascontiguousarray(d * (1.0 + 2.0 * d))  # a distorted distance
    _, dmins, _ = oracle.inisgdm(x)
    dxyz = (1.0 / (nx - 1), 1.0 / (ny - 1), 1.0 / (nz - 1))
    deltamin = float(dmins.min())
    ref = oracle.ls_reinit(phi, jbn, dmins, dxyz, deltamin)
    op = Operator(*shape, "c6c")
    p = dev(phi)
    op.reinit_level_set(p, dev(jbn), dev(dmins), dxyz, deltamin)
    torch.cuda.synchronize()
    err = normwise_rel(p.cpu().numpy(), ref)
    print(f"  reinit {shape}: {err:.1e}", flush=True)
    assert err <= 1e-12
    assert normwise_rel(ref, phi) > 1e-6  # the iterations moved phi


def test_reinit_plane_fixed_point_128():
    n = 128
    op = Operator(n, n, n, "c6c")
    x = fields.grid_uniform(n, n, n, device=torch.device("cuda"))
    jbn = op.jacobeano(x)
    phi = (0.6 * (x[0] - 0.45) + 0.8 * (x[2] - 0.55)).contiguous()
    p = phi.clone()
    h = 1.0 / (n - 1)
    op.reinit_level_set(p, jbn, torch.full_like(phi, h), (h, h, h), h)
    # the faces' mirror kinks reach about 16 cells in at this size (the
    # oracle gives 4e-9 at 11 cells, 6e-13 at 16, 5e-16 at 20)
    core = (slice(20, -20),) * 3
    err = float((p - phi)[core].abs().max())
    print(f"  plane 128^3 core change: {err:.1e}", flush=True)
    assert err < 1e-13


def test_reinit_argument_checks():
    from emix3d_amd._lib import EmxError
    op = Operator(8, 8, 8, "c6c")
    a = torch.zeros((8, 8, 8), dtype=torch.float64, device="cuda")
    j = torch.zeros((11, 8, 8, 8), dtype=torch.float64, device="cuda")
    with pytest.raises(EmxError):
        op.reinit_level_set(a, j, a, (0.1, 0.1, 0.0), 0.1)
