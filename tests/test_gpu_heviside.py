"""emx_heviside (the smoothed Heaviside of solver/advectLS.f90:29-59) on the
GPU against the oracle: outside the band bit-identical (0 and 1), inside
within 4.5e-16 absolute (2 ulp at 1; the device's double sine and the host
libm's may differ in the last bit); NaN phi keeps hevi; in place; argument
errors."""
import numpy as np
import pytest
import torch

from emix3d_amd import Operator
from emix3d_amd._lib import EmxError

pytestmark = pytest.mark.gpu
DMAX = 0.021


def phi_field(shape, seed):
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(*(np.linspace(0, 1, n) for n in (nz, ny, nx)), indexing="ij")
    phi = np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2) - 0.31
    phi += 0.1 * DMAX * rng.standard_normal(phi.shape)
    eps = 1.5 * DMAX
    phi.reshape(-1)[:5] = [eps, -eps, 0.0, np.nan, np.nextafter(eps, 1.0)]
    return phi


@pytest.mark.parametrize("shape", [(33, 17, 19), (128, 128, 128)])
def test_heviside_vs_oracle(oracle, shape):
    print(f"  heviside {shape}", flush=True)
    phi = phi_field(shape, 3)
    prev = np.full(phi.shape, 5.0)
    ref = oracle.heviside(phi, DMAX, hevi=prev)
    op = Operator(*shape, "c6c")
    out = torch.from_numpy(prev.copy()).cuda()
    op.heviside(torch.from_numpy(phi).cuda(), DMAX, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    band = np.abs(phi) <= 1.5 * DMAX
    assert band.sum() > 100
    np.testing.assert_array_equal(got[~band], ref[~band])  # 0, 1 and the NaN point's 5.0
    err = np.abs(got[band] - ref[band]).max()
    print(f"  in-band max abs diff {err:.2e} (tol 4.5e-16)", flush=True)
    assert err <= 4.5e-16


def test_heviside_in_place_and_errors(oracle):
    shape = (20, 18, 16)
    phi = phi_field(shape, 4)
    op = Operator(*shape, "c6c")
    t = torch.from_numpy(np.nan_to_num(phi, nan=0.0)).cuda()
    ref = oracle.heviside(np.nan_to_num(phi, nan=0.0), DMAX)
    op.heviside(t, DMAX, t)
    torch.cuda.synchronize()
    assert np.abs(t.cpu().numpy() - ref).max() <= 4.5e-16
    for bad in (0.0, -1.0, float("inf"), float("nan")):
        with pytest.raises(EmxError):
            op.heviside(t, bad)
