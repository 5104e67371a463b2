"""The oracle's RUNGEK glue (boundary values, conversions, velocity clip,
rebuild) against the reference Fortran's golden vectors and closed forms.

tests/golden/ref_glue_g9x8x7.npz holds frontera_z_val + frontera_y_val,
varc_to_varnc and varnc_to_varc of the flang-built reference
(tests/golden/make_golden.py).  frontera_x_val's source file also holds a
routine that needs Intel's ifport module, so it is not built: its faces are
pinned by the z/y reference values they copy and by idempotence.  The clip
and the rebuild are inline loops of solver/rungek_ico.f90 (no callable
routine): the rebuild is pinned by varnc_to_varc, which the reference writes
with the same statements, and the clip by its closed forms.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_glue_g9x8x7.npz")


@pytest.fixture(scope="module")
def z():
    return np.load(GOLD)


def test_conversions_match_reference(oracle, z):
    mach = float(z["mach"])
    u, pres, temp, conc = oracle.varc_to_varnc(z["um1"], mach)
    for name, got in (("u", u), ("pres", pres), ("temp", temp), ("conc", conc)):
        assert np.array_equal(got, z[name]), name
    um0, pres0 = oracle.varnc_to_varc(z["uin"], z["pin"], z["tin"], mach)
    assert np.array_equal(um0, z["um0"])
    assert np.array_equal(pres0, z["pres0"])


def test_boundary_values_match_reference(oracle, z):
    um1 = z["um1"]
    got = oracle.frontera_val(um1)
    ref_zy = z["frontera_zy"]
    # everything off the two x faces is the z and y boundary values
    assert np.array_equal(got[..., 1:-1], ref_zy[..., 1:-1])
    # the x faces copy density and scalar from the z/y-processed neighbour
    # column and zero the momentum; on the edges shared with the z and y
    # faces this holds only when z and y run first
    for face, nb in ((0, 1), (-1, -2)):
        for m in (0, 4):
            assert np.array_equal(got[m, :, :, face], ref_zy[m, :, :, nb])
        assert not got[1:4, :, :, face].any()
    # boundary values are a projection: applying them twice changes nothing
    assert np.array_equal(oracle.frontera_val(got), got)
    # the interior is untouched
    assert np.array_equal(got[:, 1:-2, 1:-1, 1:-1], um1[:, 1:-2, 1:-1, 1:-1])


def test_rebuild_is_varnc_to_varc(oracle, z):
    mach = float(z["mach"])
    um1 = np.array(z["um1"])
    got = oracle.rk_rebuild(um1, z["uin"], z["pin"], mach)
    assert np.array_equal(got[:4], z["um0"][:4])
    assert np.array_equal(got[4], um1[4])


def test_clip_closed_forms(oracle):
    rng = np.random.default_rng(5)
    nz, ny, nx = 6, 5, 7
    mach = 0.25
    um1 = rng.standard_normal((5, nz, ny, nx))
    um1[0] = 0.5 + rng.random((nz, ny, nx))
    um1[1:4] *= 8.0  # velocities up to ~|40|: many beyond the clip
    mask = (rng.random((nz, ny, nx)) > 0.3).astype(np.float64)
    out, u, pres, temp, conc = oracle.rk_clip(um1, mask, mach)
    raw = um1[1:4] / um1[0]
    inside = np.abs(raw) < 6.0
    assert inside.any() and (~inside).any()
    # |u| <= 6 up to the rounding of (6*rho)/rho, zero where masked
    assert np.all(np.abs(u) <= 6.0 * (1 + 2e-16))
    assert not u[:, mask == 0].any() and not out[1:4][:, mask == 0].any()
    keep = inside & (mask == 1)[None]
    assert np.allclose(u[keep], raw[keep], rtol=4e-16, atol=0)
    clipped = (~inside) & (mask == 1)[None]
    assert np.allclose(np.abs(u[clipped]), 6.0, rtol=4e-16, atol=0)
    # the conversions are the reference's (pinned above)
    ru, rp, rt, rc = oracle.varc_to_varnc(out, mach)
    assert np.array_equal(ru, u) and np.array_equal(rp, pres)
    assert np.array_equal(rt, temp) and np.array_equal(rc, conc)
    assert np.array_equal(out[0], um1[0]) and np.array_equal(out[4], um1[4])
