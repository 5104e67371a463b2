"""The oracle's smoothed Heaviside (orc_heviside, solver/advectLS.f90:29-59)
pinned on the CPU:

  * against the reference's own heviside, compiled unmodified from
    solver/advectLS.f90 into the drop-in program (oracle/build_ref.sh), run
    in its CPU-only heviside mode: bit-identical;
  * against what the formula fixes: H(0) = 1/2, H(-phi) = 1 - H(phi),
    H(eps/2) = 3/4 + 1/(2 pi) with eps = 1.5 deltamax, H = 0 / 1 beyond the
    band, and slope 1/eps at 0 and 0 at the band's edges (a C1 blend);
  * planted mistakes (band 1.0 instead of 1.5, the sine term's sign or its
    2 pi) fail those checks;
  * a NaN phi leaves hevi as it was, as the reference's three ifs do.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

EXE = os.path.join(ROOT, "oracle", "_ref", "emx_dropin")
DMAX = 0.037


def closed_form_errors(H, band_used_by_checks=1.5):
    eps = band_used_by_checks * DMAX
    errs = {}
    errs["zero"] = abs(H(np.array([0.0]))[0] - 0.5)
    x = np.linspace(-2 * eps, 2 * eps, 401)
    errs["odd"] = np.abs(H(x) + H(-x) - 1.0).max()
    errs["half band"] = abs(H(np.array([eps / 2]))[0] - (0.75 + 1.0 / (2 * np.pi)))
    out = np.array([-1.0, -eps * 1.0001, eps * 1.0001, 3.0])
    errs["outside"] = np.abs(H(out) - np.array([0.0, 0.0, 1.0, 1.0])).max()
    h = 1e-6 * eps
    slope0 = (H(np.array([h]))[0] - H(np.array([-h]))[0]) / (2 * h)
    errs["slope at 0"] = abs(slope0 * eps - 1.0)
    edge = (H(np.array([eps]))[0] - H(np.array([eps - 1e-4 * eps]))[0]) / (1e-4 * eps)
    errs["slope at edge"] = abs(edge * eps)
    return errs


def test_closed_forms(oracle):
    errs = closed_form_errors(lambda p: oracle.heviside(p, DMAX))
    print(errs)
    assert errs["zero"] == 0.0
    assert errs["odd"] <= 2e-16
    assert errs["half band"] <= 2e-16
    assert errs["outside"] == 0.0
    assert errs["slope at 0"] <= 1e-8
    assert errs["slope at edge"] <= 1e-3


def test_planted_mistakes_fail(oracle):
    """The closed-form checks reject a wrong band and a wrong sine term
    (the planted variants are restated here, on the CPU)."""
    assert closed_form_errors(lambda p: oracle.heviside(p, DMAX, band=1.0))["half band"] > 1e-3

    def variant(p, sign=1.0, twopi=2 * np.pi):
        eps = 1.5 * DMAX
        return np.clip(0.5 + p / (2 * eps) + sign * np.sin(np.pi * p / eps) / twopi, 0.0, 1.0) * \
            (np.abs(p) <= eps) + (p > eps)

    assert abs(closed_form_errors(lambda p: variant(p))["half band"]) <= 2e-16  # the restatement
    assert closed_form_errors(lambda p: variant(p, sign=-1.0))["half band"] > 1e-3
    assert closed_form_errors(lambda p: variant(p, twopi=np.pi))["slope at 0"] > 1e-3


def test_nan_keeps_previous_value(oracle):
    phi = np.array([np.nan, 0.0, np.nan])
    got = oracle.heviside(phi, DMAX, hevi=np.array([7.0, 7.0, -3.0]))
    assert got[0] == 7.0 and got[1] == 0.5 and got[2] == -3.0


@pytest.mark.skipif(not os.path.exists(EXE), reason="drop-in program not built (oracle/_ref)")
def test_matches_reference_fortran(oracle, tmp_path):
    nx, ny, nz = 17, 13, 11
    rng = np.random.default_rng(5)
    eps = 1.5 * DMAX
    # a signed distance to a sphere, spanning the band, plus exact band
    # edges, zeros and NaNs
    z, y, x = np.meshgrid(*(np.linspace(0, 1, n) for n in (nz, ny, nx)), indexing="ij")
    phi = np.sqrt((x - 0.5) ** 2 + (y - 0.45) ** 2 + (z - 0.55) ** 2) - 0.3
    phi = phi + 0.2 * eps * rng.standard_normal(phi.shape)
    flat = phi.reshape(-1)
    flat[:6] = [eps, -eps, 0.0, np.nextafter(eps, 1.0), -np.nextafter(eps, 1.0), np.nan]
    hevi0 = rng.standard_normal(phi.shape)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as fh:
        np.array([nx, ny, nz], dtype=np.int32).tofile(fh)
        np.array([DMAX], dtype=np.float64).tofile(fh)
        phi.astype(np.float64).tofile(fh)
        hevi0.astype(np.float64).tofile(fh)
    r = subprocess.run([EXE, str(fin), str(fout), "heviside"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr[-500:]
    ref = np.fromfile(fout, dtype=np.float64).reshape(phi.shape)
    got = oracle.heviside(phi, DMAX, hevi=hevi0)
    inband = (np.abs(phi) <= eps).sum()
    assert inband > 200, inband
    np.testing.assert_array_equal(got, ref)
