"""The reference's on-disk formats (host code of the library, no GPU).

tests/golden/io/ holds files written by the reference's own GRAVA
(solver/gravar_incomp.f90, flang build) for the state in ref_io.npz:
  * the library reads them back to the exact inputs;
  * the library's writers produce byte-identical files;
  * the reference's own LEER_CAMPOS / LEER_MALLA / LEER_TIME read files the
    library wrote (where oracle/_ref is built);
  * malformed files and mismatched headers are errors, not garbage.
"""
import os

import numpy as np
import pytest

from emix3d_amd import fortran_io as fio
from emix3d_amd._lib import EmxError
from oracle.pyoracle import RefFortran

GOLD = os.path.join(os.path.dirname(__file__), "golden")
IO = os.path.join(GOLD, "io")


@pytest.fixture(scope="module")
def z():
    return np.load(os.path.join(GOLD, "ref_io.npz"))


def test_reads_reference_files(z):
    assert fio.field_dims(os.path.join(IO, "field_03.007")) == (5, 7, 6, 5)
    f = fio.read_field(os.path.join(IO, "field_03.007"))
    for k in ("u", "temp", "pres"):
        assert np.array_equal(f[k], z[k]), k
    dto, it = fio.read_time(os.path.join(IO, "time_03.007"))
    assert dto == float(z["dto"]) and it == int(z["it"])
    m = fio.read_scalar(os.path.join(IO, "impellermask_03.007"), z["mask"].shape, fio.HEADER_NONE)
    assert np.array_equal(m, z["mask"])


def test_writes_reference_bytes(z, tmp_path):
    fio.write_field(tmp_path / "f", z["u"], z["temp"], z["pres"])
    fio.write_time(tmp_path / "t", float(z["dto"]), int(z["it"]))
    fio.write_scalar(tmp_path / "m", z["mask"], fio.HEADER_NONE)
    for mine, theirs in (("f", "field_03.007"), ("t", "time_03.007"),
                         ("m", "impellermask_03.007")):
        assert (tmp_path / mine).read_bytes() == open(os.path.join(IO, theirs), "rb").read()


@pytest.mark.skipif(not RefFortran.available("c6f"), reason="oracle/_ref not built")
def test_reference_reads_library_files(z, tmp_path):
    ref = RefFortran("c6f")
    nz, ny, nx = z["temp"].shape
    rng = np.random.default_rng(4)
    x = rng.random((3, nz, ny, nx))
    sob = rng.standard_normal((nz, ny, nx))
    fio.write_field(tmp_path / "myfield", z["u"], z["temp"], z["pres"])
    fio.write_grid(tmp_path / "mygrid", x)
    fio.write_scalar(tmp_path / "sobject.field", sob, fio.HEADER_DIMS)
    fio.write_time(tmp_path / "mytime", 3.5, 17)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        u, temp, pres, um0, s = ref.leer_campos("myfield", 0.3, (nx, ny, nz))
        xr = ref.leer_malla("mygrid", (nx, ny, nz))
        dto, it = ref.leer_time("mytime")
    finally:
        os.chdir(cwd)
    assert np.array_equal(u, z["u"]) and np.array_equal(temp, z["temp"])
    assert np.array_equal(pres, z["pres"]) and np.array_equal(s, sob)
    assert np.array_equal(xr, x) and (dto, it) == (3.5, 17)
    # LEER_CAMPOS ends with VARNC_TO_VARC: the library's conversion, pinned
    # bitwise by the glue tests, gives the same um0
    assert np.array_equal(um0[0], pres * 0.3 * 0.3)


def test_malformed_and_mismatched(z, tmp_path):
    good = os.path.join(IO, "field_03.007")
    data = open(good, "rb").read()
    (tmp_path / "trunc").write_bytes(data[:-10])
    with pytest.raises(EmxError) as e:
        fio.read_field(tmp_path / "trunc")
    assert e.value.status == -6
    bad = bytearray(data)
    bad[-4] ^= 0xFF  # trailing marker of the last record
    (tmp_path / "marker").write_bytes(bytes(bad))
    with pytest.raises(EmxError):
        fio.read_field(tmp_path / "marker")
    cont = bytearray(data)
    cont[32:36] = (-1680).to_bytes(4, "little", signed=True)  # a continued record
    (tmp_path / "cont").write_bytes(bytes(cont))
    with pytest.raises(EmxError):
        fio.read_field(tmp_path / "cont")
    fio.write_field(tmp_path / "nd4", z["u"], z["temp"], z["pres"], nd=4)
    with pytest.raises(EmxError) as e:
        fio.read_field(tmp_path / "nd4")
    assert e.value.status == -3
    with pytest.raises(EmxError) as e:
        fio.read_scalar(os.path.join(IO, "impellermask_03.007"), (5, 6, 8), fio.HEADER_NONE)
    assert e.value.status == -3
    with pytest.raises(EmxError) as e:
        fio.read_time(tmp_path / "missing")
    assert e.value.status == -6
