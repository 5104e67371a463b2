"""The reference's on-disk formats through the library's C ABI (host
memory): Fortran unformatted sequential field, grid, time and scalar files,
as GRAVA writes them (solver/gravar_incomp.f90) and LEER_MALLA /
LEER_CAMPOS / LEER_TIME / LEER_MASCARA read them (solver/leer_incomp.f90),
so that the fields of a reference run can feed the device operator
(load_run).  Arrays are numpy, (ncomp, nz, ny, nx) or (nz, ny, nx)."""
import ctypes

import numpy as np

from . import _lib

HEADER_NONE, HEADER_DIMS, HEADER_COUNT_DIMS = 0, 1, 2


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _b(path):
    return str(path).encode()


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def field_dims(path):
    """(nd, nx, ny, nz) of a field or grid file's header."""
    nd, nx, ny, nz = (ctypes.c_int() for _ in range(4))
    _lib.call("emx_io_field_dims", _b(path), ctypes.byref(nd), ctypes.byref(nx),
              ctypes.byref(ny), ctypes.byref(nz))
    return nd.value, nx.value, ny.value, nz.value


def read_field(path):
    """u (3 comps), temp, pres of a field file (LEER_CAMPOS)."""
    _, nx, ny, nz = field_dims(path)
    u = np.empty((3, nz, ny, nx))
    temp, pres = np.empty((nz, ny, nx)), np.empty((nz, ny, nx))
    _lib.call("emx_io_read_field", _b(path), nx, ny, nz, _p(u), _p(temp), _p(pres))
    return {"u": u, "temp": temp, "pres": pres}


def write_field(path, u, temp, pres, nd=5):
    """A field file as GRAVA writes it."""
    u, temp, pres = _c64(u), _c64(temp), _c64(pres)
    nz, ny, nx = temp.shape
    _lib.call("emx_io_write_field", _b(path), nd, nx, ny, nz, _p(u), _p(temp), _p(pres))


def read_grid(path):
    """The coordinates x (3 comps) of a grid file (LEER_MALLA)."""
    _, nx, ny, nz = field_dims(path)
    x = np.empty((3, nz, ny, nx))
    _lib.call("emx_io_read_grid", _b(path), nx, ny, nz, _p(x))
    return x


def write_grid(path, x, nvar=3):
    x = _c64(x)
    nz, ny, nx = x.shape[1:]
    _lib.call("emx_io_write_grid", _b(path), nvar, nx, ny, nz, _p(x))


def read_time(path):
    dto, it = ctypes.c_double(), ctypes.c_int()
    _lib.call("emx_io_read_time", _b(path), ctypes.byref(dto), ctypes.byref(it))
    return dto.value, it.value


def write_time(path, dto, it):
    _lib.call("emx_io_write_time", _b(path), float(dto), int(it))


def read_scalar(path, shape, header=HEADER_DIMS):
    """One scalar field; shape = (nz, ny, nx).  header: HEADER_NONE
    (impellermask), HEADER_DIMS (sobject.field), HEADER_COUNT_DIMS
    (mez.mask)."""
    nz, ny, nx = shape
    a = np.empty((nz, ny, nx))
    _lib.call("emx_io_read_scalar", _b(path), header, nx, ny, nz, _p(a))
    return a


def write_scalar(path, a, header=HEADER_DIMS):
    a = _c64(a)
    nz, ny, nx = a.shape
    _lib.call("emx_io_write_scalar", _b(path), header, nx, ny, nz, _p(a))


def load_run(field, grid, mach, sobject=None, scheme="c6c", device="cuda"):
    """The state of a reference run on the device: the grid and its metrics
    (JACOBEANO), u, temp, pres and the conservative variables
    (varnc_to_varc), and sobject when its file is given.  Returns the
    operator and a dict of device tensors."""
    import torch

    from .api import Operator
    f = read_field(field)
    x = read_grid(grid)
    if x.shape[1:] != f["temp"].shape:
        raise ValueError(f"grid {x.shape[1:]} and field {f['temp'].shape} differ")
    nz, ny, nx = f["temp"].shape
    dev = torch.device(device)
    op = Operator(nx, ny, nz, scheme)
    t = {k: torch.from_numpy(v).to(dev) for k, v in f.items()}
    t["x"] = torch.from_numpy(x).to(dev)
    t["jbn"] = op.jacobeano(t["x"])
    t["pres0"] = torch.empty_like(t["pres"])
    t["um"] = op.varnc_to_varc(t["u"], t["pres"], t["temp"], mach, pres0=t["pres0"])
    if sobject is not None:
        t["sobject"] = torch.from_numpy(read_scalar(sobject, (nz, ny, nx))).to(dev)
    return op, t
