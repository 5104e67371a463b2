"""Write the golden vectors of the reference Fortran hot path.

Runs the reference's own Fortran (built from /root/reference by
oracle/build_ref.sh into oracle/_ref/) on seeded inputs and stores inputs
and outputs as .npz fixtures next to this script.  Run in the build
container only (the reference tree is not shipped):

    python tests/golden/make_golden.py          # every fixture
    python tests/golden/make_golden.py glue     # the RUNGEK glue fixture only
    python tests/golden/make_golden.py filter   # the compact filter fixture only
    python tests/golden/make_golden.py io       # the GRAVA files only
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from emix3d_amd import fields  # noqa: E402
from oracle.pyoracle import RefFortran  # noqa: E402

GRIDS = [
    # name, (nx, ny, nz), curvilinear, turbulent, seed
    ("g12x10x9_tanh", (12, 10, 9), True, True, 20261202),
    ("g17x13x11_unif", (17, 13, 11), False, False, 20261203),
]
LINE_N = {"c4": [4, 5, 6, 7, 8, 16, 33], "c6f": [5, 6, 7, 8, 16, 33]}


def np64(t):
    return np.ascontiguousarray(t.cpu().numpy(), dtype=np.float64)


def make_lines(scheme):
    ref = RefFortran(scheme)
    rng = np.random.default_rng(7)
    out = {}
    for n in LINE_N[scheme]:
        ref.init(n, n, n)
        f = rng.standard_normal(n)
        out[f"in_{n}"] = f
        for d in range(3):
            out[f"out_{n}_d{d}"] = ref.derivnper(f, d + 1)
    return out


def make_grid(scheme, name, shape, curv, turb, seed):
    ref = RefFortran(scheme)
    nx, ny, nz = shape
    flds = fields.stage_fields(nx, ny, nz, seed, curvilinear=curv, turbulent=turb)
    x, u, temp, pres, um = (np64(flds[k]) for k in ("x", "u", "temp", "pres", "um"))
    ref.init(nx, ny, nz)
    jbn = ref.jacobeano(x)
    dvel, dtemp = ref.deriv_vel(u, temp)
    sgd, dmins, dmax = ref.inisgdm(x)
    c4 = 1.0 / fields.REYNOLDS
    amut, vis, vis5 = ref.sgdm(u, dvel, jbn, sgd, c4)
    out = dict(x=x, u=u, temp=temp, pres=pres, um=um, jbn=jbn, dvel=dvel, dtemp=dtemp,
               sgd=sgd, dmins=dmins, deltamax=np.array(dmax), c4=np.array(c4),
               amut=amut, vis=vis, vis5=vis5)
    for neq in range(1, 6):
        flux = ref.flujos(neq, um, u, pres, vis, jbn, dvel)
        rs, rsv = ref.divergencia(neq, flux, jbn)
        out[f"flux{neq}"] = flux
        out[f"rs{neq}"] = rs
        out[f"rsv{neq}"] = rsv
    return out


def make_glue():
    """RUNGEK glue of the reference (frontera_z_val + frontera_y_val,
    varc_to_varnc, varnc_to_varc) on a seeded conservative state."""
    ref = RefFortran("c6f")
    nx, ny, nz = 9, 8, 7
    ref.init(nx, ny, nz)
    rng = np.random.default_rng(20261211)
    um1 = rng.standard_normal((5, nz, ny, nx))
    um1[0] = 1.0 + 0.2 * rng.random((nz, ny, nx))
    mach = 0.3
    out = dict(um1=um1, mach=np.array(mach))
    out["frontera_zy"] = ref.frontera_zy(um1)
    out["u"], out["pres"], out["temp"], out["conc"] = ref.varc_to_varnc(um1, mach)
    uin = rng.standard_normal((3, nz, ny, nx))
    pin = 1.0 + rng.random((nz, ny, nx))
    tin = rng.standard_normal((nz, ny, nx))
    out.update(uin=uin, pin=pin, tin=tin)
    out["um0"], out["pres0"] = ref.varnc_to_varc(uin, pin, tin, mach)
    return out


FILTER_CASES = [
    # name, (nx, ny, nz), itf values, cfilt, stored fields of um1
    ("g12x11x11", (12, 11, 11), (1, 2, 3, 4, 5, 6), 0.49, (0, 2)),
    ("g21x12x13", (21, 12, 13), (1,), 0.45, (1, 4)),
    ("g9x8x7", (9, 8, 7), (1, 4), 0.49, (0, 1, 2, 3, 4)),
]


def make_filter():
    """FILTRADO(itf) of the reference on a seeded conservative state over a
    stretched grid, for every direction order, two cfilt values and a grid
    whose lines are all boundary rows.  The five fields are filtered
    independently, so a case stores the fields listed (and, for its first
    itf, the u, pres, temp FILTRADO recomputes); jbn is stored as J, 1/J,
    the only components the filter reads."""
    ref = RefFortran("c6f")
    out = {}
    for name, shape, itfs, cfilt, keep in FILTER_CASES:
        nx, ny, nz = shape
        ref.init(nx, ny, nz)
        flds = fields.stage_fields(nx, ny, nz, 20261212, curvilinear=True, turbulent=False)
        jbn = ref.jacobeano(np64(flds["x"]))
        jbn[:9] = 0.0
        rng = np.random.default_rng(sum(shape))
        um1 = rng.standard_normal((5, nz, ny, nx))
        um1[0] = 1.0 + 0.2 * rng.random((nz, ny, nx))
        out[f"{name}_um1"] = um1
        out[f"{name}_jac"] = jbn[9:11]
        out[f"{name}_cfilt"] = np.array(cfilt)
        out[f"{name}_keep"] = np.array(keep)
        for itf in itfs:
            f, u, pres, temp = ref.filtrado(itf, um1, jbn, cfilt, 0.3)
            out[f"{name}_itf{itf}_um1"] = f[list(keep)]
            if itf == itfs[0]:
                out[f"{name}_itf{itf}_u"] = u
                out[f"{name}_itf{itf}_pres"] = pres
                out[f"{name}_itf{itf}_temp"] = temp
    return out


def make_io():
    """Files written by the reference's GRAVA (field_03.007, time_03.007,
    impellermask_03.007) for a seeded 7x6x5 state, copied next to this
    script under io/, with the inputs in ref_io.npz."""
    import shutil
    import tempfile
    ref = RefFortran("c6f")
    nx, ny, nz = 7, 6, 5
    rng = np.random.default_rng(20261213)
    u = rng.standard_normal((3, nz, ny, nx))
    temp = rng.standard_normal((nz, ny, nx))
    pres = 1.0 + rng.random((nz, ny, nx))
    mask = (rng.random((nz, ny, nx)) > 0.3).astype(np.float64)
    dto, it = 12.625, 4321
    out_dir = os.path.join(HERE, "io")
    os.makedirs(out_dir, exist_ok=True)
    cwd = os.getcwd()
    with tempfile.TemporaryDirectory() as tmp:
        os.chdir(tmp)
        try:
            ref.grava(u, temp, pres, mask, dto, it, 3, 7)
        finally:
            os.chdir(cwd)
        for name in ("field_03.007", "time_03.007", "impellermask_03.007"):
            shutil.copy(os.path.join(tmp, name), os.path.join(out_dir, name))
    np.savez_compressed(os.path.join(HERE, "ref_io.npz"), u=u, temp=temp, pres=pres, mask=mask,
                        dto=np.array(dto), it=np.array(it))


def main():
    if sys.argv[1:] == ["io"]:
        make_io()
        print("io files written to", HERE)
        return
    if sys.argv[1:] == ["glue"]:
        np.savez_compressed(os.path.join(HERE, "ref_glue_g9x8x7.npz"), **make_glue())
        print("glue vectors written to", HERE)
        return
    if sys.argv[1:] == ["filter"]:
        np.savez_compressed(os.path.join(HERE, "ref_filter.npz"), **make_filter())
        print("filter vectors written to", HERE)
        return
    np.savez_compressed(os.path.join(HERE, "ref_glue_g9x8x7.npz"), **make_glue())
    np.savez_compressed(os.path.join(HERE, "ref_filter.npz"), **make_filter())
    make_io()
    for scheme in ("c4", "c6f"):
        if not RefFortran.available(scheme):
            raise SystemExit("oracle/_ref is not built: run `make -C oracle ref`")
        np.savez_compressed(os.path.join(HERE, f"ref_{scheme}_lines.npz"), **make_lines(scheme))
        for name, shape, curv, turb, seed in GRIDS:
            np.savez_compressed(os.path.join(HERE, f"ref_{scheme}_{name}.npz"),
                                **make_grid(scheme, name, shape, curv, turb, seed))
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main()
