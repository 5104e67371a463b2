"""ctypes access to the CPU oracle and to the reference Fortran build.

TEST INFRASTRUCTURE ONLY -- imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, never by the product package.

Array convention (same memory as the reference's Fortran column-major
(nx,ny,nz[,c]) arrays): numpy C-contiguous float64 arrays of shape
(nz, ny, nx) for one field and (c, nz, ny, nx) for c components.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
C4, C6F, C6C = 0, 1, 2
SCHEMES = {"c4": C4, "c6f": C6F, "c6c": C6C}

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_L = ctypes.c_long


def _ptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "float64 C-contiguous"
    return a.ctypes.data_as(_P)


def build():
    """Compile the C restatement (and the reference, where present)."""
    subprocess.run(["make", "-s", "-C", HERE, "libemx_oracle.so"], check=True)
    subprocess.run([os.path.join(HERE, "build_ref.sh")], check=True,
                   stdout=subprocess.DEVNULL)


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "libemx_oracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        sig = {
            "orc_coeffs": [_I, _I, _P, _P, _P],
            "orc_deriv_line": [_I, _I, _P, _L, _D, _P, _L],
            "orc_deriv3d": [_I, _I, _I, _I, _I, _P, _D, _P],
            "orc_jacobeano": [_I, _I, _I, _I, _P, _P],
            "orc_deriv_vel": [_I, _I, _I, _I, _P, _P, _P, _P, _D, _D, _D],
            "orc_flujos": [_I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P],
            "orc_divergencia": [_I, _I, _I, _I, _P, _P, _D, _D, _D, _P, _P],
            "orc_inisgdm": [_I, _I, _I, _P, _P, _P, _P],
            "orc_sgdm": [_I, _I, _I, _P, _P, _P, _P, _D, _P, _P, _P],
            "orc_ls_extrapolate": [_I, _I, _I, _P, _P, _P, _D, _D, _D, _D],
            "orc_ls_advect_weno5_padded": [_I, _I, _I, _P, _P, _P, _D, _D, _D, _D],
            "orc_ls_advect_weno5": [_I, _I, _I, _P, _P, _P, _D, _D, _D, _D],
            "orc_pad_mirror": [_I, _I, _I, _I, _P, _P],
            "orc_weno5_derivs": [_I, _I, _I, _P, _P, _D, _D, _D, _P, _P],
            "orc_dynsmag": [_I, _I, _I, _P, _P, _P, _P],
            "orc_dynsmag_cfg": [_I, _I, _I, _P, _P, _P, _P, _D, _D, _D],
            "orc_rk_update": [_I, _L, _P, _P, _P, _P, _P, _D],
            "orc_frontera_val": [_I, _I, _I, _P],
            "orc_ls_ghost_mirror": [_I, _I, _I, _I, _P],
            "orc_ls_reinit": [_I, _I, _I, _P, _P, _P, _D, _D, _D, _D],
            "orc_heviside": [ctypes.c_long, _P, _D, _D, _P],
            "orc_filter_coeffs": [_I, _I, _D, _P, _P, _P],
            "orc_filtronper": [_I, _P, _P, _P, _P, _D],
            "orc_filtrado": [_I, _I, _I, _I, _P, _P, _D, _D, _P, _P, _P],
            "orc_varc_to_varnc": [_L, _P, _D, _P, _P, _P, _P],
            "orc_varnc_to_varc": [_L, _P, _P, _P, _D, _P, _P],
            "orc_rk_clip": [_L, _P, _P, _D, _P, _P, _P, _P],
            "orc_rk_rebuild": [_L, _P, _P, _P, _D],
            "orc_mlp_forward": [_P, _P, _P, _L, _P],
            "orc_mlp_grad": [_P, _P, _P, _P, _L, _P],
            "orc_mlp_features": [_I, _I, _I, _P, _P, _P, _P],
            "orc_deriv3d_adj": [_I, _I, _I, _I, _I, _P, _D, _P],
            "orc_deriv_vel_adj": [_I, _I, _I, _I, _P, _P, _P, _P, _P],
            "orc_divergencia_adj": [_I, _I, _I, _I, _P, _P, _P, _D, _D, _D, _P],
            "orc_flujos_adj": [_I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
            "orc_flux_div_all_adj": [_I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _D, _D, _D,
                                     _P, _P, _P, _P, _P],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = None if name == "orc_pad_mirror" else _I
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        raise ValueError(f"oracle {what} failed with status {rc}")


def dims(a):
    nz, ny, nx = a.shape[-3:]
    return nx, ny, nz


def coeffs(scheme, n):
    a, b, c = (np.zeros(n) for _ in range(3))
    _check(lib().orc_coeffs(scheme, n, _ptr(a), _ptr(b), _ptr(c)), "coeffs")
    return a, b, c


def deriv_line(scheme, f, inv_h):
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.empty_like(f)
    _check(lib().orc_deriv_line(scheme, f.size, _ptr(f), 1, inv_h, _ptr(out), 1), "deriv_line")
    return out


def deriv3d(scheme, d, f, inv_h):
    nx, ny, nz = dims(f)
    out = np.empty_like(f)
    _check(lib().orc_deriv3d(scheme, d, nx, ny, nz, _ptr(f), inv_h, _ptr(out)), "deriv3d")
    return out


def deriv3d_adj(scheme, d, g, inv_h):
    """Transpose of deriv3d along d."""
    nx, ny, nz = dims(g)
    g = np.ascontiguousarray(g, dtype=np.float64)
    out = np.empty_like(g)
    _check(lib().orc_deriv3d_adj(scheme, d, nx, ny, nz, _ptr(g), inv_h, _ptr(out)),
           "deriv3d_adj")
    return out


def deriv_vel_adj(scheme, dvel_bar, dtemp_bar, ih):
    """Transpose of deriv_vel: (u_bar[3], temp_bar or None)."""
    nx, ny, nz = dims(dvel_bar)
    u_bar = np.empty((3, nz, ny, nx))
    temp_bar = None if dtemp_bar is None else np.empty((nz, ny, nx))
    ihs = (ctypes.c_double * 3)(*ih)
    _check(lib().orc_deriv_vel_adj(scheme, nx, ny, nz, _ptr(dvel_bar),
                                   None if dtemp_bar is None else _ptr(dtemp_bar),
                                   ctypes.cast(ihs, _P), _ptr(u_bar),
                                   None if temp_bar is None else _ptr(temp_bar)),
           "deriv_vel_adj")
    return u_bar, temp_bar


def divergencia_adj(scheme, rs_bar, rsv_bar, jbn, ih):
    """Transpose of divergencia with respect to the 6 flux fields."""
    nx, ny, nz = dims(rs_bar)
    fb = np.empty((6, nz, ny, nx))
    _check(lib().orc_divergencia_adj(scheme, nx, ny, nz, _ptr(rs_bar), _ptr(rsv_bar),
                                     _ptr(jbn), ih[0], ih[1], ih[2], _ptr(fb)),
           "divergencia_adj")
    return fb


def flujos_adj(neq, um, u, vis, jbn, dvel, flux_bar, bars):
    """Adds the transpose of flujos(neq)'s linearisation to bars (dict of
    um_bar, u_bar, pres_bar, vis_bar, dvel_bar)."""
    nx, ny, nz = dims(u)
    _check(lib().orc_flujos_adj(neq, nx, ny, nz, _ptr(um), _ptr(u), _ptr(vis), _ptr(jbn),
                                _ptr(dvel), _ptr(flux_bar), _ptr(bars["um_bar"]),
                                _ptr(bars["u_bar"]), _ptr(bars["pres_bar"]),
                                _ptr(bars["vis_bar"]), _ptr(bars["dvel_bar"])), "flujos_adj")
    return bars


def flux_div_all_adj(scheme, rs_bar, rsv_bar, um, u, vis, jbn, dvel, ih):
    """Transpose of the five-equation flux divergence's linearisation."""
    nx, ny, nz = dims(u)
    bars = {"um_bar": np.empty((5, nz, ny, nx)), "u_bar": np.empty((3, nz, ny, nx)),
            "pres_bar": np.empty((nz, ny, nx)), "vis_bar": np.empty((nz, ny, nx)),
            "dvel_bar": np.empty((9, nz, ny, nx))}
    _check(lib().orc_flux_div_all_adj(scheme, nx, ny, nz, _ptr(rs_bar), _ptr(rsv_bar),
                                      _ptr(um), _ptr(u), _ptr(vis), _ptr(jbn), _ptr(dvel),
                                      ih[0], ih[1], ih[2], _ptr(bars["um_bar"]),
                                      _ptr(bars["u_bar"]), _ptr(bars["pres_bar"]),
                                      _ptr(bars["vis_bar"]), _ptr(bars["dvel_bar"])),
           "flux_div_all_adj")
    return bars


def rk_update(sub, um0, um, rs, rsv, dt):
    """SSP-RK3 update of sub-step sub (um may be None for sub-step 1)."""
    um0 = np.ascontiguousarray(um0, dtype=np.float64)
    out = np.empty_like(um0)
    umc = None if um is None else np.ascontiguousarray(um, dtype=np.float64)
    _check(lib().orc_rk_update(sub, um0.size, _ptr(out), _ptr(um0),
                               None if umc is None else _ptr(umc),
                               _ptr(np.ascontiguousarray(rs, dtype=np.float64)),
                               _ptr(np.ascontiguousarray(rsv, dtype=np.float64)), dt), "rk_update")
    return out


def frontera_val(um1):
    """Boundary values z, y, x on a copy of um1 (5,nz,ny,nx)."""
    out = np.array(um1, dtype=np.float64, copy=True)
    nz, ny, nx = out.shape[1:]
    _check(lib().orc_frontera_val(nx, ny, nz, _ptr(out)), "frontera_val")
    return out


def varc_to_varnc(um1, mach):
    um1 = np.ascontiguousarray(um1, dtype=np.float64)
    n = um1[0].size
    u = np.empty((3,) + um1.shape[1:])
    pres, temp, conc = (np.empty(um1.shape[1:]) for _ in range(3))
    _check(lib().orc_varc_to_varnc(n, _ptr(um1), mach, _ptr(u), _ptr(pres), _ptr(temp),
                                   _ptr(conc)), "varc_to_varnc")
    return u, pres, temp, conc


def varnc_to_varc(u, pres, temp, mach):
    u = np.ascontiguousarray(u, dtype=np.float64)
    n = u[0].size
    um0 = np.empty((5,) + u.shape[1:])
    pres0 = np.empty(u.shape[1:])
    _check(lib().orc_varnc_to_varc(n, _ptr(u), _ptr(np.ascontiguousarray(pres)),
                                   _ptr(np.ascontiguousarray(temp)), mach, _ptr(um0),
                                   _ptr(pres0)), "varnc_to_varc")
    return um0, pres0


def rk_clip(um1, mask, mach):
    """Returns (um1 after the clip, u, pres, temp, conc)."""
    um1 = np.array(um1, dtype=np.float64, copy=True)
    n = um1[0].size
    u = np.empty((3,) + um1.shape[1:])
    pres, temp, conc = (np.empty(um1.shape[1:]) for _ in range(3))
    _check(lib().orc_rk_clip(n, _ptr(um1), _ptr(np.ascontiguousarray(mask, dtype=np.float64)),
                             mach, _ptr(u), _ptr(pres), _ptr(temp), _ptr(conc)), "rk_clip")
    return um1, u, pres, temp, conc


def rk_rebuild(um1, u, pres, mach):
    um1 = np.array(um1, dtype=np.float64, copy=True)
    n = um1[0].size
    _check(lib().orc_rk_rebuild(n, _ptr(um1), _ptr(np.ascontiguousarray(u)),
                                _ptr(np.ascontiguousarray(pres)), mach), "rk_rebuild")
    return um1


def filter_coeffs(periodic, n, cfilt):
    a, b, c = np.empty(n), np.empty(n), np.empty(n)
    _check(lib().orc_filter_coeffs(int(periodic), n, cfilt, _ptr(a), _ptr(b), _ptr(c)),
           "filter_coeffs")
    return a, b, c


def filtronper(line, periodic, cfilt):
    """One line through filtronper with INIFILTRO's arrays."""
    u = np.array(line, dtype=np.float64, copy=True)
    a, b, c = filter_coeffs(periodic, u.size, cfilt)
    _check(lib().orc_filtronper(u.size, _ptr(u), _ptr(a), _ptr(b), _ptr(c), cfilt), "filtronper")
    return u


def filtrado(itf, um1, jbn, cfilt, mach):
    """FILTRADO(itf) on a copy of um1; returns um1, u, pres, temp."""
    out = np.array(um1, dtype=np.float64, copy=True)
    nz, ny, nx = out.shape[1:]
    u = np.empty((3,) + out.shape[1:])
    pres, temp = np.empty(out.shape[1:]), np.empty(out.shape[1:])
    _check(lib().orc_filtrado(itf, nx, ny, nz, _ptr(out), _ptr(np.ascontiguousarray(jbn)), cfilt,
                              mach, _ptr(u), _ptr(pres), _ptr(temp)), "filtrado")
    return out, u, pres, temp


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_P)


def mlp_forward(params, norm, feat):
    """FP64 network output for feature rows feat [n][10] (FP32 parameters)."""
    (p, pp), (q, pq) = _f32(params), _f32(norm)
    feat = np.ascontiguousarray(feat, dtype=np.float64)
    nu = np.empty(feat.shape[0])
    _check(lib().orc_mlp_forward(pp, pq, _ptr(feat), feat.shape[0], _ptr(nu)), "mlp_forward")
    return nu


def mlp_grad(params, norm, feat, dnu):
    """sum_p dnu_p * d nu_p / d params (FP64, 9089 values)."""
    (p, pp), (q, pq) = _f32(params), _f32(norm)
    feat = np.ascontiguousarray(feat, dtype=np.float64)
    dnu = np.ascontiguousarray(dnu, dtype=np.float64)
    g = np.empty(9089)
    _check(lib().orc_mlp_grad(pp, pq, _ptr(feat), _ptr(dnu), feat.shape[0], _ptr(g)), "mlp_grad")
    return g


def mlp_features(dvel, jbn, delta):
    """[N][10] features: physical gradient du_a/dx_b (3a+b), delta."""
    nx, ny, nz = dims(delta)
    feat = np.empty((nx * ny * nz, 10))
    _check(lib().orc_mlp_features(nx, ny, nz, _ptr(dvel), _ptr(jbn), _ptr(delta), _ptr(feat)),
           "mlp_features")
    return feat


def jacobeano(scheme, x):
    nx, ny, nz = dims(x)
    jbn = np.zeros((11, nz, ny, nx))
    _check(lib().orc_jacobeano(scheme, nx, ny, nz, _ptr(x), _ptr(jbn)), "jacobeano")
    return jbn


def deriv_vel(scheme, u, temp, ih):
    nx, ny, nz = dims(u)
    dvel = np.empty((9, nz, ny, nx))
    dtemp = np.empty((3, nz, ny, nx))
    _check(lib().orc_deriv_vel(scheme, nx, ny, nz, _ptr(u), _ptr(temp), _ptr(dvel),
                               _ptr(dtemp), ih[0], ih[1], ih[2]), "deriv_vel")
    return dvel, dtemp


def flujos(neq, um, u, pres, vis, jbn, dvel):
    nx, ny, nz = dims(u)
    flux = np.empty((6, nz, ny, nx))
    _check(lib().orc_flujos(neq, nx, ny, nz, _ptr(um), _ptr(u), _ptr(pres), _ptr(vis),
                            _ptr(jbn), _ptr(dvel), _ptr(flux)), "flujos")
    return flux


def divergencia(scheme, flux, jbn, ih):
    nx, ny, nz = dims(flux)
    rs = np.empty((nz, ny, nx))
    rsv = np.empty((nz, ny, nx))
    _check(lib().orc_divergencia(scheme, nx, ny, nz, _ptr(flux), _ptr(jbn), ih[0], ih[1],
                                 ih[2], _ptr(rs), _ptr(rsv)), "divergencia")
    return rs, rsv


def inisgdm(x):
    nx, ny, nz = dims(x)
    sgd = np.empty((7, nz, ny, nx))
    dmins = np.empty((nz, ny, nx))
    dmax = ctypes.c_double(0.0)
    _check(lib().orc_inisgdm(nx, ny, nz, _ptr(x), _ptr(sgd), _ptr(dmins),
                             ctypes.byref(dmax)), "inisgdm")
    return sgd, dmins, dmax.value


def sgdm(u, dvel, jbn, sgd, c4):
    nx, ny, nz = dims(u)
    amut, vis, vis5 = (np.empty((nz, ny, nx)) for _ in range(3))
    _check(lib().orc_sgdm(nx, ny, nz, _ptr(u), _ptr(dvel), _ptr(jbn), _ptr(sgd), c4,
                          _ptr(amut), _ptr(vis), _ptr(vis5)), "sgdm")
    return amut, vis, vis5


def ls_extrapolate(pres, jbn, sobject, dxyz, deltamin):
    nx, ny, nz = dims(pres)
    out = np.array(pres, dtype=np.float64, copy=True)
    _check(lib().orc_ls_extrapolate(nx, ny, nz, _ptr(out), _ptr(jbn), _ptr(sobject),
                                    dxyz[0], dxyz[1], dxyz[2], deltamin), "ls_extrapolate")
    return out


def ls_ghost_mirror(padded, g=5):
    """Ghost layers of a padded array (nzG, nyG, nxG) refreshed from its
    interior (Z, then Y, then X mirror), on a copy."""
    out = np.array(padded, dtype=np.float64, copy=True)
    nzg, nyg, nxg = out.shape
    _check(lib().orc_ls_ghost_mirror(nxg, nyg, nzg, g, _ptr(out)), "ls_ghost_mirror")
    return out


def ls_reinit(phi, jbn, dmin, dxyz, deltamin):
    nx, ny, nz = dims(phi)
    out = np.array(phi, dtype=np.float64, copy=True)
    _check(lib().orc_ls_reinit(nx, ny, nz, _ptr(out), _ptr(np.ascontiguousarray(jbn)),
                               _ptr(np.ascontiguousarray(dmin)), dxyz[0], dxyz[1], dxyz[2],
                               deltamin), "ls_reinit")
    return out


def heviside(phi, deltamax, band=1.5, hevi=None):
    """Smoothed Heaviside; band overrides the 1.5 of eps = 1.5 deltamax to
    plant a mistake in tests; hevi (default zeros) holds the values a NaN
    phi leaves unchanged."""
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    out = np.zeros_like(phi) if hevi is None else np.array(hevi, dtype=np.float64, copy=True)
    _check(lib().orc_heviside(phi.size, _ptr(phi), float(deltamax), float(band), _ptr(out)),
           "heviside")
    return out


def ls_advect_weno5(phi, vel, jbn, dxyz, dt):
    nx, ny, nz = dims(phi)
    out = np.array(phi, dtype=np.float64, copy=True)
    _check(lib().orc_ls_advect_weno5(nx, ny, nz, _ptr(out), _ptr(vel), _ptr(jbn),
                                     dxyz[0], dxyz[1], dxyz[2], dt), "ls_advect_weno5")
    return out


def ls_advect_weno5_padded(phi, vel, jbn, dxyz, dt):
    nx, ny, nz = dims(phi)
    out = np.array(phi, dtype=np.float64, copy=True)
    _check(lib().orc_ls_advect_weno5_padded(nx, ny, nz, _ptr(out), _ptr(vel), _ptr(jbn),
                                            dxyz[0], dxyz[1], dxyz[2], dt), "weno5_padded")
    return out


def weno5_derivs(phi, jbn, dxyz):
    """Left/right-biased WENO5 derivatives of phi on a padded grid."""
    nx, ny, nz = dims(phi)
    left = np.empty((3, nz, ny, nx))
    right = np.empty((3, nz, ny, nx))
    _check(lib().orc_weno5_derivs(nx, ny, nz, _ptr(phi), _ptr(jbn), dxyz[0], dxyz[1], dxyz[2],
                                  _ptr(left), _ptr(right)), "weno5_derivs")
    return left, right


def pad_mirror(a, g=5):
    nx, ny, nz = dims(a)
    out = np.empty((nz + 2 * g, ny + 2 * g, nx + 2 * g))
    lib().orc_pad_mirror(nx, ny, nz, g, _ptr(np.ascontiguousarray(a)), _ptr(out))
    return out


def dynsmag(u, gphys, delta, ratio2=None, ws=None, offw=None):
    """Dynamic Smagorinsky nu_t; ratio2, ws, offw override the model's
    constants (4, 1/4, 2) to plant mistakes in tests."""
    nx, ny, nz = dims(u)
    nut = np.empty((nz, ny, nx))
    u, gphys, delta = (np.ascontiguousarray(a, dtype=np.float64) for a in (u, gphys, delta))
    if ratio2 is None and ws is None and offw is None:
        rc = lib().orc_dynsmag(nx, ny, nz, _ptr(u), _ptr(gphys), _ptr(delta), _ptr(nut))
    else:
        rc = lib().orc_dynsmag_cfg(nx, ny, nz, _ptr(u), _ptr(gphys), _ptr(delta), _ptr(nut),
                                   4.0 if ratio2 is None else ratio2,
                                   0.25 if ws is None else ws, 2.0 if offw is None else offw)
    _check(rc, "dynsmag")
    return nut


# --------------------------------------------------------------------------
# The reference Fortran, built by build_ref.sh into _ref/.

class RefFortran:
    """The reference's own Fortran hot path (flang build), one scheme per
    library: 'c4' = solver/derivada.f90, 'c6f' = derivada_incomp.f90.
    The module state is global, so one instance per library."""

    _cache = {}

    def __new__(cls, scheme):
        if scheme not in cls._cache:
            obj = super().__new__(cls)
            path = os.path.join(HERE, "_ref", f"libemx_ref_{scheme}.so")
            if not os.path.exists(path):
                raise FileNotFoundError(path)
            obj.L = ctypes.CDLL(path)
            obj.scheme = scheme
            obj.shape = None
            cls._cache[scheme] = obj
        return cls._cache[scheme]

    @staticmethod
    def available(scheme="c4"):
        return os.path.exists(os.path.join(HERE, "_ref", f"libemx_ref_{scheme}.so"))

    def init(self, nx, ny, nz):
        if self.shape != (nx, ny, nz):
            self.L.ref_init(_I(nx), _I(ny), _I(nz))
            self.shape = (nx, ny, nz)

    def derivnper(self, line, idir):
        line = np.array(line, dtype=np.float64, copy=True)
        self.L.ref_derivnper(_I(line.size), _ptr(line), _I(idir))
        return line

    def frontera_zy(self, um1):
        """frontera_z_val then frontera_y_val on a copy of um1 (5,nz,ny,nx)."""
        self.init(*dims(um1))
        out = np.array(um1, dtype=np.float64, copy=True)
        self.L.ref_frontera_zy(_ptr(out))
        return out

    def varc_to_varnc(self, um1, mach):
        self.init(*dims(um1))
        u = np.empty((3,) + um1.shape[1:])
        pres, temp, conc = (np.empty(um1.shape[1:]) for _ in range(3))
        self.L.ref_varc_to_varnc(_ptr(np.ascontiguousarray(um1)), ctypes.c_double(mach),
                                 _ptr(u), _ptr(pres), _ptr(temp), _ptr(conc))
        return u, pres, temp, conc

    def varnc_to_varc(self, u, pres, temp, mach):
        self.init(*dims(u))
        um0 = np.empty((5,) + u.shape[1:])
        pres0 = np.empty(u.shape[1:])
        self.L.ref_varnc_to_varc(_ptr(np.ascontiguousarray(u)), _ptr(np.ascontiguousarray(pres)),
                                 _ptr(np.ascontiguousarray(temp)), ctypes.c_double(mach),
                                 _ptr(um0), _ptr(pres0))
        return um0, pres0

    def filtrado(self, itf, um1, jbn, cfilt, mach):
        """INIFILTRO(cfilt) then FILTRADO(itf); returns um1, u, pres, temp."""
        self.init(*dims(um1))
        out = np.array(um1, dtype=np.float64, copy=True)
        u = np.empty((3,) + um1.shape[1:])
        temp, pres = np.empty(um1.shape[1:]), np.empty(um1.shape[1:])
        self.L.ref_filtrado(_I(itf), _ptr(out), _ptr(np.ascontiguousarray(jbn)),
                            ctypes.c_double(cfilt), ctypes.c_double(mach), _ptr(u), _ptr(temp),
                            _ptr(pres))
        return out, u, pres, temp

    def grava(self, u, temp, pres, mask, dto, it, serie, sample):
        """GRAVA in the current working directory: field_SS.NNN,
        time_SS.NNN, impellermask_SS.NNN."""
        self.init(*dims(u))
        self.L.ref_grava(_ptr(np.ascontiguousarray(u)), _ptr(np.ascontiguousarray(temp)),
                         _ptr(np.ascontiguousarray(pres)), _ptr(np.ascontiguousarray(mask)),
                         ctypes.c_double(dto), _I(it), _I(serie), _I(sample))
        self.L.ref_flush()

    def leer_campos(self, fname, mach, shape):
        """LEER_CAMPOS of fname (and sobject.field) in the current working
        directory; returns u, temp, pres, um0 (VARNC_TO_VARC), sobject."""
        nx, ny, nz = shape
        self.init(nx, ny, nz)
        u = np.empty((3, nz, ny, nx))
        temp, pres, sob = (np.empty((nz, ny, nx)) for _ in range(3))
        um0 = np.empty((5, nz, ny, nx))
        b = fname.encode()
        self.L.ref_leer_campos(ctypes.c_char_p(b), _I(len(b)), ctypes.c_double(mach), _ptr(u),
                               _ptr(temp), _ptr(pres), _ptr(um0), _ptr(sob))
        self.L.ref_flush()
        return u, temp, pres, um0, sob

    def leer_malla(self, fname, shape):
        nx, ny, nz = shape
        self.init(nx, ny, nz)
        x = np.empty((3, nz, ny, nx))
        b = fname.encode()
        self.L.ref_leer_malla(ctypes.c_char_p(b), _I(len(b)), _ptr(x))
        self.L.ref_flush()
        return x

    def leer_time(self, fname):
        dto, it = ctypes.c_double(), ctypes.c_int()
        b = fname.encode()
        self.L.ref_leer_time(ctypes.c_char_p(b), _I(len(b)), ctypes.byref(dto), ctypes.byref(it))
        self.L.ref_flush()
        return dto.value, it.value

    def jacobeano(self, x):
        self.init(*dims(x))
        jbn = np.empty((11,) + x.shape[1:])
        self.L.ref_jacobeano(_ptr(x), _ptr(jbn))
        return jbn

    def deriv_vel(self, u, temp):
        self.init(*dims(u))
        dvel = np.empty((9,) + u.shape[1:])
        dtemp = np.empty((3,) + u.shape[1:])
        self.L.ref_deriv_vel(_ptr(u), _ptr(temp), _ptr(dvel), _ptr(dtemp))
        return dvel, dtemp

    def flujos(self, neq, um, u, pres, vis, jbn, dvel):
        self.init(*dims(u))
        flux = np.empty((6,) + u.shape[1:])
        self.L.ref_flujos(_I(neq), _ptr(um), _ptr(u), _ptr(pres), _ptr(vis), _ptr(jbn),
                          _ptr(dvel), _ptr(flux))
        return flux

    def divergencia(self, neq, flux, jbn):
        self.init(*dims(flux))
        rs = np.empty(flux.shape[1:])
        rsv = np.empty(flux.shape[1:])
        self.L.ref_divergencia(_I(neq), _ptr(flux), _ptr(jbn), _ptr(rs), _ptr(rsv))
        return rs, rsv

    def inisgdm(self, x):
        self.init(*dims(x))
        sgd = np.empty((7,) + x.shape[1:])
        dmins = np.empty(x.shape[1:])
        dmax = ctypes.c_double(0.0)
        self.L.ref_inisgdm(_ptr(x), _ptr(sgd), _ptr(dmins), ctypes.byref(dmax))
        return sgd, dmins, dmax.value

    def sgdm(self, u, dvel, jbn, sgd, c4):
        self.init(*dims(u))
        amut, vis, vis5 = (np.empty(u.shape[1:]) for _ in range(3))
        self.L.ref_sgdm(_ptr(u), _ptr(dvel), _ptr(jbn), _ptr(sgd), _D(c4), _ptr(amut),
                        _ptr(vis), _ptr(vis5))
        return amut, vis, vis5

    # whole RK-stage operator (RUNGEK sub-step body, for the CPU baseline)
    def stage_load(self, um, u, temp, pres, jbn, sgd, c4):
        self.init(*dims(u))
        self._keep = (um, u, temp, pres, jbn, sgd)
        self.L.ref_stage_load(_ptr(um), _ptr(u), _ptr(temp), _ptr(pres), _ptr(jbn),
                              _ptr(sgd), _D(c4))

    def stage_run(self):
        self.L.ref_stage_run()

    def flush(self):
        """Flush the Fortran runtime's buffered standard output."""
        self.L.ref_flush()

    def stage_fetch(self):
        shp = self._keep[1].shape[1:]
        rs = np.empty((5,) + shp)
        rsv = np.empty((5,) + shp)
        amut, vis = np.empty(shp), np.empty(shp)
        dvel = np.empty((9,) + shp)
        self.L.ref_stage_fetch(_ptr(rs), _ptr(rsv), _ptr(amut), _ptr(vis), _ptr(dvel))
        return rs, rsv, amut, vis, dvel
