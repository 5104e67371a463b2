"""Python host interface over libemix3d_hip.so.

Method names mirror the reference's Fortran subroutines on the hot path
(solver/deriv_vel.f90 DERIV_VEL, solver/flujos_compact.f90 flujos,
solver/divergencia.f90 divergencia, solver/sgdm_incomp.f90 viscosidad_t,
solver/jacob_malla.f90 JACOBEANO, solver/advectLS.f90 advectLevelSet) and
keep their argument meaning.  Fields are torch float64 tensors on the GPU,
C-contiguous with shape (nz, ny, nx) or (ncomp, nz, ny, nx): the same memory
as the reference's column-major (nx, ny, nz[, ncomp]) arrays.  PyTorch only
provides device memory and the stream; every operator runs in the HIP
kernels of the library.
"""
import ctypes

import torch

from . import _lib

SCHEMES = {"c4": 0, "c6f": 1, "c6c": 2}
KIND_WAVE, KIND_STRIDED, KIND_STRIDED32, KIND_STRIDED16 = 0, 1, 2, 3
PLAN_DERIV, PLAN_FLUX = 0, 1
EMX_MLP_NPARAM, EMX_MLP_NNORM, EMX_MLP_ACT = 9089, 21, 193


def _ptr(t):
    if t is None:
        return None
    if not isinstance(t, torch.Tensor):
        raise TypeError("expected a torch tensor")
    if t.dtype != torch.float64:
        raise TypeError(f"expected float64, got {t.dtype}")
    if not t.is_cuda:
        raise ValueError("expected a GPU tensor (the operator has no CPU path)")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return ctypes.c_void_p(t.data_ptr())


def _fptr(t, numel=None):
    """Pointer of a contiguous float32 GPU tensor (of numel elements)."""
    if not isinstance(t, torch.Tensor) or t.dtype != torch.float32:
        raise TypeError("expected a float32 torch tensor")
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError("expected a contiguous GPU tensor")
    if numel is not None and t.numel() != numel:
        raise ValueError(f"expected {numel} elements, got {t.numel()}")
    return ctypes.c_void_p(t.data_ptr())


def _cur_stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def mlp_forward_f32(params, norm, feat):
    """FP32 forward of the closure network on feature rows [n][10]:
    (pre-activations [n][193], nu [n])."""
    n = feat.shape[0]
    act = torch.empty((n, EMX_MLP_ACT), dtype=torch.float32, device=feat.device)
    nu = torch.empty(n, dtype=torch.float32, device=feat.device)
    _lib.call("emx_mlp_forward_f32", _fptr(params, EMX_MLP_NPARAM), _fptr(norm, EMX_MLP_NNORM),
              _fptr(feat, n * 10), n, _fptr(act), _fptr(nu), _cur_stream())
    return act, nu


def mlp_backward_f32(params, norm, feat, act, dnu, work=None):
    """Parameter gradient sum_p dnu_p d nu_p / d params (FP32, 9089)."""
    n = feat.shape[0]
    need = int(_lib.lib().emx_mlp_backward_work(n))
    if work is None or work.numel() < need:
        work = torch.empty(need, dtype=torch.float32, device=feat.device)
    grad = torch.empty(EMX_MLP_NPARAM, dtype=torch.float32, device=feat.device)
    _lib.call("emx_mlp_backward_f32", _fptr(params, EMX_MLP_NPARAM), _fptr(norm, EMX_MLP_NNORM),
              _fptr(feat, n * 10), _fptr(act, n * EMX_MLP_ACT), _fptr(dnu, n), n, _fptr(grad),
              _fptr(work), work.numel(), _cur_stream())
    return grad, work


def adam_f32(params, grad, m, v, step, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step on the closure parameters (in place)."""
    n = params.numel()
    _lib.call("emx_adam_f32", _fptr(params), _fptr(grad, n), _fptr(m, n), _fptr(v, n), n,
              float(lr), float(beta1), float(beta2), float(eps), int(step), _cur_stream())


class Operator:
    """One grid (nx, ny, nz) and compact scheme: an emx_ctx."""

    def __init__(self, nx, ny, nz, scheme="c6c", device=None):
        if isinstance(scheme, str):
            scheme = SCHEMES[scheme.lower()]
        self.shape = (nx, ny, nz)
        self.scheme = scheme
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None
                                   else torch.device(device).index or 0)
        self._ctx = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.call("emx_create", ctypes.byref(self._ctx), nx, ny, nz, scheme,
                      ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))

    def close(self):
        if self._ctx:
            _lib.lib().emx_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ utils
    def _stream(self):
        _lib.call("emx_set_stream", self._ctx,
                  ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream))

    def _empty(self, ncomp=None):
        nx, ny, nz = self.shape
        shape = (nz, ny, nx) if ncomp is None else (ncomp, nz, ny, nx)
        return torch.empty(shape, dtype=torch.float64, device=self.device)

    def _check(self, t, ncomp=None, name="field"):
        nx, ny, nz = self.shape
        want = (nz, ny, nx) if ncomp is None else (ncomp, nz, ny, nx)
        if tuple(t.shape) != want:
            raise ValueError(f"{name}: shape {tuple(t.shape)} != {want}")
        return _ptr(t)

    def default_ih(self):
        nx, ny, nz = self.shape
        return (float(nx - 1), float(ny - 1), float(nz - 1))

    def set_plan(self, direction, kind, chunk, op=PLAN_DERIV):
        """Line-solver shape of an operation class (PLAN_DERIV: derivatives,
        gradient, metrics; PLAN_FLUX: flux divergence) along a direction."""
        _lib.call("emx_set_plan", self._ctx, op, direction, kind, chunk)

    def get_plan(self, direction, op=PLAN_DERIV):
        k, c = ctypes.c_int(), ctypes.c_int()
        _lib.call("emx_get_plan", self._ctx, op, direction, ctypes.byref(k), ctypes.byref(c))
        return k.value, c.value

    # ---------------------------------------------------------------- timing
    def timing_begin(self, label=None):
        """Start recording every launch of this context (or, with label,
        only the launches of that kernel) with HIP events on its stream
        (emx_timing_select, emx_timing_begin)."""
        _lib.call("emx_timing_select", self._ctx, None if label is None else label.encode())
        _lib.call("emx_timing_begin", self._ctx)

    def timing_end(self):
        """Stop recording; returns one dict per (kernel, direction), largest
        total time first: label, dir (-1 for pointwise kernels), launches,
        ms (summed), bytes (summed algorithmic bytes)."""
        n = ctypes.c_int()
        _lib.call("emx_timing_end", self._ctx, ctypes.byref(n))
        out = []
        for i in range(n.value):
            lab, d, cnt = ctypes.c_char_p(), ctypes.c_int(), ctypes.c_int()
            ms, by = ctypes.c_double(), ctypes.c_double()
            _lib.call("emx_timing_entry", self._ctx, i, ctypes.byref(lab), ctypes.byref(d),
                      ctypes.byref(cnt), ctypes.byref(ms), ctypes.byref(by))
            out.append({"label": lab.value.decode(), "dir": d.value, "launches": cnt.value,
                        "ms": ms.value, "bytes": by.value})
        return out

    # -------------------------------------------------------------- operators
    def derivative(self, f, direction, inv_h=None, out=None):
        """Compact first derivative along direction 0/1/2 (derivnper per line)."""
        ih = self.default_ih()[direction] if inv_h is None else inv_h
        out = self._empty() if out is None else out
        self._stream()
        _lib.call("emx_deriv", self._ctx, direction, self._check(f), self._check(out), ih)
        return out

    def deriv_vel(self, u, temp=None, ih=None, dvel=None, dtemp=None):
        """DERIV_VEL: dvel (9 comps) and dtemp (3 comps) in computational space."""
        ih = self.default_ih() if ih is None else ih
        dvel = self._empty(9) if dvel is None else dvel
        if temp is not None and dtemp is None:
            dtemp = self._empty(3)
        self._stream()
        _lib.call("emx_grad", self._ctx, self._check(u, 3, "u"),
                  None if temp is None else self._check(temp),
                  self._check(dvel, 9, "dvel"),
                  None if dtemp is None else self._check(dtemp, 3, "dtemp"), *ih)
        return dvel, dtemp

    def grad_div(self, u, ih=None, jbn=None, dvel=None, div=None):
        """Velocity gradient (9 comps) and the velocity divergence in
        divergencia's conservative form, in one line pass per direction."""
        ih = self.default_ih() if ih is None else ih
        dvel = self._empty(9) if dvel is None else dvel
        div = self._empty() if div is None else div
        self._stream()
        _lib.call("emx_grad_div", self._ctx, self._check(u, 3, "u"), self._check(jbn, 11, "jbn"),
                  self._check(dvel, 9, "dvel"), self._check(div), *ih)
        return dvel, div

    def jacobeano(self, x, jbn=None):
        """JACOBEANO: the 11 metric components from the coordinates x (3 comps)."""
        jbn = self._empty(11) if jbn is None else jbn
        self._stream()
        _lib.call("emx_metrics", self._ctx, self._check(x, 3, "x"), self._check(jbn, 11, "jbn"))
        return jbn

    def inisgdm(self, x):
        """INISGDM: the 7 SGS geometry fields, dmins and deltamax."""
        sgd = self._empty(7)
        dmins = self._empty()
        dmax = ctypes.c_double()
        self._stream()
        _lib.call("emx_sgs_geometry", self._ctx, self._check(x, 3, "x"), _ptr(sgd),
                  _ptr(dmins), ctypes.byref(dmax))
        return sgd, dmins, dmax.value

    def viscosidad_t(self, u, dvel, jbn, sgd, c4, amut=None, vis=None, vis5=None,
                     want_vis5=True):
        """viscosidad_t + viscosidad: amut, vis (= c4 + amut) and vis5."""
        amut = self._empty() if amut is None else amut
        vis = self._empty() if vis is None else vis
        if want_vis5 and vis5 is None:
            vis5 = self._empty()
        self._stream()
        _lib.call("emx_sgdm", self._ctx, self._check(u, 3, "u"), self._check(dvel, 9, "dvel"),
                  self._check(jbn, 11, "jbn"), self._check(sgd, 7, "sgd"), float(c4),
                  _ptr(amut), _ptr(vis), _ptr(vis5))
        return amut, vis, vis5

    def flujos(self, neq, um, u, pres, vis, jbn, dvel, flux=None):
        """flujos(neq): the 6 flux fields e, f, g, ev, fv, gv."""
        flux = self._empty(6) if flux is None else flux
        self._stream()
        _lib.call("emx_flux", self._ctx, int(neq), self._check(um, 5, "um"),
                  self._check(u, 3, "u"), self._check(pres), self._check(vis),
                  self._check(jbn, 11, "jbn"), self._check(dvel, 9, "dvel"),
                  self._check(flux, 6, "flux"))
        return flux

    def divergencia(self, flux, jbn, ih=None, rs=None, rsv=None):
        """divergencia: rs and rsv from the 6 flux fields (not modified)."""
        ih = self.default_ih() if ih is None else ih
        rs = self._empty() if rs is None else rs
        rsv = self._empty() if rsv is None else rsv
        self._stream()
        _lib.call("emx_divergence", self._ctx, self._check(flux, 6, "flux"),
                  self._check(jbn, 11, "jbn"), *ih, _ptr(rs), _ptr(rsv))
        return rs, rsv

    def flujos_divergencia(self, neq, um, u, pres, vis, jbn, dvel, ih=None, rs=None,
                           rsv=None):
        """flujos(neq) then divergencia(neq), fused (no flux fields stored)."""
        ih = self.default_ih() if ih is None else ih
        rs = self._empty() if rs is None else rs
        rsv = self._empty() if rsv is None else rsv
        self._stream()
        _lib.call("emx_flux_div", self._ctx, int(neq), self._check(um, 5, "um"),
                  self._check(u, 3, "u"), self._check(pres), _ptr(vis),
                  self._check(jbn, 11, "jbn"), _ptr(dvel), *ih, _ptr(rs), _ptr(rsv))
        return rs, rsv

    def flujos_divergencia_all(self, um, u, pres, vis, jbn, dvel, ih=None, rs=None,
                               rsv=None):
        """flujos(m) + divergencia(m) for m = 1..5 (the RUNGEK equation loop)
        in three directional passes; rs, rsv: 5 components."""
        ih = self.default_ih() if ih is None else ih
        rs = self._empty(5) if rs is None else rs
        rsv = self._empty(5) if rsv is None else rsv
        self._stream()
        _lib.call("emx_flux_div_all", self._ctx, self._check(um, 5, "um"),
                  self._check(u, 3, "u"), self._check(pres), self._check(vis),
                  self._check(jbn, 11, "jbn"), self._check(dvel, 9, "dvel"), *ih,
                  self._check(rs, 5, "rs"), self._check(rsv, 5, "rsv"))
        return rs, rsv

    # ------------------------------------------------ eddy-viscosity closures
    def dynsmag(self, u, dvel, jbn, delta, nut=None, c4=0.0, vis=None):
        """Dynamic Smagorinsky nu_t (Germano-Lilly, box test filter of ratio
        2); the target of the learned closure.  With vis, also
        vis = c4 + nu_t."""
        nut = self._empty(1)[0] if nut is None else nut
        self._stream()
        _lib.call("emx_dynsmag", self._ctx, self._check(u, 3, "u"), self._check(dvel, 9, "dvel"),
                  self._check(jbn, 11, "jbn"), self._check(delta), self._check(nut), float(c4),
                  None if vis is None else self._check(vis))
        return nut

    def mlp_nut(self, params, norm, dvel, jbn, delta, nut=None, c4=0.0, vis=None):
        """nu_t of the learned closure at every point (BF16 matrix cores);
        with vis, also vis = c4 + nu_t."""
        nut = self._empty(1)[0] if nut is None else nut
        self._stream()
        _lib.call("emx_mlp_nut", self._ctx, _fptr(params, EMX_MLP_NPARAM),
                  _fptr(norm, EMX_MLP_NNORM), self._check(dvel, 9, "dvel"),
                  self._check(jbn, 11, "jbn"), self._check(delta), self._check(nut), float(c4),
                  None if vis is None else self._check(vis))
        return nut

    def mlp_gather(self, dvel, jbn, delta, idx, feat=None):
        """FP32 features [n][10] of the points idx (int64 linear indices)."""
        n = idx.numel()
        if idx.dtype != torch.int64 or not idx.is_contiguous():
            raise ValueError("idx: contiguous int64")
        feat = torch.empty((n, 10), dtype=torch.float32, device=self.device) if feat is None else feat
        self._stream()
        _lib.call("emx_mlp_gather", self._ctx, self._check(dvel, 9, "dvel"),
                  self._check(jbn, 11, "jbn"), self._check(delta), ctypes.c_void_p(idx.data_ptr()),
                  n, _fptr(feat, n * 10))
        return feat

    # ------------------------------------------------------------- adjoints
    def derivative_adj(self, g, direction, inv_h=None, out=None):
        """Transpose of derivative(): out = D^T g along direction."""
        ih = self.default_ih()[direction] if inv_h is None else inv_h
        out = self._empty(1)[0] if out is None else out
        self._stream()
        _lib.call("emx_deriv_adj", self._ctx, direction, self._check(g), self._check(out),
                  float(ih))
        return out

    def deriv_vel_adj(self, dvel_bar, dtemp_bar=None, ih=None, u_bar=None, temp_bar=None):
        """Transpose of deriv_vel(): (u_bar[3], temp_bar or None)."""
        ih = self.default_ih() if ih is None else ih
        u_bar = self._empty(3) if u_bar is None else u_bar
        if dtemp_bar is not None and temp_bar is None:
            temp_bar = self._empty(1)[0]
        self._stream()
        _lib.call("emx_grad_adj", self._ctx, self._check(dvel_bar, 9, "dvel_bar"),
                  None if dtemp_bar is None else self._check(dtemp_bar, 3, "dtemp_bar"),
                  self._check(u_bar, 3, "u_bar"),
                  None if dtemp_bar is None else self._check(temp_bar), *ih)
        return u_bar, (temp_bar if dtemp_bar is not None else None)

    def divergencia_adj(self, rs_bar, rsv_bar, jbn, ih=None, flux_bar=None):
        """Transpose of divergencia() with respect to its six flux fields."""
        ih = self.default_ih() if ih is None else ih
        flux_bar = self._empty(6) if flux_bar is None else flux_bar
        self._stream()
        _lib.call("emx_divergence_adj", self._ctx, self._check(rs_bar), self._check(rsv_bar),
                  self._check(jbn, 11, "jbn"), *ih, self._check(flux_bar, 6, "flux_bar"))
        return flux_bar

    def flujos_divergencia_all_adj(self, rs_bar, rsv_bar, um, u, vis, jbn, dvel, ih=None,
                                   out=None):
        """Transpose of the linearisation of flujos_divergencia_all() about
        (um, u, vis, dvel); returns dict um_bar(5), u_bar(3), pres_bar,
        vis_bar, dvel_bar(9)."""
        ih = self.default_ih() if ih is None else ih
        if out is None:
            out = {"um_bar": self._empty(5), "u_bar": self._empty(3),
                   "pres_bar": self._empty(1)[0], "vis_bar": self._empty(1)[0],
                   "dvel_bar": self._empty(9)}
        self._stream()
        _lib.call("emx_flux_div_all_adj", self._ctx, self._check(rs_bar, 5, "rs_bar"),
                  self._check(rsv_bar, 5, "rsv_bar"), self._check(um, 5, "um"),
                  self._check(u, 3, "u"), self._check(vis), self._check(jbn, 11, "jbn"),
                  self._check(dvel, 9, "dvel"), *ih, self._check(out["um_bar"], 5, "um_bar"),
                  self._check(out["u_bar"], 3, "u_bar"), self._check(out["pres_bar"]),
                  self._check(out["vis_bar"]), self._check(out["dvel_bar"], 9, "dvel_bar"))
        return out

    def rk_update(self, sub, um0, um, rs, rsv, dt, um1=None):
        """SSP-RK3 update of sub-step 1..3 (solver/rungek_ico.f90:87-91,
        152-153, 216-219); um may be None for sub-step 1."""
        um1 = self._empty(5) if um1 is None else um1
        self._stream()
        _lib.call("emx_rk_update", self._ctx, int(sub), self._check(um1, 5, "um1"),
                  self._check(um0, 5, "um0"), None if um is None else self._check(um, 5, "um"),
                  self._check(rs, 5, "rs"), self._check(rsv, 5, "rsv"), float(dt))
        return um1

    def frontera_val(self, um1):
        """Boundary values z, y, x of um1 in place (frontera_{z,y,x}_val)."""
        self._stream()
        _lib.call("emx_frontera_val", self._ctx, self._check(um1, 5, "um1"))
        return um1

    def varc_to_varnc(self, um1, mach, u=None, pres=None, temp=None, conc=None):
        """Primitive variables from um1 (varc_to_varnc); returns u, pres,
        temp (conc filled when given)."""
        u = self._empty(3) if u is None else u
        pres = self._empty() if pres is None else pres
        temp = self._empty() if temp is None else temp
        self._stream()
        _lib.call("emx_varc_to_varnc", self._ctx, self._check(um1, 5, "um1"), float(mach),
                  self._check(u, 3, "u"), self._check(pres), self._check(temp),
                  None if conc is None else self._check(conc))
        return u, pres, temp

    def varnc_to_varc(self, u, pres, temp, mach, um0=None, pres0=None):
        """Conservative variables from u, pres, temp (varnc_to_varc)."""
        um0 = self._empty(5) if um0 is None else um0
        self._stream()
        _lib.call("emx_varnc_to_varc", self._ctx, self._check(u, 3, "u"), self._check(pres),
                  self._check(temp), float(mach), self._check(um0, 5, "um0"),
                  None if pres0 is None else self._check(pres0))
        return um0

    def rk_clip(self, um1, mask, mach, u, pres, temp, conc=None):
        """RUNGEK velocity clip of um1 in place, primitive variables out."""
        self._stream()
        _lib.call("emx_rk_clip", self._ctx, self._check(um1, 5, "um1"), self._check(mask),
                  float(mach), self._check(u, 3, "u"), self._check(pres), self._check(temp),
                  None if conc is None else self._check(conc))
        return um1

    def rk_rebuild(self, um1, u, pres, mach):
        """um1(1..4) rebuilt from pres and u, in place."""
        self._stream()
        _lib.call("emx_rk_rebuild", self._ctx, self._check(um1, 5, "um1"), self._check(u, 3, "u"),
                  self._check(pres), float(mach))
        return um1

    def filtrado(self, itf, um1, jbn, cfilt, mach, u=None, pres=None, temp=None):
        """Compact filter FILTRADO(itf) of um1 in place; returns u, pres,
        temp recomputed from the filtered variables."""
        u = self._empty(3) if u is None else u
        pres = self._empty() if pres is None else pres
        temp = self._empty() if temp is None else temp
        self._stream()
        _lib.call("emx_filtrado", self._ctx, int(itf), self._check(um1, 5, "um1"),
                  self._check(jbn, 11, "jbn"), float(cfilt), float(mach), self._check(u, 3, "u"),
                  self._check(pres), self._check(temp))
        return u, pres, temp

    def advect_level_set(self, pres, jbn, sobject, dxyz, deltamin):
        """advectLevelSet: extrapolate pres into the solid (in place)."""
        self._stream()
        _lib.call("emx_ls_extrapolate", self._ctx, self._check(pres), self._check(jbn, 11, "jbn"),
                  self._check(sobject), *map(float, dxyz), float(deltamin))
        return pres

    def reinit_level_set(self, phi, jbn, dmin, dxyz, deltamin):
        """Level-set reinitialisation cuReinitLsJB of phi (in place)."""
        self._stream()
        _lib.call("emx_ls_reinit", self._ctx, self._check(phi), self._check(jbn, 11, "jbn"),
                  self._check(dmin), *map(float, dxyz), float(deltamin))
        return phi

    def heviside(self, phi, deltamax, hevi=None):
        """heviside: the smoothed Heaviside of phi over |phi| <= 1.5
        deltamax (solver/advectLS.f90:29-59); hevi may be phi (in place)."""
        hevi = self._empty() if hevi is None else hevi
        self._stream()
        _lib.call("emx_heviside", self._ctx, self._check(phi), float(deltamax), self._check(hevi))
        return hevi

    def advect_weno5(self, phi, vel, jbn, dxyz, dt):
        """One SSP-RK3 WENO5 advection step of phi (in place)."""
        self._stream()
        _lib.call("emx_ls_advect_weno5", self._ctx, self._check(phi), self._check(vel, 3, "vel"),
                  self._check(jbn, 11, "jbn"), *map(float, dxyz), float(dt))
        return phi

    def rk_stage(self, um, u, temp, pres, jbn, sgd, c4, ih=None, work=None):
        """The spatial operator of one RUNGEK sub-step
        (solver/rungek_ico.f90:79-95): DERIV_VEL, viscosidad_t, viscosidad and
        flujos+divergencia for the five equations.  Returns a dict of device
        tensors (rs, rsv: 5 comps each; dvel, dtemp, amut, vis)."""
        w = self.stage_buffers() if work is None else work
        self.deriv_vel(u, temp, ih, dvel=w["dvel"], dtemp=w["dtemp"])
        self.viscosidad_t(u, w["dvel"], jbn, sgd, c4, amut=w["amut"], vis=w["vis"],
                          want_vis5=False)
        self.flujos_divergencia_all(um, u, pres, w["vis"], jbn, w["dvel"], ih,
                                    rs=w["rs"], rsv=w["rsv"])
        return w

    def stage_buffers(self):
        return {"dvel": self._empty(9), "dtemp": self._empty(3), "amut": self._empty(),
                "vis": self._empty(), "rs": self._empty(5), "rsv": self._empty(5)}

    def rk_substep(self, sub, um0, um, u, temp, pres, jbn, sgd, c4, dt, mach, mask,
                   conc=None, sobject=None, dxyz=None, deltamin=None, work=None, um1=None):
        """One RUNGEK sub-step on the device (solver/rungek_ico.f90:79-141,
        143-197, 201-265): the spatial operator on (um, u, temp, pres), the
        SSP-RK3 update into um1, boundary values z/y/x, the velocity clip
        with mask, the pressure extrapolation into the solid when sobject is
        given (advectLevelSet(0)), and um1 rebuilt from pres and u.  u, pres,
        temp (and conc) are updated in place; returns um1.  The impeller
        forcing (rushton) is not applied.  For sub-step 1 the reference
        first copies um0 into um, so pass um = um0."""
        w = self.rk_stage(um, u, temp, pres, jbn, sgd, c4, work=work)
        um1 = self.rk_update(sub, um0, None if sub == 1 else um, w["rs"], w["rsv"], dt, um1=um1)
        self.frontera_val(um1)
        self.rk_clip(um1, mask, mach, u, pres, temp, conc)
        if sobject is not None:
            self.advect_level_set(pres, jbn, sobject, dxyz, deltamin)
        self.rk_rebuild(um1, u, pres, mach)
        return um1


class Batch:
    """Many grids ("boxes") stepped together (emx_batch): per stage the
    gradient fused with the velocity divergence and one WENO5 / SSP-RK3
    level-set step for every box, one launch per phase over all boxes,
    replayed from a HIP graph."""

    def __init__(self, dims, scheme="c6c"):
        if isinstance(scheme, str):
            scheme = SCHEMES[scheme.lower()]
        self.dims = [tuple(int(v) for v in d) for d in dims]
        arr = (ctypes.c_int * (3 * len(self.dims)))(*[v for d in self.dims for v in d])
        self._b = ctypes.c_void_p()
        self._keep = {}
        _lib.call("emx_batch_create", ctypes.byref(self._b), len(self.dims), arr, scheme,
                  _cur_stream())

    def close(self):
        if self._b:
            _lib.lib().emx_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bind(self, i, u, jbn, phi, dvel, div, ih, dxyz, dt):
        nx, ny, nz = self.dims[i]
        for t, nc, name in ((u, 3, "u"), (jbn, 11, "jbn"), (phi, None, "phi"), (dvel, 9, "dvel"),
                            (div, None, "div")):
            want = (nz, ny, nx) if nc is None else (nc, nz, ny, nx)
            if tuple(t.shape) != want:
                raise ValueError(f"box {i} {name}: shape {tuple(t.shape)} != {want}")
        self._keep[i] = (u, jbn, phi, dvel, div)  # the batch holds raw pointers
        _lib.call("emx_batch_bind", self._b, i, _ptr(u), _ptr(jbn), _ptr(phi), _ptr(dvel),
                  _ptr(div), *map(float, ih), *map(float, dxyz), float(dt))

    def set_plan(self, direction, lines_per_wave, chunk):
        """Line-solver shape of one direction: lines_per_wave 16/32/64 (the
        strided kind) or 0 (one wavefront per line, direction 0 only)."""
        _lib.call("emx_batch_set_plan", self._b, direction, lines_per_wave, chunk)

    def get_plan(self, direction):
        lg, c = ctypes.c_int(), ctypes.c_int()
        _lib.call("emx_batch_get_plan", self._b, direction, ctypes.byref(lg), ctypes.byref(c))
        return lg.value, c.value

    def tune(self, reps=5):
        _lib.call("emx_batch_set_stream", self._b, _cur_stream())
        _lib.call("emx_batch_tune", self._b, int(reps))

    def stage(self, graph=True):
        _lib.call("emx_batch_set_stream", self._b, _cur_stream())
        _lib.call("emx_batch_stage", self._b, 1 if graph else 0)
