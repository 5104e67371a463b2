"""Synthetic inputs of the benchmark and test workloads (FP64).

Every generator is seeded (numpy PCG64) and returns torch float64 tensors of
shape (ncomp, nz, ny, nx) -- the reference's column-major (nx, ny, nz, ncomp)
memory -- on the requested device, so the same inputs can be built on the
host for the oracle and on the GPU for the kernels.
"""
import math

import numpy as np
import torch

MACH = 0.1          # acoustic scaling of the pressure: um1 = pres*Mach^2
REYNOLDS = 1000.0   # c4 = 1/Re (src/MSVSingle/iniconst_incomp.f90:11)


def _axes(nx, ny, nz, device):
    xi = torch.linspace(0.0, 1.0, nx, dtype=torch.float64, device=device)
    eta = torch.linspace(0.0, 1.0, ny, dtype=torch.float64, device=device)
    zeta = torch.linspace(0.0, 1.0, nz, dtype=torch.float64, device=device)
    return xi, eta, zeta


def grid_uniform(nx, ny, nz, lengths=(1.0, 1.0, 1.0), device="cpu"):
    """x_i = L*i/(n-1) per axis (metrics exact, J constant)."""
    xi, eta, zeta = _axes(nx, ny, nz, device)
    return _tensor_grid(lengths[0] * xi, lengths[1] * eta, lengths[2] * zeta)


def stretch_tanh(s, beta, length=1.0):
    """x(s) = L/2 * (1 + tanh(beta*(2s-1))/tanh(beta)): clustering at both walls."""
    return 0.5 * length * (1.0 + torch.tanh(beta * (2.0 * s - 1.0)) / math.tanh(beta))


def grid_tanh(nx, ny, nz, beta=2.0, lengths=(1.0, 1.0, 1.0), device="cpu"):
    """Orthogonal grid with per-axis tanh stretching toward the walls."""
    xi, eta, zeta = _axes(nx, ny, nz, device)
    return _tensor_grid(stretch_tanh(xi, beta, lengths[0]), stretch_tanh(eta, beta, lengths[1]),
                        stretch_tanh(zeta, beta, lengths[2]))


def _tensor_grid(x1, y1, z1):
    nz, ny, nx = z1.numel(), y1.numel(), x1.numel()
    x = torch.empty((3, nz, ny, nx), dtype=torch.float64, device=x1.device)
    x[0] = x1.view(1, 1, nx)
    x[1] = y1.view(1, ny, 1)
    x[2] = z1.view(nz, 1, 1)
    return x


def velocity_sines(x, seed, modes=8):
    """u_c = sum over axes a and m=1..modes of A sin(2 pi m x_a + phase),
    A ~ U[0,1), phase ~ U[0, 2 pi), drawn per component, axis and mode."""
    rng = np.random.default_rng(seed)
    A = rng.uniform(0.0, 1.0, size=(3, 3, modes))
    ph = rng.uniform(0.0, 2.0 * np.pi, size=(3, 3, modes))
    u = torch.zeros_like(x)
    for c in range(3):
        for a in range(3):
            for m in range(modes):
                u[c] += float(A[c, a, m]) * torch.sin(2.0 * math.pi * (m + 1) * x[a] + float(ph[c, a, m]))
    return u


def velocity_turbulent(x, seed, kmax=32, nmodes=96, jet=True, lengths=(1.0, 1.0, 1.0)):
    """Solenoidal random Fourier (Kraichnan) field with an E(k) ~ k^(-5/3)
    spectrum for 1 <= |k| <= kmax (in units of 2 pi / L), evaluated directly
    at the (stretched) grid points, plus an optional six-blade rotating radial
    jet at mid-height (tip speed 1, blade height 0.2 H) like a disk impeller."""
    rng = np.random.default_rng(seed)
    kmag = np.exp(rng.uniform(0.0, np.log(kmax), size=nmodes))
    dirs = rng.normal(size=(nmodes, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    kvec = kmag[:, None] * dirs * 2.0 * np.pi
    # sigma perpendicular to k
    r = rng.normal(size=(nmodes, 3))
    sig = np.cross(dirs, r)
    sig /= np.linalg.norm(sig, axis=1, keepdims=True)
    # amplitude: E(k) dk with log-uniform sampling -> a^2 ~ k^(-5/3) * k
    amp = kmag ** (-5.0 / 6.0) * np.sqrt(kmag)
    amp *= 0.3 / np.sqrt((amp ** 2).sum() / 2.0)
    phase = rng.uniform(0.0, 2.0 * np.pi, size=nmodes)
    u = torch.zeros_like(x)
    for m in range(nmodes):
        arg = (float(kvec[m, 0]) * x[0] + float(kvec[m, 1]) * x[1] + float(kvec[m, 2]) * x[2]
               + float(phase[m]))
        cm = float(amp[m]) * torch.cos(arg)
        for c in range(3):
            u[c] += float(sig[m, c]) * cm
    if jet:
        u += impeller_jet(x, lengths)
    return u


def impeller_jet(x, lengths=(1.0, 1.0, 1.0), rtip_frac=0.25, blades=6):
    """Rotating radial jet of a six-blade disk impeller at mid-height: solid
    rotation u_theta = r/R_tip inside the tip radius (tip speed 1), a decaying
    swirl and radial outflow outside, modulated by the blade passage and
    confined to the blade height 0.4H..0.6H with smooth edges."""
    Lx, Ly, Lz = lengths
    xc, yc = 0.5 * Lx, 0.5 * Ly
    rt = rtip_frac * min(Lx, Ly)
    dx, dy = x[0] - xc, x[1] - yc
    r = torch.sqrt(dx * dx + dy * dy) + 1e-300
    th = torch.atan2(dy, dx)
    zc = (x[2] - 0.5 * Lz) / (0.1 * Lz)
    band = torch.exp(-zc ** 4)
    blade = 0.75 + 0.25 * torch.cos(blades * th)
    inside = r <= rt
    ut = torch.where(inside, r / rt, (rt / r) ** 2) * band * blade
    ur = torch.where(inside, 0.2 * (r / rt) ** 2, 0.2 * (rt / r)) * band * blade
    out = torch.zeros_like(x)
    out[0] = ur * dx / r - ut * dy / r
    out[1] = ur * dy / r + ut * dx / r
    return out


def scalar_smooth(x, seed, base=1.0, amp=0.1, modes=4):
    """A smooth seeded scalar field (temperature, pressure perturbation)."""
    rng = np.random.default_rng(seed)
    A = rng.uniform(0.0, 1.0, size=(3, modes))
    ph = rng.uniform(0.0, 2.0 * np.pi, size=(3, modes))
    s = torch.full_like(x[0], base)
    for a in range(3):
        for m in range(modes):
            s += amp / (m + 1) * float(A[a, m]) * torch.sin(2.0 * math.pi * (m + 1) * x[a] + float(ph[a, m]))
    return s


def conservative(u, temp, pres, mach=MACH):
    """varnc_to_varc (src/MSVSingle/herramientas_incomp2.f90:44-52):
    um1 = pres*Mach^2, um2..4 = u*um1, um5 = temp."""
    um = torch.empty((5,) + tuple(u.shape[1:]), dtype=torch.float64, device=u.device)
    um[0] = pres * mach * mach
    um[4] = temp
    um[1] = u[0] * um[0]
    um[2] = u[1] * um[0]
    um[3] = u[2] * um[0]
    return um


def stage_fields(nx, ny, nz, seed, curvilinear=True, turbulent=True, device="cpu"):
    """Inputs of one RK-stage operator evaluation: grid x, u, temp, pres, um."""
    x = grid_tanh(nx, ny, nz, device=device) if curvilinear else grid_uniform(nx, ny, nz, device=device)
    u = velocity_turbulent(x, seed) if turbulent else velocity_sines(x, seed)
    temp = scalar_smooth(x, seed + 1, base=1.0, amp=0.1)
    pres = scalar_smooth(x, seed + 2, base=1.0 / (MACH * MACH), amp=1.0)
    um = conservative(u, temp, pres)
    return {"x": x, "u": u, "temp": temp, "pres": pres, "um": um}


def level_set_sphere_surface(x, seed, lengths=(1.0, 1.0, 1.0), radius_frac=0.25,
                             surface_frac=0.8):
    """phi = signed distance to the union of a sphere (radius 0.25 L, seeded
    centre) with the region below the free surface z = 0.8 H; phi < 0 inside
    the liquid."""
    rng = np.random.default_rng(seed)
    L = min(lengths)
    c = [float(rng.uniform(0.3, 0.7) * lengths[a]) for a in range(3)]
    d = torch.sqrt((x[0] - c[0]) ** 2 + (x[1] - c[1]) ** 2 + (x[2] - c[2]) ** 2) - radius_frac * L
    s = x[2] - surface_frac * lengths[2]
    return torch.minimum(d, s)


def batch_dims(seed, nbox=48, lo=(32, 32, 32), hi=(192, 128, 96)):
    """Box sizes of the batched multiphase stage: nbox sizes drawn uniformly
    from {lo..hi} per axis (seeded), with the first box lo^3 and the second
    the largest, hi."""
    rng = np.random.default_rng(seed)
    dims = [tuple(lo), tuple(hi)]
    while len(dims) < nbox:
        dims.append(tuple(int(rng.integers(lo[a], hi[a] + 1)) for a in range(3)))
    return dims[:nbox]


def batch_box(dims, seed, device="cpu"):
    """One box of the batched stage on the unit cube: grid, the sine-mode
    velocity, the sphere-plus-free-surface level set, the inverse spacings,
    the spacings and the time step dt = 0.5 min(dx) / max|u| (CFL 0.5)."""
    nx, ny, nz = dims
    x = grid_uniform(nx, ny, nz, device=device)
    u = velocity_sines(x, seed)
    phi = level_set_sphere_surface(x, seed)
    ih = (nx - 1.0, ny - 1.0, nz - 1.0)
    dxyz = (1.0 / (nx - 1), 1.0 / (ny - 1), 1.0 / (nz - 1))
    dt = 0.5 * min(dxyz) / float(u.abs().max())
    return {"x": x, "u": u, "phi": phi, "ih": ih, "dxyz": dxyz, "dt": dt}
