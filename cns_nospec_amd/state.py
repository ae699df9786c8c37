"""Seeded initial states for the CNS_NoSpec time step.

Two generators:

* :func:`sinusoid_state` -- the synthetic low-Mach state used for every
  parity and benchmark case: periodic domain [-0.1, 0.1]^3, dx = 0.2/n
  (fortran90/main.f90:93-106), primitives ``f = f0 + A*S/max|S|`` where S is
  a seeded sum of products of sines over wavenumbers 1..kmax with random
  phases; rho and p: f0 = 1, A = 0.05; u, v, w: f0 = 0, A = mach*sqrt(gamma).
  Conserved ``E = p/(gamma-1) + rho*(u^2+v^2+w^2)/2``.

* :func:`cgs_state` -- the reference's own initial condition in CGS units
  (fortran90/init_data.f90:43-85: rho ~ 1e-3, e ~ 2.5e9, |u| ~ 1e4).

Both return a flat float64 ghosted U (5 components) in the ABI layout with the
ghost cells already holding their periodic images.
"""
from __future__ import annotations

import math

import numpy as np

from .layout import NC, BoxLayout

GAMMA = 1.4
PROB_LO, PROB_HI = -0.1, 0.1
ETA, ALAM, CFL = 1.8e-4, 1.5e2, 0.5  # fortran90/main.f90:55-57


def box_dx(n):
    """Cell size of the periodic domain for interior extent n (fortran90/main.f90:104-106)."""
    if isinstance(n, int):
        n = (n, n, n)
    return tuple((PROB_HI - PROB_LO) / v for v in n)


def _draw_modes(rng, kmax, nmodes):
    modes = []
    for _ in range(nmodes):
        k = rng.integers(1, kmax + 1, size=3)
        ph = rng.uniform(0.0, 2.0 * math.pi, size=3)
        a = rng.uniform(0.5, 1.0)
        modes.append((k, ph, a))
    return modes


def sinusoid_primitives(n, seed, mach=0.085, kmax=2, nmodes=4):
    """rho, u, v, w, p on the interior, each shaped [k][j][i]."""
    if isinstance(n, int):
        n = (n, n, n)
    rng = np.random.default_rng(seed)
    coords = [(np.arange(v) + 0.5) / v for v in n]  # (x - lo)/L at cell centres
    c0 = math.sqrt(GAMMA)
    spec = {"rho": (1.0, 0.05), "u": (0.0, mach * c0), "v": (0.0, mach * c0),
            "w": (0.0, mach * c0), "p": (1.0, 0.05)}
    out = {}
    for name in ("rho", "u", "v", "w", "p"):
        f0, amp = spec[name]
        s = np.zeros((n[2], n[1], n[0]))
        for k, ph, a in _draw_modes(rng, kmax, nmodes):
            fx = np.sin(2 * math.pi * k[0] * coords[0] + ph[0])
            fy = np.sin(2 * math.pi * k[1] * coords[1] + ph[1])
            fz = np.sin(2 * math.pi * k[2] * coords[2] + ph[2])
            s += a * fz[:, None, None] * fy[None, :, None] * fx[None, None, :]
        smax = np.abs(s).max()
        out[name] = f0 + amp * s / (smax if smax > 0 else 1.0)
    return out


def conserved_from_primitives(rho, u, v, w, p):
    """[rho, rho u, rho v, rho w, rho E] from primitives (ideal gas, gamma = 1.4)."""
    E = p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return np.stack([rho, rho * u, rho * v, rho * w, E])


def fill_periodic(layout: BoxLayout, gv):
    """Closed-form periodic ghost fill of a ghosted [l][k][j][i] view, in place:
    ghost(g) = interior(((g - ng) mod n) + ng) in every direction."""
    ng, n = layout.ng, layout.n
    idx = [((np.arange(n[d] + 2 * ng) - ng) % n[d]) + ng for d in range(3)]
    full = gv[:, idx[2]][:, :, idx[1]][:, :, :, idx[0]]
    gv[...] = full
    return gv


def pack_ghosted(layout: BoxLayout, interior):
    """Flat ghosted array from interior components [l][k][j][i], ghosts filled."""
    ncomp = interior.shape[0]
    flat = layout.new_ghosted(ncomp)
    gv = layout.gview(flat, ncomp)
    layout.interior_of(gv)[...] = interior
    fill_periodic(layout, gv)
    return flat


def sinusoid_state(layout: BoxLayout, seed, mach=0.085, kmax=2, nmodes=4):
    """Flat ghosted U of the seeded sinusoidal state (see module docstring)."""
    prim = sinusoid_primitives(layout.n, seed, mach=mach, kmax=kmax, nmodes=nmodes)
    cons = conserved_from_primitives(prim["rho"], prim["u"], prim["v"], prim["w"], prim["p"])
    return pack_ghosted(layout, cons)


def device_sinusoid_state(layout: BoxLayout, seed, device, mach=0.085, kmax=2, nmodes=4, out=None):
    """:func:`sinusoid_state` evaluated on the device with torch (same modes and
    phases; the sines may differ from numpy's by an ulp).  Interior cells only are
    written; ghosts are left for the step's own fill.  Returns a flat float64
    device tensor in the ghosted layout."""
    import torch

    n = layout.n
    rng = np.random.default_rng(seed)
    c0 = math.sqrt(GAMMA)
    spec = (("rho", 1.0, 0.05), ("u", 0.0, mach * c0), ("v", 0.0, mach * c0),
            ("w", 0.0, mach * c0), ("p", 1.0, 0.05))
    crd = [(torch.arange(v, dtype=torch.float64, device=device) + 0.5) / v for v in n]
    prim = {}
    for name, f0, amp in spec:
        s = torch.zeros((n[2], n[1], n[0]), dtype=torch.float64, device=device)
        for k, ph, a in _draw_modes(rng, kmax, nmodes):
            fx = torch.sin(2 * math.pi * int(k[0]) * crd[0] + float(ph[0]))
            fy = torch.sin(2 * math.pi * int(k[1]) * crd[1] + float(ph[1]))
            fz = torch.sin(2 * math.pi * int(k[2]) * crd[2] + float(ph[2]))
            s.add_(float(a) * fz[:, None, None] * fy[None, :, None] * fx[None, None, :])
        smax = float(s.abs().max())
        s.mul_(amp / (smax if smax > 0 else 1.0)).add_(f0)
        prim[name] = s
    if out is None:
        out = torch.zeros(layout.ghosted_len(NC), dtype=torch.float64, device=device)
    g, ng = layout.dimg, layout.ng
    v = out.view(NC, g[2], g[1], layout.xg)[:, ng:ng + n[2], ng:ng + n[1], ng:ng + n[0]]
    rho, u, w_, p = prim["rho"], prim["u"], prim["w"], prim["p"]
    vv = prim["v"]
    v[0].copy_(rho)
    v[1].copy_(rho * u)
    v[2].copy_(rho * vv)
    v[3].copy_(rho * w_)
    v[4].copy_(p / (GAMMA - 1.0) + 0.5 * rho * (u * u + vv * vv + w_ * w_))
    return out


def cgs_state(layout: BoxLayout):
    """fortran90/init_data.f90:43-85 on a box whose Fortran lo is 0 (cell index * dx / scale)."""
    n = layout.n
    dx = box_dx(n)
    twopi = 2.0 * 3.141592653589793238462643383279502884197
    scale = [(PROB_HI - PROB_LO) / twopi] * 3
    x = np.arange(n[0]) * dx[0] / scale[0]
    y = np.arange(n[1]) * dx[1] / scale[1]
    z = np.arange(n[2]) * dx[2] / scale[2]
    zz, yy, xx = np.meshgrid(z, y, x, indexing="ij")
    uvel = 1.1e4 * np.sin(1 * xx) * np.sin(2 * yy) * np.sin(3 * zz)
    vvel = 1.0e4 * np.sin(2 * xx) * np.sin(4 * yy) * np.sin(1 * zz)
    wvel = 1.2e4 * np.sin(3 * xx) * np.cos(2 * yy) * np.sin(2 * zz)
    rholoc = 1.0e-3 + 1.0e-5 * np.sin(1 * xx) * np.cos(2 * yy) * np.cos(3 * zz)
    eloc = 2.5e9 + 1.0e-3 * np.sin(2 * xx) * np.cos(2 * yy) * np.sin(2 * zz)
    cons = np.stack([rholoc, rholoc * uvel, rholoc * vvel, rholoc * wvel,
                     rholoc * (eloc + (uvel ** 2 + vvel ** 2 + wvel ** 2) / 2)])
    return pack_ghosted(layout, cons)


__all__ = ["GAMMA", "ETA", "ALAM", "CFL", "NC", "box_dx", "sinusoid_primitives",
           "conserved_from_primitives", "fill_periodic", "pack_ghosted",
           "sinusoid_state", "cgs_state"]
