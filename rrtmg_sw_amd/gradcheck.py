"""Finite-difference check of the adjoint along a direction, on the GPU's own exact-exponential
forward (the function the adjoint differentiates).

The exact-exponential forward is piecewise smooth: its tables are interpolated linearly within
cells, and reftra's conservative / general branch switch is a small jump. A central difference
whose step crosses one measures neither side's derivative, so each check compares the adjoint's
directional derivative with three quotients -- the central difference and the one-sided
Richardson extrapolations 2 D(s/2) - D(s) of either side -- at the step s and, where none
agrees, at s / 8; each with an allowance of its rounding error (the FP64 floor of J over the
step).
"""
import numpy as np

KINDS = ("tlay", "h2ovmr", "o3vmr", "clwp", "ciwp")
ALB = ("asdir", "asdif", "aldir", "aldif")
HEATFAC = 9.8066 * 86400.0 / (1004.64 * 100.0)


def forward(api, cols):
    """Total-sky fluxes and heating rates of the exact-exponential forward (host arrays)."""
    import torch
    d = api.device_inputs(cols)
    out = api.mcica_rrtmg_sw(cols, d, precision="exact")
    torch.cuda.synchronize()
    return {k: api.to_host(out[k]) for k in ("swuflx", "swdflx", "swhr")}


def objective(o, st, sh):
    """J = sum seed_toa swuflx(top) + sum seed_hr swhr, per column."""
    L = o["swhr"].shape[1]
    return st * o["swuflx"][:, L] + (sh * o["swhr"]).sum(axis=1)


def floor(o, plev, st, sh):
    """FP64 rounding floor of J per column: J is a weighted sum of the level fluxes
    (swhr(k) = (net(k+1) - net(k)) heatfac / pdp(k), the top layer's zero); a few ulps of each
    flux times its weight, added in quadrature."""
    L = o["swhr"].shape[1]
    pdp = plev[:, :-1] - plev[:, 1:]
    w = sh * HEATFAC / pdp
    w[:, L - 1] = 0.0
    wn = np.zeros((w.shape[0], L + 1))
    wn[:, :L] -= w
    wn[:, 1:] += w
    lw = np.abs(wn)
    lw[:, L] += np.abs(st)
    mag = np.abs(o["swuflx"]) + np.abs(o["swdflx"])
    return 3e-16 * np.sqrt(((lw * mag) ** 2).sum(axis=1))


def shifted(cols, dirs, s):
    c = cols.subset(np.arange(cols.ncol))   # a copy
    for k, dx in dirs.items():
        getattr(c, k)[...] += s * dx
    return c


def directional(api, cols, dirs, objs, adjoint):
    """For each objective name -> (seed_toa (ncol,), seed_hr (ncol, nlay)) in objs: the relative
    error of the adjoint's directional derivative <dJ/dx, dirs> per column against the best of
    the quotients, and how many columns needed a one-sided or a shorter step.
    adjoint(seed_toa, seed_hr) returns the gradients as host arrays."""
    o0 = forward(api, cols)
    f = {}
    for scale in (1.0, 0.125):
        for s in (1.0, -1.0, 0.5, -0.5):
            f[scale * s] = forward(api, shifted(cols, dirs, scale * s))
    out = {}
    for name, (st, sh) in objs.items():
        g = adjoint(st, sh)
        lin = sum((g[k] * dirs[k]).sum(axis=1) for k in KINDS)
        lin = lin + sum(g["alb"][:, m] * dirs[k] for m, k in enumerate(ALB))
        J0 = objective(o0, st, sh)
        fl = floor(o0, cols.plev, st, sh)
        r = []   # central, forward, backward at s; the same at s / 8
        for scale in (1.0, 0.125):
            J = {s: objective(f[scale * s], st, sh) for s in (1.0, -1.0, 0.5, -0.5)}
            cen = (J[1.0] - J[-1.0]) / (2 * scale)
            fwd = (2 * (J[0.5] - J0) / 0.5 - (J[1.0] - J0)) / scale
            bwd = (2 * (J0 - J[-0.5]) / 0.5 - (J0 - J[-1.0])) / scale
            for q, c in ((cen, 1.0), (fwd, 10.0), (bwd, 10.0)):
                # the error beyond the quotient's rounding allowance, relative to the quotient
                r.append(np.maximum(np.abs(lin - q) - c * fl / scale, 0.0) / np.abs(q))
        r = np.array(r)
        best = r.min(axis=0)
        used = np.where(r[0] <= 1e-4, 0, np.where(r[1:3].min(axis=0) <= 1e-4, 1, 2))
        out[name] = (best, int((used == 1).sum()), int((used == 2).sum()))
    return out
