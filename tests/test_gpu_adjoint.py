"""Hand-written adjoint of the fused McICA path against finite differences of the FP64
exact-exponential forward (the function the adjoint differentiates), masks fixed; and that
forward against the oracle's exact-exponential mode.

Objectives: J1 = TOA upward flux, J2 = sum of the layer heating rates (total sky), on 64 columns
sampled from the 4096 x 60 exponential-random-overlap + aerosol workload (the cfg3 sample).
Finite differences: step h = 1e-4 |x| for every non-zero T, H2O, O3, LWP, IWP and the four
albedos of each sampled column, evaluated as one batched forward call (one column per
perturbation; the McICA streams depend on pressure only, so every copy carries the same masks).

The forward is piecewise smooth: its tables are interpolated linearly in cells (15 K temperature
cells, key-species bins, ...), so where a cell boundary lies within h of x the central
difference straddles a kink and measures neither side's derivative. Each element therefore has
three quotients: the central difference, and on each side the Richardson extrapolation
2 D(h/2) - D(h) of the one-sided differences, which is exact to O(h^2) where that side holds no
kink. The bar (written here): for every element the adjoint is within 1e-4 relative of one of
them, plus the quotient's rounding allowance (floor / h, where the floor is the FP64 rounding
level of J); each column's gradient vector per input kind is within 1e-4 relative of the
quotients so chosen. Elements whose difference quotient cannot resolve the derivative to 1e-5
(max(|adjoint|, |central|) 2h below 1e5 times the floor: mostly the upper layers, to which the
objectives are all but insensitive) are left out and counted; their share of sum |dJ/dx| h must
stay below 1e-3.
"""
import numpy as np
import pytest

import oracle_bind as ob
from rrtmg_sw_amd.workloads import workload

pytestmark = pytest.mark.gpu

KINDS = ("tlay", "h2ovmr", "o3vmr", "clwp", "ciwp")
ALB = ("asdir", "asdif", "aldir", "aldif")
NCOL_FD = 64


def _batch(cols, pert):
    """Columns: column i of `cols` repeated per perturbation (i, kind, layer, delta) in `pert`."""
    idx = np.array([p[0] for p in pert])
    b = cols.subset(idx)
    for j, (i, kind, lay, dx) in enumerate(pert):
        if kind is None:
            continue
        a = getattr(b, kind)
        if lay is None:
            a[j] += dx
        else:
            a[j, lay] += dx
    return b


HEATFAC = 9.8066 * 86400.0 / (1004.64 * 100.0)


def _forward(api, cols):
    import torch
    d = api.device_inputs(cols)
    out = api.mcica_rrtmg_sw(cols, d, precision="exact")
    torch.cuda.synchronize()
    return {k: api.to_host(out[k]) for k in ("swuflx", "swdflx", "swhr")}


def _objective(o, st, sh):
    """J = sum seed_toa swuflx(top) + sum seed_hr swhr, per column."""
    L = o["swhr"].shape[1]
    return st * o["swuflx"][:, L] + (sh * o["swhr"]).sum(axis=1)


def _floor(o, plev, st, sh):
    """FP64 rounding floor of J per column: J = sum over levels of weights times the up and
    down fluxes (swhr(k) = (net(k+1) - net(k)) heatfac / pdp(k), the top layer's zero), a few
    ulps of each flux times its weight, added in quadrature."""
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


def _fd_sample():
    full = workload("expran4096x60")
    return full.subset(np.arange(0, full.ncol, full.ncol // NCOL_FD)[:NCOL_FD])


def _quotients(api, cols, elems, scale, objs):
    """Per objective: central, forward and backward (Richardson) quotients of every element at
    step scale * h, and the rounding floor of J per element."""
    steps = (1.0, -1.0, 0.5, -0.5)
    pert = [(i, k, lay, s * scale * h) for (i, k, lay, h) in elems for s in steps]
    o = _forward(api, _batch(cols, pert))
    o0 = _forward(api, cols)
    col = np.array([e[0] for e in elems])
    h = scale * np.array([e[3] for e in elems])
    res = {}
    for name, (st, sh) in objs.items():
        rep = np.repeat(col, 4)
        f = _objective(o, st[rep], sh[rep]).reshape(-1, 4)
        f0 = _objective(o0, st, sh)[col]
        fl = _floor(o0, cols.plev, st, sh)[col]
        cen = (f[:, 0] - f[:, 1]) / (2 * h)
        fwd = 2 * (f[:, 2] - f0) / (h / 2) - (f[:, 0] - f0) / h
        bwd = 2 * (f0 - f[:, 3]) / (h / 2) - (f0 - f[:, 1]) / h
        res[name] = (cen, fwd, bwd, fl, h)
    return res


def _within(adj, q):
    cen, fwd, bwd, fl, h = q
    ok_c = np.abs(adj - cen) <= 1e-4 * np.abs(cen) + fl / h
    ok_f = np.abs(adj - fwd) <= 1e-4 * np.abs(fwd) + 10 * fl / h
    ok_b = np.abs(adj - bwd) <= 1e-4 * np.abs(bwd) + 10 * fl / h
    return ok_c, ok_f, ok_b


def test_adjoint_matches_finite_differences(gpu_lib):
    """Objectives: J1 the TOA upward flux, J2 the heating-rate sum (cfg3's two), J3 the heating
    rates weighted by random factors in [0.5, 1.5] times pdp / heatfac (the absorbed flux of each
    layer; its rounding floor is that of a flux, so nearly every element resolves, which J2's
    heatfac / pdp floor does not allow for most temperatures). Where an element meets none of its
    three quotients at step h (a jump within the step on both sides: reftra's conservative /
    general branch switch of some g-point, a discontinuity of the reference's algorithm), the
    quotients are taken again at h / 8. The unresolved share of sum |dJ/dx| h must stay below
    1e-3 for J1 and J3; for J2 the elements beyond the FD's reach are covered in aggregate by the
    directional test."""
    import torch
    api = gpu_lib
    cols = _fd_sample()
    N, L = cols.ncol, cols.nlay
    elems = []
    for i in range(N):
        for kind in KINDS:
            x = getattr(cols, kind)
            for lay in range(L):
                if x[i, lay] != 0.0:
                    elems.append((i, kind, lay, 1e-4 * abs(x[i, lay])))
        for kind in ALB:
            elems.append((i, kind, None, 1e-4 * abs(getattr(cols, kind)[i])))
    pdp = cols.plev[:, :-1] - cols.plev[:, 1:]
    w3 = np.random.default_rng(17).uniform(0.5, 1.5, (N, L)) * pdp / HEATFAC
    objs = {1: (np.ones(N), np.zeros((N, L))), 2: (np.zeros(N), np.ones((N, L))),
            3: (np.zeros(N), w3)}
    q1 = _quotients(api, cols, elems, 1.0, objs)

    d = api.device_inputs(cols)
    adj_of = {}
    for name, (st, sh) in objs.items():
        g = api.mcica_rrtmg_sw_adjoint(cols, d, api.to_device(st), api.to_device(sh))
        torch.cuda.synchronize()
        gh = {k: api.to_host(v) for k, v in g.items()}
        adj_of[name] = np.array([gh["alb"][i, ALB.index(k)] if lay is None else gh[k][i, lay]
                                 for (i, k, lay, _) in elems])
    col = np.array([e[0] for e in elems])
    kind_of = np.array(["alb" if e[2] is None else e[1] for e in elems])
    h = np.array([e[3] for e in elems])
    # elements meeting no quotient at h, for any objective: quotients again at h / 8
    miss = np.zeros(len(elems), dtype=bool)
    for name in objs:
        cen, fwd, bwd, fl, _ = q1[name]
        resolved = np.maximum(np.abs(adj_of[name]), np.abs(cen)) * 2 * h > 1e5 * fl
        ok = _within(adj_of[name], q1[name])
        miss |= resolved & ~(ok[0] | ok[1] | ok[2])
    sub = np.flatnonzero(miss)
    q8 = _quotients(api, cols, [elems[n] for n in sub], 0.125, objs) if len(sub) else None
    failures = []
    for name in objs:
        adj = adj_of[name]
        cen, fwd, bwd, fl, _ = q1[name]
        resolved = np.maximum(np.abs(adj), np.abs(cen)) * 2 * h > 1e5 * fl
        ok_c, ok_f, ok_b = _within(adj, q1[name])
        ref = np.where(ok_c, cen, np.where(ok_f, fwd, np.where(ok_b, bwd, cen)))
        good = ok_c | ok_f | ok_b
        fine = np.zeros(len(elems), dtype=bool)
        if q8 is not None:
            c8, f8, b8, _, _ = q8[name]
            o8 = _within(adj[sub], q8[name])
            g8 = (o8[0] | o8[1] | o8[2]) & ~good[sub]
            fine[sub] = g8
            r8 = np.where(o8[0], c8, np.where(o8[1], f8, np.where(o8[2], b8, c8)))
            ref[sub] = np.where(g8, r8, ref[sub])
            good = good | fine
        mag = np.maximum(np.abs(adj), np.abs(cen)) * h
        for kind in KINDS + ("alb",):
            sel = kind_of == kind
            keep = sel & resolved
            ntot, nskip = int(keep.sum()), int((sel & ~resolved).sum())
            bad = keep & ~good
            one_sided = int((keep & ~ok_c & (ok_f | ok_b)).sum())
            nfine = int((keep & fine).sum())
            share = mag[sel & ~resolved].sum() / max(mag[sel].sum(), 1e-300)
            worst = 0.0
            for i in range(N):
                m = keep & (col == i)
                if m.any() and np.linalg.norm(ref[m]) > 0:
                    worst = max(worst, np.linalg.norm(adj[m] - ref[m]) / np.linalg.norm(ref[m]))
            print(f"J{name} d/d{kind}: {ntot} resolved, {nskip} below the FD floor (share of "
                  f"sum |dJ/dx| h {share:.1e}); central within 1e-4: "
                  f"{ntot - one_sided - nfine - int(bad.sum())}, one-sided: {one_sided}, at h/8: "
                  f"{nfine}, off: {int(bad.sum())}; worst column relative error {worst:.2e}")
            for n in np.flatnonzero(bad)[:5]:
                i, k, lay, hh = elems[n]
                print(f"    off: column {i} {k} layer {lay}: adjoint {adj[n]:.9e}, central "
                      f"{cen[n]:.9e}, forward {fwd[n]:.9e}, backward {bwd[n]:.9e}, h {hh:.3e}, "
                      f"floor/h {fl[n] / hh:.2e}")
            assert ntot > 0, (name, kind)
            if name != 2:
                assert share <= 1e-3, (name, kind, share)
            if bad.any() or worst > 1e-4:
                failures.append((name, kind, int(bad.sum()), worst))
    assert not failures, failures


def test_exact_forward_matches_oracle(gpu_lib):
    """The exact-exponential GPU forward (the function the finite differences above probe)
    against the oracle's exact-exponential mode: a flux or heating-rate error of the smooth
    forward would otherwise be differentiated consistently and go unseen. The two differ in
    rounding only (the kernel's reftra is an algebraically equal rearrangement), so the bar
    is far below the path's 0.1 W/m2 / 0.01 K/day: 1e-6 W/m2, 1e-4 K/day."""
    import torch
    api = gpu_lib
    worst = {}
    for name, step in (("expran4096x60", 16), ("mcica8192x72", 64)):
        full = workload(name)
        cols = full.subset(np.arange(0, full.ncol, step))
        d = api.device_inputs(cols)
        out = api.mcica_rrtmg_sw(cols, d, precision="exact")
        torch.cuda.synchronize()
        mc, _ = ob.run_mcica(cols)
        ob.orc().orc_set_exact_exp(1)
        try:
            ref = ob.run_rrtmg(cols, mc)
        finally:
            ob.orc().orc_set_exact_exp(0)
        e = {k: float(np.abs(api.to_host(out[k]) - ref[k]).max()) for k in out}
        print(f"exact forward vs oracle (exact mode), {name} ({cols.ncol} columns):", e)
        worst[name] = e
        for k in ("swuflx", "swdflx", "swuflxc", "swdflxc"):
            assert e[k] <= 1e-6, (name, k, e[k])
        for k in ("swhr", "swhrc"):
            assert e[k] <= 1e-4, (name, k, e[k])


def test_exact_forward_close_to_table_forward(gpu_lib):
    """The exact-exponential forward differs from the table forward only by the table's
    quantisation (a few 1e-4 of a transmittance)."""
    import torch
    api = gpu_lib
    cols = workload("expran4096x60", ncol=256)
    d = api.device_inputs(cols)
    a = api.mcica_rrtmg_sw(cols, d, precision="fp64")
    b = api.mcica_rrtmg_sw(cols, d, precision="exact")
    torch.cuda.synchronize()
    e = {k: float((a[k] - b[k]).abs().max().item()) for k in a}
    print("exact vs table forward:", e)
    assert max(e[k] for k in ("swuflx", "swdflx")) < 5.0


def test_adjoint_directional_derivative(gpu_lib):
    """<dJ/dx, dx> against finite differences along a random relative direction touching every
    input at once (dx = 1e-4 |x| r, r uniform in [-1, 1]): the sum over all elements stands far
    above the rounding floor, also for the heating-rate sum whose elements mostly do not. The
    quotients are those of gradcheck.directional; each column within 1e-4 relative of one."""
    import torch
    from rrtmg_sw_amd import gradcheck
    api = gpu_lib
    cols = _fd_sample()
    N, L = cols.ncol, cols.nlay
    rng = np.random.default_rng(7)
    dirs = {k: 1e-4 * getattr(cols, k) * rng.uniform(-1, 1, getattr(cols, k).shape)
            for k in KINDS + ALB}
    d = api.device_inputs(cols)

    def adjoint(st, sh):
        g = api.mcica_rrtmg_sw_adjoint(cols, d, api.to_device(st), api.to_device(sh))
        torch.cuda.synchronize()
        return {k: api.to_host(v) for k, v in g.items()}

    objs = {1: (np.ones(N), np.zeros((N, L))), 2: (np.zeros(N), np.ones((N, L)))}
    for name, (rel, n1, n8) in gradcheck.directional(api, cols, dirs, objs, adjoint).items():
        print(f"J{name} directional over {N} columns: relative error median {np.median(rel):.2e}, "
              f"max {rel.max():.2e}; {n1} columns on a one-sided quotient, {n8} at step / 8")
        assert rel.max() <= 1e-4, (name, np.sort(rel)[-5:])


def test_autograd_function_matches_adjoint(gpu_lib):
    """The torch.autograd.Function harness returns the adjoint's gradients for a weighted loss."""
    import torch
    from rrtmg_sw_amd.torch_adjoint import McicaShortwave
    api = gpu_lib
    cols = workload("expran4096x60", ncol=64)
    d = api.device_inputs(cols)
    xs = [d[k].clone().requires_grad_(True) for k in KINDS]
    alb = torch.stack([d[k] for k in ALB]).clone().requires_grad_(True)
    toa, hr = McicaShortwave.apply(*xs, alb, cols, d)
    w1 = torch.linspace(0.5, 1.5, cols.ncol, dtype=torch.float64, device="cuda")
    w2 = torch.rand(hr.shape, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(3))
    loss = (w1 * toa).sum() + (w2 * hr).sum()
    loss.backward()
    g = api.mcica_rrtmg_sw_adjoint(cols, d, w1.contiguous(), w2.contiguous())
    torch.cuda.synchronize()
    for k, x in zip(KINDS, xs):
        assert torch.allclose(x.grad, g[k], rtol=1e-12, atol=0), k
    assert torch.allclose(alb.grad, g["alb"], rtol=1e-12, atol=0)


def test_checkpointing_matches_store(gpu_lib):
    """Checkpointing every 16 layers recomputes the same sweep states: identical gradients."""
    import torch
    api = gpu_lib
    cols = workload("expran4096x60", ncol=128)
    d = api.device_inputs(cols)
    st = torch.ones(cols.ncol, dtype=torch.float64, device="cuda")
    sh = torch.rand(cols.nlay, cols.ncol, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(5))
    a = api.mcica_rrtmg_sw_adjoint(cols, d, st, sh, ckpt=0)
    b = api.mcica_rrtmg_sw_adjoint(cols, d, st, sh, ckpt=16)
    c = api.mcica_rrtmg_sw_adjoint(cols, d, st, sh, ckpt=7)
    torch.cuda.synchronize()
    for k in a:
        assert torch.equal(a[k], b[k]), k
        assert torch.equal(a[k], c[k]), k
