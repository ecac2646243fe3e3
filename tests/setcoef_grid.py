"""Inputs that put setcoef_sw's index decisions within a few ulps of every switch.

setcoef_sw (src/rrtmg_sw_setcoef.f90:169-226) derives its interpolation indices from
    jp      = int(36 - 5 (log p + 0.04))          clamped to 1..58
    jt, jt1 = int(3 + (T - tref(jp | jp+1)) / 15)  clamped to 1..4
    indfor  = int((332 - T) / 36)                  clamped to 1..2   (log p > 4.56)
    indself = int((T - 188) / 7.2) - 7             clamped to 1..9   (log p > 4.56)
and counts laytrop (log p > 4.56) and laylow (log p >= 6.62). Each is a step function of p or
T. The pressure switches are found here by bisection over the ordered bit patterns of positive
doubles with the oracle's own evaluation (the host libm's log, as the reference's); the
temperature switches are the points where the quotients reach an integer. The grid holds the
64 doubles on either side of every switch.
"""
import numpy as np

import oracle_bind as ob

NEAR = 32           # doubles on each side of a switch
TREF = np.array([
    2.9420e+02, 2.8799e+02, 2.7894e+02, 2.6925e+02, 2.5983e+02, 2.5017e+02, 2.4077e+02,
    2.3179e+02, 2.2306e+02, 2.1578e+02, 2.1570e+02, 2.1570e+02, 2.1570e+02, 2.1706e+02,
    2.1858e+02, 2.2018e+02, 2.2174e+02, 2.2328e+02, 2.2479e+02, 2.2655e+02, 2.2834e+02,
    2.3113e+02, 2.3401e+02, 2.3703e+02, 2.4022e+02, 2.4371e+02, 2.4726e+02, 2.5085e+02,
    2.5457e+02, 2.5832e+02, 2.6216e+02, 2.6606e+02, 2.6999e+02, 2.7340e+02, 2.7536e+02,
    2.7568e+02, 2.7372e+02, 2.7163e+02, 2.6955e+02, 2.6593e+02, 2.6211e+02, 2.5828e+02,
    2.5360e+02, 2.4854e+02, 2.4348e+02, 2.3809e+02, 2.3206e+02, 2.2603e+02, 2.2000e+02,
    2.1435e+02, 2.0887e+02, 2.0340e+02, 1.9792e+02, 1.9290e+02, 1.8809e+02, 1.8329e+02,
    1.7849e+02, 1.7394e+02, 1.7212e+02])   # src/rrtmg_sw_setcoef.f90:347-360


def profiles(play, tlay):
    """setcoef inputs for layers (ncol, nlay) with pressures play and temperatures tlay: falling
    level pressures (so every dry-air amount is positive; setcoef_sw does not relate them to
    play) and a moist tropical mixture."""
    play = np.asfortranarray(play, dtype=np.float64)
    tlay = np.asfortranarray(tlay, dtype=np.float64)
    n, L = play.shape
    plev = np.asfortranarray(np.broadcast_to(1100.0 * 0.95 ** np.arange(L + 1), (n, L + 1)))
    f = lambda v: np.full((n, L), v, order="F")
    return dict(play=play, plev=plev, tlay=tlay, h2ovmr=f(1e-2), o3vmr=f(5e-8), co2vmr=f(4e-4),
                ch4vmr=f(1.8e-6), o2vmr=f(0.209))


def _decide(p):
    """jp, the troposphere test and the lowest-layer test of each pressure (oracle)."""
    p = np.asarray(p, dtype=np.float64).reshape(-1, 1)
    o = ob.run_setcoef(p.shape[0], 1, **profiles(p, np.full(p.shape, 250.0)))
    return o["jp"][:, 0], o["laytrop"], o["laylow"]


def _bits(x):
    return np.asarray(x, dtype=np.float64).view(np.uint64)


def _double(b):
    return np.asarray(b, dtype=np.uint64).view(np.float64)


def pressure_switches():
    """{name: p}: the smallest double pressure at which each decision has switched:
    ("jp", k) for k = 2..58 (jp < k from there on), "trop" (log p > 4.56), "low"
    (log p >= 6.62)."""
    names = [("jp", k) for k in range(2, 59)] + ["trop", "low"]
    n = len(names)
    lo = _bits(np.full(n, 1e-300))
    hi = _bits(np.full(n, 1e300))
    while np.any(hi - lo > 1):
        mid = lo + (hi - lo) // 2
        jp, trop, low = _decide(_double(mid))
        done = np.empty(n, dtype=bool)
        for i, nm in enumerate(names):
            if nm == "trop":
                done[i] = trop[i] == 1
            elif nm == "low":
                done[i] = low[i] == 1
            else:
                done[i] = jp[i] < nm[1]
        hi = np.where(done, mid, hi)
        lo = np.where(done, lo, mid)
    return dict(zip(names, _double(hi)))


def around(x, n=NEAR):
    """The 2n doubles nearest to x: n below it and x with the n - 1 above."""
    b = _bits(np.atleast_1d(x))[:, None] + np.arange(-n, n, dtype=np.int64).astype(np.uint64)
    return _double(b)


def pressure_grid(sw):
    """(columns, 2 NEAR) pressures: one column per pressure switch, T = 250 K."""
    p = np.concatenate([around(v) for v in sw.values()])
    return profiles(p, np.full(p.shape, 250.0))


def temperature_grid():
    """Columns of 2 NEAR temperatures around every temperature switch: jt and jt1 for each of
    the 58 pressure cells (p at the cell's geometric middle), indfor and indself in the
    troposphere (p = 500 hPa)."""
    preflog = 6.96 - 0.2 * np.arange(59)   # src/rrtmg_sw_setcoef.f90:335-346
    ps, ts = [], []
    for jp in range(1, 59):
        pmid = np.exp(0.5 * (preflog[jp - 1] + preflog[jp]))
        for tref in (TREF[jp - 1], TREF[jp]):
            for m in (2, 3, 4):   # 3 + (T - tref) / 15 reaches m
                t = around(tref + 15.0 * (m - 3))
                ts.append(t), ps.append(np.full(t.shape, pmid))
    for t0 in [260.0, 296.0] + [188.0 + 7.2 * m for m in range(8, 17)]:
        t = around(t0)
        ts.append(t), ps.append(np.full(t.shape, 500.0))
    return profiles(np.concatenate(ps), np.concatenate(ts))
