"""Pins of the C oracle that do not go through the compiled reference:

* the reference's own column-mode example suite (run_examples_std_atm, six clear-sky IATM = 0
  cases, the overcast non-McICA case MLS-cld6-imca0-icld2 (cldprop_sw with ice and liquid
  clouds, spcvrt_sw) and the four McICA cases (200 samples of mcica_subcol_sw with the Mersenne
  Twister averaged; maximum-random and exponential-random overlap, decorrelation length given
  or from latitude), committed as tests/golden/column_kats.npz by
  tests/golden/make_column_kats.py):
  broadband and per-band up / diffuse-down / direct-down / down / net flux and heating rate at
  every level, to the printed precision (fluxes 1e-4 W/m2, heating rates 1e-5 K/day);
* MT19937's published known answer (the C++ standard's std::mt19937 check: the 10000th output
  of the default seed 5489 is 4123659995) and getRandomReal = u / (2^32 - 1);
* reftra_sw against an independent solution of the PIFM two-stream equations (matrix
  exponential of the linear ODE system, exact exponentials), plus its special cases;
* vrtqdr_sw against a dense linear solve of the adding-method interface equations, and flux
  conservation for non-absorbing layers over a white surface;
* setcoef_sw's interpolation indices against their defining bracket conditions.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest
import scipy.linalg

import oracle_bind as ob
from rrtmg_sw_amd.workloads import workload

HERE = os.path.dirname(os.path.abspath(__file__))
_dp = ob._p
KATS = np.load(os.path.join(HERE, "golden", "column_kats.npz"))


def _column(name):
    g = lambda k: KATS[f"{name}/{k}"]  # noqa: E731
    L = int(g("nlay"))
    out = np.zeros((15, 6, L + 1))
    lib = ob.orc()
    lib.orc_column_sw.restype = C.c_int
    keep = [g("pavel").copy(), g("tavel").copy(), g("pz").copy(), g("coldry").copy(),
            np.asfortranarray(g("wkl")), g("semiss").copy(), g("bndsolvar").copy(),
            g("indsolvar").copy(), np.array([float(g("solcycfrac"))]),
            np.asfortranarray(g("tauaer")), np.asfortranarray(g("ssaaer")),
            np.asfortranarray(g("asmaer"))]
    common = [
        C.c_int(L), _dp(keep[0]), _dp(keep[1]), _dp(keep[2]), _dp(keep[3]), _dp(keep[4]),
        _dp(keep[5]), C.c_double(float(g("zenith"))), C.c_int(int(g("juldat"))),
        C.c_double(float(g("scon"))), C.c_int(int(g("isolvar"))), _dp(keep[6]), _dp(keep[7]),
        _dp(keep[8]), C.c_int(int(g("iaer"))), _dp(keep[9]), _dp(keep[10]), _dp(keep[11]),
        C.c_int(int(g("idelm"))), C.c_double(1004.0)]
    if f"{name}/inflag" in KATS and int(g("imca")) == 1:
        # imca = 1: 200 McICA samples (Mersenne Twister), alpha from get_alpha for icld 4/5
        icld = int(g("icld"))
        cl = [np.ascontiguousarray(g(k), dtype=np.float64)
              for k in ("cldfrac", "ciwp", "clwp", "rei", "rel")]
        cl += [np.asfortranarray(g(k), dtype=np.float64) for k in ("tauc", "ssac", "asmc", "fsfc")]
        alpha = np.zeros(L)
        if icld in (4, 5):
            dz = np.ascontiguousarray(np.diff(g("altz")) * 1000.0)     # km -> m
            lat = np.array([float(g("lat"))])
            assert lib.orc_get_alpha(1, 1, L, icld, int(g("idcor")), C.c_double(float(g("decorr_con"))),
                                     _dp(dz), _dp(lat), int(g("juldat_alpha")), _dp(cl[0]), _dp(alpha)) == 0
            keep += [dz, lat]
        keep += cl + [alpha]
        lib.orc_column_sw_mcica.restype = C.c_int
        rc = lib.orc_column_sw_mcica(*common, C.c_int(int(g("inflag"))), C.c_int(int(g("iceflag"))),
                                     C.c_int(int(g("liqflag"))), *[_dp(a) for a in cl], C.c_int(icld),
                                     C.c_int(200), _dp(alpha), out.ctypes.data_as(ob._dp))
    elif f"{name}/inflag" in KATS:
        # imca = 0 cloudy case: cldprop_sw + spcvrt_sw on overcast layers
        cl = [np.ascontiguousarray(g(k), dtype=np.float64)
              for k in ("cldfrac", "ciwp", "clwp", "rei", "rel")]
        cl += [np.asfortranarray(g(k), dtype=np.float64) for k in ("tauc", "ssac", "asmc", "fsfc")]
        keep += cl
        lib.orc_column_sw_cld.restype = C.c_int
        rc = lib.orc_column_sw_cld(*common, C.c_int(int(g("inflag"))), C.c_int(int(g("iceflag"))),
                                   C.c_int(int(g("liqflag"))), *[_dp(a) for a in cl],
                                   out.ctypes.data_as(ob._dp))
    else:
        rc = lib.orc_column_sw(*common, out.ctypes.data_as(ob._dp))
    assert rc == 0, ob.orc().orc_last_error()
    return out, g("expected"), g("have")


@pytest.mark.parametrize("name", [str(n) for n in KATS["cases"]])
def test_column_mode_known_answers(name):
    """Every printed block (broadband and, where the case prints them, bands 16-29) of the
    reference's output matches the oracle to half a unit in the last printed digit."""
    out, exp, have = _column(name)
    assert have[0]
    for b in range(15):
        if not have[b]:
            continue
        o = out[b].T                       # (nlay+1, 6), surface first
        e = exp[b]
        flux = np.abs(o[:, :5] - e[:, :5]).max()
        hr = np.abs(o[:, 5] - e[:, 5]).max()
        assert flux <= 0.5e-4 + 2e-6, (name, b, flux)
        assert hr <= 0.5e-5 + 2e-7, (name, b, hr)


def test_column_known_answers_catch_a_changed_table():
    """The fixture check is sharp: scaling the incoming flux by 1e-5 (a solar-source error far
    below anything visible in a plot) already breaks the printed-precision match."""
    out, exp, _ = _column("MLS-clr")
    o = out[0].T
    assert np.abs(o[:, 3] * (1 + 1e-5) - exp[0][:, 3]).max() > 0.5e-4 + 2e-6


def test_mt19937_known_answer():
    lib = ob.orc()
    assert lib.orc_mt_first_int(5489, 10000) == 4123659995
    assert lib.orc_mt_first_int(5489, 1) == 3499211612   # first output of the default seed
    r = np.zeros(5)
    lib.orc_mt_reals(5489, 5, _dp(r))
    ints = [lib.orc_mt_first_int(5489, n) for n in range(1, 6)]
    assert np.array_equal(r, np.array(ints, dtype=np.float64) / (2.0 ** 32 - 1.0))


# ---------------------------------------------------------------------------------------
# reftra_sw
# ---------------------------------------------------------------------------------------
def _reftra(tau, w, g, mu):
    n = len(tau)
    chk = np.ones(n, dtype=np.int32)
    out = [np.zeros(n + 1) for _ in range(4)]
    ob.orc().orc_reftra(C.c_int(n), chk.ctypes.data_as(ob._ip), _dp(np.asarray(g, float)),
                        C.c_double(mu), _dp(np.asarray(tau, float)), _dp(np.asarray(w, float)),
                        *[_dp(o) for o in out])
    return [o[:n] for o in out]   # R, Rd, T (total, collimated), Td


def _two_stream(tau, w, g, mu):
    """Layer response of the PIFM two-stream equations (gamma1 = (8 - w(5 + 3g))/4,
    gamma2 = 3w(1 - g)/4, gamma3 = (2 - 3g mu)/4, gamma4 = 1 - gamma3), optical depth
    increasing downward, no surface:
        dU/dt = g1 U - g2 D - w g3 S(t)/mu,   dD/dt = g2 U - g1 D + w g4 S(t)/mu,
    S(t) = exp(-t/mu) the direct flux for unit incident direct flux. Solved through the
    matrix exponential of the homogeneous system plus the exponential particular solution.
    Returns R (U(0), collimated), T (D(tau) + S(tau), collimated), Rd, Td (diffuse D(0)=1)."""
    g1 = (8.0 - w * (5.0 + 3.0 * g)) / 4.0
    g2 = 3.0 * w * (1.0 - g) / 4.0
    g3 = (2.0 - 3.0 * g * mu) / 4.0
    g4 = 1.0 - g3
    A = np.array([[g1, -g2], [g2, -g1]])
    M = scipy.linalg.expm(A * tau)
    b = np.array([-w * g3 / mu, w * g4 / mu])
    # particular solution c exp(-t/mu): (-1/mu) c = A c + b
    c = np.linalg.solve(-(A + np.eye(2) / mu), b)
    e = math.exp(-tau / mu)

    def solve(y0_known_D, with_beam):
        # unknown U(0); y(0) = (U0, D0); y(tau) = M (y0 - cp) + cp e, cp = c if beam
        cp = c if with_beam else np.zeros(2)
        # U(tau) = 0:  M[0,0] U0 + M[0,1] D0 - (M cp)[0] + cp[0] e = 0
        rhs = -(M[0, 1] * y0_known_D) + (M @ cp)[0] - cp[0] * e
        U0 = rhs / M[0, 0]
        y0 = np.array([U0, y0_known_D])
        yt = M @ (y0 - cp) + cp * e
        return U0, yt[1]

    R, Tdif = solve(0.0, True)
    Rd, Td = solve(1.0, False)
    return R, Tdif + e, Rd, Td


def _delta_scaled(tau0, w0, g0):
    """Delta-scaled layer optics (f = g^2, src/rrtmg_sw_spcvmc.f90:492-496): reftra_sw
    recovers the unscaled single-scattering albedo from them to pick its branch
    (src/rrtmg_sw_reftra.f90:180-188), so test inputs are built this way."""
    f = g0 * g0
    return (1 - w0 * f) * tau0, (w0 - w0 * f) / (1 - w0 * f), (g0 - f) / (1 - f)


GRID = [_delta_scaled(tau, w, g) + (mu,) for tau in (0.1, 0.8, 4.0) for w in (0.3, 0.9, 0.995)
        for g in (0.0, 0.5, 0.85) for mu in (0.35, 0.7, 1.0)]


def test_reftra_matches_two_stream_solution():
    """The oracle uses a 10001-entry exponential table and a quadratic expansion below 0.06,
    so it agrees with the exact-exponential solution to about 1e-3; a wrong sign, a dropped
    term or a swapped gamma shows up as an error of order 0.01-1."""
    worst = 0.0
    for tau, w, g, mu in GRID:
        k = math.sqrt(((8 - w * (5 + 3 * g)) / 4) ** 2 - (3 * w * (1 - g) / 4) ** 2)
        if abs(1 - k * mu) < 0.05:
            continue   # the closed form's removable singularity k mu = 1 (guarded in the reference)
        R, Rd, T, Td = [v[0] for v in _reftra([tau], [w], [g], mu)]
        eR, eT, eRd, eTd = _two_stream(tau, w, g, mu)
        d = max(abs(R - eR), abs(T - eT), abs(Rd - eRd), abs(Td - eTd))
        worst = max(worst, d)
        assert d < 2e-3, (tau, w, g, mu, (R, eR), (T, eT), (Rd, eRd), (Td, eTd))
    assert worst > 0.0


def test_reftra_exact_exponential_matches_two_stream_solution():
    """In its exact-exponential mode (the forward the adjoint differentiates) the oracle's reftra
    is the PIFM two-stream solution itself: agreement to rounding (1e-11) instead of the table's
    1e-3, on the general branch (away from the removable singularity k mu = 1) and on the
    conservative one (w = 1, checked through R + T = 1 and the exact solution)."""
    ob.orc().orc_set_exact_exp(1)
    try:
        worst = 0.0
        for tau, w, g, mu in GRID + [(t, 1.0, g, mu) for t in (0.1, 2.0) for g in (0.0, 0.7)
                                     for mu in (0.4, 1.0)]:
            k = math.sqrt(max(((8 - w * (5 + 3 * g)) / 4) ** 2 - (3 * w * (1 - g) / 4) ** 2, 0.0))
            if w < 1.0 and abs(1 - k * mu) < 0.05:
                continue
            R, Rd, T, Td = [v[0] for v in _reftra([tau], [w], [g], mu)]
            eR, eT, eRd, eTd = _two_stream(tau, w, g, mu)
            d = max(abs(R - eR), abs(T - eT), abs(Rd - eRd), abs(Td - eTd))
            worst = max(worst, d)
            assert d < 1e-11, (tau, w, g, mu, (R, eR), (T, eT), (Rd, eRd), (Td, eTd))
    finally:
        ob.orc().orc_set_exact_exp(0)
    assert worst > 0.0


def _two_stream_beam(tau, w, g, mu):
    """The same layer response from the system augmented with the direct beam, y = (U, D, S),
    dS/dt = -S/mu: one 3 x 3 matrix exponential, with no particular-solution denominator, so it
    holds at the resonance k mu = 1 too. Returns R, T (collimated, T = D(tau) + S(tau))."""
    g1 = (8.0 - w * (5.0 + 3.0 * g)) / 4.0
    g2 = 3.0 * w * (1.0 - g) / 4.0
    g3 = (2.0 - 3.0 * g * mu) / 4.0
    g4 = 1.0 - g3
    A = np.array([[g1, -g2, -w * g3 / mu], [g2, -g1, w * g4 / mu], [0.0, 0.0, -1.0 / mu]])
    M = scipy.linalg.expm(A * tau)
    U0 = -M[0, 2] / M[0, 0]   # U(tau) = 0 with D(0) = 0, S(0) = 1
    y = M @ np.array([U0, 0.0, 1.0])
    return U0, y[1] + y[2]


def _closed_form_mp(tau, w, g, mu, digits=60):
    """The reference's closed-form general solution (src/rrtmg_sw_reftra.f90:233-322: zr1..zr5,
    zt1..zt3, zdenr) in 60-digit arithmetic: its cancellation near k mu = 1 then costs digits the
    result does not need. Returns R, T."""
    import mpmath as mp
    mp.mp.dps = digits
    tau, w, g, mu = mp.mpf(tau), mp.mpf(w), mp.mpf(g), mp.mpf(mu)
    g1 = (8 - w * (5 + 3 * g)) / 4
    g2 = 3 * w * (1 - g) / 4
    g3 = (2 - 3 * g * mu) / 4
    g4 = 1 - g3
    a1, a2 = g1 * g4 + g2 * g3, g1 * g3 + g2 * g4
    k = mp.sqrt(g1 * g1 - g2 * g2)
    rp = k * mu
    r1, r2, r3 = (1 - rp) * (a2 + k * g3), (1 + rp) * (a2 - k * g3), 2 * k * (g3 - a2 * mu)
    r4, r5 = (1 - rp * rp) * (k + g1), (1 - rp * rp) * (k - g1)
    t1, t2, t3 = (1 + rp) * (a1 + k * g4), (1 - rp) * (a1 - k * g4), 2 * k * (g4 + a1 * mu)
    em1, em2 = mp.exp(-k * tau), mp.exp(-tau / mu)
    den = r4 / em1 + r5 * em1
    R = w * (r1 / em1 - r2 * em1 - r3 * em2) / den
    T = em2 - em2 * w * (t1 / em1 - t2 * em1 - t3 / em2) / den
    return float(R), float(T)


def test_reftra_exact_exponential_through_the_resonance():
    """At k mu0 = 1 the closed form's particular solution divides by 1 - (k mu0)^2, which vanishes
    with its numerators. The exact-exponential mode evaluates the general solution there in a
    form without the cancellation (|1 - k mu0| < 1e-3), so it is the two-stream solution to
    rounding at the resonance, next to it, and on both sides of the window's edge: against the
    augmented matrix exponential (an independent solution of the ODEs, for layers it resolves,
    k tau <= 8) and against the closed form in 60-digit arithmetic (every thickness; at the
    resonance itself 1e-30 from it). The reference's expressions in FP64 lose about
    1e-16 / |1 - (k mu0)^2| there (1e-8 at 1e-9 from the resonance, 2e-11 at 1e-6) and return
    R = 1e-8 where the denominator vanishes."""
    ob.orc().orc_set_exact_exp(1)
    try:
        worst, n = 0.0, 0
        for w0, g0 in ((0.5, 0.5), (0.2, 0.3), (0.4, 0.6)):
            _, w, g = _delta_scaled(1.0, w0, g0)
            g1 = (8 - w * (5 + 3 * g)) / 4
            k = math.sqrt(g1 ** 2 - (3 * w * (1 - g) / 4) ** 2)
            for tau in (1e-3, 0.1, 1.0, 4.0, 30.0, 250.0):
                for d in (0.0, 1e-12, -1e-12, 1e-9, -1e-9, 1e-6, -1e-6, 1e-4, -1e-4,
                          1e-3 - 1e-9, -1e-3 + 1e-9, 1e-3 + 1e-9, -1e-3 - 1e-9, 1e-2, -1e-2):
                    mu = (1.0 + d) / k
                    R, _, T, _ = [v[0] for v in _reftra([tau], [w], [g], mu)]
                    refs = [_closed_form_mp(tau, w, g, mu if d != 0.0 else (1.0 + 1e-30) / k)]
                    if k * tau <= 8.0:
                        refs.append(_two_stream_beam(tau, w, g, mu))
                    for eR, eT in refs:
                        err = max(abs(R - eR), abs(T - eT))
                        worst = max(worst, err)
                        n += 1
                        assert err < 1e-11, (w, g, tau, d, (R, eR), (T, eT))
    finally:
        ob.orc().orc_set_exact_exp(0)
    assert n == 3 * 15 * (6 + 4) and worst > 0.0


def test_exact_exponential_forward_near_table_forward():
    """End to end, the exact-exponential forward differs from the reference's table forward only
    by the table's quantisation: under 5 W/m2, and not at all where the mode is off again."""
    cols = workload("expran4096x60", ncol=8)
    mc, _ = ob.run_mcica(cols)
    a = ob.run_rrtmg(cols, mc)
    ob.orc().orc_set_exact_exp(1)
    try:
        b = ob.run_rrtmg(cols, mc)
    finally:
        ob.orc().orc_set_exact_exp(0)
    c = ob.run_rrtmg(cols, mc)
    d = max(float(np.abs(a[k] - b[k]).max()) for k in ("swuflx", "swdflx"))
    assert 1e-6 < d < 5.0, d
    for k in ("swuflx", "swdflx", "swhr"):
        assert np.array_equal(a[k], c[k]), k


def test_reftra_special_cases():
    # conservative scattering (w = 1): R + T = 1 for both incidences (src/rrtmg_sw_reftra.f90:213,218)
    for tau, g, mu in [(0.5, 0.0, 0.5), (3.0, 0.8, 0.9), (20.0, 0.6, 0.2)]:
        R, Rd, T, Td = [v[0] for v in _reftra([tau], [1.0], [g], mu)]
        assert R + T == pytest.approx(1.0, abs=1e-15)
        assert Rd + Td == pytest.approx(1.0, abs=1e-15)
        eR, eT, eRd, eTd = _two_stream(tau, 1.0 - 1e-12, g, mu)
        assert abs(R - eR) < 2e-3 and abs(Rd - eRd) < 2e-3
    # vanishing optical depth: R = 0, T = 1
    R, Rd, T, Td = [v[0] for v in _reftra([1e-12], [0.5], [0.3], 0.6)]
    assert abs(R) < 1e-10 and abs(T - 1) < 1e-10 and abs(Rd) < 1e-10 and abs(Td - 1) < 1e-10
    # absorbing layers: 0 <= R, T and R + T <= 1
    for tau, w, g, mu in GRID:
        R, Rd, T, Td = [v[0] for v in _reftra([tau], [w], [g], mu)]
        assert -1e-12 <= R and -1e-12 <= T and R + T <= 1 + 1e-12
        assert -1e-12 <= Rd and -1e-12 <= Td and Rd + Td <= 1 + 1e-12


# ---------------------------------------------------------------------------------------
# vrtqdr_sw
# ---------------------------------------------------------------------------------------
def _vrtqdr(R, Rd, T, Td, dbt, alb_dir, alb_dif):
    L = len(R)
    n1 = L + 1
    pref, prefd = np.append(R, alb_dir), np.append(Rd, alb_dif)
    ptra, ptrad = np.append(T, 0.0), np.append(Td, 0.0)
    pdbt = np.append(dbt, 0.0)
    ptdbt = np.ones(n1)
    for k in range(L):
        ptdbt[k + 1] = ptdbt[k] * dbt[k]
    prdnd, prup, prupd = np.zeros(n1), pref.copy(), prefd.copy()
    pfd, pfu = np.zeros(n1), np.zeros(n1)
    ob.orc().orc_vrtqdr(C.c_int(L), _dp(pref), _dp(prefd), _dp(ptra), _dp(ptrad), _dp(pdbt),
                        _dp(prdnd), _dp(prup), _dp(prupd), _dp(ptdbt), _dp(pfd), _dp(pfu))
    return pfd, pfu, ptdbt


def _dense(R, Rd, T, Td, dbt, alb_dir, alb_dif):
    """Interface equations of a stack of layers (level 0 = top) lit by a unit direct beam:
    S_{k+1} = S_k dbt_k; Dtot_{k+1} = S_k T_k + (Dtot_k - S_k) Td_k + U_{k+1} Rd_k;
    U_k = S_k R_k + (Dtot_k - S_k) Rd_k + U_{k+1} Td_k; U_L = S_L a_dir + (Dtot_L - S_L) a_dif;
    Dtot_0 = 1. Unknowns Dtot_0..L, U_0..L, solved as one dense system."""
    L = len(R)
    S = np.ones(L + 1)
    for k in range(L):
        S[k + 1] = S[k] * dbt[k]
    n = 2 * (L + 1)
    iD = lambda k: k          # noqa: E731
    iU = lambda k: L + 1 + k  # noqa: E731
    A = np.zeros((n, n))
    b = np.zeros(n)
    r = 0
    A[r, iD(0)] = 1.0
    b[r] = 1.0
    r += 1
    for k in range(L):
        A[r, iD(k + 1)] = 1.0
        A[r, iD(k)] = -Td[k]
        A[r, iU(k + 1)] = -Rd[k]
        b[r] = S[k] * T[k] - S[k] * Td[k]
        r += 1
        A[r, iU(k)] = 1.0
        A[r, iD(k)] = -Rd[k]
        A[r, iU(k + 1)] = -Td[k]
        b[r] = S[k] * R[k] - S[k] * Rd[k]
        r += 1
    A[r, iU(L)] = 1.0
    A[r, iD(L)] = -alb_dif
    b[r] = S[L] * alb_dir - S[L] * alb_dif
    x = np.linalg.solve(A, b)
    return x[:L + 1], x[L + 1:]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_vrtqdr_matches_dense_solve(seed):
    rng = np.random.default_rng(seed)
    L = 6
    tau = rng.uniform(0.01, 3.0, L)
    w = rng.uniform(0.2, 0.999, L)
    g = rng.uniform(0.0, 0.9, L)
    mu = 0.6
    R, Rd, T, Td = _reftra(tau, w, g, mu)
    dbt = np.exp(-tau / mu)
    pfd, pfu, _ = _vrtqdr(R, Rd, T, Td, dbt, 0.3, 0.25)
    D, U = _dense(R, Rd, T, Td, dbt, 0.3, 0.25)
    assert np.allclose(pfd, D, rtol=1e-12, atol=1e-14)
    assert np.allclose(pfu, U, rtol=1e-12, atol=1e-14)


def test_vrtqdr_conserves_flux_without_absorption():
    """Non-absorbing layers (R + T = 1, Rd + Td = 1) over a white surface: the net flux is zero
    at every level."""
    L = 5
    tau = np.array([0.2, 1.0, 3.0, 0.05, 7.0])
    R, Rd, T, Td = _reftra(tau, np.ones(L), np.full(L, 0.7), 0.5)
    dbt = np.exp(-tau / 0.5)
    pfd, pfu, _ = _vrtqdr(R, Rd, T, Td, dbt, 1.0, 1.0)
    assert np.abs(pfd - pfu).max() < 1e-12


# ---------------------------------------------------------------------------------------
# setcoef_sw
# ---------------------------------------------------------------------------------------
PREFLOG = np.array([6.96 - 0.2 * i for i in range(59)])


def test_setcoef_indices_bracket_their_definitions():
    """jp brackets ln p between reference levels (fp in [0, 1] where not clamped), jt/jt1 bracket
    T in 15 K steps around the reference temperatures (ft in [0, 1) where not clamped), the
    interpolation weights fac00..fac11 sum to one, and laytrop counts ln p > 4.56."""
    p = np.exp(np.linspace(math.log(1050.0), math.log(0.02), 57))
    L = len(p)
    lib = ob.orc()
    for T0 in (180.0, 215.0, 250.0, 290.0, 330.0):
        t = np.full(L, T0) + np.linspace(-5, 5, L)
        coldry = np.full(L, 2e23)
        wkl = np.zeros((39, L), order="F")
        wkl[0] = coldry * 1e-3
        wkl[1] = coldry * 4e-4
        wkl[2] = coldry * 1e-6
        wkl[5] = coldry * 1.8e-6
        wkl[6] = coldry * 0.209
        iout = np.zeros(3 + 5 * L, dtype=np.int32)
        dout = np.zeros(17 * L)
        lib.orc_setcoef(C.c_int(L), _dp(p), _dp(t), _dp(coldry), _dp(wkl),
                        iout.ctypes.data_as(ob._ip), _dp(dout))
        laytrop = iout[0]
        jp, jt, jt1 = iout[3:3 + L], iout[3 + L:3 + 2 * L], iout[3 + 2 * L:3 + 3 * L]
        fac = dout.reshape(17, L)
        plog = np.log(p)
        assert laytrop == int((plog > 4.56).sum())
        for k in range(L):
            fp = 5.0 * (PREFLOG[jp[k] - 1] - plog[k])
            if 1 < jp[k] < 58:
                assert -1e-9 <= fp <= 1 + 1e-9, (k, jp[k], fp)
            assert abs(fac[8, k] + fac[9, k] + fac[10, k] + fac[11, k] - 1.0) < 1e-12 or not (1 < jp[k] < 58)
            assert 1 <= jt[k] <= 4 and 1 <= jt1[k] <= 4
