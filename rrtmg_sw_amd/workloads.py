"""Synthetic input columns for the GCM-mode shortwave path, drawn from one seed.

Every array is returned in the reference's Fortran layout (column-major numpy arrays, column
index fastest): profiles (ncol, nlay), levels (ncol, nlay+1), per-band cloud inputs
(nbndsw, ncol, nlay), aerosol (ncol, nlay, nbndsw) -- the dummies of ``rrtmg_sw`` and
``mcica_subcol_sw`` (src/rrtmg_sw_rad.f90:220-388, src/mcica_subcol_gen_sw.f90:200-246).

Named workloads:
  clear64x60     64 clear-sky columns x 60 layers, FP64 reference case (icld=0, iaer=0).
  mcica8192x72   8192 columns x 72 layers, McICA maximum-random overlap (icld=2) with the
                 KISS generator; cloud fraction U(0,1) on 30% of the layers between 900 and
                 200 hPa; LWP lognormal (median 50 g/m2, sigma 1.0) where T > 263 K, IWP
                 lognormal (median 20 g/m2, sigma 1.2) elsewhere; liquid radius U(5,20) um,
                 ice generalized effective size U(20,120) um.
  expran4096x60  4096 columns x 60 layers, exponential-random overlap (icld=5, decorrelation
                 length U(1,4) km per column) plus aerosols (iaer=10) below 3 km.
  expran65536x137  as expran4096x60 at 65536 columns x 137 layers.
  overcast8192x72  the columns of mcica8192x72 for the non-McICA driver, which takes clear and
                 overcast layers only: every cloudy layer has cloud fraction 1.

Choices the workload descriptions leave open, applied identically to every implementation:
level pressures log-spaced from the surface value to 0.01 hPa, layer pressure the geometric
mean of its levels, heights from the hypsometric equation (Rd = 287.04, g = 9.80665) with a
6.5 K/km lapse rate from a surface temperature U(250, 305) K clamped at 190 K, H2O falling
with a 2 km scale height from U(1e-3, 3e-2), O3 a Gaussian in ln p of peak 8e-6 at 10 hPa
(sigma 1 in ln p), one albedo U(0.05, 0.8) per column for all four albedos, cos(SZA)
U(0.1, 1), scon = 1361 W/m2 with isolvar = 0. Cloudy workloads jitter the surface pressure by
1e-4 relative so that every column seeds its own KISS stream (the generator seeds from the
fractional Pa of the bottom four layers, src/mcica_subcol_gen_sw.f90:471-479).
"""
from dataclasses import dataclass, field

import numpy as np

NGPT = 112
NBND = 14
NAER = 6
RD, G0 = 287.04, 9.80665
CPDAIR = 1004.64

WORKLOADS = {
    "clear64x60": dict(ncol=64, nlay=60, icld=0, iaer=0, seed=0),
    "mcica8192x72": dict(ncol=8192, nlay=72, icld=2, iaer=0, seed=1),
    "expran4096x60": dict(ncol=4096, nlay=60, icld=5, iaer=10, seed=3),
    "expran65536x137": dict(ncol=65536, nlay=137, icld=5, iaer=10, seed=4),
    "overcast8192x72": dict(ncol=8192, nlay=72, icld=2, iaer=0, seed=1, overcast=True),
}


@dataclass
class Columns:
    """Inputs of mcica_subcol_sw + rrtmg_sw for ncol columns (Fortran-ordered arrays)."""
    ncol: int
    nlay: int
    icld: int
    iaer: int
    play: np.ndarray
    plev: np.ndarray
    tlay: np.ndarray
    tlev: np.ndarray
    tsfc: np.ndarray
    h2ovmr: np.ndarray
    o3vmr: np.ndarray
    co2vmr: np.ndarray
    ch4vmr: np.ndarray
    n2ovmr: np.ndarray
    o2vmr: np.ndarray
    asdir: np.ndarray
    asdif: np.ndarray
    aldir: np.ndarray
    aldif: np.ndarray
    coszen: np.ndarray
    adjes: float = 1.0
    dyofyr: int = 0
    scon: float = 1361.0
    isolvar: int = 0
    inflgsw: int = 2
    iceflgsw: int = 3
    liqflgsw: int = 1
    # cloud physical inputs of mcica_subcol_sw
    cldfrac: np.ndarray = None
    ciwp: np.ndarray = None
    clwp: np.ndarray = None
    rei: np.ndarray = None
    rel: np.ndarray = None
    tauc: np.ndarray = None
    ssac: np.ndarray = None
    asmc: np.ndarray = None
    fsfc: np.ndarray = None
    alpha: np.ndarray = None
    dz: np.ndarray = None
    decorr_len: np.ndarray = None
    permuteseed: int = 112
    irng: int = 0
    # aerosol
    tauaer: np.ndarray = None
    ssaaer: np.ndarray = None
    asmaer: np.ndarray = None
    ecaer: np.ndarray = None
    extra: dict = field(default_factory=dict)

    def subset(self, cols):
        """Columns `cols` (a slice or index array) as a new Columns."""
        out = {}
        for k, v in self.__dict__.items():
            if isinstance(v, np.ndarray):
                if k in ("tauc", "ssac", "asmc", "fsfc"):
                    out[k] = np.asfortranarray(v[:, cols])
                else:
                    out[k] = np.asfortranarray(v[cols])
            else:
                out[k] = v
        n = out["play"].shape[0]
        out["ncol"] = n
        return Columns(**out)


def _f(a):
    return np.asfortranarray(a, dtype=np.float64)


def make_columns(ncol, nlay, icld=0, iaer=0, seed=0, psfc=1013.0, ptop=0.01, jitter=None,
                 decorr=(1000.0, 4000.0), permuteseed=112, irng=0):
    """Draw `ncol` columns of `nlay` layers. Draw order (one numpy Generator(seed)):
    T0, q0, albedo, coszen, [surface-pressure jitter], [cloud mask, cloud fraction,
    LWP, IWP, rel, rei], [decorrelation length], [AOD, SSA, g]."""
    rng = np.random.default_rng(seed)
    C, L = ncol, nlay
    T0 = rng.uniform(250.0, 305.0, C)
    q0 = rng.uniform(1e-3, 3e-2, C)
    alb = rng.uniform(0.05, 0.8, C)
    mu0 = rng.uniform(0.1, 1.0, C)
    cloudy = icld != 0
    if jitter is None:
        jitter = cloudy
    ps = np.full(C, psfc)
    if jitter:
        ps = psfc * (1.0 + 1e-4 * rng.uniform(-1.0, 1.0, C))
    k = np.arange(L + 1)[None, :]
    lnp = np.log(ps)[:, None] + k / L * (np.log(ptop) - np.log(ps)[:, None])
    plev = np.exp(lnp)
    play = np.sqrt(plev[:, :-1] * plev[:, 1:])
    # heights: integrate the hypsometric equation level by level with the lapse-rate T
    zlev = np.zeros((C, L + 1))
    tlev = np.zeros((C, L + 1))
    tlev[:, 0] = T0
    for j in range(L):
        tl = tlev[:, j]
        zlev[:, j + 1] = zlev[:, j] + RD * tl / G0 * np.log(plev[:, j] / plev[:, j + 1])
        tlev[:, j + 1] = np.maximum(T0 - 6.5e-3 * zlev[:, j + 1], 190.0)
    zlay = zlev[:, :-1] + RD * tlev[:, :-1] / G0 * np.log(plev[:, :-1] / play)
    tlay = np.maximum(T0[:, None] - 6.5e-3 * zlay, 190.0)
    dz = zlev[:, 1:] - zlev[:, :-1]
    h2o = q0[:, None] * np.exp(-zlay / 2000.0)
    o3 = 8e-6 * np.exp(-0.5 * np.log(play / 10.0) ** 2)
    ones = np.ones((C, L), order="F")
    cols = Columns(
        ncol=C, nlay=L, icld=icld, iaer=iaer,
        play=_f(play), plev=_f(plev), tlay=_f(tlay), tlev=_f(tlev), tsfc=_f(T0),
        h2ovmr=_f(h2o), o3vmr=_f(o3), co2vmr=_f(4e-4 * ones), ch4vmr=_f(1.8e-6 * ones),
        n2ovmr=_f(3.2e-7 * ones), o2vmr=_f(0.209 * ones),
        asdir=_f(alb), asdif=_f(alb), aldir=_f(alb), aldif=_f(alb), coszen=_f(mu0),
        permuteseed=permuteseed, irng=irng,
    )
    cols.dz = _f(dz)
    cldfrac = np.zeros((C, L))
    clwp = np.zeros((C, L))
    ciwp = np.zeros((C, L))
    rel = np.zeros((C, L))
    rei = np.zeros((C, L))
    if cloudy:
        band = (play >= 200.0) & (play <= 900.0)
        mask = band & (rng.uniform(size=(C, L)) < 0.3)
        frac = rng.uniform(0.0, 1.0, (C, L))
        cldfrac = np.where(mask, frac, 0.0)
        lwp = rng.lognormal(np.log(50.0), 1.0, (C, L))
        iwp = rng.lognormal(np.log(20.0), 1.2, (C, L))
        warm = tlay > 263.0
        clwp = np.where(mask & warm, lwp, 0.0)
        ciwp = np.where(mask & ~warm, iwp, 0.0)
        rel_d = rng.uniform(5.0, 20.0, (C, L))
        rei_d = rng.uniform(20.0, 120.0, (C, L))
        rel = np.where(mask, rel_d, 10.0)
        rei = np.where(mask, rei_d, 60.0)
    cols.cldfrac, cols.clwp, cols.ciwp = _f(cldfrac), _f(clwp), _f(ciwp)
    cols.rel, cols.rei = _f(rel), _f(rei)
    cols.tauc = np.zeros((NBND, C, L), order="F")
    cols.ssac = np.ones((NBND, C, L), order="F")
    cols.asmc = np.zeros((NBND, C, L), order="F")
    cols.fsfc = np.zeros((NBND, C, L), order="F")
    cols.alpha = np.zeros((C, L), order="F")
    if icld in (4, 5):
        cols.decorr_len = _f(rng.uniform(decorr[0], decorr[1], C))
        cols.alpha = _f(alpha_for(cols))
    # (ncol, nlay, nbnd) arrays are built as C-ordered (nbnd, nlay, ncol) and transposed: the
    # Fortran layout without a copy
    tauaer = np.zeros((C, L, NBND), order="F")
    ssaaer = np.ones((C, L, NBND), order="F")
    asmaer = np.zeros((C, L, NBND), order="F")
    if iaer == 10:
        aod = rng.lognormal(np.log(0.1), 0.5, (C, NBND))
        ssa = rng.uniform(0.85, 0.99, (C, NBND))
        g = rng.uniform(0.6, 0.75, (C, NBND))
        low = zlay < 3000.0
        w = np.where(low, dz, 0.0)
        w = w / np.maximum(w.sum(axis=1, keepdims=True), 1e-30)
        tauaer = np.multiply(w.T[None, :, :], aod.T[:, None, :], order="C").T
        ssaaer = np.ascontiguousarray(np.broadcast_to(ssa.T[:, None, :], (NBND, L, C))).T
        asmaer = np.ascontiguousarray(np.broadcast_to(g.T[:, None, :], (NBND, L, C))).T
    cols.tauaer, cols.ssaaer, cols.asmaer = tauaer, ssaaer, asmaer
    cols.ecaer = np.zeros((C, L, NAER), order="F")
    return cols


def alpha_for(cols):
    """get_alpha (src/mcica_subcol_gen_sw.f90:66-178) with a per-column decorrelation length
    (idcor = 0): the reference takes one scalar length per call, so this is the result of one
    call per column. Kept on the host, as get_alpha is host-side in the reference API."""
    C, L = cols.ncol, cols.nlay
    inv = 1.0 / cols.decorr_len
    a = np.zeros((C, L))
    a[:, 1:] = np.exp(-(0.5 * (cols.dz[:, 1:] + cols.dz[:, :-1])) * inv[:, None])
    if cols.icld == 5:
        cf = cols.cldfrac
        a[:, 1:] = np.where((cf[:, 1:] == 0.0) & (cf[:, :-1] > 0.0), 0.0, a[:, 1:])
    return a


def overcast(cols):
    """The non-McICA driver's clouds from these columns: src/rrtmg_sw_rad.nomcica.f90:617-622
    stops on a cloud fraction between 1e-6 and 1 - 1e-6, so every layer with a cloud fraction
    above zero becomes overcast (fraction 1); water paths and radii are kept."""
    out = Columns(**{k: v for k, v in cols.__dict__.items()})
    out.cldfrac = _f(np.where(cols.cldfrac > 0.0, 1.0, 0.0))
    return out


def extreme(seed=95):
    """Eight 60-layer McICA + aerosol columns up to 0.001 hPa at the edges of the tables:
    temperatures below and above the reference grid (150 K, 350 K, a 140-360 K ramp, 330-350 K),
    water vapour at 1e-12, 0.06 and exactly 0, ozone at 0 and 1e-4."""
    cols = make_columns(8, 60, icld=2, iaer=10, seed=seed, ptop=0.001)
    r = np.random.default_rng(seed)
    t = cols.tlay.copy()
    t[0], t[1] = 150.0, 350.0
    t[2] = np.linspace(140.0, 360.0, 60)
    t[3] = 330.0 + 20.0 * r.uniform(size=60)
    q = cols.h2ovmr.copy()
    q[4], q[5], q[6] = 1e-12, 0.06, 0.0
    o = cols.o3vmr.copy()
    o[6], o[7] = 0.0, 1e-4
    cols.tlay, cols.h2ovmr, cols.o3vmr = _f(t), _f(q), _f(o)
    return cols


def with_cloud_optics(cols, inflag, iceflag, seed=0):
    """The columns' clouds under another cloud-optics option: ice sizes clipped into the bounds of
    iceflag 1 (13-130 um), 2 (5-131 um) or 3 (5-140 um); for inflag 0 per-band optical depth,
    single-scattering albedo, asymmetry and forward fraction (asymmetry squared) drawn for every
    cloudy layer, as a GCM passing its own cloud optics would."""
    cols.inflgsw, cols.iceflgsw = inflag, iceflag
    lo, hi = {1: (13.0, 130.0), 2: (5.0, 131.0), 3: (5.0, 140.0)}[iceflag]
    cols.rei = _f(np.clip(cols.rei, lo, hi))
    if inflag == 0:
        r = np.random.default_rng(seed)
        cl = (cols.cldfrac > 0)[None]
        cols.tauc = _f(np.where(cl, r.uniform(0.1, 20.0, cols.tauc.shape), 0.0))
        cols.ssac = _f(np.where(cl, r.uniform(0.8, 1.0, cols.ssac.shape), 1.0))
        cols.asmc = _f(np.where(cl, r.uniform(0.7, 0.9, cols.asmc.shape), 0.0))
        cols.fsfc = _f(cols.asmc ** 2)
    return cols


def workload(name, ncol=None):
    """Columns of a named workload (optionally only its first `ncol` columns)."""
    w = dict(WORKLOADS[name])
    cols = make_columns(w["ncol"], w["nlay"], icld=w["icld"], iaer=w["iaer"], seed=w["seed"])
    if w.get("overcast"):
        cols = overcast(cols)
    if ncol is not None and ncol < cols.ncol:
        cols = cols.subset(slice(0, ncol))
    return cols
