"""Write tests/golden/column_kats.npz from the reference's own column-mode example suite.

The reference ships eleven clear-sky single-column cases with their outputs
(run_examples_std_atm/input_rrtm_*-clr*, output_rrtm_*-clr*, in_aer_rrtm-aer12; run by
run_examples_std_atm/script.run_std_atm); the six whose layers are given directly (IATM = 0)
are kept. This script reads those files as data: the
INPUT_RRTM records the column program reads (src/rrtmg_sw.1col.f90:1364-1766, formats
:1870-1877 and :1321-1326), the IN_AER_RRTM records for iaer = 10 (readaer, :1985-2127), and
every flux block of each OUTPUT_RRTM, and copies each case's input files (data) to
tests/golden/column_inputs/<case>/ under the names the column program reads. It stores, per case, the layer quantities the column
program hands to setcoef_sw / spcvmc_sw (pavel, tavel, pz, coldry, wkl after the
mixing-ratio conversion of :1733-1752, semiss, zenith, solar-variability inputs, aerosol
optics) and the expected outputs (broadband and per-band up, diffuse down, direct down, down,
net flux and heating rate at every level).

    python tests/golden/make_column_kats.py [/root/reference]
"""
import math
import os
import sys

import numpy as np

# The five ...-sza45-isolvar* cases describe their atmosphere with IATM = 1 (layers built by
# the LBLRTM-style generator src/rrtatm.f from a model atmosphere); they are not restated
# here, so only the six IATM = 0 cases (layer amounts given directly) become fixtures.
# The one cloudy case without McICA (imca = 0: cldprop_sw + spcvrt_sw, overcast layers only)
# is kept too; its input file and cloud file are named apart from its output file.
# The four McICA cases (imca = 1: 200 samples averaged) are kept as well; the input and cloud
# files of each output follow run_examples_std_atm/script.run_std_atm.
CASES = ["MLS-clr", "MLS-clr-aer12", "MLS-clr-sza65", "MLW-clr", "SAW-clr", "TROP-clr",
         "MLS-cld6-imca0-icld2", "MLS-cld5-imca1-icld2", "MLS-cld7-imca1-icld2",
         "MLS-cld7-imca1-icld5-idcor0", "MLS-cld7-imca1-icld5-idcor1"]
AER = {"MLS-clr-aer12": "in_aer_rrtm-aer12"}
INPUT = {"MLS-cld6-imca0-icld2": "input_rrtm_MLS-cld-imca0-icld2",
         "MLS-cld5-imca1-icld2": "input_rrtm_MLS-cld-imca1-icld2",
         "MLS-cld7-imca1-icld2": "input_rrtm_MLS-cld-imca1-icld2",
         "MLS-cld7-imca1-icld5-idcor0": "input_rrtm_MLS-cld-imca1-icld5-idcor0",
         "MLS-cld7-imca1-icld5-idcor1": "input_rrtm_MLS-cld-imca1-icld5-idcor1"}
CLD = {"MLS-cld6-imca0-icld2": "in_cld_rrtm-cld6", "MLS-cld5-imca1-icld2": "in_cld_rrtm-cld5",
       "MLS-cld7-imca1-icld2": "in_cld_rrtm-cld7", "MLS-cld7-imca1-icld5-idcor0": "in_cld_rrtm-cld7",
       "MLS-cld7-imca1-icld5-idcor1": "in_cld_rrtm-cld7"}
IB1, IB2 = 16, 29


def fields(line, spec):
    """Fixed-column read: spec is a list of (kind, width) with kind 'x' (skip), 'i', 'f', 'a'.
    Blank numeric fields read as 0, as Fortran formatted input does."""
    out, pos = [], 0
    line = line.rstrip("\n")
    for kind, w in spec:
        txt = line[pos:pos + w] if pos < len(line) else ""
        pos += w
        if kind == "x":
            continue
        t = txt.strip()
        if kind == "a":
            out.append(txt)
        elif kind == "i":
            out.append(int(t) if t else 0)
        else:
            out.append(float(t) if t else 0.0)
    return out


def read_input(path):
    lines = open(path).read().split("\n")
    i = next(k for k, l in enumerate(lines) if l.startswith("$"))
    # 9011 (18x,i2,29x,i1,32x,i1,1x,i1,2x,i3,3x,i1,i1,3x,i1,i1)
    iaer, iatm, iscat, istrm, iout, imca, icld, idelm, icos = fields(
        lines[i + 1], [("x", 18), ("i", 2), ("x", 29), ("i", 1), ("x", 32), ("i", 1), ("x", 1),
                       ("i", 1), ("x", 2), ("i", 3), ("x", 3), ("i", 1), ("i", 1), ("x", 3),
                       ("i", 1), ("i", 1)])
    assert iatm == 0, path
    # 9020 (12x, i3, 3x, f7.4, 3x, i2, f10.4, f10.5, 14f10.5)
    v = fields(lines[i + 2], [("x", 12), ("i", 3), ("x", 3), ("f", 7), ("x", 3), ("i", 2),
                              ("f", 10), ("f", 10)] + [("f", 10)] * 14)
    juldat, sza, isolvar, scon, solcycfrac = v[:5]
    solvar = np.array(v[5:], dtype=np.float64)
    # 9012 (11x,i1,2x,i1,14f5.3)
    v = fields(lines[i + 3], [("x", 11), ("i", 1), ("x", 2), ("i", 1)] + [("f", 5)] * 14)
    iemis, semiss = v[0], np.array(v[2:], dtype=np.float64)
    if iemis == 0:
        semiss[:] = 1.0
    elif iemis == 1:
        semiss[:] = semiss[0]
    k = i + 4
    idcor, decorr_con, lat = 0, 0.0, 0.0
    if icld in (4, 5):      # 9014 (8x,i2), then 9015 (f10.3) decorrelation length or latitude
        idcor = fields(lines[k], [("x", 8), ("i", 2)])[0]
        val = fields(lines[k + 1], [("f", 10)])[0]
        decorr_con, lat = (val, 0.0) if idcor == 0 else (0.0, val)
        k += 2
    # the decorrelation length of idcor = 1 takes day 94 for juldat = 0 (:1692), after the
    # Earth-Sun factor was computed from juldat itself
    juldat_alpha = 94 if (juldat == 0 and idcor == 1) else juldat
    # 9013 (1x,i1,i3,i5)
    iform, nlay, nmol = fields(lines[k], [("x", 1), ("i", 1), ("i", 3), ("i", 5)])
    nmol = nmol or 7
    assert iform == 1 and nmol == 7, path
    f1 = [("f", 15), ("f", 10), ("f", 10), ("a", 3), ("i", 2), ("x", 1),
          ("f", 7), ("f", 8), ("f", 7), ("f", 7), ("f", 8), ("f", 7)]
    f2 = [("f", 15), ("f", 10), ("f", 10), ("a", 3), ("i", 2), ("x", 23), ("f", 7), ("f", 8), ("f", 7)]
    f3 = [("f", 15)] * 8
    pavel, tavel, pz, altz = np.zeros(nlay), np.zeros(nlay), np.zeros(nlay + 1), np.zeros(nlay + 1)
    wkl, wbrodl = np.zeros((7, nlay)), np.zeros(nlay)
    k = k + 1
    for l in range(nlay):
        if l == 0:
            v = fields(lines[k], f1)
            pavel[0], tavel[0] = v[0], v[1]
            altz[0], pz[0], altz[1], pz[1] = v[5], v[6], v[8], v[9]
        else:
            v = fields(lines[k], f2)
            pavel[l], tavel[l], altz[l + 1], pz[l + 1] = v[0], v[1], v[5], v[6]
        w = fields(lines[k + 1], f3)
        wkl[:, l], wbrodl[l] = w[:7], w[7]
        k += 2
    # mixing-ratio input (:1733-1752)
    coldry = np.zeros(nlay)
    imix = all(wkl[m, 0] <= 1.0 for m in range(nmol))
    for l in range(nlay):
        summol = 0.0
        for m in range(1, nmol):
            summol = summol + wkl[m, l]
        if imix:
            coldry[l] = wbrodl[l] / (1.0 - summol)
            wkl[:, l] = coldry[l] * wkl[:, l]
        else:
            coldry[l] = wbrodl[l] + summol
    # solar-variability inputs as readprof passes them on (:1391-1427)
    indsolvar = np.ones(2)
    bndsolvar = np.ones(14)
    if isolvar in (1, 2):
        indsolvar = solvar[0:2].copy()
    if isolvar in (-1, 3):
        bndsolvar = solvar.copy()
    zenith = math.cos(sza * (2.0 * math.asin(1.0)) / 180.0)
    return dict(iaer=iaer, iout=iout, imca=imca, icld=icld, idelm=idelm, idcor=idcor,
                decorr_con=decorr_con, lat=lat, juldat=juldat, juldat_alpha=juldat_alpha, sza=sza,
                isolvar=isolvar, scon=scon, solcycfrac=solcycfrac, indsolvar=indsolvar,
                bndsolvar=bndsolvar, semiss=semiss, nlay=nlay, pavel=pavel, tavel=tavel,
                pz=pz, altz=altz, coldry=coldry, wkl=wkl, zenith=zenith)


def read_aer(path, nlay, zenith):
    """readaer (src/rrtmg_sw.1col.f90:1985-2127) for nstr = 2."""
    rnu0 = np.array([2903., 3601., 4310., 4892., 5623., 6872., 7872., 10590., 14420., 18970.,
                     25015., 30390., 43507., 1412.])
    rnu1 = {23: 10530.7, 24: 14293.3, 25: 18678.0, 26: 24475.1}
    f_od0 = {23: 0.1084, 24: 0.167, 25: 0.245, 26: 0.3611}
    f_od1 = {23: 0.3144, 24: 0.4822, 25: 0.7013, 26: 1.0239}
    eps = 1.e-10
    lines = open(path).read().split("\n")
    naer = fields(lines[0], [("x", 3), ("i", 2)])[0]
    tau = np.zeros((nlay, 14))
    ssa_l = np.zeros((nlay, 14))
    asm_l = np.zeros((nlay, 14))
    aod = np.zeros(nlay + 1)
    k = 1
    for _ in range(naer):
        v = fields(lines[k], [("x", 2), ("i", 3), ("x", 4), ("i", 1), ("x", 4), ("i", 1), ("x", 4),
                              ("i", 1), ("f", 8), ("f", 8), ("f", 8)])
        k += 1
        nl, iaod, issa, iasym = v[:4]
        aerpar = list(v[4:7])
        specfac = np.zeros(14)
        if iaod == 0:
            if aerpar[1] < eps:
                aerpar[1] = 1.0
            for ib in range(14):
                rl = 10000.0 / rnu0[ib]
                specfac[ib] = (aerpar[1] + aerpar[2] * rl) / ((aerpar[1] + aerpar[2] - 1.0) + rl ** aerpar[0])
        lays = []
        for _ in range(nl):
            v = fields(lines[k], [("x", 2), ("i", 3)] + [("f", 7)] * 14)
            k += 1
            lay, aod1 = v[0], v[1:]
            lays.append(lay)
            assert aod[lay] < eps
            if iaod == 0:
                aod[lay] = aod1[0]
                tau[lay - 1, :] = aod[lay] * specfac
            else:
                for ib in range(14):
                    aod[lay] = max(aod[lay], aod1[ib])
                    tau[lay - 1, ib] = aod1[ib]
        for b in range(23, 27):
            if iaod == 0:
                ib = b - IB1
                od = sum(tau[l - 1, ib] for l in lays) / zenith
                rnu = rnu0[ib] + (rnu1[b] - rnu0[ib]) * (od - f_od0[b]) / (f_od1[b] - f_od0[b])
                rl = 10000.0 / rnu
                sf = (aerpar[1] + aerpar[2] * rl) / ((aerpar[1] + aerpar[2] - 1.) + rl ** aerpar[0])
                for l in lays:
                    tau[l - 1, ib] = tau[l - 1, ib] * sf / specfac[ib]
        ssa = np.array(fields(lines[k], [("f", 5)] * 14))
        k += 1
        for l in lays:
            ssa_l[l - 1, :] = ssa[0] if issa == 0 else ssa
        assert iasym < 2
        asym = np.array(fields(lines[k], [("f", 5)] * 14))
        k += 1
        for l in lays:
            asm_l[l - 1, :] = asym[0] if iasym == 0 else asym
    # layers with no aerosol keep ssa = 0 (common-block default) but tau = 0
    return tau, ssa_l, asm_l


def read_cld(path, nlay):
    """readcld (src/rrtmg_sw.1col.f90:1884-1958) for inflag 2, with the conversion of
    :1807-1834 to ice and liquid water paths; arrays from the surface up."""
    lines = open(path).read().split("\n")
    inflag, iceflag, liqflag = fields(lines[0], [("x", 3), ("i", 2), ("x", 4), ("i", 1), ("x", 4), ("i", 1)])
    assert inflag == 2, path
    cldfrac, cwp, fice, rei, rel = (np.zeros(nlay) for _ in range(5))
    for line in lines[1:]:
        ctest, lay, frac, d1, d2, d3, d4 = fields(line, [("a", 1), ("x", 1), ("i", 3)] + [("f", 10)] * 5)
        if ctest == "%" or not line.strip():
            break
        cldfrac[lay - 1], cwp[lay - 1], fice[lay - 1], rei[lay - 1], rel[lay - 1] = frac, d1, d2, d3, d4
    ciwp = cwp * fice
    clwp = cwp * (1.0 - fice)
    return dict(inflag=inflag, iceflag=iceflag, liqflag=liqflag, cldfrac=cldfrac, ciwp=ciwp,
                clwp=clwp, rei=rei, rel=rel, tauc=np.zeros((14, nlay)), ssac=np.ones((14, nlay)),
                asmc=np.zeros((14, nlay)), fsfc=np.zeros((14, nlay)))


def read_output(path, nlay):
    """All flux blocks: dict wavenumber range -> (nlay+1, 6) from the surface up."""
    lines = open(path).read().split("\n")
    blocks = []
    for k, l in enumerate(lines):
        if "Wavenumbers:" in l:
            rng = l.split(":")[1].split("cm-1")[0]
            w1, w2 = [float(x.strip().rstrip(".")) for x in rng.split(" - ")]
            rows = np.zeros((nlay + 1, 6))
            for r in range(nlay + 1):
                t = lines[k + 3 + r].split()
                lev = int(t[0])
                rows[lev] = [float(x) for x in t[2:8]]
            blocks.append((w1, w2, rows))
    return blocks


WAVENUM2 = [3250., 4000., 4650., 5150., 6150., 7700., 8050., 12850., 16000., 22650., 29000.,
            38000., 50000., 2600.]


def main(ref):
    d = os.path.join(ref, "run_examples_std_atm")
    out = {}
    for name in CASES:
        c = read_input(os.path.join(d, INPUT.get(name, "input_rrtm_" + name)))
        L = c["nlay"]
        if c["iaer"] == 10:
            tau, ssa, asm = read_aer(os.path.join(d, AER[name]), L, c["zenith"])
        else:
            tau, ssa, asm = np.zeros((L, 14)), np.ones((L, 14)), np.zeros((L, 14))
        blocks = read_output(os.path.join(d, "output_rrtm_" + name), L)
        exp = np.full((15, L + 1, 6), np.nan)
        have = np.zeros(15, dtype=bool)
        for w1, w2, rows in blocks:
            if (w1, w2) == (820.0, 50000.0):
                b = 0
            else:
                b = 1 + WAVENUM2.index(w2)
            exp[b], have[b] = rows, True
        if name in CLD:
            c.update(read_cld(os.path.join(d, CLD[name]), L))
        for k, v in c.items():
            out[f"{name}/{k}"] = np.asarray(v)
        out[f"{name}/tauaer"] = tau
        out[f"{name}/ssaaer"] = ssa
        out[f"{name}/asmaer"] = asm
        out[f"{name}/expected"] = exp
        out[f"{name}/have"] = have
        print(f"{name}: nlay {L} iaer {c['iaer']} isolvar {c['isolvar']} scon {c['scon']} "
              f"sza {c['sza']} juldat {c['juldat']} blocks {int(have.sum())}")
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "column_kats.npz"),
                        cases=np.array(CASES), **out)
    # the example suite's own input files (data), under the names the column program reads
    import shutil
    for name in CASES:
        dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "column_inputs", name)
        os.makedirs(dst, exist_ok=True)
        shutil.copyfile(os.path.join(d, INPUT.get(name, "input_rrtm_" + name)), os.path.join(dst, "INPUT_RRTM"))
        if name in CLD:
            shutil.copyfile(os.path.join(d, CLD[name]), os.path.join(dst, "IN_CLD_RRTM"))
        if name in AER:
            shutil.copyfile(os.path.join(d, AER[name]), os.path.join(dst, "IN_AER_RRTM"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
