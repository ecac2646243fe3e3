"""The boundary grid of the setcoef parity test, checked on the CPU against the oracle and the
closed form of each switch.

* The pressure switches found by bisection lie where the mathematics puts them:
  jp < k from log p = (36 - k) / 5 - 0.04 on, log p > 4.56, log p >= 6.62 (to 1e-13).
* Around every switch the oracle's decision is a single step: the 32 doubles below it on one
  side, the switch and the 31 above it on the other (the host log is monotone there).
* Around every temperature switch jt, jt1, indfor and indself each step once, by one, upward.
"""
import numpy as np

import oracle_bind as ob
import setcoef_grid as sg


def test_pressure_switches_closed_form():
    sw = sg.pressure_switches()
    for name, p in sw.items():
        if name == "trop":
            want = np.exp(4.56)
        elif name == "low":
            want = np.exp(6.62)
        else:
            want = np.exp((36.0 - name[1]) / 5.0 - 0.04)
        assert abs(p / want - 1.0) < 1e-13, (name, p, want)


def test_pressure_switches_single_step():
    sw = sg.pressure_switches()
    side = np.arange(-sg.NEAR, sg.NEAR) >= 0
    for name, p in sw.items():
        jp, trop, low = sg._decide(sg.around(p)[0])
        if name == "trop":
            got = trop == 1
        elif name == "low":
            got = low == 1
        else:
            got = jp < name[1]
            # jp is k just below the switch and k - 1 from it on
            assert np.array_equal(jp, np.where(side, name[1] - 1, name[1])), name
        assert np.array_equal(got, side), name


def test_temperature_switches_single_step():
    g = sg.temperature_grid()
    n, L = g["play"].shape   # one column per switch, its 64 temperatures as layers
    o = ob.run_setcoef(n, L, **g)
    # (the switch of a rounded quotient lies within an ulp or two of the constructed point)
    def one_step(blk, lo, hi):
        k = int(np.argmax(blk != blk[0]))
        return (blk[0] == lo and blk[-1] == hi and abs(k - sg.NEAR) <= 2
                and (blk[:k] == lo).all() and (blk[k:] == hi).all())
    # jt (tref of jp) and jt1 (tref of jp + 1) rise by one at every switch of their block
    for b in range(58 * 6):
        blk = o["jt" if (b // 3) % 2 == 0 else "jt1"][b]
        m = 2 + b % 3
        assert one_step(blk, m - 1, m), b
    # indfor falls at T = 260 K, indself rises at 188 + 7.2 m for m = 9..16 (500 hPa)
    rest = slice(58 * 6, n)
    fo, se = o["indfor"][rest], o["indself"][rest]
    assert one_step(fo[0], 2, 1)
    assert (fo[1] == 1).all()                       # 296 K: int 1 -> 0 is clamped to 1
    assert (se[2] == 1).all()                       # m = 8: 0 -> 1 is clamped to 1
    for i, m in enumerate(range(9, 17)):
        assert one_step(se[3 + i], m - 8, m - 7), m
