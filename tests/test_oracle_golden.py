"""The C oracle against golden fixtures written by the compiled reference
(tests/golden/make_golden.py): bit-for-bit equal fluxes, heating rates and McICA masks, and the
224 -> 112 reduced k-tables equal to the reference's by SHA-256. These pin the oracle where
oracle/_ref is not built."""
import os

import numpy as np
import pytest

import oracle_bind as ob
from golden.make_golden import CASES, NOMCICA_CASES, build

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_reference_golden(name):
    kind, kw = CASES[name]
    cols = build(kind, kw)
    g = np.load(os.path.join(HERE, "golden", f"{name}.npz"))
    if cols.icld:
        mc, _ = ob.run_mcica(cols, "orc")
        bits = np.packbits(mc["cldfmcl"].astype(bool).ravel(order="F"))
        assert np.array_equal(bits, g["cldfmcl_bits"])
    else:
        mc = ob.empty_mcica(cols)
    out = ob.run_rrtmg(cols, mc, which="orc", icld=min(cols.icld, 3))
    for k in ("swuflx", "swdflx", "swuflxc", "swdflxc", "swhr", "swhrc"):
        assert np.array_equal(out[k], g[k]), k


@pytest.mark.parametrize("name", sorted(NOMCICA_CASES))
def test_oracle_nomcica_matches_reference_golden(name):
    kind, kw = NOMCICA_CASES[name]
    cols = build(kind, kw)
    g = np.load(os.path.join(HERE, "golden", f"{name}.npz"))
    out, _, _ = ob.run_nomcica(cols, "orc")
    for k in ("swuflx", "swdflx", "swuflxc", "swdflxc", "swhr", "swhrc"):
        assert np.array_equal(out[k], g[k]), k


def test_oracle_reduced_tables_match_reference_digests():
    import hashlib
    import json
    want = json.load(open(os.path.join(HERE, "golden", "reduced_tables.json")))
    assert len(want) > 100
    for name, d in want.items():
        o = ob.orc_table(name)
        assert o is not None and o.size == d["n"], name
        assert hashlib.sha256(o.tobytes()).hexdigest() == d["sha256"], name
