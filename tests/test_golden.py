"""Golden fixtures under tests/golden: the reference's WENO known-answer
vectors (against the oracle here, against the device in test_gpu_parity) and
the oracle's own totals on small seeded runs (guarding the oracle against
unintended changes; regenerate with tests/golden/make_fixtures.py only after a
deliberate change that its pins still pass)."""
import json
import math
import os
import sys

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(HERE, name)) as f:
        return json.load(f)


def test_oracle_matches_reference_weno_vectors(orc):
    kat = load("weno_kat.json")
    L = orc.lib()
    for c in kat["weno3"]:
        assert abs(L.or_weno3(c["phim1"], c["phi"], c["phip1"]) - c["expected"]) <= kat["tolerance_abs"]
    for c in kat["weno5"]:
        r = L.or_weno5(c["phim2"], c["phim1"], c["phi"], c["phip1"], c["phip2"])
        assert abs(r - c["expected"]) <= kat["tolerance_abs"]
        assert r == c["expected"]  # bit for bit: the last bit fixes the summation order


def test_oracle_totals_unchanged():
    sys.path.insert(0, HERE)
    import make_fixtures

    ref = load("oracle_totals.json")["cases"]
    now = make_fixtures.compute()
    for name, r in ref.items():
        for key in ("sum", "sum_abs"):
            for a, b in zip(now[name][key], r[key]):
                assert a == b or abs(a - b) <= 1e-14 * max(abs(b), 1.0), (name, key, a, b)
