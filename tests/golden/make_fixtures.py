"""Writes tests/golden/oracle_totals.json: per-variable totals (sum and sum of
absolute values over interior cells) of the oracle after a few steps of small
seeded configurations.  They pin the oracle against unintended changes; the GPU
parity tests compare the HIP path with the live oracle bit for bit and with
these totals.

    python tests/golden/make_fixtures.py
"""
import json
import math
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import exocubed_amd as ex  # noqa: E402
from helpers import atm_config, interior, sw_config  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402

CASES = {
    "sw_roe_8x8x3_5steps": (lambda: sw_config(ex, nx=8, rsolver=0), 0, 5),
    "sw_lmars_10x10x1_5steps": (lambda: sw_config(ex, nx=10, rsolver=1, nx1=1), 0, 5),
    "atm_lmars_implicit_8x8x8_3steps": (lambda: atm_config(ex, nx=8, nx1=8, rsolver=3), 1, 3),
    "atm_hllc_implicit_6x6x10_3steps": (lambda: atm_config(ex, nx=6, nx1=10, rsolver=2), 1, 3),
}
SEED = 20261202


def totals(cfg, kind, nsteps):
    u0, s0 = ex.init_state(cfg, kind, SEED, ())
    o = Oracle(cfg)
    o.upload(u0, s0)
    o.step(nsteps)
    u = o.download()[0]
    o.close()
    ui = interior(u, cfg.nx1)
    return {"sum": [math.fsum(ui[:, n].ravel()) for n in range(ui.shape[1])],
            "sum_abs": [math.fsum(abs(x) for x in ui[:, n].ravel()) for n in range(ui.shape[1])]}


def compute():
    return {name: totals(make(), kind, n) for name, (make, kind, n) in CASES.items()}


if __name__ == "__main__":
    out = {"seed": SEED, "cases": compute()}
    with open(os.path.join(HERE, "oracle_totals.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", os.path.join(HERE, "oracle_totals.json"))
