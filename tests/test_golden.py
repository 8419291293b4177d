"""The oracle reproduces the committed golden fixtures (tests/golden/) bit for
bit: a regression anchor for the CPU restatement.  The fixtures were written
by tests/golden/make_golden.py; the manifest holds the SHA-256 of each case's
outputs."""
import hashlib
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MANIFEST = json.load(open(os.path.join(GOLD, "manifest.json")))


def load(case):
    z = np.load(os.path.join(GOLD, case["name"] + ".npz"))   # allow_pickle=False (default)
    return z["params"], z["inputs"], z["outputs"]


def oracle_run(O, case, params, inputs, nsteps):
    m, st, bc = case["model"], case["stencil"], case["bc"]
    if m == 1:
        a, b = inputs[0].copy(), inputs[1].copy()
        for s in range(nsteps):
            a, b = O.sf_step(a, b, params, stencil=st, bc=bc, seed=case["seed"], step=s)
        return np.stack([a, b])
    if m == 2:
        e = inputs.copy()
        for _ in range(nsteps):
            e = O.gg_step(e, params, stencil=st, bc=bc)
        return e
    c, e1, e2 = (x.copy() for x in inputs)
    for _ in range(nsteps):
        c, e1, e2 = O.sn_step(c, e1, e2, params, bc=bc)
    return np.stack([c, e1, e2])


@pytest.mark.parametrize("case", MANIFEST["cases"], ids=lambda c: c["name"])
def test_fixture_digest(case):
    _, _, outputs = load(case)
    h = hashlib.sha256()
    for a in outputs:
        h.update(np.ascontiguousarray(a, np.float64).tobytes())
    assert h.hexdigest() == case["sha256_outputs"]


@pytest.mark.parametrize("case", MANIFEST["cases"], ids=lambda c: c["name"])
def test_oracle_reproduces_fixture(O, case):
    params, inputs, outputs = load(case)
    got = oracle_run(O, case, params, inputs, MANIFEST["nsteps"])
    assert np.array_equal(got, outputs)
