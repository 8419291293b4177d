"""The GPU engine (FP64 path, through the C ABI) against the committed golden
fixtures of tests/golden/: every case after the fixture's step count, to the
FP64 parity bar (1e-12 relative per step, so 5 steps here: 1e-11)."""
import json
import os

import numpy as np
import pytest

from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MANIFEST = json.load(open(os.path.join(GOLD, "manifest.json")))


@pytest.mark.parametrize("case", MANIFEST["cases"], ids=lambda c: c["name"])
def test_gpu_matches_fixture(amps, case):
    z = np.load(os.path.join(GOLD, case["name"] + ".npz"))
    params, inputs, outputs = z["params"], z["inputs"], z["outputs"]
    cfg = amps.Config.default(case["model"])
    cfg.nx = cfg.ny = MANIFEST["n"]
    cfg.mode = amps.PURE_FP64
    cfg.stencil, cfg.bc, cfg.seed, cfg.q = case["stencil"], case["bc"], case["seed"], case["q"]
    for i in range(amps.NPARAMS):
        cfg.params[i] = params[i]
    with amps.Context(cfg) as ctx:
        ctx.init()
        for f in range(inputs.shape[0]):
            ctx.set_field(f, inputs[f])
        ctx.step(MANIFEST["nsteps"])
        for f in range(outputs.shape[0]):
            assert rel_err(ctx.get_field(f), outputs[f]) <= 1e-11, (case["name"], f)
