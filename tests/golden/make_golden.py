"""Writes the golden FP64 fixtures of tests/golden/ from the oracle.

Each fixture is a small seeded case of one model: the inputs (fields,
parameters, stencil, boundary, seed), the fields after NSTEPS oracle steps,
and the SHA-256 of the output arrays.  tests/test_golden.py checks that the
oracle still reproduces them bit for bit (a regression anchor for the
restatement) and tests/test_gpu_golden.py that the GPU engine matches them to
the FP64 parity bar.  Run:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from tests.helpers import smooth_fields  # noqa: E402

NSTEPS = 5
N = 48


def digest(arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a, np.float64).tobytes())
    return h.hexdigest()


def cases():
    out = []
    # snowflake: two parameter sets, cross-2 periodic and 5-point zero-flux with noise
    for k, stencil, bc, noise in ((2, O.STENCIL_REF, O.BC_PERIODIC, 0.0), (4, O.STENCIL_LAP5, O.BC_NEUMANN, 0.01)):
        p = O.sf_params_ref(k, N, N)
        p[O.SF["NOISE"]] = noise
        phi, T = smooth_fields(N, N, seed=10 + k)
        T = 0.5 * T
        a, b = phi.copy(), T.copy()
        for s in range(NSTEPS):
            a, b = O.sf_step(a, b, p, stencil=stencil, bc=bc, seed=1, step=s)
        out.append(dict(name=f"snowflake_k{k}_st{stencil}_bc{bc}", model=1, stencil=stencil, bc=bc, seed=1, q=2,
                        params=p, inputs=[phi, T], outputs=[a, b]))
    # grain growth: q = 3, Moore periodic
    rng = np.random.default_rng(3)
    p = O.gg_params_ref(1, N, N)
    eta = np.clip(rng.uniform(-0.1, 1.1, (3, N, N)), 0, 1)
    e = eta.copy()
    for _ in range(NSTEPS):
        e = O.gg_step(e, p, stencil=O.STENCIL_REF, bc=O.BC_PERIODIC)
    out.append(dict(name="graingrowth_q3_moore", model=2, stencil=O.STENCIL_REF, bc=O.BC_PERIODIC, seed=1, q=3,
                    params=p, inputs=list(eta), outputs=list(e)))
    # sintering: two tanh particles, zero flux
    y, x = np.mgrid[0:N, 0:N].astype(np.float64)
    d1 = 0.5 * (1 + np.tanh((10 - np.hypot(x - 15, y - 24)) / 2.0))
    d2 = 0.5 * (1 + np.tanh((10 - np.hypot(x - 34, y - 25)) / 2.0))
    c = np.maximum(d1, d2)
    p = O.sn_params_ref(1, N, N)
    cc, a1, a2 = c.copy(), d1.copy(), d2.copy()
    for _ in range(NSTEPS):
        cc, a1, a2 = O.sn_step(cc, a1, a2, p, bc=O.BC_NEUMANN)
    out.append(dict(name="sintering_neumann", model=3, stencil=0, bc=O.BC_NEUMANN, seed=1, q=2,
                    params=p, inputs=[c, d1, d2], outputs=[cc, a1, a2]))
    return out


def main():
    manifest = {"nsteps": NSTEPS, "n": N, "cases": []}
    for c in cases():
        path = os.path.join(HERE, c["name"] + ".npz")
        np.savez_compressed(path, params=np.asarray(c["params"], np.float64),
                            inputs=np.stack(c["inputs"]), outputs=np.stack(c["outputs"]))
        manifest["cases"].append({k: c[k] for k in ("name", "model", "stencil", "bc", "seed", "q")} |
                                 {"sha256_outputs": digest(c["outputs"])})
        print("wrote", path)
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1)


if __name__ == "__main__":
    main()
