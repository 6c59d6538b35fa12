"""Write the oracle's golden vectors used by the GPU parity tests.

Run from the repository root:  python tests/golden/make_golden.py
The fixtures are produced by the FP64 CPU oracle (oracle/), itself pinned by
the closed-form checks in tests/test_oracle.py.  Nothing here imports or runs
the reference.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.optimize_ref import cantilever, simp_ref  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def cantilever_trajectory(shape, iters, r_mm, rtol, name):
    g, fill, active, frozen, fixed, F = cantilever(shape)
    rec = []
    r = simp_ref(g, fill, active, frozen, fixed, {"c": F}, {"c": 1.0}, 1.0, 0.3, 0.3, iters=iters, r_mm=r_mm,
                 tol=0.0, emin=1e-9, rtol=rtol, record=rec, log=lambda *a: None)
    hist = r["history"]
    np.savez_compressed(
        os.path.join(HERE, name),
        shape=np.array(shape), iters=np.array(iters), r_mm=np.array(r_mm), rtol=np.array(rtol),
        compliance=np.array([h["compliance"][0] for h in hist]),
        volume=np.array([h["volumeFraction"] for h in hist]),
        change=np.array([h["change"] for h in hist]),
        grey=np.array([h["grey"] for h in hist]),
        p=np.array([h["p"] for h in hist]), beta=np.array([h["beta"] for h in hist]),
        oc_steps=np.array([h["ocSteps"] for h in hist]),
        x_final=r["x"],
        x_iter10=rec[9]["xPhys"] if len(rec) >= 10 else rec[-1]["xPhys"])
    print("wrote", name, "final compliance", hist[-1]["compliance"][0])


if __name__ == "__main__":
    which = sys.argv[1] if len(sys.argv) > 1 else "all"
    if which in ("all", "small"):
        cantilever_trajectory((16, 8, 8), 30, 1.5, 1e-11, "cantilever_16x8x8_it30.npz")
    if which in ("all", "c1"):
        cantilever_trajectory((48, 24, 24), 50, 1.5, 1e-11, "cantilever_48x24x24_it50.npz")
