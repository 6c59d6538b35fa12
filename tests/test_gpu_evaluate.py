"""The numeric core of run.evaluate (anneal/run.py:55-93) on the drop-in FEM vs the FP64 oracle.

evaluate re-solves the as-built design (penalty 1, Ee = E0 * fill) for every
load case at rtol 1e-9 and reports compliance, maximum displacement, the
99.5th percentile of element von Mises stress away from anchors and, with a
build direction, of the tension across layers.  Every call it makes on FEM
(assemble, reduce_load, solve, expand, von_mises, stress, normal_stress) is
the device path here.
"""
import numpy as np
import pytest

from gpu_util import oracle_fem, random_problem

pytestmark = pytest.mark.gpu


def _evaluate(fem, g, fill, loads, E0, bdir):
    K = fem.assemble(E0 * fill[fem.act])
    cases = list(loads)
    Us = fem.solve(K, [fem.reduce_load(loads[c]) for c in cases], rtol=1e-9)
    out = {}
    for c, u_r in zip(cases, Us):
        u = fem.expand(u_r)
        mag = np.linalg.norm(u.reshape(-1, 3), axis=1)
        vm = fem.von_mises(u, E0)
        sz = np.maximum(type(fem).normal_stress(fem.stress(u, E0), bdir), 0.0)
        out[c] = {"compliance": float(fem.reduce_load(loads[c]) @ u_r), "maxDisplacement": float(mag.max()),
                  "peakStress": float(np.percentile(vm, 99.5)), "peakLayerTension": float(np.percentile(sz, 99.5)),
                  "u": u, "vm": vm}
    return out


def test_evaluate_core_matches_oracle():
    from anneal_amd.fem import FEM
    g, active, fixed, _ = random_problem((24, 12, 12), 31)
    rng = np.random.default_rng(3)
    fill = np.clip(rng.uniform(-0.2, 1.2, g.nelem), 0.05, 1.0)
    E0 = 2100.0
    loads = {}
    for name, (node, d) in {"down": (g.nx + (g.nx + 1) * (6 + 13 * 6), (0, 0, -1.0)),
                            "side": (g.nx + (g.nx + 1) * (3 + 13 * 9), (0, 1.0, 0.3))}.items():
        F = np.zeros(3 * g.nnode)
        F[3 * node:3 * node + 3] = d
        loads[name] = F
    bdir = (0.0, 0.3, 1.0)
    got = _evaluate(FEM(g, active, fixed, 0.3), g, fill, loads, E0, bdir)
    ref = _evaluate(oracle_fem(g, active, fixed), g, fill, loads, E0, bdir)
    for c in loads:
        a, b = got[c], ref[c]
        print(c, {k: (a[k], b[k]) for k in ("compliance", "maxDisplacement", "peakStress", "peakLayerTension")},
              flush=True)
        assert abs(a["compliance"] - b["compliance"]) <= 1e-8 * abs(b["compliance"])
        assert np.abs(a["u"] - b["u"]).max() <= 1e-7 * np.abs(b["u"]).max()
        assert abs(a["maxDisplacement"] - b["maxDisplacement"]) <= 1e-7 * b["maxDisplacement"]
        assert np.abs(a["vm"] - b["vm"]).max() <= 1e-6 * b["vm"].max()
        assert abs(a["peakStress"] - b["peakStress"]) <= 1e-6 * b["peakStress"]
        assert abs(a["peakLayerTension"] - b["peakLayerTension"]) <= 1e-6 * max(b["peakLayerTension"], 1e-30)
