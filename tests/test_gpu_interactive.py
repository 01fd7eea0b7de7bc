"""Interactive updates between steps (the interactor bridge): new solvent
accessible surfaces (InteractorFreeSASA::syncParticleStateData) and new
double-membrane parameters (MDDriver custom data dmou/dmol/dmtc) take effect
at the next step without a re-setup, also when steps are replayed from a
captured HIP graph, and the forces then equal the FP64 oracle's with the new
values (1e-5)."""
import sys

import numpy as np
import pytest

from biospring_amd import workloads
from biospring_amd.network import SpringNetwork
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _impala(seed=5, n=3000):
    s = workloads.generate(2, seed=seed, n=n)
    p = workloads.impala_params(s, "double")
    p.rigidbody_enable = 0
    p.viscosity_enable = 0
    return s, p


def _force_err(F, s, p):
    o = po.Oracle(s, params=p)
    Fd, _ = o.compute_forces(po.F64)
    o.close()
    return np.linalg.norm(F - Fd, axis=1).max() / np.linalg.norm(Fd, axis=1).max()


def test_sasa_update_reaches_impala():
    s, p = _impala()
    _log("interactive sasa: n=3000")
    net = SpringNetwork.from_system(s, params=p)
    F0, _ = net.computeForces()
    sasa = (s.sasa * 1.7 + 3.0).astype(np.float32)
    net.setSolventAccessibilitySurfaces(sasa)
    F1, _ = net.computeForces()
    net.close()
    assert np.abs(F1 - F0).max() > 1e-3
    s.sasa = sasa
    assert _force_err(F1, s, p) <= 1e-5


def test_membrane_update_reaches_graph_replays():
    s, p = _impala(seed=6)
    out = {}
    for g in (16, 0):
        q = p.copy()
        q.graph_steps = g
        net = SpringNetwork.from_system(s, params=q)
        _log(f"interactive membrane: graph_steps={g}")
        net.run(40)
        net.setMembrane(upper_offset=22.0, lower_curvature=-1.0 / 150.0)
        net.run(40)
        out[g] = net.getState()[0].copy()
        if g == 0:
            F, _ = net.computeForces()
            x, v = net.getState()
        net.close()
    assert np.array_equal(out[16], out[0])
    p2 = p.copy()
    p2.imp_upper_offset = 22.0
    p2.imp_lower_curvature = np.float32(-1.0 / 150.0)
    s.x = x.reshape(-1).astype(np.float32)
    s.v = v.reshape(-1).astype(np.float32)
    assert _force_err(F, s, p2) <= 1e-5
