"""A call stops at its first failing step (SpringNetwork.cpp:222-229 ends the
program there): a non-finite position returns the state of the step that
produced it, and a neighbour list that overflows is lengthened and the call
resumes from the step where it stopped, so that no step runs with a
truncated list.  Both conditions are forced from the host (an infinite
external force; a setState that packs the beads densely): no kernel is made
to fail.
"""
import os
import sys
import tempfile

import numpy as np
import pytest

from biospring_amd import _abi, workloads
from biospring_amd.network import SpringNetwork
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_nonfinite_stops_at_the_failing_step():
    s = workloads.generate(0, seed=1, n=2000)
    ext = np.zeros((s.n, 3), np.float32)
    ext[17, 0] = np.inf
    net = SpringNetwork.from_system(s)
    net.setForce(ext)
    with pytest.raises(_abi.EngineError) as ei:
        net.run(10)
    assert ei.value.code == _abi.BS_ERR_NONFINITE and net.lastNonfiniteParticle() == 17
    assert net.counters()["steps"] == 1
    x, v = net.getState()
    # the same single step on a fresh network: every other particle matches
    ref = SpringNetwork.from_system(s)
    ref.setForce(ext)
    with pytest.raises(_abi.EngineError):
        ref.run(1)
    xr, vr = ref.getState()
    keep = np.arange(s.n) != 17
    assert np.array_equal(x[keep], xr[keep]) and np.array_equal(v[keep], vr[keep])
    assert not np.isfinite(x[17]).all()
    # the report stays until a new state is set
    with pytest.raises(_abi.EngineError):
        net.run(3)
    net.setState(s.x.reshape(-1, 3), np.zeros((s.n, 3), np.float32))
    net.setForce(None)
    net.run(3)
    assert np.isfinite(net.getState()[0]).all()
    net.close()
    ref.close()


def _packed(s, f=0.62):
    x = s.x.reshape(-1, 3).astype(np.float64)
    c = x.mean(axis=0)
    return (c + (x - c) * f).astype(np.float32)


def _soft(s):
    """Parameters under which packed beads stay put: the linear steric wall
    (forces of order C_spring per A of overlap) at a 12 A cutoff, no springs
    (packing would stretch them far beyond rest), no Coulomb."""
    p = s.params.copy()
    p.spring_enable = 0
    p.electrostatic_enable = 0
    p.steric_mode = _abi.STERIC_LINEAR
    p.viscosity_enable = 1
    return p


def test_overflow_lengthens_lists_and_resumes():
    """The beads are packed to about 4x the density the lists were sized for:
    the next rebuild overflows; the call lengthens the lists, rebuilds and
    completes every step.  The result follows the FP64 oracle as closely as
    the reference's float path does."""
    s = workloads.generate(1, seed=3, n=4000)
    p = _soft(s)
    net = SpringNetwork.from_system(s, params=p)
    net.setup()
    cap0 = net.counters()["list_cap"]
    xp = _packed(s)
    v0 = np.zeros((s.n, 3), np.float32)
    net.setState(xp, v0)
    net.run(15)
    c = net.counters()
    _log(f"list capacity {cap0} -> {c['list_cap']}, longest {c['max_list']}, steps {c['steps']}")
    assert c["list_cap"] > cap0 and c["steps"] == 15
    x, _ = net.getState()
    o64, o32 = po.Oracle(s, params=p), po.Oracle(s, params=p)
    for o in (o64, o32):
        o.set_state(xp, v0)
    o64.step(po.F64, 15)
    o32.step(po.F32, 15)
    xd, _ = o64.state(po.F64)
    xf, _ = o32.state(po.F32)
    dev, dref = np.abs(x - xd).max(), np.abs(xf - xd).max()
    _log(f"  15 steps after the overflow: gpu-f64 {dev:.3g}, f32-f64 {dref:.3g}")
    assert dev <= 2 * dref + 1e-5
    net.close()


def test_overflow_keeps_trajectory_frames():
    """Frames due after the stop are taken again once the call resumes: the
    energy table has one row per due iteration, in order, as without an
    overflow."""
    s = workloads.generate(1, seed=4, n=3000)
    p = _soft(s)
    xp = _packed(s)
    v0 = np.zeros((s.n, 3), np.float32)
    rows = {}
    with tempfile.TemporaryDirectory() as d:
        for tag, pre in (("overflow", True), ("plain", False)):
            net = SpringNetwork.from_system(s, params=p)
            if pre:
                net.setup()  # lists sized for the loose start
                net.setState(xp, v0)
            else:
                net.setParticles(xp, v0, *s.particle_arrays()[2:])
                net.setSprings(*s.spring_arrays())
            path = os.path.join(d, f"{tag}.csv")
            net.addTrajectory(_abi.TRAJ_CSV, path, 3)
            net.run(12)
            net.closeTrajectories()
            with open(path) as fh:  # step, frame rate (timing: left out), energies
                rows[tag] = [[c for k, c in enumerate(ln.split("\t")) if k != 1]
                             for ln in fh.read().splitlines()[1:]]
            net.close()
    _log("frames", rows)
    assert [r[0] for r in rows["plain"]] == ["0", "3", "6", "9"]
    assert rows["overflow"] == rows["plain"]
