"""Batches of independent systems (workload 3 at reduced size) on the GPU:
each system against its own FP64 oracle -- forces, per-system energies and a
short trajectory -- with the float-faithful oracle as the yardstick."""
import sys

import numpy as np
import pytest

from biospring_amd import workloads
from biospring_amd.network import SpringNetwork
from oracle import pyoracle as po
from tests.parity import ENERGY_RTOL, FORCE_RTOL, energy_error, force_errors

pytestmark = pytest.mark.gpu


def _log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_batch_matches_each_system_alone():
    s, off, parts = workloads.batch(12, seed=5, nmax=2500)
    _log("batch n", s.n, "m", s.m, "sizes", np.diff(off).tolist())
    net = SpringNetwork.from_system(s)
    net.setSystems(off)
    F, E = net.computeForces()
    Es = net.getSystemEnergies()
    assert np.allclose(Es.sum(0), E, rtol=1e-12, atol=1e-12)
    for k, p in enumerate(parts):
        o = po.Oracle(p)
        Fd, Ed = o.compute_forces(po.F64)
        sl = slice(off[k], off[k + 1])
        per, glob = force_errors(F[sl], Fd, (p.flags & 1) == 0)
        assert per <= FORCE_RTOL, (k, per)
        for slot in range(6):
            assert energy_error(Es[k, slot], Ed[slot], floor=1e-3) <= ENERGY_RTOL, (k, slot, Es[k, slot], Ed[slot])
    net.run(40)
    x, _ = net.getState()
    Es = net.getSystemEnergies()
    worst = 0.0
    for k, p in enumerate(parts):
        o64, o32 = po.Oracle(p), po.Oracle(p)
        o64.step(po.F64, 40)
        o32.step(po.F32, 40)
        xd, _ = o64.state(po.F64)
        xf, _ = o32.state(po.F32)
        sl = slice(off[k], off[k + 1])
        dev = np.abs(x[sl] - xd).max()
        dref = np.abs(xf - xd).max()
        worst = max(worst, dev / max(dref, 1e-12))
        assert dev <= 3 * dref + 1e-6, (k, dev, dref)
    _log(f"  40 steps: worst gpu/float-path deviation ratio {worst:.2f}")


def test_batch_graph_replay_matches_direct_launches():
    s, off, _ = workloads.batch(6, seed=6, nmax=1500)
    a = SpringNetwork.from_system(s)
    a.setSystems(off)
    p = s.params.copy()
    p.graph_steps = 4
    b = SpringNetwork.from_system(s, params=p)
    b.setSystems(off)
    for n in (7, 10, 1):
        a.run(n)
        b.run(n)
        xa, va = a.getState()
        xb, vb = b.getState()
        assert np.array_equal(xa, xb) and np.array_equal(va, vb), n
        assert np.array_equal(a.getSystemEnergies(), b.getSystemEnergies()), n


def test_autotune_picks_a_valid_configuration():
    """bs_autotune (lanes per spring row, graph replay, pair layout) leaves the
    state alone and the tuned handle steps to the same positions as an
    untuned one (to FP64-summation rounding)."""
    s, off, _ = workloads.batch(6, seed=8, nmax=1500)
    a = SpringNetwork.from_system(s)
    a.setSystems(off)
    x0, v0 = a.getState()
    g, gr, lay = a.autotune(5)
    _log(f"autotune: lanes {g}, graph {gr}, layout {lay}")
    assert g in (4, 8, 16, 32) and gr in (0, 16) and lay in (0, 2)
    x1, v1 = a.getState()
    assert np.array_equal(x0, x1) and np.array_equal(v0, v1)
    b = SpringNetwork.from_system(s)
    b.setSystems(off)
    a.run(40)
    b.run(40)
    xa, _ = a.getState()
    xb, _ = b.getState()
    assert np.abs(xa - xb).max() <= 1e-4


def test_systems_never_pair_after_moving_onto_each_other():
    """Independent systems stay independent however far they move from where
    the setup placed them on the engine's binning grid: system 1 is
    translated by exactly the lattice spacing between the systems' binning
    sites, onto system 0's site; each system's forces and energies still
    match its own FP64 oracle."""
    s, off, parts = workloads.batch(3, seed=9, nmin=800, nmax=1500)
    net = SpringNetwork.from_system(s)
    net.setSystems(off)
    net.setup()
    x, v = net.getState()
    p = s.params
    ext = max(float(np.ptp(part.x.reshape(-1, 3).astype(np.float64), axis=0).max()) for part in parts)
    radius = max(p.steric_cutoff if p.steric_enable else 0.0,
                 p.electrostatic_cutoff if p.electrostatic_enable else 0.0) + max(p.list_skin, 0.0)
    spacing = np.ceil(ext + 2.0 * radius + 20.0)
    x = x.copy()
    x[off[1]:off[2], 0] -= np.float32(spacing)
    _log(f"moving system 1 by -{spacing} A in x")
    net.setState(x, v)
    F, E = net.computeForces()
    Es = net.getSystemEnergies()
    for k, part in enumerate(parts):
        o = po.Oracle(part)
        sl = slice(off[k], off[k + 1])
        o.set_state(x[sl], v[sl])
        Fd, Ed = o.compute_forces(po.F64)
        o.close()
        per, _ = force_errors(F[sl], Fd, (part.flags & 1) == 0)
        _log(f"  system {k}: per-particle {per:.3g}")
        assert per <= FORCE_RTOL, (k, per)
        for slot in range(6):
            assert energy_error(Es[k, slot], Ed[slot], floor=1e-3) <= ENERGY_RTOL, (k, slot, Es[k, slot], Ed[slot])
    net.close()
