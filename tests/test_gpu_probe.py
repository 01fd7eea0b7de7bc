"""The probe particle on the GPU (SpringNetwork.cpp:154-182,
Particle.cpp:434-465) at sizes the golden fixtures do not reach: the spring
kernel's block sums of the probe's force over many workgroups, graph replay,
a re-setup between calls, and the refusals (batches, the adjoint).
"""
import sys

import numpy as np
import pytest

from biospring_amd import _abi, workloads
from biospring_amd.network import SpringNetwork
from oracle import pyoracle as po
from tests.parity import ENERGY_RTOL, FORCE_RTOL, energy_error, force_errors

pytestmark = pytest.mark.gpu


def _log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _system(n=20000, mode=1, ddd=1):
    s = workloads.generate(1, seed=41, n=n)
    p = s.params.copy()
    p.steric_mode = mode
    p.dielectric_model = ddd
    x = s.x.reshape(-1, 3)
    p.probe_enable, p.probe_enable_steric, p.probe_enable_electrostatic = 1, 1, 1
    p.probe_x, p.probe_y, p.probe_z = float(x[:, 0].mean()), float(x[:, 1].mean()), float(x[:, 2].max() + 4.0)
    p.probe_mass, p.probe_radius, p.probe_epsilon, p.probe_charge = 20.0, 3.5, 0.7, 2.0
    return s, p


@pytest.mark.parametrize("mode,ddd", [(1, 1), (0, 0), (2, 0), (3, 1)])
def test_probe_forces_energies_vs_fp64(mode, ddd):
    s, p = _system(mode=mode, ddd=ddd)
    q = p.copy()
    q.viscosity_enable = 0
    _log("probe forces: n", s.n, "mode", mode, "ddd", ddd)
    net = SpringNetwork.from_system(s, params=q)
    F, E = net.computeForces()
    o = po.Oracle(s, params=q)
    Fd, Ed = o.compute_forces(po.F64)
    _, _, fpd = o.probe(po.F64)
    o.close()
    per, glob = force_errors(F, Fd, (s.flags & 1) == 0)
    _log(f"  per-particle {per:.3g} global {glob:.3g}")
    assert per <= FORCE_RTOL
    for slot in range(6):
        assert energy_error(E[slot], Ed[slot], floor=1e-3) <= ENERGY_RTOL, (slot, E[slot], Ed[slot])
    x, v, f = net.getProbeParticle()
    # a force evaluation leaves the probe in place
    assert np.array_equal(x, np.float32([p.probe_x, p.probe_y, p.probe_z])) and not v.any()
    err = np.abs(f - fpd).max() / np.abs(fpd).max()
    _log(f"  probe force {f} vs {fpd}: {err:.3g}")
    assert err <= 1e-5
    net.close()


@pytest.mark.parametrize("graph", [0, 8])
def test_probe_trajectory_vs_fp64(graph):
    """50 steps with the probe: particles and probe follow the FP64 oracle as
    closely as the reference's float path does; graph replay gives the same
    result bit for bit."""
    s, p = _system(n=6000)
    p.graph_steps = graph
    net = SpringNetwork.from_system(s, params=p)
    net.run(20)
    net.run(30)
    x, _ = net.getState()
    pg = np.concatenate(net.getProbeParticle())
    E = net.getEnergies()
    o64, o32 = po.Oracle(s, params=p), po.Oracle(s, params=p)
    _, Ed = o64.step(po.F64, 50)
    _, Ef = o32.step(po.F32, 50)
    xd, _ = o64.state(po.F64)
    xf, _ = o32.state(po.F32)
    pd = np.concatenate(o64.probe(po.F64))
    pf = np.concatenate(o32.probe(po.F32))
    dev, dref = np.abs(x - xd).max(), np.abs(xf - xd).max()
    _log(f"  graph {graph}: particles gpu-f64 {dev:.3g} f32-f64 {dref:.3g}")
    assert dev <= 2 * dref + 1e-6
    for sl in (slice(0, 3), slice(3, 6), slice(6, 9)):
        eg, er = np.abs(pg[sl] - pd[sl]).max(), np.abs(pf[sl] - pd[sl]).max()
        _log(f"  probe {sl}: gpu-f64 {eg:.3g} f32-f64 {er:.3g}")
        assert eg <= 2 * er + 1e-5 * max(np.abs(pd[sl]).max(), 1e-9)
    moved = np.abs(pd[0:3] - np.array([p.probe_x, p.probe_y, p.probe_z])).max()
    assert moved > 1e3 * max(np.abs(pg[0:3] - pd[0:3]).max(), 1e-7), moved
    # 50 steps apart the trajectories differ by ~2e-4 A, so the step energies
    # are held to the float path's own distance from FP64 (or 1e-4)
    for slot in (_abi.E_STERIC, _abi.E_ELECTROSTATIC, _abi.E_KINETIC):
        err, err_ref = energy_error(E[slot], Ed[slot], floor=1e-3), energy_error(Ef[slot], Ed[slot], floor=1e-3)
        _log(f"  energy {slot}: gpu-f64 {err:.3g} f32-f64 {err_ref:.3g}")
        assert err <= max(1e-4, 2 * err_ref), (slot, E[slot], Ed[slot], Ef[slot])
    o64.close()
    o32.close()
    if graph:
        q = p.copy()
        q.graph_steps = 0
        ref = SpringNetwork.from_system(s, params=q)
        ref.run(20)
        ref.run(30)
        assert np.array_equal(ref.getState()[0], x)
        assert np.array_equal(np.concatenate(ref.getProbeParticle()), pg)
        ref.close()
    net.close()


def test_probe_survives_resetup_and_refusals():
    s, p = _system(n=3000)
    net = SpringNetwork.from_system(s, params=p)
    net.run(10)
    a = np.concatenate(net.getProbeParticle())
    # new parameters: the engine re-runs its setup lazily; the probe keeps its state
    q = p.copy()
    q.list_skin = p.list_skin + 0.25
    net.setParams(q)
    b = np.concatenate(net.getProbeParticle())
    assert np.array_equal(a, b)
    net.run(5)
    ref = SpringNetwork.from_system(s, params=p)
    ref.run(15)
    # the lists are rebuilt at another step, so the pair sums may round apart
    d = np.abs(np.concatenate(ref.getProbeParticle())[0:3] - np.concatenate(net.getProbeParticle())[0:3]).max()
    assert d <= 1e-5, d
    ref.close()
    # the adjoint has no probe term
    with pytest.raises(_abi.EngineError) as ei:
        net.adjoint(2, s.x.reshape(-1, 3))
    assert ei.value.code == _abi.BS_ERR_ARG
    net.close()
    # a batch of systems has no single probe
    bs, off, _ = workloads.batch(2, seed=3, nmin=300, nmax=600)
    pb = bs.params.copy()
    pb.probe_enable, pb.probe_enable_steric = 1, 1
    b2 = SpringNetwork.from_system(bs, params=pb)
    with pytest.raises(_abi.EngineError) as ei:
        b2.setSystems(off)
        b2.setup()
    assert ei.value.code == _abi.BS_ERR_ARG
    b2.close()
    # without the probe, bs_get_probe refuses
    q = p.copy()
    q.probe_enable = 0
    c = SpringNetwork.from_system(s, params=q)
    with pytest.raises(_abi.EngineError) as ei:
        c.getProbeParticle()
    assert ei.value.code == _abi.BS_ERR_STATE
    c.close()
