"""GPU parity of the paths the other files leave open: finite external forces
(the interactor's setForce), compensated positions over 1000 steps, and the
250k-bead headline network at its full size.
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


def _ext_forces(s, seed=11, frac=0.5):
    """Finite external forces on a random half of the beads, of the size of
    the typical conservative force (Da A fs^-2), as an interactor sets them
    (SpringNetwork::setForce, SpringNetwork.cpp:436-442)."""
    rng = np.random.default_rng(seed)
    o = po.Oracle(s)
    F0, _ = o.compute_forces(po.F64)
    o.close()
    scale = float(np.median(np.linalg.norm(F0, axis=1)))
    ext = (rng.standard_normal((s.n, 3)) * scale).astype(np.float32)
    ext[rng.random(s.n) > frac] = 0.0
    return ext


def test_external_force_finite():
    """Finite external forces are added before the force evaluation: the
    total forces equal the FP64 oracle's with the same external force, and a
    20-step trajectory driven by them follows the FP64 oracle's as closely as
    the reference's own float path does."""
    s = workloads.generate(1, seed=2, n=3000)
    p = s.params.copy()
    p.viscosity_enable = 0  # compare conservative + external forces
    ext = _ext_forces(s)
    _log("ext force: n", s.n, "nonzero", int((np.abs(ext).sum(1) > 0).sum()))
    net = SpringNetwork.from_system(s, params=p)
    net.setForce(ext)
    F, E = net.computeForces()
    o = po.Oracle(s, params=p)
    Fd, Ed = o.compute_forces(po.F64, ext=ext)
    o.close()
    dyn = (s.flags & 1) == 0
    per, glob = force_errors(F, Fd, dyn)
    _log(f"  forces with ext: per-particle {per:.3g} global {glob:.3g}")
    assert per <= FORCE_RTOL
    # the external force really is in the result: without it the forces differ
    o = po.Oracle(s, params=p)
    Fn, _ = o.compute_forces(po.F64)
    o.close()
    assert np.abs(Fd - Fn)[dyn].max() > 0.1 * np.abs(ext).max()
    for slot in range(6):
        assert energy_error(E[slot], Ed[slot], floor=1e-3) <= ENERGY_RTOL, (slot, E[slot], Ed[slot])
    net.close()

    # stepped with the reference's default viscous integrator
    net = SpringNetwork.from_system(s)
    net.setForce(ext)
    net.run(20)
    x, _ = net.getState()
    net.close()
    o64, o32 = po.Oracle(s), po.Oracle(s)
    o64.step(po.F64, 20, ext=ext)
    o32.step(po.F32, 20, ext=ext)
    xd, _ = o64.state(po.F64)
    xf, _ = o32.state(po.F32)
    o64.close()
    o32.close()
    dev, dref = np.abs(x - xd).max(), np.abs(xf - xd).max()
    # the drive is visible: the trajectory differs from the undriven one
    o = po.Oracle(s)
    o.step(po.F64, 20)
    xu, _ = o.state(po.F64)
    o.close()
    _log(f"  20 driven steps: gpu max dev {dev:.3e}, float path {dref:.3e}, drive moved {np.abs(xd - xu).max():.3e} A")
    assert np.abs(xd - xu).max() > 100 * max(dev, 1e-7)
    assert dev <= 2 * dref + 1e-6


def test_compensated_positions_1000_steps():
    """compensated_positions keeps a float low word per coordinate: after
    1000 steps of a 1,000-bead headline network the read-back positions
    differ from the FP64 oracle's trajectory by their own float rounding
    (half an ulp of the coordinate) plus at most 5e-7 A, while the plain FP32
    state has drifted by orders of magnitude more."""
    s = workloads.generate(1, seed=1, n=1000)
    runs = {}
    for comp in (0, 1):
        p = s.params.copy()
        p.compensated_positions = comp
        net = SpringNetwork.from_system(s, params=p)
        net.run(1000)
        runs[comp] = net.getState()[0].astype(np.float64)
        net.close()
    o = po.Oracle(s)
    o.step(po.F64, 1000)
    xd, _ = o.state(po.F64)
    o.close()
    xd = xd.astype(np.float64)
    half_ulp = 0.5 * np.spacing(np.abs(xd).astype(np.float32)).astype(np.float64)
    dc = np.abs(runs[1] - xd)
    dp = np.linalg.norm(runs[0] - xd, axis=1).max()
    dmax = np.linalg.norm(runs[1] - xd, axis=1).max()
    _log(f"  1000 steps: compensated max dev {dmax:.3e} A, plain {dp:.3e} A; "
         f"worst excess over half an ulp {(dc - half_ulp).max():.3e} A")
    assert (dc <= half_ulp + 5e-7).all()
    assert dmax <= 1.5e-6
    assert dp >= 10 * dmax


def test_headline_250k_forces_and_cells():
    """The headline network at its full size (250,000 beads, springs < 8 A,
    Amber 12-6 + Coulomb eps(r) = r at 12 A): the reference's cell lists bit
    for bit, and every particle's force and the six energies within 1e-5 of
    the FP64 oracle, at the start and after 20 steps."""
    s = workloads.generate(1, seed=1)
    _log("headline: n", s.n, "springs", s.m)
    net = SpringNetwork.from_system(s)
    net.setup()
    o = po.Oracle(s)
    g = net.getCells(_abi.NS_STERIC)
    r = o.cells(_abi.NS_STERIC)
    for key in ("dims", "box", "cell_of", "start", "members"):
        assert np.array_equal(np.asarray(g[key]), np.asarray(r[key])), key
    p = s.params.copy()
    p.viscosity_enable = 0
    for steps in (0, 20):
        if steps:
            net.setParams(s.params)
            net.run(steps)
            x, v = net.getState()
            o.set_state(x, v)
        net.setParams(p)
        o.set_params(p)
        F, E = net.computeForces()
        Fd, Ed = o.compute_forces(po.F64)
        per, glob = force_errors(F, Fd, (s.flags & 1) == 0)
        _log(f"  after {steps} steps: per-particle {per:.3g} global {glob:.3g}")
        assert per <= FORCE_RTOL
        for slot in range(6):
            err = energy_error(E[slot], Ed[slot], floor=1e-3)
            assert err <= ENERGY_RTOL, (steps, slot, E[slot], Ed[slot], err)
    o.close()
    net.close()
