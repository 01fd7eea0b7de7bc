"""Pins the oracle's probe particle (SpringNetwork.cpp:154-182,749-766,
Particle.cpp:434-465) beyond the reference's own trajectories
(tests/golden/probe600.npz, probe_amber800.npz, checked bit for bit in
test_oracle.py):

* FP64 forces are energy gradients: the particles' forces include the probe
  pair terms of the total energy, and the probe's own force is minus the
  gradient of that energy in the probe's position;
* Newton's third law: what the probe adds to the particles is the opposite of
  the force on the probe (viscosity off);
* the probe's Euler step from rest is the closed form x1 = x0 + dt^2 F/m;
* the probe never enters the pair searches (pair sets with and without it are
  equal) and takes no part in a term left off.
"""
import numpy as np
import pytest

from biospring_amd import _abi, workloads
from oracle import pyoracle as po


def _probe_system(mode=1, steric=1, elec=1, charge=0.7):
    s = workloads.generate(1, seed=31, n=400)
    p = s.params.copy()
    p.dielectric_model = 0
    p.steric_mode = mode
    p.viscosity_enable = 0
    x = s.x.reshape(-1, 3)
    p.probe_enable, p.probe_enable_steric, p.probe_enable_electrostatic = 1, steric, elec
    p.probe_x, p.probe_y, p.probe_z = float(x[:, 0].max() + 3.0), float(x[:, 1].mean()), float(x[:, 2].mean())
    p.probe_mass, p.probe_radius, p.probe_epsilon, p.probe_charge = 7.0, 3.0, 0.6, charge
    return s, p


def _energy(o):
    _, E = o.compute_forces(po.F64)
    return E[_abi.E_SPRING] + E[_abi.E_ELECTROSTATIC] + E[_abi.E_STERIC]


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_probe_forces_are_energy_gradients(mode):
    s, p = _probe_system(mode=mode)
    c = po.constants()["spring"]
    o = po.Oracle(s, params=p)
    x0 = s.x.reshape(-1, 3).astype(np.float64)
    F, _ = o.compute_forces(po.F64)
    _, _, fprobe = o.probe(po.F64)
    h = 1e-5
    # the beads nearest the probe carry its force
    pp = np.array([p.probe_x, p.probe_y, p.probe_z], np.float64)
    near = np.argsort(np.linalg.norm(x0 - pp, axis=1))[:6]
    pairs0 = [o.pairs(k, mode=po.F64) for k in (0, 1)]
    checked = 0
    for i in near:
        for a in range(3):
            e = []
            for sgn in (1.0, -1.0):
                xx = x0.copy()
                xx[i, a] += sgn * h
                o.set_state_f64(xx)
                if not all(np.array_equal(o.pairs(k, mode=po.F64), pairs0[k]) for k in (0, 1)):
                    break
                e.append(_energy(o))
            if len(e) < 2:
                continue
            want = -(e[0] - e[1]) / (2 * h) * c
            scale = max(abs(F[i, a]), 1e-3 * np.abs(F[i]).max(), 1e-12)
            assert abs(F[i, a] - want) <= 2e-5 * scale + 1e-9 * np.abs(F).max(), (i, a, F[i, a], want)
            checked += 1
    assert checked >= 12
    o.close()
    # the probe's force: minus the gradient in the probe's own position
    for a in range(3):
        e = []
        for sgn in (1.0, -1.0):
            q = p.copy()
            setattr(q, ("probe_x", "probe_y", "probe_z")[a], float(pp[a] + sgn * 1e-3))
            oo = po.Oracle(s, params=q)
            e.append(_energy(oo))
            oo.close()
        # the stored position is a float: take the actual offsets
        dx = (np.float32(pp[a] + 1e-3) - np.float32(pp[a] - 1e-3)).astype(np.float64)
        want = -(e[0] - e[1]) / dx * c
        assert abs(fprobe[a] - want) <= 1e-4 * np.abs(fprobe).max() + 1e-12, (a, fprobe[a], want)


@pytest.mark.parametrize("mode", [po.F32, po.F64])
def test_probe_newton_third_law(mode):
    s, p = _probe_system()
    q = p.copy()
    q.probe_enable = 0
    o1, o0 = po.Oracle(s, params=p), po.Oracle(s, params=q)
    F1, E1 = o1.compute_forces(mode)
    F0, E0 = o0.compute_forces(mode)
    _, _, fp = o1.probe(mode)
    dyn = (s.flags & 1) == 0
    add = (F1 - F0)[dyn].sum(axis=0)
    tol = (1e-5 if mode == po.F32 else 1e-10) * np.abs(F1 - F0).sum()
    assert np.abs(add + fp).max() <= tol, (add, fp)
    assert np.abs(fp).max() > 0.0
    # the probe's pair energies are added in full to the steric and Coulomb totals
    assert E1[_abi.E_STERIC] != E0[_abi.E_STERIC] and E1[_abi.E_ELECTROSTATIC] != E0[_abi.E_ELECTROSTATIC]
    # the pair sets are those without the probe (it is never a neighbour)
    for k in (0, 1):
        assert np.array_equal(o1.pairs(k, mode=mode), o0.pairs(k, mode=mode))
    o1.close()
    o0.close()


def test_probe_terms_switch_independently():
    s, p = _probe_system(steric=1, elec=0)
    q = p.copy()
    q.probe_enable = 0
    o1, o0 = po.Oracle(s, params=p), po.Oracle(s, params=q)
    _, E1 = o1.compute_forces(po.F64)
    _, E0 = o0.compute_forces(po.F64)
    assert E1[_abi.E_ELECTROSTATIC] == E0[_abi.E_ELECTROSTATIC]
    assert E1[_abi.E_STERIC] != E0[_abi.E_STERIC]
    # an uncharged probe with the Coulomb term on adds no Coulomb energy
    s2, p2 = _probe_system(steric=0, elec=1, charge=0.0)
    q2 = p2.copy()
    q2.probe_enable = 0
    a, b = po.Oracle(s2, params=p2), po.Oracle(s2, params=q2)
    Fa, Ea = a.compute_forces(po.F64)
    Fb, Eb = b.compute_forces(po.F64)
    assert np.array_equal(Fa, Fb) and Ea[_abi.E_ELECTROSTATIC] == Eb[_abi.E_ELECTROSTATIC]


def test_probe_euler_closed_form():
    """From rest, one step moves the probe by dt^2 F / m with its force at
    the start (viscosity on: -gamma v0 = 0), and its kinetic energy is added
    to the step's."""
    s, p = _probe_system()
    p.viscosity_enable = 1
    o = po.Oracle(s, params=p)
    o.compute_forces(po.F64)
    x0, v0, f0 = o.probe(po.F64)
    assert np.all(v0 == 0.0)
    _, E1 = o.step(po.F64, 1)
    x1, v1, f1 = o.probe(po.F64)
    dt, m = float(p.timestep), float(p.probe_mass)
    assert np.allclose(f1, f0, rtol=0, atol=1e-15 * np.abs(f0).max())
    assert np.allclose(v1, dt * f0 / m, rtol=1e-14, atol=0)
    assert np.allclose(x1, x0 + dt * dt * f0 / m, rtol=1e-14, atol=0)
    ck = po.constants()["kinetic"]
    ke = 0.5 * m * float(v1 @ v1) * ck
    _, v = o.state(po.F64)
    dyn = (s.flags & 1) == 0
    ke_particles = float(np.sum(0.5 * s.mass[dyn].astype(np.float64) * np.sum(v[dyn] ** 2, axis=1)) * ck)
    assert ke > 0.0
    assert E1[_abi.E_KINETIC] == pytest.approx(ke + ke_particles, rel=1e-12)
    o.close()
