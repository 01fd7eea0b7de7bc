"""Spring network from a distance cutoff on the GPU (bs_springs_from_cutoff,
in place of Topology::add_springs_from_cutoff, src/topology/Topology.hpp:
105-138).  The bar is bit-exact: the same (first, second) set as the
reference's golden fixtures and as the oracle, and bit-identical float rest
lengths.  At full size (250,000 beads, springs < 8 A) the GPU's set is also
compared with the workload generator's own springs, built by the same rule on
the host."""
import sys

import numpy as np
import pytest

from biospring_amd import _abi, workloads
from biospring_amd.network import SpringNetwork
from oracle import pyoracle as po
from tests.test_topology import CASES, brute_force, check_merged, load_case, same_springs

pytestmark = pytest.mark.gpu


def _log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _gpu_springs(x, flags, cutoff, stiffness=1.0, params=None):
    net = SpringNetwork(params if params is not None else workloads.params(1))
    net.setParticles(x, flags=flags)
    m = net.springsFromCutoff(cutoff, stiffness)
    a, b, l0, k = net.getSprings()
    assert m == len(a)
    assert np.all(k == np.float32(stiffness))
    return net, (a, b, l0)


@pytest.mark.parametrize("name", CASES)
def test_matches_reference_fixture(name):
    c = load_case(name)
    _log(f"topology fixture {name}: n={len(c['in_flags'])}")
    net, got = _gpu_springs(c["in_x"], c["in_flags"], float(c["cutoff"]), 0.7)
    same_springs(got, (c["ref_id1"], c["ref_id2"], c["ref_l0"]))
    key = got[0].astype(np.int64) * len(c["in_flags"]) + got[1]
    assert np.all(np.diff(key) > 0)  # ordered by (first, second)
    net.close()


def test_keeps_loaded_springs_like_the_reference():
    """Springs loaded before (bs_set_springs) stay first with their own rest
    length and stiffness; their pairs get no second spring; positions are
    the current ones after stepping."""
    c = load_case("topo_merge800")
    _log("topology merge fixture")
    net = SpringNetwork(workloads.params(1))
    net.setParticles(c["in_x"], flags=c["in_flags"])
    net.setSprings(c["pre_id1"], c["pre_id2"], c["pre_l0"], c["pre_k"])
    net.run(3)
    net.setState(c["in_x"], None)
    added = net.springsFromCutoff(float(c["cutoff"]), float(c["stiffness"]))
    got = net.getSprings()
    assert added == len(got[0]) - len(c["pre_id1"])
    check_merged(got, c)
    # after steps the construction sees the moved positions
    net.run(5)
    x, _ = net.getState()
    net.setSprings([], [], [], [])
    net.springsFromCutoff(float(c["cutoff"]))
    same_springs(net.getSprings()[:3], po.springs_from_cutoff(x, c["in_flags"], float(c["cutoff"])))
    net.close()


@pytest.mark.parametrize("seed,n,cutoff,span,static_frac", [(11, 20000, 7.1, 60.0, 0.0), (12, 20000, 9.3, 60.0, 0.2),
                                                            (13, 8000, 12.0, 35.0, 0.5), (14, 3000, 3.0, 400.0, 0.1)])
def test_matches_oracle_random(seed, n, cutoff, span, static_frac):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-span, span, (n, 3)).astype(np.float32)
    flags = np.where(rng.random(n) < static_frac, _abi.FLAG_STATIC, 0).astype(np.uint8)
    _log(f"topology random: seed={seed} n={n} cutoff={cutoff}")
    net, got = _gpu_springs(x, flags, cutoff)
    same_springs(got, po.springs_from_cutoff(x, flags, cutoff))
    net.close()


def _clumps():
    # 300 clumps of 10 particles spread over 6,000 A: more cells a side than
    # the grid's 256, so cells grow beyond the cutoff
    rng = np.random.default_rng(3)
    c = rng.uniform(-3000, 3000, (300, 1, 3))
    return (c + rng.uniform(-4, 4, (300, 10, 3))).reshape(-1, 3).astype(np.float32)


def test_degenerate_layouts():
    # a line of particles (zero extent on two axes), a lone particle, all
    # particles at one point, clumps over a box of 256+ cells a side, and a
    # crowd in which every pair is linked
    cases = [(np.stack([np.arange(500, dtype=np.float32) * 0.7, np.zeros(500, np.float32),
                        np.zeros(500, np.float32)], 1), 2.0),
             (np.array([[1.0, 2.0, 3.0]], np.float32), 5.0),
             (np.full((40, 3), 3.25, np.float32), 1.0),
             (_clumps(), 6.0),
             # rows of up to 1,499 springs: longer than a wave sorts in LDS
             (np.random.default_rng(4).uniform(-1.5, 1.5, (1500, 3)).astype(np.float32), 10.0)]
    for x, rc in cases:
        _log(f"topology degenerate: n={len(x)} cutoff={rc}")
        net, got = _gpu_springs(x, None, rc)
        same_springs(got, brute_force(x, None, rc))
        net.close()
    net = SpringNetwork(workloads.params(1))
    net.setParticles(np.zeros((4, 3), np.float32))
    for bad in (0.0, 5e-7, float("nan"), float("inf")):
        with pytest.raises(_abi.EngineError):
            net.springsFromCutoff(bad)
    net.close()


def test_full_size_matches_generator_and_oracle():
    s = workloads.generate(1, seed=1)
    _log(f"topology full size: n={s.n} springs={s.m}")
    net, got = _gpu_springs(s.x, s.flags, 8.0)
    a, b, l0 = got
    assert np.array_equal(a, s.id1) and np.array_equal(b, s.id2)
    assert np.array_equal(l0.view(np.uint32), s.l0.view(np.uint32))
    same_springs(got, po.springs_from_cutoff(s.x, s.flags, 8.0))
    net.close()


def test_constructed_network_steps_like_the_oracle():
    """Construction followed by the step: forces of the GPU-built network
    agree with the FP64 oracle on the oracle-built one (1e-5), so the loaded
    spring table is the one the step uses."""
    s = workloads.generate(1, seed=21, n=3000)
    p = s.params.copy()
    p.viscosity_enable = 0
    net = SpringNetwork(p)
    net.setParticles(*s.particle_arrays())
    m = net.springsFromCutoff(9.5, 0.8)
    a, b, l0 = po.springs_from_cutoff(s.x, s.flags, 9.5)
    assert m == len(a)
    t = workloads.System(x=s.x, v=s.v, mass=s.mass, charge=s.charge, radius=s.radius, eps=s.eps, sasa=s.sasa,
                         transfer=s.transfer, hydro=s.hydro, flags=s.flags, id1=a, id2=b, l0=l0,
                         k=np.full(len(a), 0.8, np.float32))
    t.params = p
    F, E = net.computeForces()
    o = po.Oracle(t, params=p)
    Fd, Ed = o.compute_forces(po.F64)
    o.close()
    err = np.linalg.norm(F - Fd, axis=1).max() / np.linalg.norm(Fd, axis=1).max()
    _log(f"  constructed network: springs {m}, force err {err:.2e}")
    assert err <= 1e-5
    assert abs(E[0] - Ed[0]) <= 1e-5 * max(abs(Ed[0]), 1e-3)
    net.run(5)
    assert np.isfinite(net.getState()[0]).all()
    net.close()


def test_device_table_steps_like_an_uploaded_one():
    """Construction leaves the spring table on the device and the setup builds
    the spring rows and exclusions from it there; the same table given from
    the host (bs_set_springs) steps to the same bits, with static beads and
    rigid bodies (whose internal springs leave the force rows), and after
    loaded springs plus a construction."""
    s = workloads.generate(0, seed=23, n=4000)
    fl = s.flags.copy()
    fl[::11] |= _abi.FLAG_STATIC
    p = s.params.copy()
    p.rigidbody_enable = 1
    bodies = (np.arange(100, 400), np.arange(2000, 2300))

    def make():
        net = SpringNetwork(p)
        net.setParticles(s.x, s.v, s.mass, s.charge, s.radius, s.eps, s.sasa, s.transfer, s.hydro, fl)
        return net

    a = make()
    m = a.springsFromCutoff(11.0, 0.7)
    for g in bodies:
        a.addRigidBody(g)
    a.run(12)
    ids = a.getSprings()
    assert len(ids[0]) == m
    b = make()
    b.setSprings(*ids)
    for g in bodies:
        b.addRigidBody(g)
    b.run(12)
    for u, v in zip(a.getState(), b.getState()):
        assert np.array_equal(u, v)
    assert np.array_equal(a.computeForces()[0], b.computeForces()[0])
    # loaded springs kept ahead of constructed ones, then stepped
    c = make()
    c.setSprings(s.id1[:500], s.id2[:500], s.l0[:500], s.k[:500])
    mc = c.springsFromCutoff(9.0)
    got = c.getSprings()
    assert len(got[0]) == 500 + mc
    assert np.array_equal(got[0][:500], s.id1[:500]) and np.array_equal(got[3][:500], s.k[:500])
    c.run(3)
    d = make()
    d.setSprings(*got)
    d.run(3)
    assert np.array_equal(c.getState()[0], d.getState()[0])
    for net in (a, b, c, d):
        net.close()


def test_dense_rows_match_fp64_oracle():
    """Rows longer than the spring-row sort's LDS stage (1,024 keys) and
    clusters whose exclusions merge thousands of words: a 2,000-bead sphere
    linked within 30 A (up to 1,999 springs a bead), then stretched by
    0.3 A (rms).  Forces and energies against the FP64 oracle on the same
    springs."""
    s = workloads.generate(0, seed=24, n=2000)
    p = s.params.copy()
    p.viscosity_enable = 0
    net = SpringNetwork(p)
    net.setParticles(*s.particle_arrays())
    m = net.springsFromCutoff(30.0, 0.05)
    xs = (s.x.reshape(-1, 3) + np.random.default_rng(5).normal(0, 0.3, (s.n, 3))).astype(np.float32)
    net.setState(xs)
    a, b, l0, k = net.getSprings()
    rows = np.bincount(np.concatenate([a, b]), minlength=s.n)
    _log(f"dense network: {m} springs, longest row {rows.max()}")
    assert rows.max() > 1024
    t = workloads.System(x=s.x, v=s.v, mass=s.mass, charge=s.charge, radius=s.radius, eps=s.eps, sasa=s.sasa,
                         transfer=s.transfer, hydro=s.hydro, flags=s.flags, id1=a, id2=b, l0=l0, k=k)
    t.params = p
    F, E = net.computeForces()
    o = po.Oracle(t, params=p)
    o.set_state(xs, None)
    Fd, Ed = o.compute_forces(po.F64)
    o.close()
    err = np.linalg.norm(F - Fd, axis=1).max() / np.linalg.norm(Fd, axis=1).max()
    _log(f"  force err {err:.2e}, spring energy {E[0]:.6g} vs {Ed[0]:.6g}")
    assert err <= 1e-5
    for c in range(3):
        assert abs(E[c] - Ed[c]) <= 1e-5 * max(abs(Ed[c]), 1e-3)
    net.close()


def test_systems_stay_apart():
    rng = np.random.default_rng(10)
    xa = rng.uniform(-30, 30, (6000, 3)).astype(np.float32)
    xb = (xa + np.float32(0.25)).astype(np.float32)
    x = np.concatenate([xa, xb])
    net = SpringNetwork(workloads.params(1))
    net.setParticles(x)
    net.setSystems(np.array([0, 6000, 12000], np.int64))
    _log("topology in two overlapping systems: n=12000")
    net.springsFromCutoff(8.0)
    a, b, l0, _ = net.getSprings()
    assert np.all((a < 6000) == (b < 6000))
    for lo, xs in ((0, xa), (6000, xb)):
        sel = a >= lo if lo else a < 6000
        same_springs((a[sel] - lo, b[sel] - lo, l0[sel]), po.springs_from_cutoff(xs, None, 8.0))
    net.run(2)
    net.close()


def test_between_groups_matches_reference_and_oracle():
    c = load_case("topo_groups900")
    _log("topology between groups: fixture")
    net = SpringNetwork(workloads.params(1))
    net.setParticles(c["in_x"], flags=c["in_flags"])
    net.springsBetweenGroupsFromCutoff(float(c["cutoff"]), c["in_groups"])
    same_springs(net.getSprings()[:3], (c["ref_id1"], c["ref_id2"], c["ref_l0"]))
    net.close()
    rng = np.random.default_rng(12)
    x = rng.uniform(-40, 40, (20000, 3)).astype(np.float32)
    g = rng.integers(0, 3, 20000).astype(np.int32)
    _log("topology between groups: n=20000, 3 groups")
    net = SpringNetwork(workloads.params(1))
    net.setParticles(x)
    net.springsBetweenGroupsFromCutoff(7.5, g)
    same_springs(net.getSprings()[:3], po.springs_from_cutoff(x, None, 7.5, groups=g))
    net.close()
