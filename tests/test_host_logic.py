"""CPU check of the engine's host logic (buffer management, call sequencing).

bs_api.cpp is compiled against a recording fake of the HIP runtime whose
launchers verify that every buffer a kernel would touch is live and large
enough (biospring_amd/csrc/hostcheck).  A violation aborts the child process,
so each sequence runs in a subprocess and the test checks its exit status.
"""
import os
import subprocess
import sys
import textwrap

import pytest
from conftest import make_locked

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "biospring_amd", "csrc")
HOSTCHECK = os.path.join(CSRC, "_build", "libbs_hostcheck.so")


@pytest.fixture(scope="module")
def hostcheck_lib():
    make_locked("-C", CSRC, "hostcheck")
    assert os.path.exists(HOSTCHECK)
    return HOSTCHECK


def _run(body, lib):
    code = textwrap.dedent(f"""
        import ctypes, sys
        sys.path.insert(0, {ROOT!r})
        import numpy as np
        from biospring_amd import _abi, workloads
        from biospring_amd.network import SpringNetwork
        L = _abi.bind(ctypes.CDLL({lib!r}))
        L.hc_live_allocations.restype = ctypes.c_int64
        """) + textwrap.dedent(body)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    return r.stdout


def test_setup_step_reparam_sequence(hostcheck_lib):
    out = _run("""
        s = workloads.generate(1, seed=3, n=3000)
        base = L.hc_live_allocations()
        net = SpringNetwork.from_system(s, library=L)
        net.setup()
        net.run(25)
        p = s.params.copy(); p.viscosity_enable = 0
        net.setParams(p)                      # frees device state, lazy re-setup
        F, E = net.computeForces()
        for kind in (0, 1, 2):
            net.getCells(kind)
        net.getPairs(0); net.getPairs(1); net.countPairs(0)
        x, v = net.getState()
        net.setState(x + 0.01, v)
        net.run(3)
        p2 = p.copy(); p2.electrostatic_cutoff = 9.0; p2.steric_cutoff = 6.0
        net.setParams(p2)                     # different list plan
        net.computeForces(); net.run(4); net.getPairs(1)
        p3 = p2.copy(); p3.compensated_positions = 1; p3.hydrophobicity_enable = 1
        net.setParams(p3)
        net.run(3); net.computeForces()
        net.profile(True); net.run(5); net.getProfile()
        net.close()
        print("leak", L.hc_live_allocations() - base)
    """, hostcheck_lib)
    assert "leak 0" in out, out


def test_small_and_degenerate_systems(hostcheck_lib):
    _run("""
        # one particle, no springs; all static; no terms enabled
        p = _abi.default_params(); p.steric_enable = 1; p.electrostatic_enable = 1
        for n in (1, 2, 5):
            net = SpringNetwork(p, library=L)
            net.setParticles(np.arange(3 * n, dtype=np.float32), charge=np.ones(n, np.float32))
            net.setup(); net.run(2); net.computeForces(); net.getPairs(0); net.getCells(1)
            net.close()
        net = SpringNetwork(p, library=L)
        net.setParticles(np.zeros(30, np.float32), flags=np.ones(10, np.uint8))
        net.run(2); net.computeForces(); net.close()
        q = _abi.default_params()
        net = SpringNetwork(q, library=L)
        s = workloads.generate(0, seed=2, n=300)
        net.setParticles(*s.particle_arrays()); net.setSprings(*s.spring_arrays())
        net.run(3); net.computeForces(); net.getCells(0); net.close()
    """, hostcheck_lib)


def test_cluster_lists_cover_reference_pairs(hostcheck_lib):
    """The host-built exclusion lists and the cluster search ranges, run
    through the checker's CPU emulation of the list build, reproduce the
    oracle's processed pair sets exactly (also after motion and a rebuild),
    for the cluster and the queue layouts."""
    _run("""
        from oracle import pyoracle as po
        for cfg, n, seed, lay in ((0, 0, 2, 0), (1, 3000, 4, 0), (1, 12000, 5, 0), (0, 0, 2, 2), (1, 12000, 5, 2)):
            s = workloads.generate(cfg, seed=seed, n=n)
            p = s.params.copy(); p.pair_layout = lay
            net = SpringNetwork.from_system(s, params=p, library=L)
            net.setup(); net.run(2)
            o = po.Oracle(s)
            for kind in (0, 1):
                assert np.array_equal(net.getPairs(kind), o.pairs(kind)), (cfg, n, kind)
            rng = np.random.default_rng(seed)
            x = s.x.reshape(-1, 3) + rng.normal(0, 0.3, (s.n, 3)).astype(np.float32)
            net.setState(x)
            o.set_state(x)
            for kind in (0, 1):
                assert np.array_equal(net.getPairs(kind), o.pairs(kind)), (cfg, n, kind, "moved")
    """, hostcheck_lib)


def test_argument_errors(hostcheck_lib):
    _run("""
        net = SpringNetwork(library=L)
        try:
            net.setSprings([0], [1], [1.0], [1.0])
            raise SystemExit("springs before particles accepted")
        except _abi.EngineError as e:
            assert e.code == _abi.BS_ERR_STATE
        net.setParticles(np.zeros(6, np.float32))
        for a, b in ((0, 2), (1, 1)):
            try:
                net.setSprings([a], [b], [1.0], [1.0])
                raise SystemExit("bad spring accepted")
            except _abi.EngineError as e:
                assert e.code == _abi.BS_ERR_ARG, e
        net.setSprings([0, 1, 0], [1, 0, 1], [1.0, 2.0, 3.0], [1.0, 1.0, 1.0])
        assert L.bs_num_springs(net._h) == 1   # duplicates ignored like addSpring
    """, hostcheck_lib)


def test_rigid_bodies_grids_sequence(hostcheck_lib):
    """Rigid bodies, field grids, burying and the insertion vector through
    setup, steps, re-parametrisation and body/grid changes."""
    out = _run("""
        s = workloads.generate(0, seed=4, n=900)
        s.flags[::17] = 1
        base = L.hc_live_allocations()
        p = s.params.copy(); p.rigidbody_enable = 1
        p.potentialgrid_enable = 1; p.densitygrid_enable = 1
        net = SpringNetwork.from_system(s, params=p, library=L)
        assert net.addRigidBody(np.arange(0, 200)) == 1
        assert net.addRigidBody(np.arange(150, 300)) == 1     # overlaps: ignored
        assert net.addRigidBody(np.arange(400, 520)) == 2
        x = s.x.reshape(-1, 3)
        org = x.min(0) - 2.0; shape = np.array([12, 10, 14]); cell = np.array([3.0, 3.5, 2.5])
        phi = np.random.default_rng(1).normal(size=shape).astype(np.float32)
        net.setGrid(_abi.GRID_POTENTIAL, org, shape, cell, phi)
        net.setGrid(_abi.GRID_DENSITY, org, shape, cell, phi * 0.5)
        net.setBurying(np.linspace(0.5, 2.0, s.n, dtype=np.float32))
        net.setup(); net.run(5); net.computeForces()
        b = net.getRigidBody(1)
        iv = net.getInsertionVector(3, 800)
        p2 = p.copy(); p2.rigidbody_enable = 0
        net.setParams(p2); net.run(3)                         # body state kept across re-setup
        net.addRigidBody(np.arange(600, 650)); net.run(2)
        net.setGrid(_abi.GRID_DENSITY, org, shape, cell, None); net.run(2)
        net.addRigidBody([]); net.run(2)                      # removes every body
        assert net.numRigidBodies() == 0
        net.close()
        print("leak", L.hc_live_allocations() - base)
    """, hostcheck_lib)
    assert "leak 0" in out, out


def test_rigid_and_grid_argument_errors(hostcheck_lib):
    _run("""
        s = workloads.generate(0, seed=5, n=100)
        net = SpringNetwork.from_system(s, library=L)
        for bad in ([0, 100], [-1], [3, 3]):
            try:
                net.addRigidBody(bad)
                raise SystemExit("accepted %r" % (bad,))
            except _abi.EngineError as e:
                assert e.code == _abi.BS_ERR_ARG
        for shape, cell in (([0, 2, 2], [1, 1, 1]), ([2, 2, 2], [1, 0, 1])):
            try:
                net.setGrid(0, [0, 0, 0], shape, cell, np.zeros(max(int(np.prod(shape)), 1), np.float32))
                raise SystemExit("accepted grid")
            except (_abi.EngineError, ValueError):
                pass
        try:
            net.getInsertionVector(0, 100)
            raise SystemExit("accepted ivec")
        except _abi.EngineError as e:
            assert e.code == _abi.BS_ERR_ARG
        net.close()
    """, hostcheck_lib)


def test_batch_of_systems_sequence(hostcheck_lib):
    out = _run("""
        s, off, parts = workloads.batch(9, seed=3, nmax=1500)
        base = L.hc_live_allocations()
        net = SpringNetwork.from_system(s, library=L)
        net.setSystems(off)
        assert net.numSystems() == 9
        net.setup(); net.run(4); net.computeForces()
        E = net.getSystemEnergies(); assert E.shape == (9, 6)
        net.addRigidBody(np.arange(off[2], off[2] + 50)); net.run(2)
        try:
            net.getCells(0)
            raise SystemExit("cells accepted for a batch")
        except _abi.EngineError:
            pass
        net.getPairs(1)
        bad = off.copy(); bad[3] = bad[2]
        try:
            net.setSystems(bad)
            raise SystemExit("accepted empty system")
        except _abi.EngineError as e:
            assert e.code == _abi.BS_ERR_ARG
        net.setSystems([0, s.n]); net.run(2)
        net.close()
        print("leak", L.hc_live_allocations() - base)
    """, hostcheck_lib)
    assert "leak 0" in out, out


def test_adjoint_sequence(hostcheck_lib):
    """Checkpointed reverse sweep: buffers, loss plumbing (the stand-in step
    leaves positions unchanged) and the edge-to-spring map (the stand-in
    counts one unit of dk per reverse step on the owning edge of each spring
    with a dynamic end)."""
    out = _run("""
        s = workloads.generate(0, seed=4, n=400)
        fl = s.flags.copy(); fl[::7] |= _abi.FLAG_STATIC; s.flags = fl
        base = L.hc_live_allocations()
        net = SpringNetwork.from_system(s, library=L)
        net.setup()
        x0, _ = net.getState()
        y = x0 + np.random.default_rng(0).normal(0, 0.3, x0.shape).astype(np.float32)
        ref = float(np.mean(np.sum((x0.astype(np.float64) - y) ** 2, axis=1)))
        a, b, _, _ = s.spring_arrays()
        dyn = (fl & _abi.FLAG_STATIC) == 0
        ends = (dyn[a] | dyn[b]).astype(int)
        for T, K in ((7, 3), (5, 0), (4, 4), (1, 9), (0, 0)):
            loss, dk, dl0, dq = net.adjoint(T, y, K)
            assert abs(loss - ref) <= 1e-12 * ref, (T, K, loss, ref)
            assert dk.shape == (net_m := int(L.bs_num_springs(net._h)),) and dq.shape == (400,)
            assert np.array_equal(dk, T * ends), (T, K)
        # two-phase form: record once, sweep twice; state changes drop the record
        # (bs_adjoint leaves one too)
        net.run(1)
        try:
            net.adjointBackward(np.zeros((400, 3)))
            raise SystemExit("backward without a record")
        except _abi.EngineError as e:
            assert e.code == _abi.BS_ERR_STATE
        net.adjointForward(6, 4)
        for seed in (np.ones((400, 3)), None):
            dk2, dl2, dq2 = net.adjointBackward(seed, np.ones((400, 3)))
            assert np.array_equal(dk2, 6 * ends)
        net.run(1)
        try:
            net.adjointBackward(np.ones((400, 3)))
            raise SystemExit("backward after a step")
        except _abi.EngineError as e:
            assert e.code == _abi.BS_ERR_STATE
        p = s.params.copy(); p.imp_enable = 1
        net.setParams(p)
        try:
            net.adjoint(3, y)
            raise SystemExit("adjoint accepted IMPALA")
        except _abi.EngineError as e:
            assert e.code == _abi.BS_ERR_ARG
        p.imp_enable = 0; p.compensated_positions = 1
        net.setParams(p)
        try:
            net.adjoint(3, y)
            raise SystemExit("adjoint accepted compensated positions")
        except _abi.EngineError as e:
            assert e.code == _abi.BS_ERR_ARG
        p.compensated_positions = 0
        net.setParams(p)
        for bad in (-1,):
            try:
                net.adjoint(bad, y)
                raise SystemExit("adjoint accepted a negative step count")
            except _abi.EngineError as e:
                assert e.code == _abi.BS_ERR_ARG
        net.addRigidBody(np.arange(10, 30))
        try:
            net.adjoint(2, y)
            raise SystemExit("adjoint accepted rigid bodies")
        except _abi.EngineError as e:
            assert e.code == _abi.BS_ERR_ARG
        net.close()
        print("leak", L.hc_live_allocations() - base)
    """, hostcheck_lib)
    assert "leak 0" in out, out


def test_noalloc_backward_allocates_nothing(hostcheck_lib):
    """bs_adjoint_backward_noalloc: no hipMalloc during the call (the
    allocating form makes some), every buffer of the sweep inside the fixed
    array (the stand-in launchers check extents against it), the same
    results as the allocating form, and a system too large for the array
    refused with BS_ERR_ARG and the size it needs."""
    out = _run("""
        L.hc_malloc_calls.restype = ctypes.c_int64
        s = workloads.generate(0, seed=4, n=400)
        fl = s.flags.copy(); fl[::7] |= _abi.FLAG_STATIC; s.flags = fl
        net = SpringNetwork.from_system(s, library=L)
        net.setup()
        net.adjointForward(5, 2)
        g = np.random.default_rng(1).normal(size=(400, 3))
        c0 = L.hc_malloc_calls()
        want = net.adjointBackward(g, g)
        c1 = L.hc_malloc_calls()
        got = net.adjointBackward(g, g, noalloc=True)
        c2 = L.hc_malloc_calls()
        assert c1 > c0 and c2 == c1, (c0, c1, c2)
        for a, b in zip(want, got):
            assert np.array_equal(a, b)
        for seeds in ((None, g), (g, None)):
            net.adjointBackward(*seeds, noalloc=True)
        assert L.hc_malloc_calls() == c2
        try:
            net.adjointBackward(None, None, noalloc=True)
            raise SystemExit("accepted no seed")
        except _abi.EngineError as e:
            assert e.code == _abi.BS_ERR_ARG
        net.profile(True)
        try:
            net.adjointBackward(g, noalloc=True)
            raise SystemExit("accepted profiling")
        except _abi.EngineError as e:
            assert e.code == _abi.BS_ERR_STATE
        net.profile(False)
        net.run(1)
        try:
            net.adjointBackward(g, noalloc=True)
            raise SystemExit("backward without a record")
        except _abi.EngineError as e:
            assert e.code == _abi.BS_ERR_STATE
        net.close()
        big = workloads.generate(1, seed=2, n=20000)
        b = SpringNetwork.from_system(big, library=L)
        b.adjointForward(1, 1)
        try:
            b.adjointBackward(np.ones((20000, 3)), noalloc=True)
            raise SystemExit("accepted a system beyond the fixed array")
        except _abi.EngineError as e:
            assert e.code == _abi.BS_ERR_ARG and "bytes of device scratch" in str(e), str(e)
        dk, _, _ = b.adjointBackward(np.ones((20000, 3)), noalloc=False)
        assert dk.shape == (big.m,)
        b.close()
        print("noalloc ok")
    """, hostcheck_lib)
    assert "noalloc ok" in out, out


def test_interactive_updates(hostcheck_lib):
    _run("""
        s = workloads.generate(2, seed=7, n=500)
        p = workloads.impala_params(s, "double")
        net = SpringNetwork(p, library=L)
        try:
            net.setSolventAccessibilitySurfaces(np.zeros(500, np.float32))
            raise SystemExit("sasa accepted before particles")
        except _abi.EngineError:
            pass
        net.setParticles(*s.particle_arrays())
        net.setSprings(*s.spring_arrays())
        net.setSolventAccessibilitySurfaces(s.sasa + 1)      # before setup: host copy only
        assert np.array_equal(net.getParticleProperties()["sasa"], (s.sasa + 1).astype(np.float32))
        net.run(3)
        net.setSolventAccessibilitySurfaces(s.sasa + 2)      # after setup: device copy too
        net.setMembrane(upper_offset=10.0)                   # NaN keeps the others
        net.run(3)
        q = s.params.copy(); q.imp_enable = 0
        net.setParams(q)
        net.setMembrane(1.0, 2.0, 3.0, 4.0)                  # ignored with IMPALA off
        net.run(1)
        net.close()
        """, hostcheck_lib)


def test_overflow_lengthens_lists_and_resumes(hostcheck_lib):
    """A neighbour list overflows mid-call (the beads are packed after the
    lists were sized): the call stops at that step, the host lengthens the
    lists and resumes, and every step of the call is completed; with the
    CPU stand-in kernels positions stay put, so each rebuild sees the packed
    state."""
    out = _run('''
        s = workloads.generate(1, seed=3, n=3000)
        p = s.params.copy(); p.spring_enable = 0
        net = SpringNetwork.from_system(s, params=p, library=L)
        net.setup()
        cap0 = net.counters()["list_cap"]
        x = s.x.reshape(-1, 3).astype(np.float64)
        c = x.mean(axis=0)
        net.setState(((x - c) * 0.6 + c).astype(np.float32), np.zeros((s.n, 3), np.float32))
        net.run(7)
        cnt = net.counters()
        assert cnt["steps"] == 7, cnt
        assert cnt["list_cap"] > cap0 and cnt["max_list"] > cap0, (cap0, cnt)
        net.run(3)
        assert net.counters()["steps"] == 10
        net.close()
        print("ok", cap0, cnt["list_cap"])
        ''', hostcheck_lib)
    assert "ok" in out
