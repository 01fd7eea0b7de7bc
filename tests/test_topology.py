"""Spring network from a distance cutoff (Topology::add_springs_from_cutoff,
src/topology/Topology.hpp:105-138), CPU side.

The oracle's restatement (oracle/bs_oracle.c, orc_springs_from_cutoff) is
pinned three ways: the golden fixtures written by the reference's own
Topology (tests/golden/topo_*.npz), the compiled reference where it is built,
and brute force over all pairs, which is what the reference's own test checks
(tests/topology/test-Topology.cpp:149-170).  The engine's host logic (grid
sizing, row offsets, the spring table it loads) runs against the recording
fake of the HIP runtime, whose kernel stand-in follows the GPU kernel's scan.

Springs are compared as sets of (first, second) with first < second, rest
lengths bit for bit: the reference lists them by first particle, then in its
cell-scan order; the engine and the oracle list them by (first, second)."""
import os
import subprocess
import sys
import textwrap

import numpy as np
import pytest

from biospring_amd import _abi
from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
CSRC = os.path.join(ROOT, "biospring_amd", "csrc")
HOSTCHECK = os.path.join(CSRC, "_build", "libbs_hostcheck.so")
CASES = ("topo_cube1000", "topo_dense1200")


def load_case(name):
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    return {k: z[k] for k in z.files}


def canonical(i1, i2, l0):
    """(first, second, rest length) with first < second, sorted."""
    a = np.minimum(i1, i2).astype(np.int64)
    b = np.maximum(i1, i2).astype(np.int64)
    o = np.lexsort((b, a))
    return a[o], b[o], np.asarray(l0, np.float32)[o]


def same_springs(got, want):
    g, w = canonical(*got), canonical(*want)
    assert len(g[0]) == len(w[0]), (len(g[0]), len(w[0]))
    assert np.array_equal(g[0], w[0]) and np.array_equal(g[1], w[1])
    assert np.array_equal(g[2].view(np.uint32), w[2].view(np.uint32))


def brute_force(x, flags, cutoff, groups=None):
    """All pairs i < j with the FP64 distance of the float positions below
    the cutoff held as a float, static-static pairs left out (and pairs of
    one group, when groups are given)."""
    x = np.asarray(x, np.float32).reshape(-1, 3).astype(np.float64)
    n = len(x)
    ii, jj = np.triu_indices(n, 1)
    d = x[ii] - x[jj]
    dist = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
    keep = dist < float(np.float32(cutoff))
    if flags is not None:
        st = (np.asarray(flags) & _abi.FLAG_STATIC) != 0
        keep &= ~(st[ii] & st[jj])
    if groups is not None:
        g = np.asarray(groups)
        keep &= g[ii] != g[jj]
    return ii[keep], jj[keep], dist[keep].astype(np.float32)


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_fixture(name):
    c = load_case(name)
    got = po.springs_from_cutoff(c["in_x"], c["in_flags"], float(c["cutoff"]))
    assert np.all(got[0] < got[1])
    same_springs(got, (c["ref_id1"], c["ref_id2"], c["ref_l0"]))
    # ordered by (first, second)
    key = got[0].astype(np.int64) * len(c["in_flags"]) + got[1]
    assert np.all(np.diff(key) > 0)


def merged_expectation(c, new):
    """Springs after construction over loaded ones: the loaded springs in
    their order, then the `new` (first, second, l0) springs whose pairs were
    not loaded, with stiffness c["stiffness"]."""
    pre = {(min(a, b), max(a, b)) for a, b in zip(c["pre_id1"].tolist(), c["pre_id2"].tolist())}
    keep = np.array([(int(a), int(b)) not in pre for a, b in zip(new[0], new[1])], bool)
    m = int(keep.sum())
    return (np.concatenate([c["pre_id1"], new[0][keep]]), np.concatenate([c["pre_id2"], new[1][keep]]),
            np.concatenate([c["pre_l0"], new[2][keep]]),
            np.concatenate([c["pre_k"], np.full(m, np.float32(c["stiffness"]))]))


def check_merged(got, c):
    """got = (id1, id2, l0, k) of the engine: the loaded springs first, in
    order and unchanged, then the new ones equal as a set to the reference's."""
    m0 = len(c["pre_id1"])
    assert len(got[0]) == len(c["ref_id1"])
    for g, key in zip(got, ("id1", "id2", "l0", "k")):
        assert np.array_equal(np.asarray(g)[:m0], c["pre_" + key]), key
        assert np.array_equal(np.asarray(c["ref_" + key])[:m0], c["pre_" + key]), key
    same_springs(tuple(np.asarray(g)[m0:] for g in got[:3]), (c["ref_id1"][m0:], c["ref_id2"][m0:], c["ref_l0"][m0:]))
    assert np.all(np.asarray(got[3])[m0:] == np.float32(c["stiffness"]))
    assert np.all(c["ref_k"][m0:] == np.float32(c["stiffness"]))


def test_oracle_merge_fixture():
    """Springs loaded before keep their rest length and stiffness (also
    beyond the cutoff and between static particles); pairs they link get no
    second spring."""
    c = load_case("topo_merge800")
    new = po.springs_from_cutoff(c["in_x"], c["in_flags"], float(c["cutoff"]))
    exp = merged_expectation(c, new)
    check_merged(exp, c)
    assert len(new[0]) - (len(exp[0]) - len(c["pre_id1"])) == 25  # the loaded pairs within the cutoff


@pytest.mark.parametrize("name", CASES)
def test_fixture_matches_brute_force(name):
    """The reference's own test asserts that every pair closer than the
    cutoff is linked; here also that no other pair is."""
    c = load_case(name)
    same_springs(brute_force(c["in_x"], c["in_flags"], float(c["cutoff"])),
                 (c["ref_id1"], c["ref_id2"], c["ref_l0"]))


def test_dense_fixture_special_pairs():
    """Pairs between 9.3 and float(9.3) are linked (the float cutoff of
    NeighborSearch), coincident particles get a zero rest length, static
    pairs none."""
    c = load_case("topo_dense1200")
    x = c["in_x"].astype(np.float64)
    a, b, l0 = canonical(c["ref_id1"], c["ref_id2"], c["ref_l0"])
    d = np.sqrt(((x[a] - x[b]) ** 2).sum(1))
    assert int(((d > 9.3) & (d < float(np.float32(9.3)))).sum()) >= 40
    assert (l0 == 0).sum() == 1
    st = (c["in_flags"] & _abi.FLAG_STATIC) != 0
    assert not np.any(st[a] & st[b])
    assert st.sum() > 100


@pytest.mark.parametrize("seed,n,cutoff,static_every", [(1, 700, 7.1, 0), (2, 900, 12.0, 3), (3, 2, 5.0, 0),
                                                        (4, 400, 300.0, 5)])
def test_oracle_matches_brute_force(seed, n, cutoff, static_every):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-30, 30, (n, 3)).astype(np.float32)
    flags = np.zeros(n, np.uint8)
    if static_every:
        flags[::static_every] = _abi.FLAG_STATIC
    same_springs(po.springs_from_cutoff(x, flags, cutoff), brute_force(x, flags, cutoff))


def test_oracle_edge_cases():
    x = np.zeros((1, 3), np.float32)
    assert len(po.springs_from_cutoff(x, None, 5.0)[0]) == 0
    assert len(po.springs_from_cutoff(np.zeros((0, 3), np.float32), None, 5.0)[0]) == 0
    # all particles at one point: every pair, rest length 0
    i1, i2, l0 = po.springs_from_cutoff(np.ones((5, 3), np.float32), None, 1.0)
    assert len(i1) == 10 and np.all(l0 == 0)
    # the reference rejects cutoffs below 1e-6 (nsearch.hpp:66-67)
    with pytest.raises(ValueError):
        po.springs_from_cutoff(np.zeros((3, 3), np.float32), None, 1e-7)


@pytest.mark.skipif(not po.ref_available(), reason="compiled reference not built (make -C oracle ref)")
@pytest.mark.parametrize("seed,n,cutoff,span", [(5, 3000, 8.0, 60.0), (6, 2500, 9.3, 40.0), (7, 1500, 4.4, 20.0)])
def test_oracle_matches_compiled_reference(seed, n, cutoff, span):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-span, span, (n, 3)).astype(np.float32)
    flags = np.where(rng.random(n) < 0.15, _abi.FLAG_STATIC, 0).astype(np.uint8)
    same_springs(po.springs_from_cutoff(x, flags, cutoff), po.ref_springs_from_cutoff(x, flags, cutoff))


@pytest.fixture(scope="module")
def hostcheck_lib():
    subprocess.run(["make", "-C", CSRC, "hostcheck"], check=True, stdout=subprocess.DEVNULL)
    assert os.path.exists(HOSTCHECK)
    return HOSTCHECK


def test_host_construction_loads_the_fixture_springs(hostcheck_lib, tmp_path):
    """bs_springs_from_cutoff through the fake runtime: the springs it
    loads equal the reference's, with the given stiffness, and the handle
    then steps with them."""
    out = tmp_path / "springs.npz"
    code = textwrap.dedent(f"""
        import ctypes, sys
        sys.path.insert(0, {ROOT!r})
        import numpy as np
        from biospring_amd import _abi, workloads
        from biospring_amd.network import SpringNetwork
        L = _abi.bind(ctypes.CDLL({hostcheck_lib!r}))
        L.hc_live_allocations.restype = ctypes.c_int64
        base = L.hc_live_allocations()
        res = {{}}
        for name in {CASES!r}:
            z = np.load({GOLDEN!r} + "/" + name + ".npz")
            x, fl = z["in_x"], z["in_flags"]
            n = len(fl)
            net = SpringNetwork(workloads.params(1), library=L)
            try:
                net.springsFromCutoff(5.0)
                raise SystemExit("construction before particles accepted")
            except _abi.EngineError:
                pass
            net.setParticles(x, flags=fl)
            for bad in (0.0, 5e-7, float("nan"), float("inf")):
                try:
                    net.springsFromCutoff(bad)
                    raise SystemExit(f"cutoff {{bad}} accepted")
                except _abi.EngineError:
                    pass
            m = net.springsFromCutoff(float(z["cutoff"]), 2.5)
            a, b, l0, k = net.getSprings()
            assert m == len(a) == net.getNumberOfSprings()
            assert np.all(k == np.float32(2.5))
            res[name + "_a"], res[name + "_b"], res[name + "_l0"] = a, b, l0
            net.run(2)                      # the loaded springs drive the step
            assert np.isfinite(net.getState()[0]).all()
            net.springsFromCutoff(3.0)      # rebuilt over the moved positions
            net.run(1)
            net.close()
        z = np.load({GOLDEN!r} + "/topo_merge800.npz")
        net = SpringNetwork(workloads.params(1), library=L)
        net.setParticles(z["in_x"], flags=z["in_flags"])
        net.setSprings(z["pre_id1"], z["pre_id2"], z["pre_l0"], z["pre_k"])
        net.run(1)                          # device state live: construction pulls it first
        net.setState(z["in_x"], None)
        added = net.springsFromCutoff(float(z["cutoff"]), float(z["stiffness"]))
        got = net.getSprings()
        assert added == len(got[0]) - len(z["pre_id1"])
        for key, g in zip(("a", "b", "l0", "k"), got):
            res["merge_" + key] = g
        net.setSprings([], [], [], [])      # m = 0 clears the table
        assert net.getNumberOfSprings() == 0
        assert net.springsFromCutoff(float(z["cutoff"])) == len(got[0]) - 25
        net.run(1)
        net.close()
        assert L.hc_live_allocations() == base
        np.savez({str(out)!r}, **res)
        """)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    got = np.load(out)
    check_merged(tuple(got["merge_" + k] for k in ("a", "b", "l0", "k")), load_case("topo_merge800"))
    for name in CASES:
        c = load_case(name)
        same_springs((got[name + "_a"], got[name + "_b"], got[name + "_l0"]),
                     (c["ref_id1"], c["ref_id2"], c["ref_l0"]))


def test_host_construction_keeps_systems_apart(hostcheck_lib, tmp_path):
    """With independent systems declared, construction links particles of
    one system only, even where two systems overlap in space; each system's
    springs are the oracle's on that system alone."""
    rng = np.random.default_rng(9)
    xa = rng.uniform(-15, 15, (700, 3)).astype(np.float32)
    xb = (xa + np.float32(0.5)).astype(np.float32)  # overlaps system A
    x = np.concatenate([xa, xb])
    np.save(tmp_path / "x.npy", x)
    code = textwrap.dedent(f"""
        import ctypes, sys
        sys.path.insert(0, {ROOT!r})
        import numpy as np
        from biospring_amd import _abi, workloads
        from biospring_amd.network import SpringNetwork
        L = _abi.bind(ctypes.CDLL({hostcheck_lib!r}))
        x = np.load({str(tmp_path / "x.npy")!r})
        net = SpringNetwork(workloads.params(99), library=L)
        net.setParticles(x)
        net.setSystems(np.array([0, 700, 1400], np.int64))
        net.springsFromCutoff(6.0)
        a, b, l0, k = net.getSprings()
        np.savez({str(tmp_path / "s.npz")!r}, a=a, b=b, l0=l0)
        net.run(1)
        net.close()
        """)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    z = np.load(tmp_path / "s.npz")
    a, b, l0 = z["a"], z["b"], z["l0"]
    assert np.all((a < 700) == (b < 700))
    for lo, xs in ((0, xa), (700, xb)):
        sel = a >= lo if lo else a < 700
        same_springs((a[sel] - lo, b[sel] - lo, l0[sel]), po.springs_from_cutoff(xs, None, 6.0))


def test_groups_fixture_oracle_and_brute_force():
    """Merging topologies (Topology::add_springs_between_topologies_from_cutoff):
    the oracle and brute force give the reference's springs, all between the
    two topologies."""
    c = load_case("topo_groups900")
    g = c["in_groups"]
    ref = (c["ref_id1"], c["ref_id2"], c["ref_l0"])
    assert np.all(g[ref[0]] != g[ref[1]])
    same_springs(po.springs_from_cutoff(c["in_x"], c["in_flags"], float(c["cutoff"]), groups=g), ref)
    same_springs(brute_force(c["in_x"], c["in_flags"], float(c["cutoff"]), groups=g), ref)


def test_host_construction_between_groups(hostcheck_lib, tmp_path):
    c = load_case("topo_groups900")
    np.savez(tmp_path / "c.npz", **c)
    code = textwrap.dedent(f"""
        import ctypes, sys
        sys.path.insert(0, {ROOT!r})
        import numpy as np
        from biospring_amd import _abi, workloads
        from biospring_amd.network import SpringNetwork
        L = _abi.bind(ctypes.CDLL({hostcheck_lib!r}))
        z = np.load({str(tmp_path / "c.npz")!r})
        net = SpringNetwork(workloads.params(99), library=L)
        net.setParticles(z["in_x"], flags=z["in_flags"])
        net.springsBetweenGroupsFromCutoff(float(z["cutoff"]), z["in_groups"], 0.5)
        a, b, l0, k = net.getSprings()
        assert np.all(k == np.float32(0.5))
        np.savez({str(tmp_path / "s.npz")!r}, a=a, b=b, l0=l0)
        net.close()
        """)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    z = np.load(tmp_path / "s.npz")
    same_springs((z["a"], z["b"], z["l0"]), (c["ref_id1"], c["ref_id2"], c["ref_l0"]))
