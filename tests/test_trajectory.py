"""Trajectory and energy output (bs_add_trajectory), CPU side.

The writers stand in for io::modern::TrajectoryManager's PDB, XTC and CSV
writers (src/IO/modern/*TrajectoryWriter.cpp) called from
SpringNetwork::idleRun.  Pinned by the reference's own output, kept as golden
fixtures (tests/golden/make_golden.py):
  - xtc_frames: XTC files from the reference's vendored xdrfile write_xtc,
    compared byte for byte (the compression is integer work: bit-exact);
  - traj300: the reference's PDB and CSV files over 7 steps of a 300-bead
    network; here the first model and the CSV header and first row, which
    precede any step, are compared as text (the later frames follow the GPU
    trajectory and are compared in test_gpu_trajectory.py).
The engine runs against the recording fake of the HIP runtime, so the host
logic (frame schedule, pinned slots, the writer thread, file handling) is the
product code.  The reference's XTC writer itself is not driven: it never opens
its XDR handle (XTCTrajectoryWriter.hpp:19-20), so xdrfile's write_xtc with
the arguments that writer passes is the yardstick."""
import ctypes
import os
import subprocess
import sys
import textwrap

import numpy as np
import pytest

from biospring_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
CSRC = os.path.join(ROOT, "biospring_amd", "csrc")
HOSTCHECK = os.path.join(CSRC, "_build", "libbs_hostcheck.so")
XTC_SETS = ("random", "pairs", "line", "wide", "nine", "ten")


def golden(name):
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    return {k: z[k] for k in z.files}


def pdb_models(text):
    """[(model number, [ATOM lines], [CONECT lines])] of a PDB trajectory."""
    models, cur = [], None
    for line in text.splitlines():
        if line.startswith("MODEL"):
            cur = (int(line[5:]), [], [])
        elif line.startswith("ATOM"):
            cur[1].append(line)
        elif line.startswith("CONECT"):
            cur[2].append(line)
        elif line.startswith("ENDMDL"):
            models.append(cur)
            cur = None
    return models


def atom_xyz(lines):
    return np.array([[float(ln[30:38]), float(ln[38:46]), float(ln[46:54])] for ln in lines])


def csv_table(text):
    lines = text.splitlines()
    return lines[0], [ln.split("\t") for ln in lines[1:] if ln]


def run_hostcheck(body, tmp_path):
    """Runs `body` in a child process with the engine built against the fake
    HIP runtime (L, SpringNetwork, np, _abi, workloads and TMP in scope)."""
    subprocess.run(["make", "-C", CSRC, "hostcheck"], check=True, stdout=subprocess.DEVNULL)
    code = textwrap.dedent(f"""
        import ctypes, os, sys
        sys.path.insert(0, {ROOT!r})
        import numpy as np
        from biospring_amd import _abi, workloads
        from biospring_amd.network import SpringNetwork
        L = _abi.bind(ctypes.CDLL({HOSTCHECK!r}))
        L.hc_live_allocations.restype = ctypes.c_int64
        TMP = {str(tmp_path)!r}
        """) + textwrap.dedent(body)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    return r.stdout


@pytest.mark.parametrize("name", XTC_SETS)
def test_xtc_bytes_match_reference_xdrfile(name, tmp_path):
    z = golden("xtc_frames")
    frames = z[f"{name}__x"]
    np.save(tmp_path / "frames.npy", frames)
    run_hostcheck("""
        frames = np.load(os.path.join(TMP, "frames.npy"))
        net = SpringNetwork(workloads.params(99), library=L)
        net.setParticles(frames[0])
        net.addTrajectory(_abi.TRAJ_XTC, os.path.join(TMP, "out.xtc"), 1)
        for k in range(3):
            net.setState(frames[k], None)
            net.writeNextStepNow()
        net.closeTrajectories()
        net.close()
        """, tmp_path)
    got = np.frombuffer((tmp_path / "out.xtc").read_bytes(), np.uint8)
    want = z[f"{name}__xtc"]
    assert len(got) == len(want)
    assert np.array_equal(got, want)


def test_pdb_and_csv_start_like_the_reference(tmp_path):
    """Model 0 (the positions before the first step, then the CONECT records
    of every spring) and the CSV header and row 0 as the reference wrote
    them; the same number of models and rows at the same iterations."""
    z = golden("traj300")
    np.savez(tmp_path / "sys.npz", **{k: v for k, v in z.items() if k.startswith("in_") or k == "params"})
    run_hostcheck("""
        z = np.load(os.path.join(TMP, "sys.npz"))
        p = _abi.Params()
        for (f, _), v in zip(_abi.Params._fields_, z["params"]):
            setattr(p, f, type(getattr(p, f))(v))
        s = workloads.System(**{k[3:]: z[k] for k in z.files if k.startswith("in_")})
        s.params = p
        net = SpringNetwork.from_system(s, library=L)
        net.addTrajectory(_abi.TRAJ_PDB, os.path.join(TMP, "t.pdb"), 2)
        net.addTrajectory(_abi.TRAJ_CSV, os.path.join(TMP, "t.csv"), 3)
        net.run(7)
        net.closeTrajectories()
        net.close()
        """, tmp_path)
    ref_pdb = pdb_models(z["pdb"].tobytes().decode())
    got_pdb = pdb_models((tmp_path / "t.pdb").read_text())
    assert [m[0] for m in got_pdb] == [m[0] for m in ref_pdb] == [0, 1, 2, 3]
    assert got_pdb[0] == ref_pdb[0]
    assert len(got_pdb[0][2]) == int(z["in_id1"].size) and all(not m[2] for m in got_pdb[1:])
    ref_head, ref_rows = csv_table(z["csv"].tobytes().decode())
    got_head, got_rows = csv_table((tmp_path / "t.csv").read_text())
    assert got_head == ref_head
    assert [r[0] for r in got_rows] == [r[0] for r in ref_rows] == ["0", "3", "6"]
    assert got_rows[0][2:] == ref_rows[0][2:]


def test_reference_frames_are_the_oracle_trajectory():
    """The fixture's frames are the float-faithful oracle's trajectory: every
    PDB coordinate as printed, and the CSV energies of the step before each
    row (E order: spring, electrostatic, steric, kinetic)."""
    from oracle import pyoracle as po
    from tests.test_oracle import params_from
    from biospring_amd import workloads
    z = golden("traj300")
    s = workloads.System(**{k[3:]: z[k] for k in z if k.startswith("in_")})
    s.params = params_from(z["params"])
    models = pdb_models(z["pdb"].tobytes().decode())
    _, rows = csv_table(z["csv"].tobytes().decode())
    o = po.Oracle(s, params=s.params)
    energies = {}
    for it in range(1, 8):
        _, E = o.step(po.F32, 1)
        energies[it] = E
        x, _ = o.state(po.F32)
        if it % 2 == 0 and it // 2 < len(models):
            want = np.array([[float("%8.3f" % np.float32(c)) for c in row] for row in x])
            assert np.array_equal(atom_xyz(models[it // 2][1]), want), it
    o.close()
    for r in rows[1:]:
        E = energies[int(r[0])]
        for col, e in zip(r[2:6], (E[3], E[0], E[2], E[1])):
            assert float(col) == pytest.approx(float(np.float32(e)), rel=6e-6), r


def test_frames_follow_the_iteration_count(tmp_path):
    """Frames at the iterations a writer's frequency divides, counted across
    calls; writeNextStepNow writes to every writer at once."""
    out = run_hostcheck("""
        s = workloads.generate(1, seed=3, n=400)
        net = SpringNetwork.from_system(s, library=L)
        net.setup()
        base = L.hc_live_allocations()
        net.addTrajectory(_abi.TRAJ_PDB, os.path.join(TMP, "a.pdb"), 3)
        net.addTrajectory(_abi.TRAJ_CSV, os.path.join(TMP, "a.csv"), 2)
        net.addTrajectory(_abi.TRAJ_XTC, os.path.join(TMP, "a.xtc"), 4)
        net.run(2); net.run(5); net.run(1)       # iterations 0..7
        net.writeNextStepNow()                   # iteration 8, every writer
        net.flushTrajectories()
        net.run(2)                               # 8, 9: PDB 9, CSV 8, XTC 8
        net.closeTrajectories()
        assert L.hc_live_allocations() == base, (L.hc_live_allocations(), base)  # slots and staging freed
        net.close()
        """, tmp_path)
    pdb = pdb_models((tmp_path / "a.pdb").read_text())
    assert [m[0] for m in pdb] == [0, 1, 2, 3, 4]          # iterations 0, 3, 6, the forced 8, then 9
    _, rows = csv_table((tmp_path / "a.csv").read_text())
    assert [r[0] for r in rows] == ["0", "2", "4", "6", "8", "8"]
    # XTC frames: iterations 0, 4, forced 8, 8 -> frame counters 0..3
    data = (tmp_path / "a.xtc").read_bytes()
    assert data.count((1995).to_bytes(4, "big")) >= 4


def test_argument_errors_and_cleanup(tmp_path):
    run_hostcheck("""
        s = workloads.generate(1, seed=4, n=200)
        base = L.hc_live_allocations()
        net = SpringNetwork(s.params, library=L)
        def bad(f, *a):
            try:
                f(*a)
            except _abi.EngineError:
                return
            raise SystemExit(f"accepted: {f.__name__}{a}")
        bad(net.addTrajectory, _abi.TRAJ_PDB, os.path.join(TMP, "x.pdb"), 1)   # no particles yet
        net.setParticles(*s.particle_arrays())
        net.setSprings(*s.spring_arrays())
        bad(net.addTrajectory, 7, os.path.join(TMP, "x.pdb"), 1)
        bad(net.addTrajectory, _abi.TRAJ_PDB, os.path.join(TMP, "x.pdb"), 0)
        bad(net.addTrajectory, _abi.TRAJ_PDB, "", 1)
        bad(net.addTrajectory, _abi.TRAJ_PDB, os.path.join(TMP, "no", "such", "dir.pdb"), 1)
        bad(net.writeNextStepNow)
        bad(net.setInsertionVector, 0, 200)
        net.addTrajectory(_abi.TRAJ_CSV, os.path.join(TMP, "c.csv"), 1)
        net.run(3)
        net.setParticles(*s.particle_arrays())    # a new particle set closes the writers
        bad(net.writeNextStepNow)
        net.close()
        assert L.hc_live_allocations() == base, (L.hc_live_allocations(), base)
        """, tmp_path)
    _, rows = csv_table((tmp_path / "c.csv").read_text())
    assert [r[0] for r in rows] == ["0", "1", "2"]


def test_labels_and_insertion_columns(tmp_path):
    """ATOM fields from bs_set_particle_labels laid out as pdb_format.hpp
    does (names under 4 characters from column 14, charges signed), and the
    CSV insertion columns (angle of x[i2] - x[i1] to the membrane normal,
    depth = the centroid's z)."""
    run_hostcheck("""
        x = np.array([[0, 0, 0], [1.5, 2.25, -3.125], [10, -20, 30], [-4, 5, 6.5]], np.float32)
        net = SpringNetwork(workloads.params(99), library=L)
        net.setParticles(x, charge=np.array([0, 1.7, -2.2, 0], np.float32))
        net.setParticleLabels(name=["CA", "OXT1", "N", "LONGNAME"], resname=["ALA", "GLY", "LYS", "H"],
                              chain=["A", "B", "", "Z"], resid=[1, 22, 333, 4444], occupancy=[1, 0.5, 0, 1],
                              tempfactor=[10.25, 0, 99.99, 1.5], element=["C", "O", "N", ""])
        net.setInsertionVector(0, 2)
        net.addTrajectory(_abi.TRAJ_PDB, os.path.join(TMP, "l.pdb"), 1)
        net.addTrajectory(_abi.TRAJ_CSV, os.path.join(TMP, "l.csv"), 1)
        net.writeNextStepNow()
        net.closeTrajectories()
        net.close()
        """, tmp_path)
    lines = [ln for ln in (tmp_path / "l.pdb").read_text().splitlines() if ln.startswith("ATOM")]
    assert lines[0] == "ATOM      1  CA  ALA A   1       0.000   0.000   0.000  1.00 10.25           C  "
    assert lines[1] == "ATOM      2 OXT1 GLY B  22       1.500   2.250  -3.125  0.50  0.00           O+1"
    assert lines[2] == "ATOM      3  N   LYS   333      10.000 -20.000  30.000  0.00 99.99           N-2"
    assert lines[3] == "ATOM      4 LONG   H Z4444      -4.000   5.000   6.500  1.00  1.50              "
    head, rows = csv_table((tmp_path / "l.csv").read_text())
    assert head.endswith("\tInsertion Angle (degrees) \tInsertion Depth (A)")
    v = np.array([10.0, -20.0, 30.0])
    angle = np.degrees(np.arccos(-v[2] / np.linalg.norm(v))) - 90.0
    assert float(rows[0][-2]) == pytest.approx(angle, rel=1e-5)
    assert float(rows[0][-1]) == pytest.approx(np.mean([0, -3.125, 30, 6.5]), rel=1e-5)
