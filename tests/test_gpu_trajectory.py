"""Trajectory and energy output on the GPU (bs_add_trajectory): frames are
packed in HBM, copied to pinned slots at step boundaries and written by a
host thread while the stream goes on stepping.

- Against the reference's own writers (golden traj300: its PDB and CSV files
  over 7 computeStep calls): every model and row at the same iterations, the
  non-coordinate PDB text identical, coordinates within the 0.001 A of the
  %8.3f field (the float trajectories differ in their last bits), energies
  within 1e-5 relative plus the 6 printed digits.
- Against the engine's own state: frames taken during graph replays and
  across calls hold exactly the positions getState returns at those
  iterations (PDB text, and XTC decoded by the reference's xdrfile when it
  is built), and the CSV rows hold the energies bs_step returns."""
import os
import sys

import numpy as np
import pytest

from biospring_amd import _abi, workloads
from biospring_amd.network import SpringNetwork
from oracle import pyoracle as po
from tests.test_oracle import params_from
from tests.test_trajectory import atom_xyz, csv_table, golden, pdb_models

pytestmark = pytest.mark.gpu


def _log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _system(z):
    s = workloads.System(**{k[3:]: z[k] for k in z if k.startswith("in_")})
    s.params = params_from(z["params"])
    return s


def test_matches_the_reference_writers(tmp_path):
    z = golden("traj300")
    s = _system(z)
    _log("trajectory vs reference writers: n=300, 7 steps")
    net = SpringNetwork.from_system(s)
    net.addTrajectory(_abi.TRAJ_PDB, str(tmp_path / "t.pdb"), int(z["pdb_frequency"]))
    net.addTrajectory(_abi.TRAJ_CSV, str(tmp_path / "t.csv"), int(z["csv_frequency"]))
    net.run(int(z["steps"]))
    net.closeTrajectories()
    net.close()
    ref = pdb_models(z["pdb"].tobytes().decode())
    got = pdb_models((tmp_path / "t.pdb").read_text())
    assert [m[0] for m in got] == [m[0] for m in ref]
    assert got[0] == ref[0]  # before the first step: identical text
    for g, r in zip(got[1:], ref[1:]):
        assert [ln[:30] + ln[54:] for ln in g[1]] == [ln[:30] + ln[54:] for ln in r[1]]
        dev = np.abs(atom_xyz(g[1]) - atom_xyz(r[1])).max()
        assert dev <= 1.0005e-3, dev
    rh, rr = csv_table(z["csv"].tobytes().decode())
    gh, gr = csv_table((tmp_path / "t.csv").read_text())
    assert gh == rh
    assert [r[0] for r in gr] == [r[0] for r in rr]
    for g, r in zip(gr, rr):
        a, b = np.array(g[2:], float), np.array(r[2:], float)  # the frame rate is a wall-clock figure
        assert np.all(np.abs(a - b) <= 1e-5 * np.abs(b) + 1e-5 * np.abs(b).max() + 1e-6), (g, r)


@pytest.mark.parametrize("xtc_every,pdb_every", [(5, 10), (20, 40)])
def test_frames_hold_the_engine_state(tmp_path, xtc_every, pdb_every):
    """Frames every few iterations over calls of 23 and 37 steps, with HIP
    graphs of 16 steps enabled (frames 20 and 40 apart leave room for graph
    replays between them); a second engine stepped to each frame's iteration
    gives the expected positions."""
    s = workloads.generate(1, seed=31, n=5000)
    p = s.params.copy()
    p.graph_steps = 16
    net = SpringNetwork.from_system(s, params=p)
    net.addTrajectory(_abi.TRAJ_XTC, str(tmp_path / "f.xtc"), xtc_every)
    net.addTrajectory(_abi.TRAJ_PDB, str(tmp_path / "f.pdb"), pdb_every)
    _log(f"frames vs state: n=5000, 60 steps, graphs of 16, XTC every {xtc_every}, PDB every {pdb_every}")
    net.run(23)
    net.run(37)
    net.closeTrajectories()
    net.close()
    chk = SpringNetwork.from_system(s, params=p)
    want = {}
    done = 0
    for it in range(0, 60, 5):
        if it > done:
            chk.run(it - done)
            done = it
        want[it] = chk.getState()[0].copy()
    chk.close()
    models = pdb_models((tmp_path / "f.pdb").read_text())
    its = list(range(0, 60, pdb_every))
    assert [m[0] for m in models] == list(range(len(its)))
    for it, m in zip(its, models):
        exp = np.array([[float("%8.3f" % c) for c in row] for row in want[it]])
        assert np.array_equal(atom_xyz(m[1]), exp), it
    if po.ref_available():
        its = list(range(0, 60, xtc_every))
        xs, steps, times, boxes, precs = po.ref_xtc_read(str(tmp_path / "f.xtc"), s.n)
        assert list(steps) == list(range(len(its))) and np.all(precs == 1000.0)
        for k, it in enumerate(its):
            nm = (want[it].astype(np.float64) / 10.0).astype(np.float32)
            assert np.abs(xs[k] - nm).max() <= 5.01e-4, it


def test_csv_rows_hold_the_step_energies(tmp_path):
    s = workloads.generate(1, seed=32, n=3000)
    net = SpringNetwork.from_system(s)
    net.addTrajectory(_abi.TRAJ_CSV, str(tmp_path / "e.csv"), 1)
    energies = []
    for _ in range(6):
        net.run(1)
        energies.append(net.getEnergies().copy())
    net.closeTrajectories()
    net.close()
    _, rows = csv_table((tmp_path / "e.csv").read_text())
    assert [r[0] for r in rows] == [str(k) for k in range(6)]
    cols = {"kinetic": 2, "spring": 3, "steric": 4, "electrostatic": 5}
    idx = {"kinetic": _abi.E_KINETIC, "spring": _abi.E_SPRING, "steric": _abi.E_STERIC,
           "electrostatic": _abi.E_ELECTROSTATIC}
    for k in range(1, 6):  # row k: the energies of step k - 1
        for name, c in cols.items():
            e = float(np.float32(energies[k - 1][idx[name]]))
            assert float(rows[k][c]) == pytest.approx(e, rel=6e-6, abs=1e-30), (k, name)
