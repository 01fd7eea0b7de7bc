"""The CPU-side code under AddressSanitizer and UndefinedBehaviorSanitizer.

Two instrumented builds (host code only):
* the engine's host logic over the CPU stand-in of the HIP runtime
  (`make -C biospring_amd/csrc hostcheck-san`): setup, steps, re-setups,
  parity hooks, rigid bodies, the probe, NetCDF reading (sound and damaged
  files), trajectory writers and topology construction;
* the oracle (`make -C oracle san`): forces, steps, pair sets, cells, rigid
  bodies, grids, the probe, the adjoint and springs from a cutoff.
Each runs in a child Python with libasan preloaded; any report aborts the
child and fails the test.
"""
import os
import subprocess
import sys
import textwrap

import pytest
from conftest import make_locked

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "biospring_amd", "csrc")
SAN_ENGINE = os.path.join(CSRC, "_build", "libbs_hostcheck_san.so")
SAN_ORACLE = os.path.join(ROOT, "oracle", "_san", "liboracle.so")


def _libasan():
    r = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = r.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.fixture(scope="module")
def asan():
    lib = _libasan()
    if lib is None:
        pytest.skip("libasan is not installed")
    make_locked("-s", "-C", CSRC, "hostcheck-san")
    make_locked("-s", "-C", os.path.join(ROOT, "oracle"), "san")
    return lib


def _run(body, lib_asan):
    code = textwrap.dedent(f"""
        import ctypes, os, sys, tempfile
        sys.path.insert(0, {ROOT!r})
        import numpy as np
        from biospring_amd import _abi, workloads
        from biospring_amd.network import SpringNetwork
        from oracle import pyoracle as po
        po.ORACLE_PATH = {SAN_ORACLE!r}
        L = _abi.bind(ctypes.CDLL({SAN_ENGINE!r}))
        """) + textwrap.dedent(body) + "\nprint('SAN-OK', flush=True)\n"
    env = dict(os.environ, LD_PRELOAD=lib_asan, OMP_NUM_THREADS="1",
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=900, env=env)
    assert r.returncode == 0 and "SAN-OK" in r.stdout, r.stdout[-3000:] + r.stderr[-6000:]


def test_engine_host_logic_under_sanitizers(asan, tmp_path):
    _run(f"""
        s = workloads.generate(1, seed=3, n=1500)
        net = SpringNetwork.from_system(s, library=L)
        net.setup(); net.run(12)
        p = s.params.copy(); p.viscosity_enable = 0
        net.setParams(p); net.computeForces()
        for kind in (0, 1): net.getCells(kind); net.getPairs(kind)
        x, v = net.getState(); net.setState(x, v); net.run(3)
        net.addRigidBody(np.arange(40)); net.run(2); net.getRigidBody(0)
        q = s.params.copy(); q.probe_enable = 1; q.probe_enable_steric = 1; q.probe_enable_electrostatic = 1
        b = SpringNetwork.from_system(s, params=q, library=L); b.run(3); b.getProbeParticle(); b.close()
        d = {str(tmp_path)!r}
        net.addTrajectory(_abi.TRAJ_PDB, os.path.join(d, 't.pdb'), 2)
        net.addTrajectory(_abi.TRAJ_CSV, os.path.join(d, 't.csv'), 2)
        net.addTrajectory(_abi.TRAJ_XTC, os.path.join(d, 't.xtc'), 2)
        net.run(5); net.closeTrajectories()
        net.saveNetCDF(os.path.join(d, 't.nc'))
        m = SpringNetwork(workloads.params(99), library=L)
        m.loadNetCDF(os.path.join(d, 't.nc'))
        raw = bytearray(open(os.path.join(d, 't.nc'), 'rb').read())
        for cut in (len(raw) // 3, len(raw) // 2, 40):
            open(os.path.join(d, 'cut.nc'), 'wb').write(bytes(raw[:cut]))
            try:
                m.loadNetCDF(os.path.join(d, 'cut.nc'))
            except _abi.EngineError:
                pass
        c = SpringNetwork(s.params, library=L)
        c.setParticles(*s.particle_arrays()); c.springsFromCutoff(8.0); c.getSprings()
        for h in (net, m, c): h.close()
        """, asan)


def test_oracle_under_sanitizers(asan):
    _run("""
        s = workloads.generate(1, seed=5, n=600)
        o = po.Oracle(s)
        for mode in (po.F32, po.F64):
            o.compute_forces(mode); o.step(mode, 2); o.pairs(0, mode=mode); o.cells(1, mode=mode)
        o.add_rigid_body(np.arange(50)); o.step(po.F32, 2); o.rigid_state(po.F32, 0)
        o.close()
        q = s.params.copy(); q.probe_enable = 1; q.probe_enable_steric = 1; q.probe_enable_electrostatic = 1
        o = po.Oracle(s, params=q); o.step(po.F32, 2); o.step(po.F64, 2); o.probe(po.F64); o.close()
        a = workloads.generate(4, seed=2, n=300)
        y = workloads.fitting_target(a, seed=2).astype(np.float64)
        o = po.Oracle(a); o.adjoint(4, y); o.close()
        po.springs_from_cutoff(s.x, s.flags, 8.0)
        """, asan)
