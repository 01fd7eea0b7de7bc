"""The reference-side adapter (biospring_amd/adapter/SpringNetworkHIP) is
compiled against the reference's own headers and linked with the
reference's objects (oracle/_ref/obj, built from its sources by
`make -C oracle ref`) and the engine's C ABI in its host-check build (a CPU
stand-in of the HIP runtime whose kernels only check their extents).  The
driver tests/adapter_check.cpp sets a SpringNetworkHIP up from a
configuration, runs initRun, a setForce and three computeStep calls, reads
positions (the probe's too), forces and energies back, and ends the run.

This checks the boundary, not the physics: the parameters the adapter
derives from the configuration and force field, the network it uploads, and
that the overridden step methods drive the engine.  Skipped where the
reference sources are absent (they are not on the GPU box).
"""
import glob
import os
import subprocess

import pytest
from conftest import make_locked

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src"
CSRC = os.path.join(ROOT, "biospring_amd", "csrc")
HOSTCHECK = os.path.join(CSRC, "_build", "libbs_hostcheck.so")
REF_OBJ = os.path.join(ROOT, "oracle", "_ref", "obj")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if not os.path.isdir(REF) or not glob.glob(os.path.join(REF_OBJ, "*.o")):
        pytest.skip("the reference sources and their objects (make -C oracle ref) are not here")
    make_locked("-C", CSRC, "hostcheck")
    exe = str(tmp_path_factory.mktemp("adapter") / "adapter_check")
    cmd = ["g++", "-std=c++20", "-O1", "-fopenmp", "-DOPENMP_SUPPORT", "-w",
           f"-I{REF}", f"-I{REF}/spn", f"-I{REF}/IO/xdrfile/src", f"-I{ROOT}/include",
           f"-I{ROOT}/biospring_amd/adapter",
           os.path.join(ROOT, "tests", "adapter_check.cpp"),
           os.path.join(ROOT, "biospring_amd", "adapter", "SpringNetworkHIP.cpp")]
    cmd += sorted(glob.glob(os.path.join(REF_OBJ, "*.o")))
    cmd += [HOSTCHECK, f"-Wl,-rpath,{os.path.dirname(HOSTCHECK)}", "-o", exe]
    subprocess.run(cmd, check=True)
    return exe


def test_adapter_drives_the_engine(driver):
    r = subprocess.run([driver], capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, OMP_NUM_THREADS="2"))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    kv = dict(line.split("=", 1) for line in r.stdout.splitlines() if "=" in line)
    # parameters from Configuration and ForceField (SpringNetwork.cpp:690-710)
    assert int(kv["param_steric_mode"]) == 1  # lennard-jones-12-6Amber
    assert float(kv["param_steric_cutoff"]) == 9.0 and float(kv["param_elec_cutoff"]) == 11.0
    assert float(kv["param_dielectric"]) == 2.0 and float(kv["param_timestep"]) == 1.0
    assert kv["param_visc"] == "1,0.5"
    assert kv["param_probe"] == "1,1,0,30,0.5"
    # the network: 216 beads plus the probe the reference appends at setup
    assert int(kv["network_particles"]) == 217 and int(kv["engine_particles"]) == 216
    ns = int(kv["network_springs"])
    assert ns > 0 and int(kv["engine_springs"]) == ns and int(kv["springs_identical"]) == ns
    # the steps ran through idleRun (iteration count) and the engine
    assert int(kv["iterations"]) == 3
    assert int(kv["position_5_unchanged"]) == 1
    assert [float(t) for t in kv["probe_position"].split(",")] == [30.0, 1.0, 1.0]
    assert kv["done"] == "1"
