"""A network loaded from a NetCDF topology (bs_load_netcdf) steps on the GPU
like the same network given array by array: a 3,000-bead workload-1 system
written with scipy.io.netcdf_file, loaded, and its forces compared with the
FP64 oracle of the arrays (1e-5), then stepped in lockstep with the
array-loaded engine (identical positions)."""
import sys

import numpy as np
import pytest

scipy_io = pytest.importorskip("scipy.io")

from biospring_amd import _abi, workloads  # noqa: E402
from biospring_amd.network import SpringNetwork  # noqa: E402
from oracle import pyoracle as po  # noqa: E402
from tests.test_netcdf import chars, write  # noqa: E402

pytestmark = pytest.mark.gpu


def test_loaded_network_steps_like_the_arrays(tmp_path):
    s = workloads.generate(1, seed=41, n=3000)
    n, m = s.n, s.m
    dims = {"particle_number": n, "spatialdim": 3, "resname_length": 4, "particlename_length": 5,
            "chainname_length": 4, "springdim": 2, "spring_number": m}
    P, S = ("particle_number",), ("spring_number",)
    v = {
        "particleids": ("i", P, np.arange(n, dtype=np.int32)),
        "resids": ("i", P, np.arange(n, dtype=np.int32)),
        "resnames": ("c", P + ("resname_length",), chars(["BEA"] * n, 4)),
        "particlenames": ("c", P + ("particlename_length",), chars(["CA"] * n, 5)),
        "chainnames": ("c", P + ("chainname_length",), chars(["A"] * n, 4)),
        "dynamicstate": ("b", P, ((s.flags & _abi.FLAG_STATIC) == 0).astype(np.int8)),
        "coordinates": ("f", P + ("spatialdim",), s.x.reshape(-1, 3)),
        "radii": ("f", P, s.radius), "mass": ("f", P, s.mass), "epsilon": ("f", P, s.eps),
        "charges": ("f", P, s.charge), "surfaceaccessibility": ("f", P, s.sasa),
        "hydrophobicityscale": ("f", P, s.transfer), "hydrophobicity": ("f", P, s.hydro),
        "springs": ("i", S + ("springdim",), np.stack([s.id1, s.id2], 1).astype(np.int32)),
        "springsstiffness": ("f", S, s.k), "springsequilibrium": ("f", S, s.l0),
    }
    path = str(tmp_path / "w1.nc")
    write(path, dims, v, version=2)
    print(f"netcdf load: n={n} springs={m}", flush=True)
    sys.stdout.flush()
    p = s.params.copy()
    p.viscosity_enable = 0
    a = SpringNetwork(p)
    a.loadNetCDF(path)
    assert a.getNumberOfParticles() == n and a.getNumberOfSprings() == m
    F, E = a.computeForces()
    o = po.Oracle(s, params=p)
    Fd, Ed = o.compute_forces(po.F64)
    o.close()
    err = np.linalg.norm(F - Fd, axis=1).max() / np.linalg.norm(Fd, axis=1).max()
    assert err <= 1e-5, err
    b = SpringNetwork.from_system(s, params=p)
    b.setState(None, np.zeros((n, 3), np.float32))  # a .nc topology carries no velocities
    a.run(10)
    b.run(10)
    assert np.array_equal(a.getState()[0], b.getState()[0])
    a.close()
    b.close()


def test_saved_network_reloads_and_steps_alike(tmp_path):
    """A network built on the GPU (springs from a cutoff), stepped, saved
    with bs_save_netcdf and loaded into a second handle continues exactly
    as the first (velocities aside: a topology file holds none)."""
    s = workloads.generate(1, seed=42, n=4000)
    p = s.params.copy()
    a = SpringNetwork(p)
    a.setParticles(*s.particle_arrays())
    a.springsFromCutoff(8.0, 0.5)
    a.run(5)
    path = str(tmp_path / "built.nc")
    a.saveNetCDF(path)
    b = SpringNetwork(p)
    b.loadNetCDF(path)
    xa, _ = a.getState()
    assert np.array_equal(b.getState()[0], xa)
    for u, v in zip(a.getSprings(), b.getSprings()):
        assert np.array_equal(u, v)
    a.setState(None, np.zeros_like(xa))
    a.run(5)
    b.run(5)
    assert np.array_equal(a.getState()[0], b.getState()[0])
    a.close()
    b.close()
