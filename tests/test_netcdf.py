"""Topology files (bs_load_netcdf), CPU side: the reference's NetCDFReader
(src/IO/NetCDFReader.cpp:8-280) followed by Topology::to_spring_network.

The checker is scipy.io.netcdf_file, an independent implementation of the
classic NetCDF format: the tests write topology files with it (both classic
versions, fixed and record dimensions, several value types) and compare what
the engine loads with what scipy reads back, after the reference reader's own
rules (a zero mass becomes 1, dynamicstate 0 means static, names cut at their
buffer length).  The reader runs inside the engine built against the fake HIP
runtime, so the product code is what is tested.  The reference itself cannot
be driven: the netCDF library is not installed here."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

scipy_io = pytest.importorskip("scipy.io")

from biospring_amd import _abi, workloads  # noqa: E402
from biospring_amd.network import SpringNetwork  # noqa: E402
from conftest import make_locked

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "biospring_amd", "csrc")
HOSTCHECK = os.path.join(CSRC, "_build", "libbs_hostcheck.so")
EXAMPLE = "/root/reference/scripts/outdated/articulatedrigidbody/example/1IYT.nc"


@pytest.fixture(scope="module")
def lib():
    make_locked("-C", CSRC, "hostcheck")
    return _abi.bind(ctypes.CDLL(HOSTCHECK))


def chars(names, width):
    a = np.zeros((len(names), width), "S1")
    for i, s in enumerate(names):
        b = s.encode()[:width]
        a[i, :len(b)] = np.frombuffer(b, "S1")
    return a


def topology(n=60, m=90, seed=0, record=False, types=None, omit=(), extra_vars=None):
    """A topology as the reference's writers lay it out (SpnBuffer), with
    random contents; types overrides a variable's NetCDF type code."""
    rng = np.random.default_rng(seed)
    types = types or {}
    pairs = set()
    while len(pairs) < m:
        a, b = rng.integers(0, n, 2)
        if a != b and (min(a, b), max(a, b)) not in pairs:
            pairs.add((int(min(a, b)), int(max(a, b))))
    ij = np.array(sorted(pairs), np.int32)
    flip = rng.random(m) < 0.5
    ij[flip] = ij[flip][:, ::-1]
    mass = rng.uniform(1, 20, n).astype(np.float32)
    mass[::7] = 0.0
    v = {
        "particleids": ("i", ("particle_number",), np.arange(n, dtype=np.int32) + 100),
        "resids": ("i", ("particle_number",), rng.integers(-5, 500, n).astype(np.int32)),
        "resnames": ("c", ("particle_number", "resname_length"),
                     chars([["ALA", "GLY", "LYSX", "W"][i % 4] for i in range(n)], 4)),
        "particlenames": ("c", ("particle_number", "particlename_length"),
                          chars([["CA", "CB", "OXT", "NZ12", "LONGN"][i % 5] for i in range(n)], 5)),
        "chainnames": ("c", ("particle_number", "chainname_length"), chars([["A", "B", "", "CH"][i % 4]
                                                                             for i in range(n)], 4)),
        "dynamicstate": ("b", ("particle_number",), (rng.random(n) < 0.8).astype(np.int8)),
        "coordinates": ("f", ("particle_number", "spatialdim"), rng.uniform(-30, 30, (n, 3)).astype(np.float32)),
        "radii": ("f", ("particle_number",), rng.uniform(1, 4, n).astype(np.float32)),
        "mass": ("f", ("particle_number",), mass),
        "epsilon": ("f", ("particle_number",), rng.uniform(0, 1, n).astype(np.float32)),
        "charges": ("f", ("particle_number",), rng.normal(0, 0.7, n).astype(np.float32)),
        "surfaceaccessibility": ("f", ("particle_number",), rng.uniform(0, 100, n).astype(np.float32)),
        "hydrophobicityscale": ("f", ("particle_number",), rng.uniform(-1, 1, n).astype(np.float32)),
        "hydrophobicity": ("f", ("particle_number",), rng.uniform(-2, 2, n).astype(np.float32)),
        "springs": ("i", ("spring_number", "springdim"), ij),
        "springsstiffness": ("f", ("spring_number",), rng.uniform(0.1, 10, m).astype(np.float32)),
        "springsequilibrium": ("f", ("spring_number",), rng.uniform(2, 9, m).astype(np.float32)),
    }
    for k, t in types.items():
        code, dims, data = v[k]
        v[k] = (t, dims, data.astype({"d": np.float64, "i": np.int32, "h": np.int16, "b": np.int8,
                                      "f": np.float32}[t]))
    for k in omit:
        v.pop(k, None)
    if extra_vars:
        v.update(extra_vars)
    dims = {"particle_number": None if record else n, "spatialdim": 3, "resname_length": 4,
            "particlename_length": 5, "chainname_length": 4, "springdim": 2, "spring_number": m}
    return dims, v


def write(path, dims, variables, version=1):
    f = scipy_io.netcdf_file(path, "w", version=version)
    for d, ln in dims.items():
        f.createDimension(d, ln)
    for k, (code, vd, data) in variables.items():
        var = f.createVariable(k, code, vd)
        var[:] = data
    f.close()


def expected(path):
    """What the reference's reader makes of a file, from scipy's reading."""
    f = scipy_io.netcdf_file(path, "r", mmap=False)
    V = f.variables
    g = lambda k: np.asarray(V[k].data)  # noqa: E731
    n = g("coordinates").shape[0]
    e = {"x": g("coordinates").astype(np.float32), "charge": g("charges").astype(np.float32),
         "radius": g("radii").astype(np.float32), "epsilon": g("epsilon").astype(np.float32)}
    mass = g("mass").astype(np.float32)
    mass[np.abs(mass) < 1e-6] = 1.0
    e["mass"] = mass
    for k, name in (("sasa", "surfaceaccessibility"), ("transfer", "hydrophobicityscale"),
                    ("hydrophobicity", "hydrophobicity")):
        e[k] = g(name).astype(np.float32) if name in V else np.zeros(n, np.float32)
    e["static"] = (g("dynamicstate").astype(np.uint8) == 0)
    if "springs" in V:
        e["springs"] = g("springs").astype(np.int64)
        e["k"] = g("springsstiffness").astype(np.float32)
        e["l0"] = g("springsequilibrium").astype(np.float32)
    f.close()
    return e


def load(lib, path):
    net = SpringNetwork(workloads.params(99), library=lib)
    warnings = net.loadNetCDF(path)
    return net, warnings


def check_loaded(net, e):
    x, v = net.getState()
    assert np.array_equal(x, e["x"]) and not v.any()
    p = net.getParticleProperties()
    for k in ("mass", "charge", "radius", "epsilon", "sasa", "transfer", "hydrophobicity"):
        assert np.array_equal(p[k], e[k]), k
    assert np.array_equal((p["flags"] & _abi.FLAG_STATIC) != 0, e["static"])
    a, b, l0, k = net.getSprings()
    assert np.array_equal(np.stack([a, b], 1), e["springs"])
    assert np.array_equal(l0, e["l0"]) and np.array_equal(k, e["k"])


@pytest.mark.parametrize("version,record", [(1, False), (2, False), (1, True), (2, True)])
def test_loads_what_scipy_reads(lib, tmp_path, version, record):
    path = str(tmp_path / "t.nc")
    write(path, *topology(record=record), version=version)
    net, warnings = load(lib, path)
    check_loaded(net, expected(path))
    # the zero masses became 1, with the reader's warning
    assert sum(w.startswith("Mass of particle") for w in warnings) == len(range(0, 60, 7))
    net.close()


def test_value_types_are_converted(lib, tmp_path):
    path = str(tmp_path / "t.nc")
    write(path, *topology(types={"coordinates": "d", "charges": "i", "radii": "h", "springs": "h",
                                 "springsequilibrium": "d", "dynamicstate": "i", "mass": "d"}))
    net, _ = load(lib, path)
    check_loaded(net, expected(path))
    net.close()


def test_optional_arrays_and_springs(lib, tmp_path):
    path = str(tmp_path / "t.nc")
    dims, v = topology(omit=("surfaceaccessibility", "hydrophobicity", "springs", "springsstiffness",
                             "springsequilibrium"))
    dims.pop("spring_number")
    dims.pop("springdim")
    write(path, dims, v)
    net, warnings = load(lib, path)
    assert 'Array "surfaceaccessibility" not found in topology file.' in warnings
    assert 'Array "hydrophobicity" not found in topology file.' in warnings
    assert 'Dimension "spring_number" not present in nc file.' in warnings
    p = net.getParticleProperties()
    assert not p["sasa"].any() and not p["hydrophobicity"].any() and p["transfer"].any()
    assert net.getNumberOfSprings() == 0
    net.close()


def test_labels_reach_the_pdb_writer(lib, tmp_path):
    path = str(tmp_path / "t.nc")
    write(path, *topology())
    net, _ = load(lib, path)
    net.addTrajectory(_abi.TRAJ_PDB, str(tmp_path / "t.pdb"), 1)
    net.writeNextStepNow()
    net.closeTrajectories()
    atoms = [ln for ln in (tmp_path / "t.pdb").read_text().splitlines() if ln.startswith("ATOM")]
    f = scipy_io.netcdf_file(path, "r", mmap=False)
    resid = f.variables["resids"].data.copy()
    f.close()
    names = ["CA", "CB", "OXT", "NZ12", "LONGN"]
    # particles whose residue and chain names fit their fields (longer ones
    # shift the line, as %3s and %1s do in pdb_format.hpp)
    for i in (0, 1, 4, 5, 8, 9):
        ln = atoms[i]
        nm = names[i % 5]
        assert ln[12:16] == (" %-3.3s" % nm if len(nm) < 4 else "%-4.4s" % nm)
        assert ln[17:20] == "%3s" % ["ALA", "GLY", "LYSX", "W"][i % 4][:3]
        assert ln[21] == (["A", "B", "", "CH"][i % 4] + " ")[0]
        assert int(ln[22:26]) == int(resid[i])
    net.close()


@pytest.mark.parametrize("case,message", [
    ("no_particle_dim", 'Dimension "particle_number" not present in nc file.'),
    ("missing_charges", 'Array "charges" not found in topology file.'),
    ("coords_ndims", '"coordinates" array has 1 dimension(s) whereas it should be 2.'),
    ("names_width", 'Dimension 1 of array "particlenames" has size 4 whereas it should be 5.'),
    ("spring_range", "spring 3: particle index out of range"),
    ("self_spring", "Cannot add a spring between a particle and itself."),
    ("repeated_spring", "Cannot add a spring that already exists."),
    ("hdf5", "NetCDF-4 (HDF5) files are not read here"),
])
def test_errors_carry_the_reference_messages(lib, tmp_path, case, message):
    path = str(tmp_path / "t.nc")
    dims, v = topology(n=20, m=10)
    if case == "no_particle_dim":
        dims["atoms"] = dims.pop("particle_number")
        v = {k: (c, tuple("atoms" if d == "particle_number" else d for d in vd), x) for k, (c, vd, x) in v.items()}
    elif case == "missing_charges":
        v.pop("charges")
    elif case == "coords_ndims":
        dims["flat"] = 60
        v["coordinates"] = ("f", ("flat",), v["coordinates"][2].reshape(-1))
    elif case == "names_width":
        dims["particlename_length"] = 4
        v["particlenames"] = ("c", ("particle_number", "particlename_length"), v["particlenames"][2][:, :4])
    elif case == "spring_range":
        v["springs"][2][3] = [1, 20]
    elif case == "self_spring":
        v["springs"][2][5] = [4, 4]
    elif case == "repeated_spring":
        v["springs"][2][6] = v["springs"][2][2][::-1]
    if case == "hdf5":
        with open(path, "wb") as fh:
            fh.write(b"\x89HDF\r\n\x1a\n" + b"\0" * 64)
    else:
        write(path, dims, v)
    net = SpringNetwork(workloads.params(99), library=lib)
    with pytest.raises(_abi.EngineError) as ei:
        net.loadNetCDF(path)
    assert message in str(ei.value)
    net.close()


@pytest.mark.skipif(not os.path.exists(EXAMPLE), reason="the reference's example topology is not here")
def test_reference_example_file(lib, tmp_path):
    """The reference's own example topology (1IYT.nc, 867 particles) stores
    4-character particle names, which its current reader refuses; with the
    names widened to 5 it loads, and every array is as scipy reads it."""
    net = SpringNetwork(workloads.params(99), library=lib)
    with pytest.raises(_abi.EngineError) as ei:
        net.loadNetCDF(EXAMPLE)
    assert 'Dimension 1 of array "particlenames" has size 4 whereas it should be 5.' in str(ei.value)
    src = scipy_io.netcdf_file(EXAMPLE, "r", mmap=False)
    dims = dict(src.dimensions)
    dims["particlename_length5"] = 5
    v = {}
    for k, var in src.variables.items():
        vd, data = var.dimensions, np.asarray(var.data)
        if k == "particlenames":
            vd = ("particle_number", "particlename_length5")
            data = np.concatenate([data, np.zeros((data.shape[0], 1), "S1")], 1)
        v[k] = (var.typecode(), vd, data)
    src.close()
    path = str(tmp_path / "ex.nc")
    write(path, dims, v, version=2)
    net.loadNetCDF(path)
    check_loaded(net, expected(path))
    assert net.getNumberOfParticles() == 867 and net.getNumberOfSprings() == 3588
    net.close()


def test_save_writes_the_reference_writer_layout(lib, tmp_path):
    """bs_save_netcdf (NetCDFWriter::writeBinary): scipy reads back the
    reference writer's dimensions, variables and attributes with the loaded
    values, and bs_load_netcdf restores the same network."""
    src = str(tmp_path / "in.nc")
    write(src, *topology(seed=3))
    net, _ = load(lib, src)
    out = str(tmp_path / "out.nc")
    net.saveNetCDF(out)
    f = scipy_io.netcdf_file(out, "r", mmap=False)
    assert f.version_byte == 2
    assert dict(f.dimensions) == {"spatialdim": 3, "particle_number": 60, "particlename_length": 5,
                                  "chainname_length": 4, "resname_length": 4, "springdim": 2, "spring_number": 90}
    V = f.variables
    assert list(V) == ["coordinates", "particleids", "charges", "radii", "epsilon", "mass", "surfaceaccessibility",
                       "hydrophobicityscale", "hydrophobicity", "resids", "dynamicstate", "particlenames", "resnames",
                       "chainnames", "springs", "springsstiffness", "springsequilibrium", "nbspringsperparticle"]
    assert V["coordinates"].units == b"angstrom" and V["mass"].long_name == b"Particle mass"
    assert np.array_equal(V["particleids"].data, np.arange(60))
    e = expected(src)
    assert np.array_equal(V["coordinates"].data, e["x"])
    assert np.array_equal(V["mass"].data, e["mass"])  # the zero masses already became 1
    assert np.array_equal(V["dynamicstate"].data == 0, e["static"])
    assert np.array_equal(V["springs"].data, e["springs"])
    per = np.bincount(e["springs"].reshape(-1), minlength=60)
    assert np.array_equal(V["nbspringsperparticle"].data, per)
    g = scipy_io.netcdf_file(src, "r", mmap=False)
    for k in ("particlenames", "resnames", "chainnames", "resids"):
        assert np.array_equal(V[k].data, g.variables[k].data), k
    g.close()
    f.close()
    net.close()
    back, _ = load(lib, out)
    check_loaded(back, expected(out))
    back.close()


def _var_headers(raw):
    """Offsets of each variable's vsize and begin fields in a classic
    (version 1 or 2) header: {name: (vsize_offset, begin_offset, begin_size)}."""
    import struct
    ver = raw[3]
    pos = 8

    def u32():
        nonlocal pos
        v = struct.unpack(">I", raw[pos:pos + 4])[0]
        pos += 4
        return v

    def name():
        nonlocal pos
        ln = u32()
        s = raw[pos:pos + ln].decode()
        pos += (ln + 3) // 4 * 4
        return s

    def atts():
        nonlocal pos
        tag, na = u32(), u32()
        for _ in range(na):
            name()
            t, ne = u32(), u32()
            pos += (ne * {1: 1, 2: 1, 3: 2, 4: 4, 5: 4, 6: 8}[t] + 3) // 4 * 4
    tag, nd = u32(), u32()
    for _ in range(nd):
        name()
        u32()
    atts()
    tag, nv = u32(), u32()
    out = {}
    for _ in range(nv):
        nm = name()
        for _ in range(u32()):
            u32()
        atts()
        u32()
        vs = pos
        u32()
        bsz = 8 if ver == 2 else 4
        out[nm] = (vs, pos, bsz)
        pos += bsz
    return out


@pytest.mark.parametrize("patch", ["begin_wraps", "begin_past_end", "vsize_wraps"])
def test_hostile_offsets_are_refused(lib, tmp_path, patch):
    """Offsets and sizes read from a damaged or hostile file are checked
    without overflow: a begin near 2^64, a begin past the end and record
    sizes whose sum wraps are refused with a message, not read."""
    import struct
    path = str(tmp_path / "t.nc")
    dims, v = topology(n=20, m=10, record=patch == "vsize_wraps")
    write(path, dims, v, version=2)
    raw = bytearray(open(path, "rb").read())
    hdr = _var_headers(bytes(raw))
    if patch == "begin_wraps":
        vs, b, sz = hdr["coordinates"]
        raw[b:b + 8] = struct.pack(">Q", 2 ** 64 - 16)
    elif patch == "begin_past_end":
        vs, b, sz = hdr["coordinates"]
        raw[b:b + 8] = struct.pack(">Q", len(raw) + 8)
    else:  # record variables: two vsize values that sum past 2^32 * k wrap the total
        recs = [k for k in ("coordinates", "charges", "radii", "mass") if k in hdr]
        for k in recs[:2]:
            vs, b, sz = hdr[k]
            raw[vs:vs + 4] = struct.pack(">I", 0xFFFFFFF0)
    open(path, "wb").write(bytes(raw))
    net = SpringNetwork(workloads.params(99), library=lib)
    with pytest.raises(_abi.EngineError) as ei:
        net.loadNetCDF(path)
    msg = str(ei.value)
    assert "beyond the end of the file" in msg or "malformed" in msg or "larger than the file" in msg, msg
    net.close()


def test_hostile_cdf5_lengths_are_refused(lib, tmp_path):
    """A 64-bit-data (CDF-5) header whose name length or attribute count is
    near 2^64 is refused as malformed or truncated, without reading past the
    buffer."""
    import struct
    for body in (
        # dimension list with one name of length 2^64 - 2
        struct.pack(">QIQQ", 0, 0x0A, 1, 2 ** 64 - 2) + b"ab\0\0",
        # global attribute list: one attribute of 2^62 doubles
        struct.pack(">QIQ", 0, 0, 0) + struct.pack(">IQQ", 0x0C, 1, 1) + b"a\0\0\0" +
        struct.pack(">IQ", 6, 2 ** 62),
    ):
        path = str(tmp_path / "h.nc")
        open(path, "wb").write(b"CDF\x05" + body + b"\0" * 16)
        net = SpringNetwork(workloads.params(99), library=lib)
        with pytest.raises(_abi.EngineError) as ei:
            net.loadNetCDF(path)
        msg = str(ei.value)
        assert "malformed" in msg or "truncated" in msg or "not present" in msg, msg
        net.close()
