"""The C-ABI library loads and exports every symbol include/bs_api.h declares;
without a GPU, bs_create refuses with BS_ERR_NOGPU (no CPU fallback)."""
import ctypes
import os
import re

import pytest

from biospring_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "bs_api.h")) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(bs_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_exports():
    assert sorted(_abi.EXPORTS) == _declared()


def test_library_exports_every_symbol():
    L = _abi.lib()
    for name in _declared():
        assert hasattr(L, name), name


def test_params_struct_matches_header():
    with open(os.path.join(ROOT, "include", "bs_api.h")) as fh:
        text = fh.read()
    body = text[text.index("typedef struct bs_params"):text.index("} bs_params;")]
    fields = re.findall(r"^\s+(?:float|int32_t)\s+(\w+);", body, flags=re.M)
    assert fields == [f for f, _ in _abi.Params._fields_]


def test_defaults_match_reference_configuration():
    p = _abi.default_params()
    # src/configuration/Configuration.hpp:153-209
    assert p.timestep == pytest.approx(0.01) and p.neighborskin == 0.0
    assert p.steric_cutoff == 1.0 and p.steric_mode == _abi.STERIC_LINEAR
    assert p.electrostatic_cutoff == 16.0 and p.electrostatic_dielectric == 1.0
    assert p.viscosity_value == 1.0 and p.viscosity_enable == 0
    assert p.hydrophobicity_cutoff == 15.0 and p.imp_scale == 1.0
    assert p.probe_enable == 0 and p.probe_enable_steric == 0 and p.probe_enable_electrostatic == 0
    assert (p.probe_x, p.probe_y, p.probe_z) == (1.0, 1.0, 1.0)
    assert (p.probe_mass, p.probe_epsilon, p.probe_radius, p.probe_charge) == (1.0, 1.0, 1.0, 0.0)


def test_create_without_gpu_refuses():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU device node exists here")
    L = _abi.lib()
    h = ctypes.c_void_p()
    p = _abi.default_params()
    assert L.bs_create(ctypes.byref(p), ctypes.byref(h)) == _abi.BS_ERR_NOGPU
    assert not h.value
