"""ctypes binding of the engine's C ABI (include/bs_api.h).

This is the binding a Python caller would add on the reference side; the C++
drop-in (a SpringNetwork subclass) binds the same symbols directly.  Loading
fails loudly when the built library is missing: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbiospring_amd.so")
GEN_PATH = os.path.join(_HERE, "libbsgen.so")

BS_OK, BS_ERR_ARG, BS_ERR_NONFINITE, BS_ERR_HIP, BS_ERR_STATE, BS_ERR_NOGPU = range(6)
STERIC_LINEAR, STERIC_AMBER_12_6, STERIC_LEWITT_8_6, STERIC_ZACHARIAS_8_6 = range(4)
DIELECTRIC_CONSTANT, DIELECTRIC_DISTANCE = 0, 1
PAIR_CLUSTERS, PAIR_ROWS, PAIR_QUEUES = 0, 1, 2
FLAG_STATIC = 1
E_SPRING, E_ELECTROSTATIC, E_STERIC, E_KINETIC, E_IMP, E_HYDROPHOBIC = range(6)
NUM_ENERGIES = 6
NS_STERIC, NS_ELECTROSTATIC, NS_HYDROPHOBIC = range(3)
PROF_NUM = 7  # BS_PROF_REBUILD, _PAIR, _SPRING, _ADJ_PAIR, _ADJ_SPRING, _TOPO_COUNT, _TOPO_FILL
GRID_POTENTIAL, GRID_DENSITY = 0, 1
TRAJ_PDB, TRAJ_XTC, TRAJ_CSV = 0, 1, 2

# Exported symbols, one per declaration in include/bs_api.h.
EXPORTS = (
    "bs_default_params", "bs_create", "bs_destroy", "bs_last_error", "bs_num_particles",
    "bs_num_springs", "bs_set_particles", "bs_set_springs", "bs_set_params", "bs_setup",
    "bs_step", "bs_compute_forces", "bs_get_state", "bs_set_state", "bs_get_energies",
    "bs_last_nonfinite_particle", "bs_get_counters", "bs_profile", "bs_get_profile",
    "bs_get_cells", "bs_get_pairs", "bs_add_rigid_body", "bs_num_rigid_bodies", "bs_get_rigid_body", "bs_get_probe",
    "bs_set_grid", "bs_set_burying", "bs_get_insertion_vector", "bs_set_systems", "bs_num_systems",
    "bs_get_system_energies", "bs_autotune", "bs_adjoint", "bs_adjoint_forward", "bs_adjoint_backward",
    "bs_adjoint_backward_noalloc",
    "bs_springs_from_cutoff", "bs_get_springs", "bs_add_trajectory", "bs_write_frame_now",
    "bs_flush_trajectories", "bs_close_trajectories", "bs_set_particle_labels", "bs_set_insertion_vector",
    "bs_load_netcdf", "bs_get_particle_properties", "bs_set_sasa", "bs_set_membrane",
    "bs_springs_between_groups_from_cutoff", "bs_save_netcdf",
)
PROF_REBUILD, PROF_PAIR, PROF_SPRING = range(3)


class Params(ctypes.Structure):
    """Mirror of bs_params (field order and types must match the header)."""
    _fields_ = [
        ("timestep", ctypes.c_float),
        ("neighborskin", ctypes.c_float),
        ("spring_enable", ctypes.c_int32),
        ("spring_scale", ctypes.c_float),
        ("steric_enable", ctypes.c_int32),
        ("steric_mode", ctypes.c_int32),
        ("steric_cutoff", ctypes.c_float),
        ("steric_gridscale", ctypes.c_float),
        ("electrostatic_enable", ctypes.c_int32),
        ("electrostatic_cutoff", ctypes.c_float),
        ("electrostatic_scale", ctypes.c_float),
        ("electrostatic_dielectric", ctypes.c_float),
        ("dielectric_model", ctypes.c_int32),
        ("viscosity_enable", ctypes.c_int32),
        ("viscosity_value", ctypes.c_float),
        ("imp_enable", ctypes.c_int32),
        ("imp_scale", ctypes.c_float),
        ("imp_upper_offset", ctypes.c_float),
        ("imp_lower_offset", ctypes.c_float),
        ("imp_upper_curvature", ctypes.c_float),
        ("imp_lower_curvature", ctypes.c_float),
        ("hydrophobicity_enable", ctypes.c_int32),
        ("hydrophobicity_cutoff", ctypes.c_float),
        ("hydrophobicity_scale", ctypes.c_float),
        ("list_skin", ctypes.c_float),
        ("rebuild_interval", ctypes.c_int32),
        ("compensated_positions", ctypes.c_int32),
        ("rigidbody_enable", ctypes.c_int32),
        ("potentialgrid_enable", ctypes.c_int32),
        ("potentialgrid_scale", ctypes.c_float),
        ("densitygrid_enable", ctypes.c_int32),
        ("densitygrid_scale", ctypes.c_float),
        ("graph_steps", ctypes.c_int32),
        ("pair_layout", ctypes.c_int32),
        ("probe_enable", ctypes.c_int32),
        ("probe_enable_electrostatic", ctypes.c_int32),
        ("probe_enable_steric", ctypes.c_int32),
        ("probe_x", ctypes.c_float),
        ("probe_y", ctypes.c_float),
        ("probe_z", ctypes.c_float),
        ("probe_mass", ctypes.c_float),
        ("probe_epsilon", ctypes.c_float),
        ("probe_radius", ctypes.c_float),
        ("probe_charge", ctypes.c_float),
    ]

    def copy(self):
        q = Params()
        ctypes.pointer(q)[0] = self
        return q

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_vp = ctypes.c_void_p


def ptr(a, ctype):
    """Pointer to a contiguous numpy array (None -> NULL)."""
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctype))


_lib = None
_gen = None


def lib():
    """The engine library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"engine library not built: {LIB_PATH} (run __graft_entry__.build())")
    _lib = bind(ctypes.CDLL(LIB_PATH))
    return _lib


def bind(L):
    """Declares the C signatures of include/bs_api.h on a loaded library."""
    L.bs_default_params.argtypes = [ctypes.POINTER(Params)]
    L.bs_default_params.restype = None
    L.bs_create.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(_vp)]
    L.bs_destroy.argtypes = [_vp]
    L.bs_last_error.argtypes = [_vp]
    L.bs_last_error.restype = ctypes.c_char_p
    L.bs_num_particles.argtypes = [_vp]
    L.bs_num_particles.restype = ctypes.c_int64
    L.bs_num_springs.argtypes = [_vp]
    L.bs_num_springs.restype = ctypes.c_int64
    L.bs_set_particles.argtypes = [_vp, ctypes.c_int64] + [_f32p] * 9 + [_u8p]
    L.bs_set_springs.argtypes = [_vp, ctypes.c_int64, _u32p, _u32p, _f32p, _f32p]
    L.bs_set_params.argtypes = [_vp, ctypes.POINTER(Params)]
    L.bs_setup.argtypes = [_vp]
    L.bs_step.argtypes = [_vp, ctypes.c_int32, _f32p]
    L.bs_compute_forces.argtypes = [_vp, _f32p, _f32p, _f64p]
    L.bs_get_state.argtypes = [_vp, _f32p, _f32p]
    L.bs_set_state.argtypes = [_vp, _f32p, _f32p]
    L.bs_get_energies.argtypes = [_vp, _f64p]
    L.bs_last_nonfinite_particle.argtypes = [_vp]
    L.bs_last_nonfinite_particle.restype = ctypes.c_int64
    L.bs_get_counters.argtypes = [_vp, _i64p]
    L.bs_profile.argtypes = [_vp, ctypes.c_int32]
    L.bs_get_profile.argtypes = [_vp, _f64p, _i64p]
    L.bs_get_cells.argtypes = [_vp, ctypes.c_int32, _i64p, _i64p, _i64p, _i32p, _i32p, _i64p, _f64p]
    L.bs_get_pairs.argtypes = [_vp, ctypes.c_int32, _u32p, ctypes.c_int64, _i64p]
    L.bs_add_rigid_body.argtypes = [_vp, ctypes.c_int64, _i64p]
    L.bs_num_rigid_bodies.argtypes = [_vp]
    L.bs_num_rigid_bodies.restype = ctypes.c_int32
    L.bs_get_probe.argtypes = [_vp, _f32p, _f32p, _f32p]
    L.bs_get_rigid_body.argtypes = [_vp, ctypes.c_int32, _f64p]
    L.bs_set_grid.argtypes = [_vp, ctypes.c_int32, _f64p, _i64p, _f64p, _f32p]
    L.bs_set_burying.argtypes = [_vp, _f32p]
    L.bs_get_insertion_vector.argtypes = [_vp, ctypes.c_int64, ctypes.c_int64, _f64p]
    L.bs_set_systems.argtypes = [_vp, ctypes.c_int32, _i64p]
    L.bs_num_systems.argtypes = [_vp]
    L.bs_num_systems.restype = ctypes.c_int32
    L.bs_get_system_energies.argtypes = [_vp, _f64p]
    L.bs_autotune.argtypes = [_vp, ctypes.c_int32, _i32p, _i32p, _i32p]
    L.bs_adjoint.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, _f32p, _f32p, _f64p, _f64p, _f64p, _f64p]
    L.bs_adjoint_forward.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, _f32p]
    L.bs_springs_from_cutoff.argtypes = [_vp, ctypes.c_double, ctypes.c_float, _i64p]
    L.bs_get_springs.argtypes = [_vp, _u32p, _u32p, _f32p, _f32p]
    L.bs_springs_between_groups_from_cutoff.argtypes = [_vp, ctypes.c_double, ctypes.c_float, _i32p, _i64p]
    L.bs_adjoint_backward.argtypes = [_vp, _f64p, _f64p, _f64p, _f64p, _f64p]
    L.bs_adjoint_backward_noalloc.argtypes = [_vp, _f64p, _f64p, _f64p, _f64p, _f64p]
    L.bs_add_trajectory.argtypes = [_vp, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int64]
    L.bs_write_frame_now.argtypes = [_vp]
    L.bs_flush_trajectories.argtypes = [_vp]
    L.bs_close_trajectories.argtypes = [_vp]
    _strs = ctypes.POINTER(ctypes.c_char_p)
    L.bs_set_particle_labels.argtypes = [_vp, _strs, _strs, _strs, _i32p, _f32p, _f32p, _strs, _i32p]
    L.bs_set_insertion_vector.argtypes = [_vp, ctypes.c_int64, ctypes.c_int64]
    L.bs_load_netcdf.argtypes = [_vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int64]
    L.bs_save_netcdf.argtypes = [_vp, ctypes.c_char_p]
    L.bs_get_particle_properties.argtypes = [_vp] + [_f32p] * 7 + [_u8p]
    L.bs_set_sasa.argtypes = [_vp, _f32p]
    L.bs_set_membrane.argtypes = [_vp] + [ctypes.c_float] * 4
    return L


def gen():
    """The host-only workload generator library."""
    global _gen
    if _gen is not None:
        return _gen
    if not os.path.exists(GEN_PATH):
        raise RuntimeError(f"generator library not built: {GEN_PATH}")
    G = ctypes.CDLL(GEN_PATH)
    G.bsg_params.argtypes = [ctypes.c_int, ctypes.POINTER(Params)]
    G.bsg_params.restype = None
    G.bsg_generate.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int64]
    G.bsg_generate.restype = _vp
    G.bsg_num_particles.argtypes = [_vp]
    G.bsg_num_particles.restype = ctypes.c_int64
    G.bsg_num_springs.argtypes = [_vp]
    G.bsg_num_springs.restype = ctypes.c_int64
    G.bsg_copy.argtypes = [_vp] + [_f32p] * 9 + [_u8p, _u32p, _u32p, _f32p, _f32p]
    G.bsg_copy.restype = None
    G.bsg_free.argtypes = [_vp]
    G.bsg_free.restype = None
    _gen = G
    return G


def default_params():
    p = Params()
    lib().bs_default_params(ctypes.byref(p))
    return p


def check(code, handle=None, what=""):
    if code != BS_OK:
        msg = lib().bs_last_error(handle).decode() if handle else ""
        raise EngineError(code, f"{what}: code {code}: {msg}")


class EngineError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def u32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.uint32)
