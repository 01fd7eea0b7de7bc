"""ctypes wrappers of the oracle (liboracle.so) and, where it was built, of the
reference's own hot path (_ref/libbsref.so).  TEST INFRASTRUCTURE ONLY: used
by tests/, the smoke check and the benchmark's CPU-baseline leg as the checker,
never by the engine.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libbsref.so")

F32, F64 = 0, 1

_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_vp = ctypes.c_void_p


def _p(a, t):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(t))


def _params_type():
    from biospring_amd._abi import Params
    return Params


_olib = None
_rlib = None


def oracle_lib():
    global _olib
    if _olib is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError(f"oracle not built: {ORACLE_PATH}")
        L = ctypes.CDLL(ORACLE_PATH)
        P = ctypes.POINTER(_params_type())
        L.orc_create.argtypes = [P, ctypes.c_int64] + [_f32p] * 9 + [_u8p, ctypes.c_int64, _u32p, _u32p, _f32p, _f32p]
        L.orc_create.restype = _vp
        L.orc_destroy.argtypes = [_vp]
        L.orc_set_params.argtypes = [_vp, P]
        L.orc_set_state.argtypes = [_vp, _f32p, _f32p]
        L.orc_set_state_d.argtypes = [_vp, _f64p, _f64p]
        L.orc_get_state.argtypes = [_vp, ctypes.c_int, _f64p, _f64p]
        L.orc_compute_forces.argtypes = [_vp, ctypes.c_int, _f64p, _f64p, _f64p]
        L.orc_step.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _f64p, _f64p]
        L.orc_pairs.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _u32p, ctypes.c_int64]
        L.orc_pairs.restype = ctypes.c_int64
        L.orc_cells.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _i64p, _i64p, _i64p, _i64p, _f64p, _i64p]
        L.orc_cells.restype = ctypes.c_int64
        L.orc_num_springs.argtypes = [_vp]
        L.orc_num_springs.restype = ctypes.c_int64
        L.orc_last_nonfinite.argtypes = [_vp]
        L.orc_last_nonfinite.restype = ctypes.c_int64
        L.orc_term_energy.argtypes = [ctypes.c_int, ctypes.c_int, _f64p, ctypes.c_int]
        L.orc_term_energy.restype = ctypes.c_double
        L.orc_term_force.argtypes = [ctypes.c_int, ctypes.c_int, _f64p, ctypes.c_int, _f64p]
        L.orc_term_force.restype = ctypes.c_double
        L.orc_constants.argtypes = [_f64p]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_add_rigid_body.argtypes = [_vp, ctypes.c_int64, _i64p]
        L.orc_add_rigid_body.restype = ctypes.c_int
        L.orc_num_bodies.argtypes = [_vp]
        L.orc_num_bodies.restype = ctypes.c_int
        L.orc_rigid_state.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _f64p]
        L.orc_rigid_state.restype = ctypes.c_int
        L.orc_set_grid.argtypes = [_vp, ctypes.c_int, _f64p, _i64p, _f64p, _f32p]
        L.orc_set_grid.restype = ctypes.c_int
        L.orc_set_burying.argtypes = [_vp, _f32p]
        L.orc_insertion.argtypes = [_vp, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, _f64p]
        L.orc_insertion.restype = ctypes.c_int
        L.orc_adjoint.argtypes = [_vp, ctypes.c_int, _f64p, _f64p, _f64p, _f64p]
        L.orc_adjoint.restype = ctypes.c_double
        L.orc_loss.argtypes = [_vp, ctypes.c_int, _f64p]
        L.orc_loss.restype = ctypes.c_double
        L.orc_set_params_arrays.argtypes = [_vp, _f32p, _f32p, _f32p]
        L.orc_hessian_vec.argtypes = [_vp, _f64p, _f64p]
        L.orc_springs_from_cutoff.argtypes = [ctypes.c_int64, _f32p, _u8p, ctypes.POINTER(ctypes.c_int32),
                                              ctypes.c_double, ctypes.c_int64, _u32p, _u32p, _f32p]
        L.orc_springs_from_cutoff.restype = ctypes.c_int64
        L.orc_probe_state.argtypes = [_vp, ctypes.c_int, _f64p]
        _olib = L
    return _olib


def ref_available():
    return os.path.exists(REF_PATH)


def ref_lib():
    global _rlib
    if _rlib is None:
        if not ref_available():
            raise RuntimeError(f"reference library not built: {REF_PATH}")
        L = ctypes.CDLL(REF_PATH)
        P = ctypes.POINTER(_params_type())
        L.ref_create.argtypes = [P, ctypes.c_int64] + [_f32p] * 9 + [_u8p, ctypes.c_int64, _u32p, _u32p, _f32p, _f32p]
        L.ref_create.restype = _vp
        L.ref_destroy.argtypes = [_vp]
        L.ref_compute_forces.argtypes = [_vp, _f32p, _f32p, _f32p]
        L.ref_step.argtypes = [_vp, ctypes.c_int, _f32p, _f32p]
        L.ref_get_state.argtypes = [_vp, _f32p, _f32p]
        L.ref_cells.argtypes = [_vp, ctypes.c_int, ctypes.c_float, ctypes.c_float, _i64p, _i64p, _i64p, _i64p,
                                _f64p, _i64p]
        L.ref_cells.restype = ctypes.c_int64
        L.ref_pairs.argtypes = [P, ctypes.c_int, ctypes.c_int64, _f32p, _f32p, _f32p, _u8p, ctypes.c_int64, _u32p,
                                _u32p, _f32p, _f32p, _u32p, ctypes.c_int64]
        L.ref_pairs.restype = ctypes.c_int64
        L.ref_set_threads.argtypes = [ctypes.c_int]
        L.ref_add_rigid_body.argtypes = [_vp, ctypes.c_int64, _i64p]
        L.ref_add_rigid_body.restype = ctypes.c_int
        L.ref_rigid_state.argtypes = [ctypes.c_int, _f64p]
        L.ref_rigid_state.restype = ctypes.c_int
        L.ref_set_grid.argtypes = [_vp, ctypes.c_int, _f64p, _i64p, _f64p, _f32p, ctypes.c_double]
        L.ref_set_burying.argtypes = [_vp, _f32p]
        L.ref_insertion.argtypes = [_vp, ctypes.c_int64, ctypes.c_int64, _f64p]
        L.ref_springs_from_cutoff.argtypes = [ctypes.c_int64, _f32p, _u8p, ctypes.POINTER(ctypes.c_int32),
                                              ctypes.c_int64, _u32p, _u32p, _f32p, _f32p, ctypes.c_double,
                                              ctypes.c_double, ctypes.c_int64, _u32p, _u32p, _f32p, _f32p]
        L.ref_springs_from_cutoff.restype = ctypes.c_int64
        L.ref_get_probe.argtypes = [_vp, _f32p]
        L.ref_get_probe.restype = ctypes.c_int
        L.ref_set_trajectories.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.c_char_p, ctypes.c_int64]
        L.ref_xtc_write.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _f32p, ctypes.c_float]
        L.ref_xtc_read.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _f32p,
                                   ctypes.POINTER(ctypes.c_int), _f32p, _f32p, _f32p]
        _rlib = L
    return _rlib


def _sys_args(system):
    c = ctypes
    s = system
    return [ctypes.c_int64(s.n), _p(s.x, c.c_float), _p(s.v, c.c_float), _p(s.mass, c.c_float),
            _p(s.charge, c.c_float), _p(s.radius, c.c_float), _p(s.eps, c.c_float), _p(s.sasa, c.c_float),
            _p(s.transfer, c.c_float), _p(s.hydro, c.c_float), _p(s.flags, c.c_uint8), ctypes.c_int64(s.m),
            _p(s.id1, c.c_uint32), _p(s.id2, c.c_uint32), _p(s.l0, c.c_float), _p(s.k, c.c_float)]


class Oracle:
    """CPU restatement of the step (mode F32: float-faithful, F64: FP64)."""

    def __init__(self, system, params=None):
        self.L = oracle_lib()
        self.sys = system
        self.params = (params if params is not None else system.params).copy()
        self.h = self.L.orc_create(ctypes.byref(self.params), *_sys_args(system))
        self.n = system.n

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    __del__ = close

    def set_params(self, params):
        self.params = params.copy()
        self.L.orc_set_params(self.h, ctypes.byref(self.params))

    def set_state(self, x=None, v=None):
        xx = None if x is None else np.ascontiguousarray(x, np.float32).reshape(-1)
        vv = None if v is None else np.ascontiguousarray(v, np.float32).reshape(-1)
        self.L.orc_set_state(self.h, _p(xx, ctypes.c_float), _p(vv, ctypes.c_float))

    def set_state_f64(self, x=None, v=None):
        xx = None if x is None else np.ascontiguousarray(x, np.float64).reshape(-1)
        vv = None if v is None else np.ascontiguousarray(v, np.float64).reshape(-1)
        self.L.orc_set_state_d(self.h, _p(xx, ctypes.c_double), _p(vv, ctypes.c_double))

    def state(self, mode):
        x = np.zeros(3 * self.n)
        v = np.zeros(3 * self.n)
        self.L.orc_get_state(self.h, mode, _p(x, ctypes.c_double), _p(v, ctypes.c_double))
        return x.reshape(-1, 3), v.reshape(-1, 3)

    def compute_forces(self, mode, ext=None):
        F = np.zeros(3 * self.n)
        E = np.zeros(6)
        e = None if ext is None else np.ascontiguousarray(ext, np.float64).reshape(-1)
        self.L.orc_compute_forces(self.h, mode, _p(e, ctypes.c_double), _p(F, ctypes.c_double),
                                  _p(E, ctypes.c_double))
        return F.reshape(-1, 3), E

    def step(self, mode, nsteps, ext=None):
        E = np.zeros(6)
        e = None if ext is None else np.ascontiguousarray(ext, np.float64).reshape(-1)
        rc = self.L.orc_step(self.h, mode, nsteps, _p(e, ctypes.c_double), _p(E, ctypes.c_double))
        return rc, E

    def probe(self, mode):
        """The probe particle: position, velocity and the force of the last
        evaluation, each (3,)."""
        out = np.zeros(9)
        self.L.orc_probe_state(self.h, mode, _p(out, ctypes.c_double))
        return out[0:3], out[3:6], out[6:9]

    def pairs(self, kind, method=0, mode=F32):
        cnt = self.L.orc_pairs(self.h, kind, method, mode, None, 0)
        out = np.zeros(2 * max(cnt, 1), np.uint32)
        self.L.orc_pairs(self.h, kind, method, mode, _p(out, ctypes.c_uint32), cnt)
        return out[:2 * cnt].reshape(-1, 2)

    def add_rigid_body(self, ids):
        a = np.ascontiguousarray(ids, np.int64)
        r = self.L.orc_add_rigid_body(self.h, a.size, _p(a, ctypes.c_int64))
        if r < 0:
            raise ValueError("rigid body id out of range")
        return r

    def rigid_state(self, mode, b):
        """X, v, L, omega (3 each) and q (x, y, z, w) of body b."""
        out = np.zeros(16)
        if self.L.orc_rigid_state(self.h, mode, b, _p(out, ctypes.c_double)) != 0:
            raise IndexError(b)
        return out

    def set_grid(self, kind, origin, shape, cell, scalar):
        o = np.ascontiguousarray(origin, np.float64)
        sh = np.ascontiguousarray(shape, np.int64)
        c = np.ascontiguousarray(cell, np.float64)
        v = None if scalar is None else np.ascontiguousarray(scalar, np.float32).reshape(-1)
        if self.L.orc_set_grid(self.h, kind, _p(o, ctypes.c_double), _p(sh, ctypes.c_int64),
                               _p(c, ctypes.c_double), _p(v, ctypes.c_float)) != 0:
            raise ValueError("bad grid")

    def set_burying(self, burying):
        b = np.ascontiguousarray(burying, np.float32)
        self.L.orc_set_burying(self.h, _p(b, ctypes.c_float))

    def insertion(self, mode, i1, i2):
        """angle, roll angle, depth, vector (3)."""
        out = np.zeros(6)
        if self.L.orc_insertion(self.h, mode, i1, i2, _p(out, ctypes.c_double)) != 0:
            raise IndexError((i1, i2))
        return out

    def adjoint(self, nsteps, target):
        """FP64 forward of nsteps from the current F64 state, then the reverse
        sweep: (loss, dL/dk, dL/dL0 per spring, dL/dq per particle)."""
        y = np.ascontiguousarray(target, np.float64).reshape(-1)
        ns = int(self.L.orc_num_springs(self.h))
        dk, dl, dq = np.zeros(max(ns, 1)), np.zeros(max(ns, 1)), np.zeros(self.n)
        loss = self.L.orc_adjoint(self.h, nsteps, _p(y, ctypes.c_double), _p(dk, ctypes.c_double),
                                  _p(dl, ctypes.c_double), _p(dq, ctypes.c_double))
        return loss, dk[:ns], dl[:ns], dq

    def loss(self, nsteps, target):
        y = np.ascontiguousarray(target, np.float64).reshape(-1)
        return self.L.orc_loss(self.h, nsteps, _p(y, ctypes.c_double))

    def set_arrays(self, k=None, l0=None, q=None):
        a = [None if v is None else np.ascontiguousarray(v, np.float32) for v in (k, l0, q)]
        self.L.orc_set_params_arrays(self.h, *[_p(v, ctypes.c_float) for v in a])

    def hessian_vec(self, w):
        ww = np.ascontiguousarray(w, np.float64).reshape(-1)
        out = np.zeros(3 * self.n)
        self.L.orc_hessian_vec(self.h, _p(ww, ctypes.c_double), _p(out, ctypes.c_double))
        return out.reshape(-1, 3)

    def cells(self, kind, mode=F32):
        nm = ctypes.c_int64()
        nc = self.L.orc_cells(self.h, kind, mode, None, None, None, None, None, ctypes.byref(nm))
        cell_of = np.zeros(self.n, np.int64)
        start = np.zeros(nc + 1, np.int64)
        mem = np.zeros(max(nm.value, 1), np.int64)
        dims = np.zeros(3, np.int64)
        box = np.zeros(6)
        self.L.orc_cells(self.h, kind, mode, _p(cell_of, ctypes.c_int64), _p(start, ctypes.c_int64),
                         _p(mem, ctypes.c_int64), _p(dims, ctypes.c_int64), _p(box, ctypes.c_double),
                         ctypes.byref(nm))
        return {"cell_of": cell_of, "start": start, "members": mem[:nm.value], "dims": dims, "box": box}


def term_energy(term, mode, args):
    a = np.ascontiguousarray(args, np.float64)
    return oracle_lib().orc_term_energy(term, mode, _p(a, ctypes.c_double), a.size)


def term_force(term, mode, args):
    a = np.ascontiguousarray(args, np.float64)
    out = np.zeros(3)
    fm = oracle_lib().orc_term_force(term, mode, _p(a, ctypes.c_double), a.size, _p(out, ctypes.c_double))
    return fm, out


def constants():
    out = np.zeros(4)
    oracle_lib().orc_constants(_p(out, ctypes.c_double))
    return {"spring": out[0], "elec": out[1], "kinetic": out[2], "coulomb": out[3]}


class Reference:
    """The reference's own SpringNetwork, compiled from its sources."""

    def __init__(self, system, params=None):
        self.L = ref_lib()
        self.sys = system
        self.params = (params if params is not None else system.params).copy()
        self.h = self.L.ref_create(ctypes.byref(self.params), *_sys_args(system))
        self.n = system.n

    def close(self):
        if self.h:
            self.L.ref_destroy(self.h)
            self.h = None

    __del__ = close

    def compute_forces(self, ext=None):
        F = np.zeros(3 * self.n, np.float32)
        E = np.zeros(6, np.float32)
        e = None if ext is None else np.ascontiguousarray(ext, np.float32).reshape(-1)
        self.L.ref_compute_forces(self.h, _p(e, ctypes.c_float), _p(F, ctypes.c_float), _p(E, ctypes.c_float))
        return F.reshape(-1, 3), E

    def step(self, nsteps, ext=None):
        E = np.zeros(6, np.float32)
        e = None if ext is None else np.ascontiguousarray(ext, np.float32).reshape(-1)
        self.L.ref_step(self.h, nsteps, _p(e, ctypes.c_float), _p(E, ctypes.c_float))
        return E

    def state(self):
        x = np.zeros(3 * self.n, np.float32)
        v = np.zeros(3 * self.n, np.float32)
        self.L.ref_get_state(self.h, _p(x, ctypes.c_float), _p(v, ctypes.c_float))
        return x.reshape(-1, 3), v.reshape(-1, 3)

    def probe(self):
        """The probe particle: position, velocity and the force of its last
        step, as float32 (9,)."""
        out = np.zeros(9, np.float32)
        if self.L.ref_get_probe(self.h, _p(out, ctypes.c_float)) != 0:
            raise RuntimeError("the probe is not enabled")
        return out

    def add_rigid_body(self, ids):
        a = np.ascontiguousarray(ids, np.int64)
        return self.L.ref_add_rigid_body(self.h, a.size, _p(a, ctypes.c_int64))

    def rigid_state(self, b):
        """X, (zeros), omega of body b."""
        out = np.zeros(9)
        if self.L.ref_rigid_state(b, _p(out, ctypes.c_double)) != 0:
            raise IndexError(b)
        return out

    def set_grid(self, kind, origin, shape, cell, scalar, scale=1.0):
        o = np.ascontiguousarray(origin, np.float64)
        sh = np.ascontiguousarray(shape, np.int64)
        c = np.ascontiguousarray(cell, np.float64)
        v = np.ascontiguousarray(scalar, np.float32).reshape(-1)
        self.L.ref_set_grid(self.h, kind, _p(o, ctypes.c_double), _p(sh, ctypes.c_int64), _p(c, ctypes.c_double),
                            _p(v, ctypes.c_float), scale)

    def set_burying(self, burying):
        b = np.ascontiguousarray(burying, np.float32)
        self.L.ref_set_burying(self.h, _p(b, ctypes.c_float))

    def insertion(self, i1, i2):
        out = np.zeros(6)
        self.L.ref_insertion(self.h, i1, i2, _p(out, ctypes.c_double))
        return out

    def cells(self, kind, cutoff, skin=0.0):
        nm = ctypes.c_int64()
        nc = self.L.ref_cells(self.h, kind, cutoff, skin, None, None, None, None, None, ctypes.byref(nm))
        cell_of = np.zeros(self.n, np.int64)
        start = np.zeros(nc + 1, np.int64)
        mem = np.zeros(max(nm.value, 1), np.int64)
        dims = np.zeros(3, np.int64)
        box = np.zeros(6)
        self.L.ref_cells(self.h, kind, cutoff, skin, _p(cell_of, ctypes.c_int64), _p(start, ctypes.c_int64),
                         _p(mem, ctypes.c_int64), _p(dims, ctypes.c_int64), _p(box, ctypes.c_double),
                         ctypes.byref(nm))
        return {"cell_of": cell_of, "start": start, "members": mem[:nm.value], "dims": dims, "box": box}


def ref_pairs(system, kind, params=None):
    L = ref_lib()
    p = (params if params is not None else system.params).copy()
    s = system
    c = ctypes
    args = [ctypes.byref(p), kind, s.n, _p(s.x, c.c_float), _p(s.charge, c.c_float), _p(s.hydro, c.c_float),
            _p(s.flags, c.c_uint8), s.m, _p(s.id1, c.c_uint32), _p(s.id2, c.c_uint32), _p(s.l0, c.c_float),
            _p(s.k, c.c_float)]
    cnt = L.ref_pairs(*args, None, 0)
    out = np.zeros(2 * max(cnt, 1), np.uint32)
    L.ref_pairs(*args, _p(out, c.c_uint32), cnt)
    return out[:2 * cnt].reshape(-1, 2)


def springs_from_cutoff(x, flags, cutoff, groups=None):
    """Oracle: springs of Topology::add_springs_from_cutoff (with groups,
    add_springs_between_topologies_from_cutoff) ordered by (first, second),
    first < second, with float rest lengths."""
    x = np.ascontiguousarray(x, np.float32).reshape(-1)
    n = x.size // 3
    fl = None if flags is None else np.ascontiguousarray(flags, np.uint8)
    gr = None if groups is None else np.ascontiguousarray(groups, np.int32)
    c = ctypes
    fn = oracle_lib().orc_springs_from_cutoff
    head = [n, _p(x, c.c_float), _p(fl, c.c_uint8), _p(gr, c.c_int32), cutoff]
    cnt = fn(*head, 0, None, None, None)
    if cnt < 0:
        raise ValueError(f"cutoff {cutoff} rejected")
    i1, i2 = np.zeros(max(cnt, 1), np.uint32), np.zeros(max(cnt, 1), np.uint32)
    l0 = np.zeros(max(cnt, 1), np.float32)
    fn(*head, cnt, _p(i1, c.c_uint32), _p(i2, c.c_uint32), _p(l0, c.c_float))
    return i1[:cnt], i2[:cnt], l0[:cnt]


def ref_springs_from_cutoff(x, flags, cutoff, stiffness=1.0, existing=None, groups=None):
    """The reference's own Topology::add_springs_from_cutoff, run after the
    `existing` springs (id1, id2, l0, k) were added with add_spring; all
    springs in its order as (first, second, float(equilibrium)), plus
    float(stiffness) when `existing` is given."""
    x = np.ascontiguousarray(x, np.float32).reshape(-1)
    n = x.size // 3
    fl = None if flags is None else np.ascontiguousarray(flags, np.uint8)
    c = ctypes
    ex = [None] * 4 if existing is None else [np.ascontiguousarray(a, t) for a, t in
                                               zip(existing, (np.uint32, np.uint32, np.float32, np.float32))]
    m0 = 0 if existing is None else len(ex[0])
    gr = None if groups is None else np.ascontiguousarray(groups, np.int32)
    head = [n, _p(x, c.c_float), _p(fl, c.c_uint8), _p(gr, c.c_int32), m0, _p(ex[0], c.c_uint32),
            _p(ex[1], c.c_uint32), _p(ex[2], c.c_float), _p(ex[3], c.c_float), cutoff, stiffness]
    fn = ref_lib().ref_springs_from_cutoff
    cnt = fn(*head, 0, None, None, None, None)
    i1, i2 = np.zeros(max(cnt, 1), np.uint32), np.zeros(max(cnt, 1), np.uint32)
    l0, k = np.zeros(max(cnt, 1), np.float32), np.zeros(max(cnt, 1), np.float32)
    fn(*head, cnt, _p(i1, c.c_uint32), _p(i2, c.c_uint32), _p(l0, c.c_float), _p(k, c.c_float))
    if existing is None:
        return i1[:cnt], i2[:cnt], l0[:cnt]
    return i1[:cnt], i2[:cnt], l0[:cnt], k[:cnt]


def ref_set_trajectories(pdb=None, pdb_frequency=100, csv=None, csv_frequency=100):
    """pdbtrajectory / csvsampling settings for the next Reference(...)."""
    enc = lambda p: None if p is None else os.fsencode(p)  # noqa: E731
    ref_lib().ref_set_trajectories(enc(pdb), pdb_frequency, enc(csv), csv_frequency)


def ref_xtc_write(path, frames_nm, precision=1000.0):
    """frames (k, n, 3) in nm through the reference's xdrfile write_xtc."""
    f = np.ascontiguousarray(frames_nm, np.float32)
    rc = ref_lib().ref_xtc_write(os.fsencode(path), f.shape[0], f.shape[1], _p(f, ctypes.c_float), precision)
    if rc != 0:
        raise RuntimeError(f"write_xtc failed ({rc})")


def ref_xtc_read(path, natoms, maxframes=1000):
    """Frames of an XTC file through the reference's xdrfile read_xtc:
    (x (k, n, 3) nm, steps, times, boxes (k, 3, 3), precisions)."""
    x = np.zeros((maxframes, natoms, 3), np.float32)
    st = np.zeros(maxframes, np.int32)
    t = np.zeros(maxframes, np.float32)
    box = np.zeros((maxframes, 3, 3), np.float32)
    pr = np.zeros(maxframes, np.float32)
    k = ref_lib().ref_xtc_read(os.fsencode(path), natoms, maxframes, _p(x, ctypes.c_float),
                               st.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _p(t, ctypes.c_float),
                               _p(box, ctypes.c_float), _p(pr, ctypes.c_float))
    if k < 0:
        raise RuntimeError(f"cannot open {path}")
    return x[:k], st[:k], t[:k], box[:k], pr[:k]
