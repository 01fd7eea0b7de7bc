"""Seeded synthetic spring networks (host generator, libbsgen.so).

Workload 0: 2,000-bead CA elastic network (chain in a 30 A sphere, springs
< 12 A, linear steric 4 A, Coulomb 16 A, 20% charged).  Workload 1: 250,000
beads in a 140 A cube (springs < 8 A, Amber 12-6 + Coulomb with eps(r) = r,
12 A cutoffs).  Workload 2: IMPALA CG beads.  Workload 3: one system of a
batch.  Workload 4: workload 1's beads at 100,000 with springs < 10 A (the
fitting network).  See biospring_amd/csrc/bs_gen.cpp.
"""
import ctypes
import hashlib
from dataclasses import dataclass, field

import numpy as np

from . import _abi


@dataclass
class System:
    x: np.ndarray
    v: np.ndarray
    mass: np.ndarray
    charge: np.ndarray
    radius: np.ndarray
    eps: np.ndarray
    sasa: np.ndarray
    transfer: np.ndarray
    hydro: np.ndarray
    flags: np.ndarray
    id1: np.ndarray
    id2: np.ndarray
    l0: np.ndarray
    k: np.ndarray
    params: _abi.Params = field(default_factory=_abi.Params)

    @property
    def n(self):
        return int(self.mass.shape[0])

    @property
    def m(self):
        return int(self.id1.shape[0])

    def digest(self):
        h = hashlib.sha256()
        for a in (self.x, self.v, self.mass, self.charge, self.radius, self.eps, self.sasa,
                  self.transfer, self.hydro, self.flags, self.id1, self.id2, self.l0, self.k):
            h.update(np.ascontiguousarray(a).tobytes())
        return h.hexdigest()[:16]

    def particle_arrays(self):
        return (self.x, self.v, self.mass, self.charge, self.radius, self.eps, self.sasa,
                self.transfer, self.hydro, self.flags)

    def spring_arrays(self):
        return (self.id1, self.id2, self.l0, self.k)


def params(config):
    p = _abi.Params()
    _abi.gen().bsg_params(config, ctypes.byref(p))
    return p


def generate(config, seed=1, n=0):
    """Generates workload `config` with `seed`; n > 0 overrides the bead count
    (same density)."""
    G = _abi.gen()
    h = G.bsg_generate(config, seed, n)
    if not h:
        raise ValueError(f"unknown workload {config}")
    try:
        nn = G.bsg_num_particles(h)
        m = G.bsg_num_springs(h)
        f = lambda k=1: np.zeros(nn * k, np.float32)
        s = System(x=f(3), v=f(3), mass=f(), charge=f(), radius=f(), eps=f(), sasa=f(),
                   transfer=f(), hydro=f(), flags=np.zeros(nn, np.uint8),
                   id1=np.zeros(m, np.uint32), id2=np.zeros(m, np.uint32),
                   l0=np.zeros(m, np.float32), k=np.zeros(m, np.float32))
        P = _abi.ptr
        c = ctypes
        G.bsg_copy(h, P(s.x, c.c_float), P(s.v, c.c_float), P(s.mass, c.c_float),
                   P(s.charge, c.c_float), P(s.radius, c.c_float), P(s.eps, c.c_float),
                   P(s.sasa, c.c_float), P(s.transfer, c.c_float), P(s.hydro, c.c_float),
                   P(s.flags, c.c_uint8), P(s.id1, c.c_uint32), P(s.id2, c.c_uint32),
                   P(s.l0, c.c_float), P(s.k, c.c_float))
    finally:
        G.bsg_free(h)
    s.params = params(config)
    return s


def rigid_groups(system, nbodies=4, size=2000):
    """Disjoint compact groups for workload 2's rigid bodies: around each of
    `nbodies` centres spread over the system's xy extent (z = its mean), the
    `size` closest beads not already taken.  Deterministic."""
    x = system.x.reshape(-1, 3).astype(np.float64)
    lo, hi = x.min(0), x.max(0)
    mid = 0.5 * (lo + hi)
    q = 0.25 * (hi - lo)
    centres = [mid + np.array([sx * q[0], sy * q[1], 0.0]) for sx, sy in ((-1, -1), (1, -1), (-1, 1), (1, 1))]
    taken = np.zeros(system.n, bool)
    groups = []
    for c in centres[:nbodies]:
        d = np.linalg.norm(x - c, axis=1)
        d[taken] = np.inf
        ids = np.sort(np.argsort(d, kind="stable")[:min(size, system.n)])
        taken[ids] = True
        groups.append(ids.astype(np.int64))
    return groups


def impala_params(system, membrane="double"):
    """Workload 2 parameters: IMPALA (flat slab, or a curved double membrane
    of tube radius 200 A with 30 A offsets) with rigid bodies enabled."""
    p = system.params.copy()
    p.imp_enable = 1
    p.rigidbody_enable = 1
    if membrane == "double":
        p.imp_upper_offset, p.imp_lower_offset = 30.0, 30.0
        p.imp_upper_curvature, p.imp_lower_curvature = 1.0 / 200.0, -1.0 / 200.0
    return p


def batch(nsys=512, seed=1, nmin=200, nmax=20000):
    """Workload 3: `nsys` independent CA networks, sizes log-uniform in
    [nmin, nmax], each generated like workload 0 from its own seed with a
    spring cutoff drawn from U[9, 14] A.  Returns the concatenated system
    and the particle offsets of the systems (nsys + 1)."""
    rng = np.random.default_rng(seed)
    sizes = np.rint(np.exp(rng.uniform(np.log(nmin), np.log(nmax), nsys))).astype(np.int64)
    parts = [generate(3, seed=int(seed) * 100003 + s, n=int(sizes[s])) for s in range(nsys)]
    off = np.zeros(nsys + 1, np.int64)
    off[1:] = np.cumsum([p.n for p in parts])
    cat = lambda k: np.concatenate([getattr(p, k) for p in parts])
    out = System(x=cat("x"), v=cat("v"), mass=cat("mass"), charge=cat("charge"), radius=cat("radius"),
                 eps=cat("eps"), sasa=cat("sasa"), transfer=cat("transfer"), hydro=cat("hydro"),
                 flags=cat("flags"),
                 id1=np.concatenate([p.id1 + np.uint32(off[s]) for s, p in enumerate(parts)]),
                 id2=np.concatenate([p.id2 + np.uint32(off[s]) for s, p in enumerate(parts)]),
                 l0=cat("l0"), k=cat("k"))
    out.params = parts[0].params.copy()
    return out, off, parts


def fitting_target(system, seed=1, modes=3, amplitude=5.0):
    """Workload 4's target conformation: the positions displaced along
    `modes` smooth low-frequency fields u_m(x) = a_m sin(k_m . x + phi_m)
    (a_m a random unit vector, k_m = 2 pi / L times a nonzero vector of
    -1/0/1 entries over the system's extent L, phi_m uniform), scaled so
    that the largest displacement is `amplitude` A."""
    rng = np.random.default_rng(seed)
    x = system.x.reshape(-1, 3).astype(np.float64)
    ext = np.maximum(x.max(0) - x.min(0), 1e-9)
    u = np.zeros_like(x)
    for _ in range(modes):
        a = rng.normal(size=3)
        a /= np.linalg.norm(a)
        k = np.zeros(3)
        while not k.any():
            k = rng.integers(-1, 2, size=3).astype(np.float64)
        ph = rng.uniform(0.0, 2.0 * np.pi)
        u += np.sin((x - x.min(0)) @ (2.0 * np.pi * k / ext) + ph)[:, None] * a
    u *= amplitude / max(np.linalg.norm(u, axis=1).max(), 1e-12)
    return (x + u).astype(np.float32)
