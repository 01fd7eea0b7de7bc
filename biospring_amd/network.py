"""Python mirror of biospring::spn::SpringNetwork's step API over the C ABI.

Method names follow the reference (src/spn/SpringNetwork.h:303-397):
setup, computeForces, computeStep, run, getParticlePosition, setForce and the
energy getters.  Every call goes through libbiospring_amd.so; there is no CPU
path.
"""
import ctypes
import os

import numpy as np

from . import _abi
from ._abi import check, f32, ptr, u32


class SpringNetwork:
    """One simulation on the current HIP device."""

    def __init__(self, params=None, library=None):
        L = library if library is not None else _abi.lib()
        self._L = L
        self._h = ctypes.c_void_p()
        self.params = params.copy() if params is not None else _abi.default_params()
        code = L.bs_create(ctypes.byref(self.params), ctypes.byref(self._h))
        if code != _abi.BS_OK:
            raise _abi.EngineError(code, f"bs_create failed with code {code} (no gfx950 device?)")
        self._ext = None
        self.n = 0

    # -- construction ---------------------------------------------------------
    def setParticles(self, x, v=None, mass=None, charge=None, radius=None, epsilon=None,
                     sasa=None, transfer=None, hydrophobicity=None, flags=None):
        x = f32(x).reshape(-1)
        n = x.size // 3
        arrs = [f32(a) for a in (v, mass, charge, radius, epsilon, sasa, transfer, hydrophobicity)]
        fl = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint8)
        c = ctypes
        check(self._L.bs_set_particles(self._h, n, ptr(x, c.c_float),
                                       *[ptr(a, c.c_float) for a in arrs], ptr(fl, c.c_uint8)),
              self._h, "bs_set_particles")
        self._keep = (x, arrs, fl)
        self.n = n

    def setSprings(self, id1, id2, equilibrium, stiffness):
        a, b, l0, k = u32(id1), u32(id2), f32(equilibrium), f32(stiffness)
        c = ctypes
        check(self._L.bs_set_springs(self._h, a.size, ptr(a, c.c_uint32), ptr(b, c.c_uint32),
                                     ptr(l0, c.c_float), ptr(k, c.c_float)), self._h, "bs_set_springs")

    def springsFromCutoff(self, cutoff, stiffness=1.0):
        """Topology::add_springs_from_cutoff on the GPU: a spring for every
        pair closer than `cutoff` at the current positions (see
        bs_springs_from_cutoff); returns the number of springs."""
        m = ctypes.c_int64()
        check(self._L.bs_springs_from_cutoff(self._h, float(cutoff), float(stiffness), ctypes.byref(m)), self._h,
              "bs_springs_from_cutoff")
        return int(m.value)

    def springsBetweenGroupsFromCutoff(self, cutoff, group, stiffness=1.0):
        """Topology::add_springs_between_topologies_from_cutoff: springs only
        between particles of different groups (topology ids)."""
        g = np.ascontiguousarray(group, np.int32)
        m = ctypes.c_int64()
        check(self._L.bs_springs_between_groups_from_cutoff(self._h, float(cutoff), float(stiffness),
                                                            ptr(g, ctypes.c_int32), ctypes.byref(m)),
              self._h, "bs_springs_between_groups_from_cutoff")
        return int(m.value)

    def getSprings(self):
        """(id1, id2, equilibrium, stiffness) of the loaded springs."""
        m = int(self._L.bs_num_springs(self._h))
        a, b = np.zeros(max(m, 1), np.uint32), np.zeros(max(m, 1), np.uint32)
        l0, k = np.zeros(max(m, 1), np.float32), np.zeros(max(m, 1), np.float32)
        c = ctypes
        check(self._L.bs_get_springs(self._h, ptr(a, c.c_uint32), ptr(b, c.c_uint32), ptr(l0, c.c_float),
                                     ptr(k, c.c_float)), self._h, "bs_get_springs")
        return a[:m], b[:m], l0[:m], k[:m]

    # -- interactive updates (the interactor bridge)
    def setSolventAccessibilitySurfaces(self, sasa):
        """InteractorFreeSASA: per-particle SASA for IMPALA from the next step."""
        a = f32(sasa)
        check(self._L.bs_set_sasa(self._h, ptr(a, ctypes.c_float)), self._h, "bs_set_sasa")

    def setMembrane(self, upper_offset=float("nan"), lower_offset=float("nan"), upper_curvature=float("nan"),
                    lower_curvature=float("nan")):
        """MDDriver custom data dmou / dmol / dmtc; NaN keeps a value."""
        check(self._L.bs_set_membrane(self._h, upper_offset, lower_offset, upper_curvature, lower_curvature),
              self._h, "bs_set_membrane")

    # -- topology files (io::readTopology + Topology::to_spring_network)
    def loadNetCDF(self, path):
        """Particles, springs and PDB labels from a .nc topology; returns the
        reader's warnings."""
        buf = ctypes.create_string_buffer(1 << 16)
        check(self._L.bs_load_netcdf(self._h, os.fsencode(path), buf, len(buf)), self._h, "bs_load_netcdf")
        self.n = int(self._L.bs_num_particles(self._h))
        return [w for w in buf.value.decode().split("\n") if w]

    def saveNetCDF(self, path):
        """NetCDFWriter::writeBinary of the loaded network at its current positions."""
        check(self._L.bs_save_netcdf(self._h, os.fsencode(path)), self._h, "bs_save_netcdf")

    def getParticleProperties(self):
        """dict of mass, charge, radius, epsilon, sasa, transfer, hydrophobicity, flags."""
        n = int(self._L.bs_num_particles(self._h))
        out = {k: np.zeros(max(n, 1), np.float32) for k in
               ("mass", "charge", "radius", "epsilon", "sasa", "transfer", "hydrophobicity")}
        out["flags"] = np.zeros(max(n, 1), np.uint8)
        c = ctypes
        check(self._L.bs_get_particle_properties(
            self._h, *[ptr(out[k], c.c_float) for k in ("mass", "charge", "radius", "epsilon", "sasa", "transfer",
                                                         "hydrophobicity")], ptr(out["flags"], c.c_uint8)),
              self._h, "bs_get_particle_properties")
        return {k: v[:n] for k, v in out.items()}

    # -- trajectory output (SpringNetwork::_setupTrajectories / writeNextStepNow)
    def addTrajectory(self, kind, path, frequency=100):
        """A PDB / XTC / CSV writer (_abi.TRAJ_*) taking a frame every
        `frequency` iterations of run()."""
        check(self._L.bs_add_trajectory(self._h, int(kind), os.fsencode(path), int(frequency)), self._h,
              "bs_add_trajectory")

    def writeNextStepNow(self):
        check(self._L.bs_write_frame_now(self._h), self._h, "bs_write_frame_now")

    def flushTrajectories(self):
        check(self._L.bs_flush_trajectories(self._h), self._h, "bs_flush_trajectories")

    def closeTrajectories(self):
        check(self._L.bs_close_trajectories(self._h), self._h, "bs_close_trajectories")

    def setParticleLabels(self, name=None, resname=None, chain=None, resid=None, occupancy=None,
                          tempfactor=None, element=None, electron_charge=None):
        """PDB fields per particle (lists of str, or arrays); None keeps the
        reference's defaults."""
        def strs(v):
            if v is None:
                return None
            arr = (ctypes.c_char_p * self.n)(*[str(t).encode() for t in v])
            keep.append(arr)
            return ctypes.cast(arr, ctypes.POINTER(ctypes.c_char_p))
        keep = []
        ri = None if resid is None else np.ascontiguousarray(resid, np.int32)
        oc = None if occupancy is None else f32(occupancy)
        tf = None if tempfactor is None else f32(tempfactor)
        ec = None if electron_charge is None else np.ascontiguousarray(electron_charge, np.int32)
        c = ctypes
        check(self._L.bs_set_particle_labels(self._h, strs(name), strs(resname), strs(chain), ptr(ri, c.c_int32),
                                             ptr(oc, c.c_float), ptr(tf, c.c_float), strs(element),
                                             ptr(ec, c.c_int32)), self._h, "bs_set_particle_labels")

    def setInsertionVector(self, i1, i2):
        check(self._L.bs_set_insertion_vector(self._h, int(i1), int(i2)), self._h, "bs_set_insertion_vector")

    def getNumberOfParticles(self):
        return int(self._L.bs_num_particles(self._h))

    def getNumberOfSprings(self):
        return int(self._L.bs_num_springs(self._h))

    @classmethod
    def from_system(cls, system, params=None, library=None):
        net = cls(params if params is not None else system.params, library=library)
        net.setParticles(*system.particle_arrays())
        net.setSprings(*system.spring_arrays())
        return net

    def setParams(self, params):
        self.params = params.copy()
        check(self._L.bs_set_params(self._h, ctypes.byref(self.params)), self._h, "bs_set_params")

    def setup(self):
        check(self._L.bs_setup(self._h), self._h, "bs_setup")

    # -- step -------------------------------------------------------------------
    def setForce(self, forces):
        """External forces (n,3) added before every subsequent step's force
        evaluation (SpringNetwork::setForce); None clears them."""
        self._ext = None if forces is None else f32(forces).reshape(-1)

    def computeStep(self):
        self.run(1)

    def run(self, nsteps):
        check(self._L.bs_step(self._h, int(nsteps), ptr(self._ext, ctypes.c_float)), self._h, "bs_step")

    def computeForces(self):
        """Total forces (n,3) and the six FP64 energies at the current state."""
        F = np.zeros(3 * self.n, np.float32)
        E = np.zeros(_abi.NUM_ENERGIES, np.float64)
        check(self._L.bs_compute_forces(self._h, ptr(self._ext, ctypes.c_float), ptr(F, ctypes.c_float),
                                        ptr(E, ctypes.c_double)), self._h, "bs_compute_forces")
        return F.reshape(-1, 3), E

    # -- state ------------------------------------------------------------------
    def getState(self):
        x = np.zeros(3 * self.n, np.float32)
        v = np.zeros(3 * self.n, np.float32)
        check(self._L.bs_get_state(self._h, ptr(x, ctypes.c_float), ptr(v, ctypes.c_float)), self._h,
              "bs_get_state")
        return x.reshape(-1, 3), v.reshape(-1, 3)

    def setState(self, x=None, v=None):
        xx = None if x is None else f32(x).reshape(-1)
        vv = None if v is None else f32(v).reshape(-1)
        check(self._L.bs_set_state(self._h, ptr(xx, ctypes.c_float), ptr(vv, ctypes.c_float)), self._h,
              "bs_set_state")

    def getParticlePosition(self, i):
        return self.getState()[0][i]

    def getEnergies(self):
        E = np.zeros(_abi.NUM_ENERGIES, np.float64)
        check(self._L.bs_get_energies(self._h, ptr(E, ctypes.c_double)), self._h, "bs_get_energies")
        return E

    def getSpringEnergy(self):
        return self.getEnergies()[_abi.E_SPRING]

    def getElectrostaticEnergy(self):
        return self.getEnergies()[_abi.E_ELECTROSTATIC]

    def getStericEnergy(self):
        return self.getEnergies()[_abi.E_STERIC]

    def getKineticEnergy(self):
        return self.getEnergies()[_abi.E_KINETIC]

    def getIMPEnergy(self):
        return self.getEnergies()[_abi.E_IMP]

    def getHydrophobicEnergy(self):
        return self.getEnergies()[_abi.E_HYDROPHOBIC]

    def counters(self):
        c = np.zeros(6, np.int64)
        check(self._L.bs_get_counters(self._h, ptr(c, ctypes.c_int64)), self._h, "bs_get_counters")
        return {"steps": int(c[0]), "rebuilds": int(c[1]), "max_list": int(c[2]), "list_cap": int(c[3]),
                "edges": int(c[4])}

    def profile(self, enable=True):
        """Per-kernel-class HIP-event timing of run() and adjoint() (resets the
        totals)."""
        check(self._L.bs_profile(self._h, 1 if enable else 0), self._h, "bs_profile")

    def getProfile(self):
        ms = np.zeros(_abi.PROF_NUM, np.float64)
        n = np.zeros(_abi.PROF_NUM, np.int64)
        check(self._L.bs_get_profile(self._h, ptr(ms, ctypes.c_double), ptr(n, ctypes.c_int64)), self._h,
              "bs_get_profile")
        names = ("rebuild", "pair", "spring", "adj_pair", "adj_spring", "topo_count", "topo_fill")
        return {names[c]: {"ms": float(ms[c]), "launches": int(n[c])} for c in range(_abi.PROF_NUM)}

    def lastNonfiniteParticle(self):
        return int(self._L.bs_last_nonfinite_particle(self._h))

    # -- parity hooks -------------------------------------------------------------
    def getCells(self, kind):
        c = ctypes
        nc, nm = c.c_int64(), c.c_int64()
        check(self._L.bs_get_cells(self._h, kind, c.byref(nc), c.byref(nm), None, None, None, None, None),
              self._h, "bs_get_cells")
        cell_of = np.zeros(self.n, np.int64)
        start = np.zeros(nc.value + 1, np.int32)
        idx = np.zeros(max(nm.value, 1), np.int32)
        dims = np.zeros(3, np.int64)
        box = np.zeros(6, np.float64)
        check(self._L.bs_get_cells(self._h, kind, c.byref(nc), c.byref(nm), ptr(cell_of, c.c_int64),
                                   ptr(start, c.c_int32), ptr(idx, c.c_int32), ptr(dims, c.c_int64),
                                   ptr(box, c.c_double)), self._h, "bs_get_cells")
        return {"cell_of": cell_of, "start": start.astype(np.int64), "members": idx[:nm.value].astype(np.int64),
                "dims": dims, "box": box}

    def countPairs(self, kind):
        """Number of processed pairs of search `kind` at the current state."""
        cnt = ctypes.c_int64()
        check(self._L.bs_get_pairs(self._h, kind, None, 0, ctypes.byref(cnt)), self._h, "bs_get_pairs")
        return int(cnt.value)

    def getPairs(self, kind):
        c = ctypes
        cnt = c.c_int64()
        check(self._L.bs_get_pairs(self._h, kind, None, 0, c.byref(cnt)), self._h, "bs_get_pairs")
        out = np.zeros(2 * max(cnt.value, 1), np.uint32)
        check(self._L.bs_get_pairs(self._h, kind, ptr(out, c.c_uint32), cnt.value, c.byref(cnt)), self._h,
              "bs_get_pairs")
        return out[:2 * cnt.value].reshape(-1, 2)

    # -- rigid bodies (RigidBodiesManager, src/rigidbody/RigidBodiesManager.cpp:11-60) --
    def addRigidBody(self, ids):
        """InitRigidBodies(this, ids): a new body unless an id is already rigid."""
        a = np.ascontiguousarray(ids, np.int64)
        check(self._L.bs_add_rigid_body(self._h, a.size, ptr(a, ctypes.c_int64)), self._h, "bs_add_rigid_body")
        return self.numRigidBodies()

    def getProbeParticle(self):
        """The probe particle (SpringNetwork::getProbeParticle): position,
        velocity and the force of its last step, each (3,) float32."""
        x, v, f = (np.zeros(3, np.float32) for _ in range(3))
        check(self._L.bs_get_probe(self._h, ptr(x, ctypes.c_float), ptr(v, ctypes.c_float), ptr(f, ctypes.c_float)),
              self._h, "bs_get_probe")
        return x, v, f

    def numRigidBodies(self):
        return int(self._L.bs_num_rigid_bodies(self._h))

    def getRigidBody(self, b):
        """X, v, L, omega (3 each) and q (x, y, z, w) of body b."""
        out = np.zeros(16)
        check(self._L.bs_get_rigid_body(self._h, b, ptr(out, ctypes.c_double)), self._h, "bs_get_rigid_body")
        return out

    # -- fields (SpringNetwork::getPotentialGrid / getDensityGrid, SpringNetwork.h:179-184) --
    def setGrid(self, kind, origin, shape, cell, scalar):
        o = np.ascontiguousarray(origin, np.float64)
        sh = np.ascontiguousarray(shape, np.int64)
        c = np.ascontiguousarray(cell, np.float64)
        v = None if scalar is None else np.ascontiguousarray(scalar, np.float32).reshape(-1)
        if v is not None and v.size != int(np.prod(sh)):
            raise ValueError("scalar size does not match the grid shape")
        check(self._L.bs_set_grid(self._h, kind, ptr(o, ctypes.c_double), ptr(sh, ctypes.c_int64),
                                  ptr(c, ctypes.c_double), ptr(v, ctypes.c_float)), self._h, "bs_set_grid")

    def setBurying(self, burying):
        b = None if burying is None else f32(burying)
        check(self._L.bs_set_burying(self._h, ptr(b, ctypes.c_float)), self._h, "bs_set_burying")

    # -- insertion vector (src/InsertionVector.cpp:18-78) --
    def getInsertionVector(self, i1, i2):
        """angle (deg), roll angle (deg), insertion depth, vector (3)."""
        out = np.zeros(6)
        check(self._L.bs_get_insertion_vector(self._h, int(i1), int(i2), ptr(out, ctypes.c_double)), self._h,
              "bs_get_insertion_vector")
        return out

    # -- batches of independent systems --
    def setSystems(self, offsets):
        """Particles [offsets[s], offsets[s+1]) form independent system s."""
        o = np.ascontiguousarray(offsets, np.int64)
        check(self._L.bs_set_systems(self._h, o.size - 1, ptr(o, ctypes.c_int64)), self._h, "bs_set_systems")

    def numSystems(self):
        return int(self._L.bs_num_systems(self._h))

    def getSystemEnergies(self):
        """(nsys, 6) energies of each system after the last step."""
        out = np.zeros(6 * self.numSystems())
        check(self._L.bs_get_system_energies(self._h, ptr(out, ctypes.c_double)), self._h,
              "bs_get_system_energies")
        return out.reshape(-1, 6)

    def autotune(self, steps=20):
        """Picks the fastest launch configuration on a copy of the system;
        returns (lanes per spring row, graph steps, pair layout)."""
        g, gr, lay = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(self._L.bs_autotune(self._h, int(steps), ctypes.byref(g), ctypes.byref(gr), ctypes.byref(lay)),
              self._h, "bs_autotune")
        return int(g.value), int(gr.value), int(lay.value)

    # -- gradients (no reference counterpart) --
    def adjoint(self, nsteps, target, checkpoint_every=0):
        """Runs nsteps steps (the state advances as with run()) and returns
        (loss, dk, dl0, dq): the loss (1/n) sum |x_T - target|^2 and its FP64
        gradients per spring (stiffness, rest length; bs_set_springs order)
        and per particle charge.  The external force set by setForce applies."""
        y = f32(target).reshape(-1)
        if y.size != 3 * self.n:
            raise ValueError("target must hold 3 coordinates per particle")
        m = int(self._L.bs_num_springs(self._h))
        dk = np.zeros(max(m, 1))
        dl0 = np.zeros(max(m, 1))
        dq = np.zeros(max(self.n, 1))
        loss = ctypes.c_double()
        c = ctypes
        check(self._L.bs_adjoint(self._h, int(nsteps), int(checkpoint_every), ptr(y, c.c_float),
                                 ptr(self._ext, c.c_float), c.byref(loss), ptr(dk, c.c_double),
                                 ptr(dl0, c.c_double), ptr(dq, c.c_double)), self._h, "bs_adjoint")
        return float(loss.value), dk[:m], dl0[:m], dq[:self.n]

    def adjointForward(self, nsteps, checkpoint_every=0):
        """Runs and records nsteps steps (the state advances) for
        adjointBackward.  Returns the record's generation number, which
        changes with every recording (see tape)."""
        check(self._L.bs_adjoint_forward(self._h, int(nsteps), int(checkpoint_every), ptr(self._ext, ctypes.c_float)),
              self._h, "bs_adjoint_forward")
        self._tape = getattr(self, "_tape", 0) + 1
        return self._tape

    @property
    def tape(self):
        """Generation number of the last adjointForward record (0: none)."""
        return getattr(self, "_tape", 0)

    def adjointBackward(self, dl_dx=None, dl_dv=None, noalloc=False):
        """(dk, dl0, dq) of a loss of the recorded final state, given dL/dx_T
        and/or dL/dv_T as (n, 3) arrays.  noalloc: the variant that takes its
        device buffers from the library's fixed array instead of allocating
        (bs_adjoint_backward_noalloc; systems up to its size)."""
        c = ctypes
        gx = None if dl_dx is None else np.ascontiguousarray(dl_dx, np.float64).reshape(-1)
        gv = None if dl_dv is None else np.ascontiguousarray(dl_dv, np.float64).reshape(-1)
        for g in (gx, gv):
            if g is not None and g.size != 3 * self.n:
                raise ValueError("gradients must hold 3 values per particle")
        m = int(self._L.bs_num_springs(self._h))
        dk, dl0, dq = np.zeros(max(m, 1)), np.zeros(max(m, 1)), np.zeros(max(self.n, 1))
        name = "bs_adjoint_backward_noalloc" if noalloc else "bs_adjoint_backward"
        check(getattr(self._L, name)(self._h, ptr(gx, c.c_double), ptr(gv, c.c_double), ptr(dk, c.c_double),
                                     ptr(dl0, c.c_double), ptr(dq, c.c_double)), self._h, name)
        return dk[:m], dl0[:m], dq[:self.n]

    def close(self):
        if self._h:
            self._L.bs_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
