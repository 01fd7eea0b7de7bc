"""GPU engine against the golden vectors the reference wrote (tests/golden/).

Per fixture: reference cell lists and processed pair sets bit for bit;
per-particle forces within 1e-5 of the FP64 oracle; energies within 1e-5 of
the FP64 oracle and no further from the reference's own float energies than
those are from FP64.  The fixtures cover every steric mode, constant
dielectric, static beads, the pseudo-hydrophobic term and IMPALA membranes.
"""
import sys

import numpy as np
import pytest

from biospring_amd import _abi
from biospring_amd.network import SpringNetwork
from oracle import pyoracle as po
from tests.parity import ENERGY_RTOL, FORCE_RTOL, energy_error, force_errors
from tests.test_oracle import CASES, load, params_from, prepare

pytestmark = pytest.mark.gpu


def _log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def prepare_net(net, z):
    """The case's rigid bodies, grids and burying on the engine."""
    if "body_ids" in z:
        net.addRigidBody(z["body_ids"])
    if "grid_phi" in z:
        net.setGrid(_abi.GRID_POTENTIAL, z["grid_origin"], z["grid_shape"], z["grid_cell"], z["grid_phi"])
        net.setGrid(_abi.GRID_DENSITY, z["grid_origin"], z["grid_shape"], z["grid_cell"], z["grid_density"])
        net.setBurying(z["burying"])
    return net


def _check(name, s, params, F_ref, E_ref, pairs_ref, z=None):
    net = SpringNetwork.from_system(s, params=params)
    o = po.Oracle(s, params=params)
    if z is not None:
        prepare_net(net, z)
        prepare(o, z)
    F, E = net.computeForces()
    Fd, Ed = o.compute_forces(po.F64)
    dyn = (s.flags & 1) == 0
    per, glob = force_errors(F, Fd, dyn)
    per_ref, _ = force_errors(F_ref, Fd, dyn)
    _log(f"  {name}: force err gpu {per:.3g} (global {glob:.3g}), reference float {per_ref:.3g}")
    assert per <= FORCE_RTOL, (name, per)
    for slot in range(6):
        err = energy_error(E[slot], Ed[slot], floor=1e-3)
        err_ref = energy_error(E_ref[slot], Ed[slot], floor=1e-3)
        assert err <= max(ENERGY_RTOL, err_ref), (name, slot, E[slot], Ed[slot], E_ref[slot])
    for kind, want in pairs_ref.items():
        got = net.getPairs(kind)
        assert got.shape == want.shape and np.array_equal(got, want), (name, kind, got.shape, want.shape)
    return net


@pytest.mark.parametrize("name", CASES)
def test_golden_case(name):
    s, z = load(name)
    _log("case", name, "n", s.n, "m", s.m)
    pairs = {k: z[f"pairs_{k}"] for k in (0, 1, 2) if f"pairs_{k}" in z}
    net = _check(name, s, s.params, z["forces"], z["energies"], pairs, z)
    for kind in (0, 1, 2):
        if f"cells_{kind}_cell_of" not in z:
            continue
        g = net.getCells(kind)
        for key in ("cell_of", "start", "members", "dims", "box"):
            assert np.array_equal(np.asarray(g[key]), z[f"cells_{kind}_{key}"]), (name, kind, key)


def test_golden_steric_variants():
    s, z = load("lj800")
    for tag in sorted({k.split("__")[0] for k in z.files if "__" in k}):
        p = params_from(z[f"{tag}__params"])
        pairs = {k: z[f"{tag}__pairs_{k}"] for k in (0, 1) if f"{tag}__pairs_{k}" in z}
        _log("variant", tag)
        _check(tag, s, p, z[f"{tag}__forces"], z[f"{tag}__energies"], pairs)


@pytest.mark.parametrize("name", ["enm600", "static800", "hydro600", "imp_curved", "rigid300", "rigidall200",
                                  "grid300", "probe600", "probe_amber800"])
def test_golden_trajectory(name):
    """Positions after the fixture's step counts: the GPU is no further from
    the FP64 oracle than the reference's float trajectory is (plus 1e-5 of the
    largest displacement), at every recorded step; rigid-body states and the
    insertion vector likewise."""
    s, z = load(name)
    net = prepare_net(SpringNetwork.from_system(s), z)
    o = prepare(po.Oracle(s), z)
    done = 0
    x0 = s.x.reshape(-1, 3).astype(np.float64)
    for st in z["steps"]:
        st = int(st)
        net.run(st - done)
        o.step(po.F64, st - done)
        done = st
        xg, _ = net.getState()
        xd, _ = o.state(po.F64)
        disp = np.abs(xd - x0).max()
        dg = np.abs(xg - xd).max()
        dr = np.abs(z[f"x_{st}"].astype(np.float64) - xd).max()
        _log(f"  {name} step {st}: gpu-f64 {dg:.3g}  ref-f64 {dr:.3g}  max displacement {disp:.3g}")
        assert dg <= 2 * dr + 1e-5 * max(disp, 1e-3), (name, st, dg, dr)
        if "body_ids" in z:
            bg, bd = net.getRigidBody(0), o.rigid_state(po.F64, 0)
            br = z[f"body_{st}"]
            for sl, rsl in ((slice(0, 3), slice(0, 3)), (slice(9, 12), slice(6, 9))):
                eg = np.abs(bg[sl] - bd[sl]).max()
                er = np.abs(br[rsl] - bd[sl]).max()
                _log(f"    body {sl}: gpu-f64 {eg:.3g} ref-f64 {er:.3g}")
                assert eg <= 2 * er + 1e-6 * max(np.abs(bd[sl]).max(), 1e-9), (name, st, sl, bg, bd)
        if "ivec_ids" in z:
            i1, i2 = (int(t) for t in z["ivec_ids"])
            ig, idd = net.getInsertionVector(i1, i2), o.insertion(po.F64, i1, i2)
            ir = z[f"ivec_{st}"]
            eg, er = np.abs(ig - idd), np.abs(ir - idd)
            _log(f"    ivec gpu-f64 {eg.max():.3g} ref-f64 {er.max():.3g}")
            assert (eg <= 2 * er + 1e-4).all(), (name, st, ig, idd, ir)
        if f"probe_{st}" in z:  # the probe: position, velocity, force of the step
            pg = np.concatenate(net.getProbeParticle()).astype(np.float64)
            pd = np.concatenate(o.probe(po.F64))
            pr = z[f"probe_{st}"].astype(np.float64)
            for sl in (slice(0, 3), slice(3, 6), slice(6, 9)):
                eg, er = np.abs(pg[sl] - pd[sl]).max(), np.abs(pr[sl] - pd[sl]).max()
                _log(f"    probe {sl}: gpu-f64 {eg:.3g} ref-f64 {er:.3g}")
                assert eg <= 2 * er + 1e-5 * max(np.abs(pd[sl]).max(), 1e-9), (name, st, sl, pg, pd, pr)
        assert np.array_equal(net.getState()[0][(s.flags & 1) == 1], x0[(s.flags & 1) == 1].astype(np.float32))
