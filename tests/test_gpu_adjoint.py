"""Adjoint of the step on the GPU (bs_adjoint): gradients of the trajectory
loss (1/n) sum |x_T - y|^2 with respect to spring stiffness, rest length and
charge.

The yardstick is the FP64 oracle's adjoint, itself pinned by central finite
differences of the FP64 loss and by H w against the force derivative.  The
GPU sweeps the FP32 trajectory in FP64, so its gradients differ from the
oracle's by the float trajectory's own error, which each case reports.  Also:
independence of the checkpoint interval, determinism, and that the handle
keeps stepping afterwards.

Central differences of the GPU's own FP32 loss are not used: a change of
loss of 1e-8 relative, which is what a 1e-3 change of one parameter makes
over 20 femtoseconds, is below the rounding noise of a float trajectory; the
differences are taken on the FP64 oracle instead."""
import sys

import numpy as np
import pytest

from biospring_amd import _abi, workloads
from biospring_amd.network import SpringNetwork
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

# max |g_gpu - g_oracle| / max |g_oracle| per gradient
GRAD_RTOL = 2e-4


def _log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _case(name):
    if name == "enm":  # linear steric + constant-dielectric Coulomb, static beads
        s = workloads.generate(0, seed=11, n=500)
        fl = s.flags.copy()
        fl[::9] |= _abi.FLAG_STATIC
        s.flags = fl
        return s, s.params.copy(), 30
    s = workloads.generate(1, seed=12, n=400)
    p = s.params.copy()
    if name == "amber_ddd":
        return s, p, 20
    p.dielectric_model = _abi.DIELECTRIC_CONSTANT
    p.steric_mode = {"lewitt": _abi.STERIC_LEWITT_8_6, "zacharias": _abi.STERIC_ZACHARIAS_8_6}[name]
    return s, p, 6 if name == "zacharias" else 12


def _target(s):
    x0 = s.x.reshape(-1, 3).astype(np.float64)
    return (x0 + 0.5 * np.sin(x0 / 7.0)).astype(np.float32)


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("name,layout", [("amber_ddd", _abi.PAIR_QUEUES), ("enm", _abi.PAIR_QUEUES),
                                         ("lewitt", _abi.PAIR_QUEUES), ("zacharias", _abi.PAIR_QUEUES),
                                         ("amber_ddd", _abi.PAIR_CLUSTERS), ("enm", _abi.PAIR_CLUSTERS)])
def test_adjoint_matches_fp64_oracle(name, layout):
    """Both pair layouts' adjoint kernels (queued and cluster lists)."""
    s, p, T = _case(name)
    p.pair_layout = layout
    y = _target(s)
    _log(f"adjoint {name}: n={s.n} springs={s.m} T={T}")
    net = SpringNetwork.from_system(s, params=p)
    loss, dk, dl, dq = net.adjoint(T, y, checkpoint_every=7)
    x_gpu, _ = net.getState()
    o = po.Oracle(s, params=p)
    lo, dko, dlo, dqo = o.adjoint(T, y.astype(np.float64))
    xo, _ = o.state(po.F64)
    traj = float(np.abs(x_gpu - xo).max())
    errs = {"k": _rel(dk, dko), "l0": _rel(dl, dlo), "q": _rel(dq, dqo)}
    _log(f"  loss {loss:.9g} vs {lo:.9g}; trajectory dev {traj:.2e} A; gradient errors {errs}")
    assert abs(loss - lo) <= 1e-5 * lo
    for g, e in errs.items():
        assert e <= GRAD_RTOL, (g, e)
    # uncharged particles have no charge gradient; static ones keep theirs
    assert np.all(dq[s.charge == 0] == 0)


def test_adjoint_stretched_start_matches_fp64_oracle():
    """3,000 beads started 0.3 A (rms) off their springs' rest positions, so
    that the spring terms carry the motion from the first step.  With this
    much motion, pairs cross the cutoffs during the run, and the gradient of a
    frozen pair set depends on the step at which each crossing falls: the
    oracle's own gradient moves by up to 1e-3 of its largest charge entry when
    its start is perturbed by 1e-5 A, the size of the float trajectory's
    error.  The bar is three times that envelope (and at least GRAD_RTOL)."""
    s = workloads.generate(1, seed=13, n=3000)
    p = s.params.copy()
    x0 = s.x.reshape(-1, 3)
    xs = (x0 + np.random.default_rng(3).normal(0, 0.3, x0.shape)).astype(np.float32)
    y = _target(s)
    T = 20
    net = SpringNetwork.from_system(s, params=p)
    net.setState(xs, None)
    loss, dk, dl, dq = net.adjoint(T, y, checkpoint_every=6)
    o = po.Oracle(s, params=p)
    o.set_state(xs, None)
    lo, dko, dlo, dqo = o.adjoint(T, y.astype(np.float64))
    op = po.Oracle(s, params=p)
    op.set_state_f64(xs.astype(np.float64) + np.random.default_rng(7).normal(0, 1e-5, xs.shape))
    _, dkp, dlp, dqp = op.adjoint(T, y.astype(np.float64))
    errs = {"k": _rel(dk, dko), "l0": _rel(dl, dlo), "q": _rel(dq, dqo)}
    env = {"k": _rel(dkp, dko), "l0": _rel(dlp, dlo), "q": _rel(dqp, dqo)}
    _log(f"  stretched start: loss {loss:.9g} vs {lo:.9g}; gradient errors {errs}; oracle envelope {env}")
    assert abs(loss - lo) <= 1e-5 * lo
    for g, e in errs.items():
        assert e <= max(GRAD_RTOL, 3 * env[g]), (g, e, env[g])


def test_adjoint_checkpoint_interval_and_determinism():
    s = workloads.generate(1, seed=14, n=6000)
    p = s.params.copy()
    y = _target(s)
    T = 24
    x0 = s.x.reshape(-1, 3)
    res = {}
    for K in (24, 1, 5, 0, 5):
        net = SpringNetwork.from_system(s, params=p)
        r = net.adjoint(T, y, checkpoint_every=K)
        if K in res:  # same interval again: bitwise the same
            assert r[0] == res[K][0]
            for a, b in zip(r[1:], res[K][1:]):
                assert np.array_equal(a, b)
        res[K] = r
        x, v = net.getState()
        assert np.isfinite(x).all() and not np.array_equal(x, x0)
        net.run(3)  # the handle keeps stepping after the sweep
        assert np.isfinite(net.getState()[0]).all()
    ref = res[24]
    for K, r in res.items():
        assert abs(r[0] - ref[0]) <= 1e-6 * ref[0]
        for g in range(1, 4):
            assert _rel(r[g], ref[g]) <= 1e-5, (K, g, _rel(r[g], ref[g]))


def test_adjoint_rejects_terms_without_adjoint():
    s = workloads.generate(2, seed=3, n=2000)
    p = workloads.impala_params(s, "flat")
    net = SpringNetwork.from_system(s, params=p)
    with pytest.raises(_abi.EngineError):
        net.adjoint(3, s.x.reshape(-1, 3))


def test_two_phase_adjoint_matches_one_call_and_is_linear_in_the_seed():
    """adjointForward + adjointBackward(dL/dx_T) reproduces bs_adjoint for the
    MSD loss, the sweep is linear in the seed (dL/dx and dL/dv), and the
    record serves several seeds."""
    s = workloads.generate(1, seed=15, n=3000)
    y = _target(s)
    T = 16
    a = SpringNetwork.from_system(s)
    loss, dk, dl, dq = a.adjoint(T, y, checkpoint_every=5)
    b = SpringNetwork.from_system(s)
    b.adjointForward(T, 5)
    x, _ = b.getState()
    g = 2.0 * (x.astype(np.float64) - y) / s.n
    dk2, dl2, dq2 = b.adjointBackward(g)
    for u, v in ((dk, dk2), (dl, dl2), (dq, dq2)):
        assert _rel(v, u) <= 1e-6
    rng = np.random.default_rng(4)
    g1, g2 = rng.normal(size=(s.n, 3)), rng.normal(size=(s.n, 3))
    v1 = rng.normal(size=(s.n, 3))
    r1 = b.adjointBackward(g1)
    r2 = b.adjointBackward(g2)
    r12 = b.adjointBackward(2.0 * g1 - 0.5 * g2)
    rv = b.adjointBackward(None, v1)
    rxv = b.adjointBackward(g1, v1)
    for i in range(3):
        assert _rel(r12[i], 2.0 * r1[i] - 0.5 * r2[i]) <= 1e-5
        assert _rel(rxv[i], r1[i] + rv[i]) <= 1e-5
    # a sweep does not depend on what earlier sweeps left (lists, buffers)
    r1b = b.adjointBackward(g1)
    for i in range(3):
        assert np.array_equal(r1b[i], r1[i]), i
    b.run(1)
    with pytest.raises(_abi.EngineError):
        b.adjointBackward(g1)


def test_torch_autograd_wrapper_gives_the_adjoint_gradients():
    torch = pytest.importorskip("torch")
    from biospring_amd.autograd import final_positions

    s = workloads.generate(1, seed=16, n=2000)
    y = _target(s)
    T = 12
    ref = SpringNetwork.from_system(s).adjoint(T, y, checkpoint_every=4)
    net = SpringNetwork.from_system(s)
    k = torch.tensor(s.k.astype(np.float64), requires_grad=True)
    q = torch.tensor(s.charge.astype(np.float64), requires_grad=True)
    x = final_positions(net, s, k=k, q=q, nsteps=T, checkpoint_every=4)
    loss = ((x - torch.from_numpy(y.astype(np.float64))) ** 2).sum(dim=1).mean()
    loss.backward()
    assert abs(loss.item() - ref[0]) <= 1e-9 * ref[0]
    assert _rel(k.grad.numpy(), ref[1]) <= 1e-6
    assert _rel(q.grad.numpy(), ref[3]) <= 1e-6


def test_torch_wrapper_refuses_a_replaced_record():
    """The handle holds one forward record: the backward of a pass whose
    record a later forward replaced raises instead of returning the later
    pass's gradients; charges are not reloaded on a network with rigid
    bodies (setParticles would drop them)."""
    torch = pytest.importorskip("torch")
    from biospring_amd.autograd import final_positions

    s = workloads.generate(1, seed=17, n=1000)
    y = torch.from_numpy(_target(s).astype(np.float64))
    net = SpringNetwork.from_system(s)
    k = torch.tensor(s.k.astype(np.float64), requires_grad=True)
    x1 = final_positions(net, s, k=k, nsteps=4)
    x2 = final_positions(net, s, k=k, nsteps=4)
    with pytest.raises(RuntimeError):
        ((x1 - y) ** 2).sum().backward()
    ((x2 - y) ** 2).sum().backward()
    assert torch.isfinite(k.grad).all()
    net.addRigidBody(np.arange(10))
    q = torch.tensor(s.charge.astype(np.float64), requires_grad=True)
    with pytest.raises(ValueError):
        final_positions(net, s, q=q, nsteps=2)
