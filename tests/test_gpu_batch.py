"""Batches: many solves of one kind in one launch per RK stage.  Every member
must compute exactly what it computes alone (same operations, same order):
states and monitors are compared bit for bit, alone against in the mix, also
when members join and leave while others run."""
import numpy as np
import pytest

from fvm2d_amd import mesh
from fvm2d_amd.gpu import FvmError
from fvm2d_amd.solver import Batch, Inputs, Solver
from tests.helpers import case_kwargs, grid

pytestmark = pytest.mark.gpu

EULER = dict(iflow=1, iflux=1, ilimit=1, cfl=1.0, maxiter=10**6, minres=1e-12)


def euler_members():
    """Three different meshes, Mach numbers and angles."""
    return [
        (mesh.o_grid(65, 17, t=0.12, first_spacing=4e-3, seed=3), dict(mach_inf=0.5, aoa_deg=2.0)),
        (mesh.o_grid(129, 33, m=0.02, p=0.4, t=0.15, seed=5), dict(mach_inf=0.8, aoa_deg=1.25)),
        (mesh.o_grid(97, 25, m=0.04, p=0.3, t=0.09, jitter=0.1, seed=7), dict(mach_inf=0.3, aoa_deg=-3.0)),
    ]


def alone(g, kw, n):
    with Solver(g, Inputs.from_kwargs(**kw)) as s:
        h, _ = s.iterate(n)
        return h, s.get_state(), s.monitor()[0]


def test_batch_same_as_alone():
    n = 48
    members = euler_members()
    ref = [alone(g, dict(EULER, **kw), n) for g, kw in members]
    solvers = [Solver(g, Inputs.from_kwargs(**dict(EULER, **kw))) for g, kw in members]
    with Batch() as b:
        for s in solvers:
            b.add(s)
        st = b.iterate(n)
        assert st == [0, 0, 0]
        for s, (h, state, mon) in zip(solvers, ref):
            got = s.get_state()
            np.testing.assert_array_equal(got["prim"], state["prim"])
            m, status = s.monitor()
            assert status == 0 and m["iter"] == n
            assert (m["fres"], m["cl"], m["cd"], m["iresi"]) == (h[-1, 1], h[-1, 4], h[-1, 5], int(h[-1, 2]))
    for s in solvers:
        s.close()


def test_batch_members_join_and_leave():
    """Continuous batching: B joins while A runs, then A leaves."""
    (ga, ka), (gb, kb), _ = euler_members()
    ka, kb = dict(EULER, **ka), dict(EULER, **kb)
    a_ref = alone(ga, ka, 40)[1]["prim"]
    b_ref = alone(gb, kb, 30)[1]["prim"]
    with Solver(ga, Inputs.from_kwargs(**ka)) as a, Solver(gb, Inputs.from_kwargs(**kb)) as bb, Batch() as b:
        b.add(a)
        b.iterate(16)
        b.add(bb)
        b.iterate(24)                   # A at 40, B at 24
        np.testing.assert_array_equal(a.get_state()["prim"], a_ref)
        b.remove(a)
        b.iterate(6)                    # B at 30
        np.testing.assert_array_equal(bb.get_state()["prim"], b_ref)
        assert a.monitor()[0]["iter"] == 40 and bb.monitor()[0]["iter"] == 30


def test_batch_members_stop_on_their_own():
    """A member that reaches MinRes stops where it stops alone; the others go on."""
    members = euler_members()[:2]
    kw0 = dict(EULER, **members[0][1])
    h0, _, _ = alone(members[0][0], kw0, 60)
    kw0["minres"] = float(h0[30, 1])          # reached at iteration 30
    h_stop, st_stop, _ = alone(members[0][0], kw0, 60)
    s0 = Solver(members[0][0], Inputs.from_kwargs(**kw0))
    s1 = Solver(members[1][0], Inputs.from_kwargs(**dict(EULER, **members[1][1])))
    with Batch() as b:
        b.add(s0)
        b.add(s1)
        st = b.iterate(60)
        assert st[0] == 1 and st[1] == 0          # converged / still running
        assert s0.monitor()[0]["iter"] == len(h_stop)
        assert s1.monitor()[0]["iter"] == 60
        np.testing.assert_array_equal(s0.get_state()["prim"], st_stop["prim"])
    s0.close()
    s1.close()


def test_batch_rejects_other_kind():
    g, kw = euler_members()[0]
    with Solver(g, Inputs.from_kwargs(**dict(EULER, **kw))) as a, \
            Solver(g, Inputs.from_kwargs(**dict(EULER, iflux=2, **kw))) as k, Batch() as b:
        b.add(a)
        with pytest.raises(FvmError):
            b.add(k)
        with pytest.raises(FvmError):
            b.add(a)                                   # already a member


def test_batch_members_share_the_workgroup_size():
    """Members run the same kernel instances: a 512-thread context does not
    join a batch of 256-thread ones, and the workgroup size the library
    reports is the one asked for, or its automatic choice (256 below 32768
    vertices)."""
    g, kw = euler_members()[0]
    with Solver(g, Inputs.from_kwargs(**dict(EULER, **kw))) as a, \
            Solver(g, Inputs.from_kwargs(**dict(EULER, block_threads=512, **kw))) as w, Batch() as b:
        assert a.layout_info()["threads"] == 256 and w.layout_info()["threads"] == 512
        b.add(a)
        with pytest.raises(FvmError):
            b.add(w)


@pytest.mark.parametrize("name,precision", [("rans_roe_97x25", 0), ("laminar_roe_97x25", 1)])
def test_batch_viscous_same_as_alone(name, precision):
    """SA (time step, SA step and copy as batched vertex kernels) and the
    split-precision laminar path, two angles each."""
    g = grid(name)
    n = 20
    kws = [case_kwargs(name, precision=precision, aoa_deg=a) for a in (0.0, 4.0)]
    ref = [alone(g, kw, n)[1] for kw in kws]
    solvers = [Solver(g, Inputs.from_kwargs(**kw)) for kw in kws]
    with Batch() as b:
        for s in solvers:
            b.add(s)
        b.iterate(n)
        for s, r in zip(solvers, ref):
            got = s.get_state()
            np.testing.assert_array_equal(got["prim"], r["prim"])
            np.testing.assert_array_equal(got["nut"], r["nut"])
    for s in solvers:
        s.close()


def test_batch_member_error_is_isolated():
    """A member that hits a reference STOP condition (here a negative
    pressure put into its state) stops with the status and index it gets
    alone; the other members run on, bit for bit as alone."""
    (ga, ka), (gb, kb), _ = euler_members()
    ka, kb = dict(EULER, **ka), dict(EULER, **kb)
    with Solver(ga, Inputs.from_kwargs(**ka)) as a:
        prim = a.get_state()["prim"].copy()
    prim[4 * 400 + 3] = -50.0                   # pressure of vertex 401
    alone_err = None
    with Solver(ga, Inputs.from_kwargs(**ka)) as a:
        a.set_state(prim)
        with pytest.raises(FvmError) as ei:
            a.iterate(20)
        alone_err = (ei.value.code, ei.value.index, a.monitor()[0]["iter"])
    b_ref = alone(gb, kb, 20)[1]["prim"]
    with Solver(ga, Inputs.from_kwargs(**ka)) as a, Solver(gb, Inputs.from_kwargs(**kb)) as b, \
            Batch() as batch:
        a.set_state(prim)
        batch.add(a)
        batch.add(b)
        st = batch.iterate(20)
        assert st[0] == alone_err[0] < 0 and st[1] == 0
        assert a.monitor()[0]["iter"] == alone_err[2]
        np.testing.assert_array_equal(b.get_state()["prim"], b_ref)
