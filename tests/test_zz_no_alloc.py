"""cyl_batch_collect_noalloc (the allocation-free collect) through the C ABI
against cyl_batch_collect on the same device state: two identical batches
are stepped, each collect is called on one of them, and rows, flags, ids and
the deletion order must be identical; two calls in a row, the second on a
smaller scenario, so that the static staging array's previous contents must
not leak into the second result."""
import ctypes as C

import numpy as np
import pytest

import cylag_amd as cy
from cylag_amd import _hip
from tests.cases import open_channel_case, settling_3d_case
from tests.test_gpu_batch import pick, streams

pytestmark = pytest.mark.gpu

MAX_ROWS = 524288   # CYL_NOALLOC_MAX_ROWS in include/cylag_hip.h


def raw_collect(b, s, allocate):
    """One collect call through the C ABI; returns every output array."""
    ps = b.particle_sets[s]
    n, dim = ps.npart, b.dim
    out = dict(position=np.full((n, dim), 7.), velocity=np.full((n, dim), 7.), us=np.full((n, dim), 7.),
               inactive=np.full(n, 7, np.int32), tri=np.full(n, 7, np.int32), ll=np.full(n, 7, np.int32),
               deleted=np.full(n, -7, np.int32))
    cnt = C.c_int(-1)
    P = _hip.ptr
    fn = b.ctx.lib.cyl_batch_collect if allocate else b.ctx.lib.cyl_batch_collect_noalloc
    rc = fn(b.ctx.h, s, n, P(out["position"]), P(out["velocity"]), P(out["us"]), P(out["inactive"], "i"),
            P(out["tri"], "i"), P(out["ll"], "i"), P(out["deleted"], "i"), n, C.byref(cnt))
    b.ctx.check(rc, "collect")
    out["ndeleted"] = cnt.value
    out["deleted"] = out["deleted"][:cnt.value].copy()
    return out


def assert_same(a, b):
    assert a["ndeleted"] == b["ndeleted"]
    for k in ("position", "velocity", "us", "inactive", "tri", "ll", "deleted"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def channel_batch():
    case = open_channel_case(npart=1200)
    base = dict(final_time=1e9, time_step=1., frozen_eulerian_fields=True, model=1, diffusion_model=1)
    mix = [(dict(horizontal_diffusivity=2.), 500, 60), (dict(horizontal_diffusivity=1.), 300, 45),
           (dict(time_scheme=2, horizontal_diffusivity=0.5), 380, 30)]
    mesh = cy.TriangularMesh(case.x, case.y, case.tri, boundary_points=case.boundary_points, nopen=case.nopen)
    fs = cy.EulerianFieldSet(mesh, trifinder=case.trifinder(), **case.fields)
    psets = [cy.LagrangianParticleSet(pick(case, n, 600 + i), dim=2) for i, (_, n, _) in enumerate(mix)]
    b = cy.BatchSolver(fs, psets, [dict(base, **p) for p, _, _ in mix], streams(len(mix)),
                       nsteps=[k for _, _, k in mix], sort_interval=4)
    b.run()
    return b


def settling_batch():
    case = settling_3d_case(npart=2400)
    lsm2 = dict(model=2, time_scheme=5, diffusion_lsm2_option=1, particle_velocity_init=1)
    base = dict(final_time=1e9, time_step=10., frozen_eulerian_fields=True, diffusion_model=3)
    mix = [(dict(model=1), 1400, 20), (lsm2, 900, 25)]
    mesh = cy.TriangularMesh(case.x, case.y, case.tri)
    fs = cy.EulerianFieldSet(mesh, trifinder=case.trifinder(), **case.fields)
    psets, lo = [], 0
    for p, n, _ in mix:
        ps = cy.LagrangianParticleSet(case.position[lo:lo + n].copy(), dim=3, **case.pset_kw)
        ps.diameter = case.diameter[lo:lo + n].copy()
        ps.volume = np.pi * ps.diameter ** 3 / 6.
        psets.append(ps)
        lo += n
    b = cy.BatchSolver(fs, psets, [dict(base, **p) for p, _, _ in mix], streams(len(mix)),
                       nsteps=[k for _, _, k in mix], sort_interval=4)
    b.run()
    return b


@pytest.mark.parametrize("make,order", [(channel_batch, [0, 1, 0]), (settling_batch, [0, 1])],
                         ids=["2d_channel_deletions", "3d_settling"])
def test_noalloc_collect_equals_collect(make, order):
    a, b = make(), make()
    try:
        for s in order:   # a larger scenario, then a smaller one (and scenario 0 again: deletions reported once)
            got, ref = raw_collect(a, s, allocate=False), raw_collect(b, s, allocate=True)
            assert_same(got, ref)
        if make is channel_batch:
            assert raw_collect(b, 0, True)["ndeleted"] == 0
            # the first report of scenario 1 held deletions, in step order
            a2, b2 = make(), make()
            got, ref = raw_collect(a2, 1, False), raw_collect(b2, 1, True)
            assert got["ndeleted"] > 10
            assert_same(got, ref)
            a2.close()
            b2.close()
    finally:
        a.close()
        b.close()


def test_noalloc_collect_through_batch_solver():
    """BatchSolver.collect(allocate=False) gives the particle sets (tags
    renumbered in deletion order) that collect() gives."""
    a, b = channel_batch(), channel_batch()
    try:
        for s in (0, 2, 1):
            pa, pb = a.collect(s, allocate=False), b.collect(s)
            for k in ("position", "velocity", "inactive", "tri", "tag"):
                np.testing.assert_array_equal(getattr(pa, k), getattr(pb, k), err_msg=k)
            assert pa.maxtag == pb.maxtag and pa.npart_active == pb.npart_active
    finally:
        a.close()
        b.close()


def test_noalloc_collect_rejects_oversized_scenario():
    case = open_channel_case(npart=10)
    mesh = cy.TriangularMesh(case.x, case.y, case.tri, boundary_points=case.boundary_points, nopen=case.nopen)
    fs = cy.EulerianFieldSet(mesh, **case.fields)
    big = cy.LagrangianParticleSet(np.full((MAX_ROWS + 1, 2), 50.3), dim=2)
    small = cy.LagrangianParticleSet(np.full((100, 2), 40.1), dim=2)
    params = dict(final_time=1e9, time_step=1., frozen_eulerian_fields=True, model=1, diffusion_model=0)
    b = cy.BatchSolver(fs, [big, small], [params, params], streams(2), nsteps=[2, 3])
    try:
        b.run()
        with pytest.raises(_hip.HipError, match="CYL_NOALLOC_MAX_ROWS"):
            raw_collect(b, 0, allocate=False)
        ref = raw_collect(b, 0, allocate=True)
        assert ref["inactive"].sum() == 0
        assert_same(raw_collect(b, 1, allocate=False), raw_collect(b, 1, allocate=True))
    finally:
        b.close()
