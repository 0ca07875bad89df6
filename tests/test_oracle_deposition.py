"""Deposition on the bed (bottom_deposition), an extension with no counterpart
in the reference: the oracle restates the reference's step, whose bottom is a
damped reflection (compute_bc.pyx:294-332), and adds the deposition branch in
its place when the parameter is on.  Pinned here by the closed form of
particles sinking at a constant settling velocity through still water over a
flat bed: a particle at z0 reaches z0 - k ws dt after k steps and deposits at
the first step that takes it below the bed, at (x0, y0, -h)."""
import numpy as np

from oracle import refpy as R
from tests.cases import still_column_case

H, WS, DT = 10., 1e-3, 100.
PARAMS = dict(model=1, time_scheme=1, diffusion_model=0, time_step=DT, final_time=DT * 120,
              frozen_eulerian_fields=True, bottom_deposition=True)


def expected_deposits(pos):
    """Iteration (1-based) and position of each particle's deposit; rows whose
    crossing lies within 1e-6 of a step boundary are flagged (rounding could
    move them by one step)."""
    steps = (pos[:, 2] + H) / (WS * DT)
    k = np.floor(steps).astype(int) + 1
    tie = np.abs(steps - np.round(steps)) < 1e-6
    return k, tie


def test_closed_form_deposition():
    case = still_column_case(h=H, ws=WS)
    r = case.run_oracle(R, PARAMS)
    ps = r.solver.p
    n = ps.s.deposited_count
    assert n == len(case.position), "every particle reaches the bed"
    assert np.all(r.inactive[:n] == 1)
    k, tie = expected_deposits(case.position)
    # records in the serial order of the loop: iterations in turn, rows in pool order
    order = np.lexsort((np.arange(len(k)), k))
    got_it = ps.deposited_iteration[:n]
    got_xyz = ps.deposited[:n]
    assert np.all(np.diff(got_it) >= 0)
    ok = ~tie[order]
    assert np.array_equal(got_it[ok], k[order][ok])
    assert np.allclose(got_xyz[:, :2], case.position[order, :2], rtol=0, atol=1e-12)
    assert np.allclose(got_xyz[:, 2], -H, rtol=0, atol=1e-12)


def test_deposition_off_keeps_the_reflection():
    """Off (the default): no record, the particles stay in the column above the bed."""
    case = still_column_case(h=H, ws=WS)
    p = dict(PARAMS, bottom_deposition=False)
    r = case.run_oracle(R, p)
    assert r.solver.p.s.deposited_count == 0
    assert np.all(r.inactive[:len(case.position)] == 0)
    assert np.all(r.position[:len(case.position), 2] >= -H - 1e-12)
