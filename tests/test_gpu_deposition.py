"""Deposition on the bed (bottom_deposition, an extension with no counterpart in
the reference, whose bottom is a damped reflection, compute_bc.pyx:294-332) on
the HIP step, against the oracle's restatement of it: the closed form of
particles sinking through still water, and the stochastic settling case
(k-epsilon K(z), per-row settling diameters) through the fused and the
generic kernels.  Identical flags, element ids, tags and stream state,
positions within 1e-9 relative, and the same deposit records (position on the
bed and iteration) in the same order."""
import numpy as np
import pytest

import cylag_amd as cy
from oracle import refpy as R
from tests.cases import settling_3d_case, still_column_case
from tests.test_gpu_parity import assert_parity
from tests.test_oracle_deposition import DT, H, PARAMS, WS, expected_deposits

pytestmark = pytest.mark.gpu


def deposits_gpu(g):
    xyz, it = g.solver.deposited
    return xyz.copy(), it.copy()


def deposits_oracle(r):
    p = r.solver.p
    n = p.s.deposited_count
    return p.deposited[:n].copy(), p.deposited_iteration[:n].copy()


def assert_same_deposits(g, r):
    gx, gi = deposits_gpu(g)
    rx, ri = deposits_oracle(r)
    assert gi.size == ri.size, "deposit counts differ: {} vs {}".format(gi.size, ri.size)
    assert np.array_equal(gi, ri), "deposit iterations differ"
    scale = np.maximum(np.abs(rx), 1.)
    assert np.max(np.abs(gx - rx) / scale, initial=0.) <= 1e-9


def test_closed_form_deposition_fused_3d(tmp_path):
    case = still_column_case(h=H, ws=WS)
    params = dict(PARAMS, output_rep=str(tmp_path))
    g = case.run_gpu(cy, params)
    r = case.run_oracle(R, params)
    assert g.solver.ctx.last_path()[0] > 0, "the fused 3D kernel runs this configuration"
    assert_parity(g, r)
    assert_same_deposits(g, r)
    gx, gi = deposits_gpu(g)
    k, tie = expected_deposits(case.position)
    order = np.lexsort((np.arange(len(k)), k))
    assert gi.size == len(case.position)
    ok = ~tie[order]
    assert np.array_equal(gi[ok], k[order][ok])
    assert np.allclose(gx[:, 2], -H, rtol=0, atol=1e-12)


CONFIGS = {
    "lsm1_fused": dict(model=1, diffusion_model=3),
    "lsm2_di1_fused": dict(model=2, time_scheme=5, diffusion_model=3, diffusion_lsm2_option=1,
                           particle_velocity_init=1),
    "lsm1_rk2_generic": dict(model=1, time_scheme=2, diffusion_model=1, horizontal_diffusivity=0.1,
                             vertical_diffusivity=1e-3),
}


@pytest.mark.parametrize("rng_method", [1, 2])
@pytest.mark.parametrize("which", list(CONFIGS))
def test_settling_with_deposition(which, rng_method, tmp_path):
    case = settling_3d_case(seed=41)
    params = dict(final_time=400., time_step=5., rng_method=rng_method, bottom_deposition=True,
                  output_rep=str(tmp_path), **CONFIGS[which])
    g = case.run_gpu(cy, params)
    r = case.run_oracle(R, params)
    fused, generic = g.solver.ctx.last_path()
    assert (fused > 0) == which.endswith("fused") and (generic > 0) == which.endswith("generic"), (fused, generic)
    assert deposits_oracle(r)[1].size > 20, "the case deposits particles"
    assert_parity(g, r)
    assert_same_deposits(g, r)


def test_deposit_file_written_by_flush(tmp_path):
    """solve() ends with a flush, which writes the records since the last flush."""
    case = still_column_case(h=H, ws=WS, npart=100)
    params = dict(PARAMS, output_rep=str(tmp_path))
    g = case.run_gpu(cy, params)
    gx, gi = deposits_gpu(g)
    rows = np.loadtxt(tmp_path / "deposited_particles.txt", delimiter=',', skiprows=2, ndmin=2)
    assert rows.shape == (gi.size, 4)
    assert np.array_equal(rows[:, :3], gx)
    assert np.array_equal(rows[:, 3].astype(int), gi)
    assert g.solver.particle_set.deposited_count == 0  # reset by the flush
