"""Autotuner for a running batch: the element re-sort interval, the resident
blocks per CU of the persistent kernels, and where the 2D fused kernels get
their normals (the pool-order generator, the slot-order generator, or inline
sampling from each slot's prepared stream position).

Each candidate setting is timed over real steps of the job (the results do
not depend on any setting), by coordinate search: the re-sort interval first,
then the source of the normals, then the blocks per CU.  The trial steps are part
of the job, so tuning costs only the steps run under slower settings.
"""
import time


def autotune(batch, sort_intervals=(4, 8, 16, 32), blocks_per_cu=(0, 3, 2), inline_normals=(0, 2, 1),
             trial_steps=32, clock=time.perf_counter, log=None):
    """Tune `batch` (a BatchSolver) in place; returns (sort_interval,
    blocks_per_cu, inline_normals, {setting: seconds per step})."""
    def left():
        return max(batch.steps_left()) if batch.nscen else 0

    table = {}

    def trial(si, bpc, inl):
        if (si, bpc, inl) in table:
            return table[(si, bpc, inl)]
        k = min(trial_steps, left())
        if k <= 0:
            return float("inf")
        batch.set_tuning(si, bpc, inl)
        t0 = clock()
        batch.run(k)
        dt = (clock() - t0) / k
        table[(si, bpc, inl)] = dt
        if log:
            log("autotune: sort every {} steps, {} blocks/CU, {} normals: {:.3f} ms/step".format(
                si, bpc or "max", ("pool-order generated", "inline", "slot-order generated")[inl], dt * 1e3))
        return dt

    cur = tuple(getattr(batch, "tuning", (sort_intervals[0], blocks_per_cu[0], inline_normals[0])))
    trial(*cur)  # warm-up under the current setting (not compared)
    table.clear()
    b0, n0 = blocks_per_cu[0], inline_normals[0]
    best_si = min(sort_intervals, key=lambda si: trial(si, b0, n0))
    best_n = min(inline_normals, key=lambda n: trial(best_si, b0, n))
    best_b = min(blocks_per_cu, key=lambda b: trial(best_si, b, best_n))
    batch.set_tuning(best_si, best_b, best_n)
    return best_si, best_b, best_n, table
