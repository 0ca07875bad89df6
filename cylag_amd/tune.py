"""Autotuner for a running batch: the element re-sort interval and the
resident blocks per CU of the persistent kernels.

Each candidate setting is timed over real steps of the job (the results do
not depend on either setting), by coordinate search: the re-sort interval
first, then the blocks per CU at the best interval.  The trial steps are part
of the job, so tuning costs only the steps run under slower settings.
"""
import time


def autotune(batch, sort_intervals=(4, 8, 16, 32), blocks_per_cu=(0, 3, 2), trial_steps=32,
             clock=time.perf_counter, log=None):
    """Tune `batch` (a BatchSolver) in place; returns (sort_interval,
    blocks_per_cu, {setting: seconds per step})."""
    def left():
        return max(batch.steps_left()) if batch.nscen else 0

    table = {}

    def trial(si, bpc):
        k = min(trial_steps, left())
        if k <= 0:
            return float("inf")
        batch.set_tuning(si, bpc)
        t0 = clock()
        batch.run(k)
        dt = (clock() - t0) / k
        table[(si, bpc)] = dt
        if log:
            log("autotune: sort every {} steps, {} blocks/CU: {:.3f} ms/step".format(si, bpc or "max", dt * 1e3))
        return dt

    cur = getattr(batch, "tuning", (sort_intervals[0], blocks_per_cu[0]))
    trial(*cur)  # warm-up under the current setting (not compared)
    table.clear()
    best_si = min(sort_intervals, key=lambda si: trial(si, blocks_per_cu[0]))
    best_b = min(blocks_per_cu, key=lambda b: table.get((best_si, b)) or trial(best_si, b))
    batch.set_tuning(best_si, best_b)
    return best_si, best_b, table
