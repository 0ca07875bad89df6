"""The autotuner's search against a stand-in batch whose step time depends on
the setting: it must keep the fastest setting and run only real steps."""
from cylag_amd.tune import autotune


class FakeBatch:
    def __init__(self, steps, cost):  # cost(sort_interval, blocks_per_cu, inline_normals)
        self.left = steps
        self.cost = cost
        self.tuning = (8, 0, 0)
        self.t = 0.
        self.ran = 0
        self.nscen = 1

    def steps_left(self):
        return [self.left]

    def set_tuning(self, si, b, n):
        self.tuning = (si, b, n)

    def run(self, k):
        k = min(k, self.left)
        self.left -= k
        self.ran += k
        self.t += k * self.cost(*self.tuning)

    def clock(self):
        return self.t


def test_autotune_picks_fastest_and_steps_are_real():
    cost = lambda si, b, n: (1.0 + abs(si - 16) * 0.01 + (0.05 if b == 0 else 0.) + (0.02 if b == 2 else 0.)
                             - (0.03 if n == 1 else 0.01 * n))
    fb = FakeBatch(10_000, cost)
    si, b, n, table = autotune(fb, trial_steps=10, clock=fb.clock)
    assert (si, b, n) == (16, 3, 1) and fb.tuning == (16, 3, 1)
    assert fb.ran == 10 * (1 + 4 + 2 + 2)
    assert min(table, key=table.get) == (16, 3, 1)


def test_autotune_short_job():
    fb = FakeBatch(15, lambda si, b, n: 1.0)
    si, b, n, _ = autotune(fb, trial_steps=10, clock=fb.clock)
    assert fb.left == 0 and fb.tuning == (si, b, n)
