"""Simulation clock (cylag/core/simutime.pyx:25-125)."""
import math

import numpy as np

INT_MAX = 2**31 - 1


class SimuTime:
    def __init__(self, print_progress, initial_time=0., final_time=0., time_step=0.,
                 maxiteration=INT_MAX):
        if final_time <= initial_time:
            raise ValueError("final_time must be greater than initial_time")
        if time_step <= 0.:
            raise ValueError("time_step must be greater than 0")
        if maxiteration <= 0:
            raise ValueError("maxiteration must be greater than 0")
        if final_time - initial_time < time_step:
            raise ValueError("final_time - initial_time must be greater than time_step")
        self.time = float(initial_time)
        self.initial_time = float(initial_time)
        self.time_step = float(time_step)
        self.final_time = float(final_time)
        self.nt = int((final_time - initial_time) / time_step)
        self.times = np.linspace(initial_time, final_time, self.nt + 1)
        self.iteration = 0
        self.maxiteration = maxiteration
        self.is_finished = False
        self.print_progress = print_progress

    def increment(self):
        self.time += self.time_step
        self.iteration += 1
        if self.time >= self.final_time or self.iteration >= self.maxiteration:
            self.is_finished = True

    def convert_time_to_jjhhmmss(self, t):
        return (int(t / 86400.0), int(math.fmod(t, 86400.0) / 3600.0),
                int(math.fmod(t, 3600.0) / 60.0), int(math.fmod(t, 60.0)))

    def progress(self):
        jj, hh, mm, ss = self.convert_time_to_jjhhmmss(self.time)
        duration = self.final_time - self.initial_time
        prog = 1. if duration <= 0. else min(1., max(0., (self.time - self.initial_time) / duration))
        percent, filled = int(prog * 100.), int(prog * 15.)
        bar = "[{}{}{}] {:3d}%".format("=" * (filled - 1), ">" if filled > 0 else "",
                                      "." * (15 - filled), percent)
        print("Iter = {:10d} ; Dt = {:4.4f}s ; Time = {:02d}d {:02d}:{:02d}:{:02d} ; {}".format(
            self.iteration, self.time_step, jj, hh, mm, ss, bar))
