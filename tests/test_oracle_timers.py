"""The oracle's timer channels (the CPU baseline reports them per step): every
channel accumulates over a step, their sum stays within the step's wall time,
and a reset zeroes them."""
import time

import numpy as np

import oracle
from tests import parity


def test_timer_channels_cover_the_step():
    sol = parity.make_oracle("ab3")
    u0 = sol.project(sol.initial_field(seed=3))
    sol.set_field(oracle.KIND_VEL, 0, u0)
    sol.timers(reset=True)
    assert all(v == 0.0 for v in sol.timers().values())
    t0 = time.perf_counter()
    for it in range(3):
        sol.step(it * 1e-3, 1e-3, it % 4, it + 1)
    wall = time.perf_counter() - t0
    t = sol.timers()
    assert set(t) == set(oracle.Solver.TIMER_CHANNELS)
    assert all(v > 0.0 for v in t.values()), t
    assert sum(t.values()) <= wall
    sol.timers(reset=True)
    assert np.allclose(list(sol.timers().values()), 0.0)
    sol.close()
