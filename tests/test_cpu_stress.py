"""Single-tick parity on the contact-heavy workload: every arena of
`workloads.stress_states` (balls at ramps, corners and goal mouths, cars on
walls and the ceiling, overlapping car spawns) is loaded into the oracle and
into the CPU build of the engine's tick, stepped one tick with the same
controls, and compared at the parity bar.  Runs on the CPU only."""
import ctypes as C

import numpy as np
import pytest

import parity_harness as P
import rsxlib as R

E = R.engine
N_ARENAS = 64


class RecordBatch:
    """The record arrays of a 1v1 batch on the host, for the workload
    generators (which read and write whole-batch record arrays)."""

    def __init__(self, n):
        self.num_arenas, self.team_size, self.num_cars = n, [1] * n, 2 * n
        self.balls = (E.BallState * n)()
        for b in self.balls:
            b.pos.z = 93.15
        self.cars = (E.CarState * (2 * n))()
        for c in self.cars:
            c.rot.forward.x = c.rot.right.y = c.rot.up.z = 1
            c.pos.z, c.boost, c.is_on_ground = 17, 33.3, 1
        self.pads = (E.BoostPadState * (34 * n))()

    def GetAllBallStates(self): return self.balls
    def SetAllBallStates(self, s): self.balls = s
    def GetAllCarStates(self): return self.cars
    def SetAllCarStates(self, s): self.cars = s
    def GetAllBoostPadStates(self): return self.pads
    def SetAllBoostPadStates(self, s): self.pads = s


@pytest.fixture(scope="module")
def stress_cases():
    from emy_s_framework_amd import workloads
    b = RecordBatch(N_ARENAS)
    workloads.stress_states(b, 4)
    lut = E.default_action_lut()
    acts = np.random.default_rng(11).integers(0, 90, 2 * N_ARENAS)
    cases = []
    sz = C.sizeof(E.BoostPadState)
    for a in range(N_ARENAS):
        pads = R.PADS.from_buffer_copy(bytes(b.pads)[a * 34 * sz:(a + 1) * 34 * sz])
        cases.append(dict(blue=1, orange=1, cars=[b.cars[2 * a], b.cars[2 * a + 1]], ball=b.balls[a], pads=pads,
                          controls=[lut[acts[2 * a]], lut[acts[2 * a + 1]]], ticks=1, tick=0))
    return cases


def test_single_tick_parity_on_contact_heavy_states(stress_cases, oracle, cpucheck):
    lib, w = cpucheck
    bad = []
    for a, case in enumerate(stress_cases):
        ref = P.run_oracle(oracle, case)
        eng = P.run_cpucheck(lib, w, case)
        fails = P.compare_tick(ref, eng, 0)
        if fails:
            bad.append((a, fails))
    assert not bad, bad[:4]
