"""The engine's tick code compiled for the host (tests/cpu_check) against the
oracle, tick by tick, on every shared case: positions, velocities and angular
velocities per vector within 1e-3 + 1e-4*|reference|, rotation entries within
1e-4, discrete state (flags, wheel contacts, pads, goal) exact."""
import pytest

import parity_harness as P
import scenarios as S

CASES = S.cases()


@pytest.mark.parametrize("name", list(CASES))
def test_cpu_build_matches_oracle(name, oracle, cpucheck):
    lib, w = cpucheck
    case = CASES[name]
    ref = P.run_oracle(oracle, case)
    eng = P.run_cpucheck(lib, w, case)
    for t in range(case["ticks"]):
        fails = P.compare_tick(ref, eng, t)
        assert not fails, f"{name}: tick {t}: {fails}"


@pytest.mark.parametrize("name", ["ball back wall", "car hits ball", "ball ramp roll"])
def test_cpu_build_contact_points_match_oracle(name, oracle, cpucheck):
    """Manifolds (unordered body pair, point count) equal the oracle's, tick by tick."""
    lib, w = cpucheck
    case = CASES[name]
    ref = P.run_oracle(oracle, case)
    eng = P.run_cpucheck(lib, w, case)
    for t in range(case["ticks"]):
        r = sorted((*sorted(m[:2]), len(m[3])) for m in ref.contacts[t] if m[3])
        e = sorted((*sorted(m[:2]), len(m[3])) for m in eng.contacts[t] if m[3])
        assert r == e, f"{name}: tick {t}"


def _boundary_cases():
    c = dict(CASES)
    # the blue car (slot 0) rests on the orange car: its wheel rays hit a car
    # whose tick runs after its own, so they read that car's body state as
    # the previous tick left it
    stack = [P.default_car(pos=(0, 0, 75), yaw=0.3, on_ground=False),
             P.default_car(pos=(10, 5, 17), yaw=0.35, team=1)]
    c["car resting on car"] = dict(blue=1, orange=1, cars=stack, ball=P.ball((2000, 2000, 93.15)),
                                   controls=[P.controls(), P.controls()], ticks=48)
    return c


BOUNDARY_CASES = _boundary_cases()


@pytest.mark.parametrize("name", list(BOUNDARY_CASES))
def test_launch_boundaries_are_invisible(name, cpucheck):
    """Reloading the arena from its stored words every 8 ticks, over scribbled
    working state (what the start of each GPU launch does), changes nothing:
    every piece of state a tick reads is stored or rebuilt by the load."""
    lib, w = cpucheck
    case = BOUNDARY_CASES[name]
    a = P.run_cpucheck(lib, w, case)
    for fill in (1, 6):
        b = P.run_cpucheck(lib, w, case, boundary_every=8, fill=fill)
        for t in range(case["ticks"]):
            assert [bytes(c) for c in a.cars[t]] == [bytes(c) for c in b.cars[t]], (fill, t)
            assert bytes(a.balls[t]) == bytes(b.balls[t]), (fill, t)
            assert bytes(a.pads[t]) == bytes(b.pads[t]), (fill, t)
