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
