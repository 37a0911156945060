"""The CPU build of the engine's tick against the oracle on samples of the
three benchmark workloads (workload_harness): 256 arenas of the random-state
1v1 workload (4096 arenas, seed 1), 128 of the mixed 1v1/2v2/3v3 workload
(16384 arenas, seed 2) and 256 of the contact-heavy 1v1 workload (65536
arenas, seed 4), each arena under its own uniform default-table actions.

Gates:
  - tick 0 (one tick from the workload state): every arena within the parity
    bar, no exemption;
  - ticks 0-7 teacher-forced (the oracle re-fed the engine's state every
    tick): every arena-tick within the bar, except where the oracle itself,
    given the same input moved by one float32 ulp, leaves the bar (those and
    the arena's later ticks are exempt, reported, and bounded to 3% of the
    arenas);
  - no tick runs out of contact-manifold storage."""
import pytest

import rsxlib as R
import workload_harness as WH

SAMPLES = [("random", 256), ("mixed", 128), ("stress", 256)]
TICKS = 8


@pytest.mark.parametrize("kind,count", SAMPLES)
def test_workload_sample_matches_oracle(kind, count, oracle, cpucheck):
    lib, world = cpucheck
    lut = WH.lut_controls()
    cases = WH.sample(kind, count)
    engs = WH.run_cpucheck_batch(lib, world, cases, TICKS, lut)
    failures, exempt, worst = WH.teacher_forced_check(oracle, cases, engs, TICKS, lut)
    first = [x for x in exempt if x[1] == 0]
    assert not failures, failures[:5]
    assert not first, f"tick-0 arenas outside the bar: {first[:5]}"
    arenas = {x[0] for x in exempt}
    assert len(arenas) <= 0.03 * count, exempt
    print(f"{kind}: {count} arenas x {TICKS} ticks, worst passing error {worst:.3f} of the bar, "
          f"exempt {len(exempt)} arena-ticks in {len(arenas)} arenas")


def test_contact_storage_suffices_on_the_workloads(cpucheck):
    """No tick of the contact-heavy and mixed samples fills every contact
    manifold the engine stores (MfList::kLive)."""
    import ctypes as C
    import parity_harness as P
    lib, world = cpucheck
    lut = WH.lut_controls()
    for kind, count in (("stress", 128), ("mixed", 64)):
        for case in WH.sample(kind, count):
            h = C.c_void_p(lib.cc_arena_new(world, case["blue"], case["orange"]))
            lib.cc_set_tick(h, case["tick"])
            for i, s in enumerate(case["cars"]):
                lib.cc_set_car(h, i, C.byref(s))
                lib.cc_set_wheels(h, i, P.WHEEL_ZERO6)
            lib.cc_set_ball(h, C.byref(case["ball"]))
            lib.cc_set_pads(h, case["pads"])
            for t in range(24):
                for i, k in enumerate(case["actions"]):
                    lib.cc_set_controls(h, i, C.byref(lut[k]))
                lib.cc_step(h, 1)
            over = lib.cc_contact_overflow(h)
            lib.cc_arena_free(h)
            assert not over, (kind, case["source"])


def test_bvh_traversal_order_is_bullets(oracle, cpucheck):
    """The engine visits a mesh's triangles in the order RocketSim's Bullet
    reports them (btQuantizedBvh::reportAabbOverlappingNodex walks subtree
    headers, btQuantizedBvh.cpp:307-325): for every soccar mesh, sorting the
    triangles by the engine's rank gives exactly the order Bullet's
    processAllTriangles reports over an AABB holding the whole mesh."""
    import ctypes as C
    import numpy as np
    lib, world = cpucheck
    rk = (C.c_int32 * 9000)()
    n = lib.cc_tri_rank(world, rk, 9000)
    rank = np.array(rk[:n])
    info = (C.c_float * (9000 * 14))()
    lib.cc_tri_info(world, info, 9000)
    mesh = np.array(info[:n * 14]).reshape(n, 14)[:, 13].astype(int)
    big = (C.c_float * 6)(-1e6, -1e6, -1e6, 1e6, 1e6, 1e6)
    base = 0
    for m in range(16):
        nt = int((mesh == m).sum())
        out = (C.c_int32 * 2000)()
        k = oracle.ref_mesh_query(m, big, out, 2000)
        assert k == nt
        np.testing.assert_array_equal(np.argsort(rank[base:base + nt]), np.array(out[:k]), err_msg=f"mesh {m}")
        base += nt


def test_oracle_visits_cars_last_added_first(oracle):
    """Arena::Step's loops over Arena::_cars run last-added first in the
    oracle's arenas (the order the engine's tick follows)."""
    import ctypes as C
    for team in (1, 2, 3):
        h = C.c_void_p(oracle.ref_arena_new(team, team))
        buf = (C.c_uint32 * 8)()
        n = oracle.ref_car_order(h, buf, 8)
        oracle.ref_arena_free(h)
        assert list(buf[:n]) == list(range(2 * team, 0, -1))


def _pileup_case():
    """Six cars (3v3) boosting into the orange goal's right corner behind the
    ball: a 3 x 2 block of cars 120 uu apart at 1200 uu/s, yawed at the
    corner, the ball at the corner wall ahead of them."""
    import numpy as np
    import parity_harness as P
    cars = []
    for k in range(6):
        i, j = k % 3, k // 3
        yaw = np.pi / 4 + 0.05 * (k - 2.5)
        cars.append(P.default_car(pos=(3350 + 120 * i + 40 * j, 4500 + 120 * j + 30 * i, 17), yaw=yaw,
                                  vel=(850, 850, 0), boost=100, team=0 if k < 3 else 1))
    pads = R.PADS()
    for p in range(34):
        pads[p].is_active = 1
    return dict(blue=3, orange=3, cars=cars, ball=P.ball((3700, 4900, 93.15), (100, 150, 0)), pads=pads,
                tick=1000, source=-1, actions=np.array([18] * 6, dtype=np.int32))


PILEUP_TICKS = 120


def test_3v3_pileup_in_a_goal_corner(oracle, cpucheck):
    """Six cars and the ball packed into a goal corner: the CPU build follows
    the oracle tick by tick (teacher-forced, same exemption rule), the pile-up
    holds many contact manifolds at once, and no tick runs out of their
    storage."""
    import ctypes as C
    lib, world = cpucheck
    lut = WH.lut_controls()
    case = _pileup_case()
    eng = WH.run_cpucheck(lib, world, case, PILEUP_TICKS, lut)
    failures, exempt, worst = WH.teacher_forced_check(oracle, [case], [eng], PILEUP_TICKS, lut)
    assert not failures, failures[:5]
    assert not [x for x in exempt if x[1] == 0]
    h = C.c_void_p(lib.cc_arena_new(world, 3, 3))
    lib.cc_set_tick(h, case["tick"])
    for i, s in enumerate(case["cars"]):
        lib.cc_set_car(h, i, C.byref(s))
    lib.cc_set_ball(h, C.byref(case["ball"]))
    lib.cc_set_pads(h, case["pads"])
    buf = (C.c_float * 8192)()
    most = 0
    for t in range(PILEUP_TICKS):
        for i in range(6):
            lib.cc_set_controls(h, i, C.byref(lut[18]))
        lib.cc_step(h, 1)
        most = max(most, len(R.parse_contacts(buf, lib.cc_dump_contacts(h, buf, 8192))))
    over = lib.cc_contact_overflow(h)
    lib.cc_arena_free(h)
    assert not over
    assert most >= 6, most  # car-car, car-ball and car-wall manifolds at once
    print(f"pile-up: most manifolds with contacts in a tick {most}, worst passing error {worst:.3f}, "
          f"exempt arena-ticks {len(exempt)} (first {exempt[:1]})")


def test_extra_cases_match_oracle(oracle, cpucheck):
    """Wheel rays on another car and on the ball, the auto-flip, and a
    demolition through its respawn (extra_cases.py), teacher-forced against
    the oracle with the same exemption rule; the respawn is reached."""
    import extra_cases as X
    lib, world = cpucheck
    lut = WH.lut_controls()
    cases = X.cases()
    engs = [WH.run_cpucheck(lib, world, c, c["ticks"], lut) for c in cases]
    failures, exempt, worst, respawned = X.check(oracle, cases, engs, lut)
    print(f"extra cases: worst passing {worst:.3f}, exempt {[(x[0], x[1]) for x in exempt if x[3] != 'after an exempt tick']}, "
          f"respawns {respawned}")
    assert not failures, failures[:5]
    assert [r[0] for r in respawned] == ["demolition respawn"]
    auto = engs[3]
    assert any(c[0].is_auto_flipping for c in auto.cars)


def _run_solver_variant(lib, world, case, ticks, lut, lds):
    """The CPU build on one workload case, the solver iterating on the
    private bodies (lds=0) or on an LDS-like column (lds=1)."""
    import ctypes as C
    import parity_harness as P
    h = C.c_void_p(lib.cc_arena_new(world, case["blue"], case["orange"]))
    lib.cc_set_solver_lds(h, lds)
    lib.cc_set_tick(h, case["tick"])
    for i, s in enumerate(case["cars"]):
        lib.cc_set_car(h, i, C.byref(s))
        lib.cc_set_wheels(h, i, P.WHEEL_ZERO6)
    lib.cc_set_ball(h, C.byref(case["ball"]))
    lib.cc_set_pads(h, case["pads"])
    out = []
    for t in range(ticks):
        for i, c in enumerate(WH.case_controls(case, t, lut)):
            lib.cc_set_controls(h, i, C.byref(c))
        lib.cc_step(h, 1)
        for i in range(case["blue"] + case["orange"]):
            s = R.CarState()
            lib.cc_get_car(h, i, C.byref(s))
            out.append(bytes(s))
        b = R.BallState()
        lib.cc_get_ball(h, C.byref(b))
        out.append(bytes(b))
    lib.cc_arena_free(h)
    return out


def test_solver_on_an_lds_column_is_bit_identical(cpucheck):
    """The GPU kernels run the solver's iterations on the bodies' hot state in
    the wavefront's LDS (solver_iterate_lds); the same code on the CPU build,
    over contact-heavy and mixed workload samples, gives every car and ball
    bit for bit what the iterations on the private bodies give."""
    lib, world = cpucheck
    lut = WH.lut_controls()
    for kind, count in (("stress", 96), ("mixed", 48)):
        for k, case in enumerate(WH.sample(kind, count)):
            a = _run_solver_variant(lib, world, case, 24, lut, 0)
            b = _run_solver_variant(lib, world, case, 24, lut, 1)
            assert a == b, (kind, k)
