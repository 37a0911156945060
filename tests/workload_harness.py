"""Samples of the benchmark workloads as one-arena oracle cases, and runners
that step the same sample on the oracle (reference CPU RocketSim), on the CPU
build of the engine's tick and on the GPU engine.

Workloads (emy-s-framework_amd/workloads.py), generated on the host at their
full size so the sample is exactly what the benchmark runs:
  random  4096 1v1 arenas, random_field_states(seed 1)
  mixed   16384 arenas, mixed_team_sizes(seed 2) + random_field_states(seed 2)
  stress  65536 1v1 arenas, stress_states(seed 4) (contact-heavy)
A sample takes every k-th arena.

Every run first steps the loaded state one tick and then loads it again, on
both sides: a world step registers each dynamic body in the broadphase
(btRSBroadphase::setAabb) where a suspension ray of the next tick looks for
it, and a state set without a step leaves the registration where the last
step put it (RocketSim and the engine alike).  The warm tick makes the
registrations, and the per-wheel values a tick carries to the next, those of
a running simulation instead of whatever the arena held before."""
import ctypes as C

import numpy as np

import parity_harness as P
import rsxlib as R

E = R.engine

SPECS = {
    "random": dict(n=4096, seed=1),
    "mixed": dict(n=16384, seed=2),
    "stress": dict(n=65536, seed=4),
}


class RecordBatch:
    """Whole-batch record arrays on the host, with the accessors the workload
    generators use."""

    def __init__(self, team_sizes):
        n = len(team_sizes)
        self.num_arenas, self.team_size = n, list(team_sizes)
        self.num_cars = 2 * sum(team_sizes)
        self.car_off = np.concatenate([[0], np.cumsum([2 * t for t in team_sizes])[:-1]]).astype(int)
        self.balls = (E.BallState * n)()
        for b in self.balls:
            b.pos.z = 93.15
        self.cars = (E.CarState * self.num_cars)()
        for a in range(n):
            t = team_sizes[a]
            for i in range(2 * t):
                c = self.cars[self.car_off[a] + i]
                c.rot.forward.x = c.rot.right.y = c.rot.up.z = 1
                c.pos.z, c.boost, c.is_on_ground = 17, 33.3, 1
                c.team = 0 if i < t else 1
                c.ball_hit_tick = c.ball_hit_extra_impulse_tick = 2**64 - 1
        self.pads = (E.BoostPadState * (34 * n))()

    def GetAllBallStates(self): return self.balls
    def SetAllBallStates(self, s): self.balls = s
    def GetAllCarStates(self): return self.cars
    def SetAllCarStates(self, s): self.cars = s
    def GetAllBoostPadStates(self): return self.pads
    def SetAllBoostPadStates(self, s): self.pads = s


def host_workload(kind):
    from emy_s_framework_amd import workloads as WL
    sp = SPECS[kind]
    n = sp["n"]
    teams = WL.mixed_team_sizes(n, sp["seed"]) if kind == "mixed" else [1] * n
    b = RecordBatch(teams)
    if kind == "stress":
        WL.stress_states(b, sp["seed"])
    else:
        WL.random_field_states(b, sp["seed"])
    return b


def sample(kind, count, tick=1000, action_seed=17):
    """`count` arenas spread evenly over the full workload, as cases for
    parity_harness (controls filled in by the runners), each with one
    uniform default-table action index per car."""
    b = host_workload(kind)
    idx = np.linspace(0, b.num_arenas - 1, count).round().astype(int)
    sz = C.sizeof(E.BoostPadState)
    raw = bytes(b.pads)
    rng = np.random.default_rng(action_seed)
    cases = []
    for a in idx:
        t = b.team_size[a]
        o = b.car_off[a]
        cars = [E.CarState.from_buffer_copy(b.cars[o + i]) for i in range(2 * t)]
        pads = R.PADS.from_buffer_copy(raw[a * 34 * sz:(a + 1) * 34 * sz])
        cases.append(dict(blue=t, orange=t, cars=cars, ball=E.BallState.from_buffer_copy(b.balls[a]), pads=pads,
                          tick=tick, source=int(a), actions=rng.integers(0, 90, 2 * t).astype(np.int32)))
    return cases


def lut_controls():
    return list(E.default_action_lut())


def case_controls(case, t, lut):
    """Controls of every car at tick t: the case's own "controls" (a list, or
    a callable of the tick) when it has them, else its table actions."""
    c = case.get("controls")
    if c is not None:
        return c(t) if callable(c) else c
    return [lut[k] for k in case["actions"]]


def run_oracle(lib, case, ticks, lut, warm=True):
    """Warm tick, reload, then `ticks` ticks under the case's actions; records
    every tick after the reload."""
    nb, no = case["blue"], case["orange"]
    h = C.c_void_p(lib.ref_arena_new(nb, no))
    try:
        lib.ref_step(h, 1)

        def load():
            lib.ref_set_tick(h, case["tick"])
            for i, s in enumerate(case["cars"]):
                lib.ref_set_car(h, i, C.byref(s))
            lib.ref_set_ball(h, C.byref(case["ball"]))
            lib.ref_set_pads(h, case["pads"])

        load()
        for i in range(nb + no):
            lib.ref_set_wheels(h, i, P.WHEEL_ZERO)
        if warm:
            for i, c in enumerate(case_controls(case, 0, lut)):
                lib.ref_set_controls(h, i, C.byref(c))
            lib.ref_step(h, 1)
            load()
        res = P.Result()
        for t in range(ticks):
            for i, c in enumerate(case_controls(case, t, lut)):
                lib.ref_set_controls(h, i, C.byref(c))
            lib.ref_step(h, 1)
            cars = []
            for i in range(nb + no):
                s = R.CarState()
                lib.ref_get_car(h, i, C.byref(s))
                cars.append(s)
            bb = R.BallState()
            lib.ref_get_ball(h, C.byref(bb))
            pp = R.PADS()
            lib.ref_get_pads(h, pp)
            res.cars.append(cars)
            res.balls.append(bb)
            res.pads.append(pp)
            res.scored.append(lib.ref_ball_scored(h))
        res.scored = list(np.maximum.accumulate(res.scored))  # the engine's goal flag is sticky
        return res
    finally:
        lib.ref_arena_free(h)


def run_cpucheck(lib, world, case, ticks, lut, warm=True):
    """The CPU build of the engine's tick, same protocol as run_oracle."""
    nb, no = case["blue"], case["orange"]
    h = C.c_void_p(lib.cc_arena_new(world, nb, no))
    try:
        def load():
            lib.cc_set_tick(h, case["tick"])
            for i, s in enumerate(case["cars"]):
                lib.cc_set_car(h, i, C.byref(s))
            lib.cc_set_ball(h, C.byref(case["ball"]))
            lib.cc_set_pads(h, case["pads"])

        load()
        for i in range(nb + no):
            lib.cc_set_wheels(h, i, P.WHEEL_ZERO6)
        if warm:
            for i, c in enumerate(case_controls(case, 0, lut)):
                lib.cc_set_controls(h, i, C.byref(c))
            lib.cc_step(h, 1)
            load()
        res = P.Result()
        scored = 0
        for t in range(ticks):
            for i, c in enumerate(case_controls(case, t, lut)):
                lib.cc_set_controls(h, i, C.byref(c))
            scored |= lib.cc_step(h, 1)
            cars = []
            for i in range(nb + no):
                s = R.CarState()
                lib.cc_get_car(h, i, C.byref(s))
                cars.append(s)
            bb = R.BallState()
            lib.cc_get_ball(h, C.byref(bb))
            pp = R.PADS()
            lib.cc_get_pads(h, pp)
            res.cars.append(cars)
            res.balls.append(bb)
            res.pads.append(pp)
            res.scored.append(scored)
        return res
    finally:
        lib.cc_arena_free(h)


def engine_batch(cases, seed=7):
    """A GPU batch holding one arena per case (team sizes as the cases')."""
    return E.RocketSimBatch(len(cases), [c["blue"] for c in cases], seed=seed)


def load_engine(batch, cases):
    n = len(cases)
    cars = batch.GetAllCarStates()
    balls = batch.GetAllBallStates()
    pads = batch.GetAllBoostPadStates()
    o = 0
    for a, c in enumerate(cases):
        for i, s in enumerate(c["cars"]):
            cars[o + i] = s
        o += len(c["cars"])
        balls[a] = c["ball"]
        C.memmove(C.addressof(pads) + a * 34 * C.sizeof(E.BoostPadState), c["pads"], 34 * C.sizeof(E.BoostPadState))
        batch.SetTickCount(a, c["tick"])
    batch.SetAllCarStates(cars)
    batch.SetAllBallStates(balls)
    batch.SetAllBoostPadStates(pads)
    assert o == batch.num_cars and n == batch.num_arenas


def register_away(batch):
    """One tick with every car high in the air, away from the field's centre.

    The oracle's arena ticks once before a case is loaded (run_oracle,
    oracle_teacher_forced) with its cars at their respawn spots (Arena::AddCar
    respawns each car, Arena.cpp:49-61), so during the warm tick every car is
    still registered in the broadphase there, while a new engine batch has
    them registered where Bullet adds them, at the origin: a car resting on a
    car near the origin would then have its warm-tick wheel rays see the other
    car on one side only, and carry different friction factors into tick 0.
    Cars in the air keep the per-wheel friction factors of a new batch (zero,
    as the oracle's zeroed wheels) and reset the suspension pushback."""
    cars = batch.GetAllCarStates()
    for k in range(batch.num_cars):
        # cars of one arena are consecutive, at most 6: distinct spots
        cars[k] = P.default_car(pos=(-2500 + 1000 * (k % 6), 3000, 1000), on_ground=False)
    batch.SetAllCarStates(cars)
    batch.SetAllCarControls((E.CarControls * batch.num_cars)())
    batch.Step(1)


def run_engine(batch, cases, ticks):
    """The GPU engine (one arena per case): a tick with the cars away
    (register_away), load, warm tick, reload, then `ticks` launches of one
    tick, recording every arena after each.  The cars get the cases' discrete
    actions through the device action table (rsx_set_actions_from_indices),
    or, when a case has its own "controls", every car's controls through
    rsx_set_all_controls."""
    lut = lut_controls()
    direct = any(c.get("controls") is not None for c in cases)

    def set_controls(t):
        if not direct:
            batch.SetDiscreteActionsFromIndices(np.concatenate([c["actions"] for c in cases]).astype(np.int32))
            return
        arr = (E.CarControls * batch.num_cars)()
        o = 0
        for c in cases:
            for i, k in enumerate(case_controls(c, t, lut)):
                arr[o + i] = k
            o += len(c["cars"])
        batch.SetAllCarControls(arr)

    register_away(batch)
    load_engine(batch, cases)
    set_controls(0)
    batch.Step(1)
    load_engine(batch, cases)
    res = [P.Result() for _ in cases]
    for t in range(ticks):
        set_controls(t)
        batch.Step(1)
        batch.Synchronize()
        cars = batch.GetAllCarStates()
        balls = batch.GetAllBallStates()
        pads = batch.GetAllBoostPadStates()
        infos = batch.GetAllArenaInfos()
        o = 0
        sz = C.sizeof(E.BoostPadState)
        raw = bytes(pads)
        for a, c in enumerate(cases):
            k = len(c["cars"])
            res[a].cars.append([E.CarState.from_buffer_copy(cars[o + i]) for i in range(k)])
            o += k
            res[a].balls.append(E.BallState.from_buffer_copy(balls[a]))
            res[a].pads.append(R.PADS.from_buffer_copy(raw[a * 34 * sz:(a + 1) * 34 * sz]))
            res[a].scored.append(int(infos[a].goal_scored))
    return res


def errors(ref, eng, t):
    """(bad: failures dict, err: largest physics error in units of the bar)."""
    fails = P.compare_tick(ref, eng, t)
    err = 0.0
    for rc, ec in zip(ref.cars[t], eng.cars[t]):
        av, bv = R.car_vec(ec), R.car_vec(rc)
        for lo, hi in ((0, 3), (12, 15), (15, 18)):
            err = max(err, P.vec_err(av[lo:hi], bv[lo:hi]))
    av, bv = R.ball_vec(eng.balls[t]), R.ball_vec(ref.balls[t])
    for lo, hi in ((0, 3), (3, 6), (6, 9)):
        err = max(err, P.vec_err(av[lo:hi], bv[lo:hi]))
    return fails, err


def run_cpucheck_batch(lib, world, cases, ticks, lut):
    """run_cpucheck over every case (the engine side of a CPU check)."""
    return [run_cpucheck(lib, world, c, ticks, lut) for c in cases]


def _nudge(v, rng):
    for ax in "xyz":
        x = np.float32(getattr(v, ax))
        setattr(v, ax, float(np.nextafter(x, np.float32(np.inf if rng.uniform() < 0.5 else -np.inf))))


def _perturbed(cars, ball, seed):
    """Copies of the records with every component of the positions,
    orientations, velocities and angular velocities moved one float32 ulp up
    or down (the size of the engine's rounding differences from the oracle)."""
    rng = np.random.default_rng(seed)
    out = []
    for s in cars:
        s2 = R.CarState.from_buffer_copy(s)
        for v in (s2.pos, s2.rot.forward, s2.rot.right, s2.rot.up, s2.vel, s2.ang_vel):
            _nudge(v, rng)
        out.append(s2)
    b = R.BallState.from_buffer_copy(ball)
    for v in (b.pos, b.vel, b.ang_vel):
        _nudge(v, rng)
    return out, b


RESPAWN_X = (-2304.0, -2688.0, 2304.0, 2688.0)  # RLConst.h CAR_RESPAWN_LOCATIONS_SOCCAR
RESPAWN_Y, RESPAWN_Z = 4608.0, 36.0


def at_respawn_spot(c):
    """A car just respawned (Car::Respawn, Car.cpp:43-56): on one of its
    team's four spots after one tick of falling, boost 33.33, not demolished."""
    y = -RESPAWN_Y if c.team == 0 else RESPAWN_Y
    return (min(abs(c.pos.x - x) for x in RESPAWN_X) < 1.0 and abs(c.pos.y - y) < 1.0 and
            abs(c.pos.z - RESPAWN_Z) < 1.0 and abs(c.boost - 100 / 3.0) < 0.01 and not c.is_demoed)


def oracle_teacher_forced(lib, case, eng, ticks, lut, perturb_at=-1, seed=0):
    """The oracle re-fed the engine's state every tick: tick 0 starts from the
    case (same warm protocol as run_oracle), tick t > 0 from the engine's
    records after tick t - 1 (cars, ball, pads, tick count); per-wheel carried
    values and broadphase registrations stay the oracle's own, from its
    previous tick.  Records each tick's result, so result t is the oracle's
    single tick from exactly the state the engine's tick t started from.
    perturb_at >= 0 moves that tick's input positions and velocities by one
    ulp (seeded), to measure how far the oracle departs from itself."""
    nb, no = case["blue"], case["orange"]
    h = C.c_void_p(lib.ref_arena_new(nb, no))
    try:
        lib.ref_step(h, 1)

        def load(tick, cars, ball, pads):
            lib.ref_set_tick(h, tick)
            for i, s in enumerate(cars):
                lib.ref_set_car(h, i, C.byref(s))
            lib.ref_set_ball(h, C.byref(ball))
            lib.ref_set_pads(h, pads)

        load(case["tick"], case["cars"], case["ball"], case["pads"])
        for i in range(nb + no):
            lib.ref_set_wheels(h, i, P.WHEEL_ZERO)
        for i, c in enumerate(case_controls(case, 0, lut)):
            lib.ref_set_controls(h, i, C.byref(c))
        lib.ref_step(h, 1)
        res = P.Result()
        for t in range(ticks):
            if t == 0:
                cars, ball, pads = case["cars"], case["ball"], case["pads"]
            else:
                cars, ball, pads = eng.cars[t - 1], eng.balls[t - 1], eng.pads[t - 1]
            if t == perturb_at:
                cars, ball = _perturbed(cars, ball, seed)
            load(case["tick"] + t, cars, ball, pads)
            for i, c in enumerate(case_controls(case, t, lut)):
                lib.ref_set_controls(h, i, C.byref(c))
            lib.ref_step(h, 1)
            out = []
            for i in range(nb + no):
                s = R.CarState()
                lib.ref_get_car(h, i, C.byref(s))
                out.append(s)
            bb = R.BallState()
            lib.ref_get_ball(h, C.byref(bb))
            pp = R.PADS()
            lib.ref_get_pads(h, pp)
            res.cars.append(out)
            res.balls.append(bb)
            res.pads.append(pp)
            res.scored.append(lib.ref_ball_scored(h))
        return res
    finally:
        lib.ref_arena_free(h)


def _scored_now(res):
    """Per-tick goal test on a result's ball records (the engine's own goal
    flag is sticky over an episode)."""
    spec = R.load_spec()
    return [int(spec.ball_scored(b)) for b in res.balls]


def teacher_forced_check(lib, cases, engs, ticks, lut, n_perturb=8):
    """Single-tick parity at every tick of the engines' runs.  Returns
    (failures, exempt, worst): an arena-tick outside the bar is exempt when the
    oracle itself, given the same input moved by one ulp, departs from its own
    result by more than the bar (the single-tick bar is then unattainable by
    any code that is not bit-identical to the oracle's x86 build).  After an
    exempt tick the oracle's carried per-wheel values and broadphase cells come
    from a result the bar could not fix, so the arena's later ticks are
    compared and reported but exempt as well.  A car respawning after a
    demolition lands on a spot each side draws at random: at that tick its
    position only has to be one of its team's spots on both sides.  The rest
    are failures."""
    failures, exempt, worst = [], [], 0.0
    for case, eng in zip(cases, engs):
        ref = oracle_teacher_forced(lib, case, eng, ticks, lut)
        e2 = P.Result()
        e2.cars, e2.balls, e2.pads, e2.scored = eng.cars, eng.balls, eng.pads, _scored_now(eng)
        tainted = False
        for t in range(ticks):
            f, e = errors(ref, e2, t)
            if f and t > 0:
                # a car respawning this tick lands on a spot each side draws
                # at random: accepted when only its position differs and both
                # sides put it on one of its team's spots
                for i, (a, b) in enumerate(zip(eng.cars[t - 1], eng.cars[t])):
                    if a.is_demoed and not b.is_demoed and all(k.startswith(f"car{i}.pos") for k in f) and \
                            at_respawn_spot(b) and at_respawn_spot(ref.cars[t][i]):
                        f = {}
            if not f:
                worst = max(worst, e) if e < 1.0 else worst
                continue
            if tainted:
                exempt.append((case["source"], t, round(e, 2), "after an exempt tick", sorted(f)[:4]))
                continue
            env = 0.0
            for k in range(n_perturb):
                r2 = oracle_teacher_forced(lib, case, eng, t + 1, lut, perturb_at=t, seed=k)
                env = max(env, errors(ref, r2, t)[1])
            rec = (case["source"], t, round(e, 2), round(env, 2), sorted(f)[:4])
            if env > 1.0:
                exempt.append(rec)
                tainted = True
            else:
                failures.append(rec)
    return failures, exempt, worst
