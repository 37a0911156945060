"""Pins the oracle (the reference CPU RocketSim built from its own sources)
before it is trusted as the parity reference: its mesh set against the
reference's golden hashes, its free flight against the closed-form Bullet
integration, its kickoff against the reference's spawn table."""
import ctypes as C
import json
import os

import numpy as np

import parity_harness as P
import rsxlib as R

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_mesh_hashes_match_reference(cpucheck):
    """The host world builder parses the 16 CMF files and hashes them with
    RocketSim's hash; the set must equal RocketSim.cpp:22-27's soccar list."""
    lib, w = cpucheck
    out = (C.c_uint32 * 16)()
    lib.cc_mesh_hashes(w, out)
    want = json.load(open(os.path.join(GOLDEN, "soccar_mesh_hashes.json")))["hashes"]
    assert sorted(out) == sorted(want)


def _ball_free_flight(z0, v0, ticks, dt=1 / 120, g=-650.0, drag=0.03):
    """Bullet's per-step order for a free body: the velocity is damped first
    (btDiscreteDynamicsWorld::predictUnconstraintMotion -> applyDamping), the
    solver then adds the external force impulse g*dt at write-back
    (btSequentialImpulseConstraintSolver m_externalForceImpulse), and the
    transform integrates the final velocity (integrateTransforms)."""
    x = np.array([0.0, 0.0, z0])
    v = np.array(v0, dtype=np.float64)
    out = []
    for _ in range(ticks):
        v = v * (1 - drag) ** dt
        v = v + np.array([0, 0, g]) * dt
        x = x + v * dt
        out.append((x.copy(), v.copy()))
    return out


def test_oracle_ball_free_flight_closed_form(oracle):
    far = [P.default_car(pos=(-3000, -4000, 17)), P.default_car(pos=(3000, 4000, 17), yaw=np.pi, team=1)]
    v0 = (100.0, 200.0, 300.0)
    case = dict(blue=1, orange=1, cars=far, ball=P.ball((0, 0, 1000), v0, (1, 2, 3)),
                controls=[P.controls(), P.controls()], ticks=60)
    ref = P.run_oracle(oracle, case)
    want = _ball_free_flight(1000.0, v0, 60)
    for t in range(60):
        b = R.ball_vec(ref.balls[t])
        x, v = want[t]
        assert np.max(np.abs(b[0:3] - x)) <= 1e-3 + 1e-4 * np.linalg.norm(x), t
        assert np.max(np.abs(b[3:6] - v)) <= 1e-3 + 1e-4 * np.linalg.norm(v), t
        np.testing.assert_allclose(b[6:9], [1, 2, 3], rtol=1e-6)


def test_oracle_wrong_gravity_sign_is_caught(oracle):
    """The closed form above is sensitive enough to catch a flipped gravity
    sign or a dropped drag term in a restatement."""
    a = _ball_free_flight(1000.0, (100, 200, 300), 60)
    b = _ball_free_flight(1000.0, (100, 200, 300), 60, g=650.0)
    c = _ball_free_flight(1000.0, (100, 200, 300), 60, drag=0.0)
    assert np.max(np.abs(a[-1][0] - b[-1][0])) > 1.0
    assert np.max(np.abs(a[-1][1] - c[-1][1])) > 1e-3 + 1e-4 * np.linalg.norm(a[-1][1])


KICKOFF_XY = {(-2048, -2560), (2048, -2560), (-256, -3840), (256, -3840), (0, -4608)}


def test_oracle_kickoff_spawns(oracle):
    """RLConst.h CAR_SPAWN_LOCATIONS_SOCCAR: blue cars take distinct spawn spots,
    orange takes the mirror image of the same spots, everyone at z=17 with
    33.3 boost."""
    for seed in range(6):
        h = C.c_void_p(oracle.ref_arena_new(3, 3))
        try:
            oracle.ref_reset_kickoff(h, seed)
            cars = []
            for i in range(6):
                s = R.CarState()
                oracle.ref_get_car(h, i, C.byref(s))
                cars.append(s)
            blue = [(round(c.pos.x), round(c.pos.y)) for c in cars[:3]]
            orange = [(round(c.pos.x), round(c.pos.y)) for c in cars[3:]]
            assert len(set(blue)) == 3 and set(blue) <= KICKOFF_XY
            # blue/orange slot pairing follows the arena's car container order,
            # so compare the sets of spots
            assert sorted(orange) == sorted((-x, -y) for x, y in blue)
            for c in cars:
                assert abs(c.pos.z - 17) < 1e-3
                assert abs(c.boost - 100 / 3) < 1e-3
            b = R.BallState()
            oracle.ref_get_ball(h, C.byref(b))
            assert abs(b.pos.z - 93.15) < 1e-3 and b.pos.x == 0 and b.pos.y == 0
        finally:
            oracle.ref_arena_free(h)
