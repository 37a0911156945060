"""Divergence of the gfx950 engine from the oracle over 120-tick rollouts.

Usage (on a machine with the GPU and the built libraries):
    python tests/divergence_curve.py [--arenas 256] [--ticks 120] [--out profiles/...md]

For each workload sample (workload_harness: the random-state 1v1 workload,
4096 arenas seed 1, and the contact-heavy 1v1 workload, 65536 arenas seed 4;
`--arenas` arenas spread evenly over each), the GPU engine and the oracle run
the same rollout free, from the same state and under the same actions (a
fresh uniform default-table action per car every 8 ticks, as an env step
draws them), one tick per launch.  Before the rollout both take the warm tick
of workload_harness (step, load the state again).

Per tick the script reports the share of arenas whose every body is within
the single-tick parity bar (1e-3 + 1e-4 |ref| per vector for position,
velocity, angular velocity; 1e-4 on rotation entries) and the median and
largest error in units of that bar.  The same numbers for the oracle against
itself, its start state moved by one float32 ulp, show how fast the dynamics
alone separate two rollouts that start a rounding error apart: the engine's
curve is read against that one.  The rollouts are reported, not gated (the
dynamics are chaotic); the single-tick bar is gated by test_gpu_workloads.py.
"""
import argparse
import ctypes as C
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import parity_harness as P  # noqa: E402
import rsxlib as R  # noqa: E402
import workload_harness as WH  # noqa: E402

E = R.engine


def _actions(cases, ticks, seed):
    rng = np.random.default_rng(seed)
    steps = (ticks + 7) // 8
    return [[rng.integers(0, 90, len(c["cars"])).astype(np.int32) for _ in range(steps)] for c in cases]


def oracle_rollout(lib, case, acts, ticks, lut, perturb_seed=None):
    nb, no = case["blue"], case["orange"]
    cars, ball = case["cars"], case["ball"]
    if perturb_seed is not None:
        cars, ball = WH._perturbed(cars, ball, perturb_seed)
    h = C.c_void_p(lib.ref_arena_new(nb, no))
    try:
        lib.ref_step(h, 1)

        def load():
            lib.ref_set_tick(h, case["tick"])
            for i, s in enumerate(cars):
                lib.ref_set_car(h, i, C.byref(s))
            lib.ref_set_ball(h, C.byref(ball))
            lib.ref_set_pads(h, case["pads"])

        load()
        for i in range(nb + no):
            lib.ref_set_wheels(h, i, P.WHEEL_ZERO)
        for i, k in enumerate(acts[0]):
            lib.ref_set_controls(h, i, C.byref(lut[k]))
        lib.ref_step(h, 1)
        load()
        res = P.Result()
        for t in range(ticks):
            for i, k in enumerate(acts[t // 8]):
                lib.ref_set_controls(h, i, C.byref(lut[k]))
            lib.ref_step(h, 1)
            out = []
            for i in range(nb + no):
                s = R.CarState()
                lib.ref_get_car(h, i, C.byref(s))
                out.append(s)
            b = R.BallState()
            lib.ref_get_ball(h, C.byref(b))
            res.cars.append(out)
            res.balls.append(b)
        return res
    finally:
        lib.ref_arena_free(h)


def engine_rollout(cases, acts, ticks):
    batch = WH.engine_batch(cases)
    try:
        WH.register_away(batch)
        WH.load_engine(batch, cases)
        batch.SetDiscreteActionsFromIndices(np.concatenate([a[0] for a in acts]))
        batch.Step(1)
        WH.load_engine(batch, cases)
        res = [P.Result() for _ in cases]
        for t in range(ticks):
            batch.SetDiscreteActionsFromIndices(np.concatenate([a[t // 8] for a in acts]))
            batch.Step(1)
            batch.Synchronize()
            cars = batch.GetAllCarStates()
            balls = batch.GetAllBallStates()
            o = 0
            for a, c in enumerate(cases):
                k = len(c["cars"])
                res[a].cars.append([E.CarState.from_buffer_copy(cars[o + i]) for i in range(k)])
                res[a].balls.append(E.BallState.from_buffer_copy(balls[a]))
                o += k
        return res
    finally:
        batch.Destroy()


def physics_error(ref, got, t):
    """Largest error of any body at tick t in units of the parity bar
    (rotation entries against their own 1e-4 bar)."""
    err = 0.0
    for rc, ec in zip(ref.cars[t], got.cars[t]):
        a, b = R.car_vec(ec), R.car_vec(rc)
        for lo, hi in ((0, 3), (12, 15), (15, 18)):
            err = max(err, P.vec_err(a[lo:hi], b[lo:hi]))
        err = max(err, float(np.max(np.abs(a[3:12] - b[3:12]))) / P.ROT_ABS)
    a, b = R.ball_vec(got.balls[t]), R.ball_vec(ref.balls[t])
    for lo, hi in ((0, 3), (3, 6), (6, 9)):
        err = max(err, P.vec_err(a[lo:hi], b[lo:hi]))
    return err


def curve(refs, gots, ticks):
    rows = []
    for t in range(ticks):
        e = np.array([physics_error(r, g, t) for r, g in zip(refs, gots)])
        rows.append((t + 1, float(np.mean(e <= 1.0)), float(np.median(e)), float(np.max(e))))
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--arenas", type=int, default=256)
    ap.add_argument("--ticks", type=int, default=120)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(HERE), "profiles",
                                                  "divergence_120_ticks_random_and_stress.md"))
    args = ap.parse_args()
    lib = R.load_oracle()
    if lib is None:
        raise SystemExit("oracle/_ref/librsx_ref.so is missing")
    lut = WH.lut_controls()
    lines = ["# Divergence of the gfx950 engine from the oracle over 120-tick rollouts", "",
             f"Command: `python tests/divergence_curve.py --arenas {args.arenas} --ticks {args.ticks}`", "",
             "Per tick: share of arenas with every body within the single-tick parity bar, median and largest "
             "error in units of the bar.  `engine` = GPU engine vs oracle; `oracle 1 ulp` = oracle vs itself "
             "started one float32 ulp apart (positions, orientations, velocities).  Free rollouts, a fresh "
             "uniform default-table action per car every 8 ticks.", ""]
    for kind in ("random", "stress"):
        t0 = time.time()
        cases = WH.sample(kind, args.arenas)
        acts = _actions(cases, args.ticks, seed=31)
        gots = engine_rollout(cases, acts, args.ticks)
        refs = [oracle_rollout(lib, c, a, args.ticks, lut) for c, a in zip(cases, acts)]
        selfs = [oracle_rollout(lib, c, a, args.ticks, lut, perturb_seed=5) for c, a in zip(cases, acts)]
        ce = curve(refs, gots, args.ticks)
        co = curve(refs, selfs, args.ticks)
        print(f"# {kind}: {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
        lines += [f"## {kind} workload ({args.arenas} arenas of {WH.SPECS[kind]['n']}, seed {WH.SPECS[kind]['seed']})",
                  "", "| tick | engine within bar | engine median | engine max | oracle 1 ulp within bar | "
                  "oracle 1 ulp median | oracle 1 ulp max |", "|---|---|---|---|---|---|---|"]
        for (t, f, m, x), (_, fo, mo, xo) in zip(ce, co):
            if t in (1, 2, 4) or t % 8 == 0 or t == args.ticks:
                lines.append(f"| {t} | {f:.3f} | {m:.3g} | {x:.3g} | {fo:.3f} | {mo:.3g} | {xo:.3g} |")
        lines.append("")
    text = "\n".join(lines) + "\n"
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    with open(args.out, "w") as f:
        f.write(text)
    print(text)


if __name__ == "__main__":
    main()
