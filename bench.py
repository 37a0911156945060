"""Env-step throughput of the gfx950 arena engine on one MI355X.

Workload (default): 65536 soccar 1v1 arenas in the contact-heavy states
(emy-s-framework_amd/workloads.py stress_states, seed 4; --workload random or
mixed for the others), a fresh uniform discrete action per car each step
(device int32 indices over the 90-entry default table), tick-skip 8 at
120 Hz, timed with exact collision and with the learned cull.  One step = every arena advances one
env step: 8 physics ticks, the four stub rewards, goal / no-touch terminals,
kickoff auto-reset of terminal arenas and each car's DefaultObs row, all in
one fused launch (rsx_step_env).  Inputs are resident in HBM when the timed
region starts.

Prints one JSON line: throughput in env-steps/s (arenas x steps / s), the
physics kernel's HBM roofline fraction measured with HIP events on the stream
the kernel runs on, and the reference CPU RocketSim timed on this host."""
import argparse
import ctypes as C
import importlib.util
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "emy-s-framework_amd")

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes of one 1v1 env step: arena state read once and written
# once (2 x 1132 B: two cars of 400 B, ball 36 B, 34 pads of 8 B, arena record
# 24 B), one int32 action per car in (2 x 4 B), per car one 89-float DefaultObs
# row and one float reward out (2 x 356 B + 2 x 4 B), one terminal byte.
ALGO_BYTES_1V1 = 2 * 1132 + 2 * 4 + 2 * 89 * 4 + 2 * 4 + 1


def algo_bytes_per_env_step(team):
    """The same convention for an arena of `team` cars per side: 400 B per
    car, 36 B ball, 272 B pads and 24 B arena record read and written, and
    per car an action in, an obs row of 51 + 19 x cars floats and a reward
    out, plus the terminal byte."""
    cars = 2 * team
    return 2 * (400 * cars + 36 + 272 + 24) + cars * 4 + cars * (51 + 19 * cars) * 4 + cars * 4 + 1


def _import_package():
    if "emy_s_framework_amd" in sys.modules:
        return sys.modules["emy_s_framework_amd"]
    spec = importlib.util.spec_from_file_location("emy_s_framework_amd", os.path.join(PKG, "__init__.py"),
                                                  submodule_search_locations=[PKG])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["emy_s_framework_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


class _HostRecords:
    """Whole-batch record arrays on the host (1v1 arenas), for the workload
    generators, which read and write every arena's records at once."""

    def __init__(self, eng, n):
        import numpy as np
        self.num_arenas, self.team_size, self.num_cars = n, [1] * n, 2 * n
        self.balls = (eng.BallState * n)()
        self.cars = (eng.CarState * (2 * n))()
        v = np.ctypeslib.as_array(self.cars)
        v["rot"]["forward"]["x"] = v["rot"]["right"]["y"] = v["rot"]["up"]["z"] = 1
        v["pos"]["z"], v["boost"], v["is_on_ground"] = 17, 33.3, 1
        v["team"] = np.arange(2 * n) & 1
        v["ball_hit_tick"] = v["ball_hit_extra_impulse_tick"] = 2**64 - 1
        self.pads = (eng.BoostPadState * (34 * n))()

    def GetAllBallStates(self): return self.balls
    def SetAllBallStates(self, s): self.balls = s
    def GetAllCarStates(self): return self.cars
    def SetAllCarStates(self, s): self.cars = s
    def GetAllBoostPadStates(self): return self.pads
    def SetAllBoostPadStates(self, s): self.pads = s


def host_info():
    """Cores the CPU baseline may use and what they are."""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    nproc = os.cpu_count() or 1
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(avail, share) if share > 0 else avail
    return {"cpu_model": model, "nproc": nproc, "affinity": avail, "omp_num_threads": share or None,
            "threads": max(1, threads)}


def cpu_baseline(seconds, threads, workload, tick_skip=8, sample_every=32):
    """The reference CPU RocketSim (oracle/_ref/librsx_ref.so) on this host,
    running the same env step the engine fuses: a uniform default-table action
    per car, Arena::Step(tick_skip), DefaultObs rows, the four stub rewards,
    goal / no-touch (10 s) terminals and the kickoff reset of terminal arenas
    (ref_env_run).  Arenas: every `sample_every`-th arena of the benchmark's
    own 65536-arena workload (per-arena cost does not depend on the batch
    size), split over `threads` threads; runs for about `seconds`.  Also times
    the C1 shape (64 1v1 arenas from kickoffs) for a few seconds."""
    path = os.path.join(ROOT, "oracle", "_ref", "librsx_ref.so")
    if not os.path.exists(path):
        return None
    eng = _import_package()
    from emy_s_framework_amd.workloads import random_field_states, stress_states
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import rlgym_spec
    lib = C.CDLL(path)
    VP = C.c_void_p
    lib.ref_init.argtypes = [C.c_char_p]
    lib.ref_arena_new.restype = VP
    lib.ref_arena_new.argtypes = [C.c_int, C.c_int]
    lib.ref_arena_free.argtypes = [VP]
    lib.ref_reset_kickoff.argtypes = [VP, C.c_int]
    lib.ref_step.argtypes = [VP, C.c_int]
    lib.ref_set_car.argtypes = [VP, C.c_int, C.POINTER(eng.CarState)]
    lib.ref_set_ball.argtypes = [VP, C.POINTER(eng.BallState)]
    lib.ref_set_pads.argtypes = [VP, C.POINTER(eng.BoostPadState)]
    lib.ref_set_tick.argtypes = [VP, C.c_uint64]
    lib.ref_env_run.restype = C.c_int64
    lib.ref_env_run.argtypes = [C.POINTER(VP), C.c_int, C.c_int, C.c_int, C.c_uint32, C.POINTER(eng.CarControls),
                                C.c_int, C.POINTER(C.c_int32), C.c_float]
    if lib.ref_init(eng.MESH_DIR.encode()) != 0:
        raise RuntimeError("oracle mesh load failed")
    lut = eng.default_action_lut()
    pad_map = (C.c_int32 * 34)(*[int(x) for x in rlgym_spec.PAD_MAP])

    def timed(n_arenas, make, secs, seed0):
        """Env steps per second of `threads` threads over n_arenas arenas:
        thread t builds arenas t, t + threads, ... with make(k) (Bullet world
        set-up is the slow part), runs one untimed step (per-thread caches),
        then chunks of 2 env steps until `secs` have passed; frees its arenas."""
        per = [list(range(t, n_arenas, threads)) for t in range(threads)]
        per = [p for p in per if p]
        nt = len(per)
        done = [0] * nt
        chunks = [0] * nt
        ready, go = threading.Barrier(nt + 1), threading.Barrier(nt + 1)
        deadline = [0.0]

        def work(t):
            hs = [make(k) for k in per[t]]
            arr = (VP * len(hs))(*hs)
            lib.ref_env_run(arr, len(hs), 1, tick_skip, seed0 + t, lut, len(lut), pad_map, 10.0)
            ready.wait()
            go.wait()
            while time.perf_counter() < deadline[0]:
                done[t] += lib.ref_env_run(arr, len(hs), 2, tick_skip, seed0 + 100 + 7919 * chunks[t] + t, lut,
                                           len(lut), pad_map, 10.0)
                chunks[t] += 1
            for h in hs:
                lib.ref_arena_free(h)

        ths = [threading.Thread(target=work, args=(t,)) for t in range(nt)]
        for th in ths:
            th.start()
        ready.wait()
        t0 = time.perf_counter()
        deadline[0] = t0 + secs
        go.wait()
        for th in ths:
            th.join()
        return sum(done), max(time.perf_counter() - t0, 1e-9), min(chunks) * 2

    # the benchmark workload, sampled
    full = 65536
    recs = _HostRecords(eng, full)
    (stress_states if workload == "stress" else random_field_states)(recs, 4)
    idx = list(range(0, full, sample_every))
    psz = C.sizeof(eng.BoostPadState)

    def make_workload(k):
        a = idx[k]
        h = lib.ref_arena_new(1, 1)
        lib.ref_step(h, 1)
        lib.ref_set_tick(h, 0)
        for i in range(2):
            lib.ref_set_car(h, i, C.byref(recs.cars[2 * a + i]))
        lib.ref_set_ball(h, C.byref(recs.balls[a]))
        lib.ref_set_pads(h, C.cast(C.addressof(recs.pads) + a * 34 * psz, C.POINTER(eng.BoostPadState)))
        return h

    def make_kickoff(k):
        h = lib.ref_arena_new(1, 1)
        lib.ref_reset_kickoff(h, k)
        return h

    total, dt, steps = timed(len(idx), make_workload, seconds, 1000)
    k_total, k_dt, k_steps = timed(64, make_kickoff, min(5.0, seconds / 3), 5000)
    info = host_info()
    return {"value": total / dt, "unit": "env-steps/s", "cores": threads, "kind": "reference",
            "sample": f"{len(idx)} arenas (every {sample_every}th of the 65536-arena {workload} workload, seed 4) x "
                      f">= {steps} env steps, full env step (8 ticks, DefaultObs, 4 stub rewards, goal / no-touch "
                      f"terminals, kickoff auto-reset), {threads} threads, {dt:.1f} s",
            "arena_ticks_per_s": total * tick_skip / dt, "env_steps_per_s_per_core": total / dt / threads,
            "host": info,
            "c1_kickoff_64": {"value": k_total / k_dt, "unit": "env-steps/s",
                              "sample": f"64 1v1 arenas from kickoffs x >= {k_steps} env steps, {threads} threads, "
                                        f"{k_dt:.1f} s"}}


def gae_bench(eng, n_players, dev, steps_per_player=64, reps=20):
    """The device GAE (rsx_gae) over one training iteration's experience
    buffer: every player's trajectory of `steps_per_player` steps ending in a
    truncation, goals scattered (p = 0.002 per step).  Algorithmic bytes per
    step: reward, value, terminal in (9 B), advantage, target, return out
    (12 B)."""
    import torch
    from emy_s_framework_amd.gae import ComputeGAEOnDevice
    T = n_players * steps_per_player
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    term = torch.zeros(T, dtype=torch.int8, device=dev)
    term[steps_per_player - 1::steps_per_player] = 2
    goal = (torch.rand(T, device=dev, generator=g) < 0.002) & (term == 0)
    term[goal] = 1
    rews = torch.randn(T, device=dev, generator=g)
    vals = torch.randn(T, device=dev, generator=g)
    tv = torch.randn(n_players, device=dev, generator=g)
    for _ in range(3):
        ComputeGAEOnDevice(rews, term, vals, tv)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        ComputeGAEOnDevice(rews, term, vals, tv)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    gbs = T * 21 / (ms * 1e-3) / 1e9
    print(f"# device GAE: {T} steps in {ms:.3f} ms ({gbs:.0f} GB/s)", file=sys.stderr, flush=True)
    return {"steps": T, "layout": f"{n_players} players x {steps_per_player} steps, each trajectory truncated at its "
                                  f"end, goals p=0.002", "ms": ms, "steps_per_s": T / (ms * 1e-3),
            "achieved_gbs": gbs, "frac_hbm_peak": gbs / HBM_PEAK_GBS, "algo_bytes_per_step": 21}


def vanta_bench(batch, apply_states, stream, sp, actions, reps=10):
    """VantaObs (RSX_OBS_VANTA, ball prediction on) over the whole batch,
    timed with HIP events on the launch stream: a build from freshly set
    states (every arena a new 360-tick rollout), then env steps each followed
    by the build (arenas whose ball moved as predicted continue their last
    rollout), against the same env steps (the first `reps` after the same
    freshly set states) with VantaObs without prediction and with
    DefaultObs.  Restores DefaultObs."""
    import torch
    eng = _import_package()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn, n):
        torch.cuda.synchronize()
        e0.record(stream)
        for i in range(n):
            fn(i)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n

    batch.ConfigureEnv(no_touch_timeout_s=10.0, auto_reset=True)
    apply_states(batch, seed=4)
    plain_ms = timed(lambda i: batch.StepEnv(actions[i % len(actions)].data_ptr(), sp), reps)
    batch.ConfigureEnv(no_touch_timeout_s=10.0, auto_reset=True, obs_builder=eng.OBS_VANTA)
    batch.ConfigureVanta(predict_ball=True)
    apply_states(batch, seed=4)
    full_ms = timed(lambda i: batch.BuildObs(sp), 1)
    step_ms = timed(lambda i: batch.StepEnv(actions[i % len(actions)].data_ptr(), sp), reps)
    batch.ConfigureVanta(predict_ball=False)
    apply_states(batch, seed=4)  # the same env steps as the two runs above
    nopred_ms = timed(lambda i: batch.StepEnv(actions[i % len(actions)].data_ptr(), sp), reps)
    batch.ConfigureEnv(no_touch_timeout_s=10.0, auto_reset=True)
    print(f"# VantaObs: build with every rollout new {full_ms:.2f} ms; env step + build {step_ms:.2f} ms "
          f"(no prediction {nopred_ms:.2f} ms, DefaultObs env step {plain_ms:.2f} ms)", file=sys.stderr, flush=True)
    return {"build_all_new_rollouts_ms": full_ms, "env_step_plus_build_ms": step_ms,
            "env_step_plus_build_no_prediction_ms": nopred_ms, "env_step_default_obs_ms": plain_ms,
            "env_steps_per_s": batch.num_arenas / (step_ms * 1e-3),
            "rollout": "360 ball-only ticks per arena (slices at 1, 2, 3 s); an arena whose ball is, bit for bit, the "
                       "one its last rollout passed at the env step's tick skip continues that rollout instead "
                       f"({reps} env steps after freshly set states, each variant from the same states)"}


def measured_traffic(workload):
    """HBM bytes per k_env_step<2> launch from the committed rocprofv3
    counter summary of this workload (FETCH_SIZE + WRITE_SIZE, copy-scaled),
    or None."""
    path = os.path.join(ROOT, "profiles", f"k_env_step_1v1_65536_{workload}_states.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--arenas", type=int, default=None, help="default 65536 (16384 for the mixed workload)")
    ap.add_argument("--tick-skip", type=int, default=8)
    ap.add_argument("--workload", choices=("stress", "random", "mixed"), default="stress")
    ap.add_argument("--graph", action=argparse.BooleanOptionalAction, default=True,
                    help="replay each env step from a captured hipGraph")
    ap.add_argument("--autotune", action=argparse.BooleanOptionalAction, default=True,
                    help="pick the launch shape with the library's autotuner first")
    ap.add_argument("--tuning", default=None,
                    help="fixed launch shape 'block,arenas_per_wave[,pool]' (skips the autotuner)")
    ap.add_argument("--cull", choices=("auto", "on", "off"), default="auto")
    ap.add_argument("--audit-steps", type=int, default=1000,
                    help="env steps of the culled engine with every culled body checked against the exact mesh "
                         "distance on every tick")
    ap.add_argument("--long-steps", type=int, default=300,
                    help="env steps of the long window (exact and culled, from fresh states); 0 skips it")
    ap.add_argument("--gae", action=argparse.BooleanOptionalAction, default=True,
                    help="also time the device GAE on the buffer of one training iteration")
    ap.add_argument("--vanta", action=argparse.BooleanOptionalAction, default=True,
                    help="also time the VantaObs build (ball prediction on) over the same arenas")
    ap.add_argument("--cull-every", type=int, default=8)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-cal", action="store_true",
                    help="only copy 1 GiB device to device three times (scales the HBM byte counters)")
    args = ap.parse_args()
    if args.copy_cal:
        import torch
        x = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print("copied 1 GiB x3", flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: the arenas are independent and every GPU box here has one GPU "
              f"(asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    import torch
    eng = _import_package()
    from emy_s_framework_amd.sdf import WEIGHTS_PATH, NeuralSdf, margins_for
    from emy_s_framework_amd.workloads import mixed_team_sizes, random_field_states, stress_states

    torch.cuda.init()
    dev = torch.device("cuda:0")
    if args.arenas is None:
        args.arenas = 16384 if args.workload == "mixed" else 65536
    if args.workload == "mixed":
        teams = mixed_team_sizes(args.arenas, seed=2)
        batch = eng.RocketSimBatch(args.arenas, teams, tick_skip=args.tick_skip, seed=2, device=0)
        apply_states = lambda b, seed: random_field_states(b, 2)  # noqa: E731
    else:
        teams = [1] * args.arenas
        batch = eng.RocketSimBatch(args.arenas, 1, tick_skip=args.tick_skip, seed=4, device=0)
        apply_states = stress_states if args.workload == "stress" else random_field_states
    algo_bytes = sum(algo_bytes_per_env_step(t) for t in teams)
    batch.ConfigureEnv(no_touch_timeout_s=10.0, auto_reset=True)
    n_cars = batch.num_cars
    stream = torch.cuda.current_stream(dev)
    sp = C.c_void_p(stream.cuda_stream)
    gen = torch.Generator(device=dev)
    gen.manual_seed(4)
    actions = [torch.randint(0, 90, (n_cars,), dtype=torch.int32, device=dev, generator=gen)
               for _ in range(8)]
    use_cull = args.cull == "on" or (args.cull == "auto" and os.path.exists(WEIGHTS_PATH))
    net = None
    margin = None
    if use_cull:
        net = NeuralSdf(batch, seed=0)
        margin = margins_for(net.load(WEIGHTS_PATH), args.cull_every)

    d_act = torch.empty(n_cars, dtype=torch.int32, device=dev)  # the captured graph's action buffer

    def tune():
        """Launch shape for the current configuration (exact or culled)."""
        if args.tuning:
            batch.SetTuning(*[int(x) for x in args.tuning.split(",")])
            return None
        if not args.autotune:
            return None
        apply_states(batch, seed=4)
        # past the first env steps (freshly placed overlapping cars settle
        # there, at twice the steady step time), as the timed steps will be
        for i in range(args.warmup):
            batch.StepEnv(actions[i % len(actions)].data_ptr(), sp)
        torch.cuda.synchronize()
        d_act.copy_(actions[0])
        res = batch.Autotune(d_act.data_ptr(), env_steps=3)
        return [{"block": b, "arenas_per_wave": a, "pool": p, "ms": ms} for b, a, p, ms in res]

    def run(tag, steps, warmup, retune=True):
        """Fresh workload states, warmup, then `steps` timed fused env steps
        (with --graph each replayed from the captured hipGraph, whose action
        buffer is refreshed by a device copy before the replay)."""
        tuning = tune() if retune else None
        apply_states(batch, seed=4)
        if args.graph:
            batch.CaptureEnvGraph(d_act.data_ptr(), 1)

        def env_step(i, ev=None):
            if args.graph:
                d_act.copy_(actions[i % len(actions)])
                if ev:
                    ev[0].record(stream)
                batch.LaunchEnvGraph(sp)
            else:
                if ev:
                    ev[0].record(stream)
                batch.StepEnv(actions[i % len(actions)].data_ptr(), sp)
            if ev:
                ev[1].record(stream)

        torch.cuda.synchronize()
        for i in range(warmup):
            env_step(i)
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(steps):
            env_step(i, evs[i])
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kern = sum(a.elapsed_time(b) for a, b in evs) / steps
        print(f"# {tag}: {steps} steps, {wall * 1e3 / steps:.2f} ms/step (kernel {kern:.2f} ms)", file=sys.stderr,
              flush=True)
        shape = batch.GetTuning()
        return {"value": args.arenas * steps / wall, "ms_per_step": wall * 1e3 / steps, "kernel_ms": kern,
                "agent_steps_per_s": n_cars * steps / wall,
                "launch_shape": {"block": shape[0], "arenas_per_wave": shape[1], "pool_narrowphase": shape[2]},
                "autotune": tuning}

    def long_window(tag):
        if args.long_steps <= 0:
            return None
        r = run(tag, args.long_steps, 0, retune=False)
        return {"steps": args.long_steps, "ms_per_step": r["ms_per_step"], "kernel_ms": r["kernel_ms"],
                "value": r["value"]}

    exact = run("exact collision", args.steps, args.warmup)
    exact_long = long_window("exact collision, long window")
    cull = audit = cull_long = None
    if use_cull:
        net.attach(margin[0], every=args.cull_every, audit=False, margin_car_uu=margin[1])
        cull = run("learned cull", args.steps, args.warmup)
        cull_long = long_window("learned cull, long window")
        net.attach(margin[0], every=args.cull_every, audit=True, margin_car_uu=margin[1])
        net.reset_stats()
        run("learned cull, audited", args.audit_steps, 0, retune=False)
        audit = net.stats()
        body_ticks = args.audit_steps * args.tick_skip * (args.arenas + n_cars)
        audit.update({"steps": args.audit_steps, "ticks_audited": args.audit_steps * args.tick_skip,
                      "margin_ball_uu": margin[0], "margin_car_uu": margin[1], "every_ticks": args.cull_every,
                      "culled_fraction": audit["culled"] / body_ticks})
        net.detach()
    gae = gae_bench(eng, n_cars, dev) if args.gae else None
    vanta = vanta_bench(batch, apply_states, stream, sp, actions) if args.vanta else None
    head = cull if cull is not None else exact
    achieved = algo_bytes / (head["kernel_ms"] * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_seconds, host_info()["threads"],
                           args.workload if args.workload != "mixed" else "random")
    wl = {"stress": "contact-heavy states (seed 4: 30% ball near corners/ramps/goal mouths, 20% cars on walls or "
                    "ceiling, 10% overlapping car spawns)",
          "random": "random field states (seed 4)",
          "mixed": "random field states (seed 2), team sizes 1v1 50% / 2v2 30% / 3v3 20% shuffled (seed 2), "
                   "one launch over all team sizes"}[args.workload]
    kinds = "1v1" if args.workload != "mixed" else "mixed 1v1/2v2/3v3"
    out = {
        "metric": "env-steps/sec @65536 1v1 arenas (neural-SDF cull vs exact), % HBM roofline",
        "value": head["value"],
        "unit": "env-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.arenas} soccar {kinds} arenas, {wl}, uniform random discrete actions, "
                               f"tick-skip {args.tick_skip}, fused env step (physics + DefaultObs + 4 stub rewards "
                               f"+ terminals + auto-reset)" + (", replayed from a hipGraph" if args.graph else ""),
                   "arenas": args.arenas, "agents": n_cars, "tick_skip": args.tick_skip,
                   "cull": (f"learned SDF cull every {args.cull_every} ticks, margins {margin[0]:.1f} uu (ball) / "
                            f"{margin[1]:.1f} uu (cars), exact "
                            f"fallback" if cull is not None else "none (exact collision)")},
        "exact": exact,
        "cull": cull,
        "exact_long": exact_long,
        "cull_long": cull_long,
        "audit": audit,
        "gae": gae,
        "vanta_obs": vanta,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(args.workload) if args.arenas == 65536 else None,
                     "kernel": "k_env_mixed" if args.workload == "mixed" else "k_env_step<2>",
                     "kernel_ms": head["kernel_ms"], "algo_bytes_per_launch": algo_bytes,
                     "algo_bytes_per_env_step_1v1": ALGO_BYTES_1V1},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    if net is not None:
        net.destroy()
    batch.Destroy()


if __name__ == "__main__":
    main()
