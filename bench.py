"""Env-step throughput of the gfx950 arena engine on one MI355X.

Workload (default): 65536 soccar 1v1 arenas from the random-state workload
(emy-s-framework_amd/workloads.py, seed 4), a fresh uniform discrete action
per car each step (device int32 indices over the 90-entry default table),
tick-skip 8 at 120 Hz, exact collision.  One step = every arena advances one
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


def cpu_baseline(seconds, threads, tick_skip=8):
    """The reference CPU RocketSim (oracle/_ref/librsx_ref.so) on this host:
    64 1v1 arenas from kickoffs, random default-table actions, Step(8) per env
    step, arenas split over `threads` threads; runs for about `seconds`."""
    path = os.path.join(ROOT, "oracle", "_ref", "librsx_ref.so")
    if not os.path.exists(path):
        return None
    eng = _import_package()
    lib = C.CDLL(path)
    lib.ref_init.argtypes = [C.c_char_p]
    lib.ref_arena_new.restype = C.c_void_p
    lib.ref_arena_new.argtypes = [C.c_int, C.c_int]
    lib.ref_arena_free.argtypes = [C.c_void_p]
    lib.ref_reset_kickoff.argtypes = [C.c_void_p, C.c_int]
    lib.ref_run_random.restype = C.c_int64
    lib.ref_run_random.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int, C.c_uint32,
                                   C.POINTER(eng.CarControls), C.c_int]
    if lib.ref_init(eng.MESH_DIR.encode()) != 0:
        raise RuntimeError("oracle mesh load failed")
    lut = eng.default_action_lut()
    n_arenas = 64
    arenas = [lib.ref_arena_new(1, 1) for _ in range(n_arenas)]
    for i, h in enumerate(arenas):
        lib.ref_reset_kickoff(h, i)
    per = [arenas[t::threads] for t in range(threads)]
    arrs = [(C.c_void_p * len(p))(*p) for p in per]
    steps = [1]
    done = [0] * threads
    per_thread_s = [0.0] * threads
    ready = threading.Barrier(threads + 1)
    go = threading.Barrier(threads + 1)

    def work(t):
        # RocketSim builds per-thread caches on a thread's first step: do that
        # (and time one step) before the timed region.
        lib.ref_run_random(arrs[t], len(per[t]), 1, tick_skip, 7 + t, lut, len(lut))
        t0 = time.perf_counter()
        lib.ref_run_random(arrs[t], len(per[t]), 2, tick_skip, 77 + t, lut, len(lut))
        per_thread_s[t] = (time.perf_counter() - t0) / 2
        ready.wait()
        go.wait()
        done[t] = lib.ref_run_random(arrs[t], len(per[t]), steps[0], tick_skip, 1000 + t, lut, len(lut))

    ths = [threading.Thread(target=work, args=(t,)) for t in range(threads)]
    for th in ths:
        th.start()
    ready.wait()
    steps[0] = max(2, int(seconds / max(max(per_thread_s), 1e-6)))
    go.wait()
    t0 = time.perf_counter()
    for th in ths:
        th.join()
    dt = time.perf_counter() - t0
    steps = steps[0]
    for h in arenas:
        lib.ref_arena_free(h)
    total = sum(done)
    return {"value": total / dt, "unit": "env-steps/s", "cores": threads, "kind": "reference",
            "sample": f"64 1v1 arenas x {steps} env steps (tick-skip 8, random default-table actions, "
                      f"physics only) on {threads} threads, {dt:.1f} s"}


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
    ap.add_argument("--audit-steps", type=int, default=8)
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
        d_act.copy_(actions[0])
        res = batch.Autotune(d_act.data_ptr(), env_steps=2)
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

    exact = run("exact collision", args.steps, args.warmup)
    cull = audit = None
    if use_cull:
        net.attach(margin[0], every=args.cull_every, audit=False, margin_car_uu=margin[1])
        cull = run("learned cull", args.steps, args.warmup)
        net.attach(margin[0], every=args.cull_every, audit=True, margin_car_uu=margin[1])
        net.reset_stats()
        run("learned cull, audited", args.audit_steps, 0, retune=False)
        audit = net.stats()
        body_ticks = args.audit_steps * args.tick_skip * (args.arenas + n_cars)
        audit.update({"steps": args.audit_steps, "margin_ball_uu": margin[0], "margin_car_uu": margin[1],
                      "every_ticks": args.cull_every,
                      "culled_fraction": audit["culled"] / body_ticks})
        net.detach()
    head = cull if cull is not None else exact
    achieved = algo_bytes / (head["kernel_ms"] * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_seconds, max(1, min(16, os.cpu_count() or 1)))
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
        "audit": audit,
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
