"""Trains the learned cull's network on the GPU and writes its weights.

    python emy-s-framework_amd/train_sdf.py --out weights.npz --report report.json

Default sizes follow the learned-cull workload: 4M training points and 1M
held-out points (70% near the mesh surface, 30% uniform in its bounding box),
labels from the exact GPU mesh query.  The report gives the held-out max, p99
and mean error, the safety margin the cull is attached with, and the cull
rate of ball and car queries on states of the kickoff workload (64 1v1
arenas, 1000 env steps of uniformly random actions, seed 0)."""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))


def _package():
    if "emy_s_framework_amd" in sys.modules:
        return sys.modules["emy_s_framework_amd"]
    spec = importlib.util.spec_from_file_location("emy_s_framework_amd", os.path.join(PKG, "__init__.py"),
                                                  submodule_search_locations=[PKG])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["emy_s_framework_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def kickoff_query_points(eng, steps=1000, arenas=64, seed=0):
    """Ball centres and car box centres (uu) along the kickoff workload."""
    import torch
    b = eng.RocketSimBatch(arenas, 1, tick_skip=8, seed=seed)
    b.ResetAllArenas()
    b.ConfigureEnv(auto_reset=True)
    rng = np.random.default_rng(seed)
    from emy_s_framework_amd.sdf import CAR_BOX_OFFSET_UU as off
    balls, cars = [], []
    dev = torch.device("cuda")
    for _ in range(steps):
        acts = torch.from_numpy(rng.integers(0, 90, b.num_cars).astype(np.int32)).to(dev)
        b.StepEnv(acts.data_ptr())
        b.Synchronize()
        for s in b.GetAllBallStates():
            balls.append((s.pos.x, s.pos.y, s.pos.z))
        for c in b.GetAllCarStates():
            f, r, u = c.rot.forward, c.rot.right, c.rot.up
            p = np.array([c.pos.x, c.pos.y, c.pos.z]) + off[0] * np.array([f.x, f.y, f.z]) + \
                off[1] * np.array([r.x, r.y, r.z]) + off[2] * np.array([u.x, u.y, u.z])
            cars.append(p)
    b.Destroy()
    return (torch.tensor(np.array(balls, np.float32), device=dev), torch.tensor(np.array(cars, np.float32), device=dev))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(PKG, "assets", "sdf_soccar.npz"))
    ap.add_argument("--report", default=None)
    ap.add_argument("--train", type=int, default=4_000_000)
    ap.add_argument("--val", type=int, default=1_000_000)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--cull-steps", type=int, default=1000)
    ap.add_argument("--remargin", default=None, metavar="WEIGHTS",
                    help="keep these weights; recompute the margins and cull rates on the held-out split")
    args = ap.parse_args()
    import torch
    eng = _package()
    from emy_s_framework_amd.sdf import NeuralSdf, cull_rates, sample_points, train, validate
    torch.cuda.init()
    host = eng.RocketSimBatch(1, 1)
    net = NeuralSdf(host, seed=args.seed)
    t0 = time.perf_counter()
    if args.remargin:
        old = net.load(args.remargin)
        pts = sample_points(net.mesh_triangles(), args.train + args.val, args.seed)[args.train:]
        xva = torch.from_numpy(pts).cuda()
        yva = torch.cat([net.exact_distance(xva[s:s + (1 << 20)]) for s in range(0, xva.shape[0], 1 << 20)])
        rep = validate(net, xva, yva)
        rep["weights_from"] = os.path.basename(args.remargin)
        rep.update({k: v for k, v in old.items() if k not in rep and np.ndim(v) == 0})
    else:
        rep = train(net, n_train=args.train, n_val=args.val, batch=args.batch, epochs=args.epochs, seed=args.seed,
                    log=lambda m: print(m, flush=True))
    rep["wall_seconds"] = time.perf_counter() - t0
    rb, rc = net.reach()
    rep["reach_ball_uu"], rep["reach_car_uu"] = rb, rc
    pts = kickoff_query_points(eng, steps=args.cull_steps)
    rep["cull_rate_kickoff"] = {f"every_{e}": cull_rates(net, pts, mb, mc) for e, mb, mc in
                                zip(rep["schedules"], rep["margin_ball_uu"], rep["margin_car_uu"])}
    net.save(args.out, margin_uu=rep["margin_uu"], max_abs_err_uu=rep["max_abs_err_uu"],
             p99_abs_err_uu=rep["p99_abs_err_uu"], schedules=np.array(rep["schedules"]),
             margin_ball_uu=np.array(rep["margin_ball_uu"]), margin_car_uu=np.array(rep["margin_car_uu"]))
    print(json.dumps(rep), flush=True)
    if args.report:
        with open(args.report, "w") as f:
            json.dump(rep, f, indent=1)
    net.destroy()
    host.Destroy()


if __name__ == "__main__":
    main()
