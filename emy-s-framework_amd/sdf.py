"""Learned broadphase cull: a small MLP distance field of the soccar collision
mesh, trained on the GPU by the library's own kernels (include/rsx_sdf.h).

The network predicts the unsigned distance (uu) from a point to the nearest
mesh triangle; the four static planes stay analytic.  Attached to an engine it
drops a body from the tick's mesh queries when prediction - margin exceeds the
body's reach (plus the distance it can have moved since the network ran); the
exact triangle tests stay the fallback, and an audit mode checks every cull
against the exact distance.

Margins.  The default margin is the largest over-estimate on held-out points
anywhere, x1.25 + 2 uu.  A tighter, per-schedule bound is also reported: a
cull made k ticks after the network ran at p0 misses a contact only if the
body came within its reach R of the mesh, which needs d(p0) <= R + k * step
(the exact distance is 1-Lipschitz), so only over-estimates at points with
d(p) <= R + k * step matter.  Measured on held-out points that bound is
2.5-3x tighter, but sampled points near the surface under-cover its sharp
features (the audit found a miss with it), so it is not the default.

Training data follow the learned-cull workload: points in the mesh bounding
box, 70% within 300 uu of the surface (a uniform point of an area-weighted
random triangle, offset along its normal by N(0, 60 uu) and redrawn beyond
300 uu), 30% uniform; labels are the exact distances from the GPU mesh query.
"""
import ctypes as C
import math
import os
import time

import numpy as np

from .batch import _check, load_library

CLAMP_UU = 1000.0          # training labels are min(distance, CLAMP_UU)
# The car's query point is its hitbox centre: the Octane hitbox offset along
# the car's (forward, right, up) axes (CarConfig.cpp:19-60), uu.
CAR_BOX_OFFSET_UU = (13.87566, 0.0, 20.755)
WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "assets", "sdf_soccar.npz")


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class NeuralSdf:
    """The network of one engine (its mesh and device)."""

    def __init__(self, batch, seed=0):
        self._lib = load_library()
        self._batch = batch
        h = C.c_void_p()
        _check(self._lib.rsx_sdf_create(batch._h, C.c_uint64(seed), C.byref(h)))
        self._h = h
        self.num_params = int(self._lib.rsx_sdf_num_params())

    def destroy(self):
        if getattr(self, "_h", None):
            self._lib.rsx_sdf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    # ---- mesh and exact distance ----
    def mesh_triangles(self):
        """(ntri, 3, 3) float32 vertices in uu."""
        n = C.c_int32()
        _check(self._lib.rsx_sdf_mesh_triangles(self._h, None, 0, C.byref(n)))
        out = np.zeros(n.value * 9, np.float32)
        _check(self._lib.rsx_sdf_mesh_triangles(self._h, out.ctypes.data_as(C.POINTER(C.c_float)), out.size,
                                                C.byref(n)))
        return out.reshape(-1, 3, 3)

    def exact_distance(self, pts):
        """Exact distance (uu) for a (n, 3) float32 CUDA tensor of points (uu)."""
        import torch
        pts = pts.contiguous()
        out = torch.empty(pts.shape[0], dtype=torch.float32, device=pts.device)
        _check(self._lib.rsx_sdf_exact_distance(self._h, _ptr(pts), pts.shape[0], _ptr(out), _stream()))
        return out

    def eval(self, pts):
        """Network distance (uu) for a (n, 3) float32 CUDA tensor of points (uu)."""
        import torch
        pts = pts.contiguous()
        out = torch.empty(pts.shape[0], dtype=torch.float32, device=pts.device)
        _check(self._lib.rsx_sdf_eval(self._h, _ptr(pts), pts.shape[0], _ptr(out), _stream()))
        return out

    # ---- training ----
    def train_step(self, pts, labels, lr, clamp_uu=CLAMP_UU, over_weight=1.0, loss_out=None):
        _check(self._lib.rsx_sdf_train_step(self._h, _ptr(pts), _ptr(labels), pts.shape[0], C.c_float(lr),
                                            C.c_float(clamp_uu), C.c_float(over_weight),
                                            _ptr(loss_out) if loss_out is not None else None, _stream()))

    def gradients(self, pts, labels, clamp_uu=CLAMP_UU, over_weight=1.0):
        """(gradient vector, loss) of the batch, without an update."""
        g = np.zeros(self.num_params, np.float32)
        loss = C.c_float()
        _check(self._lib.rsx_sdf_gradients(self._h, _ptr(pts), _ptr(labels), pts.shape[0], C.c_float(clamp_uu),
                                           C.c_float(over_weight), g.ctypes.data_as(C.POINTER(C.c_float)),
                                           C.byref(loss)))
        return g, float(loss.value)

    def get_params(self):
        p = np.zeros(self.num_params, np.float32)
        _check(self._lib.rsx_sdf_get_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float))))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.num_params
        _check(self._lib.rsx_sdf_set_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float))))

    def save(self, path, **meta):
        np.savez(path, params=self.get_params(), **{k: np.asarray(v) for k, v in meta.items()})

    def load(self, path):
        with np.load(path, allow_pickle=False) as z:
            self.set_params(z["params"])
            return {k: z[k].item() if z[k].ndim == 0 else z[k] for k in z.files if k != "params"}

    # ---- cull ----
    def attach(self, margin_uu, every=2, audit=False, margin_car_uu=None):
        """margin_uu: the ball's margin (and the cars' unless margin_car_uu)."""
        mc = margin_uu if margin_car_uu is None else margin_car_uu
        _check(self._lib.rsx_attach_cull(self._batch._h, self._h, C.c_float(margin_uu), C.c_float(mc), int(every),
                                         1 if audit else 0))

    def detach(self):
        _check(self._lib.rsx_attach_cull(self._batch._h, None, C.c_float(0.0), C.c_float(0.0), 1, 0))

    def stats(self):
        out = (C.c_uint64 * 3)()
        _check(self._lib.rsx_get_cull_stats(self._h, C.byref(out)))
        return {"queries": int(out[0]), "culled": int(out[1]), "defects": int(out[2])}

    def reset_stats(self):
        _check(self._lib.rsx_reset_cull_stats(self._h))

    def reach(self):
        b, c = C.c_float(), C.c_float()
        _check(self._lib.rsx_cull_reach(self._h, C.byref(b), C.byref(c)))
        return float(b.value), float(c.value)


def sample_points(tris, n, seed, near_frac=0.7, sigma=60.0, max_offset=300.0):
    """(n, 3) float32 points (uu) of the training distribution."""
    rng = np.random.default_rng(seed)
    a, b, c = tris[:, 0].astype(np.float64), tris[:, 1].astype(np.float64), tris[:, 2].astype(np.float64)
    cr = np.cross(b - a, c - a)
    area = 0.5 * np.linalg.norm(cr, axis=1)
    nrm = cr / np.maximum(2.0 * area, 1e-12)[:, None]
    n_near = int(round(n * near_frac))
    t = rng.choice(len(tris), size=n_near, p=area / area.sum())
    r1, r2 = np.sqrt(rng.random(n_near)), rng.random(n_near)
    surf = (1 - r1)[:, None] * a[t] + (r1 * (1 - r2))[:, None] * b[t] + (r1 * r2)[:, None] * c[t]
    off = rng.normal(0.0, sigma, n_near)
    bad = np.abs(off) > max_offset
    while bad.any():
        off[bad] = rng.normal(0.0, sigma, bad.sum())
        bad = np.abs(off) > max_offset
    near = surf + off[:, None] * nrm[t]
    lo, hi = tris.reshape(-1, 3).min(0), tris.reshape(-1, 3).max(0)
    uni = lo + (hi - lo) * rng.random((n - n_near, 3))
    pts = np.concatenate([near, uni]).astype(np.float32)
    return pts[rng.permutation(n)]


def train(sdf, n_train=4_000_000, n_val=1_000_000, batch=8192, epochs=40, lr_max=3e-3, lr_min=3e-5, seed=3,
          over_weight=4.0, log=print):
    """Trains on GPU-labelled points; returns the validation report."""
    import torch
    dev = torch.device("cuda")
    tris = sdf.mesh_triangles()
    t0 = time.perf_counter()
    pts = torch.from_numpy(sample_points(tris, n_train + n_val, seed)).to(dev)
    labels = torch.empty(pts.shape[0], dtype=torch.float32, device=dev)
    for s in range(0, pts.shape[0], 1 << 20):
        labels[s:s + (1 << 20)] = sdf.exact_distance(pts[s:s + (1 << 20)])
    torch.cuda.synchronize()
    t_label = time.perf_counter() - t0
    xtr, ytr = pts[:n_train], labels[:n_train]
    xva, yva = pts[n_train:], labels[n_train:]
    steps_per_epoch = n_train // batch
    total = steps_per_epoch * epochs
    loss = torch.zeros(1, dtype=torch.float32, device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    t1 = time.perf_counter()
    step = 0
    for ep in range(epochs):
        perm = torch.randperm(n_train, device=dev, generator=gen)
        run = 0.0
        for k in range(steps_per_epoch):
            idx = perm[k * batch:(k + 1) * batch]
            xb, yb = xtr[idx].contiguous(), ytr[idx].contiguous()
            lr = lr_min + 0.5 * (lr_max - lr_min) * (1 + math.cos(math.pi * step / total))
            sdf.train_step(xb, yb, lr, CLAMP_UU, over_weight, loss)
            step += 1
            if k % 64 == 0:
                run = float(loss.item())
        if log:
            log(f"epoch {ep + 1}/{epochs}: batch loss {run * 100:.2f} uu (weighted L1)")
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    rep = validate(sdf, xva, yva)
    rep.update({"train_points": n_train, "val_points": n_val, "batch": batch, "epochs": epochs, "steps": step,
                "over_weight": over_weight,
                "label_seconds": t_label, "train_seconds": t_train})
    return rep


# Bounds on a body's motion in one tick the tick kernels assume (uu): ball, car.
STEP_BALL_UU, STEP_CAR_UU = 100.0, 50.0
SCHEDULES = (1, 2, 4, 8)


def cull_margins(pred, labels, reach, step, every):
    """Margin for one body kind and schedule: max over k < every of
    max(pred - R_k) over points with exact distance <= R_k = reach + k * step,
    x1.25 + 2 uu (held-out points stand in for the whole field)."""
    worst = 0.0
    for k in range(every):
        rk = reach + k * step
        sel = labels <= rk
        if bool(sel.any()):
            worst = max(worst, float((pred[sel] - rk).max().item()))
    return worst * 1.25 + 2.0


def validate(sdf, pts, labels):
    """Errors of the network on held-out points against min(exact, CLAMP_UU);
    the largest over-estimate anywhere ("margin_uu", the bound without regard
    to distance) and the per-schedule margins of the ball and the cars."""
    import torch
    pred = torch.cat([sdf.eval(pts[s:s + (1 << 20)]) for s in range(0, pts.shape[0], 1 << 20)])
    t = torch.clamp(labels, max=CLAMP_UU)
    err = (pred - t)
    inside = labels < CLAMP_UU
    over = err[inside].max().item()
    a = err.abs()
    rb, rc = sdf.reach()
    return {"max_abs_err_uu": a.max().item(), "p99_abs_err_uu": torch.quantile(a[:1 << 20], 0.99).item(),
            "mean_abs_err_uu": a.mean().item(), "max_overestimate_uu": over,
            "margin_uu": max(0.0, over) * 1.25 + 2.0,
            "schedules": list(SCHEDULES),
            "margin_ball_uu": [cull_margins(pred, labels, rb, STEP_BALL_UU, e) for e in SCHEDULES],
            "margin_car_uu": [cull_margins(pred, labels, rc, STEP_CAR_UU, e) for e in SCHEDULES]}


def margins_for(meta, every, refined=False):
    """(ball, car) margins to attach the stored weights with.  By default the
    largest held-out over-estimate anywhere, x1.25 + 2 uu, for both bodies:
    the per-schedule margins (refined=True) rest on held-out points near the
    surface only, and the audit has caught a miss with them at a schedule of
    2 ticks, so they are reported but not used by default."""
    if not refined:
        m = float(meta["margin_uu"])
        return m, m
    sch = [int(x) for x in np.atleast_1d(meta["schedules"])]
    if every > sch[-1]:
        raise ValueError(f"no margin stored for a schedule of {every} ticks")
    i = next(k for k, e in enumerate(sch) if e >= every)
    return float(np.atleast_1d(meta["margin_ball_uu"])[i]), float(np.atleast_1d(meta["margin_car_uu"])[i])


def cull_rates(sdf, batch_states_points, margin_ball_uu, margin_car_uu):
    """Fraction of query points the cull drops right after the network ran:
    dict with 'ball' and 'car' given (n, 3) CUDA tensors of ball centres and
    car box centres (uu)."""
    rb, rc = sdf.reach()
    out = {}
    for name, pts, reach, m in (("ball", batch_states_points[0], rb, margin_ball_uu),
                                ("car", batch_states_points[1], rc, margin_car_uu)):
        d = sdf.eval(pts)
        out[name] = float(((d - m) > reach).float().mean().item())
    return out
