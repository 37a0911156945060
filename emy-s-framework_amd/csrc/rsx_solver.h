// rsx_solver.h — the sequential-impulse contact solver of one tick, with the
// RocketSim modifications to Bullet 3.24:
//   * split impulse always on (penetration threshold 1e30, erp2 0.8),
//   * no positional term folded into the velocity rhs,
//   * ball-world points marked "special": their normal rows only take part in
//     the split-impulse (position) pass, and one averaged row per body carries
//     the velocity response.
// Reference: btSequentialImpulseConstraintSolver.cpp:45-105 (row solvers),
// :283-313 (split impulse row), :458-491 (initSolverBody), :517-606
// (friction rows), :795-983 (contact rows), :1003-1211 (convertContact,
// convertContactInner, convertContactSpecial), :1422-1599 (setup),
// :1601-1829 (iterations), :1831-1927 (write-back); btSolverBody.h:130-300.
#pragma once

#include "rsx_collide.h"

namespace rsx {

struct SBody {
    V3 dlin, dang, push, turn;
    V3 lin, ang, ext_f, ext_t;
    float inv_mass;
    M3 inv_i;
    Xf xf;
    bool dyn;        // has an original rigid body with mass
    bool used;
};

struct Row {
    V3 n1, rc1, n2, rc2, angA, angB;
    float applied, push, rhs, rhs_pen, jac_inv, lo, hi, friction;
    int8_t a, b;
    bool special;
    int16_t fidx;
};

constexpr int kMaxRows = 40;

template <int MAXC>
struct Solver {
    SBody body[MAXC + 2];  // 0 ball, 1.. cars, MAXC+1 fixed
    Row row[kMaxRows];
    Row fric[kMaxRows];
    int nrow, nfric;
    // per-body special accumulation (btCollisionObject::m_specialResolveInfo)
    int nspecial[MAXC + 1];
    float sp_friction[MAXC + 1], sp_restitution[MAXC + 1], sp_dist[MAXC + 1];
    V3 sp_normal[MAXC + 1];
};

RSX_HD void sbody_init_fixed(SBody& s) {
    V3 z = V3{0, 0, 0};
    s.dlin = s.dang = s.push = s.turn = z;
    s.lin = s.ang = s.ext_f = s.ext_t = z;
    s.inv_mass = 0;
    s.inv_i = M3{{z, z, z}};
    s.xf = xf_identity();
    s.dyn = false;
    s.used = true;
}

RSX_HD void sbody_init(SBody& s, const Xf& xf, V3 lin, V3 ang, V3 force, V3 torque, float inv_mass, const M3& inv_i,
                       float dt) {
    V3 z = V3{0, 0, 0};
    s.dlin = s.dang = s.push = s.turn = z;
    s.xf = xf;
    s.inv_mass = inv_mass;
    s.inv_i = inv_i;
    s.lin = lin;
    s.ang = ang;
    s.ext_f = force * inv_mass * dt;
    s.ext_t = torque * inv_i * dt;
    s.dyn = true;
    s.used = true;
}

RSX_HD float restitution_curve(float rel_vel, float restitution) {
    if (fabsf(rel_vel) < kRestitutionVelThreshold) return 0.f;
    return restitution * -rel_vel;
}

template <int MAXC>
RSX_HD void setup_contact_row(Solver<MAXC>& S, Row& out, int ia, int ib, V3 normal, float dist, float friction,
                              float restitution, bool special, V3 rel1, V3 rel2, float dt) {
    // the bodies read before the row is written, the row built in registers
    // and stored once (row and body stores cannot then reorder the reads)
    const SBody A = S.body[ia];
    const SBody B = S.body[ib];
    Row r{};
    r.a = (int8_t)ia;
    r.b = (int8_t)ib;
    const float relaxation = 1.f;
    float inv_dt = 1.f / dt;
    const float erp = kErp2;
    V3 ta = cross(rel1, normal);
    r.angA = A.dyn ? (A.inv_i * ta) : V3{0, 0, 0};
    V3 tb = cross(rel2, normal);
    r.angB = B.dyn ? (B.inv_i * (-tb)) : V3{0, 0, 0};
    float d0 = 0.f, d1 = 0.f;
    if (A.dyn) d0 = A.inv_mass + dot(normal, cross(r.angA, rel1));
    if (B.dyn) d1 = B.inv_mass + dot(normal, cross(-r.angB, rel2));
    r.jac_inv = relaxation / (d0 + d1 + 0.f);
    if (A.dyn) {
        r.n1 = normal;
        r.rc1 = ta;
    } else {
        r.n1 = r.rc1 = V3{0, 0, 0};
    }
    if (B.dyn) {
        r.n2 = -normal;
        r.rc2 = -tb;
    } else {
        r.n2 = r.rc2 = V3{0, 0, 0};
    }
    float penetration = dist + 0.f;
    {
        V3 v1 = A.dyn ? (A.lin + cross(A.ang, rel1)) : V3{0, 0, 0};
        V3 v2 = B.dyn ? (B.lin + cross(B.ang, rel2)) : V3{0, 0, 0};
        float rv = dot(normal, v1 - v2);
        r.friction = friction;
        float rest = restitution_curve(rv, restitution);
        restitution = rest <= 0.f ? 0.f : rest;
    }
    r.applied = 0.f;  // warm start: fresh manifolds carry no impulse
    r.push = 0.f;
    {
        V3 efa = A.dyn ? A.ext_f : V3{0, 0, 0};
        V3 eta = A.dyn ? A.ext_t : V3{0, 0, 0};
        V3 efb = B.dyn ? B.ext_f : V3{0, 0, 0};
        V3 etb = B.dyn ? B.ext_t : V3{0, 0, 0};
        float v1 = dot(r.n1, A.lin + efa) + dot(r.rc1, A.ang + eta);
        float v2 = dot(r.n2, B.lin + efb) + dot(r.rc2, B.ang + etb);
        float rv = v1 + v2;
        float pos_err = 0.f;
        float vel_err = restitution - rv;
        if (penetration > 0) pos_err = 0;
        else pos_err = -penetration * erp * inv_dt;
        r.rhs = vel_err * r.jac_inv;
        r.rhs_pen = pos_err * r.jac_inv;
        r.lo = 0;
        r.hi = 1e10f;
    }
    r.special = special;
    out = r;
}

template <int MAXC>
RSX_HD void add_friction_row(Solver<MAXC>& S, int& nfric, V3 dir, int ia, int ib, float friction, V3 rel1, V3 rel2,
                             int fidx) {
    if (nfric >= kMaxRows) return;
    Row& out = S.fric[nfric++];
    const SBody A = S.body[ia];
    const SBody B = S.body[ib];
    Row r{};
    r.a = (int8_t)ia;
    r.b = (int8_t)ib;
    r.fidx = (int16_t)fidx;
    r.friction = friction;
    r.applied = 0.f;
    r.push = 0.f;
    r.special = false;
    if (A.dyn) {
        r.n1 = dir;
        r.rc1 = cross(rel1, r.n1);
        r.angA = A.inv_i * r.rc1;
    } else {
        r.n1 = r.rc1 = r.angA = V3{0, 0, 0};
    }
    if (B.dyn) {
        r.n2 = -dir;
        r.rc2 = cross(rel2, r.n2);
        r.angB = B.inv_i * r.rc2;
    } else {
        r.n2 = r.rc2 = r.angB = V3{0, 0, 0};
    }
    float d0 = 0.f, d1 = 0.f;
    if (A.dyn) d0 = A.inv_mass + dot(dir, cross(r.angA, rel1));
    if (B.dyn) d1 = B.inv_mass + dot(dir, cross(-r.angB, rel2));
    r.jac_inv = 1.f / (d0 + d1);
    float v1 = dot(r.n1, A.dyn ? A.lin + A.ext_f : V3{0, 0, 0}) + dot(r.rc1, A.dyn ? A.ang : V3{0, 0, 0});
    float v2 = dot(r.n2, B.dyn ? B.lin + B.ext_f : V3{0, 0, 0}) + dot(r.rc2, B.dyn ? B.ang : V3{0, 0, 0});
    float rv = v1 + v2;
    float vel_err = 0.f - rv;
    r.rhs = 0.f + vel_err * r.jac_inv;
    r.rhs_pen = 0.f;
    r.lo = -friction;
    r.hi = friction;
    out = r;
}

// convertContactInner without rolling friction (all rolling frictions are 0)
template <int MAXC>
RSX_HD void add_contact_friction(Solver<MAXC>& S, int& nfric, int ia, int ib, V3 normal, float friction, V3 rel1,
                                 V3 rel2, int fidx) {
    SBody& A = S.body[ia];
    SBody& B = S.body[ib];
    V3 v1 = A.dyn ? (A.lin + A.ext_f + cross(A.ang + A.ext_t, rel1)) : V3{0, 0, 0};
    V3 v2 = B.dyn ? (B.lin + B.ext_f + cross(B.ang + B.ext_t, rel2)) : V3{0, 0, 0};
    V3 vel = v1 - v2;
    float rv = dot(normal, vel);
    V3 lat = vel - normal * rv;
    float l2 = len2(lat);
    if (l2 > kEps) {
        lat *= 1.f / sqrtf(l2);
        add_friction_row(S, nfric, lat, ia, ib, friction, rel1, rel2, fidx);
    } else {
        V3 p, q;
        plane_space(normal, p, q);
        add_friction_row(S, nfric, p, ia, ib, friction, rel1, rel2, fidx);
    }
}

RSX_HD void sbody_apply(SBody& s, V3 lin_comp, V3 ang_comp, float imp) {
    if (s.dyn) {
        s.dlin += lin_comp * imp;
        s.dang += ang_comp * imp;
    }
}
RSX_HD void sbody_apply_push(SBody& s, V3 lin_comp, V3 ang_comp, float imp) {
    if (s.dyn) {
        s.push += lin_comp * imp;
        s.turn += ang_comp * imp;
    }
}

template <int MAXC>
RSX_HD float solve_split(Solver<MAXC>& S, Row& c) {
    float delta = 0.f;
    if (c.rhs_pen != 0.f) {
        SBody& A = S.body[c.a];
        SBody& B = S.body[c.b];
        delta = c.rhs_pen - c.push * 0.f;
        const float dv1 = dot(c.n1, A.push) + dot(c.rc1, A.turn);
        const float dv2 = dot(c.n2, B.push) + dot(c.rc2, B.turn);
        delta -= dv1 * c.jac_inv;
        delta -= dv2 * c.jac_inv;
        const float sum = c.push + delta;
        if (sum < c.lo) {
            delta = c.lo - c.push;
            c.push = c.lo;
        } else {
            c.push = sum;
        }
        sbody_apply_push(A, c.n1 * A.inv_mass, c.angA, delta);
        sbody_apply_push(B, c.n2 * B.inv_mass, c.angB, delta);
    }
    return (float)((double)delta * (1. / (double)c.jac_inv));
}

template <int MAXC>
RSX_HD void solve_row(Solver<MAXC>& S, Row& c, bool upper) {
    SBody& A = S.body[c.a];
    SBody& B = S.body[c.b];
    float delta = c.rhs - c.applied * 0.f;
    const float dv1 = dot(c.n1, A.dlin) + dot(c.rc1, A.dang);
    const float dv2 = dot(c.n2, B.dlin) + dot(c.rc2, B.dang);
    delta -= dv1 * c.jac_inv;
    delta -= dv2 * c.jac_inv;
    const float sum = c.applied + delta;
    if (sum < c.lo) {
        delta = c.lo - c.applied;
        c.applied = c.lo;
    } else if (upper && sum > c.hi) {
        delta = c.hi - c.applied;
        c.applied = c.hi;
    } else {
        c.applied = sum;
    }
    sbody_apply(A, c.n1 * A.inv_mass, c.angA, delta);
    sbody_apply(B, c.n2 * B.inv_mass, c.angB, delta);
}

// One Gauss-Seidel sweep over rows[0, n): solve(home, c) solves register
// copy c of row `home` and writes back the impulses it changed.  The next
// row's copy is loaded before the current one is solved (a row changes only
// its own impulses), so the rows' private-memory loads overlap the previous
// row's solve.
template <typename Solve>
RSX_HD void sweep_rows(Row* rows, int n, Solve&& solve) {
    if (n <= 0) return;
    Row cur = rows[0];
    for (int j = 0; j < n; j++) {
        const Row nxt = rows[j + 1 < n ? j + 1 : j];
        solve(rows[j], cur);
        cur = nxt;
    }
}

template <int MAXC>
RSX_HD void solver_iterate(Solver<MAXC>& S) {
    for (int it = 0; it < kSolverIterations; it++) {
        float res = 0.f;
        for (int j = 0; j < S.nrow; j++) {
            float r = solve_split(S, S.row[j]);
            res = maxf(res, r * r);
        }
        if (res <= 0.f || it >= (kSolverIterations - 1)) break;
    }
    for (int it = 0; it < kSolverIterations; it++) {
        for (int j = 0; j < S.nrow; j++) {
            if (S.row[j].special) continue;
            solve_row(S, S.row[j], false);
        }
        for (int j = 0; j < S.nfric; j++) {
            Row& f = S.fric[j];
            float total = S.row[f.fidx].applied;
            if (total > 0.f) {
                f.lo = -(f.friction * total);
                f.hi = f.friction * total;
                solve_row(S, f, true);
            }
        }
    }
}

// ---- iterations with the bodies' hot state in LDS ---------------------------
// The iterations read and update, per row, the two bodies' delta velocities
// (or split-impulse push and turn), inverse mass and dynamic flag.  Kept in
// private memory, every row waits for the previous row's stores to come back
// from the cache; in LDS a lane's accesses complete in order at LDS latency.
// Each lane owns a column of the wavefront's LDS: field f of body b at
// col[(b * kHotFields + f) * 64].  The arithmetic is the same, in the same
// order, as solve_split / solve_row on the private bodies.
constexpr int kHotFields = 14;  // dlin 3, dang 3, push 3, turn 3, inv_mass, dyn
template <int MAXC>
constexpr int solver_lds_bytes() {
    return (MAXC + 2) * kHotFields * 64 * (int)sizeof(float);
}

struct HotBodies {
    float* col;
    RSX_HD float& f(int b, int k) const { return col[(b * kHotFields + k) * 64]; }
    RSX_HD V3 v(int b, int k) const { return V3{f(b, k), f(b, k + 1), f(b, k + 2)}; }
    RSX_HD void set(int b, int k, V3 x) const {
        f(b, k) = x.x;
        f(b, k + 1) = x.y;
        f(b, k + 2) = x.z;
    }
};
constexpr int kHotDlin = 0, kHotDang = 3, kHotPush = 6, kHotTurn = 9, kHotInvMass = 12, kHotDyn = 13;

RSX_HD void hot_apply(const HotBodies& H, int b, int lin_k, int ang_k, V3 lin_comp, V3 ang_comp, float imp) {
    if (H.f(b, kHotDyn) != 0.f) {
        H.set(b, lin_k, H.v(b, lin_k) + lin_comp * imp);
        H.set(b, ang_k, H.v(b, ang_k) + ang_comp * imp);
    }
}

RSX_HD float hot_solve_split(const HotBodies& H, Row& c) {
    float delta = 0.f;
    if (c.rhs_pen != 0.f) {
        delta = c.rhs_pen - c.push * 0.f;
        const float dv1 = dot(c.n1, H.v(c.a, kHotPush)) + dot(c.rc1, H.v(c.a, kHotTurn));
        const float dv2 = dot(c.n2, H.v(c.b, kHotPush)) + dot(c.rc2, H.v(c.b, kHotTurn));
        delta -= dv1 * c.jac_inv;
        delta -= dv2 * c.jac_inv;
        const float sum = c.push + delta;
        if (sum < c.lo) {
            delta = c.lo - c.push;
            c.push = c.lo;
        } else {
            c.push = sum;
        }
        hot_apply(H, c.a, kHotPush, kHotTurn, c.n1 * H.f(c.a, kHotInvMass), c.angA, delta);
        hot_apply(H, c.b, kHotPush, kHotTurn, c.n2 * H.f(c.b, kHotInvMass), c.angB, delta);
    }
    return (float)((double)delta * (1. / (double)c.jac_inv));
}

RSX_HD void hot_solve_row(const HotBodies& H, Row& c, bool upper) {
    float delta = c.rhs - c.applied * 0.f;
    const float dv1 = dot(c.n1, H.v(c.a, kHotDlin)) + dot(c.rc1, H.v(c.a, kHotDang));
    const float dv2 = dot(c.n2, H.v(c.b, kHotDlin)) + dot(c.rc2, H.v(c.b, kHotDang));
    delta -= dv1 * c.jac_inv;
    delta -= dv2 * c.jac_inv;
    const float sum = c.applied + delta;
    if (sum < c.lo) {
        delta = c.lo - c.applied;
        c.applied = c.lo;
    } else if (upper && sum > c.hi) {
        delta = c.hi - c.applied;
        c.applied = c.hi;
    } else {
        c.applied = sum;
    }
    hot_apply(H, c.a, kHotDlin, kHotDang, c.n1 * H.f(c.a, kHotInvMass), c.angA, delta);
    hot_apply(H, c.b, kHotDlin, kHotDang, c.n2 * H.f(c.b, kHotInvMass), c.angB, delta);
}

// solver_iterate with the used bodies' hot state copied into the lane's LDS
// column and back.
template <int MAXC>
RSX_HD void solver_iterate_lds(Solver<MAXC>& S, float* col) {
    const HotBodies H{col};
    for (int b = 0; b < MAXC + 2; b++) {
        const SBody& s = S.body[b];
        if (!s.used) continue;
        H.set(b, kHotDlin, s.dlin);
        H.set(b, kHotDang, s.dang);
        H.set(b, kHotPush, s.push);
        H.set(b, kHotTurn, s.turn);
        H.f(b, kHotInvMass) = s.inv_mass;
        H.f(b, kHotDyn) = s.dyn ? 1.f : 0.f;
    }
    const int nr = S.nrow, nf = S.nfric;
    for (int it = 0; it < kSolverIterations; it++) {
        float res = 0.f;
        sweep_rows(S.row, nr, [&](Row& home, Row& c) {
            const float r = hot_solve_split(H, c);
            home.push = c.push;
            res = maxf(res, r * r);
        });
        if (res <= 0.f || it >= (kSolverIterations - 1)) break;
    }
    for (int it = 0; it < kSolverIterations; it++) {
        sweep_rows(S.row, nr, [&](Row& home, Row& c) {
            if (c.special) return;
            hot_solve_row(H, c, false);
            home.applied = c.applied;
        });
        sweep_rows(S.fric, nf, [&](Row& home, Row& c) {
            const float total = S.row[c.fidx].applied;
            if (total > 0.f) {
                c.lo = -(c.friction * total);
                c.hi = c.friction * total;
                hot_solve_row(H, c, true);
                home.lo = c.lo;
                home.hi = c.hi;
                home.applied = c.applied;
            }
        });
    }
    for (int b = 0; b < MAXC + 2; b++) {
        SBody& s = S.body[b];
        if (!s.used) continue;
        s.dlin = H.v(b, kHotDlin);
        s.dang = H.v(b, kHotDang);
        s.push = H.v(b, kHotPush);
        s.turn = H.v(b, kHotTurn);
    }
}

// btSolverBody::writebackVelocityAndTransform + writeBackBodies
RSX_HD void sbody_writeback(SBody& s, bool no_rot, float dt, Xf& xf_out, V3& lin_out, V3& ang_out) {
    s.lin += s.dlin;
    s.ang += s.dang;
    if (s.push.x != 0.f || s.push.y != 0 || s.push.z != 0 || s.turn.x != 0.f || s.turn.y != 0 || s.turn.z != 0) {
        if (no_rot) {
            Xf n;
            n.o = s.xf.o + s.push * dt;
            n.b = s.xf.b;
            s.xf = n;
        } else {
            s.xf = integrate_transform(s.xf, s.push, s.turn * kSplitImpulseTurnErp, dt);
        }
    }
    lin_out = s.lin + s.ext_f;
    ang_out = s.ang + s.ext_t;
    xf_out = s.xf;
}

}  // namespace rsx
