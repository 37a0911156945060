// rsx_boxbox.h — car box vs car box contacts: Bullet's btBoxBoxDetector (the
// Open Dynamics Engine dBoxBox separating-axis test with face clipping and
// edge-edge closest approach), which RocketSim reaches for car pairs through
// its compound-shape algorithms.
//
// Reference behaviour restated here (bullet3-3.24/BulletCollision/CollisionDispatch/):
//   dLineClosestApproach, intersectRectQuad2, cullPoints2, dBoxBox2,
//   btBoxBoxDetector::getClosestPoints          btBoxBoxDetector.cpp:60-767
//   btBoxBoxCollisionAlgorithm::processCollision btBoxBoxCollisionAlgorithm.cpp:45-80
//
// Matrices are Bullet bases (rows); the detector's "R[i*4+j]" element is row
// i, column j, and its column dot products keep the x, y, z summation order.
//
// This file is an altered restatement of zlib-licensed upstream code, kept
// with its notice as that licence requires (the alterations: one GPU lane per
// query, explicit stacks and index-linked pools instead of recursion and
// pointers, the engine's own types):
//
//   Box-Box collision detection re-distributed under the ZLib license with
//   permission from Russell L. Smith.  Original version is from Open Dynamics
//   Engine, Copyright (C) 2001,2002 Russell L. Smith.  All rights reserved.
//   Bullet Continuous Collision Detection and Physics Library, Bullet is
//   Copyright (c) 2003-2006 Erwin Coumans  https://bulletphysics.org
//
//   This software is provided 'as-is', without any express or implied
//   warranty.  In no event will the authors be held liable for any damages
//   arising from the use of this software.  Permission is granted to anyone
//   to use this software for any purpose, including commercial applications,
//   and to alter it and redistribute it freely, subject to the following
//   restrictions:
//   1. The origin of this software must not be misrepresented; you must not
//      claim that you wrote the original software.  If you use this software
//      in a product, an acknowledgment in the product documentation would be
//      appreciated but is not required.
//   2. Altered source versions must be plainly marked as such, and must not
//      be misrepresented as being the original software.
//   3. This notice may not be removed or altered from any source distribution.
#pragma once

#include "rsx_math.h"

namespace rsx {
namespace boxbox {

struct Out {
    int n;
    V3 normal_on_b[4], point[4];
    float depth[4];
};

RSX_HD void add(Out& o, V3 n, V3 p, float d) {
    if (o.n < 4) {
        o.normal_on_b[o.n] = n;
        o.point[o.n] = p;
        o.depth[o.n] = d;
        o.n++;
    }
}

RSX_HD float rc(const M3& R, int i, int j) { return idx(R.r[i], j); }
// column j of R dotted with vector v (dDOT41 / dDOT14)
RSX_HD float cdot(const M3& R, int j, V3 v) { return rc(R, 0, j) * v.x + rc(R, 1, j) * v.y + rc(R, 2, j) * v.z; }
// column a of R1 dotted with column b of R2 (dDOT44)
RSX_HD float ccdot(const M3& R1, int a, const M3& R2, int b) {
    return rc(R1, 0, a) * rc(R2, 0, b) + rc(R1, 1, a) * rc(R2, 1, b) + rc(R1, 2, a) * rc(R2, 2, b);
}

RSX_HD void line_closest_approach(V3 pa, V3 ua, V3 pb, V3 ub, float& alpha, float& beta) {
    V3 p = V3{pb.x - pa.x, pb.y - pa.y, pb.z - pa.z};
    float uaub = dot(ua, ub);
    float q1 = dot(ua, p);
    float q2 = -dot(ub, p);
    float d = 1 - uaub * uaub;
    if (d <= 0.0001f) {
        alpha = 0;
        beta = 0;
    } else {
        d = 1.f / d;
        alpha = (q1 + uaub * q2) * d;
        beta = (uaub * q1 + q2) * d;
    }
}

RSX_HD int intersect_rect_quad(const float h[2], float p[8], float ret[16]) {
    int nq = 4, nr = 0;
    float buffer[16];
    float* q = p;
    float* r = ret;
    for (int dir = 0; dir <= 1; dir++) {
        for (int sign = -1; sign <= 1; sign += 2) {
            float* pq = q;
            float* pr = r;
            nr = 0;
            for (int i = nq; i > 0; i--) {
                if (sign * pq[dir] < h[dir]) {
                    pr[0] = pq[0];
                    pr[1] = pq[1];
                    pr += 2;
                    nr++;
                    if (nr & 8) {
                        q = r;
                        goto done;
                    }
                }
                float* nextq = (i > 1) ? pq + 2 : q;
                if ((sign * pq[dir] < h[dir]) ^ (sign * nextq[dir] < h[dir])) {
                    pr[1 - dir] = pq[1 - dir] + (nextq[1 - dir] - pq[1 - dir]) / (nextq[dir] - pq[dir]) * (sign * h[dir] - pq[dir]);
                    pr[dir] = sign * h[dir];
                    pr += 2;
                    nr++;
                    if (nr & 8) {
                        q = r;
                        goto done;
                    }
                }
                pq += 2;
            }
            q = r;
            r = (q == ret) ? buffer : ret;
            nq = nr;
        }
    }
done:
    if (q != ret)
        for (int i = 0; i < nr * 2; i++) ret[i] = q[i];
    return nr;
}

RSX_HD void cull_points(int n, const float p[], int m, int i0, int iret[]) {
    const float kPi = 3.14159265f;  // M__PI
    float a, cx, cy, q;
    if (n == 1) {
        cx = p[0];
        cy = p[1];
    } else if (n == 2) {
        cx = 0.5f * (p[0] + p[2]);
        cy = 0.5f * (p[1] + p[3]);
    } else {
        a = 0;
        cx = 0;
        cy = 0;
        for (int i = 0; i < (n - 1); i++) {
            q = p[i * 2] * p[i * 2 + 3] - p[i * 2 + 2] * p[i * 2 + 1];
            a += q;
            cx += q * (p[i * 2] + p[i * 2 + 2]);
            cy += q * (p[i * 2 + 1] + p[i * 2 + 3]);
        }
        q = p[n * 2 - 2] * p[1] - p[0] * p[n * 2 - 1];
        if (fabsf(a + q) > 1.1920928955078125e-07f)
            a = 1.f / (3.0f * (a + q));
        else
            a = 1e18f;
        cx = a * (cx + q * (p[n * 2 - 2] + p[0]));
        cy = a * (cy + q * (p[n * 2 - 1] + p[1]));
    }
    float A[8];
    for (int i = 0; i < n; i++) A[i] = atan2f(p[i * 2 + 1] - cy, p[i * 2] - cx);
    int avail[8];
    for (int i = 0; i < n; i++) avail[i] = 1;
    avail[i0] = 0;
    iret[0] = i0;
    int* out = iret + 1;
    for (int j = 1; j < m; j++) {
        a = (float)j * (2 * kPi / m) + A[i0];
        if (a > kPi) a -= 2 * kPi;
        float maxdiff = 1e9f, diff;
        *out = i0;
        for (int i = 0; i < n; i++) {
            if (avail[i]) {
                diff = fabsf(A[i] - a);
                if (diff > kPi) diff = 2 * kPi - diff;
                if (diff < maxdiff) {
                    maxdiff = diff;
                    *out = i;
                }
            }
        }
        avail[*out] = 0;
        out++;
    }
}

// dBoxBox2 with maxc = 4; box sizes are full side lengths.  Contacts are
// reported as (normal on B, point on B, signed distance) like
// btDiscreteCollisionDetectorInterface::Result::addContactPoint.
RSX_HD int box_box(V3 p1, const M3& R1, V3 side1, V3 p2, const M3& R2, V3 side2, Out& out) {
    const float fudge_factor = 1.05f;
    V3 normalC = V3{0, 0, 0};
    int normalR = -1;  // column of R1 (0..2) or of R2 (3..5)
    float A[3], B[3];
    int invert_normal, code;
    V3 p = p2 - p1;
    V3 pp = V3{cdot(R1, 0, p), cdot(R1, 1, p), cdot(R1, 2, p)};
    A[0] = side1.x * 0.5f;
    A[1] = side1.y * 0.5f;
    A[2] = side1.z * 0.5f;
    B[0] = side2.x * 0.5f;
    B[1] = side2.y * 0.5f;
    B[2] = side2.z * 0.5f;
    float R11 = ccdot(R1, 0, R2, 0), R12 = ccdot(R1, 0, R2, 1), R13 = ccdot(R1, 0, R2, 2);
    float R21 = ccdot(R1, 1, R2, 0), R22 = ccdot(R1, 1, R2, 1), R23 = ccdot(R1, 1, R2, 2);
    float R31 = ccdot(R1, 2, R2, 0), R32 = ccdot(R1, 2, R2, 1), R33 = ccdot(R1, 2, R2, 2);
    float Q11 = fabsf(R11), Q12 = fabsf(R12), Q13 = fabsf(R13);
    float Q21 = fabsf(R21), Q22 = fabsf(R22), Q23 = fabsf(R23);
    float Q31 = fabsf(R31), Q32 = fabsf(R32), Q33 = fabsf(R33);
    float s = -3.40282347e+38f, s2, l;
    invert_normal = 0;
    code = 0;
#define RSX_TST_FACE(expr1, expr2, norm, cc) \
    s2 = fabsf(expr1) - (expr2);             \
    if (s2 > 0) return 0;                    \
    if (s2 > s) {                            \
        s = s2;                              \
        normalR = (norm);                    \
        invert_normal = ((expr1) < 0);       \
        code = (cc);                         \
    }
    RSX_TST_FACE(pp.x, (A[0] + B[0] * Q11 + B[1] * Q12 + B[2] * Q13), 0, 1);
    RSX_TST_FACE(pp.y, (A[1] + B[0] * Q21 + B[1] * Q22 + B[2] * Q23), 1, 2);
    RSX_TST_FACE(pp.z, (A[2] + B[0] * Q31 + B[1] * Q32 + B[2] * Q33), 2, 3);
    RSX_TST_FACE(cdot(R2, 0, p), (A[0] * Q11 + A[1] * Q21 + A[2] * Q31 + B[0]), 3, 4);
    RSX_TST_FACE(cdot(R2, 1, p), (A[0] * Q12 + A[1] * Q22 + A[2] * Q32 + B[1]), 4, 5);
    RSX_TST_FACE(cdot(R2, 2, p), (A[0] * Q13 + A[1] * Q23 + A[2] * Q33 + B[2]), 5, 6);
#undef RSX_TST_FACE
#define RSX_TST_EDGE(expr1, expr2, n1, n2, n3, cc)        \
    s2 = fabsf(expr1) - (expr2);                          \
    if (s2 > 1.1920928955078125e-07f) return 0;           \
    l = sqrtf((n1) * (n1) + (n2) * (n2) + (n3) * (n3));   \
    if (l > 1.1920928955078125e-07f) {                    \
        s2 /= l;                                          \
        if (s2 * fudge_factor > s) {                      \
            s = s2;                                       \
            normalR = -1;                                 \
            normalC.x = (n1) / l;                         \
            normalC.y = (n2) / l;                         \
            normalC.z = (n3) / l;                         \
            invert_normal = ((expr1) < 0);                \
            code = (cc);                                  \
        }                                                 \
    }
    const float fudge2 = 1.0e-5f;
    Q11 += fudge2;
    Q12 += fudge2;
    Q13 += fudge2;
    Q21 += fudge2;
    Q22 += fudge2;
    Q23 += fudge2;
    Q31 += fudge2;
    Q32 += fudge2;
    Q33 += fudge2;
    RSX_TST_EDGE(pp.z * R21 - pp.y * R31, (A[1] * Q31 + A[2] * Q21 + B[1] * Q13 + B[2] * Q12), 0, -R31, R21, 7);
    RSX_TST_EDGE(pp.z * R22 - pp.y * R32, (A[1] * Q32 + A[2] * Q22 + B[0] * Q13 + B[2] * Q11), 0, -R32, R22, 8);
    RSX_TST_EDGE(pp.z * R23 - pp.y * R33, (A[1] * Q33 + A[2] * Q23 + B[0] * Q12 + B[1] * Q11), 0, -R33, R23, 9);
    RSX_TST_EDGE(pp.x * R31 - pp.z * R11, (A[0] * Q31 + A[2] * Q11 + B[1] * Q23 + B[2] * Q22), R31, 0, -R11, 10);
    RSX_TST_EDGE(pp.x * R32 - pp.z * R12, (A[0] * Q32 + A[2] * Q12 + B[0] * Q23 + B[2] * Q21), R32, 0, -R12, 11);
    RSX_TST_EDGE(pp.x * R33 - pp.z * R13, (A[0] * Q33 + A[2] * Q13 + B[0] * Q22 + B[1] * Q21), R33, 0, -R13, 12);
    RSX_TST_EDGE(pp.y * R11 - pp.x * R21, (A[0] * Q21 + A[1] * Q11 + B[1] * Q33 + B[2] * Q32), -R21, R11, 0, 13);
    RSX_TST_EDGE(pp.y * R12 - pp.x * R22, (A[0] * Q22 + A[1] * Q12 + B[0] * Q33 + B[2] * Q31), -R22, R12, 0, 14);
    RSX_TST_EDGE(pp.y * R13 - pp.x * R23, (A[0] * Q23 + A[1] * Q13 + B[0] * Q32 + B[1] * Q31), -R23, R13, 0, 15);
#undef RSX_TST_EDGE
    if (!code) return 0;
    V3 normal;
    if (normalR >= 0) {
        const M3& Rn = normalR < 3 ? R1 : R2;
        int cn = normalR % 3;
        normal = V3{rc(Rn, 0, cn), rc(Rn, 1, cn), rc(Rn, 2, cn)};
    } else {
        normal = R1 * normalC;
    }
    if (invert_normal) normal = V3{-normal.x, -normal.y, -normal.z};
    const float depth = -s;

    if (code > 6) {
        // edge-edge
        V3 pa = p1;
        for (int j = 0; j < 3; j++) {
            float sign = (cdot(R1, j, normal) > 0) ? 1.0f : -1.0f;
            pa.x += sign * A[j] * rc(R1, 0, j);
            pa.y += sign * A[j] * rc(R1, 1, j);
            pa.z += sign * A[j] * rc(R1, 2, j);
        }
        V3 pb = p2;
        for (int j = 0; j < 3; j++) {
            float sign = (cdot(R2, j, normal) > 0) ? -1.0f : 1.0f;
            pb.x += sign * B[j] * rc(R2, 0, j);
            pb.y += sign * B[j] * rc(R2, 1, j);
            pb.z += sign * B[j] * rc(R2, 2, j);
        }
        float alpha, beta;
        const int ca = (code - 7) / 3, cb = (code - 7) % 3;
        V3 ua = V3{rc(R1, 0, ca), rc(R1, 1, ca), rc(R1, 2, ca)};
        V3 ub = V3{rc(R2, 0, cb), rc(R2, 1, cb), rc(R2, 2, cb)};
        line_closest_approach(pa, ua, pb, ub, alpha, beta);
        pb = V3{pb.x + ub.x * beta, pb.y + ub.y * beta, pb.z + ub.z * beta};
        add(out, -normal, pb, -depth);
        return 1;
    }

    // face: find the incident face of the other box and clip it
    const M3& Ra = code <= 3 ? R1 : R2;
    const M3& Rb = code <= 3 ? R2 : R1;
    const V3 pa = code <= 3 ? p1 : p2;
    const V3 pb = code <= 3 ? p2 : p1;
    const float* Sa = code <= 3 ? A : B;
    const float* Sb = code <= 3 ? B : A;
    V3 normal2 = code <= 3 ? normal : V3{-normal.x, -normal.y, -normal.z};
    V3 nr = V3{cdot(Rb, 0, normal2), cdot(Rb, 1, normal2), cdot(Rb, 2, normal2)};
    V3 anr = V3{fabsf(nr.x), fabsf(nr.y), fabsf(nr.z)};
    int lanr, a1, a2;
    if (anr.y > anr.x) {
        if (anr.y > anr.z) {
            a1 = 0;
            lanr = 1;
            a2 = 2;
        } else {
            a1 = 0;
            a2 = 1;
            lanr = 2;
        }
    } else {
        if (anr.x > anr.z) {
            lanr = 0;
            a1 = 1;
            a2 = 2;
        } else {
            a1 = 0;
            a2 = 1;
            lanr = 2;
        }
    }
    V3 center;
    if (idx(nr, lanr) < 0) {
        center.x = pb.x - pa.x + Sb[lanr] * rc(Rb, 0, lanr);
        center.y = pb.y - pa.y + Sb[lanr] * rc(Rb, 1, lanr);
        center.z = pb.z - pa.z + Sb[lanr] * rc(Rb, 2, lanr);
    } else {
        center.x = pb.x - pa.x - Sb[lanr] * rc(Rb, 0, lanr);
        center.y = pb.y - pa.y - Sb[lanr] * rc(Rb, 1, lanr);
        center.z = pb.z - pa.z - Sb[lanr] * rc(Rb, 2, lanr);
    }
    int codeN = code <= 3 ? code - 1 : code - 4;
    int code1, code2;
    if (codeN == 0) {
        code1 = 1;
        code2 = 2;
    } else if (codeN == 1) {
        code1 = 0;
        code2 = 2;
    } else {
        code1 = 0;
        code2 = 1;
    }
    float quad[8];
    float c1 = cdot(Ra, code1, center), c2 = cdot(Ra, code2, center);
    float m11 = ccdot(Ra, code1, Rb, a1), m12 = ccdot(Ra, code1, Rb, a2);
    float m21 = ccdot(Ra, code2, Rb, a1), m22 = ccdot(Ra, code2, Rb, a2);
    {
        float k1 = m11 * Sb[a1], k2 = m21 * Sb[a1], k3 = m12 * Sb[a2], k4 = m22 * Sb[a2];
        quad[0] = c1 - k1 - k3;
        quad[1] = c2 - k2 - k4;
        quad[2] = c1 - k1 + k3;
        quad[3] = c2 - k2 + k4;
        quad[4] = c1 + k1 + k3;
        quad[5] = c2 + k2 + k4;
        quad[6] = c1 + k1 - k3;
        quad[7] = c2 + k2 - k4;
    }
    float rect[2] = {Sa[code1], Sa[code2]};
    float ret[16];
    int n = intersect_rect_quad(rect, quad, ret);
    if (n < 1) return 0;
    float point[3 * 8];
    float dep[8];
    float det1 = 1.f / (m11 * m22 - m12 * m21);
    m11 *= det1;
    m12 *= det1;
    m21 *= det1;
    m22 *= det1;
    int cnum = 0;
    for (int j = 0; j < n; j++) {
        float k1 = m22 * (ret[j * 2] - c1) - m12 * (ret[j * 2 + 1] - c2);
        float k2 = -m21 * (ret[j * 2] - c1) + m11 * (ret[j * 2 + 1] - c2);
        point[cnum * 3 + 0] = center.x + k1 * rc(Rb, 0, a1) + k2 * rc(Rb, 0, a2);
        point[cnum * 3 + 1] = center.y + k1 * rc(Rb, 1, a1) + k2 * rc(Rb, 1, a2);
        point[cnum * 3 + 2] = center.z + k1 * rc(Rb, 2, a1) + k2 * rc(Rb, 2, a2);
        dep[cnum] = Sa[codeN] - (normal2.x * point[cnum * 3] + normal2.y * point[cnum * 3 + 1] +
                                 normal2.z * point[cnum * 3 + 2]);
        if (dep[cnum] >= 0) {
            ret[cnum * 2] = ret[j * 2];
            ret[cnum * 2 + 1] = ret[j * 2 + 1];
            cnum++;
        }
    }
    if (cnum < 1) return 0;
    int maxc = 4;
    if (maxc > cnum) maxc = cnum;
    if (maxc < 1) maxc = 1;
    if (cnum <= maxc) {
        for (int j = 0; j < cnum; j++) {
            V3 pw = V3{point[j * 3] + pa.x, point[j * 3 + 1] + pa.y, point[j * 3 + 2] + pa.z};
            if (code >= 4) pw = V3{pw.x - normal.x * dep[j], pw.y - normal.y * dep[j], pw.z - normal.z * dep[j]};
            add(out, -normal, pw, -dep[j]);
        }
    } else {
        int i1 = 0;
        float maxdepth = dep[0];
        for (int i = 1; i < cnum; i++) {
            if (dep[i] > maxdepth) {
                maxdepth = dep[i];
                i1 = i;
            }
        }
        int iret[8];
        cull_points(cnum, ret, maxc, i1, iret);
        for (int j = 0; j < maxc; j++) {
            V3 pw = V3{point[iret[j] * 3] + pa.x, point[iret[j] * 3 + 1] + pa.y, point[iret[j] * 3 + 2] + pa.z};
            if (code < 4)
                add(out, -normal, pw, -dep[iret[j]]);
            else
                add(out, -normal, pw - normal * dep[iret[j]], -dep[iret[j]]);
        }
        cnum = maxc;
    }
    return cnum;
}

}  // namespace boxbox
}  // namespace rsx
