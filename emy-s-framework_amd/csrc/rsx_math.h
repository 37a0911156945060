// rsx_math.h — small-vector algebra shared by the gfx950 kernels and the host
// build of the same code.
//
// The arena engine must agree with the CPU RocketSim to 1e-4 relative after a
// tick.  To stay within rounding of Bullet 3.24's LinearMath, every operation
// here evaluates its terms in the order Bullet's scalar path does (for example
// a dot product is ((x*x' + y*y') + z*z'), a matrix-vector product is three row
// dots, a division of a vector by s multiplies by 1/s).  The whole project is
// compiled with floating-point contraction off so no FMA changes the rounding.
// (On x86-64 RocketSim's Bullet takes some SSE paths, e.g. normalize() via a
// refined rsqrt, so agreement is to rounding, not to the bit.)
//
// Layouts follow Bullet's conventions: M3 stores rows; the rotation of a body
// has the body's forward/right/up axes as its COLUMNS; Xf is basis + origin.
// (Reference map: LinearMath/btVector3.h, btMatrix3x3.h, btQuaternion.h,
// btTransform.h, btTransformUtil.h in bullet3-3.24.)
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RSX_HD __host__ __device__ __forceinline__
#else
#define RSX_HD inline
#endif

namespace rsx {

constexpr float kPi = 3.1415926535897932384626433832795029f;
constexpr float kHalfPi = kPi * 0.5f;
constexpr float kEps = 1.1920928955078125e-07f;  // FLT_EPSILON
constexpr float kLargeFloat = 1e18f;
constexpr float kSqrt12 = 0.7071067811865475244008443621048490f;

struct V3 {
    float x, y, z;
};

RSX_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
RSX_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
RSX_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
RSX_HD V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
RSX_HD V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
RSX_HD V3 operator*(float s, V3 a) { return V3{a.x * s, a.y * s, a.z * s}; }
RSX_HD V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
// Bullet divides a vector by s as a multiplication by (1/s).
RSX_HD V3 operator/(V3 a, float s) { return a * (1.0f / s); }
RSX_HD V3& operator+=(V3& a, V3 b) { a = a + b; return a; }
RSX_HD V3& operator-=(V3& a, V3 b) { a = a - b; return a; }
RSX_HD V3& operator*=(V3& a, float s) { a = a * s; return a; }
RSX_HD bool operator==(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

RSX_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RSX_HD V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
RSX_HD float len2(V3 a) { return dot(a, a); }
RSX_HD float len(V3 a) { return sqrtf(len2(a)); }
RSX_HD bool is_zero(V3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
RSX_HD bool fuzzy_zero(V3 a) { return len2(a) < kEps * kEps; }
RSX_HD float idx(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
RSX_HD void set_idx(V3& a, int i, float v) {
    if (i == 0) a.x = v; else if (i == 1) a.y = v; else a.z = v;
}

// btVector3::normalize(): *this /= length()
RSX_HD V3 normalized(V3 a) { return a / len(a); }
// btVector3::safeNormalize()
RSX_HD V3 safe_normalized(V3 a) {
    float l2 = len2(a);
    if (l2 >= kEps * kEps) return a / sqrtf(l2);
    return V3{1.f, 0.f, 0.f};
}
// RocketSim's Vec::Normalized (Math/MathTypes/MathTypes.h:88-95): divide
// each component by the length unless the length is ~0.
RSX_HD V3 rl_norm(V3 v) {
    float l2 = v.x * v.x + v.y * v.y + v.z * v.z;
    float l = l2 > 0 ? sqrtf(l2) : 0.f;
    if (l > 1.1920929e-07f * 1.1920929e-07f) return V3{v.x / l, v.y / l, v.z / l};
    return V3{0, 0, 0};
}
RSX_HD float dist2(V3 a, V3 b) { return len2(b - a); }

RSX_HD V3 vmin(V3 a, V3 b) {
    return V3{a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y, a.z < b.z ? a.z : b.z};
}
RSX_HD V3 vmax(V3 a, V3 b) {
    return V3{a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y, a.z > b.z ? a.z : b.z};
}
RSX_HD V3 vabs(V3 a) { return V3{fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }

// btPlaneSpace1: two unit vectors orthogonal to unit n.
RSX_HD void plane_space(V3 n, V3& p, V3& q) {
    if (fabsf(n.z) > kSqrt12) {
        float a = n.y * n.y + n.z * n.z;
        float k = 1.0f / sqrtf(a);
        p = V3{0.f, -n.z * k, n.y * k};
        q = V3{a * k, -n.x * p.z, n.x * p.y};
    } else {
        float a = n.x * n.x + n.y * n.y;
        float k = 1.0f / sqrtf(a);
        p = V3{-n.y * k, n.x * k, 0.f};
        q = V3{-n.z * p.y, n.z * p.x, a * k};
    }
}

struct M3 {
    V3 r[3];  // rows
};

RSX_HD M3 m3(float xx, float xy, float xz, float yx, float yy, float yz, float zx, float zy, float zz) {
    return M3{{V3{xx, xy, xz}, V3{yx, yy, yz}, V3{zx, zy, zz}}};
}
RSX_HD M3 m3_identity() { return m3(1, 0, 0, 0, 1, 0, 0, 0, 1); }
RSX_HD V3 col(const M3& m, int c) {
    return c == 0 ? V3{m.r[0].x, m.r[1].x, m.r[2].x}
                  : (c == 1 ? V3{m.r[0].y, m.r[1].y, m.r[2].y} : V3{m.r[0].z, m.r[1].z, m.r[2].z});
}
// Body axes as columns: forward = col 0, right = col 1, up = col 2.
RSX_HD M3 m3_from_axes(V3 f, V3 r, V3 u) { return m3(f.x, r.x, u.x, f.y, r.y, u.y, f.z, r.z, u.z); }
RSX_HD V3 operator*(const M3& m, V3 v) { return V3{dot(m.r[0], v), dot(m.r[1], v), dot(m.r[2], v)}; }
RSX_HD float tdotx(const M3& m, V3 v) { return m.r[0].x * v.x + m.r[1].x * v.y + m.r[2].x * v.z; }
RSX_HD float tdoty(const M3& m, V3 v) { return m.r[0].y * v.x + m.r[1].y * v.y + m.r[2].y * v.z; }
RSX_HD float tdotz(const M3& m, V3 v) { return m.r[0].z * v.x + m.r[1].z * v.y + m.r[2].z * v.z; }
// v * M (row vector times matrix)
RSX_HD V3 operator*(V3 v, const M3& m) { return V3{tdotx(m, v), tdoty(m, v), tdotz(m, v)}; }
RSX_HD M3 operator*(const M3& a, const M3& b) {
    return M3{{V3{tdotx(b, a.r[0]), tdoty(b, a.r[0]), tdotz(b, a.r[0])},
               V3{tdotx(b, a.r[1]), tdoty(b, a.r[1]), tdotz(b, a.r[1])},
               V3{tdotx(b, a.r[2]), tdoty(b, a.r[2]), tdotz(b, a.r[2])}}};
}
RSX_HD M3 transpose(const M3& m) {
    return m3(m.r[0].x, m.r[1].x, m.r[2].x, m.r[0].y, m.r[1].y, m.r[2].y, m.r[0].z, m.r[1].z, m.r[2].z);
}
// btMatrix3x3::scaled: M * diag(s)
RSX_HD M3 scaled(const M3& m, V3 s) {
    return m3(m.r[0].x * s.x, m.r[0].y * s.y, m.r[0].z * s.z, m.r[1].x * s.x, m.r[1].y * s.y, m.r[1].z * s.z,
              m.r[2].x * s.x, m.r[2].y * s.y, m.r[2].z * s.z);
}
// btMatrix3x3::transposeTimes: transpose(a) * b
RSX_HD M3 transpose_times(const M3& a, const M3& b) {
    const V3* e = a.r;
    const V3* m = b.r;
    return m3(e[0].x * m[0].x + e[1].x * m[1].x + e[2].x * m[2].x, e[0].x * m[0].y + e[1].x * m[1].y + e[2].x * m[2].y,
              e[0].x * m[0].z + e[1].x * m[1].z + e[2].x * m[2].z, e[0].y * m[0].x + e[1].y * m[1].x + e[2].y * m[2].x,
              e[0].y * m[0].y + e[1].y * m[1].y + e[2].y * m[2].y, e[0].y * m[0].z + e[1].y * m[1].z + e[2].y * m[2].z,
              e[0].z * m[0].x + e[1].z * m[1].x + e[2].z * m[2].x, e[0].z * m[0].y + e[1].z * m[1].y + e[2].z * m[2].y,
              e[0].z * m[0].z + e[1].z * m[1].z + e[2].z * m[2].z);
}
RSX_HD float mget(const M3& m, int i, int j) { return idx(m.r[i], j); }
RSX_HD float cofac(const M3& m, int r1, int c1, int r2, int c2) {
    return mget(m, r1, c1) * mget(m, r2, c2) - mget(m, r1, c2) * mget(m, r2, c1);
}
// btMatrix3x3::inverse (cofactor form)
RSX_HD M3 inverse(const M3& m) {
    V3 co = V3{cofac(m, 1, 1, 2, 2), cofac(m, 1, 2, 2, 0), cofac(m, 1, 0, 2, 1)};
    float det = dot(m.r[0], co);
    float s = 1.0f / det;
    return m3(co.x * s, cofac(m, 0, 2, 2, 1) * s, cofac(m, 0, 1, 1, 2) * s, co.y * s, cofac(m, 0, 0, 2, 2) * s,
              cofac(m, 0, 2, 1, 0) * s, co.z * s, cofac(m, 0, 1, 2, 0) * s, cofac(m, 0, 0, 1, 1) * s);
}

struct Q4 {
    float x, y, z, w;
};
RSX_HD float qlen2(Q4 q) { return q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w; }
RSX_HD Q4 qscale(Q4 q, float s) { return Q4{q.x * s, q.y * s, q.z * s, q.w * s}; }
RSX_HD Q4 qmul(Q4 a, Q4 b) {
    return Q4{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z,
              a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
RSX_HD Q4 qinverse(Q4 q) { return Q4{-q.x, -q.y, -q.z, q.w}; }
// btQuaternion(axis, angle)
RSX_HD Q4 q_axis_angle(V3 axis, float angle) {
    float d = len(axis);
    float s = sinf(angle * 0.5f) / d;
    return Q4{axis.x * s, axis.y * s, axis.z * s, cosf(angle * 0.5f)};
}
// btQuaternion::safeNormalize
RSX_HD Q4 q_safe_normalized(Q4 q) {
    float l2 = qlen2(q);
    if (l2 > kEps) return qscale(q, 1.0f / sqrtf(qlen2(q)));
    return q;
}
// btMatrix3x3::setRotation (scalar path)
RSX_HD M3 m3_from_quat(Q4 q) {
    float d = qlen2(q);
    float s = 2.0f / d;
    float xs = q.x * s, ys = q.y * s, zs = q.z * s;
    float wx = q.w * xs, wy = q.w * ys, wz = q.w * zs;
    float xx = q.x * xs, xy = q.x * ys, xz = q.x * zs;
    float yy = q.y * ys, yz = q.y * zs, zz = q.z * zs;
    return m3(1.0f - (yy + zz), xy - wz, xz + wy, xy + wz, 1.0f - (xx + zz), yz - wx, xz - wy, yz + wx,
              1.0f - (xx + yy));
}
// btMatrix3x3::getRotation (scalar path)
RSX_HD Q4 quat_from_m3(const M3& m) {
    float trace = m.r[0].x + m.r[1].y + m.r[2].z;
    float t[4];
    if (trace > 0.0f) {
        float s = sqrtf(trace + 1.0f);
        t[3] = s * 0.5f;
        s = 0.5f / s;
        t[0] = (m.r[2].y - m.r[1].z) * s;
        t[1] = (m.r[0].z - m.r[2].x) * s;
        t[2] = (m.r[1].x - m.r[0].y) * s;
    } else {
        int i = m.r[0].x < m.r[1].y ? (m.r[1].y < m.r[2].z ? 2 : 1) : (m.r[0].x < m.r[2].z ? 2 : 0);
        int j = (i + 1) % 3;
        int k = (i + 2) % 3;
        float s = sqrtf(mget(m, i, i) - mget(m, j, j) - mget(m, k, k) + 1.0f);
        t[i] = s * 0.5f;
        s = 0.5f / s;
        t[3] = (mget(m, k, j) - mget(m, j, k)) * s;
        t[j] = (mget(m, j, i) + mget(m, i, j)) * s;
        t[k] = (mget(m, k, i) + mget(m, i, k)) * s;
    }
    return Q4{t[0], t[1], t[2], t[3]};
}

// btMatrix3x3::setEulerZYX(eulerX, eulerY, eulerZ)
RSX_HD M3 m3_euler_zyx(float ex, float ey, float ez) {
    float ci = cosf(ex), cj = cosf(ey), ch = cosf(ez);
    float si = sinf(ex), sj = sinf(ey), sh = sinf(ez);
    float cc = ci * ch, cs = ci * sh, sc = si * ch, ss = si * sh;
    return m3(cj * ch, sj * sc - cs, sj * cc + ss, cj * sh, sj * ss + cc, sj * cs - sc, -sj, cj * si, cj * ci);
}

struct Xf {
    M3 b;
    V3 o;
};
RSX_HD Xf xf_identity() { return Xf{m3_identity(), V3{0, 0, 0}}; }
// btTransform::operator()(x): x.dot3(rows) + origin
RSX_HD V3 xform(const Xf& t, V3 x) { return V3{dot(x, t.b.r[0]), dot(x, t.b.r[1]), dot(x, t.b.r[2])} + t.o; }
// btTransform::invXform
RSX_HD V3 inv_xform(const Xf& t, V3 x) { return transpose(t.b) * (x - t.o); }
// btTransform::inverse
RSX_HD Xf xf_inverse(const Xf& t) {
    M3 inv = transpose(t.b);
    return Xf{inv, inv * (-t.o)};
}
// btTransform::operator*(t)
RSX_HD Xf xf_mul(const Xf& a, const Xf& b) { return Xf{a.b * b.b, xform(a, b.o)}; }
// btTransform::inverseTimes(t) = a^-1 * b
RSX_HD Xf xf_inverse_times(const Xf& a, const Xf& b) {
    V3 v = b.o - a.o;
    return Xf{transpose_times(a.b, b.b), v * a.b};
}

// btTransformUtil::integrateTransform (exponential map, angular clamp)
RSX_HD Xf integrate_transform(const Xf& cur, V3 linvel, V3 angvel, float dt) {
    Xf out;
    out.o = cur.o + linvel * dt;
    float angle2 = len2(angvel);
    float angle = 0.f;
    if (angle2 > kEps) angle = sqrtf(angle2);
    const float kMaxAngle = 0.5f * kHalfPi;
    if (angle * dt > kMaxAngle) angle = kMaxAngle / dt;
    V3 axis;
    if (angle < 0.001f) {
        axis = angvel * (0.5f * dt - (dt * dt * dt) * (0.020833333333f) * angle * angle);
    } else {
        axis = angvel * (sinf(0.5f * angle * dt) / angle);
    }
    Q4 dorn = Q4{axis.x, axis.y, axis.z, cosf(angle * dt * 0.5f)};
    Q4 orn0 = quat_from_m3(cur.b);
    Q4 pred = q_safe_normalized(qmul(dorn, orn0));
    if (qlen2(pred) > kEps) {
        out.b = m3_from_quat(pred);
    } else {
        out.b = cur.b;
    }
    return out;
}

// Piecewise-linear curve lookup with RocketSim's semantics
// (RocketSim/src/Math/Math.cpp:5-35, LinearPieceCurve::GetOutput).
struct Curve {
    int n;
    float x[6];
    float y[6];
};
RSX_HD float curve_eval(const Curve& c, float input) {
    if (c.n == 0) return 1.0f;
    if (input <= c.x[0]) return c.y[0];
    for (int i = 1; i < c.n; i++) {
        if (c.x[i] > input) {
            float range = c.x[i] - c.x[i - 1];
            float diff = c.y[i] - c.y[i - 1];
            float f = (input - c.x[i - 1]) / range;
            return c.y[i - 1] + diff * f;
        }
    }
    return c.y[c.n - 1];
}

RSX_HD float clampf(float v, float lo, float hi) {
    // RS_CLAMP(val, min, max) = RS_MIN(RS_MAX(val, min), max)
    float a = (v > lo) ? v : lo;
    return (a < hi) ? a : hi;
}
RSX_HD float maxf(float a, float b) { return (a > b) ? a : b; }  // RS_MAX
RSX_HD float minf(float a, float b) { return (a < b) ? a : b; }  // RS_MIN
RSX_HD int sgnf(float v) { return (v > 0) - (v < 0); }           // RS_SGN

}  // namespace rsx
