// Float math with the operation order of Jolt's SSE code, for host and device.
//
// Every helper states the reference it matches (paths relative to the
// reference's Jolt/ directory).  Compiled with -ffp-contract=off and without
// fast-math, each expression rounds exactly like the reference's SSE build
// with FMA contraction off, so results agree bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define XHD __host__ __device__ __forceinline__

namespace xpbd {

struct V3 { float x, y, z; };

XHD V3 v3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
XHD V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
XHD V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
XHD V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }        // Vec3.inl:518
XHD V3 operator*(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }        // Vec3.inl:535
XHD V3 operator/(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }        // Vec3.inl:552
XHD V3 mulv(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }          // Vec3.inl:500
// Unary minus is 0 - v (Vec3.inl:669-672)
XHD V3 neg(V3 a) { return v3(0.0f - a.x, 0.0f - a.y, 0.0f - a.z); }
// Dot = (x*x' + y*y') + z*z' (Vec3.inl:901-934, ReduceSum)
XHD float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
XHD float len_sq(V3 a) { return dot(a, a); }
XHD float len(V3 a) { return sqrtf(len_sq(a)); }                                  // Vec3.inl:951
// Cross (Vec3.inl:855-863): per component mul, mul, sub
XHD V3 cross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
// _mm_min_ps / _mm_max_ps lane semantics
XHD float smin(float a, float b) { return a < b ? a : b; }
XHD float smax(float a, float b) { return a > b ? a : b; }
XHD V3 vmin(V3 a, V3 b) { return v3(smin(a.x, b.x), smin(a.y, b.y), smin(a.z, b.z)); }
XHD V3 vmax(V3 a, V3 b) { return v3(smax(a.x, b.x), smax(a.y, b.y), smax(a.z, b.z)); }
XHD float comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
// Abs = max(0 - v, v) (Vec3.inl:817-822)
XHD V3 vabs(V3 a) { return vmax(neg(a), a); }
// GetSign: sign bit or'ed onto 1.0 (Vec3.inl:1161-1166)
XHD float sign_bit_one(float v) { return __builtin_signbit(v) ? -1.0f : 1.0f; }
XHD V3 vsign(V3 a) { return v3(sign_bit_one(a.x), sign_bit_one(a.y), sign_bit_one(a.z)); }
// GetLowestComponentIndex (Vec3.inl:809-812)
XHD int lowest_index(V3 a) { return a.x < a.y ? (a.z < a.x ? 2 : 0) : (a.z < a.y ? 2 : 1); }
// Sign(0) = +1 (Math/Math.h:116-119)
XHD float sign_of(float v) { return v < 0.0f ? -1.0f : 1.0f; }

constexpr float kPi = 3.14159265358979323846f;     // JPH_PI

// ACosApproximate (Math/Trigonometry.h:48-65)
XHD float acos_approx(float x)
{
    float ax = fabsf(x);
    float abs_x = ax < 1.0f ? ax : 1.0f;
    float val = sqrtf(1.0f - abs_x) * (kPi / 2 - 0.175394f * abs_x);
    return x < 0.0f ? kPi - val : val;
}

// Column-major 4x4: c[col][row]
struct M44 { float c[4][4]; };

XHD V3 col3(const M44 &m, int i) { return v3(m.c[i][0], m.c[i][1], m.c[i][2]); }

// Mat44::sRotationTranslation (Mat44.inl:85-154; the scalar branch is written to
// equal the SSE4.1 one)
XHD M44 rotation_translation(float qx, float qy, float qz, float qw, V3 t)
{
    float tx = qx + qx, ty = qy + qy, tz = qz + qz;
    float xx = tx * qx, yy = ty * qy, zz = tz * qz;
    float xy = tx * qy, xz = tx * qz, xw = tx * qw;
    float yz = ty * qz, yw = ty * qw, zw = tz * qw;
    M44 m;
    m.c[0][0] = (1.0f - yy) - zz; m.c[0][1] = xy + zw; m.c[0][2] = xz - yw; m.c[0][3] = 0.0f;
    m.c[1][0] = xy - zw; m.c[1][1] = (1.0f - zz) - xx; m.c[1][2] = yz + xw; m.c[1][3] = 0.0f;
    m.c[2][0] = xz + yw; m.c[2][1] = yz - xw; m.c[2][2] = (1.0f - xx) - yy; m.c[2][3] = 0.0f;
    m.c[3][0] = t.x; m.c[3][1] = t.y; m.c[3][2] = t.z; m.c[3][3] = 1.0f;
    return m;
}

// Mat44 * Vec3 (Mat44.inl:295-301): ((c0*x + c1*y) + c2*z) + c3
XHD V3 mul_point(const M44 &m, V3 v)
{
    return v3(((m.c[0][0] * v.x + m.c[1][0] * v.y) + m.c[2][0] * v.z) + m.c[3][0],
              ((m.c[0][1] * v.x + m.c[1][1] * v.y) + m.c[2][1] * v.z) + m.c[3][1],
              ((m.c[0][2] * v.x + m.c[1][2] * v.y) + m.c[2][2] * v.z) + m.c[3][2]);
}

// Mat44::Multiply3x3 (Mat44.inl:384-389)
XHD V3 mul3x3(const M44 &m, V3 v)
{
    return v3((m.c[0][0] * v.x + m.c[1][0] * v.y) + m.c[2][0] * v.z,
              (m.c[0][1] * v.x + m.c[1][1] * v.y) + m.c[2][1] * v.z,
              (m.c[0][2] * v.x + m.c[1][2] * v.y) + m.c[2][2] * v.z);
}

// Mat44::Multiply3x3Transposed (Mat44.inl:423-426) = Transposed3x3().Multiply3x3()
XHD V3 mul3x3_transposed(const M44 &m, V3 v)
{
    return v3((m.c[0][0] * v.x + m.c[0][1] * v.y) + m.c[0][2] * v.z,
              (m.c[1][0] * v.x + m.c[1][1] * v.y) + m.c[1][2] * v.z,
              (m.c[2][0] * v.x + m.c[2][1] * v.y) + m.c[2][2] * v.z);
}

// Mat44 * Mat44 (Mat44.inl:241-252)
XHD M44 mul(const M44 &a, const M44 &b)
{
    M44 r;
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 4; ++k)
            r.c[i][k] = ((a.c[0][k] * b.c[i][0] + a.c[1][k] * b.c[i][1]) + a.c[2][k] * b.c[i][2]) + a.c[3][k] * b.c[i][3];
    return r;
}

// InversedRotationTranslation (Mat44.inl:944-949)
XHD M44 inversed_rotation_translation(const M44 &m)
{
    M44 r;
    for (int i = 0; i < 3; ++i)
    {
        for (int k = 0; k < 3; ++k)
            r.c[i][k] = m.c[k][i];
        r.c[i][3] = 0.0f;
    }
    r.c[3][0] = 0.0f; r.c[3][1] = 0.0f; r.c[3][2] = 0.0f; r.c[3][3] = 1.0f;
    V3 t = neg(mul3x3(r, col3(m, 3)));
    r.c[3][0] = t.x; r.c[3][1] = t.y; r.c[3][2] = t.z;
    return r;
}

// Mat44::GetQuaternion (Mat44.inl:998-1050)
XHD void get_quaternion(const M44 &m, float q[4])
{
    float tr = m.c[0][0] + m.c[1][1] + m.c[2][2];
    if (tr >= 0.0f)
    {
        float s = sqrtf(tr + 1.0f);
        float is = 0.5f / s;
        q[0] = (m.c[1][2] - m.c[2][1]) * is; q[1] = (m.c[2][0] - m.c[0][2]) * is;
        q[2] = (m.c[0][1] - m.c[1][0]) * is; q[3] = 0.5f * s;
        return;
    }
    int i = 0;
    if (m.c[1][1] > m.c[0][0]) i = 1;
    // m(i, i) by selection: an indexed read would put the matrix on the stack
    if (m.c[2][2] > (i == 1 ? m.c[1][1] : m.c[0][0])) i = 2;
    if (i == 0)
    {
        float s = sqrtf(m.c[0][0] - (m.c[1][1] + m.c[2][2]) + 1.0f);
        float is = 0.5f / s;
        q[0] = 0.5f * s; q[1] = (m.c[1][0] + m.c[0][1]) * is; q[2] = (m.c[0][2] + m.c[2][0]) * is; q[3] = (m.c[1][2] - m.c[2][1]) * is;
    }
    else if (i == 1)
    {
        float s = sqrtf(m.c[1][1] - (m.c[2][2] + m.c[0][0]) + 1.0f);
        float is = 0.5f / s;
        q[0] = (m.c[1][0] + m.c[0][1]) * is; q[1] = 0.5f * s; q[2] = (m.c[2][1] + m.c[1][2]) * is; q[3] = (m.c[2][0] - m.c[0][2]) * is;
    }
    else
    {
        float s = sqrtf(m.c[2][2] - (m.c[0][0] + m.c[1][1]) + 1.0f);
        float is = 0.5f / s;
        q[0] = (m.c[0][2] + m.c[2][0]) * is; q[1] = (m.c[2][1] + m.c[1][2]) * is; q[2] = 0.5f * s; q[3] = (m.c[0][1] - m.c[1][0]) * is;
    }
}

// Quaternion {x, y, z, w} with the SSE build's operation order (Quat.inl, Vec4.inl).
struct Q4 { float x, y, z, w; };

XHD Q4 q4(float x, float y, float z, float w) { Q4 r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }
XHD Q4 operator+(Q4 a, Q4 b) { return q4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
XHD Q4 operator-(Q4 a, Q4 b) { return q4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
XHD Q4 operator*(Q4 a, float s) { return q4(a.x * s, a.y * s, a.z * s, a.w * s); }        // Quat.h:157-158
XHD Q4 operator/(Q4 a, float s) { return q4(a.x / s, a.y / s, a.z / s, a.w / s); }        // Vec4.inl:602-606
// Conjugated: sign flips of x, y, z (Quat.h:185)
XHD Q4 conjugated(Q4 a) { return q4(-a.x, -a.y, -a.z, a.w); }
// Vec4::Dot = ReduceSum of the products: (x + y) + (z + w) (Vec4.inl:988-1010)
XHD float dot4(Q4 a, Q4 b) { return (a.x * b.x + a.y * b.y) + (a.z * b.z + a.w * b.w); }
// Normalized = v / Length, a division per component (Vec4.inl:1022-1030)
XHD Q4 normalized(Q4 a) { return a / sqrtf(dot4(a, a)); }
// Quat * Quat without FMA (Quat.inl:7-48):
// [(aw+bz)+(dx-cy), (bw+cx)+(dy-az), (cw+ay)+(dz-bx), -(ax+by)+(dw-cz)] for lhs {a,b,c,d}, rhs {x,y,z,w}
XHD Q4 quat_mul(Q4 l, Q4 r)
{
    return q4((l.x * r.w + l.y * r.z) + (l.w * r.x - l.z * r.y),
              (l.y * r.w + l.z * r.x) + (l.w * r.y - l.x * r.z),
              (l.z * r.w + l.x * r.y) + (l.w * r.z - l.y * r.x),
              -(l.x * r.x + l.y * r.y) + (l.w * r.w - l.z * r.z));
}
// sMultiplyImaginary: the quaternion {v, 0} times q (Quat.inl:90-118):
// [(aw+bz)-cy, (bw+cx)-az, (cw+ay)-bx, -(ax+by)-cz]
XHD Q4 mul_imaginary(V3 v, Q4 q)
{
    return q4((v.x * q.w + v.y * q.z) - v.z * q.y,
              (v.y * q.w + v.z * q.x) - v.x * q.z,
              (v.z * q.w + v.x * q.y) - v.y * q.x,
              -(v.x * q.x + v.y * q.y) - v.z * q.z);
}
// Quat * Vec3 (Quat.inl:383-405): p + 2 (w (q x p) + q x (q x p)), each cross product formed
// as the SSE build forms it (a swizzled difference of products)
XHD V3 quat_rotate(Q4 q, V3 p)
{
    const V3 a = v3(p.z * q.y - q.z * p.y, p.x * q.z - q.x * p.z, p.y * q.x - q.y * p.x);
    const V3 b = v3(a.z * q.y - q.z * a.y, a.x * q.z - q.x * a.z, a.y * q.x - q.y * a.x);
    const V3 v = q.w * a + b;
    return p + (v + v);
}
// RotateAxisZ: q * (0, 0, 1) written out (Quat.inl:434-440)
XHD V3 rotate_axis_z(Q4 q)
{
    const float tz = q.z + q.z, tw = q.w + q.w;
    return v3(tz * q.x + tw * q.y, tz * q.y + -(tw * q.x), (tz * q.z + tw * q.w) - 1.0f);
}

// Plane (Geometry/Plane.h): normal + constant
struct Pl { V3 n; float c; };
// sFromPointAndNormal (Plane.h:21)
XHD Pl plane_from_point_normal(V3 p, V3 n) { Pl r; r.n = n; r.c = -dot(n, p); return r; }
// GetTransformed (Plane.h:42-46)
XHD Pl plane_transformed(const Pl &p, const M44 &m)
{
    V3 tn = mul3x3(m, p.n);
    Pl r; r.n = tn; r.c = p.c - dot(col3(m, 3), tn);
    return r;
}
// SignedDistance (Plane.h:57)
XHD float signed_distance(const Pl &p, V3 x) { return dot(x, p.n) + p.c; }

} // namespace xpbd
