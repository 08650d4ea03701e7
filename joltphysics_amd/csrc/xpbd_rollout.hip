// Differentiable rollout of the soft-body update for bodies with edge, dihedral bend, volume and
// LRA constraints: float and double forward passes and the adjoint.
//
// The forward pass follows the update of the main kernels step for step
// (SoftBodyMotionProperties.cpp: InitializeUpdateContext :937-960,
// IntegratePositions :325-350, ApplyEdgeConstraints :573-601, the velocity
// update :700-710 and the velocity clamp of UpdateSoftBodyState :852-857), so
// its float instantiation rounds exactly like them.  One workgroup runs one
// body with its positions in LDS.
//
// Adjoint of one sub-step (w = inverse mass, h = sub-step, d = damping):
//   forward   v1 = w>0 ? (v0 + g h) d : v0,  y0 = x0 + v1 h,  y = C(y0),
//             v2 = w>0 ? (y - x0) / h : v1,  x1 = y   (+ clamp of v2 on the last
//             iteration of a frame)
//   reverse   ybar = x1bar + [w>0] v2bar / h,  y0bar = C^T ybar,
//             x0bar = y0bar - [w>0] v2bar / h,
//             v1bar = [w=0] v2bar + y0bar h,  v0bar = w>0 ? v1bar d : v1bar.
// C is the sequence of edge projections; C^T runs the levels in reverse, each
// edge from its recorded pre-projection positions:
//   delta = x1 - x0, L = |delta|, s = w0 + w1 + alpha / h^2, c = L - rest,
//   corr = delta c / (L s);  x0' = x0 + w0 corr,  x1' = x1 - w1 corr
//   cbar = w0 x0'bar - w1 x1'bar,  n = delta / L
//   deltabar = ((c / L) cbar + (rest / L) n (n . cbar)) / s
//   x0bar = x0'bar - deltabar,  x1bar = x1'bar + deltabar,
//   alphabar += (cbar . corr) (-1 / (h^2 s)).
#include <mutex>

#include "xpbd_device.h"

namespace xpbd {

constexpr int kRolloutBlock = 256;

template <class T> struct R3 { T x, y, z; };
template <class T> struct alignas(sizeof(T) * 4) R4 { T x, y, z, w; };

template <class T> __device__ __forceinline__ R3<T> r3(T x, T y, T z) { R3<T> r; r.x = x; r.y = y; r.z = z; return r; }
template <class T> __device__ __forceinline__ R4<T> r4(R3<T> a, T w) { R4<T> r; r.x = a.x; r.y = a.y; r.z = a.z; r.w = w; return r; }
template <class T> __device__ __forceinline__ R3<T> xyz3(R4<T> p) { return r3(p.x, p.y, p.z); }
template <class T> __device__ __forceinline__ R3<T> operator+(R3<T> a, R3<T> b) { return r3(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> __device__ __forceinline__ R3<T> operator-(R3<T> a, R3<T> b) { return r3(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T> __device__ __forceinline__ R3<T> operator*(R3<T> a, T s) { return r3(a.x * s, a.y * s, a.z * s); }
template <class T> __device__ __forceinline__ R3<T> operator*(T s, R3<T> a) { return r3(s * a.x, s * a.y, s * a.z); }
template <class T> __device__ __forceinline__ R3<T> operator/(R3<T> a, T s) { return r3(a.x / s, a.y / s, a.z / s); }
// Dot = (x*x' + y*y') + z*z' (Vec3.inl:901-934)
template <class T> __device__ __forceinline__ T rdot(R3<T> a, R3<T> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ float rsqrt_exact(float v) { return sqrtf(v); }
__device__ __forceinline__ double rsqrt_exact(double v) { return sqrt(v); }

template <class T> __device__ __forceinline__ R4<T> to_r4(float4 p) { R4<T> r; r.x = (T)p.x; r.y = (T)p.y; r.z = (T)p.z; r.w = (T)p.w; return r; }
__device__ __forceinline__ float4 to_f4(R4<float> p) { return make_float4(p.x, p.y, p.z, p.w); }

// Per-launch constants of InitializeUpdateContext for one collision step per
// frame and identity body rotation (SoftBodyMotionProperties.cpp:937-960).
template <class T>
struct StepConstants
{
    T dt_sub, inv_dt_sq, damping, max_lin_sq, impulse;
    R3<T> sub_step_gravity;

    __device__ StepConstants(const RolloutDev &R, uint32_t nv)
    {
        const T step_dt = (T)R.dt / (T)1;                          // PhysicsSystem.cpp:215, one collision step
        dt_sub = step_dt / (T)R.iterations;
        inv_dt_sq = (T)1 / (dt_sub * dt_sub);
        const T d = (T)1 - (T)R.linear_damping * dt_sub;
        damping = (T)0 > d ? (T)0 : d;
        const T gf = (T)R.gravity_factor;
        const R3<T> g_local = r3(gf * (T)R.gravity[0], gf * (T)R.gravity[1], gf * (T)R.gravity[2]);
        sub_step_gravity = g_local * dt_sub;
        const T nvf = (T)nv;
        impulse = ((T)0 * dt_sub) / (nvf > (T)1 ? nvf : (T)1);    // accumulated force is zero
        max_lin_sq = (T)R.max_linear_velocity * (T)R.max_linear_velocity;
    }

    // IntegratePositions for one vertex: returns v1 (x0, w from p)
    __device__ __forceinline__ R3<T> integrate_velocity(R3<T> v, T w) const
    {
        if (w > (T)0)
        {
            v = v + (sub_step_gravity + r3(impulse * w, impulse * w, impulse * w));
            v = v * damping;
        }
        return v;
    }
};

// ApplyEdgeConstraints body (SoftBodyMotionProperties.cpp:583-600)
template <class T>
__device__ __forceinline__ void edge_project(R4<T> *X, uint32_t i0, uint32_t i1, T rest, T compliance, T inv_dt_sq)
{
    const R4<T> p0 = X[i0], p1 = X[i1];
    const R3<T> x0 = xyz3(p0), x1 = xyz3(p1);
    const T w0 = p0.w, w1 = p1.w;
    const R3<T> delta = x1 - x0;
    const T length = rsqrt_exact(rdot(delta, delta));
    const T denom = length * ((w0 + w1) + compliance * inv_dt_sq);
    if (denom < (T)1.0e-12f)
        return;
    const R3<T> correction = (delta * (length - rest)) / denom;
    X[i0] = r4(x0 + w0 * correction, w0);
    X[i1] = r4(x1 - w1 * correction, w1);
}

// ApplyLRAConstraints body (SoftBodyMotionProperties.cpp:685-697): only vertex 1 moves
template <class T>
__device__ __forceinline__ void lra_project(R4<T> *X, uint32_t i0, uint32_t i1, T max_distance)
{
    const R3<T> x0 = xyz3(X[i0]);
    const R4<T> p1 = X[i1];
    const R3<T> delta = xyz3(p1) - x0;
    const T delta_len_sq = rdot(delta, delta);
    if (delta_len_sq > max_distance * max_distance)
        X[i1] = r4(x0 + (delta * max_distance) / rsqrt_exact(delta_len_sq), p1.w);
}

// Adjoint of lra_project from its pre-projection positions q0, q1: when the constraint acts,
// x1' = x0 + m u with u = delta / |delta|, so with a1 = x1'bar
//   x1bar = (m / |delta|) (a1 - u (u . a1)),  x0bar = x0'bar + a1 - x1bar.
// Vertex 0 is only read in the forward pass, so the LRAs of one level may share it (all LRAs
// of a pinned corner do): their x0bar terms are added atomically.  Nothing else of the level
// touches it.
__device__ __forceinline__ void lra_adjoint(float4 *A, uint32_t i0, uint32_t i1, float4 q0, float4 q1, float max_distance)
{
    const R3<float> delta = r3(q1.x - q0.x, q1.y - q0.y, q1.z - q0.z);
    const float delta_len_sq = rdot(delta, delta);
    if (!(delta_len_sq > max_distance * max_distance))
        return;
    const float length = sqrtf(delta_len_sq);
    const R3<float> u = delta / length;
    const float4 a1 = A[i1];
    const R3<float> ab1 = r3(a1.x, a1.y, a1.z);
    const R3<float> g = (max_distance / length) * (ab1 - rdot(u, ab1) * u);
    A[i1] = make_float4(g.x, g.y, g.z, a1.w);
    atomicAdd(&A[i0].x, ab1.x - g.x);
    atomicAdd(&A[i0].y, ab1.y - g.y);
    atomicAdd(&A[i0].z, ab1.z - g.z);
}

// Record type of the level programs (bits 14-15 of the first word): 3 for an LRA, 0 for a
// volume (vertices 2 and 3 in the second word, as the update kernels unpack them)
__device__ __forceinline__ bool is_lra_record(const uint4 &q) { return ((q.x >> 14) & 3u) == 3u; }
__device__ __forceinline__ bool is_volume_record(const uint4 &q) { return ((q.x >> 14) & 3u) == 0u; }

template <class T> __device__ __forceinline__ R3<T> rcross(R3<T> a, R3<T> b)
{
    return r3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);       // Vec3.inl:855-863
}

// Forward-mode dual numbers: a value and its derivative along one input direction.  The
// adjoint of a dihedral bend is Jacobian-transpose times the output adjoints, its Jacobian
// built column by column by running the projection on duals.
template <class T> struct Dual { T v, d; };
template <class T> __device__ __forceinline__ Dual<T> dual(T v, T d = (T)0) { Dual<T> r; r.v = v; r.d = d; return r; }
template <class T> __device__ __forceinline__ Dual<T> operator+(Dual<T> a, Dual<T> b) { return dual(a.v + b.v, a.d + b.d); }
template <class T> __device__ __forceinline__ Dual<T> operator-(Dual<T> a, Dual<T> b) { return dual(a.v - b.v, a.d - b.d); }
template <class T> __device__ __forceinline__ Dual<T> operator*(Dual<T> a, Dual<T> b) { return dual(a.v * b.v, a.d * b.v + a.v * b.d); }
template <class T> __device__ __forceinline__ Dual<T> operator/(Dual<T> a, Dual<T> b)
{
    return dual(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v));
}
// sqrt on duals; at 0 (a hinge whose cosine rounds to exactly 1, the cusp of ACosApproximate)
// the derivative is taken as 0, a subgradient, where the formula would give 0 / 0.
template <class T> __device__ __forceinline__ Dual<T> rsqrt_exact(Dual<T> a)
{
    const T s = rsqrt_exact(a.v);
    return dual(s, s > (T)0 ? a.d / ((T)2 * s) : (T)0);
}
template <class T> __device__ __forceinline__ T value_of(T a) { return a; }
template <class T> __device__ __forceinline__ T value_of(Dual<T> a) { return a.v; }
// A constant in the arithmetic of S (a float literal converts exactly; pi rounds to the type)
template <class S> struct Constant;
template <> struct Constant<float> { static __device__ __forceinline__ float of(double c) { return (float)c; } };
template <> struct Constant<double> { static __device__ __forceinline__ double of(double c) { return c; } };
template <> struct Constant<Dual<double>> { static __device__ __forceinline__ Dual<double> of(double c) { return dual(c); } };
template <class S> __device__ __forceinline__ S constant(double c) { return Constant<S>::of(c); }

// ACosApproximate (Math/Trigonometry.h:48-65)
template <class S> __device__ __forceinline__ S acos_approx_t(S x)
{
    const S one = constant<S>(1.0f);
    const S ax = value_of(x) < 0.0f ? constant<S>(0.0f) - x : x;
    const S abs_x = value_of(ax) < 1.0f ? ax : one;
    const S half_pi = constant<S>(3.14159265358979323846) / constant<S>(2.0f);
    const S v = rsqrt_exact(one - abs_x) * (half_pi - constant<S>(0.175394f) * abs_x);
    return value_of(x) < 0.0f ? constant<S>(3.14159265358979323846) - v : v;
}

// ApplyDihedralBendConstraints body (SoftBodyMotionProperties.cpp:367-456) on positions x[4] of
// scalar type S (float, or duals for the adjoint), writing the projected positions to y[4];
// returns false where the constraint does nothing (y = x then).
template <class S, class P>
__device__ __forceinline__ bool bend_generic(const R3<S> *x, const float *w, float compliance, float initial_angle, P inv_dt_sq,
                                             R3<S> *y)
{
    for (int k = 0; k < 4; ++k)
        y[k] = x[k];
    const R3<S> e = x[1] - x[0];
    const S e_len = rsqrt_exact(rdot(e, e));
    if (value_of(e_len) < 1.0e-6f)
        return false;
    const R3<S> x1x2 = x[2] - x[1], x1x3 = x[3] - x[1];
    R3<S> n1 = rcross(x[2] - x[0], x1x2);
    R3<S> n2 = rcross(x1x3, x[3] - x[0]);
    const S n1sq = rdot(n1, n1), n2sq = rdot(n2, n2);
    const S n1n2 = n1sq * n2sq;
    if (value_of(n1n2) < 1.0e-24f)
        return false;
    const S sign = constant<S>(value_of(rdot(rcross(n2, n1), e)) < 0.0f ? -1.0f : 1.0f);
    const S d = rdot(n1, n2) / rsqrt_exact(n1n2);
    S c = sign * acos_approx_t(d) - constant<S>(initial_angle);
    const S pi = constant<S>(3.14159265358979323846);
    if (value_of(c) > value_of(pi))
        c = c - constant<S>(2.0f) * pi;
    else if (value_of(c) < -value_of(pi))
        c = c + constant<S>(2.0f) * pi;
    n1 = n1 / n1sq;
    n2 = n2 / n2sq;
    const R3<S> d0c = (rdot(x1x2, e) * n1 + rdot(x1x3, e) * n2) / e_len;
    const R3<S> d2c = e_len * n1;
    const R3<S> d3c = e_len * n2;
    const R3<S> d1c = ((r3(constant<S>(0.0f), constant<S>(0.0f), constant<S>(0.0f)) - d0c) - d2c) - d3c;
    const S w0 = constant<S>(w[0]), w1 = constant<S>(w[1]), w2 = constant<S>(w[2]), w3 = constant<S>(w[3]);
    const S denom = (((w0 * rdot(d0c, d0c) + w1 * rdot(d1c, d1c)) + w2 * rdot(d2c, d2c)) + w3 * rdot(d3c, d3c))
                    + constant<S>(compliance) * constant<S>(inv_dt_sq);
    if (value_of(denom) < 1.0e-12f)
        return false;
    const S ml = c / denom;
    y[0] = x[0] - (ml * w0) * d0c;
    y[1] = x[1] - (ml * w1) * d1c;
    y[2] = x[2] - (ml * w2) * d2c;
    y[3] = x[3] - (ml * w3) * d3c;
    return true;
}

template <class T>
__device__ __forceinline__ void bend_project(R4<T> *X, uint32_t i0, uint32_t i1, uint32_t i2, uint32_t i3, float compliance,
                                             float initial_angle, T inv_dt_sq)
{
    const uint32_t idx[4] = { i0, i1, i2, i3 };
    R3<T> x[4], y[4];
    float w[4];
    for (int k = 0; k < 4; ++k)
    {
        x[k] = xyz3(X[idx[k]]);
        w[k] = (float)X[idx[k]].w;
    }
    if (bend_generic<T, T>(x, w, compliance, initial_angle, inv_dt_sq, y))
        for (int k = 0; k < 4; ++k)
            X[idx[k]] = r4(y[k], X[idx[k]].w);
}

// Adjoint of bend_project from its pre-projection positions q[4]: xbar = J^T ybar, the
// Jacobian J of the projection taken one input component at a time on duals.  The Jacobian is
// evaluated in double at the float positions: near a flat hinge the derivative of
// ACosApproximate, 1 / (2 sqrt(1 - |x|)), magnifies the rounding of 1 - |x| and float duals lose
// a few digits there.
__device__ __forceinline__ void bend_adjoint(float4 *A, uint32_t i0, uint32_t i1, uint32_t i2, uint32_t i3, const float4 *q,
                                             float compliance, float initial_angle, float inv_dt_sq)
{
    const uint32_t idx[4] = { i0, i1, i2, i3 };
    float w[4];
    R3<float> ybar[4];
    for (int k = 0; k < 4; ++k)
    {
        w[k] = q[k].w;
        const float4 a = A[idx[k]];
        ybar[k] = r3(a.x, a.y, a.z);
    }
    double xbar[12];
    for (int col = 0; col < 12; ++col)
    {
        R3<Dual<double>> x[4], y[4];
        for (int k = 0; k < 4; ++k)
            x[k] = r3(dual((double)q[k].x, col == 3 * k ? 1.0 : 0.0), dual((double)q[k].y, col == 3 * k + 1 ? 1.0 : 0.0),
                      dual((double)q[k].z, col == 3 * k + 2 ? 1.0 : 0.0));
        bend_generic<Dual<double>, double>(x, w, compliance, initial_angle, (double)inv_dt_sq, y);
        double acc = 0.0;
        for (int k = 0; k < 4; ++k)
            acc += (y[k].x.d * (double)ybar[k].x + y[k].y.d * (double)ybar[k].y) + y[k].z.d * (double)ybar[k].z;
        xbar[col] = acc;
    }
    for (int k = 0; k < 4; ++k)
    {
        const float4 a = A[idx[k]];
        A[idx[k]] = make_float4((float)xbar[3 * k], (float)xbar[3 * k + 1], (float)xbar[3 * k + 2], a.w);
    }
}

// Record type: 1 for a dihedral bend, whose second-lane half (bit 15 of the second word) the
// rollout skips (it runs each bend on one lane)
__device__ __forceinline__ bool is_bend_record(const uint4 &q) { return ((q.x >> 14) & 3u) == 1u; }
__device__ __forceinline__ bool is_second_half(const uint4 &q) { return (q.y & 0x8000u) != 0u; }

// ApplyVolumeConstraints body (SoftBodyMotionProperties.cpp:467-506)
template <class T>
__device__ __forceinline__ void volume_project(R4<T> *X, uint32_t i1, uint32_t i2, uint32_t i3, uint32_t i4, T six_rest_volume,
                                               T compliance, T inv_dt_sq)
{
    const R4<T> p1 = X[i1], p2 = X[i2], p3 = X[i3], p4 = X[i4];
    const R3<T> x1 = xyz3(p1), x2 = xyz3(p2), x3 = xyz3(p3), x4 = xyz3(p4);
    const R3<T> x1x2 = x2 - x1, x1x3 = x3 - x1, x1x4 = x4 - x1;
    const T six = rdot(rcross(x1x2, x1x3), x1x4);
    const T c = (six < (T)0 ? -six : six) - six_rest_volume;
    const R3<T> d1c = rcross(x4 - x2, x3 - x2), d2c = rcross(x1x3, x1x4), d3c = rcross(x1x4, x1x2), d4c = rcross(x1x2, x1x3);
    const T w1 = p1.w, w2 = p2.w, w3 = p3.w, w4 = p4.w;
    const T denom = (((w1 * rdot(d1c, d1c) + w2 * rdot(d2c, d2c)) + w3 * rdot(d3c, d3c)) + w4 * rdot(d4c, d4c)) + compliance * inv_dt_sq;
    if (denom < (T)1.0e-12f)
        return;
    const T minus_lambda = c / denom;
    X[i1] = r4(x1 - (minus_lambda * w1) * d1c, w1);
    X[i2] = r4(x2 - (minus_lambda * w2) * d2c, w2);
    X[i3] = r4(x3 - (minus_lambda * w3) * d3c, w3);
    X[i4] = r4(x4 - (minus_lambda * w4) * d4c, w4);
}

// Adjoint of volume_project from its pre-projection positions q1..q4 (outputs y_i = x_i - mu w_i d_i,
// mu = c / D, c = |6V| - 6V0, D = sum w_i |d_i|^2 + alpha / h^2, d_i = grad_i 6V):
//   s = sum w_i d_i . ybar_i,  mubar = -s,  cbar = mubar / D,  Dbar = -mubar mu / D,
//   dbar_i = -mu w_i ybar_i + 2 w_i Dbar d_i,  xbar_j = ybar_j + cbar sign(6V) d_j, and dbar_i back
//   through the cross products that form d_i (for a x b: abar = b x dbar, bbar = dbar x a).
// The volume's compliance is not a parameter of the rollout.
__device__ __forceinline__ void volume_adjoint(float4 *A, uint32_t i1, uint32_t i2, uint32_t i3, uint32_t i4, const float4 *q,
                                               float six_rest_volume, float compliance, float inv_dt_sq)
{
    const R3<float> x1 = xyz3(to_r4<float>(q[0])), x2 = xyz3(to_r4<float>(q[1])), x3 = xyz3(to_r4<float>(q[2])),
                    x4 = xyz3(to_r4<float>(q[3]));
    const float w1 = q[0].w, w2 = q[1].w, w3 = q[2].w, w4 = q[3].w;
    const R3<float> x1x2 = x2 - x1, x1x3 = x3 - x1, x1x4 = x4 - x1, x2x4 = x4 - x2, x2x3 = x3 - x2;
    const float six = rdot(rcross(x1x2, x1x3), x1x4);
    const float c = fabsf(six) - six_rest_volume;
    const R3<float> d1 = rcross(x2x4, x2x3), d2 = rcross(x1x3, x1x4), d3 = rcross(x1x4, x1x2), d4 = rcross(x1x2, x1x3);
    const float denom = (((w1 * rdot(d1, d1) + w2 * rdot(d2, d2)) + w3 * rdot(d3, d3)) + w4 * rdot(d4, d4)) + compliance * inv_dt_sq;
    if (denom < 1.0e-12f)
        return;
    const float mu = c / denom;
    const float4 a1 = A[i1], a2 = A[i2], a3 = A[i3], a4 = A[i4];
    const R3<float> y1 = xyz3(to_r4<float>(a1)), y2 = xyz3(to_r4<float>(a2)), y3 = xyz3(to_r4<float>(a3)), y4 = xyz3(to_r4<float>(a4));
    const float s = ((w1 * rdot(d1, y1) + w2 * rdot(d2, y2)) + w3 * rdot(d3, y3)) + w4 * rdot(d4, y4);
    const float mubar = -s;
    const float cbar = (mubar / denom) * (six < 0.0f ? -1.0f : 1.0f);
    const float dbar_den = 2.0f * (-mubar * mu / denom);
    const R3<float> e1 = (-mu * w1) * y1 + (dbar_den * w1) * d1, e2 = (-mu * w2) * y2 + (dbar_den * w2) * d2;
    const R3<float> e3 = (-mu * w3) * y3 + (dbar_den * w3) * d3, e4 = (-mu * w4) * y4 + (dbar_den * w4) * d4;
    R3<float> b1 = y1 + cbar * d1, b2 = y2 + cbar * d2, b3 = y3 + cbar * d3, b4 = y4 + cbar * d4;
    // d1 = (x4 - x2) x (x3 - x2)
    R3<float> g = rcross(x2x3, e1), h = rcross(e1, x2x4);
    b4 = b4 + g; b3 = b3 + h; b2 = b2 - g - h;
    // d2 = (x3 - x1) x (x4 - x1)
    g = rcross(x1x4, e2); h = rcross(e2, x1x3);
    b3 = b3 + g; b4 = b4 + h; b1 = b1 - g - h;
    // d3 = (x4 - x1) x (x2 - x1)
    g = rcross(x1x2, e3); h = rcross(e3, x1x4);
    b4 = b4 + g; b2 = b2 + h; b1 = b1 - g - h;
    // d4 = (x2 - x1) x (x3 - x1)
    g = rcross(x1x3, e4); h = rcross(e4, x1x2);
    b2 = b2 + g; b3 = b3 + h; b1 = b1 - g - h;
    A[i1] = make_float4(b1.x, b1.y, b1.z, a1.w);
    A[i2] = make_float4(b2.x, b2.y, b2.z, a2.w);
    A[i3] = make_float4(b3.x, b3.y, b3.z, a3.w);
    A[i4] = make_float4(b4.x, b4.y, b4.z, a4.w);
}

// Adjoint of edge_project from its pre-projection positions q0, q1.
__device__ __forceinline__ void edge_adjoint(float4 *A, uint32_t i0, uint32_t i1, float4 q0, float4 q1, float rest,
                                             float compliance, float inv_dt_sq, float &alpha_bar)
{
    const R3<float> x0 = r3(q0.x, q0.y, q0.z), x1 = r3(q1.x, q1.y, q1.z);
    const float w0 = q0.w, w1 = q1.w;
    const R3<float> delta = x1 - x0;
    const float length = sqrtf(rdot(delta, delta));
    const float s = (w0 + w1) + compliance * inv_dt_sq;
    const float denom = length * s;
    if (denom < 1.0e-12f)
        return;
    const float c = length - rest;
    const R3<float> correction = (delta * c) / denom;
    const float4 a0 = A[i0], a1 = A[i1];
    const R3<float> ab0 = r3(a0.x, a0.y, a0.z), ab1 = r3(a1.x, a1.y, a1.z);
    const R3<float> cbar = w0 * ab0 - w1 * ab1;
    const R3<float> n = delta / length;
    const R3<float> dbar = ((c / length) * cbar + ((rest / length) * rdot(n, cbar)) * n) / s;
    A[i0] = make_float4(ab0.x - dbar.x, ab0.y - dbar.y, ab0.z - dbar.z, a0.w);
    A[i1] = make_float4(ab1.x + dbar.x, ab1.y + dbar.y, ab1.z + dbar.z, a1.w);
    alpha_bar += rdot(cbar, correction) * (-inv_dt_sq / s);
}

template <class T> __device__ __forceinline__ R4<T> *state_vel(const RolloutDev &R);
template <> __device__ __forceinline__ R4<float> *state_vel<float>(const RolloutDev &R) { return reinterpret_cast<R4<float> *>(R.vel); }
template <> __device__ __forceinline__ R4<double> *state_vel<double>(const RolloutDev &R) { return reinterpret_cast<R4<double> *>(R.vel64); }
template <class T> __device__ __forceinline__ R4<T> *state_prev(const RolloutDev &R);
template <> __device__ __forceinline__ R4<float> *state_prev<float>(const RolloutDev &R) { return reinterpret_cast<R4<float> *>(R.prev); }
template <> __device__ __forceinline__ R4<double> *state_prev<double>(const RolloutDev &R) { return reinterpret_cast<R4<double> *>(R.prev64); }

// ApplyPressure (SoftBodyMotionProperties.cpp:291-323) on the sub-step start positions X (LDS)
// and the velocities V, with the operation order of the update kernel: GetVolumeTimesSix
// (:107-118) summed in face order by one lane from per-face terms (terms: num_faces + 2 slots),
// then each vertex adds its faces' impulses w * c * (x2 - x1) x (x3 - x1) in face order.
template <class T>
__device__ __forceinline__ void pressure_apply(const RolloutBodyDev &B, const StepConstants<T> &K, const R4<T> *X, R4<T> *V,
                                               T *terms, uint32_t tid)
{
    const uint32_t nf = B.num_faces, nv = B.nv;
    for (uint32_t f = tid; f < nf; f += kRolloutBlock)
    {
        const uint32_t *t = B.faces + 3u * f;
        terms[f] = rdot(rcross(xyz3(X[t[0]]), xyz3(X[t[1]])), xyz3(X[t[2]]));
    }
    __syncthreads();
    if (tid == 0)
    {
        T six_volume = (T)0;
        for (uint32_t f = 0; f < nf; ++f)
            six_volume += terms[f];
        terms[nf] = six_volume;
    }
    __syncthreads();
    const T six_volume = terms[nf];
    if (six_volume > (T)0)
    {
        const T coefficient = (T)B.pressure * K.dt_sub / six_volume;
        for (uint32_t i = tid; i < nv; i += kRolloutBlock)
        {
            const uint32_t fb = B.vf_begin[i], fe = B.vf_begin[i + 1];
            if (fb == fe)
                continue;
            const T w = X[i].w;
            R3<T> v = xyz3(V[i]);
            for (uint32_t k = fb; k < fe; ++k)
            {
                const uint32_t *t = B.faces + 3u * B.vf_list[k];
                const R3<T> x1 = xyz3(X[t[0]]);
                v = v + w * (coefficient * rcross(xyz3(X[t[1]]) - x1, xyz3(X[t[2]]) - x1));
            }
            V[i] = r4(v, (T)0);
        }
    }
    __syncthreads();
}

__device__ __forceinline__ void atomic_add3(float4 *A, uint32_t i, const R3<float> &d)
{
    atomicAdd(&A[i].x, d.x);
    atomicAdd(&A[i].y, d.y);
    atomicAdd(&A[i].z, d.z);
}

// Adjoint of pressure_apply at the sub-step start positions X0 (+ inverse mass).  The
// velocities pass through unchanged (vbar: their adjoints after the pressure step); with
// c = p h / S, S = sum_f x1 . (x2 x x3), n_f = (x2 - x1) x (x3 - x1) and
// ibar_f = sum_{i in f} w_i vbar_i: cbar = sum_f ibar_f . n_f, nbar_f = c ibar_f (into
// x1, x2, x3 through n_f = a x b: abar = b x nbar, bbar = nbar x a), Sbar = -cbar c / S (into
// x1, x2, x3 through S: x2 x x3, x3 x x1, x1 x x2).  A (LDS) holds the position adjoints; faces
// sharing a vertex add to it atomically.
__device__ __forceinline__ void pressure_adjoint(const RolloutBodyDev &B, float h, float4 *A, const float4 *vbar, const float4 *X0,
                                                 float *terms, uint32_t tid)
{
    const uint32_t nf = B.num_faces;
    auto x_of = [&](uint32_t i) { const float4 q = X0[i]; return r3(q.x, q.y, q.z); };
    for (uint32_t f = tid; f < nf; f += kRolloutBlock)
    {
        const uint32_t *t = B.faces + 3u * f;
        terms[f] = rdot(rcross(x_of(t[0]), x_of(t[1])), x_of(t[2]));
    }
    if (tid == 0)
        terms[nf + 1] = 0.0f;
    __syncthreads();
    if (tid == 0)
    {
        float six_volume = 0.0f;
        for (uint32_t f = 0; f < nf; ++f)
            six_volume += terms[f];
        terms[nf] = six_volume;
    }
    __syncthreads();
    const float S = terms[nf];
    if (S > 0.0f)
    {
        const float c = B.pressure * h / S;
        float cb = 0.0f;
        for (uint32_t f = tid; f < nf; f += kRolloutBlock)
        {
            const uint32_t *t = B.faces + 3u * f;
            const R3<float> x1 = x_of(t[0]), a = x_of(t[1]) - x1, b = x_of(t[2]) - x1;
            R3<float> ib = r3(0.0f, 0.0f, 0.0f);
            for (int j = 0; j < 3; ++j)
            {
                const float4 vb = vbar[t[j]];
                ib = ib + X0[t[j]].w * r3(vb.x, vb.y, vb.z);
            }
            cb += rdot(ib, rcross(a, b));
            const R3<float> nb = c * ib, ab = rcross(b, nb), bb = rcross(nb, a);
            atomic_add3(A, t[0], r3(0.0f, 0.0f, 0.0f) - (ab + bb));
            atomic_add3(A, t[1], ab);
            atomic_add3(A, t[2], bb);
        }
        atomicAdd(&terms[nf + 1], cb);
        __syncthreads();
        const float Sb = -terms[nf + 1] * c / S;
        for (uint32_t f = tid; f < nf; f += kRolloutBlock)
        {
            const uint32_t *t = B.faces + 3u * f;
            const R3<float> x1 = x_of(t[0]), x2 = x_of(t[1]), x3 = x_of(t[2]);
            atomic_add3(A, t[0], Sb * rcross(x2, x3));
            atomic_add3(A, t[1], Sb * rcross(x3, x1));
            atomic_add3(A, t[2], Sb * rcross(x1, x2));
        }
    }
    __syncthreads();
}

// Pressure scratch after the LDS positions (and edge adjoints): num_faces + 2 slots of T
template <class T> __device__ __forceinline__ T *pressure_terms(unsigned char *lds, uint32_t offset) { return reinterpret_cast<T *>(lds + offset); }

// One sub-step of the forward pass on the LDS positions X; V / P are the
// body's velocity / previous-position arrays.  The compliance of record r is
// comp(r).
template <class T, class Comp>
__device__ __forceinline__ void forward_substep(const RolloutBodyDev &B, const StepConstants<T> &K, R4<T> *X, R4<T> *V, R4<T> *P,
                                                bool last, uint32_t tid, Comp comp)
{
    const uint32_t nv = B.nv;
    for (uint32_t i = tid; i < nv; i += kRolloutBlock)
    {
        const R4<T> p = X[i];
        const R3<T> v = K.integrate_velocity(xyz3(V[i]), p.w);
        const R3<T> position = xyz3(p);
        P[i] = r4(position, (T)0);
        X[i] = r4(position + v * K.dt_sub, p.w);
        V[i] = r4(v, (T)0);
    }
    __syncthreads();
    if (B.edges_only)
        run_program_global<kRolloutBlock, false>(B.recs, B.lvl, B.nlev, tid, [&](const uint4 &q, uint32_t) {
            if (q.x != kNopRecord)
                edge_project<T>(X, q.x & 0x3fffu, q.x >> 16, (T)__uint_as_float(q.z), comp(q), K.inv_dt_sq);
        });
    else
        run_program_global<kRolloutBlock, false>(B.recs, B.lvl, B.nlev, tid, [&](const uint4 &q, uint32_t) {
            if (q.x == kNopRecord)
                return;
            if (is_lra_record(q))
                lra_project<T>(X, q.x & 0x3fffu, q.x >> 16, (T)__uint_as_float(q.z));
            else if (is_bend_record(q))
            {
                if (!is_second_half(q))
                    bend_project<T>(X, q.x & 0x3fffu, q.x >> 16, q.y & 0x7fffu, q.y >> 16, __uint_as_float(q.z),
                                    __uint_as_float(q.w), K.inv_dt_sq);
            }
            else if (is_volume_record(q))
                volume_project<T>(X, q.x & 0x3fffu, q.x >> 16, q.y & 0x7fffu, q.y >> 16, (T)__uint_as_float(q.z),
                                  (T)__uint_as_float(q.w), K.inv_dt_sq);
            else
                edge_project<T>(X, q.x & 0x3fffu, q.x >> 16, (T)__uint_as_float(q.z), comp(q), K.inv_dt_sq);
        });
    for (uint32_t i = tid; i < nv; i += kRolloutBlock)
    {
        const R4<T> p = X[i];
        R3<T> v = xyz3(V[i]);
        if (p.w > (T)0)
            v = (xyz3(p) - xyz3(P[i])) / K.dt_sub;
        if (last)
        {
            const T v_sq = rdot(v, v);
            if (v_sq > K.max_lin_sq)
                v = v * rsqrt_exact(K.max_lin_sq / v_sq);
        }
        V[i] = r4(v, (T)0);
    }
    __syncthreads();
}

// Forward pass.  The float pass first writes each record's compliance into
// the record (.w), where the backward pass reads it, and stores the (x, v)
// checkpoints; the double pass reads the compliance array directly.
template <class T>
__global__ __launch_bounds__(kRolloutBlock) void rollout_forward_kernel(RolloutDev R, const T *__restrict__ compliance,
                                                                      const T *__restrict__ velocity, T *__restrict__ positions)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    R4<T> *X = reinterpret_cast<R4<T> *>(lds_raw);
    const uint32_t tid = threadIdx.x;
    const RolloutBodyDev B = R.bodies[blockIdx.x];
    const uint32_t nv = B.nv;
    T *terms = pressure_terms<T>(lds_raw, nv * (uint32_t)sizeof(R4<T>));
    R4<T> *V = state_vel<T>(R) + B.voff;
    R4<T> *P = state_prev<T>(R) + B.voff;
    constexpr bool kFloat = sizeof(T) == 4;
    if (kFloat)
    {
        uint4 *recs = const_cast<uint4 *>(B.recs);
        for (uint32_t r = tid; r < B.num_records; r += kRolloutBlock)
        {
            uint4 q = recs[r];
            if (q.x != kNopRecord && !is_lra_record(q) && !is_volume_record(q) && !is_bend_record(q))
            {
                q.w = __float_as_uint((float)compliance[B.eoff + q.y]);
                recs[r] = q;
            }
        }
    }
    for (uint32_t i = tid; i < nv; i += kRolloutBlock)
    {
        X[i] = to_r4<T>(B.x0[i]);
        const uint32_t g = 3u * (B.voff + i);
        R4<T> v; v.x = velocity[g]; v.y = velocity[g + 1]; v.z = velocity[g + 2]; v.w = (T)0;
        V[i] = v;
    }
    __syncthreads();
    const StepConstants<T> K(R, nv);
    auto comp = [&](const uint4 &q) -> T {
        if (kFloat)
            return (T)__uint_as_float(q.w);
        return compliance[B.eoff + q.y];
    };
    for (uint32_t frame = 0; frame < R.frames; ++frame)
    {
        if (kFloat && frame % R.interval == 0)
        {
            const size_t base = (size_t)(frame / R.interval) * R.total_nv + B.voff;
            for (uint32_t i = tid; i < nv; i += kRolloutBlock)
            {
                const R4<T> p = X[i], v = V[i];
                R.ckpt_x[base + i] = make_float4((float)p.x, (float)p.y, (float)p.z, (float)p.w);
                R.ckpt_v[base + i] = make_float4((float)v.x, (float)v.y, (float)v.z, 0.0f);
            }
        }
        for (uint32_t it = 0; it < R.iterations; ++it)
        {
            if (B.num_faces > 0)
                pressure_apply<T>(B, K, X, V, terms, tid);
            forward_substep<T>(B, K, X, V, P, it + 1 == R.iterations, tid, comp);
        }
    }
    for (uint32_t i = tid; i < nv; i += kRolloutBlock)
    {
        const uint32_t g = 3u * (B.voff + i);
        const R4<T> p = X[i];
        positions[g] = p.x; positions[g + 1] = p.y; positions[g + 2] = p.z;
    }
}

// Adjoint of the last float forward pass.
__global__ __launch_bounds__(kRolloutBlock) void rollout_backward_kernel(RolloutDev R, const float *__restrict__ grad_positions,
                                                                       float *__restrict__ grad_compliance,
                                                                       float *__restrict__ grad_velocity)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const uint32_t tid = threadIdx.x;
    const RolloutBodyDev B = R.bodies[blockIdx.x];
    const uint32_t nv = B.nv, ne = B.num_edges;
    float4 *A = reinterpret_cast<float4 *>(lds_raw);                 // positions, then their adjoints
    R4<float> *X = reinterpret_cast<R4<float> *>(lds_raw);
    float *alpha_bar = reinterpret_cast<float *>(lds_raw + 16u * nv);
    float *terms = pressure_terms<float>(lds_raw, 16u * nv + ((4u * ne + 15u) & ~15u));
    const bool has_pressure = B.num_faces > 0;
    R4<float> *V = state_vel<float>(R) + B.voff;
    R4<float> *P = state_prev<float>(R) + B.voff;
    float4 *adj_x = R.adj_x + B.voff, *adj_v = R.adj_v + B.voff;
    float4 *tape = R.rec_tape + 4u * (size_t)B.rbase;

    for (uint32_t e = tid; e < ne; e += kRolloutBlock)
        alpha_bar[e] = 0.0f;
    for (uint32_t i = tid; i < nv; i += kRolloutBlock)
    {
        const uint32_t g = 3u * (B.voff + i);
        adj_x[i] = make_float4(grad_positions[g], grad_positions[g + 1], grad_positions[g + 2], 0.0f);
        adj_v[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    const StepConstants<float> K(R, nv);
    const float h = K.dt_sub;
    auto comp = [](const uint4 &q) -> float { return __uint_as_float(q.w); };
    const uint32_t I = R.iterations;
    const uint32_t num_ckpt = (R.frames + R.interval - 1) / R.interval;
    __syncthreads();

    for (uint32_t seg = num_ckpt; seg-- > 0;)
    {
        const uint32_t f0 = seg * R.interval;
        const uint32_t f1 = f0 + R.interval < R.frames ? f0 + R.interval : R.frames;
        const uint32_t nsub = (f1 - f0) * I;
        // recompute the segment from its checkpoint, keeping every sub-step's start state
        const size_t cbase = (size_t)seg * R.total_nv + B.voff;
        for (uint32_t i = tid; i < nv; i += kRolloutBlock)
        {
            const float4 p = R.ckpt_x[cbase + i], v = R.ckpt_v[cbase + i];
            X[i] = to_r4<float>(p);
            V[i] = to_r4<float>(v);
        }
        __syncthreads();
        for (uint32_t s = 0; s < nsub; ++s)
        {
            const size_t sbase = (size_t)s * R.total_nv + B.voff;
            // the tape keeps the velocity after the pressure step
            if (has_pressure)
                pressure_apply<float>(B, K, X, V, terms, tid);
            for (uint32_t i = tid; i < nv; i += kRolloutBlock)
            {
                R.seg_x[sbase + i] = A[i];
                R.seg_v[sbase + i] = to_f4(V[i]);
            }
            forward_substep<float>(B, K, X, V, P, s % I == I - 1, tid, comp);
        }
        // reverse the segment sub-step by sub-step
        for (uint32_t s = nsub; s-- > 0;)
        {
            const size_t sbase = (size_t)s * R.total_nv + B.voff;
            const bool last = s % I == I - 1;
            // re-run the sub-step, recording every edge's pre-projection positions
            for (uint32_t i = tid; i < nv; i += kRolloutBlock)
            {
                const float4 p = R.seg_x[sbase + i], v0 = R.seg_v[sbase + i];
                const R3<float> v1 = K.integrate_velocity(r3(v0.x, v0.y, v0.z), p.w);
                X[i] = r4(r3(p.x, p.y, p.z) + v1 * h, p.w);
            }
            __syncthreads();
            if (B.edges_only)
                run_program_global<kRolloutBlock, false>(B.recs, B.lvl, B.nlev, tid, [&](const uint4 &q, uint32_t r) {
                    if (q.x == kNopRecord)
                        return;
                    const uint32_t i0 = q.x & 0x3fffu, i1 = q.x >> 16;
                    tape[4u * r] = A[i0];
                    tape[4u * r + 1u] = A[i1];
                    edge_project<float>(X, i0, i1, __uint_as_float(q.z), __uint_as_float(q.w), K.inv_dt_sq);
                });
            else
                run_program_global<kRolloutBlock, false>(B.recs, B.lvl, B.nlev, tid, [&](const uint4 &q, uint32_t r) {
                    if (q.x == kNopRecord)
                        return;
                    const uint32_t i0 = q.x & 0x3fffu, i1 = q.x >> 16;
                    tape[4u * r] = A[i0];
                    tape[4u * r + 1u] = A[i1];
                    if (is_volume_record(q))
                    {
                        const uint32_t i2 = q.y & 0x7fffu, i3 = q.y >> 16;
                        tape[4u * r + 2u] = A[i2];
                        tape[4u * r + 3u] = A[i3];
                        volume_project<float>(X, i0, i1, i2, i3, __uint_as_float(q.z), __uint_as_float(q.w), K.inv_dt_sq);
                    }
                    else if (is_bend_record(q))
                    {
                        if (!is_second_half(q))
                        {
                            const uint32_t i2 = q.y & 0x7fffu, i3 = q.y >> 16;
                            tape[4u * r + 2u] = A[i2];
                            tape[4u * r + 3u] = A[i3];
                            bend_project<float>(X, i0, i1, i2, i3, __uint_as_float(q.z), __uint_as_float(q.w), K.inv_dt_sq);
                        }
                    }
                    else if (is_lra_record(q))
                        lra_project<float>(X, i0, i1, __uint_as_float(q.z));
                    else
                        edge_project<float>(X, i0, i1, __uint_as_float(q.z), __uint_as_float(q.w), K.inv_dt_sq);
                });
            // adjoint of the velocity update (and of the clamp after the last iteration)
            for (uint32_t i = tid; i < nv; i += kRolloutBlock)
            {
                const float4 p = R.seg_x[sbase + i], v0 = R.seg_v[sbase + i];
                const float w = p.w;
                const R3<float> x0 = r3(p.x, p.y, p.z);
                const R3<float> v1 = K.integrate_velocity(r3(v0.x, v0.y, v0.z), w);
                const R3<float> y = xyz3(X[i]);
                const R3<float> v2 = w > 0.0f ? (y - x0) / h : v1;
                const float4 vbf = adj_v[i], xbf = adj_x[i];
                R3<float> vb = r3(vbf.x, vbf.y, vbf.z);
                if (last)
                {
                    const float v_sq = rdot(v2, v2);
                    if (v_sq > K.max_lin_sq)
                    {
                        const float l = sqrtf(v_sq);
                        const R3<float> n = v2 / l;
                        vb = (sqrtf(K.max_lin_sq) / l) * (vb - rdot(n, vb) * n);
                    }
                }
                adj_v[i] = make_float4(vb.x, vb.y, vb.z, 0.0f);
                R3<float> yb = r3(xbf.x, xbf.y, xbf.z);
                if (w > 0.0f)
                    yb = yb + vb / h;
                A[i] = make_float4(yb.x, yb.y, yb.z, w);
            }
            __syncthreads();
            if (B.edges_only)
                run_program_global<kRolloutBlock, true>(B.recs, B.lvl, B.nlev, tid, [&](const uint4 &q, uint32_t r) {
                    if (q.x != kNopRecord)
                        edge_adjoint(A, q.x & 0x3fffu, q.x >> 16, tape[4u * r], tape[4u * r + 1u], __uint_as_float(q.z),
                                     __uint_as_float(q.w), K.inv_dt_sq, alpha_bar[q.y]);
                });
            else
                run_program_global<kRolloutBlock, true>(B.recs, B.lvl, B.nlev, tid, [&](const uint4 &q, uint32_t r) {
                    if (q.x == kNopRecord)
                        return;
                    if (is_lra_record(q))
                    {
                        lra_adjoint(A, q.x & 0x3fffu, q.x >> 16, tape[4u * r], tape[4u * r + 1u], __uint_as_float(q.z));
                        return;
                    }
                    if (is_bend_record(q))
                    {
                        if (!is_second_half(q))
                            bend_adjoint(A, q.x & 0x3fffu, q.x >> 16, q.y & 0x7fffu, q.y >> 16, tape + 4u * r, __uint_as_float(q.z),
                                         __uint_as_float(q.w), K.inv_dt_sq);
                        return;
                    }
                    if (is_volume_record(q))
                    {
                        volume_adjoint(A, q.x & 0x3fffu, q.x >> 16, q.y & 0x7fffu, q.y >> 16, tape + 4u * r, __uint_as_float(q.z),
                                       __uint_as_float(q.w), K.inv_dt_sq);
                        return;
                    }
                    edge_adjoint(A, q.x & 0x3fffu, q.x >> 16, tape[4u * r], tape[4u * r + 1u], __uint_as_float(q.z),
                                 __uint_as_float(q.w), K.inv_dt_sq, alpha_bar[q.y]);
                });
            // adjoint of prev = x0, y0 = x0 + v1 h and the integration of v
            for (uint32_t i = tid; i < nv; i += kRolloutBlock)
            {
                const float4 yb4 = A[i], vbf = adj_v[i];
                const float w = yb4.w;
                const R3<float> y0b = r3(yb4.x, yb4.y, yb4.z), vb = r3(vbf.x, vbf.y, vbf.z);
                R3<float> x0b = y0b, v1b = y0b * h;
                if (w > 0.0f)
                    x0b = x0b - vb / h;
                else
                    v1b = v1b + vb;
                const R3<float> v0b = w > 0.0f ? v1b * K.damping : v1b;
                if (has_pressure)
                    A[i] = make_float4(x0b.x, x0b.y, x0b.z, w);
                else
                    adj_x[i] = make_float4(x0b.x, x0b.y, x0b.z, 0.0f);
                adj_v[i] = make_float4(v0b.x, v0b.y, v0b.z, 0.0f);
            }
            __syncthreads();
            if (has_pressure)
            {
                pressure_adjoint(B, h, A, adj_v, R.seg_x + sbase, terms, tid);
                for (uint32_t i = tid; i < nv; i += kRolloutBlock)
                {
                    const float4 a = A[i];
                    adj_x[i] = make_float4(a.x, a.y, a.z, 0.0f);
                }
                __syncthreads();
            }
        }
    }
    for (uint32_t i = tid; i < nv; i += kRolloutBlock)
    {
        const uint32_t g = 3u * (B.voff + i);
        const float4 v = adj_v[i];
        grad_velocity[g] = v.x; grad_velocity[g + 1] = v.y; grad_velocity[g + 2] = v.z;
    }
    for (uint32_t e = tid; e < ne; e += kRolloutBlock)
        grad_compliance[B.eoff + e] = alpha_bar[e];
}

// nf: faces under pressure (0: none), which add num_faces + 2 scratch slots
uint32_t rollout_forward_lds_bytes(uint32_t nv, uint32_t nf, bool f64)
{
    return nv * (f64 ? 32u : 16u) + (nf ? (nf + 2u) * (f64 ? 8u : 4u) : 0u);
}
uint32_t rollout_backward_lds_bytes(uint32_t nv, uint32_t ne, uint32_t nf)
{
    return 16u * nv + ((4u * ne + 15u) & ~15u) + (nf ? 4u * (nf + 2u) : 0u);
}

template <class K>
static hipError_t allow_lds(K kernel)
{
    return hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 1024);
}

hipError_t launch_rollout_forward_f32(const RolloutDev &R, const float *compliance, const float *velocity, float *positions,
                                      uint32_t lds, hipStream_t stream)
{
    static hipError_t configured = allow_lds(rollout_forward_kernel<float>);
    if (configured != hipSuccess)
        return configured;
    hipLaunchKernelGGL(rollout_forward_kernel<float>, dim3(R.num_bodies), dim3(kRolloutBlock), lds, stream, R, compliance,
                       velocity, positions);
    return hipGetLastError();
}

// ---------------------------------------------------------------- no-allocation forward
// The float forward keeps each vertex's velocity and previous position in global memory
// while it runs (R.vel, R.prev; read back only by the same pass).  The allocation-free
// entry point takes them from this fixed array instead: velocities in the first
// kNoAllocMaxVertices entries, previous positions in the second.  Every call on a GPU
// shares that GPU's copy; the pass writes every entry it reads first, so nothing needs
// clearing between calls.
constexpr uint32_t kNoAllocMaxVertices = 262144;
static __device__ __align__(256) float4 scratch[2u * kNoAllocMaxVertices];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
float4 *g_scratch[kMaxScratchDevices] = {};
}

uint32_t rollout_noalloc_max_vertices() { return kNoAllocMaxVertices; }

// Looks up this device's copy of `scratch` (the host-side name is not a device address);
// called once per device by context creation, with the device current.
hipError_t rollout_scratch_setup(int device)
{
    if (device < 0 || device >= kMaxScratchDevices)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch[device])
        return hipSuccess;
    float4 *p = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess)
        return e;
    g_scratch[device] = p;
    return hipSuccess;
}

// The float forward with R.vel / R.prev in this device's scratch array.  The caller has
// checked R.total_nv against kNoAllocMaxVertices; allocates nothing.
hipError_t launch_rollout_forward_f32_noalloc(int device, RolloutDev R, const float *compliance, const float *velocity,
                                              float *positions, uint32_t lds, hipStream_t stream)
{
    if (device < 0 || device >= kMaxScratchDevices || R.total_nv > kNoAllocMaxVertices)
        return hipErrorInvalidValue;
    float4 *base;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mutex);
        base = g_scratch[device];
    }
    if (!base)
        return hipErrorNotInitialized;
    R.vel = base;
    R.prev = base + kNoAllocMaxVertices;
    return launch_rollout_forward_f32(R, compliance, velocity, positions, lds, stream);
}

hipError_t launch_rollout_forward_f64(const RolloutDev &R, const double *compliance, const double *velocity, double *positions,
                                      uint32_t lds, hipStream_t stream)
{
    static hipError_t configured = allow_lds(rollout_forward_kernel<double>);
    if (configured != hipSuccess)
        return configured;
    hipLaunchKernelGGL(rollout_forward_kernel<double>, dim3(R.num_bodies), dim3(kRolloutBlock), lds, stream, R, compliance,
                       velocity, positions);
    return hipGetLastError();
}

hipError_t launch_rollout_backward(const RolloutDev &R, const float *grad_positions, float *grad_compliance, float *grad_velocity,
                                   uint32_t lds, hipStream_t stream)
{
    static hipError_t configured = allow_lds(rollout_backward_kernel);
    if (configured != hipSuccess)
        return configured;
    hipLaunchKernelGGL(rollout_backward_kernel, dim3(R.num_bodies), dim3(kRolloutBlock), lds, stream, R, grad_positions,
                       grad_compliance, grad_velocity);
    return hipGetLastError();
}

} // namespace xpbd
