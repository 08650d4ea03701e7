// CPU restatement of the differentiable rollout and its adjoint.
//
// TEST INFRASTRUCTURE ONLY -- this is the oracle for the GPU rollout
// (include/jolt_xpbd.h, xpbd_rollout_*).  It is loaded by tests/ as the
// checker; the product never links, loads or calls it.
//
// Scope: one body with edge, dihedral bend, volume and LRA constraints, no colliders, no pressure, one
// collision step per frame, identity rotation.  The update is restated from
// the reference (paths relative to its Jolt/ directory):
//   InitializeUpdateContext        Physics/SoftBody/SoftBodyMotionProperties.cpp:937-960
//   IntegratePositions             :325-350
//   ProcessGroup :1030-1060 -- the update groups in order, in each its volumes
//       (ApplyVolumeConstraints :458-507), its bends (ApplyDihedralBendConstraints
//       :367-456), its edges (ApplyEdgeConstraints :573-601),
//       then its LRAs (ApplyLRAConstraints :679-698), each in array order: a
//       sequential Gauss-Seidel sweep
//   velocity update                :700-710 (no colliding shape)
//   velocity clamp                 UpdateSoftBodyState :852-857
// The forward pass is templated (float: the FP32 update, double: the FP64
// re-run).  The adjoint is FP64 and works from a full tape of every sub-step:
// its state, every edge's pre-projection positions and the projected positions.
// It is pinned by the tests against central finite differences of the FP64
// forward pass.  It can also run on the tape of the FP32 forward pass (the
// FP32 program's gradient in FP64 arithmetic), which separates the rounding of
// a float adjoint from the FP32 trajectory's own departure from the FP64 one.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

template <class F> struct V { F x, y, z; };
template <class F> inline V<F> mk(F x, F y, F z) { return V<F>{ x, y, z }; }
template <class F> inline V<F> operator+(V<F> a, V<F> b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class F> inline V<F> operator-(V<F> a, V<F> b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class F> inline V<F> operator*(V<F> a, F s) { return mk(a.x * s, a.y * s, a.z * s); }
template <class F> inline V<F> operator*(F s, V<F> a) { return mk(s * a.x, s * a.y, s * a.z); }
template <class F> inline V<F> operator/(V<F> a, F s) { return mk(a.x / s, a.y / s, a.z / s); }
template <class F> inline F dot(V<F> a, V<F> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }   // Vec3.inl:901-934
template <class F> inline V<F> cross(V<F> a, V<F> b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

// Forward-mode dual numbers (value, derivative) for the bend's Jacobian in the adjoint
struct D { double v, d; };
inline D operator+(D a, D b) { return { a.v + b.v, a.d + b.d }; }
inline D operator-(D a, D b) { return { a.v - b.v, a.d - b.d }; }
inline D operator*(D a, D b) { return { a.v * b.v, a.d * b.v + a.v * b.d }; }
inline D operator/(D a, D b) { return { a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v) }; }
inline D Sqrt(D a) { const double r = std::sqrt(a.v); return { r, r > 0.0 ? a.d / (2.0 * r) : 0.0 }; }   // 0 at the cusp
inline float Sqrt(float a) { return std::sqrt(a); }
inline double Sqrt(double a) { return std::sqrt(a); }
inline double Val(D a) { return a.v; }
inline double Val(double a) { return a; }
inline float Val(float a) { return a; }
template <class S> inline S Cst(double c);
template <> inline float Cst<float>(double c) { return (float)c; }
template <> inline double Cst<double>(double c) { return c; }
template <> inline D Cst<D>(double c) { return { c, 0.0 }; }

// ACosApproximate (Math/Trigonometry.h:48-65)
template <class S> S ACosApprox(S x)
{
    const S ax = Val(x) < 0.0 ? Cst<S>(0.0) - x : x;
    const S abs_x = Val(ax) < 1.0 ? ax : Cst<S>(1.0);
    const S val = Sqrt(Cst<S>(1.0) - abs_x) * (Cst<S>(3.14159265358979323846) / Cst<S>(2.0) - Cst<S>(0.175394f) * abs_x);
    return Val(x) < 0.0 ? Cst<S>(3.14159265358979323846) - val : val;
}

// ApplyDihedralBendConstraints body (SoftBodyMotionProperties.cpp:376-455) on x[4], into y[4]
template <class S, class P>
void Bend(const V<S> *x, const float *w, float compliance, float initial_angle, P inv_dt_sq, V<S> *y)
{
    for (int k = 0; k < 4; ++k)
        y[k] = x[k];
    const V<S> e = x[1] - x[0];
    const S e_len = Sqrt(dot(e, e));
    if (Val(e_len) < 1.0e-6f)
        return;
    const V<S> x1x2 = x[2] - x[1], x1x3 = x[3] - x[1];
    V<S> n1 = cross(x[2] - x[0], x1x2), n2 = cross(x1x3, x[3] - x[0]);
    const S n1sq = dot(n1, n1), n2sq = dot(n2, n2), n1n2 = n1sq * n2sq;
    if (Val(n1n2) < 1.0e-24f)
        return;
    const S sign = Cst<S>(Val(dot(cross(n2, n1), e)) < 0.0 ? -1.0 : 1.0);
    const S dd = dot(n1, n2) / Sqrt(n1n2);
    S c = sign * ACosApprox(dd) - Cst<S>(initial_angle);
    const S pi = Cst<S>(3.14159265358979323846);
    if (Val(c) > Val(pi))
        c = c - Cst<S>(2.0) * pi;
    else if (Val(c) < -Val(pi))
        c = c + Cst<S>(2.0) * pi;
    n1 = n1 / n1sq;
    n2 = n2 / n2sq;
    const V<S> d0c = (dot(x1x2, e) * n1 + dot(x1x3, e) * n2) / e_len, d2c = e_len * n1, d3c = e_len * n2;
    const V<S> zero = mk(Cst<S>(0.0), Cst<S>(0.0), Cst<S>(0.0));
    const V<S> d1c = ((zero - d0c) - d2c) - d3c;
    const S w0 = Cst<S>(w[0]), w1 = Cst<S>(w[1]), w2 = Cst<S>(w[2]), w3 = Cst<S>(w[3]);
    const S denom = (((w0 * dot(d0c, d0c) + w1 * dot(d1c, d1c)) + w2 * dot(d2c, d2c)) + w3 * dot(d3c, d3c))
                    + Cst<S>(compliance) * Cst<S>((double)inv_dt_sq);
    if (Val(denom) < 1.0e-12f)
        return;
    const S ml = c / denom;
    y[0] = x[0] - (ml * w0) * d0c;
    y[1] = x[1] - (ml * w1) * d1c;
    y[2] = x[2] - (ml * w2) * d2c;
    y[3] = x[3] - (ml * w3) * d3c;
}

struct Spec
{
    uint32_t nv, ne;
    const float *pos;       // 3 per vertex
    const float *inv_mass;
    const uint32_t *ev;     // 2 per edge
    const float *rest;
    uint32_t iterations, frames;
    float dt, linear_damping, max_linear_velocity, gravity[3], gravity_factor;
    const uint32_t *lv;     // 2 per LRA
    const float *max_distance;
    const uint32_t *tv;     // 4 per volume
    const float *six_rest, *vol_compliance;
    const uint32_t *bv;     // 4 per dihedral bend
    const float *bend_compliance, *initial_angle;
    uint32_t nf = 0;        // faces (3 vertices each) of a body under pressure
    const uint32_t *fv = nullptr;
    float pressure = 0.0f;
    // processing order: (0, edge index), (1, LRA index), (2, volume index) or (3, bend index)
    // per projection
    std::vector<std::pair<int, uint32_t>> seq;
};

template <class F>
struct Consts
{
    F dt_sub, inv_dt_sq, damping, max_lin_sq, impulse;
    V<F> sub_step_gravity;
    explicit Consts(const Spec &s)
    {
        const F step_dt = (F)s.dt / (F)1;                      // PhysicsSystem.cpp:215, one collision step
        dt_sub = step_dt / (F)s.iterations;                    // SoftBodyMotionProperties.cpp:953
        inv_dt_sq = (F)1 / (dt_sub * dt_sub);
        const F d = (F)1 - (F)s.linear_damping * dt_sub;       // :331, MotionProperties max(0, 1 - c dt)
        damping = (F)0 > d ? (F)0 : d;
        const F gf = (F)s.gravity_factor;
        const V<F> g = mk(gf * (F)s.gravity[0], gf * (F)s.gravity[1], gf * (F)s.gravity[2]);   // :949, identity rotation
        sub_step_gravity = g * dt_sub;                         // :334
        const F nvf = (F)s.nv;
        impulse = ((F)0 * dt_sub) / (nvf > (F)1 ? nvf : (F)1); // :335, no accumulated force
        max_lin_sq = (F)s.max_linear_velocity * (F)s.max_linear_velocity;
    }
    V<F> integrate(V<F> v, F w) const                          // :336-345
    {
        if (w > (F)0)
        {
            v = v + (sub_step_gravity + mk(impulse * w, impulse * w, impulse * w));
            v = v * damping;
        }
        return v;
    }
};

// ApplyPressure (SoftBodyMotionProperties.cpp:291-323) with GetVolumeTimesSix (:107-118): the
// volume sums face by face in face order, and each face's impulse adds to its vertices'
// velocities in face order.
template <class F>
void Pressure(const Spec &s, const Consts<F> &K, const std::vector<V<F>> &x, const std::vector<F> &w, std::vector<V<F>> &v)
{
    if (!(s.pressure > 0.0f) || s.nf == 0)
        return;
    F six_volume = (F)0;
    for (uint32_t f = 0; f < s.nf; ++f)
        six_volume += dot(cross(x[s.fv[3 * f]], x[s.fv[3 * f + 1]]), x[s.fv[3 * f + 2]]);
    if (!(six_volume > (F)0))
        return;
    const F coefficient = (F)s.pressure * K.dt_sub / six_volume;
    for (uint32_t f = 0; f < s.nf; ++f)
    {
        const uint32_t *t = s.fv + 3 * f;
        const V<F> impulse = coefficient * cross(x[t[1]] - x[t[0]], x[t[2]] - x[t[0]]);
        for (int j = 0; j < 3; ++j)
            v[t[j]] = v[t[j]] + w[t[j]] * impulse;
    }
}

// Adjoint of Pressure at positions x0: the velocities pass through unchanged; with
// c = p h / S, S = sum_f x1 . (x2 x x3), n_f = (x2 - x1) x (x3 - x1) and
// ibar_f = sum_{i in f} w_i vbar_i: cbar = sum_f ibar_f . n_f, nbar_f = c ibar_f,
// Sbar = -cbar c / S, and xbar gains nbar_f through n_f and Sbar through S.
void PressureAdjoint(const Spec &s, const Consts<double> &K, const double *x0, const std::vector<V<double>> &vb,
                     std::vector<V<double>> &xb)
{
    if (!(s.pressure > 0.0f) || s.nf == 0)
        return;
    auto X = [&](uint32_t i) { return mk(x0[3 * i], x0[3 * i + 1], x0[3 * i + 2]); };
    double S = 0.0;
    for (uint32_t f = 0; f < s.nf; ++f)
        S += dot(cross(X(s.fv[3 * f]), X(s.fv[3 * f + 1])), X(s.fv[3 * f + 2]));
    if (!(S > 0.0))
        return;
    const double c = (double)s.pressure * K.dt_sub / S;
    double cbar = 0.0;
    for (uint32_t f = 0; f < s.nf; ++f)
    {
        const uint32_t *t = s.fv + 3 * f;
        const V<double> a = X(t[1]) - X(t[0]), b = X(t[2]) - X(t[0]);
        V<double> ib = mk(0.0, 0.0, 0.0);
        for (int j = 0; j < 3; ++j)
            ib = ib + (double)s.inv_mass[t[j]] * vb[t[j]];
        cbar += dot(ib, cross(a, b));
        const V<double> nb = c * ib, ab = cross(b, nb), bb = cross(nb, a);
        xb[t[0]] = xb[t[0]] - (ab + bb);
        xb[t[1]] = xb[t[1]] + ab;
        xb[t[2]] = xb[t[2]] + bb;
    }
    const double Sb = -cbar * c / S;
    for (uint32_t f = 0; f < s.nf; ++f)
    {
        const uint32_t *t = s.fv + 3 * f;
        const V<double> x1 = X(t[0]), x2 = X(t[1]), x3 = X(t[2]);
        xb[t[0]] = xb[t[0]] + Sb * cross(x2, x3);
        xb[t[1]] = xb[t[1]] + Sb * cross(x3, x1);
        xb[t[2]] = xb[t[2]] + Sb * cross(x1, x2);
    }
}

struct Tape
{
    std::vector<double> x0, v0, pre, y;     // per sub-step: 3 nv, 3 nv (after pressure), 12 per projection, 3 nv
    std::vector<uint8_t> acted;             // per sub-step and projection: an LRA moved its vertex (decided
                                            // in the forward pass's precision)
};

// Forward pass; with a tape, records every sub-step (double only).
template <class F>
void Forward(const Spec &s, const F *compliance, const F *velocity, F *out, Tape *tape)
{
    const Consts<F> K(s);
    const uint32_t nv = s.nv, ne = s.ne;
    std::vector<V<F>> x(nv), v(nv), prev(nv);
    std::vector<F> w(nv);
    for (uint32_t i = 0; i < nv; ++i)
    {
        x[i] = mk((F)s.pos[3 * i], (F)s.pos[3 * i + 1], (F)s.pos[3 * i + 2]);
        w[i] = (F)s.inv_mass[i];
        v[i] = mk(velocity[3 * i], velocity[3 * i + 1], velocity[3 * i + 2]);
    }
    for (uint32_t frame = 0; frame < s.frames; ++frame)
        for (uint32_t it = 0; it < s.iterations; ++it)
        {
            Pressure<F>(s, K, x, w, v);
            if (tape)
                for (uint32_t i = 0; i < nv; ++i)
                {
                    const double a[6] = { (double)x[i].x, (double)x[i].y, (double)x[i].z, (double)v[i].x, (double)v[i].y, (double)v[i].z };
                    tape->x0.insert(tape->x0.end(), a, a + 3);
                    tape->v0.insert(tape->v0.end(), a + 3, a + 6);
                }
            for (uint32_t i = 0; i < nv; ++i)                  // IntegratePositions
            {
                v[i] = K.integrate(v[i], w[i]);
                prev[i] = x[i];
                x[i] = x[i] + v[i] * K.dt_sub;
            }
            for (const auto &op : s.seq)
            {
                const uint32_t e = op.second;
                const uint32_t *vs = op.first == 3 ? s.bv + 4 * e : op.first == 2 ? s.tv + 4 * e : op.first == 1 ? s.lv + 2 * e : s.ev + 2 * e;
                const uint32_t i0 = vs[0], i1 = vs[1];
                if (tape)
                    for (int j = 0; j < 4; ++j)
                    {
                        const uint32_t vj = j < (op.first >= 2 ? 4 : 2) ? vs[j] : vs[0];
                        const double a[3] = { (double)x[vj].x, (double)x[vj].y, (double)x[vj].z };
                        tape->pre.insert(tape->pre.end(), a, a + 3);
                    }
                if (op.first == 3)                              // ApplyDihedralBendConstraints :376-455
                {
                    if (tape)
                        tape->acted.push_back(0);
                    const V<F> xs[4] = { x[vs[0]], x[vs[1]], x[vs[2]], x[vs[3]] };
                    const float ws[4] = { (float)w[vs[0]], (float)w[vs[1]], (float)w[vs[2]], (float)w[vs[3]] };
                    V<F> ys[4];
                    Bend<F, F>(xs, ws, s.bend_compliance[e], s.initial_angle[e], K.inv_dt_sq, ys);
                    for (int j = 0; j < 4; ++j)
                        x[vs[j]] = ys[j];
                    continue;
                }
                if (op.first == 2)                              // ApplyVolumeConstraints :467-506
                {
                    if (tape)
                        tape->acted.push_back(0);
                    const uint32_t i2 = vs[2], i3 = vs[3];
                    const V<F> x1x2 = x[i1] - x[i0], x1x3 = x[i2] - x[i0], x1x4 = x[i3] - x[i0];
                    const F six = dot(cross(x1x2, x1x3), x1x4);
                    const F c = std::abs(six) - (F)s.six_rest[e];
                    const V<F> d1 = cross(x[i3] - x[i1], x[i2] - x[i1]), d2 = cross(x1x3, x1x4), d3 = cross(x1x4, x1x2), d4 = cross(x1x2, x1x3);
                    const F denom = (((w[i0] * dot(d1, d1) + w[i1] * dot(d2, d2)) + w[i2] * dot(d3, d3)) + w[i3] * dot(d4, d4))
                                    + (F)s.vol_compliance[e] * K.inv_dt_sq;
                    if (denom < (F)1.0e-12f)
                        continue;
                    const F ml = c / denom;
                    x[i0] = x[i0] - (ml * w[i0]) * d1;
                    x[i1] = x[i1] - (ml * w[i1]) * d2;
                    x[i2] = x[i2] - (ml * w[i2]) * d3;
                    x[i3] = x[i3] - (ml * w[i3]) * d4;
                    continue;
                }
                if (op.first)                                   // ApplyLRAConstraints :685-697
                {
                    const V<F> d = x[i1] - x[i0];
                    const F d_sq = dot(d, d), m = (F)s.max_distance[e];
                    const bool act = d_sq > m * m;
                    if (tape)
                        tape->acted.push_back(act ? 1 : 0);
                    if (act)
                        x[i1] = x[i0] + (d * m) / std::sqrt(d_sq);
                    continue;
                }
                if (tape)
                    tape->acted.push_back(0);
                // ApplyEdgeConstraints :583-600
                const V<F> delta = x[i1] - x[i0];
                const F length = std::sqrt(dot(delta, delta));
                const F denom = length * ((w[i0] + w[i1]) + compliance[e] * K.inv_dt_sq);
                if (denom < (F)1.0e-12f)
                    continue;
                const V<F> correction = (delta * (length - (F)s.rest[e])) / denom;
                x[i0] = x[i0] + w[i0] * correction;
                x[i1] = x[i1] - w[i1] * correction;
            }
            if (tape)
                for (uint32_t i = 0; i < nv; ++i)
                {
                    const double a[3] = { (double)x[i].x, (double)x[i].y, (double)x[i].z };
                    tape->y.insert(tape->y.end(), a, a + 3);
                }
            const bool last = it + 1 == s.iterations;
            for (uint32_t i = 0; i < nv; ++i)                  // velocity update :700-710, clamp :852-857
            {
                if (w[i] > (F)0)
                    v[i] = (x[i] - prev[i]) / K.dt_sub;
                if (last)
                {
                    const F v_sq = dot(v[i], v[i]);
                    if (v_sq > K.max_lin_sq)
                        v[i] = v[i] * std::sqrt(K.max_lin_sq / v_sq);
                }
            }
        }
    for (uint32_t i = 0; i < nv; ++i)
    {
        out[3 * i] = x[i].x; out[3 * i + 1] = x[i].y; out[3 * i + 2] = x[i].z;
    }
}

// Reverse pass over the tape, in FP64.  forward_precision 64: the tape of the
// FP64 forward pass (the FP64 program's gradient); 32: the tape of the FP32
// forward pass (the FP32 program's Jacobians, evaluated in FP64 arithmetic).
void Adjoint(const Spec &s, const double *compliance, const double *velocity, const double *grad_pos, double *grad_comp,
             double *grad_vel, double *out_pos, int forward_precision)
{
    Tape tape;
    if (forward_precision == 32)
    {
        std::vector<float> c(compliance, compliance + s.ne), v(velocity, velocity + 3 * (size_t)s.nv), out(3 * (size_t)s.nv);
        Forward<float>(s, c.data(), v.data(), out.data(), &tape);
        for (size_t k = 0; k < out.size(); ++k)
            out_pos[k] = out[k];
    }
    else
        Forward<double>(s, compliance, velocity, out_pos, &tape);
    const Consts<double> K(s);
    const uint32_t nv = s.nv, ne = s.ne, I = s.iterations;
    const double h = K.dt_sub;
    std::vector<V<double>> xb(nv), vb(nv), A(nv);
    for (uint32_t i = 0; i < nv; ++i)
    {
        xb[i] = mk(grad_pos[3 * i], grad_pos[3 * i + 1], grad_pos[3 * i + 2]);
        vb[i] = mk(0.0, 0.0, 0.0);
    }
    for (uint32_t e = 0; e < ne; ++e)
        grad_comp[e] = 0.0;
    const uint32_t nsub = s.frames * I;
    for (uint32_t sub = nsub; sub-- > 0;)
    {
        const bool last = sub % I == I - 1;
        const double *x0 = &tape.x0[(size_t)sub * 3 * nv], *v0 = &tape.v0[(size_t)sub * 3 * nv];
        const double *y = &tape.y[(size_t)sub * 3 * nv], *pre = &tape.pre[(size_t)sub * 12 * s.seq.size()];
        for (uint32_t i = 0; i < nv; ++i)
        {
            const double w = s.inv_mass[i];
            const V<double> vv1 = K.integrate(mk(v0[3 * i], v0[3 * i + 1], v0[3 * i + 2]), w);
            const V<double> yy = mk(y[3 * i], y[3 * i + 1], y[3 * i + 2]), xx0 = mk(x0[3 * i], x0[3 * i + 1], x0[3 * i + 2]);
            const V<double> v2 = w > 0.0 ? (yy - xx0) / h : vv1;
            if (last)
            {
                const double v_sq = dot(v2, v2);
                if (v_sq > K.max_lin_sq)
                {
                    const double l = std::sqrt(v_sq);
                    const V<double> n = v2 / l;
                    vb[i] = (std::sqrt(K.max_lin_sq) / l) * (vb[i] - dot(n, vb[i]) * n);
                }
            }
            A[i] = w > 0.0 ? xb[i] + vb[i] / h : xb[i];
        }
        for (size_t k = s.seq.size(); k-- > 0;)
        {
            const uint32_t e = s.seq[k].second;
            const double *q = pre + 12 * k;
            const V<double> p0 = mk(q[0], q[1], q[2]), p1 = mk(q[3], q[4], q[5]);
            if (s.seq[k].first == 3)
            {
                // xbar = J^T ybar, J column by column on duals from the pre-projection positions
                const uint32_t *vs = s.bv + 4 * e;
                const float ws[4] = { (float)s.inv_mass[vs[0]], (float)s.inv_mass[vs[1]], (float)s.inv_mass[vs[2]], (float)s.inv_mass[vs[3]] };
                double xbar[12];
                for (int col = 0; col < 12; ++col)
                {
                    V<D> xs[4], ys[4];
                    for (int j = 0; j < 4; ++j)
                        xs[j] = mk(D{ q[3 * j], col == 3 * j ? 1.0 : 0.0 }, D{ q[3 * j + 1], col == 3 * j + 1 ? 1.0 : 0.0 },
                                   D{ q[3 * j + 2], col == 3 * j + 2 ? 1.0 : 0.0 });
                    Bend<D, double>(xs, ws, s.bend_compliance[e], s.initial_angle[e], K.inv_dt_sq, ys);
                    double acc = 0.0;
                    for (int j = 0; j < 4; ++j)
                        acc += ys[j].x.d * A[vs[j]].x + ys[j].y.d * A[vs[j]].y + ys[j].z.d * A[vs[j]].z;
                    xbar[col] = acc;
                }
                for (int j = 0; j < 4; ++j)
                    A[vs[j]] = mk(xbar[3 * j], xbar[3 * j + 1], xbar[3 * j + 2]);
                continue;
            }
            if (s.seq[k].first == 2)
            {
                // y_i = x_i - mu w_i d_i, mu = c / D, c = |6V| - 6V0, d_i = grad_i 6V: see the device's volume_adjoint
                const uint32_t *vs = s.tv + 4 * e;
                const V<double> p2 = mk(q[6], q[7], q[8]), p3 = mk(q[9], q[10], q[11]);
                const double w0 = s.inv_mass[vs[0]], w1 = s.inv_mass[vs[1]], w2 = s.inv_mass[vs[2]], w3 = s.inv_mass[vs[3]];
                const V<double> a12 = p1 - p0, a13 = p2 - p0, a14 = p3 - p0, a24 = p3 - p1, a23 = p2 - p1;
                const double six = dot(cross(a12, a13), a14);
                const V<double> d1 = cross(a24, a23), d2 = cross(a13, a14), d3 = cross(a14, a12), d4 = cross(a12, a13);
                const double D = w0 * dot(d1, d1) + w1 * dot(d2, d2) + w2 * dot(d3, d3) + w3 * dot(d4, d4)
                                 + (double)s.vol_compliance[e] * K.inv_dt_sq;
                if (D < (double)1.0e-12f)
                    continue;
                const double mu = (std::abs(six) - (double)s.six_rest[e]) / D;
                const V<double> y0 = A[vs[0]], y1 = A[vs[1]], y2 = A[vs[2]], y3 = A[vs[3]];
                const double mubar = -(w0 * dot(d1, y0) + w1 * dot(d2, y1) + w2 * dot(d3, y2) + w3 * dot(d4, y3));
                const double cbar = mubar / D * (six < 0.0 ? -1.0 : 1.0), Dbar2 = 2.0 * (-mubar * mu / D);
                const V<double> e1 = (-mu * w0) * y0 + (Dbar2 * w0) * d1, e2 = (-mu * w1) * y1 + (Dbar2 * w1) * d2;
                const V<double> e3 = (-mu * w2) * y2 + (Dbar2 * w2) * d3, e4 = (-mu * w3) * y3 + (Dbar2 * w3) * d4;
                V<double> b0 = y0 + cbar * d1, b1 = y1 + cbar * d2, b2 = y2 + cbar * d3, b3 = y3 + cbar * d4;
                V<double> g = cross(a23, e1), hh = cross(e1, a24);          // d1 = (x4 - x2) x (x3 - x2)
                b3 = b3 + g; b2 = b2 + hh; b1 = b1 - g - hh;
                g = cross(a14, e2); hh = cross(e2, a13);                    // d2 = (x3 - x1) x (x4 - x1)
                b2 = b2 + g; b3 = b3 + hh; b0 = b0 - g - hh;
                g = cross(a12, e3); hh = cross(e3, a14);                    // d3 = (x4 - x1) x (x2 - x1)
                b3 = b3 + g; b1 = b1 + hh; b0 = b0 - g - hh;
                g = cross(a13, e4); hh = cross(e4, a12);                    // d4 = (x2 - x1) x (x3 - x1)
                b1 = b1 + g; b2 = b2 + hh; b0 = b0 - g - hh;
                A[vs[0]] = b0; A[vs[1]] = b1; A[vs[2]] = b2; A[vs[3]] = b3;
                continue;
            }
            if (s.seq[k].first)
            {
                // x1' = x0 + m u, u = d / |d|:  x1bar = (m / |d|)(a1 - u (u . a1)),  x0bar += a1 - x1bar
                const uint32_t j0 = s.lv[2 * e], j1 = s.lv[2 * e + 1];
                const V<double> d = p1 - p0;
                const double d_sq = dot(d, d), m = (double)s.max_distance[e];
                if (!tape.acted[(size_t)sub * s.seq.size() + k])
                    continue;
                const double L = std::sqrt(d_sq);
                const V<double> u = d / L;
                const V<double> g = (m / L) * (A[j1] - dot(u, A[j1]) * u);
                A[j0] = A[j0] + A[j1] - g;
                A[j1] = g;
                continue;
            }
            const uint32_t i0 = s.ev[2 * e], i1 = s.ev[2 * e + 1];
            const double w0 = s.inv_mass[i0], w1 = s.inv_mass[i1];
            const V<double> delta = p1 - p0;
            const double L = std::sqrt(dot(delta, delta));
            const double sc = (w0 + w1) + compliance[e] * K.inv_dt_sq;
            if (L * sc < (double)1.0e-12f)
                continue;
            const double c = L - (double)s.rest[e];
            const V<double> corr = (delta * c) / (L * sc);
            const V<double> cbar = w0 * A[i0] - w1 * A[i1];
            const V<double> n = delta / L;
            const V<double> dbar = ((c / L) * cbar + ((double)s.rest[e] / L * dot(n, cbar)) * n) / sc;
            A[i0] = A[i0] - dbar;
            A[i1] = A[i1] + dbar;
            grad_comp[e] += dot(cbar, corr) * (-K.inv_dt_sq / sc);
        }
        for (uint32_t i = 0; i < nv; ++i)
        {
            const double w = s.inv_mass[i];
            V<double> x0b = A[i], v1b = A[i] * h;
            if (w > 0.0)
                x0b = x0b - vb[i] / h;
            else
                v1b = v1b + vb[i];
            xb[i] = x0b;
            vb[i] = w > 0.0 ? v1b * K.damping : v1b;
        }
        PressureAdjoint(s, K, x0, vb, xb);
    }
    for (uint32_t i = 0; i < nv; ++i)
    {
        grad_vel[3 * i] = vb[i].x; grad_vel[3 * i + 1] = vb[i].y; grad_vel[3 * i + 2] = vb[i].z;
    }
}

// groups: ng update groups' end indices (edge end, LRA end per group); ng 0: the edges in
// array order, no LRA
Spec MakeSpec(uint32_t nv, const float *pos, const float *inv_mass, uint32_t ne, const uint32_t *ev, const float *rest,
              uint32_t iterations, uint32_t frames, float dt, float linear_damping, float max_linear_velocity,
              const float *gravity, float gravity_factor, uint32_t nl = 0, const uint32_t *lv = nullptr,
              const float *max_distance = nullptr, uint32_t nt = 0, const uint32_t *tv = nullptr, const float *six_rest = nullptr,
              const float *vol_compliance = nullptr, uint32_t nb = 0, const uint32_t *bv = nullptr,
              const float *bend_compliance = nullptr, const float *initial_angle = nullptr, uint32_t ng = 0,
              const uint32_t *groups = nullptr, uint32_t nf = 0, const uint32_t *fv = nullptr, float pressure = 0.0f)
{
    Spec s;
    s.nf = nf; s.fv = fv; s.pressure = pressure;
    s.nv = nv; s.ne = ne; s.pos = pos; s.inv_mass = inv_mass; s.ev = ev; s.rest = rest;
    s.lv = lv; s.max_distance = max_distance;
    s.tv = tv; s.six_rest = six_rest; s.vol_compliance = vol_compliance;
    s.bv = bv; s.bend_compliance = bend_compliance; s.initial_angle = initial_angle;
    if (ng == 0)
        for (uint32_t e = 0; e < ne; ++e)
            s.seq.push_back({ 0, e });
    uint32_t t0 = 0, b0 = 0, e0 = 0, l0 = 0;
    for (uint32_t g = 0; g < ng; ++g)
    {
        for (uint32_t t = t0; t < groups[4 * g] && t < nt; ++t)
            s.seq.push_back({ 2, t });
        for (uint32_t b = b0; b < groups[4 * g + 1] && b < nb; ++b)
            s.seq.push_back({ 3, b });
        for (uint32_t e = e0; e < groups[4 * g + 2] && e < ne; ++e)
            s.seq.push_back({ 0, e });
        for (uint32_t l = l0; l < groups[4 * g + 3] && l < nl; ++l)
            if (lv[2 * l] != lv[2 * l + 1])         // an LRA of a vertex to itself never moves it
                s.seq.push_back({ 1, l });
        t0 = groups[4 * g];
        b0 = groups[4 * g + 1];
        e0 = groups[4 * g + 2];
        l0 = groups[4 * g + 3];
    }
    s.iterations = iterations; s.frames = frames; s.dt = dt; s.linear_damping = linear_damping;
    s.max_linear_velocity = max_linear_velocity;
    s.gravity[0] = gravity[0]; s.gravity[1] = gravity[1]; s.gravity[2] = gravity[2];
    s.gravity_factor = gravity_factor;
    return s;
}

} // namespace

extern "C" {

// precision 32: FP32 forward (parameters rounded to float); 64: FP64.
int xa_rollout_forward(int precision, uint32_t nv, const float *pos, const float *inv_mass, uint32_t ne, const uint32_t *ev,
                       const float *rest, uint32_t iterations, uint32_t frames, float dt, float linear_damping,
                       float max_linear_velocity, const float *gravity, float gravity_factor, const double *compliance,
                       const double *velocity, double *positions)
{
    const Spec s = MakeSpec(nv, pos, inv_mass, ne, ev, rest, iterations, frames, dt, linear_damping, max_linear_velocity,
                            gravity, gravity_factor);
    if (precision == 64)
    {
        Forward<double>(s, compliance, velocity, positions, nullptr);
        return 0;
    }
    if (precision != 32)
        return -1;
    std::vector<float> c(compliance, compliance + ne), v(velocity, velocity + 3 * (size_t)nv), out(3 * (size_t)nv);
    Forward<float>(s, c.data(), v.data(), out.data(), nullptr);
    for (size_t k = 0; k < out.size(); ++k)
        positions[k] = out[k];
    return 0;
}

int xa_rollout_gradient(int forward_precision, uint32_t nv, const float *pos, const float *inv_mass, uint32_t ne,
                        const uint32_t *ev, const float *rest, uint32_t iterations, uint32_t frames, float dt,
                        float linear_damping, float max_linear_velocity, const float *gravity, float gravity_factor,
                        const double *compliance, const double *velocity, const double *grad_positions,
                        double *grad_compliance, double *grad_velocity, double *positions)
{
    if (forward_precision != 32 && forward_precision != 64)
        return -1;
    const Spec s = MakeSpec(nv, pos, inv_mass, ne, ev, rest, iterations, frames, dt, linear_damping, max_linear_velocity,
                            gravity, gravity_factor);
    Adjoint(s, compliance, velocity, grad_positions, grad_compliance, grad_velocity, positions, forward_precision);
    return 0;
}

// The two entry points above for a body with bend, volume and LRA constraints: nl LRAs (2
// vertices each, max distance), nt volumes (4 vertices, six rest volume, compliance each), nb
// dihedral bends (4 vertices, compliance, initial angle each), ng update groups (volume end,
// bend end, edge end, LRA end each).
int xa_rollout_forward_lra(int precision, uint32_t nv, const float *pos, const float *inv_mass, uint32_t ne, const uint32_t *ev,
                           const float *rest, uint32_t nl, const uint32_t *lv, const float *max_distance, uint32_t nt,
                           const uint32_t *tv, const float *six_rest, const float *vol_compliance, uint32_t nb,
                           const uint32_t *bv, const float *bend_compliance, const float *initial_angle, uint32_t ng,
                           const uint32_t *groups, uint32_t nf, const uint32_t *fv, float pressure, uint32_t iterations, uint32_t frames, float dt, float linear_damping,
                           float max_linear_velocity, const float *gravity, float gravity_factor, const double *compliance,
                           const double *velocity, double *positions)
{
    const Spec s = MakeSpec(nv, pos, inv_mass, ne, ev, rest, iterations, frames, dt, linear_damping, max_linear_velocity,
                            gravity, gravity_factor, nl, lv, max_distance, nt, tv, six_rest, vol_compliance, nb, bv, bend_compliance, initial_angle, ng, groups, nf, fv, pressure);
    if (precision == 64)
    {
        Forward<double>(s, compliance, velocity, positions, nullptr);
        return 0;
    }
    if (precision != 32)
        return -1;
    std::vector<float> c(compliance, compliance + ne), v(velocity, velocity + 3 * (size_t)nv), out(3 * (size_t)nv);
    Forward<float>(s, c.data(), v.data(), out.data(), nullptr);
    for (size_t k = 0; k < out.size(); ++k)
        positions[k] = out[k];
    return 0;
}

int xa_rollout_gradient_lra(int forward_precision, uint32_t nv, const float *pos, const float *inv_mass, uint32_t ne,
                            const uint32_t *ev, const float *rest, uint32_t nl, const uint32_t *lv, const float *max_distance,
                            uint32_t nt, const uint32_t *tv, const float *six_rest, const float *vol_compliance,
                            uint32_t nb, const uint32_t *bv, const float *bend_compliance, const float *initial_angle,
                            uint32_t ng, const uint32_t *groups, uint32_t nf, const uint32_t *fv, float pressure, uint32_t iterations, uint32_t frames, float dt,
                            float linear_damping, float max_linear_velocity, const float *gravity, float gravity_factor,
                            const double *compliance, const double *velocity, const double *grad_positions,
                            double *grad_compliance, double *grad_velocity, double *positions)
{
    if (forward_precision != 32 && forward_precision != 64)
        return -1;
    const Spec s = MakeSpec(nv, pos, inv_mass, ne, ev, rest, iterations, frames, dt, linear_damping, max_linear_velocity,
                            gravity, gravity_factor, nl, lv, max_distance, nt, tv, six_rest, vol_compliance, nb, bv, bend_compliance, initial_angle, ng, groups, nf, fv, pressure);
    Adjoint(s, compliance, velocity, grad_positions, grad_compliance, grad_velocity, positions, forward_precision);
    return 0;
}

} // extern "C"
