// CPU restatement of Jolt Physics' soft-body XPBD update (reference v5.5.1).
//
// TEST INFRASTRUCTURE ONLY -- this is the oracle.  It is loaded by tests/ and
// by bench.py's cpu_baseline leg as the checker; the product never links,
// loads or calls it.
//
// Templated on the arithmetic type F:
//   F = float  : follows the reference's operation order so that, built with
//                -ffp-contract=off, it is bit-exact with the reference compiled
//                the same way (pinned against oracle/_ref by the tests);
//   F = double : the same algorithm in FP64, the "FP64 CPU re-run" used to state
//                the FP32 error of both the reference and the GPU path.
//
// Not restated here (their parity is anchored on the reference library built under
// oracle/_ref and on closed forms in the tests): skinned constraints, Cosserat rods,
// per-pair contact settings of a contact listener, and sensors (skipped: they never
// change the simulation).
//
// Every function cites the reference lines it restates.  Paths are relative to
// the reference's Jolt/ directory.  The SSE semantics that fix the float
// results (Math/Vec3.inl, Math/Mat44.inl) are restated in the small vector
// helpers below:
//   Dot(a,b)   = (a.x*b.x + a.y*b.y) + a.z*b.z            Vec3.inl:901-934 (ReduceSum)
//   Cross      = componentwise mul/sub, no FMA             Vec3.inl:855-863
//   -v         = 0 - v                                     Vec3.inl:669-672
//   sMin/sMax  = (a < b ? a : b) / (a > b ? a : b)         _mm_min_ps/_mm_max_ps
//   Mat44*Vec3 = ((c0*x + c1*y) + c2*z) + c3               Mat44.inl:295-301
//   Multiply3x3= (c0*x + c1*y) + c2*z                      Mat44.inl:384-389
//   Mat44*Mat44= per column ((c0*m.x + c1*m.y) + c2*m.z) + c3*m.w   Mat44.inl:241-252

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

using u32 = uint32_t;

// ------------------------------------------------------------------ vectors
template <class F> struct V3 { F x, y, z; };

template <class F> inline V3<F> operator+(V3<F> a, V3<F> b) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
template <class F> inline V3<F> operator-(V3<F> a, V3<F> b) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
template <class F> inline V3<F> operator*(V3<F> a, F s) { return { a.x * s, a.y * s, a.z * s }; }
template <class F> inline V3<F> operator*(F s, V3<F> a) { return { s * a.x, s * a.y, s * a.z }; }
template <class F> inline V3<F> operator/(V3<F> a, F s) { return { a.x / s, a.y / s, a.z / s }; }
template <class F> inline V3<F> Neg(V3<F> a) { return { F(0) - a.x, F(0) - a.y, F(0) - a.z }; }
template <class F> inline V3<F> Mul(V3<F> a, V3<F> b) { return { a.x * b.x, a.y * b.y, a.z * b.z }; }
template <class F> inline F Dot(V3<F> a, V3<F> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class F> inline F LenSq(V3<F> a) { return Dot(a, a); }
template <class F> inline F Len(V3<F> a) { return std::sqrt(LenSq(a)); }
template <class F> inline V3<F> Cross(V3<F> a, V3<F> b)
{
	return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x };
}
template <class F> inline F SMin(F a, F b) { return a < b ? a : b; }
template <class F> inline F SMax(F a, F b) { return a > b ? a : b; }
template <class F> inline V3<F> VMin(V3<F> a, V3<F> b) { return { SMin(a.x, b.x), SMin(a.y, b.y), SMin(a.z, b.z) }; }
template <class F> inline V3<F> VMax(V3<F> a, V3<F> b) { return { SMax(a.x, b.x), SMax(a.y, b.y), SMax(a.z, b.z) }; }
template <class F> inline F Comp(V3<F> a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
// Abs: max(0 - v, v) (Vec3.inl:817-822)
template <class F> inline V3<F> VAbs(V3<F> a) { return VMax(Neg(a), a); }
// GetSign: sign bit of v or'ed onto 1.0 (Vec3.inl:1161-1166)
template <class F> inline V3<F> VSign(V3<F> a)
{
	return { std::signbit(a.x) ? F(-1) : F(1), std::signbit(a.y) ? F(-1) : F(1), std::signbit(a.z) ? F(-1) : F(1) };
}
// GetLowestComponentIndex (Vec3.inl:809-812)
template <class F> inline int LowestIndex(V3<F> a) { return a.x < a.y ? (a.z < a.x ? 2 : 0) : (a.z < a.y ? 2 : 1); }

template <class F> struct Constants;
template <> struct Constants<float> { static constexpr float kPi = 3.14159265358979323846f; };
template <> struct Constants<double> { static constexpr double kPi = 3.14159265358979323846; };

// Sign(0) = +1 (Math/Math.h:116-119)
template <class F> inline F SignOf(F v) { return v < F(0) ? F(-1) : F(1); }
// Vec3::NormalizedOr (Vec3.inl:978-1000): zero_value when |v|^2 <= FLT_MIN
template <class F> inline V3<F> NormalizedOr(V3<F> v, V3<F> zero_value)
{
	const F l2 = LenSq(v);
	return l2 <= F(FLT_MIN) ? zero_value : v / std::sqrt(l2);
}

// ClosestPoint::GetClosestPointOnLine with the line's GetBaryCentricCoordinates
// (ClosestPoint.h:15-41, 129-151): closest point to the origin on segment (a, b);
// a degenerate segment (|b - a|^2 < FLT_EPSILON^2) falls back to the nearer end.
template <class F> inline V3<F> ClosestPointOnLine(V3<F> a, V3<F> b)
{
	V3<F> ab = b - a;
	F denominator = LenSq(ab);
	F u, v;
	if (denominator < F(FLT_EPSILON) * F(FLT_EPSILON))
	{
		if (LenSq(a) < LenSq(b)) { u = 1; v = 0; }
		else { u = 0; v = 1; }
	}
	else
	{
		v = -Dot(a, ab) / denominator;
		u = F(1) - v;
	}
	if (v <= F(0))
		return a;
	if (u <= F(0))
		return b;
	return u * a + v * b;
}

// ACosApproximate (Math/Trigonometry.h:48-65)
template <class F> inline F ACosApprox(F x)
{
	F abs_x = std::min(std::abs(x), F(1));
	F val = std::sqrt(F(1) - abs_x) * (Constants<F>::kPi / F(2) - F(0.175394f) * abs_x);
	return x < F(0) ? Constants<F>::kPi - val : val;
}

// ------------------------------------------------------------------ matrices
// Column major: c[col] = {x, y, z, w}
template <class F> struct M44 { F c[4][4]; };

template <class F> inline V3<F> Col3(const M44<F> &m, int i) { return { m.c[i][0], m.c[i][1], m.c[i][2] }; }

// Mat44::sRotation(Quat) (Mat44.inl:85-141, scalar branch documented equal to the SSE4.1 one)
template <class F> M44<F> RotationTranslation(const F q[4], V3<F> t)
{
	F x = q[0], y = q[1], z = q[2], w = q[3];
	F tx = x + x, ty = y + y, tz = z + z;
	F xx = tx * x, yy = ty * y, zz = tz * z;
	F xy = tx * y, xz = tx * z, xw = tx * w;
	F yz = ty * z, yw = ty * w, zw = tz * w;
	M44<F> m;
	m.c[0][0] = (F(1) - yy) - zz; m.c[0][1] = xy + zw; m.c[0][2] = xz - yw; m.c[0][3] = 0;
	m.c[1][0] = xy - zw; m.c[1][1] = (F(1) - zz) - xx; m.c[1][2] = yz + xw; m.c[1][3] = 0;
	m.c[2][0] = xz + yw; m.c[2][1] = yz - xw; m.c[2][2] = (F(1) - xx) - yy; m.c[2][3] = 0;
	m.c[3][0] = t.x; m.c[3][1] = t.y; m.c[3][2] = t.z; m.c[3][3] = 1;	// sRotationTranslation (Mat44.inl:149-154)
	return m;
}

template <class F> inline V3<F> MulPoint(const M44<F> &m, V3<F> v)
{
	V3<F> r;
	for (int k = 0; k < 3; ++k)
		(&r.x)[k] = ((m.c[0][k] * v.x + m.c[1][k] * v.y) + m.c[2][k] * v.z) + m.c[3][k];
	return r;
}

template <class F> inline V3<F> Mul3x3(const M44<F> &m, V3<F> v)
{
	V3<F> r;
	for (int k = 0; k < 3; ++k)
		(&r.x)[k] = (m.c[0][k] * v.x + m.c[1][k] * v.y) + m.c[2][k] * v.z;
	return r;
}

template <class F> M44<F> Mul(const M44<F> &a, const M44<F> &b)
{
	M44<F> r;
	for (int i = 0; i < 4; ++i)
		for (int k = 0; k < 4; ++k)
			r.c[i][k] = ((a.c[0][k] * b.c[i][0] + a.c[1][k] * b.c[i][1]) + a.c[2][k] * b.c[i][2]) + a.c[3][k] * b.c[i][3];
	return r;
}

// Transposed3x3 (Mat44.inl:612-630): rotation transposed, translation column (0,0,0,1)
template <class F> M44<F> Transposed3x3(const M44<F> &m)
{
	M44<F> r;
	for (int i = 0; i < 3; ++i)
	{
		for (int k = 0; k < 3; ++k)
			r.c[i][k] = m.c[k][i];
		r.c[i][3] = 0;
	}
	r.c[3][0] = r.c[3][1] = r.c[3][2] = 0; r.c[3][3] = 1;
	return r;
}

// InversedRotationTranslation (Mat44.inl:944-949)
template <class F> M44<F> InvRT(const M44<F> &m)
{
	M44<F> r = Transposed3x3(m);
	V3<F> t = Neg(Mul3x3(r, Col3(m, 3)));
	r.c[3][0] = t.x; r.c[3][1] = t.y; r.c[3][2] = t.z; r.c[3][3] = 1;
	return r;
}

// GetQuaternion (Mat44.inl:998-1050)
template <class F> void GetQuaternion(const M44<F> &m, F q[4])
{
	auto e = [&](int col, int row) { return m.c[col][row]; };
	F tr = e(0, 0) + e(1, 1) + e(2, 2);
	if (tr >= F(0))
	{
		F s = std::sqrt(tr + F(1));
		F is = F(0.5) / s;
		q[0] = (e(1, 2) - e(2, 1)) * is; q[1] = (e(2, 0) - e(0, 2)) * is; q[2] = (e(0, 1) - e(1, 0)) * is; q[3] = F(0.5) * s;
		return;
	}
	int i = 0;
	if (e(1, 1) > e(0, 0)) i = 1;
	if (e(2, 2) > e(i, i)) i = 2;
	if (i == 0)
	{
		F s = std::sqrt(e(0, 0) - (e(1, 1) + e(2, 2)) + F(1));
		F is = F(0.5) / s;
		q[0] = F(0.5) * s; q[1] = (e(1, 0) + e(0, 1)) * is; q[2] = (e(0, 2) + e(2, 0)) * is; q[3] = (e(1, 2) - e(2, 1)) * is;
	}
	else if (i == 1)
	{
		F s = std::sqrt(e(1, 1) - (e(2, 2) + e(0, 0)) + F(1));
		F is = F(0.5) / s;
		q[0] = (e(1, 0) + e(0, 1)) * is; q[1] = F(0.5) * s; q[2] = (e(2, 1) + e(1, 2)) * is; q[3] = (e(2, 0) - e(0, 2)) * is;
	}
	else
	{
		F s = std::sqrt(e(2, 2) - (e(0, 0) + e(1, 1)) + F(1));
		F is = F(0.5) / s;
		q[0] = (e(0, 2) + e(2, 0)) * is; q[1] = (e(2, 1) + e(1, 2)) * is; q[2] = F(0.5) * s; q[3] = (e(0, 1) - e(1, 0)) * is;
	}
}

// Quat * Quat (Quat.inl:7-48, no FMA): x = (a w + b z) + (d x - c y), y = (b w + c x) + (d y - a z),
// z = (c w + a y) + (d z - b x), w = -(a x + b y) + (d w - c z) for lhs {a b c d}, rhs {x y z w}
template <class F> void QuatMul(const F l[4], const F r[4], F out[4])
{
	const F a = l[0], b = l[1], c = l[2], d = l[3];
	const F x = r[0], y = r[1], z = r[2], w = r[3];
	out[0] = (a * w + b * z) + (d * x - c * y);
	out[1] = (b * w + c * x) + (d * y - a * z);
	out[2] = (c * w + a * y) + (d * z - b * x);
	out[3] = -(a * x + b * y) + (d * w - c * z);
}

// Quat * Vec3 (Quat.inl:383-405): v + 2 (w (u x v) + u x (u x v)) for u = the vector part, both
// cross products as the difference of the swizzled products the SSE code forms
template <class F> V3<F> QuatRotate(const F q[4], V3<F> v)
{
	auto cross = [&](V3<F> a) { return V3<F>{ a.z * q[1] - q[2] * a.y, a.x * q[2] - q[0] * a.z, a.y * q[0] - q[1] * a.x }; };
	const V3<F> uv = cross(v);
	const V3<F> uuv = cross(uv);
	const V3<F> t = q[3] * uv + uuv;
	return v + (t + t);
}

// ------------------------------------------------------------------ planes / boxes
template <class F> struct PlaneT { V3<F> n; F c; };

// Plane::sFromPointAndNormal (Geometry/Plane.h:21)
template <class F> inline PlaneT<F> PlaneFromPointNormal(V3<F> p, V3<F> n) { return { n, -Dot(n, p) }; }
// Plane::GetTransformed (Geometry/Plane.h:42-46)
template <class F> inline PlaneT<F> PlaneTransformed(const PlaneT<F> &p, const M44<F> &m)
{
	V3<F> tn = Mul3x3(m, p.n);
	return { tn, p.c - Dot(Col3(m, 3), tn) };
}
// Plane::SignedDistance (Geometry/Plane.h:57)
template <class F> inline F SignedDistance(const PlaneT<F> &p, V3<F> x) { return Dot(x, p.n) + p.c; }

template <class F> struct Box3 { V3<F> mn, mx; };
template <class F> inline Box3<F> EmptyBox() { F m = F(FLT_MAX); return { { m, m, m }, { -m, -m, -m } }; }
template <class F> inline void Encapsulate(Box3<F> &b, V3<F> p) { b.mn = VMin(b.mn, p); b.mx = VMax(b.mx, p); }
// AABox::Transformed (Geometry/AABox.h:193-213)
template <class F> Box3<F> Transformed(const Box3<F> &b, const M44<F> &m)
{
	V3<F> nmin = Col3(m, 3), nmax = Col3(m, 3);
	for (int c = 0; c < 3; ++c)
	{
		V3<F> col = Col3(m, c);
		V3<F> a = col * Comp(b.mn, c);
		V3<F> bb = col * Comp(b.mx, c);
		nmin = nmin + VMin(a, bb);
		nmax = nmax + VMax(a, bb);
	}
	return { nmin, nmax };
}
// AABox::Overlaps (Geometry/AABox.h:164-167)
template <class F> inline bool Overlaps(const Box3<F> &a, const Box3<F> &b)
{
	return !(a.mn.x > b.mx.x || a.mn.y > b.mx.y || a.mn.z > b.mx.z || a.mx.x < b.mn.x || a.mx.y < b.mn.y || a.mx.z < b.mn.z);
}

// ------------------------------------------------------------------ records
struct VertexRec { float pos[3], vel[3], inv_mass; };
struct FaceRec { u32 v[3], material; };
struct EdgeRec { u32 v[2]; float rest, compliance; };
struct BendRec { u32 v[4]; float compliance, initial_angle; };
struct VolumeRec { u32 v[4]; float six_rest_volume, compliance; };
struct LRARec { u32 v[2]; float max_distance; };
struct GroupRec { u32 edge_end, lra_end, rod_ss_end, rod_bt_end, bend_end, volume_end, skinned_end; };

// ------------------------------------------------------------------ world
struct Collider
{
	int type;				// 0 plane, 1 sphere, 2 box, 3 capsule, 4 cylinder, 5 tapered capsule, 6 convex hull
	float params[8];		// as the shape holds them (tapered capsule: top / bottom centre, radii, sin, tan)
	// convex hull as the shape holds it: points relative to the centre of mass (xyz),
	// planes (normal xyz, constant), per face first entry / count into index
	std::vector<float> hull_points, hull_planes;
	std::vector<u32> face_first, face_count, index;
	float pos[3], rot[4];
	float friction, restitution;
	float bounds[6];		// world space bounds of the static body (min xyz, max xyz)
	// motion (EMotionType 0 static, 1 kinematic, 2 dynamic); dynamic: mass properties and
	// world velocities, fixed for an update call
	int motion = 0;
	float inv_mass = 0, inv_inertia_diag[3] = { 0, 0, 0 }, inertia_rot[4] = { 0, 0, 0, 1 };
	float lin_vel[3] = { 0, 0, 0 }, ang_vel[3] = { 0, 0, 0 };
	int is_sensor = 0;     // sensors never collide (DetermineCollidingShapes keeps them apart)
	float scale[3] = { 1, 1, 1 };   // ScaledShape around the shape (ScaledShape.cpp:142-145)
	// triangle tree (type 8) as MeshShape walks it: per node the four children's bounds
	// (min xyz, max xyz, 4 each) and references; triangles 9 floats each
	std::vector<float> tree_bounds;
	std::vector<u32> tree_child;
	std::vector<float> tree_tris;
	u32 tree_root = 0;
	bool tree_height_field = false;	// invalid child bounds (min y > max y) are never visited
	// a compound or decorated body: its leaf shapes in collection order, each with the steps
	// that place it relative to the body's centre of mass (kind 1 bounds test against v[0..5],
	// 2 p += q * v, 3 q = q * v, 4 p -= q * v); the leaves carry no pose or material of their own
	struct Op { u32 kind; float v[6]; };
	std::vector<Op> ops;
	std::vector<Collider> parts;
};

template <class F> struct Leaf
{
	// the body's leaf shapes (LeafShape, SoftBodyMotionProperties.h:187-195): transform in
	// soft-body local space and shape
	std::vector<std::pair<M44<F>, const Collider *>> shapes;
	size_t col_index;
	F friction, restitution;
	// CollidingShape fields of a dynamic body (SoftBodyMotionProperties.h:190-212)
	bool dynamic = false;
	F inv_mass = 0;
	M44<F> inv_inertia;		// translation column zero
	V3<F> com_translation, lin, ang, orig_lin, orig_ang;
	bool update_velocities = false;
};

template <class F> struct Body
{
	// shared settings (after Optimize)
	std::vector<FaceRec> faces;
	std::vector<EdgeRec> edges;
	std::vector<BendRec> bends;
	std::vector<VolumeRec> volumes;
	std::vector<LRARec> lras;
	std::vector<GroupRec> groups;

	// vertex state (SoftBodyVertex.h:34-41)
	std::vector<V3<F>> x, prev, v;
	std::vector<F> w;
	std::vector<PlaneT<F>> plane;
	std::vector<int> shape_idx;
	std::vector<F> largest;
	std::vector<char> contact;

	// creation settings (SoftBodyCreationSettings.h:61-72)
	u32 iterations;
	float linear_damping, max_linear_velocity, restitution, friction, pressure, gravity_factor, vertex_radius;
	bool update_position, allow_sleeping;

	// body
	V3<F> position;
	F rotation[4];
	Box3<F> local_bounds, local_predicted_bounds;
	F sleep_timer = 0;
	bool active = true;
	V3<F> linear_velocity{ 0, 0, 0 }, angular_velocity{ 0, 0, 0 };

	// per step context (SoftBodyUpdateContext.h:56-60)
	V3<F> gravity;
	F dt_sub;
	V3<F> disp_gravity;
	std::vector<Leaf<F>> leaves;

	// UpdateRigidBodyVelocities output per collider, summed over the steps of an update call
	std::vector<V3<F>> dyn_lin, dyn_ang;
	std::vector<int> dyn_updated;
};

template <class F> struct World
{
	V3<F> gravity{ 0, F(-9.81f), 0 };
	std::vector<Collider> colliders;
	std::vector<Body<F>> bodies;
};

// ClosestPoint::GetClosestPointOnTriangle (ClosestPoint.h:157-318, MustIncludeC false): the
// point of triangle (a, b, c) nearest the origin and its feature (bit 0 a, bit 1 b, bit 2 c,
// 7 the interior).  Without FMA, CrossPrecise is Cross.
template <class F> V3<F> ClosestPointOnTriangle(V3<F> in_a, V3<F> in_b, V3<F> in_c, u32 &out_set)
{
	auto clamp01 = [](F v) { F lo = v > F(0) ? v : F(0); return lo < F(1) ? lo : F(1); };
	V3<F> ba = in_a - in_b, bc = in_c - in_b;
	bool swap_ac = Dot(bc, bc) < Dot(ba, ba);
	V3<F> a = swap_ac ? in_c : in_a;
	V3<F> c = swap_ac ? in_a : in_c;
	V3<F> ab = in_b - a;
	V3<F> ac = c - a;
	V3<F> n = Cross(ab, ac);
	F n_len_sq = LenSq(n);
	if (n_len_sq < F(1.0e-10f))
	{
		u32 closest_set = 4;
		V3<F> closest_point = in_c;
		F best = LenSq(in_c);
		F a_len_sq = LenSq(in_a);
		if (a_len_sq < best) { closest_set = 1; closest_point = in_a; best = a_len_sq; }
		F b_len_sq = LenSq(in_b);
		if (b_len_sq < best) { closest_set = 2; closest_point = in_b; best = b_len_sq; }
		const F eps_sq = F(FLT_EPSILON) * F(FLT_EPSILON);
		F ac_len_sq = LenSq(ac);
		if (ac_len_sq > eps_sq)
		{
			F v = clamp01(-Dot(a, ac) / ac_len_sq);
			V3<F> q = a + v * ac;
			F d = LenSq(q);
			if (d < best) { closest_set = 5; closest_point = q; best = d; }
		}
		V3<F> bc2 = in_c - in_b;
		F bc_len_sq = LenSq(bc2);
		if (bc_len_sq > eps_sq)
		{
			F v = clamp01(-Dot(in_b, bc2) / bc_len_sq);
			V3<F> q = in_b + v * bc2;
			F d = LenSq(q);
			if (d < best) { closest_set = 6; closest_point = q; best = d; }
		}
		V3<F> ab2 = in_b - in_a;
		F ab_len_sq = LenSq(ab2);
		if (ab_len_sq > eps_sq)
		{
			F v = clamp01(-Dot(in_a, ab2) / ab_len_sq);
			V3<F> q = in_a + v * ab2;
			F d = LenSq(q);
			if (d < best) { closest_set = 3; closest_point = q; best = d; }
		}
		out_set = closest_set;
		return closest_point;
	}
	V3<F> ap = Neg(a);
	F d1 = Dot(ab, ap), d2 = Dot(ac, ap);
	if (d1 <= F(0) && d2 <= F(0)) { out_set = swap_ac ? 4 : 1; return a; }
	V3<F> bp = Neg(in_b);
	F d3 = Dot(ab, bp), d4 = Dot(ac, bp);
	if (d3 >= F(0) && d4 <= d3) { out_set = 2; return in_b; }
	if (d1 * d4 <= d3 * d2 && d1 >= F(0) && d3 <= F(0)) { F v = d1 / (d1 - d3); out_set = swap_ac ? 6 : 3; return a + v * ab; }
	V3<F> cp = Neg(c);
	F d5 = Dot(ab, cp), d6 = Dot(ac, cp);
	if (d6 >= F(0) && d5 <= d6) { out_set = swap_ac ? 1 : 4; return c; }
	if (d5 * d2 <= d1 * d6 && d2 >= F(0) && d6 <= F(0)) { F w = d2 / (d2 - d6); out_set = 5; return a + w * ac; }
	F d4_d3 = d4 - d3, d5_d6 = d5 - d6;
	if (d3 * d6 <= d5 * d4 && d4_d3 >= F(0) && d5_d6 >= F(0)) { F w = d4_d3 / (d4_d3 + d5_d6); out_set = swap_ac ? 3 : 6; return in_b + w * (c - in_b); }
	out_set = 7;
	return (n * Dot((a + in_b) + c, n)) / (F(3.0f) * n_len_sq);
}

template <class F> struct TriangleHit
{
	V3<F> v0, v1, v2, closest;
	F dist_sq = F(FLT_MAX);
	u32 set = 0;
};

// The closest triangle of a triangle tree to `local` as MeshShape's walk finds it
// (MeshShape.cpp:833-885, NodeCodecQuadTreeHalfFloat.h:245-318, SortReverseAndStore.h:15-33):
// recursion restated as the reference's explicit stack, children sorted far to near by the
// Vec4::sSort4Reverse network, pruned against the closest distance so far.
template <class F> TriangleHit<F> WalkTriangleTree(const Collider &c, V3<F> scale, V3<F> local)
{
	TriangleHit<F> hit;
	std::vector<u32> stack{ c.tree_root };
	std::vector<F> dist{ F(0) };
	int top = 0;
	do
	{
		const u32 ref = stack[top];
		const u32 kind = ref & 0xc0000000u;
		if (kind == 0u)
		{
			const float *bd = &c.tree_bounds[24 * size_t(ref)];
			F d[4];
			u32 id[4];
			for (int k = 0; k < 4; ++k)
			{
				F lo[3], hi[3];
				const F sc[3] = { scale.x, scale.y, scale.z };
				for (int a = 0; a < 3; ++a)
				{
					F s0 = sc[a] * F(bd[4 * a + k]), s1 = sc[a] * F(bd[12 + 4 * a + k]);
					lo[a] = SMin(s0, s1);
					hi[a] = SMax(s0, s1);
				}
				F cx = SMin(SMax(local.x, lo[0]), hi[0]) - local.x;
				F cy = SMin(SMax(local.y, lo[1]), hi[1]) - local.y;
				F cz = SMin(SMax(local.z, lo[2]), hi[2]) - local.z;
				d[k] = (cx * cx + cy * cy) + cz * cz;
				// a height field's block without collision (HeightFieldShape.cpp:2268-2269)
				if (c.tree_height_field && !(bd[4 + k] <= bd[16 + k]))
					d[k] = F(FLT_MAX);
				id[k] = c.tree_child[4 * size_t(ref) + k];
			}
			auto exchange = [&](int p, int q) { if (d[q] > d[p]) { std::swap(d[p], d[q]); std::swap(id[p], id[q]); } };
			exchange(0, 2); exchange(1, 3);
			exchange(0, 1); exchange(2, 3);
			exchange(1, 2);
			int n = 0;
			for (int k = 0; k < 4; ++k)
				n += d[k] < hit.dist_sq ? 1 : 0;
			if (size_t(top + n) > stack.size())
			{
				stack.resize(top + n);
				dist.resize(top + n);
			}
			for (int j = 0; j < n; ++j)
			{
				stack[top + j] = id[4 - n + j];
				dist[top + j] = d[4 - n + j];
			}
			top += n;
		}
		else if (kind == 0x40000000u)
		{
			const u32 first = ref & 0x00ffffffu, count = (ref >> 24) & 0x3fu;
			for (u32 t = first; t < first + count; ++t)
			{
				const float *tv = &c.tree_tris[9 * size_t(t)];
				V3<F> v0 = Mul(scale, V3<F>{ F(tv[0]), F(tv[1]), F(tv[2]) });
				V3<F> v1 = Mul(scale, V3<F>{ F(tv[3]), F(tv[4]), F(tv[5]) });
				V3<F> v2 = Mul(scale, V3<F>{ F(tv[6]), F(tv[7]), F(tv[8]) });
				u32 set;
				V3<F> cp = ClosestPointOnTriangle(v0 - local, v1 - local, v2 - local, set);
				F d2 = LenSq(cp);
				if (d2 < hit.dist_sq)
				{
					hit.v0 = v0; hit.v1 = v1; hit.v2 = v2;
					hit.closest = cp;
					hit.dist_sq = d2;
					hit.set = set;
				}
			}
		}
		do
			--top;
		while (top >= 0 && !(dist[top] < hit.dist_sq));
	} while (top >= 0);
	return hit;
}

// ------------------------------------------------------------------ shape collision
// Restates Shape::CollideSoftBodyVertices for the four primitive shapes, with
// UpdatePenetration / SetCollision of CollideSoftBodyVertexIterator.h:83-97.
template <class F> void CollideVertices(Body<F> &b, const Collider &c, const M44<F> &T, int shape_index)
{
	const size_t nv = b.x.size();
	const V3<F> scale{ F(c.scale[0]), F(c.scale[1]), F(c.scale[2]) };
	const V3<F> abs_scale = VAbs(scale);
	auto update = [&](size_t i, F pen) {
		if (b.largest[i] >= pen)
			return false;
		b.largest[i] = pen;
		return true;
	};
	switch (c.type)
	{
	case 0: // PlaneShape.cpp:243-256 with Plane::Scaled (Plane.h:49-54) and GetTransformed
		{
			PlaneT<F> p{ { F(c.params[0]), F(c.params[1]), F(c.params[2]) }, F(c.params[3]) };
			V3<F> sn{ p.n.x / scale.x, p.n.y / scale.y, p.n.z / scale.z };
			F snl = Len(sn);
			PlaneT<F> scaled{ sn / snl, p.c / snl };
			PlaneT<F> plane = PlaneTransformed(scaled, T);
			for (size_t i = 0; i < nv; ++i)
				if (b.w[i] > F(0))
				{
					F pen = -SignedDistance(plane, b.x[i]);
					if (update(i, pen)) { b.plane[i] = plane; b.shape_idx[i] = shape_index; }
				}
		}
		break;
	case 1: // SphereShape.cpp:279-301
		{
			V3<F> center = Col3(T, 3);
			F radius = abs_scale.x * F(c.params[0]);	// SphereShape::GetScaledRadius
			for (size_t i = 0; i < nv; ++i)
				if (b.w[i] > F(0))
				{
					V3<F> delta = b.x[i] - center;
					F distance = Len(delta);
					F pen = radius - distance;
					if (update(i, pen))
					{
						V3<F> normal = distance > F(0) ? delta / distance : V3<F>{ 0, 1, 0 };
						V3<F> point = center + radius * normal;
						b.plane[i] = PlaneFromPointNormal(point, normal);
						b.shape_idx[i] = shape_index;
					}
				}
		}
		break;
	case 2: // BoxShape.cpp:234-283
		{
			M44<F> inv = InvRT(T);
			V3<F> he{ F(c.params[0]), F(c.params[1]), F(c.params[2]) };
			he = Mul(abs_scale, he);
			for (size_t i = 0; i < nv; ++i)
				if (b.w[i] > F(0))
				{
					V3<F> local = MulPoint(inv, b.x[i]);
					V3<F> clamped = VMax(VMin(local, he), Neg(he));
					if (clamped.x == local.x && clamped.y == local.y && clamped.z == local.z)
					{
						V3<F> delta = he - VAbs(local);
						int index = LowestIndex(delta);
						F pen = Comp(delta, index);
						if (update(i, pen))
						{
							V3<F> axis{ F(index == 0), F(index == 1), F(index == 2) };
							V3<F> normal = Mul(VSign(local), axis);
							V3<F> point = Mul(normal, he);
							b.plane[i] = PlaneTransformed(PlaneFromPointNormal(point, normal), T);
							b.shape_idx[i] = shape_index;
						}
					}
					else
					{
						V3<F> normal = local - clamped;
						F nl = Len(normal);
						F pen = -nl;
						if (update(i, pen))
						{
							normal = normal / nl;
							b.plane[i] = PlaneTransformed(PlaneFromPointNormal(clamped, normal), T);
							b.shape_idx[i] = shape_index;
						}
					}
				}
		}
		break;
	case 3: // CapsuleShape.cpp:324-375
		{
			M44<F> inv = InvRT(T);
			F s_xz = std::abs(scale.x);
			F hh = s_xz * F(c.params[0]);
			F radius = s_xz * F(c.params[1]);
			for (size_t i = 0; i < nv; ++i)
				if (b.w[i] > F(0))
				{
					V3<F> local = MulPoint(inv, b.x[i]);
					if (std::abs(local.y) <= hh)
					{
						V3<F> normal{ local.x, 0, local.z };
						F nl = Len(normal);
						F pen = radius - nl;
						if (update(i, pen))
						{
							normal = nl > F(0) ? normal / nl : V3<F>{ 1, 0, 0 };
							V3<F> point = radius * normal;
							b.plane[i] = PlaneTransformed(PlaneFromPointNormal(point, normal), T);
							b.shape_idx[i] = shape_index;
						}
					}
					else
					{
						V3<F> center{ 0, SignOf(local.y) * hh, 0 };
						V3<F> delta = local - center;
						F distance = Len(delta);
						F pen = radius - distance;
						if (update(i, pen))
						{
							V3<F> normal = delta / distance;
							V3<F> point = center + radius * normal;
							b.plane[i] = PlaneTransformed(PlaneFromPointNormal(point, normal), T);
							b.shape_idx[i] = shape_index;
						}
					}
				}
		}
		break;
	case 4: // CylinderShape.cpp:317-366
		{
			M44<F> inv = InvRT(T);
			F half_height = abs_scale.y * F(c.params[0]);
			F radius = abs_scale.x * F(c.params[1]);
			for (size_t i = 0; i < nv; ++i)
				if (b.w[i] > F(0))
				{
					V3<F> local = MulPoint(inv, b.x[i]);
					V3<F> side_normal{ local.x, 0, local.z };
					F side_normal_length = Len(side_normal);
					F side_penetration = radius - side_normal_length;
					F top_penetration = half_height - std::abs(local.y);
					V3<F> point, normal;
					if (side_penetration < F(0) && top_penetration < F(0))
					{
						point = side_normal * (radius / side_normal_length) + V3<F>{ 0, half_height * SignOf(local.y), 0 };
						normal = NormalizedOr(local - point, V3<F>{ 0, 1, 0 });
					}
					else if (side_penetration < top_penetration)
					{
						normal = side_normal_length > F(0) ? side_normal / side_normal_length : V3<F>{ 1, 0, 0 };
						point = radius * normal;
					}
					else
					{
						normal = V3<F>{ 0, SignOf(local.y), 0 };
						point = half_height * normal;
					}
					PlaneT<F> plane = PlaneFromPointNormal(point, normal);
					F pen = -SignedDistance(plane, local);
					if (update(i, pen))
					{
						b.plane[i] = PlaneTransformed(plane, T);
						b.shape_idx[i] = shape_index;
					}
				}
		}
		break;
	case 5: // TaperedCapsuleShape.cpp:304-371
		{
			M44<F> inv = InvRT(T);
			V3<F> top_center{ 0, abs_scale.y * F(c.params[0]), 0 }, bottom_center{ 0, abs_scale.y * F(c.params[1]), 0 };
			F top_radius = abs_scale.x * F(c.params[2]), bottom_radius = abs_scale.x * F(c.params[3]);
			F sin_alpha = F(c.params[4]), tan_alpha = F(c.params[5]);
			V3<F> scale_y_flip{ 1, SignOf(scale.y), 1 };
			for (size_t i = 0; i < nv; ++i)
				if (b.w[i] > F(0))
				{
					V3<F> local = Mul(scale_y_flip, MulPoint(inv, b.x[i]));
					V3<F> position, normal;
					V3<F> to_top = local - top_center;
					F to_top_len = Len(to_top);
					if (to_top.y >= sin_alpha * to_top_len)
					{
						normal = to_top_len != F(0) ? to_top / to_top_len : V3<F>{ 0, 1, 0 };
						position = top_center + top_radius * normal;
					}
					else
					{
						V3<F> to_bottom = local - bottom_center;
						F to_bottom_len = Len(to_bottom);
						if (to_bottom.y <= sin_alpha * to_bottom_len)
							normal = to_bottom_len != F(0) ? to_bottom / to_bottom_len : V3<F>{ F(0) - F(0), F(0) - F(1), F(0) - F(0) };
						else
						{
							normal = NormalizedOr(V3<F>{ local.x, 0, local.z }, V3<F>{ 1, 0, 0 });
							normal.y = tan_alpha;
							normal = NormalizedOr(normal, V3<F>{ 1, 0, 0 });
						}
						position = bottom_center + bottom_radius * normal;
					}
					PlaneT<F> plane = PlaneFromPointNormal(position, normal);
					F pen = -SignedDistance(plane, local);
					if (update(i, pen))
					{
						plane.n = Mul(scale_y_flip, plane.n);
						b.plane[i] = PlaneTransformed(plane, T);
						b.shape_idx[i] = shape_index;
					}
				}
		}
		break;
	case 6: // ConvexHullShape.cpp:1060-1164
		{
			M44<F> inv = InvRT(T);
			const size_t nf = c.face_count.size();
			const V3<F> inv_scale{ F(1) / scale.x, F(1) / scale.y, F(1) / scale.z };	// Vec3::Reciprocal
			const bool is_not_scaled = LenSq(scale - V3<F>{ 1, 1, 1 }) <= F(1.0e-8f);	// ScaleHelpers::IsNotScaled
			const int negatives = int(scale.x < F(0)) + int(scale.y < F(0)) + int(scale.z < F(0));
			const F scale_flip = (negatives & 1) ? F(-1) : F(1);						// ScaleHelpers::IsInsideOut
			auto point = [&](u32 k) { return V3<F>{ F(c.hull_points[3 * k]), F(c.hull_points[3 * k + 1]), F(c.hull_points[3 * k + 2]) }; };
			for (size_t i = 0; i < nv; ++i)
				if (b.w[i] > F(0))
				{
					V3<F> local = MulPoint(inv, b.x[i]);
					// most facing plane, first of equals (strict >)
					F max_distance = -F(FLT_MAX);
					V3<F> max_plane_normal{ 0, 0, 0 };
					size_t max_plane_idx = 0;
					for (size_t f = 0; f < nf; ++f)
					{
						V3<F> n{ F(c.hull_planes[4 * f]), F(c.hull_planes[4 * f + 1]), F(c.hull_planes[4 * f + 2]) };
						F distance;
						V3<F> plane_normal;
						if (is_not_scaled)
						{
							PlaneT<F> p{ n, F(c.hull_planes[4 * f + 3]) };
							distance = SignedDistance(p, local);
							plane_normal = n;
						}
						else
						{
							// the plane rebuilt from the scaled normal and the scaled first point of the face
							V3<F> sn = Mul(inv_scale, n);
							plane_normal = sn / Len(sn);
							V3<F> plane_point = Mul(scale, point(c.index[c.face_first[f]]));
							distance = Dot(plane_normal, local - plane_point);
						}
						if (distance > max_distance)
						{
							max_distance = distance;
							max_plane_normal = plane_normal;
							max_plane_idx = f;
						}
					}
					V3<F> closest_point = (-max_distance) * max_plane_normal;
					const bool is_outside = max_distance > F(0);
					if (is_outside)
					{
						// edges of the face the vertex lies beyond; nearest wins, first of equals
						F closest_point_dist_sq = F(FLT_MAX);
						const u32 first = c.face_first[max_plane_idx], count = c.face_count[max_plane_idx];
						for (u32 k = 0; k < count; ++k)
						{
							V3<F> p1 = Mul(scale, point(c.index[first + k]));
							V3<F> p2 = Mul(scale, point(c.index[first + (k + 1 == count ? 0 : k + 1)]));
							V3<F> edge_normal = Cross(p2 - p1, max_plane_normal);
							if (scale_flip * Dot(edge_normal, local - p1) > F(0))
							{
								V3<F> closest = ClosestPointOnLine(p1 - local, p2 - local);
								F distance_sq = LenSq(closest);
								if (distance_sq < closest_point_dist_sq)
								{
									closest_point_dist_sq = distance_sq;
									closest_point = closest;
								}
							}
						}
					}
					V3<F> normal = Neg(closest_point);
					F normal_length = Len(normal);
					F penetration = normal_length;
					if (is_outside)
						penetration = -penetration;
					else
						normal = Neg(normal);
					if (update(i, penetration))
					{
						normal = normal_length > F(1.0e-12f) ? normal / normal_length : max_plane_normal;
						b.plane[i] = PlaneTransformed(PlaneFromPointNormal(local + closest_point, normal), T);
						b.shape_idx[i] = shape_index;
					}
				}
		}
		break;
	case 7: // TaperedCylinderShape.cpp:390-506 with GetScaled (:181-199)
		{
			M44<F> inv = InvRT(T);
			F top = scale.y * F(c.params[0]), bottom = scale.y * F(c.params[1]);
			F top_radius = abs_scale.x * F(c.params[2]), bottom_radius = abs_scale.x * F(c.params[3]);
			if (bottom > top)
			{
				std::swap(top, bottom);
				std::swap(top_radius, bottom_radius);
			}
			const V3<F> top_3d{ 0, top, 0 }, bottom_3d{ 0, bottom, 0 };
			for (size_t i = 0; i < nv; ++i)
				if (b.w[i] > F(0))
				{
					V3<F> local = MulPoint(inv, b.x[i]);
					V3<F> normal_xz = NormalizedOr(Mul(V3<F>{ 1, 0, 1 }, local), V3<F>{ 1, 0, 0 });
					F tan_alpha = (bottom_radius - top_radius) / (top - bottom);
					V3<F> sn{ normal_xz.x, tan_alpha, normal_xz.z };
					V3<F> side_normal = sn / Len(sn);
					V3<F> side_support_top = normal_xz * top_radius + top_3d;
					F side_penetration = Dot(side_support_top - local, side_normal);
					F top_penetration = top - local.y;
					F bottom_penetration = local.y - bottom;
					F min_top_bottom_penetration = std::min(top_penetration, bottom_penetration);
					V3<F> point, normal;
					if (side_penetration < F(0) || min_top_bottom_penetration < F(0))
					{
						V3<F> side_support_bottom = normal_xz * bottom_radius + bottom_3d;
						V3<F> bottom_to_top = side_support_top - side_support_bottom;
						F fraction = Dot(local - side_support_bottom, bottom_to_top);
						F distance_to_axis = Dot(normal_xz, local);
						bool inside_top_radius = distance_to_axis <= top_radius;
						bool inside_bottom_radius = distance_to_axis <= bottom_radius;
						if (fraction >= LenSq(bottom_to_top) && !inside_top_radius)
						{
							point = side_support_top;
							normal = NormalizedOr(local - point, V3<F>{ 0, 1, 0 });
						}
						else if (fraction < F(0) && !inside_bottom_radius)
						{
							point = side_support_bottom;
							normal = NormalizedOr(local - point, V3<F>{ 0, 1, 0 });
						}
						else if (top_penetration < F(0) && inside_top_radius)
						{
							point = top_3d;
							normal = V3<F>{ 0, 1, 0 };
						}
						else if (bottom_penetration < F(0) && inside_bottom_radius)
						{
							point = bottom_3d;
							normal = V3<F>{ 0, -1, 0 };
						}
						else
						{
							point = side_support_top;
							normal = side_normal;
						}
					}
					else if (side_penetration < min_top_bottom_penetration)
					{
						point = side_support_top;
						normal = side_normal;
					}
					else if (top_penetration < bottom_penetration)
					{
						point = top_3d;
						normal = V3<F>{ 0, 1, 0 };
					}
					else
					{
						point = bottom_3d;
						normal = V3<F>{ 0, -1, 0 };
					}
					PlaneT<F> plane = PlaneFromPointNormal(point, normal);
					F pen = -SignedDistance(plane, local);
					if (update(i, pen))
					{
						b.plane[i] = PlaneTransformed(plane, T);
						b.shape_idx[i] = shape_index;
					}
				}
		}
		break;
	case 8: // MeshShape.cpp:833-885 / TriangleShape.cpp:272-283 through CollideSoftBodyVerticesVsTriangles
		{
			M44<F> inv = InvRT(T);
			const int negatives = int(scale.x < F(0)) + int(scale.y < F(0)) + int(scale.z < F(0));
			const F normal_sign = (negatives & 1) ? F(-1) : F(1);
			for (size_t i = 0; i < nv; ++i)
				if (b.w[i] > F(0))
				{
					const V3<F> local = MulPoint(inv, b.x[i]);
					TriangleHit<F> hit = WalkTriangleTree(c, scale, local);
					if (!(hit.dist_sq < F(FLT_MAX)))
						continue;
					// FinishVertex: back to the soft body's space
					V3<F> v0 = MulPoint(T, hit.v0), v1 = MulPoint(T, hit.v1), v2 = MulPoint(T, hit.v2);
					V3<F> triangle_normal = normal_sign * NormalizedOr(Cross(v1 - v0, v2 - v0), V3<F>{ 0, 1, 0 });
					if (hit.set == 7u)
					{
						// interior: the triangle's plane, at most sTriangleThickness (0.1) deep
						F penetration = Dot(triangle_normal, v0 - b.x[i]);
						if (penetration < F(0.1f) && update(i, penetration))
						{
							b.plane[i] = PlaneFromPointNormal(v0, triangle_normal);
							b.shape_idx[i] = shape_index;
						}
					}
					else
					{
						// edge or vertex: the plane through the closest point facing the vertex
						V3<F> closest_point = MulPoint(T, local + hit.closest);
						V3<F> normal = b.x[i] - closest_point;
						if (Dot(normal, triangle_normal) > F(0))
						{
							F normal_length = Len(normal);
							F penetration = -normal_length;
							if (update(i, penetration))
							{
								b.plane[i] = PlaneFromPointNormal(closest_point, normal_length > F(0) ? normal / normal_length : triangle_normal);
								b.shape_idx[i] = shape_index;
							}
						}
					}
				}
		}
		break;
	}
}

// ------------------------------------------------------------------ solver pieces
// GetVolumeTimesSix (SoftBodyMotionProperties.cpp:107-118), sequential sum
template <class F> F VolumeTimesSix(const Body<F> &b)
{
	F six = 0;
	for (const FaceRec &f : b.faces)
		six += Dot(Cross(b.x[f.v[0]], b.x[f.v[1]]), b.x[f.v[2]]);
	return six;
}

// ApplyPressure (SoftBodyMotionProperties.cpp:291-323)
template <class F> void ApplyPressure(Body<F> &b)
{
	F dt = b.dt_sub;
	F coef_p = F(b.pressure);
	if (coef_p > F(0))
	{
		F six = VolumeTimesSix(b);
		if (six > F(0))
		{
			F coefficient = coef_p * dt / six;
			for (const FaceRec &f : b.faces)
			{
				V3<F> x1 = b.x[f.v[0]], x2 = b.x[f.v[1]], x3 = b.x[f.v[2]];
				V3<F> impulse = coefficient * Cross(x2 - x1, x3 - x1);
				for (u32 i : f.v)
					b.v[i] = b.v[i] + b.w[i] * impulse;
			}
		}
	}
}

// IntegratePositions (SoftBodyMotionProperties.cpp:325-365); accumulated force is zero
template <class F> void Integrate(Body<F> &b)
{
	F dt = b.dt_sub;
	F damping = std::max(F(0), F(1) - F(b.linear_damping) * dt);
	V3<F> gdt = b.gravity * dt;
	V3<F> imp = V3<F>{ 0, 0, 0 } * dt / std::max(F(b.x.size()), F(1));
	for (size_t i = 0; i < b.x.size(); ++i)
	{
		if (b.w[i] > F(0))
		{
			b.v[i] = b.v[i] + (gdt + imp * b.w[i]);
			b.v[i] = b.v[i] * damping;
		}
		V3<F> p = b.x[i];
		b.prev[i] = p;
		b.x[i] = p + b.v[i] * dt;
	}
}

// ApplyDihedralBendConstraints (SoftBodyMotionProperties.cpp:367-456)
template <class F> void ApplyBends(Body<F> &b, u32 begin, u32 end)
{
	F inv_dt_sq = F(1) / (b.dt_sub * b.dt_sub);
	const F pi = Constants<F>::kPi;
	for (u32 k = begin; k < end; ++k)
	{
		const BendRec &r = b.bends[k];
		V3<F> x0 = b.x[r.v[0]], x1 = b.x[r.v[1]], x2 = b.x[r.v[2]], x3 = b.x[r.v[3]];
		V3<F> e = x1 - x0;
		F e_len = Len(e);
		if (e_len < F(1.0e-6f))
			continue;
		V3<F> x1x2 = x2 - x1, x1x3 = x3 - x1;
		V3<F> n1 = Cross(x2 - x0, x1x2);
		V3<F> n2 = Cross(x1x3, x3 - x0);
		F n1sq = LenSq(n1), n2sq = LenSq(n2);
		F n1n2 = n1sq * n2sq;
		if (n1n2 < F(1.0e-24f))
			continue;
		F sign = SignOf(Dot(Cross(n2, n1), e));
		F d = Dot(n1, n2) / std::sqrt(n1n2);
		F c = sign * ACosApprox(d) - F(r.initial_angle);
		if (c > pi)
			c -= F(2) * pi;
		else if (c < -pi)
			c += F(2) * pi;
		n1 = n1 / n1sq;
		n2 = n2 / n2sq;
		V3<F> d0c = (Dot(x1x2, e) * n1 + Dot(x1x3, e) * n2) / e_len;
		V3<F> d2c = e_len * n1;
		V3<F> d3c = e_len * n2;
		V3<F> d1c = (Neg(d0c) - d2c) - d3c;
		F w0 = b.w[r.v[0]], w1 = b.w[r.v[1]], w2 = b.w[r.v[2]], w3 = b.w[r.v[3]];
		F denom = (((w0 * LenSq(d0c) + w1 * LenSq(d1c)) + w2 * LenSq(d2c)) + w3 * LenSq(d3c)) + F(r.compliance) * inv_dt_sq;
		if (denom < F(1.0e-12f))
			continue;
		F ml = c / denom;
		b.x[r.v[0]] = x0 - (ml * w0) * d0c;
		b.x[r.v[1]] = x1 - (ml * w1) * d1c;
		b.x[r.v[2]] = x2 - (ml * w2) * d2c;
		b.x[r.v[3]] = x3 - (ml * w3) * d3c;
	}
}

// ApplyVolumeConstraints (SoftBodyMotionProperties.cpp:458-507)
template <class F> void ApplyVolumes(Body<F> &b, u32 begin, u32 end)
{
	F inv_dt_sq = F(1) / (b.dt_sub * b.dt_sub);
	for (u32 k = begin; k < end; ++k)
	{
		const VolumeRec &r = b.volumes[k];
		V3<F> x1 = b.x[r.v[0]], x2 = b.x[r.v[1]], x3 = b.x[r.v[2]], x4 = b.x[r.v[3]];
		V3<F> x1x2 = x2 - x1, x1x3 = x3 - x1, x1x4 = x4 - x1;
		F c = std::abs(Dot(Cross(x1x2, x1x3), x1x4)) - F(r.six_rest_volume);
		V3<F> d1c = Cross(x4 - x2, x3 - x2);
		V3<F> d2c = Cross(x1x3, x1x4);
		V3<F> d3c = Cross(x1x4, x1x2);
		V3<F> d4c = Cross(x1x2, x1x3);
		F w1 = b.w[r.v[0]], w2 = b.w[r.v[1]], w3 = b.w[r.v[2]], w4 = b.w[r.v[3]];
		F denom = (((w1 * LenSq(d1c) + w2 * LenSq(d2c)) + w3 * LenSq(d3c)) + w4 * LenSq(d4c)) + F(r.compliance) * inv_dt_sq;
		if (denom < F(1.0e-12f))
			continue;
		F ml = c / denom;
		b.x[r.v[0]] = x1 - (ml * w1) * d1c;
		b.x[r.v[1]] = x2 - (ml * w2) * d2c;
		b.x[r.v[2]] = x3 - (ml * w3) * d3c;
		b.x[r.v[3]] = x4 - (ml * w4) * d4c;
	}
}

// ApplyEdgeConstraints (SoftBodyMotionProperties.cpp:573-601)
template <class F> void ApplyEdges(Body<F> &b, u32 begin, u32 end)
{
	F inv_dt_sq = F(1) / (b.dt_sub * b.dt_sub);
	for (u32 k = begin; k < end; ++k)
	{
		const EdgeRec &r = b.edges[k];
		V3<F> x0 = b.x[r.v[0]], x1 = b.x[r.v[1]];
		V3<F> delta = x1 - x0;
		F length = Len(delta);
		F w0 = b.w[r.v[0]], w1 = b.w[r.v[1]];
		F denom = length * ((w0 + w1) + F(r.compliance) * inv_dt_sq);
		if (denom < F(1.0e-12f))
			continue;
		V3<F> corr = (delta * (length - F(r.rest))) / denom;
		b.x[r.v[0]] = x0 + w0 * corr;
		b.x[r.v[1]] = x1 - w1 * corr;
	}
}

// ApplyLRAConstraints (SoftBodyMotionProperties.cpp:679-698)
template <class F> void ApplyLRAs(Body<F> &b, u32 begin, u32 end)
{
	for (u32 k = begin; k < end; ++k)
	{
		const LRARec &r = b.lras[k];
		V3<F> x0 = b.x[r.v[0]];
		V3<F> delta = b.x[r.v[1]] - x0;
		F dsq = LenSq(delta);
		F md = F(r.max_distance);
		if (dsq > md * md)
			b.x[r.v[1]] = x0 + (delta * md) / std::sqrt(dsq);
	}
}

// ProcessGroup (SoftBodyMotionProperties.cpp:1030-1055): volume -> bend -> skin -> edge -> rods -> LRA
template <class F> void ProcessGroup(Body<F> &b, size_t g)
{
	GroupRec start{ 0, 0, 0, 0, 0, 0, 0 };
	const GroupRec &p = g > 0 ? b.groups[g - 1] : start;
	const GroupRec &c = b.groups[g];
	ApplyVolumes(b, p.volume_end, c.volume_end);
	ApplyBends(b, p.bend_end, c.bend_end);
	ApplyEdges(b, p.edge_end, c.edge_end);
	ApplyLRAs(b, p.lra_end, c.lra_end);
}

// Dynamic-body branch of ApplyCollisionConstraintsAndUpdateVelocities
// (SoftBodyMotionProperties.cpp:752-797), every contact-settings scale 1.
template <class F> void DynamicContact(Leaf<F> &cs, V3<F> x, V3<F> &v, V3<F> prev_v, V3<F> n, F pd, F w, F dt, F restitution_threshold)
{
	V3<F> r2 = x - cs.com_translation;
	V3<F> v2 = cs.lin + Cross(cs.ang, r2); // CollidingShape::GetPointVelocity
	V3<F> relative_velocity = v - v2;
	V3<F> v_normal = n * Dot(n, relative_velocity);
	V3<F> v_tangential = relative_velocity - v_normal;
	F v_tangential_length = Len(v_tangential);
	F vertex_inv_mass = F(1) * w;
	V3<F> r2_cross_n = Cross(r2, n);
	F w2 = cs.inv_mass + Dot(r2_cross_n, MulPoint(cs.inv_inertia, r2_cross_n));
	F w1_plus_w2 = vertex_inv_mass + w2;
	if (w1_plus_w2 > F(0))
	{
		V3<F> dv;
		if (v_tangential_length > F(0))
			dv = v_tangential * std::min(cs.friction * pd / (v_tangential_length * dt), F(1));
		else
			dv = V3<F>{ 0, 0, 0 };
		dv = dv + v_normal;
		F prev_v_normal = Dot(prev_v - v2, n);
		if (prev_v_normal < restitution_threshold)
			dv = dv + (cs.restitution * prev_v_normal) * n;
		V3<F> p = dv / w1_plus_w2;
		v = v - p * vertex_inv_mass;
		cs.lin = cs.lin + p * cs.inv_mass;
		cs.ang = cs.ang + MulPoint(cs.inv_inertia, Cross(r2, p));
		cs.update_velocities = true;
	}
}

// ApplyCollisionConstraintsAndUpdateVelocities (SoftBodyMotionProperties.cpp:700-825): the
// static / kinematic branch, and the dynamic branch in vertex order.
template <class F> void CollisionAndVelocity(Body<F> &b)
{
	F dt = b.dt_sub;
	F restitution_threshold = F(-2) * Len(b.gravity) * dt;
	F r_vtx = F(b.vertex_radius);
	for (size_t i = 0; i < b.x.size(); ++i)
		if (b.w[i] > F(0))
		{
			V3<F> prev_v = b.v[i];
			b.v[i] = (b.x[i] - b.prev[i]) / dt;
			if (b.shape_idx[i] >= 0)
			{
				F pd = -SignedDistance(b.plane[i], b.x[i]) + r_vtx;
				if (pd > F(0))
				{
					b.contact[i] = 1;
					Leaf<F> &cs = b.leaves[b.shape_idx[i]];
					V3<F> n = b.plane[i].n;
					b.x[i] = b.x[i] + n * pd;
					if (cs.dynamic)
					{
						DynamicContact(cs, b.x[i], b.v[i], prev_v, n, pd, b.w[i], dt, restitution_threshold);
						continue;
					}
					// static or kinematic body; mSoftBodyInvMassScale = 1 > 0
					V3<F> vn = n * Dot(n, b.v[i]);
					V3<F> vt = b.v[i] - vn;
					F vtl = Len(vt);
					if (vtl > F(0))
						b.v[i] = b.v[i] - vt * std::min(cs.friction * pd / (vtl * dt), F(1));
					b.v[i] = b.v[i] - vn;
					F pvn = Dot(prev_v, n);
					if (pvn < restitution_threshold)
						b.v[i] = b.v[i] - (cs.restitution * pvn) * n;
				}
			}
		}
}

// UpdateSoftBodyState (SoftBodyMotionProperties.cpp:827-921); step_dt = context mDeltaTime
template <class F> V3<F> UpdateState(Body<F> &b, F step_dt, const M44<F> &com)
{
	F max_lin_sq = F(b.max_linear_velocity) * F(b.max_linear_velocity);
	F max_v_sq = 0;
	V3<F> lin{ 0, 0, 0 }, ang{ 0, 0, 0 };
	b.local_bounds = EmptyBox<F>();
	b.local_predicted_bounds = EmptyBox<F>();
	for (size_t i = 0; i < b.x.size(); ++i)
	{
		F v_sq = LenSq(b.v[i]);
		max_v_sq = std::max(max_v_sq, v_sq);
		if (v_sq > max_lin_sq)
			b.v[i] = b.v[i] * std::sqrt(max_lin_sq / v_sq);
		lin = lin + b.v[i];
		ang = ang + Cross(b.x[i], b.v[i]);
		Encapsulate(b.local_bounds, b.x[i]);
		Encapsulate(b.local_predicted_bounds, (b.x[i] + b.v[i] * step_dt) + b.disp_gravity);
		b.largest[i] = F(-FLT_MAX);
		b.shape_idx[i] = -1;
		b.contact[i] = 0;
	}
	F nvd = F(std::max(int(b.x.size()), 1));
	b.linear_velocity = Mul3x3(com, lin / nvd);
	b.angular_velocity = Mul3x3(com, ang / nvd);
	V3<F> delta_world{ 0, 0, 0 };
	if (b.update_position)
	{
		V3<F> delta = F(0.5) * (b.local_bounds.mn + b.local_bounds.mx);
		delta_world = Mul3x3(com, delta);
		for (size_t i = 0; i < b.x.size(); ++i)
			b.x[i] = b.x[i] - delta;
		b.local_bounds.mn = b.local_bounds.mn + Neg(delta);
		b.local_bounds.mx = b.local_bounds.mx + Neg(delta);
		b.local_predicted_bounds.mn = b.local_predicted_bounds.mn + Neg(delta);
		b.local_predicted_bounds.mx = b.local_predicted_bounds.mx + Neg(delta);
	}
	if (b.allow_sleeping)
	{
		if (max_v_sq > F(0.03f))
			b.sleep_timer = 0;
		else
		{
			b.sleep_timer += step_dt;
			if (b.sleep_timer >= F(0.5f))
				b.active = false;	// ECanSleep::CanSleep -> deactivated (PhysicsSystem.cpp:2866-2876)
		}
	}
	return delta_world;
}

// AABox4VsBox (Geometry/AABox4.h:86-185) for a single box: do the bounds [bmin, bmax] of a compound
// child overlap the oriented box `orientation` * [-half, half]?  Separating axis test over the
// box axes A, the oriented box axes B and their 9 cross products, with 1e-6 added to every
// |rotation| entry.  R(i, j) below is row i, column j of the orientation.
template <class F> bool BoxVsOrientedBox(const M44<F> &orientation, V3<F> half, const F bmin[3], const F bmax[3])
{
	auto R = [&](int row, int col) { return orientation.c[col][row]; };
	const F eps = F(1.0e-6f);
	F abs_r[3][3];		// abs_r[axis of B][component]
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
			abs_r[i][j] = std::abs(R(j, i)) + eps;
	F t[3], a[3];
	for (int i = 0; i < 3; ++i)
	{
		t[i] = orientation.c[3][i] - F(0.5f) * (bmin[i] + bmax[i]);
		a[i] = F(0.5f) * (bmax[i] - bmin[i]);
	}
	const F h[3] = { half.x, half.y, half.z };
	bool overlap = true;
	auto test = [&](F dist, F ra, F rb) { overlap = overlap && std::abs(dist) <= ra + rb; };
	for (int i = 0; i < 3; ++i)
		test(t[i], a[i], (h[0] * abs_r[0][i] + h[1] * abs_r[1][i]) + h[2] * abs_r[2][i]);
	for (int i = 0; i < 3; ++i)
		test((t[0] * R(0, i) + t[1] * R(1, i)) + t[2] * R(2, i), (a[0] * abs_r[i][0] + a[1] * abs_r[i][1]) + a[2] * abs_r[i][2], h[i]);
	// A_i x B_j for i, j = 0..2: with (i1, i2) and (j1, j2) the other two axes of each box
	const int other[3][2] = { { 1, 2 }, { 0, 2 }, { 0, 1 } };
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
		{
			const int i1 = other[i][0], i2 = other[i][1], j1 = other[j][0], j2 = other[j][1];
			const F ra = a[i1] * abs_r[j][i2] + a[i2] * abs_r[j][i1];
			const F rb = h[j1] * abs_r[j2][i] + h[j2] * abs_r[j1][i];
			// the distance along A_i x B_j: t[i2] R(i1, j) - t[i1] R(i2, j), signs as the
			// reference orders them (t2 R1j - t1 R2j, t0 R2j - t2 R0j, t1 R0j - t0 R1j)
			const F dist = i == 1 ? t[0] * R(2, j) - t[2] * R(0, j) : t[i2] * R(i1, j) - t[i1] * R(i2, j);
			test(dist, ra, rb);
		}
	return overlap;
}

// CollectTransformedShapes from a compound or decorated body to one of its leaves
// (CompoundShapeVisitors.h:226-272, StaticCompoundShape.cpp:543-570, MutableCompoundShape.cpp:443-470,
// RotatedTranslatedShape.cpp:169-176, OffsetCenterOfMassShape.cpp:135-142): starting from the
// body's rotation q and position p relative to the soft body, apply the part's steps; a bounds
// step builds the compound's view of the soft body's local bounds `box`,
// OrientedBox(Mat44::sInverseRotationTranslation(q, p), box) (Mat44.inl:156-161, OrientedBox.h:27),
// and drops the part when the child's bounds miss it.
template <class F> bool PlacePart(const Collider &part, const Box3<F> &box, F q[4], V3<F> &p)
{
	for (const Collider::Op &op : part.ops)
		switch (op.kind)
		{
		case 1:
			{
				const F conj[4] = { -q[0], -q[1], -q[2], q[3] };
				M44<F> m = RotationTranslation(conj, V3<F>{ 0, 0, 0 });
				const V3<F> mt = Neg(Mul3x3(m, p));
				const V3<F> center = F(0.5f) * (box.mn + box.mx);
				const V3<F> half = F(0.5f) * (box.mx - box.mn);
				const V3<F> origin = mt + Mul3x3(m, center);
				m.c[3][0] = origin.x; m.c[3][1] = origin.y; m.c[3][2] = origin.z;
				const F bmin[3] = { F(op.v[0]), F(op.v[1]), F(op.v[2]) }, bmax[3] = { F(op.v[3]), F(op.v[4]), F(op.v[5]) };
				if (!BoxVsOrientedBox(m, half, bmin, bmax))
					return false;
			}
			break;
		case 2:
			p = p + QuatRotate(q, V3<F>{ F(op.v[0]), F(op.v[1]), F(op.v[2]) });
			break;
		case 3:
			{
				const F r[4] = { F(op.v[0]), F(op.v[1]), F(op.v[2]), F(op.v[3]) };
				F out[4];
				QuatMul(q, r, out);
				for (int k = 0; k < 4; ++k)
					q[k] = out[k];
			}
			break;
		default:
			p = p - QuatRotate(q, V3<F>{ F(op.v[0]), F(op.v[1]), F(op.v[2]) });
			break;
		}
	return true;
}

// One collision step of one body: InitializeUpdateContext + DetermineCollidingShapes
// + ParallelUpdate state machine + JobSoftBodyFinalize position update
// (SoftBodyMotionProperties.cpp:937-960, 120-255, 969-1115; PhysicsSystem.cpp:2842-2864).
template <class F> void StepBody(World<F> &w, Body<F> &b, F step_dt)
{
	M44<F> com = RotationTranslation(b.rotation, b.position);
	// InitializeUpdateContext: gravity to local space, sub step, gravity displacement
	V3<F> g = F(b.gravity_factor) * w.gravity;
	b.gravity = Mul3x3(Transposed3x3(com), g);
	b.dt_sub = step_dt / F(b.iterations);
	F ni = F(b.iterations);
	b.disp_gravity = (((F(0.5f) * ni) * F(b.iterations + 1)) * (b.dt_sub * b.dt_sub)) * b.gravity;

	// DetermineCollidingShapes (static bodies only)
	Box3<F> lb = b.local_bounds;
	lb.mn = VMin(lb.mn, b.local_predicted_bounds.mn);
	lb.mx = VMax(lb.mx, b.local_predicted_bounds.mx);
	V3<F> rv{ F(b.vertex_radius), F(b.vertex_radius), F(b.vertex_radius) };
	lb.mn = lb.mn - rv;
	lb.mx = lb.mx + rv;
	Box3<F> wb = Transformed(lb, com);
	M44<F> inv = InvRT(com);
	b.leaves.clear();
	for (const Collider &c : w.colliders)
	{
		Box3<F> cb{ { F(c.bounds[0]), F(c.bounds[1]), F(c.bounds[2]) }, { F(c.bounds[3]), F(c.bounds[4]), F(c.bounds[5]) } };
		if (c.is_sensor || !Overlaps(wb, cb))
			continue;
		F q[4] = { F(c.rot[0]), F(c.rot[1]), F(c.rot[2]), F(c.rot[3]) };
		M44<F> ccom = RotationTranslation(q, V3<F>{ F(c.pos[0]), F(c.pos[1]), F(c.pos[2]) });
		M44<F> rel = Mul(inv, ccom);
		F lq[4];
		GetQuaternion(rel, lq);
		Leaf<F> leaf;
		if (c.parts.empty())
			leaf.shapes.push_back({ RotationTranslation(lq, Col3(rel, 3)), &c });
		else
			for (const Collider &part : c.parts)
			{
				F pq[4] = { lq[0], lq[1], lq[2], lq[3] };
				V3<F> pp = Col3(rel, 3);
				if (PlacePart(part, lb, pq, pp))
					leaf.shapes.push_back({ RotationTranslation(pq, pp), &part });
			}
		if (leaf.shapes.empty())
			continue;		// no leaf collected: the body is left out (SoftBodyMotionProperties.cpp:189-190)
		leaf.col_index = size_t(&c - w.colliders.data());
		leaf.friction = std::sqrt(F(b.friction) * F(c.friction));		// ContactConstraintManager.h:554
		leaf.restitution = std::max(F(b.restitution), F(c.restitution)); // ContactConstraintManager.h:555
		if (c.motion == 2)
		{
			// SoftBodyMotionProperties.cpp:224-233 with MotionProperties::GetInverseInertiaForRotation
			// (MotionProperties.inl:66-82): rotation = rel.Multiply3x3(sRotation(mInertiaRotation)),
			// inverse inertia = rotation.Multiply3x3RightTransposed(diag-scaled columns of rotation)
			leaf.dynamic = true;
			leaf.inv_mass = F(c.inv_mass);
			F iq[4] = { F(c.inertia_rot[0]), F(c.inertia_rot[1]), F(c.inertia_rot[2]), F(c.inertia_rot[3]) };
			M44<F> ir = RotationTranslation(iq, V3<F>{ 0, 0, 0 });
			V3<F> rc[3], sc[3];
			for (int k = 0; k < 3; ++k)
				rc[k] = (Col3(rel, 0) * ir.c[k][0] + Col3(rel, 1) * ir.c[k][1]) + Col3(rel, 2) * ir.c[k][2];
			for (int k = 0; k < 3; ++k)
				sc[k] = F(c.inv_inertia_diag[k]) * rc[k];
			for (int j = 0; j < 3; ++j)
			{
				V3<F> col = (rc[0] * Comp(sc[0], j) + rc[1] * Comp(sc[1], j)) + rc[2] * Comp(sc[2], j);
				leaf.inv_inertia.c[j][0] = col.x; leaf.inv_inertia.c[j][1] = col.y; leaf.inv_inertia.c[j][2] = col.z;
				leaf.inv_inertia.c[j][3] = 0;
			}
			leaf.inv_inertia.c[3][0] = 0; leaf.inv_inertia.c[3][1] = 0; leaf.inv_inertia.c[3][2] = 0; leaf.inv_inertia.c[3][3] = 1;
			leaf.com_translation = Col3(rel, 3);
			leaf.lin = leaf.orig_lin = Mul3x3(inv, V3<F>{ F(c.lin_vel[0]), F(c.lin_vel[1]), F(c.lin_vel[2]) });
			leaf.ang = leaf.orig_ang = Mul3x3(inv, V3<F>{ F(c.ang_vel[0]), F(c.ang_vel[1]), F(c.ang_vel[2]) });
		}
		b.leaves.push_back(leaf);
	}

	// DetermineCollisionPlanes (SoftBodyMotionProperties.cpp:257-265)
	for (size_t s = 0; s < b.leaves.size(); ++s)
		for (const auto &shape : b.leaves[s].shapes)
			CollideVertices(b, *shape.second, shape.first, int(s));

	// Iterations: StartFirstIteration, then ParallelApplyConstraints loop
	ApplyPressure(b);
	Integrate(b);
	for (u32 it = 0; it < b.iterations; ++it)
	{
		for (size_t g = 0; g + 1 < b.groups.size(); ++g)
			ProcessGroup(b, g);
		if (!b.groups.empty())
			ProcessGroup(b, b.groups.size() - 1);
		CollisionAndVelocity(b);
		if (it + 1 < b.iterations)
		{
			ApplyPressure(b);
			Integrate(b);
		}
	}
	V3<F> delta_world = UpdateState(b, step_dt, com);
	// UpdateRigidBodyVelocities (SoftBodyMotionProperties.cpp:923-935): world-space changes
	for (const Leaf<F> &cs : b.leaves)
		if (cs.dynamic && cs.update_velocities)
		{
			b.dyn_lin[cs.col_index] = b.dyn_lin[cs.col_index] + Mul3x3(com, cs.lin - cs.orig_lin);
			b.dyn_ang[cs.col_index] = b.dyn_ang[cs.col_index] + Mul3x3(com, cs.ang - cs.orig_ang);
			b.dyn_updated[cs.col_index] = 1;
		}
	// JobSoftBodyFinalize: SetPositionAndRotationInternal(GetPosition() + delta)
	b.position = b.position + delta_world;
}

template <class F> struct Api
{
	static void *Create() { return new World<F>; }
};

} // namespace

// ====================================================================== C API
// precision: 32 -> float restatement, 64 -> double restatement.
struct XrWorld
{
	int precision;
	World<float> f;
	World<double> d;
};

template <class F> static World<F> &Get(XrWorld *w);
template <> World<float> &Get<float>(XrWorld *w) { return w->f; }
template <> World<double> &Get<double>(XrWorld *w) { return w->d; }

template <class F> static int AddBody(XrWorld *xw, const VertexRec *verts, u32 nv, const FaceRec *faces, u32 nf, const EdgeRec *edges, u32 ne,
									  const BendRec *bends, u32 nb, const VolumeRec *vols, u32 nt, const LRARec *lras, u32 nl,
									  const GroupRec *groups, u32 ng, const float *pos, const float *rot, const float *p)
{
	World<F> &w = Get<F>(xw);
	Body<F> b;
	b.faces.assign(faces, faces + nf);
	b.edges.assign(edges, edges + ne);
	b.bends.assign(bends, bends + nb);
	b.volumes.assign(vols, vols + nt);
	b.lras.assign(lras, lras + nl);
	b.groups.assign(groups, groups + ng);
	b.iterations = u32(p[0]);
	b.linear_damping = p[1]; b.max_linear_velocity = p[2]; b.restitution = p[3]; b.friction = p[4];
	b.pressure = p[5]; b.gravity_factor = p[6]; b.vertex_radius = p[7];
	b.update_position = p[8] != 0.0f;
	bool make_identity = p[9] != 0.0f;
	b.allow_sleeping = p[10] != 0.0f;
	b.position = { F(pos[0]), F(pos[1]), F(pos[2]) };
	F q[4] = { F(rot[0]), F(rot[1]), F(rot[2]), F(rot[3]) };
	// SoftBodyMotionProperties::Initialize (SoftBodyMotionProperties.cpp:59-105)
	M44<F> R;
	F ident[4] = { 0, 0, 0, 1 };
	R = RotationTranslation(make_identity ? q : ident, V3<F>{ 0, 0, 0 });
	for (int k = 0; k < 4; ++k)
		b.rotation[k] = make_identity ? ident[k] : q[k];
	b.x.resize(nv); b.prev.resize(nv); b.v.resize(nv); b.w.resize(nv);
	b.plane.assign(nv, PlaneT<F>{ { 0, 0, 0 }, 0 });
	b.shape_idx.assign(nv, -1); b.largest.assign(nv, F(-FLT_MAX)); b.contact.assign(nv, 0);
	b.local_bounds = EmptyBox<F>();
	for (u32 i = 0; i < nv; ++i)
	{
		V3<F> xp{ F(verts[i].pos[0]), F(verts[i].pos[1]), F(verts[i].pos[2]) };
		V3<F> vv{ F(verts[i].vel[0]), F(verts[i].vel[1]), F(verts[i].vel[2]) };
		b.x[i] = b.prev[i] = MulPoint(R, xp);
		b.v[i] = Mul3x3(R, vv);
		b.w[i] = F(verts[i].inv_mass);
		Encapsulate(b.local_bounds, b.x[i]);
	}
	b.local_predicted_bounds = b.local_bounds;
	w.bodies.push_back(std::move(b));
	return int(w.bodies.size() - 1);
}

template <class F> static void UpdateWorld(XrWorld *xw, float dt, int steps)
{
	World<F> &w = Get<F>(xw);
	// PhysicsSystem::Update: step delta time = dt / collision steps (PhysicsSystem.cpp:215)
	F step_dt = F(dt) / F(steps);
	for (Body<F> &b : w.bodies)
	{
		b.dyn_lin.assign(w.colliders.size(), V3<F>{ 0, 0, 0 });
		b.dyn_ang.assign(w.colliders.size(), V3<F>{ 0, 0, 0 });
		b.dyn_updated.assign(w.colliders.size(), 0);
	}
	for (int s = 0; s < steps; ++s)
		for (Body<F> &b : w.bodies)
			if (b.active)
				StepBody(w, b, step_dt);
}

template <class F> static int ReadBody(XrWorld *xw, int bi, double *pos, double *prev, double *vel, double *plane, int *idx, double *pen, int *contact, double *bpos)
{
	World<F> &w = Get<F>(xw);
	if (bi < 0 || bi >= int(w.bodies.size()))
		return -1;
	const Body<F> &b = w.bodies[bi];
	for (size_t i = 0; i < b.x.size(); ++i)
	{
		if (pos) { pos[3 * i] = b.x[i].x; pos[3 * i + 1] = b.x[i].y; pos[3 * i + 2] = b.x[i].z; }
		if (prev) { prev[3 * i] = b.prev[i].x; prev[3 * i + 1] = b.prev[i].y; prev[3 * i + 2] = b.prev[i].z; }
		if (vel) { vel[3 * i] = b.v[i].x; vel[3 * i + 1] = b.v[i].y; vel[3 * i + 2] = b.v[i].z; }
		if (plane) { plane[4 * i] = b.plane[i].n.x; plane[4 * i + 1] = b.plane[i].n.y; plane[4 * i + 2] = b.plane[i].n.z; plane[4 * i + 3] = b.plane[i].c; }
		if (idx) idx[i] = b.shape_idx[i];
		if (pen) pen[i] = b.largest[i];
		if (contact) contact[i] = b.contact[i];
	}
	if (bpos) { bpos[0] = b.position.x; bpos[1] = b.position.y; bpos[2] = b.position.z; }
	return int(b.x.size());
}

template <class F> static int WriteBody(XrWorld *xw, int bi, const double *pos, const double *vel)
{
	World<F> &w = Get<F>(xw);
	if (bi < 0 || bi >= int(w.bodies.size()))
		return -1;
	Body<F> &b = w.bodies[bi];
	for (size_t i = 0; i < b.x.size(); ++i)
	{
		if (pos) b.x[i] = { F(pos[3 * i]), F(pos[3 * i + 1]), F(pos[3 * i + 2]) };
		if (vel) b.v[i] = { F(vel[3 * i]), F(vel[3 * i + 1]), F(vel[3 * i + 2]) };
	}
	return 0;
}

// World-space velocity changes body `body` applied to each of the n colliders in the last
// update call (3 doubles each), and whether it touched it.
template <class F> static int VelocityDeltas(XrWorld *xw, int bi, int n, double *lin, double *ang, int *updated)
{
	World<F> &w = Get<F>(xw);
	if (bi < 0 || bi >= int(w.bodies.size()) || n != int(w.colliders.size()))
		return -1;
	const Body<F> &b = w.bodies[bi];
	for (int c = 0; c < n; ++c)
	{
		const bool have = size_t(c) < b.dyn_updated.size();
		V3<F> l = have ? b.dyn_lin[c] : V3<F>{ 0, 0, 0 }, a = have ? b.dyn_ang[c] : V3<F>{ 0, 0, 0 };
		lin[3 * c] = l.x; lin[3 * c + 1] = l.y; lin[3 * c + 2] = l.z;
		ang[3 * c] = a.x; ang[3 * c + 1] = a.y; ang[3 * c + 2] = a.z;
		updated[c] = have ? b.dyn_updated[c] : 0;
	}
	return 0;
}

extern "C" {

void *xr_world_create(int precision)
{
	if (precision != 32 && precision != 64)
		return nullptr;
	XrWorld *w = new XrWorld;
	w->precision = precision;
	return w;
}

void xr_world_free(void *w) { delete (XrWorld *)w; }

void xr_world_set_gravity(void *wp, float gx, float gy, float gz)
{
	XrWorld *w = (XrWorld *)wp;
	w->f.gravity = { gx, gy, gz };
	w->d.gravity = { gx, gy, gz };
}

// type 0 plane {nx,ny,nz,c,halfExtent}, 1 sphere {r}, 2 box {hx,hy,hz}, 3 capsule {halfHeight, r},
// 4 cylinder {halfHeight, r}, 5 tapered capsule {topCenter, bottomCenter, topRadius, bottomRadius,
// sinAlpha, tanAlpha} (as the shape holds them); pos: the centre of mass;
// bounds: the static body's world-space bounds (min xyz, max xyz).
int xr_world_add_collider(void *wp, int type, const float *params, const float *pos, const float *rot, float friction, float restitution, const float *bounds)
{
	XrWorld *w = (XrWorld *)wp;
	if (type < 0 || type > 8)
		return -1;
	Collider c{};
	c.type = type;
	int np = type == 0 ? 5 : type == 1 ? 1 : type == 2 ? 3 : type == 5 ? 6 : type == 6 || type == 8 ? 0 : type == 7 ? 4 : 2;
	for (int i = 0; i < np; ++i) c.params[i] = params[i];
	for (int i = 0; i < 3; ++i) c.pos[i] = pos[i];
	for (int i = 0; i < 4; ++i) c.rot[i] = rot[i];
	for (int i = 0; i < 6; ++i) c.bounds[i] = bounds[i];
	c.friction = friction;
	c.restitution = restitution;
	w->f.colliders.push_back(c);
	w->d.colliders.push_back(c);
	return int(w->f.colliders.size() - 1);
}

// Scale of collider `collider` (a ScaledShape around its shape)
int xr_world_set_collider_scale(void *wp, int collider, const float *scale3)
{
	XrWorld *w = (XrWorld *)wp;
	if (collider < 0 || size_t(collider) >= w->f.colliders.size())
		return -1;
	for (int k = 0; k < 3; ++k)
	{
		w->f.colliders[collider].scale[k] = scale3[k];
		w->d.colliders[collider].scale[k] = scale3[k];
	}
	return 0;
}

// The collider itself (part < 0) or one of its parts, in the float and the double world
static bool Targets(XrWorld *w, int collider, int part, Collider *out[2])
{
	if (collider < 0 || size_t(collider) >= w->f.colliders.size())
		return false;
	Collider &f = w->f.colliders[collider], &d = w->d.colliders[collider];
	if (part < 0)
	{
		out[0] = &f; out[1] = &d;
		return true;
	}
	if (size_t(part) >= f.parts.size())
		return false;
	out[0] = &f.parts[part]; out[1] = &d.parts[part];
	return true;
}

// Triangle tree of a mesh shape (type 8), of collider `collider` (part < 0) or of one of its parts:
// the root reference, nn nodes of 24 bound floats (min x, y, z, max x, y, z of the four children)
// and 4 child references each, nt triangles of 9 floats; flags bit 0: a height field's hierarchy.
int xr_world_set_part_tree(void *wp, int collider, int part, u32 root, const float *bounds, const u32 *children, u32 nn,
						   const float *tris, u32 nt, u32 flags)
{
	Collider *targets[2];
	if (!Targets((XrWorld *)wp, collider, part, targets) || targets[0]->type != 8)
		return -1;
	for (Collider *c : targets)
	{
		c->tree_root = root;
		c->tree_bounds.assign(bounds, bounds + 24 * size_t(nn));
		c->tree_child.assign(children, children + 4 * size_t(nn));
		c->tree_tris.assign(tris, tris + 9 * size_t(nt));
		c->tree_height_field = (flags & 1u) != 0;
	}
	return 0;
}

int xr_world_set_tree(void *wp, int collider, u32 root, const float *bounds, const u32 *children, u32 nn, const float *tris, u32 nt,
					  u32 flags)
{
	return xr_world_set_part_tree(wp, collider, -1, root, bounds, children, nn, tris, nt, flags);
}

// Hull data of a convex hull shape (type 6), of collider `collider` (part < 0) or of one of its
// parts: np points relative to the centre of mass (xyz), nf planes (normal xyz, constant), per
// face first entry and vertex count into the ni concatenated face vertex indices.
int xr_world_set_part_hull(void *wp, int collider, int part, const float *points, u32 np, const float *planes, u32 nf,
						   const u32 *face_first, const u32 *face_count, const u32 *index, u32 ni)
{
	Collider *targets[2];
	if (!Targets((XrWorld *)wp, collider, part, targets) || targets[0]->type != 6)
		return -1;
	for (u32 f = 0; f < nf; ++f)
		if (face_count[f] < 2 || size_t(face_first[f]) + face_count[f] > ni)
			return -1;
	for (u32 k = 0; k < ni; ++k)
		if (index[k] >= np)
			return -1;
	for (Collider *c : targets)
	{
		c->hull_points.assign(points, points + 3 * np);
		c->hull_planes.assign(planes, planes + 4 * nf);
		c->face_first.assign(face_first, face_first + nf);
		c->face_count.assign(face_count, face_count + nf);
		c->index.assign(index, index + ni);
	}
	return 0;
}

int xr_world_set_hull(void *wp, int collider, const float *points, u32 np, const float *planes, u32 nf,
					  const u32 *face_first, const u32 *face_count, const u32 *index, u32 ni)
{
	return xr_world_set_part_hull(wp, collider, -1, points, np, planes, nf, face_first, face_count, index, ni);
}

// The leaf transforms (16 floats each, column major) that a soft body at (soft_pos3, soft_rot4)
// with local bounds box6 (min xyz, max xyz) collects from collider `collider`, in the float
// world; the number of leaves (at most max_leaves written), or -1.
int xr_collect_leaves(void *wp, int collider, const float *soft_pos3, const float *soft_rot4, const float *box6, float *out,
					  int max_leaves)
{
	XrWorld *w = (XrWorld *)wp;
	if (collider < 0 || size_t(collider) >= w->f.colliders.size())
		return -1;
	const Collider &c = w->f.colliders[collider];
	const float sq[4] = { soft_rot4[0], soft_rot4[1], soft_rot4[2], soft_rot4[3] };
	M44<float> com = RotationTranslation(sq, V3<float>{ soft_pos3[0], soft_pos3[1], soft_pos3[2] });
	const float cq[4] = { c.rot[0], c.rot[1], c.rot[2], c.rot[3] };
	M44<float> rel = Mul(InvRT(com), RotationTranslation(cq, V3<float>{ c.pos[0], c.pos[1], c.pos[2] }));
	float lq[4];
	GetQuaternion(rel, lq);
	const Box3<float> box{ { box6[0], box6[1], box6[2] }, { box6[3], box6[4], box6[5] } };
	std::vector<M44<float>> leaves;
	if (c.parts.empty())
		leaves.push_back(RotationTranslation(lq, Col3(rel, 3)));
	for (const Collider &part : c.parts)
	{
		float pq[4] = { lq[0], lq[1], lq[2], lq[3] };
		V3<float> pp = Col3(rel, 3);
		if (PlacePart(part, box, pq, pp))
			leaves.push_back(RotationTranslation(pq, pp));
	}
	for (size_t n = 0; n < leaves.size() && int(n) < max_leaves; ++n)
		for (int col = 0; col < 4; ++col)
			for (int r = 0; r < 4; ++r)
				out[16 * n + 4 * col + r] = leaves[n].c[col][r];
	return int(leaves.size());
}

// A leaf shape of compound collider `collider`, appended in collection order: type and params as
// xr_world_add_collider takes them, scale[3], and num_ops placement steps of 7 words each (kind as
// a 32-bit integer, then 6 floats).  Returns the part's index.
int xr_world_add_part(void *wp, int collider, int type, const float *params, const float *scale3, u32 num_ops, const void *ops)
{
	XrWorld *w = (XrWorld *)wp;
	if (collider < 0 || size_t(collider) >= w->f.colliders.size() || type < 0 || type > 8)
		return -1;
	Collider part{};
	part.type = type;
	for (int i = 0; i < 8; ++i) part.params[i] = params[i];
	for (int i = 0; i < 3; ++i) part.scale[i] = scale3[i];
	const unsigned char *bytes = (const unsigned char *)ops;
	for (u32 k = 0; k < num_ops; ++k)
	{
		Collider::Op op;
		memcpy(&op.kind, bytes + 28 * size_t(k), 4);
		memcpy(op.v, bytes + 28 * size_t(k) + 4, 24);
		if (op.kind < 1 || op.kind > 4)
			return -1;
		part.ops.push_back(op);
	}
	w->f.colliders[collider].parts.push_back(part);
	w->d.colliders[collider].parts.push_back(part);
	return int(w->f.colliders[collider].parts.size() - 1);
}

// p: {numIterations, linearDamping, maxLinearVelocity, restitution, friction,
// pressure, gravityFactor, vertexRadius, updatePosition, makeRotationIdentity, allowSleeping}
int xr_world_add_soft_body(void *wp, const void *verts, u32 nv, const void *faces4, u32 nf, const void *edges, u32 ne,
						   const void *bends, u32 nb, const void *vols, u32 nt, const void *lras, u32 nl,
						   const void *groups7, u32 ng, const float *pos, const float *rot, const float *p)
{
	XrWorld *w = (XrWorld *)wp;
	int a = AddBody<float>(w, (const VertexRec *)verts, nv, (const FaceRec *)faces4, nf, (const EdgeRec *)edges, ne, (const BendRec *)bends, nb,
						   (const VolumeRec *)vols, nt, (const LRARec *)lras, nl, (const GroupRec *)groups7, ng, pos, rot, p);
	int b = AddBody<double>(w, (const VertexRec *)verts, nv, (const FaceRec *)faces4, nf, (const EdgeRec *)edges, ne, (const BendRec *)bends, nb,
							(const VolumeRec *)vols, nt, (const LRARec *)lras, nl, (const GroupRec *)groups7, ng, pos, rot, p);
	return a == b ? a : -1;
}

double xr_world_update(void *wp, float dt, int collision_steps, int frames)
{
	XrWorld *w = (XrWorld *)wp;
	auto t0 = std::chrono::high_resolution_clock::now();
	for (int i = 0; i < frames; ++i)
	{
		if (w->precision == 32)
			UpdateWorld<float>(w, dt, collision_steps);
		else
			UpdateWorld<double>(w, dt, collision_steps);
	}
	return std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
}

int xr_body_read(void *wp, int body, double *pos, double *prev, double *vel, double *plane, int *idx, double *pen, int *contact, double *bpos)
{
	XrWorld *w = (XrWorld *)wp;
	return w->precision == 32 ? ReadBody<float>(w, body, pos, prev, vel, plane, idx, pen, contact, bpos)
							  : ReadBody<double>(w, body, pos, prev, vel, plane, idx, pen, contact, bpos);
}

int xr_body_write(void *wp, int body, const double *pos, const double *vel)
{
	XrWorld *w = (XrWorld *)wp;
	return w->precision == 32 ? WriteBody<float>(w, body, pos, vel) : WriteBody<double>(w, body, pos, vel);
}

// Marks collider `collider` as a sensor: it takes no part in the collision.
int xr_world_set_sensor(void *wp, int collider)
{
	XrWorld *w = (XrWorld *)wp;
	if (collider < 0 || size_t(collider) >= w->f.colliders.size())
		return -1;
	w->f.colliders[collider].is_sensor = 1;
	w->d.colliders[collider].is_sensor = 1;
	return 0;
}

// Motion of collider `collider`: EMotionType (0 static, 1 kinematic, 2 dynamic) and, for a
// dynamic body, inverse mass, inverse inertia diagonal, inertia rotation, world velocities.
int xr_world_set_motion(void *wp, int collider, int motion, float inv_mass, const float *inv_inertia_diag,
						const float *inertia_rot, const float *lin_vel, const float *ang_vel)
{
	XrWorld *w = (XrWorld *)wp;
	if (collider < 0 || size_t(collider) >= w->f.colliders.size() || motion < 0 || motion > 2)
		return -1;
	Collider *targets[2] = { &w->f.colliders[collider], &w->d.colliders[collider] };
	for (Collider *c : targets)
	{
		c->motion = motion;
		c->inv_mass = inv_mass;
		for (int k = 0; k < 3; ++k)
		{
			c->inv_inertia_diag[k] = inv_inertia_diag[k];
			c->lin_vel[k] = lin_vel[k];
			c->ang_vel[k] = ang_vel[k];
		}
		for (int k = 0; k < 4; ++k)
			c->inertia_rot[k] = inertia_rot[k];
	}
	return 0;
}

int xr_body_velocity_deltas(void *wp, int body, int n, double *lin, double *ang, int *updated)
{
	XrWorld *w = (XrWorld *)wp;
	return w->precision == 32 ? VelocityDeltas<float>(w, body, n, lin, ang, updated) : VelocityDeltas<double>(w, body, n, lin, ang, updated);
}

// ACosApproximate restated (for the closed-form pins)
float xr_acos_approximate(float x) { return ACosApprox<float>(x); }

} // extern "C"
