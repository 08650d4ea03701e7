/*
 * jolt_xpbd.h -- C ABI of the MI355X (gfx950) soft-body XPBD update.
 *
 * A drop-in for the soft-body hot path of Jolt Physics v5.5.1: the work that
 * SoftBodyMotionProperties::ParallelUpdate performs for every soft body inside
 * PhysicsSystem::Update (Jolt/Physics/PhysicsSystem.cpp:2701-2897) and inside
 * the public single-body SoftBodyMotionProperties::CustomUpdate
 * (Jolt/Physics/SoftBody/SoftBodyMotionProperties.cpp:1232-1256).
 *
 * All arrays crossing this boundary use the reference's own in-memory record
 * layouts (Jolt/Physics/SoftBody/SoftBodySharedSettings.h), so an integration
 * passes SoftBodySharedSettings::mEdgeConstraints.data() etc. unchanged:
 *   Vertex 28 B {float pos[3], vel[3], invMass}          (.h:127-138)
 *   Face 16 B {uint32 v[3], materialIndex}               (.h:141-154)
 *   Edge 16 B {uint32 v[2]; float restLength, compliance} (.h:157-171)
 *   DihedralBend 24 B {uint32 v[4]; float compliance, initialAngle} (.h:191-205)
 *   Volume 24 B {uint32 v[4]; float sixRestVolume, compliance}      (.h:208-222)
 *   LRA 12 B {uint32 v[2]; float maxDistance}            (.h:289-303)
 *   UpdateGroup 28 B {uint32 edgeEnd, lraEnd, rodStretchShearEnd,
 *                     rodBendTwistEnd, dihedralBendEnd, volumeEnd, skinnedEnd} (.h:374-383)
 *
 * Conventions: every function returns XPBD_OK (0) or a negative XPBD_E*
 * code and never aborts; xpbd_last_error() describes the last failure of the
 * calling thread.  The context owns all device memory and one HIP stream;
 * caller buffers are only read/written during the call.  One host thread per
 * context.
 */
#ifndef JOLT_XPBD_H
#define JOLT_XPBD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XPBD_OK 0
#define XPBD_E_INVALID -1       /* bad argument / handle / record */
#define XPBD_E_UNSUPPORTED -2   /* feature outside this build (skinned, rods, ...) */
#define XPBD_E_DEVICE -3        /* HIP runtime error */
#define XPBD_E_NOMEM -4         /* device or host allocation failed */
#define XPBD_E_CAPACITY -5      /* a compile-time capacity limit was exceeded */

/* Collider shape types */
#define XPBD_SHAPE_PLANE 0      /* params {nx, ny, nz, c, halfExtent}  PlaneShape */
#define XPBD_SHAPE_SPHERE 1     /* params {radius}                     SphereShape */
#define XPBD_SHAPE_BOX 2        /* params {hx, hy, hz}                 BoxShape */
#define XPBD_SHAPE_CAPSULE 3    /* params {halfHeightOfCylinder, radius} CapsuleShape */
#define XPBD_SHAPE_CYLINDER 4   /* params {halfHeight, radius}         CylinderShape */
/* TaperedCapsuleShape: params {topCenter, bottomCenter, topRadius, bottomRadius, sinAlpha,
 * tanAlpha} as the shape holds them (TaperedCapsuleShape.cpp:86-131: the centres are relative
 * to the shape's centre of mass; tanAlpha = Tan(ASin(sinAlpha)) with Jolt's approximations) */
#define XPBD_SHAPE_TAPERED_CAPSULE 5
/* ConvexHullShape: no params; the hull's own data in xpbd_collider_desc.hull
 * (ConvexHullShape.cpp:1060-1164) */
#define XPBD_SHAPE_CONVEX_HULL 6
/* TaperedCylinderShape: params {top, bottom, topRadius, bottomRadius} as the shape holds them
 * (TaperedCylinderShape.cpp:108-130: top and bottom relative to the shape's centre of mass;
 * SaveBinaryState writes them); TaperedCylinderShape.cpp:390-506 */
#define XPBD_SHAPE_TAPERED_CYLINDER 7
/* TriangleShape (TriangleShape.cpp:272-283), MeshShape (MeshShape.cpp:833-885) and
 * HeightFieldShape (HeightFieldShape.cpp:2242-2293): no params;
 * xpbd_collider_desc.tree holds the triangles as the shape's tree walk visits them
 * (CollideSoftBodyVerticesVsTriangles.h).  A triangle is a tree whose root is a leaf holding
 * its one triangle. */
#define XPBD_SHAPE_MESH 8

/* Child references of an xpbd_tree_node (and xpbd_triangle_tree.root) */
#define XPBD_TREE_NODE 0x00000000u      /* | node index (bits 0-29) */
#define XPBD_TREE_LEAF 0x40000000u      /* | triangle count << 24 (1-63) | first triangle (bits 0-23) */
#define XPBD_TREE_EMPTY 0x80000000u     /* a padding child: never visited */

/* One node of a four-way bounding-volume tree: the bounds of its four children, per axis,
 * decoded to float exactly as the reference decodes them (NodeCodecQuadTreeHalfFloat.h:245-
 * 283: half floats widened to float), and the children's references in stored order. */
typedef struct xpbd_tree_node {
    float min_x[4], min_y[4], min_z[4];
    float max_x[4], max_y[4], max_z[4];
    uint32_t child[4];
} xpbd_tree_node;

/* A MeshShape's tree as NodeCodecQuadTreeHalfFloat::DecodingContext::WalkTree walks it: the
 * root reference, the nodes, and the triangles of every leaf in the order the leaf's block
 * decodes them (TriangleCodecIndexed8BitPackSOA4Flags, vertices in the shape's space, 9
 * floats per triangle).  Per vertex the device walks it closest child first, with the
 * reference's sort of the child distances (Vec4::sSort4Reverse) and its pruning, so the
 * triangle it keeps is the one the reference keeps, ties included. */
typedef struct xpbd_triangle_tree {
    uint32_t root;
    uint32_t num_nodes;
    const xpbd_tree_node *nodes;
    uint32_t num_triangles;
    const float *triangles;         /* num_triangles x 9: v0, v1, v2 */
    uint32_t flags;                 /* XPBD_TREE_HEIGHT_FIELD or 0 */
} xpbd_triangle_tree;

/* xpbd_triangle_tree.flags: a HeightFieldShape's range-block hierarchy (HeightFieldShape.cpp:
 * 1745-2093, 2242-2293): a child whose unscaled bounds have min y > max y (a block without
 * collision) gets distance FLT_MAX and is never visited. */
#define XPBD_TREE_HEIGHT_FIELD 1u

/* Largest depth of a triangle tree (levels of nodes below the root): the per-vertex walk keeps
 * its stack in 3 x depth + 4 entries. */
#define XPBD_TREE_MAX_DEPTH 20

/* A convex hull as ConvexHullShape holds it: points relative to its centre of mass
 * (GetPoint), one plane per face (GetPlanes: normal xyz, constant) and each face's
 * vertex indices in counter-clockwise order (GetFaceVertices), concatenated. */
typedef struct xpbd_convex_hull {
    uint32_t num_points;
    const float *points;            /* num_points x 3 */
    uint32_t num_faces;
    const float *planes;            /* num_faces x 4 */
    const uint32_t *face_first;     /* num_faces: first entry of the face in vertex_index */
    const uint32_t *face_count;     /* num_faces: vertices of the face */
    uint32_t num_vertex_index;
    const uint32_t *vertex_index;   /* indices into points */
} xpbd_convex_hull;

typedef struct xpbd_context xpbd_context;
typedef struct xpbd_settings xpbd_settings;

/* Last error message of the calling thread (never NULL). */
const char *xpbd_last_error(void);

/* Library / device management.  device: HIP ordinal. */
int xpbd_context_create(int device, xpbd_context **out_ctx);
int xpbd_context_destroy(xpbd_context *ctx);
/* The context's HIP stream (hipStream_t), for callers that time or order work. */
void *xpbd_context_stream(xpbd_context *ctx);

/*
 * Shared settings: replaces SoftBodySharedSettings after Optimize()
 * (SoftBodySharedSettings.cpp:613-1182).  Builds the order-preserving level
 * schedule of the update groups and uploads everything.  Settings are
 * immutable and shared by any number of bodies (like the reference's
 * RefConst<SoftBodySharedSettings>, SoftBodyCreationSettings.h:49).
 */
typedef struct xpbd_settings_desc {
    const void *vertices;       uint32_t num_vertices;
    const void *faces;          uint32_t num_faces;
    const void *edges;          uint32_t num_edges;
    const void *dihedral_bends; uint32_t num_dihedral_bends;
    const void *volumes;        uint32_t num_volumes;
    const void *lras;           uint32_t num_lras;
    const void *update_groups;  uint32_t num_update_groups;
    /* skinned constraints: Skinned records (SoftBodySharedSettings.h:252-285, 52 B:
     * vertex, 4 x {invBindIndex, weight}, maxDistance, backStopDistance,
     * backStopRadius, normalInfo), at most one per vertex */
    uint32_t num_skinned;
    /* Cosserat rods (SoftBodySharedSettings.h:305-342), arrays in rod_stretch_shear /
     * rod_bend_twist below, after CalculateRodProperties and Optimize */
    uint32_t num_rod_stretch_shear;
    uint32_t num_rod_bend_twist;
    const void *skinned;
    const uint32_t *skinned_normals;    uint32_t num_skinned_normals;   /* mSkinnedConstraintNormals */
    const void *inv_bind_matrices;      uint32_t num_inv_bind_matrices; /* InvBind records (80 B: jointIndex,
                                                                           12 B padding, Mat44 column major) */
    /* RodStretchShear records (48 B: vertex[2], length, invMass, compliance, 12 B padding,
     * bishop quaternion xyzw) and RodBendTwist records (32 B: rod[2], compliance, 4 B
     * padding, omega0 quaternion xyzw), in the reference's memory layout */
    const void *rod_stretch_shear;
    const void *rod_bend_twist;
} xpbd_settings_desc;

int xpbd_settings_create(xpbd_context *ctx, const xpbd_settings_desc *desc, xpbd_settings **out_settings);
/*
 * Throughput schedules.  XPBD_SCHEDULE_EXACT is xpbd_settings_create's: every update group's
 * constraints in the reference's order (SoftBodyMotionProperties.cpp:1030-1055).  The coloured
 * schedules project the constraints of each run of one type in a group (all groups, or only the
 * non-parallel last one) colour by colour: a constraint takes the lowest colour that no earlier
 * constraint of the run holds on a shared vertex either of them writes.  The result equals the
 * reference's on settings whose constraints are sorted (stably) by that colour within each run
 * of each group, a reordering SoftBodySharedSettings::Optimize does not fix either; it differs
 * from the reference on the given order by the rounding of the reordered Gauss-Seidel sweep.
 * Groups with rod constraints keep the exact order.  No reference counterpart: for callers that
 * trade the reference's constraint order for fewer level passes (many more constraints per pass).
 */
#define XPBD_SCHEDULE_EXACT 0
#define XPBD_SCHEDULE_COLOURED_TAIL 1
#define XPBD_SCHEDULE_COLOURED 2
int xpbd_settings_create_scheduled(xpbd_context *ctx, const xpbd_settings_desc *desc, int32_t schedule,
                                   xpbd_settings **out_settings);
int xpbd_settings_release(xpbd_settings *settings);
/* Schedule statistics: out[0]=parallel groups, out[1]=max levels of a parallel
 * group, out[2]=levels of the non-parallel group, out[3]=constraints of the
 * non-parallel group, out[4]=total constraints, out[5]=1 if the body keeps all
 * positions in LDS, out[6]/out[7]=parallel / non-parallel windows. */
int xpbd_settings_stats(const xpbd_settings *settings, uint32_t out[8]);
/* xpbd_schedule_build (below) with one of the XPBD_SCHEDULE_* schedules.  No reference counterpart. */
int xpbd_schedule_build_scheduled(const xpbd_settings_desc *desc, int32_t schedule, xpbd_settings **out);
/*
 * The level programs of settings built on the host alone, as xpbd_settings_create would
 * upload them (no device; for checks of the schedule).  xpbd_schedule_build returns settings
 * usable only with xpbd_schedule_sizes / xpbd_schedule_read / xpbd_settings_release.
 * sizes: out[0] records (4 x u32 each, packed as in xpbd_types.h), out[1] level offsets,
 * out[2] windows (4 x u32: vertex list begin, count (0: identity), first level, levels),
 * out[3] parallel-group windows (the first ones), out[4] vertex list entries, out[5]
 * coloured records (8 x u32: v0-v3, type, p0, p1, 0), out[6] colour offsets, out[7] 1 if
 * resident, out[8] / out[9] records and level offsets of the identity program (body vertex
 * numbers; the program bundles merge), empty for bodies of 16384 vertices or more.
 * rec_index: the constraint index (in its type's array) of every record, ~0 for padding.
 * No reference counterpart.
 */
int xpbd_schedule_build(const xpbd_settings_desc *desc, xpbd_settings **out_settings);
int xpbd_schedule_sizes(const xpbd_settings *settings, uint32_t out[10]);
int xpbd_schedule_read(const xpbd_settings *settings, uint32_t *recs, uint32_t *rec_index, uint32_t *lvl, uint32_t *windows,
                       uint32_t *vlist, uint32_t *colour_recs, uint32_t *colour_begin, uint32_t *identity_recs,
                       uint32_t *identity_rec_index, uint32_t *identity_level_begin);

/* Team statistics of a streamed body's settings (0 for LDS-resident ones): out[0]=parallel-group
 * windows, out[1]=windows of the non-parallel group, out[2]=1 if that group can be coloured
 * (it holds only volumes, bends, edges and LRAs), out[3]=its colours.  No reference
 * counterpart (a property of this build's schedule). */
int xpbd_settings_team_stats(const xpbd_settings *settings, uint32_t out[4]);

/*
 * Host-side settings builder: for callers that have a mesh but not the
 * reference library.  Produces exactly the arrays the reference's
 * SoftBodySharedSettings would hold after the requested steps (same
 * constraints, same order, same update groups):
 *   CreateConstraints     (SoftBodySharedSettings.cpp:189-369)
 *   CalculateEdgeLengths  (:371-379)
 *   CalculateVolumeConstraintVolumes (:536-555)
 *   CalculateLRALengths   (:483-491)
 *   CalculateBendConstraintConstants (:493-534)
 *   Optimize              (:613-1182)
 * Runs on the CPU; no device needed.
 */
#define XPBD_BUILD_CREATE_CONSTRAINTS 1
#define XPBD_BUILD_EDGE_LENGTHS 2
#define XPBD_BUILD_VOLUMES 4
#define XPBD_BUILD_LRA_LENGTHS 8
#define XPBD_BUILD_OPTIMIZE 16
#define XPBD_BUILD_BEND_CONSTANTS 32

/* SoftBodySharedSettings::VertexAttributes (SoftBodySharedSettings.h:39-50);
 * lra_type: 0 none, 1 Euclidean, 2 geodesic (ELRAType). */
typedef struct xpbd_vertex_attributes {
    float compliance;
    float shear_compliance;
    float bend_compliance;
    int32_t lra_type;
    float lra_max_distance_multiplier;
} xpbd_vertex_attributes;

typedef struct xpbd_build_desc {
    const void *vertices;       uint32_t num_vertices;   /* Vertex records */
    const uint32_t *faces;      uint32_t num_faces;      /* 3 indices per face (AddFace order) */
    const void *edges;          uint32_t num_edges;      /* explicit Edge records */
    const void *volumes;        uint32_t num_volumes;    /* explicit Volume records */
    const void *lras;           uint32_t num_lras;       /* explicit LRA records (mLRAConstraints before the steps) */
    const xpbd_vertex_attributes *attributes; uint32_t num_attributes;
    int32_t bend_type;          /* EBendType: 0 none, 1 distance, 2 dihedral */
    float angle_tolerance;      /* radians (reference default DegreesToRadians(8)) */
    float lra_max_distance_multiplier;  /* for XPBD_BUILD_LRA_LENGTHS */
    uint32_t flags;             /* XPBD_BUILD_* steps, applied in the order listed above */
} xpbd_build_desc;

typedef struct xpbd_built_settings xpbd_built_settings;

int xpbd_settings_build(const xpbd_build_desc *desc, xpbd_built_settings **out_built);
/* counts[7] = vertices, faces, edges, dihedral bends, volumes, LRAs, update groups */
int xpbd_built_counts(const xpbd_built_settings *built, uint32_t counts[7]);
/* Copies the arrays out in the reference's record layouts (NULL skips). */
int xpbd_built_read(const xpbd_built_settings *built, void *vertices, void *faces, void *edges, void *dihedral_bends,
                    void *volumes, void *lras, void *update_groups);
int xpbd_built_free(xpbd_built_settings *built);
/* Last error of the settings builder on the calling thread. */
const char *xpbd_build_last_error(void);

/*
 * Soft body: replaces BodyInterface::CreateSoftBody + SoftBodyMotionProperties::
 * Initialize (Body.cpp:466-492, SoftBodyMotionProperties.cpp:59-105) for the
 * fields of SoftBodyCreationSettings that the update uses (.h:61-72).
 */
typedef struct xpbd_body_desc {
    xpbd_settings *settings;
    float position[3];
    float rotation[4];          /* quaternion x, y, z, w */
    uint32_t num_iterations;
    float linear_damping;
    float max_linear_velocity;
    float restitution;
    float friction;
    float pressure;
    float gravity_factor;
    float vertex_radius;
    int32_t update_position;
    int32_t make_rotation_identity;
    int32_t allow_sleeping;
} xpbd_body_desc;

int xpbd_body_create(xpbd_context *ctx, const xpbd_body_desc *desc, int32_t *out_body);

/* Placement steps from a collider body's centre of mass down to one leaf shape of a compound or
 * decorated shape, as CollectTransformedShapes takes them (SoftBodyMotionProperties.cpp:174-188):
 * each collision step the device starts from the body's position p and rotation q relative to the
 * soft body and applies the steps in order; the leaf's transform is then (q, p). */
#define XPBD_OP_CULL 1          /* leave the part out unless the soft body's local bounds, as an oriented box
                                   in the current (p, q) space, overlap the box v[0..5] (min xyz, max xyz): a
                                   compound child's bounds (or a static compound tree node's, one step per
                                   level), times the compound's scale (CompoundShapeVisitors.h:226-256,
                                   AABox4.h:86-185) */
#define XPBD_OP_TRANSLATE 2     /* p = p + q * v[0..2]: a compound child's scaled centre of mass
                                   (CompoundShapeVisitors.h:266) */
#define XPBD_OP_ROTATE 3        /* q = q * v[0..3] (x y z w): a compound child's or a RotatedTranslatedShape's
                                   rotation (CompoundShapeVisitors.h:267, RotatedTranslatedShape.cpp:175) */
#define XPBD_OP_OFFSET 4        /* p = p - q * v[0..2]: an OffsetCenterOfMassShape's scaled offset
                                   (OffsetCenterOfMassShape.cpp:141) */
#define XPBD_MAX_PART_OPS 1024

typedef struct xpbd_part_op {
    uint32_t kind;
    float v[6];
} xpbd_part_op;

/* One leaf shape of a collider body, in the order the reference collects them (StaticCompoundShape
 * walks its tree last child first, MutableCompoundShape its sub shapes in index order;
 * RotatedTranslatedShape, OffsetCenterOfMassShape and ScaledShape pass through to their inner
 * shape).  type, params, scale, hull and tree as in xpbd_collider_desc; ops: the placement steps
 * (copied by xpbd_set_colliders). */
typedef struct xpbd_collider_part {
    int32_t type;
    float params[8];
    float scale[3];
    const xpbd_convex_hull *hull;
    const xpbd_triangle_tree *tree;
    uint32_t num_ops;
    const xpbd_part_op *ops;
} xpbd_collider_part;

/*
 * Static colliders of the world, in broadphase order (the order in which
 * PhysicsSystem::GetBroadPhaseQuery().CollideAABox reports them,
 * SoftBodyMotionProperties.cpp:253).  Each step every soft body keeps the
 * colliders whose world_bounds overlap its predicted bounds and transforms them
 * into its local space exactly as DetermineCollidingShapes does
 * (SoftBodyMotionProperties.cpp:120-255).  Friction / restitution are the
 * collider body's own; they are combined with the soft body's as
 * sqrt(f1*f2) / max(r1,r2) (ContactConstraintManager.h:554-555).
 */
typedef struct xpbd_collider_desc {
    int32_t type;
    float params[8];
    float position[3];          /* centre of mass (Body::GetCenterOfMassPosition()); the body
                                   position for every shape but the tapered capsule */
    float rotation[4];
    float friction;
    float restitution;
    float world_bounds[6];      /* Body::GetWorldSpaceBounds(): min xyz, max xyz */
    const xpbd_convex_hull *hull;   /* XPBD_SHAPE_CONVEX_HULL only (copied by xpbd_set_colliders) */
    /* Motion (Body::GetMotionType(), EMotionType: 0 static, 1 kinematic, 2 dynamic).  Static and
     * kinematic bodies take the static-body contact branch; a dynamic body shares each contact
     * impulse with the vertex (SoftBodyMotionProperties.cpp:752-797), as without a contact
     * listener (every SoftBodyContactSettings scale 1, all angular DOFs allowed).  Dynamic only: */
    int32_t motion_type;
    float inv_mass;                 /* MotionProperties::GetInverseMass() */
    float inv_inertia_diagonal[3];  /* MotionProperties::GetInverseInertiaDiagonal() */
    float inertia_rotation[4];      /* MotionProperties::GetInertiaRotation(), x y z w */
    float linear_velocity[3];       /* world space; held fixed for the whole xpbd_update call */
    float angular_velocity[3];
    /* Body::IsSensor(): the soft body does not collide with it; each step it records whether a
     * vertex penetrates it (DetermineSensorCollisions, SoftBodyMotionProperties.cpp:267-289;
     * xpbd_body_sensor_contacts).  The reference considers sensors only with a contact listener. */
    int32_t is_sensor;
    /* The caller's identifier of the collider body (e.g. BodyID::GetIndexAndSequenceNumber()).
     * A table of xpbd_body_set_contact_settings holds only while every index of the collider
     * list keeps its identifier, shape type, motion type and sensor flag. */
    uint32_t user_id;
    /* The shape's scale, as a ScaledShape around it gives it (ScaledShape.cpp:142-150:
     * CollideSoftBodyVertices of the inner shape with inScale * mScale); {0, 0, 0} stands for
     * {1, 1, 1}.  Sphere, capsule and tapered capsule take a uniform scale, cylinder and
     * tapered cylinder one uniform in x and z (IsValidScale of the shape); `position` is the
     * centre of mass of the scaled shape. */
    float scale[3];
    const xpbd_triangle_tree *tree;     /* XPBD_SHAPE_MESH only (copied by xpbd_set_colliders) */
    /* A compound or decorated body (num_parts > 0): its leaf shapes, copied by xpbd_set_colliders;
     * type, params, scale, hull and tree above are then unused.  The body's leaves share its
     * motion, material and contact settings; a vertex's contact names the body
     * (CollidingShape, SoftBodyMotionProperties.h:200-212).  num_parts 0: the one shape above. */
    uint32_t num_parts;
    const xpbd_collider_part *parts;
} xpbd_collider_desc;

int xpbd_set_colliders(xpbd_context *ctx, const xpbd_collider_desc *colliders, uint32_t num_colliders);
int xpbd_set_gravity(xpbd_context *ctx, const float gravity[3]);

/*
 * Advances every active soft body by `frames` calls of
 * PhysicsSystem::Update(dt, collision_steps) (PhysicsSystem.cpp:176): per
 * collision step InitializeUpdateContext, DetermineCollidingShapes,
 * DetermineCollisionPlanes, num_iterations x (pressure, integrate, update
 * groups, collision + velocity update), UpdateSoftBodyState and the position
 * update of JobSoftBodyFinalize.  Asynchronous on the context stream.
 */
int xpbd_update(xpbd_context *ctx, float dt, int32_t collision_steps, int32_t frames);

/*
 * Large bodies.  A streamed body (over 8192 vertices) whose parallel update groups fill at
 * least `windows` LDS windows runs on a team of workgroups, one per window: the parallel
 * groups' windows run concurrently (SoftBodyMotionProperties.cpp:1057-1115 spreads the same
 * groups over its job threads), the non-parallel group follows on one of them, and the
 * per-vertex phases are split.  Results are bit-identical to the one-workgroup path.
 * Bodies with rods or pressure, and contexts with dynamic or sensor colliders, keep the
 * one-workgroup path.  Default 3; 0 never forms teams.  No reference counterpart (a
 * scheduling choice).
 */
int xpbd_context_set_team_min_windows(xpbd_context *ctx, uint32_t windows);
/* Workgroups per team at least (0, the default: one per parallel-group window, two for a body
 * with the coloured non-parallel group; at most 224 in all).  Members beyond the windows take part in the per-vertex phases and the coloured
 * non-parallel group.  No reference counterpart. */
int xpbd_context_set_team_size(xpbd_context *ctx, uint32_t workgroups);

/*
 * How a body runs its non-parallel update group (the reference's last group, applied in
 * order by one thread, SoftBodyMotionProperties.cpp:1105-1115):
 *   XPBD_TAIL_EXACT     in the reference's order (default; bit-exact);
 *   XPBD_TAIL_COLOURED  re-coloured: within each constraint type, in the reference's order,
 *                       each constraint takes the lowest colour no earlier constraint holds on
 *                       a shared vertex either of them writes; the body runs on a team and
 *                       projects one colour at a time across all members.  A different
 *                       Gauss-Seidel order: NOT bit-exact with the reference (a throughput mode
 *                       for bodies whose non-parallel group is long).  Streamed bodies whose
 *                       non-parallel group holds only volumes, bends, edges and LRAs.
 * No reference counterpart.
 */
#define XPBD_TAIL_EXACT 0
#define XPBD_TAIL_COLOURED 1
int xpbd_body_set_tail_mode(xpbd_context *ctx, int32_t body, int32_t mode);
int xpbd_synchronize(xpbd_context *ctx);

/*
 * Vertex state in body-local space (SoftBodyMotionProperties::GetVertices,
 * SoftBodyMotionProperties.h:56-57).  Any pointer may be NULL.
 *   pos/prev/vel: V x 3 floats; plane: V x 4 (normal, constant);
 *   shape_idx: V int32 (-1 none); contact: V int32.
 * body_pos[3], body_rot[4]: body position and rotation.  Synchronizes.
 */
int xpbd_body_read(xpbd_context *ctx, int32_t body, float *pos, float *prev, float *vel, float *plane,
                   int32_t *shape_idx, int32_t *contact, float *body_pos, float *body_rot);
/* Overwrites positions / velocities (body-local space).  NULL skips. */
int xpbd_body_write(xpbd_context *ctx, int32_t body, const float *pos, const float *vel);
/* Body summary: out[0..2] linear velocity, out[3..5] angular velocity (world),
 * out[6..11] local bounds, out[12] active flag, out[13] sleep timer. */
int xpbd_body_state(xpbd_context *ctx, int32_t body, float out[14]);
/* Where the last xpbd_update ran a body: out[0] = XPBD_PLACED_BUNDLE (LDS-resident, sharing a
 * workgroup), XPBD_PLACED_STREAMED (one workgroup, windows from HBM) or XPBD_PLACED_TEAM;
 * out[1] = threads per workgroup; out[2] = bundle or team index (-1 streamed); out[3] =
 * bodies in the bundle, or workgroups in the team.  No reference counterpart. */
#define XPBD_PLACED_BUNDLE 0
#define XPBD_PLACED_STREAMED 1
#define XPBD_PLACED_TEAM 2
int xpbd_body_placement(xpbd_context *ctx, int32_t body, int32_t out[4]);

/*
 * Per-pair contact settings of a body: what SoftBodyContactListener::
 * OnSoftBodyContactValidate (SoftBodyContactListener.h:19-60) decides for the pair
 * (body, collider), applied by DetermineCollidingShapes (SoftBodyMotionProperties.cpp:
 * 151-216).  settings: num_colliders x {mInvMassScale1 (soft body), mInvMassScale2,
 * mInvInertiaScale2 (the collider), accept (1 AcceptContact, 0 RejectContact)}, in the
 * order of xpbd_set_colliders; a colliding pair without interaction (scale1 0 and, for a
 * dynamic collider, scale2 0) is skipped as the reference skips it.  The table holds for
 * every step of the following updates (the adapter asks the listener once per pair);
 * changing a pair's mIsSensor is not supported.  settings NULL removes the table (no
 * listener: every pair with unit scales; sensors are still detected and reported, so pass
 * sensor colliders only when a contact listener exists, as the reference considers sensors
 * only with one, SoftBodyMotionProperties.cpp:154-158).  A later xpbd_set_colliders whose
 * list differs in count, or at any index in user_id, shape type, motion type or sensor flag,
 * makes xpbd_update fail until the table is set again.
 */
int xpbd_body_set_contact_settings(xpbd_context *ctx, int32_t body, uint32_t num_colliders, const float *settings);

/*
 * Cosserat rod states of a body: SoftBodyMotionProperties::GetRodRotation /
 * GetRodAngularVelocity (SoftBodyMotionProperties.h:64-65) for all rods at once.
 * rotations: num_rods x 4 (xyzw, relative to the centre of mass transform);
 * angular_velocities: num_rods x 3; either may be NULL.  num_rods must equal the
 * settings' rod stretch shear constraints.
 */
int xpbd_body_rod_states(xpbd_context *ctx, int32_t body, uint32_t num_rods, float *rotations, float *angular_velocities);

/*
 * What soft body `body` did to the dynamic colliders during the last xpbd_update
 * (SoftBodyMotionProperties::UpdateRigidBodyVelocities, SoftBodyMotionProperties.cpp:923-935):
 * per collider of the last xpbd_set_colliders, the world-space linear and angular
 * velocity changes (3 floats each) and whether the body touched it (updated: the
 * reference's mUpdateVelocities; the caller then applies
 * BodyInterface::AddLinearAndAngularVelocity).  Summed over the call's collision steps;
 * zero for static and kinematic colliders.  Any output may be NULL.
 */
int xpbd_body_velocity_deltas(xpbd_context *ctx, int32_t body, uint32_t num_colliders, float *linear, float *angular,
                              int32_t *updated);

/*
 * The contacts of soft body `body` in the last collision step of the last xpbd_update, as
 * the contact listener's SoftBodyManifold presents them to OnSoftBodyContactAdded
 * (SoftBodyMotionProperties.cpp:832-845, SoftBodyManifold.h:24-40): per vertex the collider
 * index (into the last xpbd_set_colliders list, the manifold's GetContactBodyID) or -1 when
 * the vertex had no contact (HasContact), and the local contact point (GetLocalContactPoint,
 * in the body space of that step, before the step re-centred the vertices; 3 floats).  The
 * contact normal is minus the collision plane's normal (xpbd_body_read).  Any output may be NULL.
 */
int xpbd_body_contacts(xpbd_context *ctx, int32_t body, float *points, int32_t *collider);

/* Sensor contacts of soft body `body` in the last collision step of the last xpbd_update:
 * per collider of the last xpbd_set_colliders 1 when it is a sensor that detected a vertex
 * (the manifold's GetNumSensorContacts / GetSensorContactBodyID list), else 0. */
int xpbd_body_sensor_contacts(xpbd_context *ctx, int32_t body, uint32_t num_colliders, int32_t *hits);

/*
 * The byte stream SoftBodyMotionProperties::SaveState writes for a soft body
 * (SoftBodyMotionProperties.cpp:1446-1473, MotionProperties.cpp:63-76): linear and
 * angular velocity (world), force and torque (zero between updates), the three sleep
 * test spheres (not used by soft bodies; written as zeros), sleep timer, allow sleeping
 * (1 byte), then position and velocity of every vertex (body space), rotation (4 floats)
 * and angular velocity of every rod, previous skinned position, skinned position and
 * normal of every vertex of a skinned body, then the local bounds and local predicted
 * bounds (min, max each); Vec3s as 3 floats.  With `out`
 * NULL only *size is set.  The body's position and rotation are Body state, not part
 * of the stream (xpbd_body_read).
 */
/*
 * SoftBodyMotionProperties::SkinVertices (SoftBodyMotionProperties.cpp:1139-1226): skins
 * the skinned vertices with `num_joints` joint matrices (16 floats each, column major,
 * in the body's centre-of-mass space), updates their skin states and normals, and with
 * hard_skin_all sets every skinned vertex to its skinned position with zero velocity
 * (otherwise, with skin constraints disabled, only the kinematic ones, max distance 0).
 */
int xpbd_body_skin_vertices(xpbd_context *ctx, int32_t body, const float *joint_matrices, uint32_t num_joints,
                            int32_t hard_skin_all);
/* SetEnableSkinConstraints / SetSkinnedMaxDistanceMultiplier (SoftBodyMotionProperties.h:93-98) */
int xpbd_body_set_skin_options(xpbd_context *ctx, int32_t body, int32_t enable_skin_constraints,
                               float max_distance_multiplier);

int xpbd_body_save_state(xpbd_context *ctx, int32_t body, void *out, size_t capacity, size_t *size);

/*
 * Restores a stream of that layout (SoftBodyMotionProperties::RestoreState,
 * SoftBodyMotionProperties.cpp:1475-1502): velocities, sleep timer and allow-sleeping,
 * vertex positions and velocities, bounds; force, torque and the sleep spheres are read
 * and ignored.  position[3] / rotation[4] (may be NULL) set the body's position and
 * rotation, as Body::RestoreState would.  Fails on a size mismatch.
 */
int xpbd_body_restore_state(xpbd_context *ctx, int32_t body, const void *in, size_t size, const float *position,
                            const float *rotation);

/*
 * Kernel timing on the context stream (hipEvents around every launch of
 * xpbd_update while enabled).  xpbd_profile_read returns the number of timed
 * launches and their summed duration in milliseconds, then resets.
 */
int xpbd_profile_enable(xpbd_context *ctx, int32_t enable);
int xpbd_profile_read(xpbd_context *ctx, int32_t *num_launches, double *total_ms);

/*
 * Differentiable rollout.  No counterpart in the reference (Jolt computes no
 * gradients); it differentiates exactly the update above for bodies with edge,
 * volume, dihedral-bend and LRA constraints and optional pressure, no colliders,
 * no skinning, no sleeping and mUpdatePosition false, one collision step per
 * frame.  Each frame runs InitializeUpdateContext
 * (SoftBodyMotionProperties.cpp:937-960), then num_iterations x (ApplyPressure
 * :291-323, IntegratePositions :325-350, each update group's volume, bend, edge
 * and LRA projections :367-507, :573-601, :679-698, the velocity update
 * :700-710) and the velocity clamp
 * of UpdateSoftBodyState (:852-857); the float forward is bit-identical to
 * xpbd_update on such a body with identity rotation.
 *
 * Parameters are the per-edge compliance (SoftBodySharedSettings::Edge::
 * mCompliance, in each body's mEdgeConstraints order, bodies concatenated) and
 * the initial vertex velocities (3 floats per vertex, bodies concatenated);
 * initial positions and inverse masses are the settings' vertices.  The output
 * is the body-local vertex positions after `frames` frames.
 *
 * The backward pass is a hand-written adjoint: (x, v) checkpoints every
 * checkpoint_interval frames, each segment recomputed forward, then every
 * substep re-run with its pre-projection positions recorded and reversed level
 * by level in reverse order.  It returns d(loss)/d(compliance) and
 * d(loss)/d(initial velocity) given d(loss)/d(final positions).
 *
 * All array arguments are device pointers; `stream` is the hipStream_t the
 * work is ordered on (NULL: the default stream, as torch's default stream is;
 * xpbd_context_stream() gives the context's own); calls are asynchronous on it
 * and allocate nothing.
 */
typedef struct xpbd_rollout xpbd_rollout;

typedef struct xpbd_rollout_desc {
    xpbd_settings *const *settings;     /* one per body: edges, volumes, dihedral bends, LRAs, faces */
    uint32_t num_bodies;
    uint32_t num_iterations;            /* SoftBodyCreationSettings::mNumIterations */
    float linear_damping;               /* mLinearDamping */
    float max_linear_velocity;          /* mMaxLinearVelocity */
    float gravity_factor;               /* mGravityFactor; gravity is the context's */
    float dt;                           /* frame time step (one collision step per frame) */
    uint32_t frames;
    uint32_t checkpoint_interval;       /* frames between (x, v) checkpoints, >= 1 */
    float pressure;                     /* mPressure (n R T) of every body; > 0: ApplyPressure on each
                                           body's faces every sub-step (SoftBodyMotionProperties.cpp:291-323) */
} xpbd_rollout_desc;

int xpbd_rollout_create(xpbd_context *ctx, const xpbd_rollout_desc *desc, xpbd_rollout **out_rollout);
int xpbd_rollout_destroy(xpbd_rollout *rollout);
/* Total edges and vertices over all bodies (parameter and output lengths). */
int xpbd_rollout_sizes(xpbd_rollout *rollout, uint32_t *num_edges, uint32_t *num_vertices);
/* float forward; stores the checkpoints the backward pass starts from. */
int xpbd_rollout_forward(xpbd_rollout *rollout, const float *compliance, const float *velocity, float *positions, void *stream);
/*
 * xpbd_rollout_forward for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  Same results and checkpoints as xpbd_rollout_forward (a backward pass may
 * follow it); the forward pass's per-vertex velocity and previous-position state lives in
 * a fixed device array of the library instead of the rollout's buffers, sized for
 * XPBD_NOALLOC_MAX_VERTICES vertices (8 MiB, held on every GPU the library runs on until
 * the process ends).  A rollout with more vertices fails with XPBD_E_CAPACITY.  Every call
 * on one GPU shares that array: calls that overlap on one GPU (other streams, other
 * threads) race.  Allocates nothing.
 */
#define XPBD_NOALLOC_MAX_VERTICES 262144
int xpbd_rollout_forward_noalloc(xpbd_rollout *rollout, const float *compliance, const float *velocity, float *positions,
                                 void *stream);
/* double forward of the same update (finite-difference checks); no checkpoints. */
int xpbd_rollout_forward_f64(xpbd_rollout *rollout, const double *compliance, const double *velocity, double *positions,
                             void *stream);
/* Adjoint of the last float forward.  grad_positions: 3 floats per vertex. */
int xpbd_rollout_backward(xpbd_rollout *rollout, const float *grad_positions, float *grad_compliance, float *grad_velocity,
                          void *stream);

#ifdef __cplusplus
}
#endif

#endif /* JOLT_XPBD_H */
