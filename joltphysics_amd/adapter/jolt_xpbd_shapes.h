// Jolt Physics shapes as xpbd_collider_desc describes them (include/jolt_xpbd.h): the shape's
// parameters, convex hull tables and triangle trees exactly as the shape holds them, for a
// caller compiled next to Jolt Physics (v5.5.1).  Header only; uses Jolt's public API, and for
// values the shapes keep private, the bytes of Shape::SaveBinaryState.
//
//     ShapeData data;                      // keeps hull tables / trees alive
//     std::string err;
//     if (!DescribeShape(body.GetShape(), desc, data, err)) ...   // sets type, params, scale,
//                                                                  // hull or tree
#pragma once

#include <Jolt/Jolt.h>
#include <Jolt/AABBTree/NodeCodec/NodeCodecQuadTreeHalfFloat.h>
#include <Jolt/AABBTree/TriangleCodec/TriangleCodecIndexed8BitPackSOA4Flags.h>
#include <Jolt/Core/StreamWrapper.h>
#include <Jolt/Math/HalfFloat.h>
#include <Jolt/Math/Trigonometry.h>
#include <Jolt/Physics/Collision/Shape/BoxShape.h>
#include <Jolt/Physics/Collision/Shape/CapsuleShape.h>
#include <Jolt/Physics/Collision/Shape/ConvexHullShape.h>
#include <Jolt/Physics/Collision/Shape/CylinderShape.h>
#include <Jolt/Physics/Collision/Shape/HeightFieldShape.h>
#include <Jolt/Geometry/AABox4.h>
#include <Jolt/Physics/Collision/Shape/MeshShape.h>
#include <Jolt/Physics/Collision/Shape/MutableCompoundShape.h>
#include <Jolt/Physics/Collision/Shape/OffsetCenterOfMassShape.h>
#include <Jolt/Physics/Collision/Shape/PlaneShape.h>
#include <Jolt/Physics/Collision/Shape/RotatedTranslatedShape.h>
#include <Jolt/Physics/Collision/Shape/ScaledShape.h>
#include <Jolt/Physics/Collision/Shape/SphereShape.h>
#include <Jolt/Physics/Collision/Shape/StaticCompoundShape.h>
#include <Jolt/Physics/Collision/Shape/TaperedCapsuleShape.h>
#include <Jolt/Physics/Collision/Shape/TaperedCylinderShape.h>
#include <Jolt/Physics/Collision/Shape/TriangleShape.h>

#include <array>
#include <cstring>
#include <memory>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "jolt_xpbd.h"

namespace JoltXpbd {

// Tables a description points into; alive until xpbd_set_colliders returns
struct ShapeData
{
    std::vector<float> points, planes;
    std::vector<uint32_t> first, count, index;
    xpbd_convex_hull hull{};
    std::vector<xpbd_tree_node> nodes;
    std::vector<float> triangles;
    xpbd_triangle_tree tree{};
    // a compound or decorated shape: its leaves, their placement steps and tables
    std::vector<xpbd_collider_part> parts;
    std::vector<std::vector<xpbd_part_op>> part_ops;
    std::vector<std::unique_ptr<ShapeData>> part_data;
};

namespace shapes_detail {

inline std::string SavedState(const JPH::Shape &shape)
{
    std::stringstream ss;
    JPH::StreamOutWrapper out(ss);
    shape.SaveBinaryState(out);
    return ss.str();
}

// Shape::SaveBinaryState writes the sub type (1 byte) and the user data (8 bytes) first
// (Shape.cpp:71-75); ConvexShape adds the density (4 bytes, ConvexShape.cpp:530-535).
constexpr size_t kShapeStateBytes = 1 + 8;
constexpr size_t kConvexShapeStateBytes = kShapeStateBytes + 4;

// Four half floats from `p` widened as the reference widens them (HalfFloat.h:195-204)
inline void Widen4(const JPH::HalfFloat *p, float out[4])
{
    uint32_t w[4] = { 0, 0, 0, 0 };
    memcpy(w, p, 8);
    const JPH::Vec4 v = JPH::HalfFloatConversion::ToFloat(JPH::UVec4(w[0], w[1], w[2], w[3]));
    for (int k = 0; k < 4; ++k)
        out[k] = v[k];
}

} // namespace shapes_detail

// TaperedCylinderShape's {top, bottom, top radius, bottom radius} as the shape holds them
// (TaperedCylinderShape.cpp:108-130, written by SaveBinaryState, :644-653)
inline bool TaperedCylinderParams(const JPH::TaperedCylinderShape &shape, float out[4])
{
    const std::string s = shapes_detail::SavedState(shape);
    const size_t at = shapes_detail::kConvexShapeStateBytes;
    if (s.size() < at + 16)
        return false;
    memcpy(out, s.data() + at, 16);
    return true;
}

// TaperedCapsuleShape's {top centre, bottom centre, top radius, bottom radius, sin alpha, tan
// alpha} as the shape holds them: SaveBinaryState writes after ConvexShape's bytes the centre of
// mass (3 floats), the top and bottom radius, top and bottom centre, convex radius, sin and tan
// of the cone angle (TaperedCapsuleShape.cpp:406-418).  GetHalfHeight() is rebuilt from the
// centres, so the centres cannot be recomputed from it exactly.
inline bool TaperedCapsuleParams(const JPH::TaperedCapsuleShape &shape, float out[6])
{
    const std::string s = shapes_detail::SavedState(shape);
    const size_t at = shapes_detail::kConvexShapeStateBytes + 12;
    if (s.size() < at + 28)
        return false;
    float v[7];     // top radius, bottom radius, top centre, bottom centre, convex radius, sin, tan
    memcpy(v, s.data() + at, 28);
    out[0] = v[2]; out[1] = v[3]; out[2] = v[0]; out[3] = v[1]; out[4] = v[5]; out[5] = v[6];
    return true;
}

// A MeshShape's tree (MeshShape.cpp:481-489: node header, triangle header, nodes and triangle
// blocks in one buffer, written by SaveBinaryState as a uint32 length and the bytes) in the
// form the device walks: nodes with widened child bounds and references, the triangles of
// each leaf decoded by the mesh's own triangle codec in block order.
inline bool FlattenMesh(const JPH::MeshShape &mesh, ShapeData &out, std::string &err)
{
    using NodeCodec = JPH::NodeCodecQuadTreeHalfFloat;
    using TriangleCodec = JPH::TriangleCodecIndexed8BitPackSOA4Flags;
    const std::string s = shapes_detail::SavedState(mesh);
    const size_t at = shapes_detail::kShapeStateBytes;
    uint32_t len = 0;
    if (s.size() < at + 4)
    {
        err = "mesh state too short";
        return false;
    }
    memcpy(&len, s.data() + at, 4);
    if (s.size() < at + 4 + size_t(len) || len < NodeCodec::HeaderSize + sizeof(TriangleCodec::TriangleHeader))
    {
        err = "mesh tree truncated";
        return false;
    }
    // 16-byte aligned copy of the tree buffer
    std::vector<JPH::Vec4> storage((len + 15) / 16 + 1);
    uint8_t *tree = reinterpret_cast<uint8_t *>(storage.data());
    memcpy(tree, s.data() + at + 4, len);
    const NodeCodec::Header *header = reinterpret_cast<const NodeCodec::Header *>(tree);
    const TriangleCodec::DecodingContext triangles(reinterpret_cast<const TriangleCodec::TriangleHeader *>(tree + NodeCodec::HeaderSize));

    out.nodes.clear();
    out.triangles.clear();
    std::unordered_map<uint32_t, uint32_t> node_index;        // node offset -> flattened index
    // references are built depth first; node slots are reserved before their children
    struct Pending { uint32_t props; uint32_t node; int child; };
    auto leaf_ref = [&](uint32_t props, uint32_t &ref) -> bool {
        const uint32_t count = props >> NodeCodec::TRIANGLE_COUNT_SHIFT;
        const size_t offset = size_t(props & NodeCodec::OFFSET_MASK) << NodeCodec::OFFSET_NON_SIGNIFICANT_BITS;
        if (count == 0 || count >= (uint32_t)NodeCodec::TRIANGLE_COUNT_MASK || offset >= len)
        {
            err = "bad triangle block";
            return false;
        }
        JPH::Vec3 v[NodeCodec::TRIANGLE_COUNT_MASK * 3];
        triangles.Unpack(tree + offset, count, v);
        const uint32_t first = uint32_t(out.triangles.size() / 9);
        for (uint32_t k = 0; k < 3 * count; ++k)
            out.triangles.insert(out.triangles.end(), { v[k].GetX(), v[k].GetY(), v[k].GetZ() });
        ref = XPBD_TREE_LEAF | (count << 24) | first;
        return true;
    };
    std::vector<Pending> todo;
    auto classify = [&](uint32_t props, uint32_t &ref, bool &is_node) -> bool {
        const uint32_t count = props >> NodeCodec::TRIANGLE_COUNT_SHIFT;
        is_node = false;
        if (count == NodeCodec::TRIANGLE_COUNT_MASK)
        {
            ref = XPBD_TREE_EMPTY;
            return true;
        }
        if (count != 0)
            return leaf_ref(props, ref);
        const size_t offset = size_t(props) << NodeCodec::OFFSET_NON_SIGNIFICANT_BITS;
        if (offset + sizeof(NodeCodec::Node) > len || node_index.count(props))
        {
            err = "bad node reference";
            return false;
        }
        is_node = true;
        ref = XPBD_TREE_NODE | uint32_t(out.nodes.size());
        node_index[props] = uint32_t(out.nodes.size());
        out.nodes.emplace_back();
        return true;
    };
    uint32_t root;
    bool root_is_node;
    if (!classify(header->mRootProperties, root, root_is_node))
        return false;
    if (root_is_node)
        todo.push_back({ header->mRootProperties, root, -1 });
    while (!todo.empty())
    {
        const Pending p = todo.back();
        todo.pop_back();
        const NodeCodec::Node *node = reinterpret_cast<const NodeCodec::Node *>(tree + (size_t(p.props) << NodeCodec::OFFSET_NON_SIGNIFICANT_BITS));
        xpbd_tree_node n;
        shapes_detail::Widen4(node->mBoundsMinX, n.min_x);
        shapes_detail::Widen4(node->mBoundsMinY, n.min_y);
        shapes_detail::Widen4(node->mBoundsMinZ, n.min_z);
        shapes_detail::Widen4(node->mBoundsMaxX, n.max_x);
        shapes_detail::Widen4(node->mBoundsMaxY, n.max_y);
        shapes_detail::Widen4(node->mBoundsMaxZ, n.max_z);
        for (int k = 0; k < 4; ++k)
        {
            bool is_node;
            if (!classify(node->mNodeProperties[k], n.child[k], is_node))
                return false;
            if (is_node)
                todo.push_back({ node->mNodeProperties[k], n.child[k], k });
        }
        out.nodes[p.node] = n;
    }
    out.tree.root = root;
    out.tree.num_nodes = uint32_t(out.nodes.size());
    out.tree.nodes = out.nodes.data();
    out.tree.num_triangles = uint32_t(out.triangles.size() / 9);
    out.tree.triangles = out.triangles.data();
    out.tree.flags = 0;
    return true;
}

// A HeightFieldShape's range-block hierarchy as HeightFieldShape::DecodingContext::
// WalkHeightField visits it (HeightFieldShape.cpp:1757-2093): a node per range block above the
// finest level, its four children's bounds from the quantised range data (internal cells in
// x and z, min / max height in y); at the finest level a node per block whose four children
// are the block's quarters, bounds from their decoded samples (inside out when none has
// collision), each a leaf of its triangles in the walk's order (rows, columns, two triangles
// per cell, cells and triangles touching a sample without collision left out).  Positions
// come from HeightFieldShape::GetPosition / IsNoCollision, which decode a sample exactly as
// the walk does; the range data from SaveBinaryState (HeightFieldShape.cpp:2660-2682).  Only
// block counts that are powers of two (the shape's efficient case) are taken.
inline bool FlattenHeightField(const JPH::HeightFieldShape &hf, ShapeData &out, std::string &err)
{
    using namespace JPH;
    const std::string s = shapes_detail::SavedState(hf);
    size_t at = shapes_detail::kShapeStateBytes;
    auto get = [&](void *dst, size_t n) -> bool {
        if (at + n > s.size())
            return false;
        memcpy(dst, s.data() + at, n);
        at += n;
        return true;
    };
    float offset[3], scale[3];
    uint32_t sample_count, block_size, mat_count, mat_bits;
    uint8_t bits_per_sample, has_heights;
    uint16_t min_sample, max_sample;
    if (!get(offset, 12) || !get(scale, 12) || !get(&sample_count, 4) || !get(&block_size, 4) || !get(&bits_per_sample, 1)
        || !get(&min_sample, 2) || !get(&max_sample, 2) || !get(&mat_count, 4))
    {
        err = "height field state too short";
        return false;
    }
    at += mat_count;
    if (!get(&mat_bits, 4) || !get(&has_heights, 1))
    {
        err = "height field state too short";
        return false;
    }
    out.nodes.clear();
    out.triangles.clear();
    out.tree.flags = XPBD_TREE_HEIGHT_FIELD;
    out.tree.root = XPBD_TREE_EMPTY;
    out.tree.num_nodes = 0;
    out.tree.nodes = nullptr;
    out.tree.num_triangles = 0;
    out.tree.triangles = nullptr;
    if (!has_heights)
        return true;                // no samples: nothing collides (HeightFieldShape.cpp:1761-1762)
    if (block_size == 0 || sample_count % block_size != 0)
    {
        err = "height field block size";
        return false;
    }
    const uint32_t num_blocks = sample_count / block_size;
    if (num_blocks < 2)
    {
        err = "height field of one block";
        return false;
    }
    const uint32_t max_level = 32 - CountLeadingZeros(num_blocks - 1);
    static const uint32_t grid_offsets[] = { 0, 1, 5, 21, 85, 341, 1365, 5461, 21845, 87381, 349525, 1398101, 5592405, 22369621 };
    if (max_level >= sizeof(grid_offsets) / sizeof(grid_offsets[0]))
    {
        err = "height field too large";
        return false;
    }
    const char *range_blocks = s.data() + at;       // uint16 min[4], max[4] per range block
    const size_t range_bytes = s.size() - at;
    const uint32_t max_stride = (num_blocks + 1) >> 1;
    const Vec4 ox = Vec4::sReplicate(offset[0]), oy = Vec4::sReplicate(offset[1]), oz = Vec4::sReplicate(offset[2]);
    const Vec4 sx = Vec4::sReplicate(scale[0]), sy = Vec4::sReplicate(scale[1]), sz = Vec4::sReplicate(scale[2]);
    const UVec4 sample_count_min_1 = UVec4::sReplicate(sample_count - 1);
    const float large = 1.0e30f;                    // cLargeFloat

    struct Item { uint32_t level, x, y, node; };
    std::vector<Item> todo;
    auto new_node = [&]() { out.nodes.emplace_back(); return uint32_t(out.nodes.size() - 1); };
    out.tree.root = XPBD_TREE_NODE | new_node();
    todo.push_back({ 0, 0, 0, out.tree.root });
    while (!todo.empty())
    {
        const Item it = todo.back();
        todo.pop_back();
        xpbd_tree_node n;
        if (it.level < max_level)
        {
            // an internal range block: HeightFieldShape.cpp:2038-2062
            const uint32_t stride = std::min(1u << it.level, max_stride);
            const size_t offset16 = size_t(grid_offsets[it.level] + stride * it.y + it.x) * 16;
            if (offset16 + 16 > range_bytes)
            {
                err = "height field range data truncated";
                return false;
            }
            uint16_t mm[8];
            memcpy(mm, range_blocks + offset16, 16);
            const Vec4 bounds_miny = oy + sy * Vec4(float(mm[0]), float(mm[1]), float(mm[2]), float(mm[3]));
            const Vec4 bounds_maxy = oy + sy * Vec4(float(mm[4]), float(mm[5]), float(mm[6]), float(mm[7]));
            const UVec4 cell = UVec4::sReplicate(block_size << (max_level - it.level - 1));
            const UVec4 two_x = UVec4::sReplicate(2 * it.x), two_y = UVec4::sReplicate(2 * it.y);
            const Vec4 bounds_minx = ox + sx * (cell * (two_x + UVec4(0, 1, 0, 1))).ToFloat();
            const Vec4 bounds_maxx = ox + sx * UVec4::sMin(cell * (two_x + UVec4(1, 2, 1, 2)), sample_count_min_1).ToFloat();
            const Vec4 bounds_minz = oz + sz * (cell * (two_y + UVec4(0, 0, 1, 1))).ToFloat();
            const Vec4 bounds_maxz = oz + sz * UVec4::sMin(cell * (two_y + UVec4(1, 1, 2, 2)), sample_count_min_1).ToFloat();
            for (int k = 0; k < 4; ++k)
            {
                n.min_x[k] = bounds_minx[k]; n.min_y[k] = bounds_miny[k]; n.min_z[k] = bounds_minz[k];
                n.max_x[k] = bounds_maxx[k]; n.max_y[k] = bounds_maxy[k]; n.max_z[k] = bounds_maxz[k];
                // With a block count that is not a power of two, the last range blocks have
                // children past the last block: the shape stores them with min > max, so the
                // walk never enters them (their distance is FLT_MAX); they get no node.
                const uint32_t cx = 2 * it.x + (k & 1), cy = 2 * it.y + (k >> 1), shift = max_level - (it.level + 1);
                if ((cx << shift) >= num_blocks || (cy << shift) >= num_blocks)
                {
                    if (mm[k] <= mm[4 + k])
                    {
                        err = "height field range block past the last block with valid bounds";
                        return false;
                    }
                    n.child[k] = XPBD_TREE_EMPTY;
                    continue;
                }
                const uint32_t child = new_node();
                n.child[k] = XPBD_TREE_NODE | child;
                todo.push_back({ it.level + 1, cx, cy, child });
            }
        }
        else
        {
            // a block at the finest level: HeightFieldShape.cpp:1807-2036
            const uint32_t min_x = it.x * block_size, min_y = it.y * block_size;
            uint32_t max_x = min_x + block_size, max_y = min_y + block_size;
            if (it.x >= num_blocks - 1)
                --max_x;                    // no next block: one less triangle
            if (it.y >= num_blocks - 1)
                --max_y;
            const uint32_t half = block_size >> 1;
            const uint32_t bsx = max_x - min_x - half, bsy = max_y - min_y - half;
            const uint32_t range[4][4] = { { 0, 0, half, half }, { half, 0, bsx, half }, { 0, half, half, bsy }, { half, half, bsx, bsy } };
            for (int k = 0; k < 4; ++k)
            {
                Vec3 vmin = Vec3::sReplicate(large), vmax = Vec3::sReplicate(-large);
                for (uint32_t vy = min_y + range[k][1]; vy <= min_y + range[k][1] + range[k][3]; ++vy)
                    for (uint32_t vx = min_x + range[k][0]; vx <= min_x + range[k][0] + range[k][2]; ++vx)
                        if (!hf.IsNoCollision(vx, vy))
                        {
                            const Vec3 p = hf.GetPosition(vx, vy);
                            vmin = Vec3::sMin(vmin, p);
                            vmax = Vec3::sMax(vmax, p);
                        }
                n.min_x[k] = vmin.GetX(); n.min_y[k] = vmin.GetY(); n.min_z[k] = vmin.GetZ();
                n.max_x[k] = vmax.GetX(); n.max_y[k] = vmax.GetY(); n.max_z[k] = vmax.GetZ();
                // the quarter's triangles (HeightFieldShape.cpp:1959-2023)
                const uint32_t first = uint32_t(out.triangles.size() / 9);
                auto put = [&](uint32_t ax, uint32_t ay, uint32_t bx, uint32_t by, uint32_t cx, uint32_t cy) {
                    for (const Vec3 &p : { hf.GetPosition(ax, ay), hf.GetPosition(bx, by), hf.GetPosition(cx, cy) })
                        out.triangles.insert(out.triangles.end(), { p.GetX(), p.GetY(), p.GetZ() });
                };
                for (uint32_t vy = min_y + range[k][1]; vy < min_y + range[k][1] + range[k][3]; ++vy)
                    for (uint32_t vx = min_x + range[k][0]; vx < min_x + range[k][0] + range[k][2]; ++vx)
                    {
                        if (hf.IsNoCollision(vx, vy) || hf.IsNoCollision(vx + 1, vy + 1))
                            continue;
                        if (!hf.IsNoCollision(vx, vy + 1))
                            put(vx, vy, vx, vy + 1, vx + 1, vy + 1);
                        if (!hf.IsNoCollision(vx + 1, vy))
                            put(vx, vy, vx + 1, vy + 1, vx + 1, vy);
                    }
                const uint32_t count = uint32_t(out.triangles.size() / 9) - first;
                if (count > 63)
                {
                    err = "height field block too large";
                    return false;
                }
                n.child[k] = count == 0 ? XPBD_TREE_EMPTY : (XPBD_TREE_LEAF | (count << 24) | first);
            }
        }
        out.nodes[it.node] = n;
    }
    out.tree.num_nodes = uint32_t(out.nodes.size());
    out.tree.nodes = out.nodes.data();
    out.tree.num_triangles = uint32_t(out.triangles.size() / 9);
    out.tree.triangles = out.triangles.data();
    return true;
}

// A TriangleShape as a tree whose root is a leaf holding its triangle (TriangleShape.cpp:
// 272-283: the vertices as stored, in the shape's centre-of-mass space)
inline void TriangleTree(const JPH::TriangleShape &shape, ShapeData &out)
{
    out.nodes.clear();
    out.triangles.clear();
    for (const JPH::Vec3 &v : { shape.GetVertex1(), shape.GetVertex2(), shape.GetVertex3() })
        out.triangles.insert(out.triangles.end(), { v.GetX(), v.GetY(), v.GetZ() });
    out.tree.root = XPBD_TREE_LEAF | (1u << 24);
    out.tree.num_nodes = 0;
    out.tree.nodes = nullptr;
    out.tree.num_triangles = 1;
    out.tree.triangles = out.triangles.data();
    out.tree.flags = 0;
}

// Type, parameters, scale and tables of the leaf shape `shape` seen at scale `scale` (convex
// shapes, triangles, meshes and height fields; hull and tree point into `data`).  False (and `err`)
// for shapes the device does not collide with.
inline bool DescribeLeaf(const JPH::Shape *shape, JPH::Vec3Arg scale, int32_t &type, float params[8], float scale_out[3],
                         const xpbd_convex_hull *&hull, const xpbd_triangle_tree *&tree, ShapeData &data, std::string &err)
{
    scale_out[0] = scale.GetX(); scale_out[1] = scale.GetY(); scale_out[2] = scale.GetZ();
    hull = nullptr;
    tree = nullptr;
    for (int k = 0; k < 8; ++k)
        params[k] = 0.0f;
    switch (shape->GetSubType())
    {
    case JPH::EShapeSubType::Plane:
        {
            const auto *s = static_cast<const JPH::PlaneShape *>(shape);
            const JPH::Plane &pl = s->GetPlane();
            type = XPBD_SHAPE_PLANE;
            params[0] = pl.GetNormal().GetX(); params[1] = pl.GetNormal().GetY(); params[2] = pl.GetNormal().GetZ();
            params[3] = pl.GetConstant(); params[4] = s->GetHalfExtent();
        }
        return true;
    case JPH::EShapeSubType::Sphere:
        type = XPBD_SHAPE_SPHERE;
        params[0] = static_cast<const JPH::SphereShape *>(shape)->GetRadius();
        return true;
    case JPH::EShapeSubType::Box:
        {
            const JPH::Vec3 he = static_cast<const JPH::BoxShape *>(shape)->GetHalfExtent();
            type = XPBD_SHAPE_BOX;
            params[0] = he.GetX(); params[1] = he.GetY(); params[2] = he.GetZ();
        }
        return true;
    case JPH::EShapeSubType::Capsule:
        {
            const auto *s = static_cast<const JPH::CapsuleShape *>(shape);
            type = XPBD_SHAPE_CAPSULE;
            params[0] = s->GetHalfHeightOfCylinder(); params[1] = s->GetRadius();
        }
        return true;
    case JPH::EShapeSubType::Cylinder:
        {
            const auto *s = static_cast<const JPH::CylinderShape *>(shape);
            type = XPBD_SHAPE_CYLINDER;
            params[0] = s->GetHalfHeight(); params[1] = s->GetRadius();
        }
        return true;
    case JPH::EShapeSubType::TaperedCapsule:
        // centres relative to the centre of mass and the cone angle, as the shape holds them
        type = XPBD_SHAPE_TAPERED_CAPSULE;
        if (!TaperedCapsuleParams(*static_cast<const JPH::TaperedCapsuleShape *>(shape), params))
        {
            err = "tapered capsule state";
            return false;
        }
        return true;
    case JPH::EShapeSubType::TaperedCylinder:
        type = XPBD_SHAPE_TAPERED_CYLINDER;
        if (!TaperedCylinderParams(*static_cast<const JPH::TaperedCylinderShape *>(shape), params))
        {
            err = "tapered cylinder state";
            return false;
        }
        return true;
    case JPH::EShapeSubType::ConvexHull:
        {
            const auto *s = static_cast<const JPH::ConvexHullShape *>(shape);
            data.points.clear(); data.planes.clear(); data.first.clear(); data.count.clear(); data.index.clear();
            for (JPH::uint i = 0; i < s->GetNumPoints(); ++i)
            {
                const JPH::Vec3 p = s->GetPoint(i);
                data.points.insert(data.points.end(), { p.GetX(), p.GetY(), p.GetZ() });
            }
            for (JPH::uint f = 0; f < s->GetNumFaces(); ++f)
            {
                const JPH::Plane &pl = s->GetPlanes()[f];
                data.planes.insert(data.planes.end(), { pl.GetNormal().GetX(), pl.GetNormal().GetY(), pl.GetNormal().GetZ(), pl.GetConstant() });
                JPH::uint vi[256];
                const JPH::uint n = s->GetFaceVertices(f, 256, vi);
                data.first.push_back((uint32_t)data.index.size());
                data.count.push_back(n);
                data.index.insert(data.index.end(), vi, vi + n);
            }
            data.hull = { s->GetNumPoints(), data.points.data(), s->GetNumFaces(), data.planes.data(), data.first.data(),
                          data.count.data(), (uint32_t)data.index.size(), data.index.data() };
            type = XPBD_SHAPE_CONVEX_HULL;
            hull = &data.hull;
        }
        return true;
    case JPH::EShapeSubType::Triangle:
        TriangleTree(*static_cast<const JPH::TriangleShape *>(shape), data);
        type = XPBD_SHAPE_MESH;
        tree = &data.tree;
        return true;
    case JPH::EShapeSubType::Mesh:
        if (!FlattenMesh(*static_cast<const JPH::MeshShape *>(shape), data, err))
            return false;
        type = XPBD_SHAPE_MESH;
        tree = &data.tree;
        return true;
    case JPH::EShapeSubType::HeightField:
        if (!FlattenHeightField(*static_cast<const JPH::HeightFieldShape *>(shape), data, err))
            return false;
        type = XPBD_SHAPE_MESH;
        tree = &data.tree;
        return true;
    default:
        err = "a soft body reaches a collider shape the device path does not support";
        return false;
    }
}

namespace shapes_detail {

inline xpbd_part_op Op(uint32_t kind, std::initializer_list<float> v)
{
    xpbd_part_op op{};
    op.kind = kind;
    int k = 0;
    for (float x : v)
        op.v[k++] = x;
    return op;
}

// One bounds step: a child's (or tree node's) box, scaled by the compound's scale the way the
// compound's visitor scales it (AABox4Scale, AABox4.h:37-52)
inline xpbd_part_op CullOp(JPH::Vec3Arg scale, const float mn[3], const float mx[3])
{
    JPH::Vec4 bmin_x(mn[0], 0, 0, 0), bmin_y(mn[1], 0, 0, 0), bmin_z(mn[2], 0, 0, 0);
    JPH::Vec4 bmax_x(mx[0], 0, 0, 0), bmax_y(mx[1], 0, 0, 0), bmax_z(mx[2], 0, 0, 0);
    JPH::Vec4 omin_x, omin_y, omin_z, omax_x, omax_y, omax_z;
    JPH::AABox4Scale(scale, bmin_x, bmin_y, bmin_z, bmax_x, bmax_y, bmax_z, omin_x, omin_y, omin_z, omax_x, omax_y, omax_z);
    return Op(XPBD_OP_CULL, { omin_x.GetX(), omin_y.GetX(), omin_z.GetX(), omax_x.GetX(), omax_y.GetX(), omax_z.GetX() });
}

// StaticCompoundShape's tree (the bytes SaveBinaryState writes after CompoundShape's: node count,
// then per node 4 x 6 half-float bounds and 4 child words, StaticCompoundShape.h:100-124; IS_SUBSHAPE
// 0x80000000, INVALID_NODE 0x7fffffff): for each sub shape in the order WalkTree visits them (a
// node's children pushed in order, the last one popped first, StaticCompoundShape.cpp:358-419), the
// bounds of every entry on its way down from the root node
struct StaticLeaf
{
    uint32_t sub_shape;
    std::vector<std::array<float, 6>> path;     // min xyz, max xyz per level, root first
};

inline bool StaticCompoundOrder(const JPH::StaticCompoundShape &shape, std::vector<StaticLeaf> &out, std::string &err)
{
    const std::string state = SavedState(shape);
    // CompoundShape::SaveBinaryState (CompoundShape.cpp:394-405): Shape's header, centre of mass,
    // local bounds, inner radius, sub shape count and per sub shape: user data, position, rotation
    size_t at = kShapeStateBytes + 12 + 24 + 4;
    uint32_t num_sub = 0;
    if (state.size() < at + 4)
    {
        err = "static compound state";
        return false;
    }
    memcpy(&num_sub, state.data() + at, 4);
    at += 4 + (size_t)num_sub * (4 + 12 + 12);
    uint32_t num_nodes = 0;
    if (state.size() < at + 4)
    {
        err = "static compound state";
        return false;
    }
    memcpy(&num_nodes, state.data() + at, 4);
    at += 4;
    constexpr size_t kNodeBytes = 24 * 2 + 4 * 4;
    if (num_nodes == 0 || state.size() < at + (size_t)num_nodes * kNodeBytes)
    {
        err = "static compound tree state";
        return false;
    }
    struct Node { float mn[3][4], mx[3][4]; uint32_t child[4]; };
    std::vector<Node> nodes(num_nodes);
    for (uint32_t n = 0; n < num_nodes; ++n)
    {
        const char *p = state.data() + at + (size_t)n * kNodeBytes;
        JPH::HalfFloat h[24];
        memcpy(h, p, sizeof(h));
        for (int a = 0; a < 3; ++a)
        {
            Widen4(h + 4 * a, nodes[n].mn[a]);
            Widen4(h + 12 + 4 * a, nodes[n].mx[a]);
        }
        memcpy(nodes[n].child, p + 48, 16);
    }
    constexpr uint32_t kSubShape = 0x80000000u, kInvalid = 0x7fffffffu;
    struct Entry { uint32_t ref; std::vector<std::array<float, 6>> path; };
    std::vector<Entry> stack{ { 0u, {} } };
    while (!stack.empty())
    {
        Entry e = std::move(stack.back());
        stack.pop_back();
        if (e.ref == kInvalid)
            continue;
        if (e.ref & kSubShape)
        {
            out.push_back({ e.ref ^ kSubShape, std::move(e.path) });
            continue;
        }
        if (e.ref >= num_nodes || e.path.size() > 64)
        {
            err = "static compound tree state";
            return false;
        }
        const Node &nd = nodes[e.ref];
        for (int k = 0; k < 4; ++k)
        {
            if (nd.child[k] == kInvalid)
                continue;
            Entry c{ nd.child[k], e.path };
            c.path.push_back({ nd.mn[0][k], nd.mn[1][k], nd.mn[2][k], nd.mx[0][k], nd.mx[1][k], nd.mx[2][k] });
            stack.push_back(std::move(c));
        }
    }
    if (out.size() != num_sub)
    {
        err = "static compound tree does not reach every sub shape once";
        return false;
    }
    return true;
}

// MutableCompoundShape's per sub shape bounds (the Bounds blocks SaveBinaryState writes after
// CompoundShape's: 6 x 4 floats per block of 4 sub shapes, MutableCompoundShape.cpp:560-568)
inline bool MutableCompoundBounds(const JPH::MutableCompoundShape &shape, std::vector<std::array<float, 6>> &out, std::string &err)
{
    const std::string state = SavedState(shape);
    size_t at = kShapeStateBytes + 12 + 24 + 4;
    uint32_t num_sub = 0;
    if (state.size() < at + 4)
    {
        err = "mutable compound state";
        return false;
    }
    memcpy(&num_sub, state.data() + at, 4);
    at += 4 + (size_t)num_sub * (4 + 12 + 12);
    const uint32_t num_blocks = (num_sub + 3) / 4;
    if (state.size() < at + (size_t)num_blocks * 96)
    {
        err = "mutable compound bounds state";
        return false;
    }
    out.resize(num_sub);
    for (uint32_t i = 0; i < num_sub; ++i)
    {
        const char *block = state.data() + at + (size_t)(i / 4) * 96;
        for (int a = 0; a < 6; ++a)
            memcpy(&out[i][a], block + 16 * a + 4 * (i % 4), 4);
    }
    return true;
}

} // namespace shapes_detail

// The leaves of `shape` at scale `scale`, after the placement steps `ops`, in the order
// CollectTransformedShapes reaches them: ScaledShape multiplies the scale (ScaledShape.cpp:147-154),
// RotatedTranslatedShape rotates (RotatedTranslatedShape.cpp:169-176), OffsetCenterOfMassShape
// shifts (OffsetCenterOfMassShape.cpp:135-142) and compounds test each child's bounds, then move
// to and rotate into the child (CompoundShapeVisitors.h:226-272).
inline bool CollectParts(const JPH::Shape *shape, JPH::Vec3Arg scale, std::vector<xpbd_part_op> &ops, ShapeData &data, std::string &err)
{
    using namespace shapes_detail;
    switch (shape->GetSubType())
    {
    case JPH::EShapeSubType::Scaled:
        {
            const auto *s = static_cast<const JPH::ScaledShape *>(shape);
            return CollectParts(s->GetInnerShape(), scale * s->GetScale(), ops, data, err);
        }
    case JPH::EShapeSubType::RotatedTranslated:
        {
            const auto *s = static_cast<const JPH::RotatedTranslatedShape *>(shape);
            const JPH::Quat r = s->GetRotation();
            ops.push_back(Op(XPBD_OP_ROTATE, { r.GetX(), r.GetY(), r.GetZ(), r.GetW() }));
            const bool ok = CollectParts(s->GetInnerShape(), s->TransformScale(scale), ops, data, err);
            ops.pop_back();
            return ok;
        }
    case JPH::EShapeSubType::OffsetCenterOfMass:
        {
            const auto *s = static_cast<const JPH::OffsetCenterOfMassShape *>(shape);
            const JPH::Vec3 off = scale * s->GetOffset();
            ops.push_back(Op(XPBD_OP_OFFSET, { off.GetX(), off.GetY(), off.GetZ() }));
            const bool ok = CollectParts(s->GetInnerShape(), scale, ops, data, err);
            ops.pop_back();
            return ok;
        }
    case JPH::EShapeSubType::StaticCompound:
    case JPH::EShapeSubType::MutableCompound:
        {
            const auto *cs = static_cast<const JPH::CompoundShape *>(shape);
            // (sub shape, bounds steps) in collection order
            std::vector<std::pair<uint32_t, std::vector<std::array<float, 6>>>> order;
            if (shape->GetSubType() == JPH::EShapeSubType::StaticCompound)
            {
                std::vector<StaticLeaf> leaves;
                if (!StaticCompoundOrder(*static_cast<const JPH::StaticCompoundShape *>(shape), leaves, err))
                    return false;
                for (StaticLeaf &l : leaves)
                    order.emplace_back(l.sub_shape, std::move(l.path));
            }
            else
            {
                std::vector<std::array<float, 6>> bounds;
                if (!MutableCompoundBounds(*static_cast<const JPH::MutableCompoundShape *>(shape), bounds, err))
                    return false;
                for (uint32_t i = 0; i < (uint32_t)bounds.size(); ++i)
                    order.push_back({ i, { bounds[i] } });
            }
            for (const auto &entry : order)
            {
                const JPH::CompoundShape::SubShape &sub = cs->GetSubShape(entry.first);
                const size_t depth = ops.size();
                for (const std::array<float, 6> &b : entry.second)
                    ops.push_back(CullOp(scale, b.data(), b.data() + 3));
                const JPH::Vec3 pos = scale * sub.GetPositionCOM();
                const JPH::Quat rot = sub.GetRotation();
                ops.push_back(Op(XPBD_OP_TRANSLATE, { pos.GetX(), pos.GetY(), pos.GetZ() }));
                ops.push_back(Op(XPBD_OP_ROTATE, { rot.GetX(), rot.GetY(), rot.GetZ(), rot.GetW() }));
                const bool ok = CollectParts(sub.mShape, sub.TransformScale(scale), ops, data, err);
                ops.resize(depth);
                if (!ok)
                    return false;
            }
            return true;
        }
    default:
        {
            if (ops.size() > XPBD_MAX_PART_OPS)
            {
                err = "a compound leaf lies more than XPBD_MAX_PART_OPS placement steps deep";
                return false;
            }
            data.part_data.push_back(std::make_unique<ShapeData>());
            xpbd_collider_part part{};
            if (!DescribeLeaf(shape, scale, part.type, part.params, part.scale, part.hull, part.tree, *data.part_data.back(), err))
                return false;
            data.part_ops.push_back(ops);
            data.parts.push_back(part);
            return true;
        }
    }
}

// Shape type, parameters, scale and tables of `shape` into `d` (position, rotation, material
// and motion are the caller's).  A leaf shape inside ScaledShape wrappers fills type, params,
// scale, hull / tree; compounds and RotatedTranslated / OffsetCenterOfMass shapes fill
// num_parts / parts.  False (and `err`) for shapes the device does not collide with.
inline bool DescribeShape(const JPH::Shape *shape, xpbd_collider_desc &d, ShapeData &data, std::string &err)
{
    JPH::Vec3 scale = JPH::Vec3::sOne();
    const JPH::Shape *leaf = shape;
    while (leaf->GetSubType() == JPH::EShapeSubType::Scaled)
    {
        const auto *s = static_cast<const JPH::ScaledShape *>(leaf);
        scale = scale * s->GetScale();
        leaf = s->GetInnerShape();
    }
    d.num_parts = 0;
    d.parts = nullptr;
    const JPH::EShapeType t = leaf->GetType();
    if (t != JPH::EShapeType::Compound && t != JPH::EShapeType::Decorated)
        return DescribeLeaf(leaf, scale, d.type, d.params, d.scale, d.hull, d.tree, data, err);
    data.parts.clear();
    data.part_ops.clear();
    data.part_data.clear();
    std::vector<xpbd_part_op> ops;
    if (!CollectParts(shape, JPH::Vec3::sOne(), ops, data, err))
        return false;
    if (data.parts.empty())
    {
        err = "compound shape without leaves";
        return false;
    }
    for (size_t k = 0; k < data.parts.size(); ++k)
    {
        data.parts[k].num_ops = (uint32_t)data.part_ops[k].size();
        data.parts[k].ops = data.part_ops[k].data();
    }
    d.type = data.parts[0].type;
    for (int k = 0; k < 8; ++k)
        d.params[k] = 0.0f;
    d.scale[0] = d.scale[1] = d.scale[2] = 1.0f;
    d.hull = nullptr;
    d.tree = nullptr;
    d.num_parts = (uint32_t)data.parts.size();
    d.parts = data.parts.data();
    return true;
}

} // namespace JoltXpbd
