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
#include <Jolt/Physics/Collision/Shape/MeshShape.h>
#include <Jolt/Physics/Collision/Shape/PlaneShape.h>
#include <Jolt/Physics/Collision/Shape/ScaledShape.h>
#include <Jolt/Physics/Collision/Shape/SphereShape.h>
#include <Jolt/Physics/Collision/Shape/TaperedCapsuleShape.h>
#include <Jolt/Physics/Collision/Shape/TaperedCylinderShape.h>
#include <Jolt/Physics/Collision/Shape/TriangleShape.h>

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
    if (num_blocks < 2 || (num_blocks & (num_blocks - 1)) != 0)
    {
        err = "height field with a block count that is not a power of two";
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
                const uint32_t child = new_node();
                n.child[k] = XPBD_TREE_NODE | child;
                todo.push_back({ it.level + 1, 2 * it.x + (k & 1), 2 * it.y + (k >> 1), child });
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

// Shape type, parameters, scale and tables of `shape` into `d` (position, rotation, material
// and motion are the caller's).  ScaledShape wrappers multiply into d.scale (ScaledShape.cpp:
// 142-150).  False (and `err`) for shapes the device does not collide with.
inline bool DescribeShape(const JPH::Shape *shape, xpbd_collider_desc &d, ShapeData &data, std::string &err)
{
    JPH::Vec3 scale = JPH::Vec3::sOne();
    while (shape->GetSubType() == JPH::EShapeSubType::Scaled)
    {
        const auto *s = static_cast<const JPH::ScaledShape *>(shape);
        scale = scale * s->GetScale();
        shape = s->GetInnerShape();
    }
    d.scale[0] = scale.GetX(); d.scale[1] = scale.GetY(); d.scale[2] = scale.GetZ();
    d.hull = nullptr;
    d.tree = nullptr;
    switch (shape->GetSubType())
    {
    case JPH::EShapeSubType::Plane:
        {
            const auto *s = static_cast<const JPH::PlaneShape *>(shape);
            const JPH::Plane &pl = s->GetPlane();
            d.type = XPBD_SHAPE_PLANE;
            d.params[0] = pl.GetNormal().GetX(); d.params[1] = pl.GetNormal().GetY(); d.params[2] = pl.GetNormal().GetZ();
            d.params[3] = pl.GetConstant(); d.params[4] = s->GetHalfExtent();
        }
        return true;
    case JPH::EShapeSubType::Sphere:
        d.type = XPBD_SHAPE_SPHERE;
        d.params[0] = static_cast<const JPH::SphereShape *>(shape)->GetRadius();
        return true;
    case JPH::EShapeSubType::Box:
        {
            const JPH::Vec3 he = static_cast<const JPH::BoxShape *>(shape)->GetHalfExtent();
            d.type = XPBD_SHAPE_BOX;
            d.params[0] = he.GetX(); d.params[1] = he.GetY(); d.params[2] = he.GetZ();
        }
        return true;
    case JPH::EShapeSubType::Capsule:
        {
            const auto *s = static_cast<const JPH::CapsuleShape *>(shape);
            d.type = XPBD_SHAPE_CAPSULE;
            d.params[0] = s->GetHalfHeightOfCylinder(); d.params[1] = s->GetRadius();
        }
        return true;
    case JPH::EShapeSubType::Cylinder:
        {
            const auto *s = static_cast<const JPH::CylinderShape *>(shape);
            d.type = XPBD_SHAPE_CYLINDER;
            d.params[0] = s->GetHalfHeight(); d.params[1] = s->GetRadius();
        }
        return true;
    case JPH::EShapeSubType::TaperedCapsule:
        {
            // centres relative to the centre of mass and the cone angle, as the shape's
            // constructor derives them (TaperedCapsuleShape.cpp:86-131)
            const auto *s = static_cast<const JPH::TaperedCapsuleShape *>(shape);
            const float half_height = s->GetHalfHeight(), top = s->GetTopRadius(), bottom = s->GetBottomRadius();
            const float top_center = half_height + 0.5f * (bottom - top);
            const float bottom_center = -half_height + 0.5f * (bottom - top);
            const float sin_alpha = (bottom - top) / (top_center - bottom_center);
            d.type = XPBD_SHAPE_TAPERED_CAPSULE;
            d.params[0] = top_center; d.params[1] = bottom_center; d.params[2] = top; d.params[3] = bottom;
            d.params[4] = sin_alpha; d.params[5] = JPH::Tan(JPH::ASin(sin_alpha));
        }
        return true;
    case JPH::EShapeSubType::TaperedCylinder:
        d.type = XPBD_SHAPE_TAPERED_CYLINDER;
        if (!TaperedCylinderParams(*static_cast<const JPH::TaperedCylinderShape *>(shape), d.params))
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
            d.type = XPBD_SHAPE_CONVEX_HULL;
            d.hull = &data.hull;
        }
        return true;
    case JPH::EShapeSubType::Triangle:
        TriangleTree(*static_cast<const JPH::TriangleShape *>(shape), data);
        d.type = XPBD_SHAPE_MESH;
        d.tree = &data.tree;
        return true;
    case JPH::EShapeSubType::Mesh:
        if (!FlattenMesh(*static_cast<const JPH::MeshShape *>(shape), data, err))
            return false;
        d.type = XPBD_SHAPE_MESH;
        d.tree = &data.tree;
        return true;
    case JPH::EShapeSubType::HeightField:
        if (!FlattenHeightField(*static_cast<const JPH::HeightFieldShape *>(shape), data, err))
            return false;
        d.type = XPBD_SHAPE_MESH;
        d.tree = &data.tree;
        return true;
    default:
        err = "a soft body reaches a collider shape the device path does not support (compound, rotated / "
              "translated or offset centre of mass shapes)";
        return false;
    }
}

} // namespace JoltXpbd
