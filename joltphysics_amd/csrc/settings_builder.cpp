// Host-side soft-body settings builder: the product's counterpart of
// SoftBodySharedSettings' constraint generation and grouping
// (Jolt/Physics/SoftBody/SoftBodySharedSettings.cpp).  It lets a caller that
// has only a mesh build the exact arrays the reference would build (same
// constraints, same order, same update groups) and feed them to
// xpbd_settings_create.
//
// The output must equal the reference's array for array, so the steps that
// decide the order of equal keys are restated as the reference performs them:
// its quicksort (Hoare partition, ninther pivot, insertion sort below 33
// elements; Core/QuickSort.h, Core/InsertionSort.h) and its binary heap
// (Core/BinaryHeap.h) used by the closest-kinematic Dijkstra search.  Float
// expressions follow the reference's operation order (jolt_math.h) and the
// file is compiled without FP contraction.

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/jolt_xpbd.h"
#include "jolt_math.h"

using namespace xpbd;

namespace {

thread_local std::string g_build_error = "no error";

int BuildFail(int code, const std::string &msg)
{
    g_build_error = msg;
    return code;
}

struct VertexRec { float pos[3], vel[3], inv_mass; };
struct FaceRec { uint32_t v[3], material; };
struct EdgeRec { uint32_t v[2]; float rest, compliance; };
struct BendRec { uint32_t v[4]; float compliance, initial_angle; };
struct VolumeRec { uint32_t v[4]; float six_rest_volume, compliance; };
struct LRARec { uint32_t v[2]; float max_distance; };
struct GroupRec { uint32_t edge_end, lra_end, rod_ss_end, rod_bt_end, bend_end, volume_end, skinned_end; };

// ------------------------------------------------------------------ sorting
// Insertion sort with the reference's step order (Core/InsertionSort.h:14-49).
template <class It, class Cmp>
void InsertionSortSteps(It begin, It end, Cmp cmp)
{
    if (begin == end)
        return;
    for (It i = begin + 1; i != end; ++i)
    {
        auto x = std::move(*i);
        if (cmp(x, *begin))
        {
            for (It j = i; j != begin; --j)
                *j = std::move(*(j - 1));
            *begin = std::move(x);
        }
        else
        {
            It j = i;
            for (It prev = j - 1; cmp(x, *prev); j = prev, --prev)
                *j = std::move(*prev);
            *j = std::move(x);
        }
    }
}

template <class It, class Cmp>
void MedianOfThree(It first, It middle, It last, Cmp cmp)
{
    if (cmp(*middle, *first)) std::swap(*first, *middle);
    if (cmp(*last, *first)) std::swap(*first, *last);
    if (cmp(*last, *middle)) std::swap(*middle, *last);
}

template <class It, class Cmp>
void Ninther(It first, It middle, It last, Cmp cmp)
{
    auto diff = (last - first) >> 3;
    auto two_diff = diff << 1;
    It mid1 = first + diff;
    MedianOfThree(first, mid1, first + two_diff, cmp);
    MedianOfThree(middle - diff, middle, middle + diff, cmp);
    It mid3 = last - diff;
    MedianOfThree(last - two_diff, mid3, last, cmp);
    MedianOfThree(mid1, middle, mid3, cmp);
}

// Quicksort with the reference's step order (Core/QuickSort.h:45-137).
template <class It, class Cmp>
void QuickSortSteps(It begin, It end, Cmp cmp)
{
    for (;;)
    {
        auto n = end - begin;
        if (n < 2)
            return;
        if (n <= 32)
        {
            InsertionSortSteps(begin, end, cmp);
            return;
        }
        It pivot_it = begin + ((n - 1) >> 1);
        Ninther(begin, pivot_it, end - 1, cmp);
        auto pivot = *pivot_it;
        It i = begin, j = end;
        for (;;)
        {
            while (cmp(*i, pivot))
                ++i;
            do
                --j;
            while (cmp(pivot, *j));
            if (i >= j)
                break;
            std::swap(*i, *j);
            ++i;
        }
        ++j;
        if (j - begin < end - j)
        {
            QuickSortSteps(begin, j, cmp);
            begin = j;
        }
        else
        {
            QuickSortSteps(j, end, cmp);
            end = j;
        }
    }
}

// Binary heap push / pop with the reference's step order (Core/BinaryHeap.h:14-92).
template <class It, class Pred>
void HeapPush(It begin, It end, Pred pred)
{
    auto current = (end - begin) - 1;
    while (current > 0)
    {
        auto parent = (current - 1) >> 1;
        if (pred(*(begin + parent), *(begin + current)))
        {
            std::swap(*(begin + parent), *(begin + current));
            current = parent;
        }
        else
            break;
    }
}

template <class It, class Pred>
void HeapPop(It begin, It end, Pred pred)
{
    std::swap(*(end - 1), *begin);
    auto count = (end - begin) - 1;
    decltype(count) largest = 0;
    for (;;)
    {
        auto child = (largest << 1) + 1;
        if (child >= count)
            break;
        auto prev_largest = largest;
        if (pred(*(begin + largest), *(begin + child)))
            largest = child;
        ++child;
        if (child < count && pred(*(begin + largest), *(begin + child)))
            largest = child;
        if (prev_largest == largest)
            break;
        std::swap(*(begin + prev_largest), *(begin + largest));
    }
}

// ------------------------------------------------------------------ sin / cos
// Cephes sinf/cosf polynomial with two-step Cody-Waite reduction, as the
// reference evaluates it for scalar inputs (Math/Vec4.inl:1171-1231).
void SinCos(float in, float &out_sin, float &out_cos)
{
    uint32_t bits;
    memcpy(&bits, &in, 4);
    uint32_t sin_sign = bits & 0x80000000u;
    uint32_t xb = bits ^ sin_sign;
    float x;
    memcpy(&x, &xb, 4);
    uint32_t quadrant = (uint32_t)(int32_t)(0.6366197723675814f * x + 0.5f);
    float fq = (float)(int32_t)quadrant;
    x = ((x - fq * 1.5703125f) - fq * 0.0004837512969970703125f) - fq * 7.549789948768648e-8f;
    float x2 = x * x;
    float taylor_cos = ((((2.443315711809948e-5f * x2 - 1.388731625493765e-3f) * x2 + 4.166664568298827e-2f) * x2) * x2 - 0.5f * x2) + 1.0f;
    float taylor_sin = ((((-1.9515295891e-4f * x2 + 8.3321608736e-3f) * x2 - 1.6666654611e-1f) * x2) * x) + x;
    uint32_t bit1 = quadrant << 31;
    uint32_t bit2 = (quadrant << 30) & 0x80000000u;
    float s = bit1 ? taylor_cos : taylor_sin;
    float c = bit1 ? taylor_sin : taylor_cos;
    uint32_t ss = sin_sign ^ bit2, cs = bit1 ^ bit2;
    uint32_t sb, cb;
    memcpy(&sb, &s, 4);
    memcpy(&cb, &c, 4);
    sb ^= ss;
    cb ^= cs;
    memcpy(&out_sin, &sb, 4);
    memcpy(&out_cos, &cb, 4);
}

inline float MinF(float a, float b) { return a < b ? a : b; }
inline float MaxF(float a, float b) { return a > b ? a : b; }
inline float Square(float a) { return a * a; }

struct Attr { float compliance, shear, bend; int32_t lra_type; float lra_mult; };

struct Closest { uint32_t vertex = 0xffffffffu; uint32_t hops = 0xffffffffu; float distance = FLT_MAX; };

struct Builder
{
    std::vector<VertexRec> vertices;
    std::vector<FaceRec> faces;
    std::vector<EdgeRec> edges;
    std::vector<BendRec> bends;
    std::vector<VolumeRec> volumes;
    std::vector<LRARec> lras;
    std::vector<GroupRec> groups;
    std::vector<Closest> closest;

    V3 X(uint32_t i) const { return v3(vertices[i].pos[0], vertices[i].pos[1], vertices[i].pos[2]); }

    // CalculateClosestKinematic (SoftBodySharedSettings.cpp:109-187)
    void CalculateClosestKinematic()
    {
        if (!closest.empty())
            return;
        const uint32_t nv = (uint32_t)vertices.size();
        closest.assign(nv, Closest());
        std::vector<std::vector<uint32_t>> conn(nv);
        for (const EdgeRec &e : edges)
        {
            conn[e.v[0]].push_back(e.v[1]);
            conn[e.v[1]].push_back(e.v[0]);
        }
        struct Open { uint32_t vertex; float distance; };
        auto less = [](const Open &a, const Open &b) { return a.distance > b.distance; };
        std::vector<Open> to_visit;
        for (uint32_t v = 0; v < nv; ++v)
            if (vertices[v].inv_mass == 0.0f)
            {
                closest[v].vertex = v;
                closest[v].hops = 0;
                closest[v].distance = 0.0f;
                to_visit.push_back({ v, 0.0f });
                HeapPush(to_visit.begin(), to_visit.end(), less);
            }
        while (!to_visit.empty())
        {
            HeapPop(to_visit.begin(), to_visit.end(), less);
            Open current = to_visit.back();
            to_visit.pop_back();
            for (uint32_t v : conn[current.vertex])
            {
                float nd = current.distance + len(X(v) - X(current.vertex));
                if (nd < closest[v].distance)
                {
                    closest[v].vertex = closest[current.vertex].vertex;
                    closest[v].hops = closest[current.vertex].hops + 1;
                    closest[v].distance = nd;
                    to_visit.push_back({ v, nd });
                    HeapPush(to_visit.begin(), to_visit.end(), less);
                }
            }
        }
    }

    // CalculateBendConstraintConstants (SoftBodySharedSettings.cpp:493-534)
    void CalculateBendConstants()
    {
        for (BendRec &b : bends)
        {
            V3 x0 = X(b.v[0]), x1 = X(b.v[1]), x2 = X(b.v[2]), x3 = X(b.v[3]);
            V3 e0 = x1 - x0, e1 = x2 - x0, e2 = x3 - x0;
            V3 n1 = cross(e0, e1);
            V3 n2 = cross(e2, e0);
            float denom = sqrtf(len_sq(n1) * len_sq(n2));
            if (denom < 1.0e-12f)
                b.initial_angle = 0.0f;
            else
            {
                float sign = sign_of(dot(cross(n2, n1), e0));
                b.initial_angle = sign * acos_approx(dot(n1, n2) / denom);
            }
        }
    }

    // CalculateEdgeLengths (:371-379), CalculateVolumeConstraintVolumes (:536-555), CalculateLRALengths (:483-491)
    void CalculateEdgeLengths()
    {
        for (EdgeRec &e : edges)
            e.rest = len(X(e.v[1]) - X(e.v[0]));
    }
    void CalculateVolumes()
    {
        for (VolumeRec &t : volumes)
        {
            V3 x1 = X(t.v[0]), x2 = X(t.v[1]), x3 = X(t.v[2]), x4 = X(t.v[3]);
            V3 x1x2 = x2 - x1, x1x3 = x3 - x1, x1x4 = x4 - x1;
            t.six_rest_volume = fabsf(dot(cross(x1x2, x1x3), x1x4));
        }
    }
    void CalculateLRALengths(float mult)
    {
        for (LRARec &l : lras)
            l.max_distance = mult * len(X(l.v[1]) - X(l.v[0]));
    }

    // CreateConstraints (SoftBodySharedSettings.cpp:189-369)
    void CreateConstraints(const Attr *attrs, uint32_t na, int bend_type, float angle_tolerance)
    {
        struct EdgeHelper { uint32_t v[2]; uint32_t edge_idx; };
        std::vector<EdgeHelper> eh;
        eh.reserve(faces.size() * 3);
        for (uint32_t fi = 0; fi < faces.size(); ++fi)
            for (int i = 0; i < 3; ++i)
            {
                uint32_t a = faces[fi].v[i], b = faces[fi].v[(i + 1) % 3];
                eh.push_back({ { std::min(a, b), std::max(a, b) }, fi * 3 + (uint32_t)i });
            }
        QuickSortSteps(eh.begin(), eh.end(), [](const EdgeHelper &l, const EdgeHelper &r) {
            return l.v[0] < r.v[0] || (l.v[0] == r.v[0] && l.v[1] < r.v[1]);
        });
        auto attr = [attrs, na](uint32_t v) -> const Attr & { return attrs[std::min(v, na - 1)]; };
        auto add_edge = [this](uint32_t a, uint32_t b, float c1, float c2) {
            if ((vertices[a].inv_mass > 0.0f || vertices[b].inv_mass > 0.0f) && c1 < FLT_MAX && c2 < FLT_MAX)
            {
                EdgeRec e;
                e.v[0] = a;
                e.v[1] = b;
                e.compliance = 0.5f * (c1 + c2);
                e.rest = len(X(b) - X(a));
                edges.push_back(e);
            }
        };
        float s, c;
        SinCos(angle_tolerance, s, c);
        const float sq_sin_tolerance = Square(s);
        const float sq_cos_tolerance = Square(c);
        edges.clear();
        for (size_t i = 0; i < eh.size(); ++i)
        {
            const EdgeHelper e0 = eh[i];
            const Attr &a0 = attr(e0.v[0]);
            const Attr &a1 = attr(e0.v[1]);
            bool is_shear = false;
            for (size_t j = i + 1; j < eh.size(); ++j)
            {
                const EdgeHelper &e1 = eh[j];
                if (e0.v[0] == e1.v[0] && e0.v[1] == e1.v[1])
                {
                    const FaceRec &f0 = faces[e0.edge_idx / 3];
                    const FaceRec &f1 = faces[e1.edge_idx / 3];
                    uint32_t vo0 = f0.v[(e0.edge_idx + 2) % 3];
                    uint32_t vo1 = f1.v[(e1.edge_idx + 2) % 3];
                    const Attr &ao0 = attr(vo0);
                    const Attr &ao1 = attr(vo1);
                    if (vo0 == vo1)
                        continue;
                    V3 n0 = cross(X(f0.v[2]) - X(f0.v[0]), X(f0.v[1]) - X(f0.v[0]));
                    V3 n1 = cross(X(f1.v[2]) - X(f1.v[0]), X(f1.v[1]) - X(f1.v[0]));
                    float n0_dot_n1 = dot(n0, n1);
                    if (n0_dot_n1 > 0.0f && Square(n0_dot_n1) > (sq_cos_tolerance * len_sq(n0)) * len_sq(n1))
                    {
                        V3 e0_dir = X(vo0) - X(e0.v[0]);
                        V3 e1_dir = X(vo1) - X(e0.v[0]);
                        if (Square(dot(e0_dir, e1_dir)) < (sq_sin_tolerance * len_sq(e0_dir)) * len_sq(e1_dir))
                        {
                            add_edge(vo0, vo1, ao0.shear, ao1.shear);
                            is_shear = true;
                        }
                    }
                    if (bend_type == 1)
                    {
                        if (!is_shear)
                            add_edge(vo0, vo1, a0.bend, a1.bend);
                    }
                    else if (bend_type == 2)
                    {
                        if ((vertices[vo0].inv_mass > 0.0f || vertices[vo1].inv_mass > 0.0f) && a0.bend < FLT_MAX && a1.bend < FLT_MAX)
                        {
                            BendRec b;
                            b.v[0] = e0.v[0]; b.v[1] = e0.v[1]; b.v[2] = vo0; b.v[3] = vo1;
                            b.compliance = 0.5f * (a0.bend + a1.bend);
                            b.initial_angle = 0.0f;
                            bends.push_back(b);
                        }
                    }
                }
                else
                {
                    i = j - 1;
                    break;
                }
            }
            add_edge(e0.v[0], e0.v[1], is_shear ? a0.shear : a0.compliance, is_shear ? a1.shear : a1.compliance);
        }
        CalculateBendConstants();

        bool has_lra = false;
        for (uint32_t k = 0; k < na; ++k)
            if (attrs[k].lra_type != 0)
            {
                has_lra = true;
                break;
            }
        if (has_lra)
        {
            CalculateClosestKinematic();
            for (uint32_t v = 0; v < (uint32_t)vertices.size(); ++v)
                if (vertices[v].inv_mass > 0.0f)
                {
                    uint32_t cv = closest[v].vertex;
                    if (cv != 0xffffffffu)
                    {
                        const Attr &va = attr(v);
                        if (va.lra_type == 1)
                            lras.push_back({ { cv, v }, va.lra_mult * len(X(cv) - X(v)) });
                        else if (va.lra_type == 2)
                            lras.push_back({ { cv, v }, va.lra_mult * closest[v].distance });
                    }
                }
        }
    }

    // Optimize (SoftBodySharedSettings.cpp:613-1182) without rods / skinning
    void Optimize()
    {
        constexpr uint32_t kBatch = 256;     // SoftBodyUpdateContext::cVertexConstraintBatch
        const uint32_t nv = (uint32_t)vertices.size();
        groups.clear();
        struct Connection { uint32_t vertex; uint32_t count; };
        std::vector<std::vector<Connection>> conn(nv);
        auto add_connection = [&conn](uint32_t a, uint32_t b) {
            for (int i = 0; i < 2; ++i)
            {
                bool found = false;
                for (Connection &c : conn[a])
                    if (c.vertex == b)
                    {
                        c.count++;
                        found = true;
                        break;
                    }
                if (!found)
                    conn[a].push_back({ b, 1 });
                std::swap(a, b);
            }
        };
        for (const EdgeRec &c : edges) add_connection(c.v[0], c.v[1]);
        for (const LRARec &c : lras) add_connection(c.v[0], c.v[1]);
        for (const BendRec &c : bends)
        {
            add_connection(c.v[0], c.v[1]); add_connection(c.v[0], c.v[2]); add_connection(c.v[0], c.v[3]);
            add_connection(c.v[1], c.v[2]); add_connection(c.v[1], c.v[3]); add_connection(c.v[2], c.v[3]);
        }
        for (const VolumeRec &c : volumes)
        {
            add_connection(c.v[0], c.v[1]); add_connection(c.v[0], c.v[2]); add_connection(c.v[0], c.v[3]);
            add_connection(c.v[1], c.v[2]); add_connection(c.v[1], c.v[3]); add_connection(c.v[2], c.v[3]);
        }

        // Greedy grouping, the reference's loop (SoftBodySharedSettings.cpp:672-774) with
        // the same choices made without its quadratic rescans:
        // * the bounds and the start vertex of the ungrouped vertices come from sorted
        //   orders (per axis, and per axis priority with index as the last key) read
        //   from lazily advanced cursors that skip grouped vertices;
        // * the vertex most connected to the current group is kept incrementally: each
        //   ungrouped neighbour of the group carries its connection count to the group,
        //   its distance to the group's first vertex and its first position in the
        //   reference's scan (group member index, connection index); the reference
        //   keeps the first of equal (count, distance) pairs in that scan, so the
        //   choice is the maximum of (count, -distance, -first position).
        std::vector<int> group_idx(nv, -1);
        int current_group_idx = 0;
        std::vector<uint32_t> current_group;

        std::vector<uint32_t> by_axis[3];
        for (int k = 0; k < 3; ++k)
        {
            by_axis[k].resize(nv);
            for (uint32_t i = 0; i < nv; ++i) by_axis[k][i] = i;
            std::sort(by_axis[k].begin(), by_axis[k].end(), [&](uint32_t a, uint32_t b) { return vertices[a].pos[k] < vertices[b].pos[k]; });
        }
        size_t lo_cursor[3] = { 0, 0, 0 }, hi_cursor[3] = { nv, nv, nv };
        std::vector<uint32_t> by_priority[9];       // index 3 * max_axis + min_axis
        size_t priority_cursor[9] = {};
        auto priority_order = [&](uint32_t a0, uint32_t a1, uint32_t a2) -> std::vector<uint32_t> & {
            std::vector<uint32_t> &o = by_priority[3 * a0 + a2];
            if (o.empty() && nv)
            {
                o.resize(nv);
                for (uint32_t i = 0; i < nv; ++i) o[i] = i;
                std::sort(o.begin(), o.end(), [&](uint32_t a, uint32_t b) {
                    const float *p = vertices[a].pos, *q = vertices[b].pos;
                    if (p[a0] < q[a0]) return true;
                    if (!(p[a0] == q[a0])) return false;
                    if (p[a1] < q[a1]) return true;
                    if (!(p[a1] == q[a1])) return false;
                    if (p[a2] < q[a2]) return true;
                    if (!(p[a2] == q[a2])) return false;
                    return a < b;
                });
            }
            return o;
        };

        // candidates of the current group
        std::vector<uint32_t> cand_num(nv, 0);
        std::vector<uint64_t> cand_first(nv, 0);
        std::vector<float> cand_dist(nv, 0.0f);
        std::vector<char> is_cand(nv, 0);
        std::vector<uint32_t> cand_list;
        auto join = [&](uint32_t u) {
            const uint32_t member = (uint32_t)current_group.size();
            current_group.push_back(u);
            group_idx[u] = current_group_idx;
            is_cand[u] = 0;
            const V3 front = X(current_group.front());
            const std::vector<Connection> &cu = conn[u];
            for (uint32_t k = 0; k < (uint32_t)cu.size(); ++k)
            {
                const uint32_t v = cu[k].vertex;
                if (group_idx[v] != -1)
                    continue;
                if (!is_cand[v])
                {
                    is_cand[v] = 1;
                    cand_num[v] = 0;
                    cand_first[v] = ((uint64_t)member << 32) | k;
                    cand_dist[v] = len_sq(X(v) - front);
                    cand_list.push_back(v);
                }
                cand_num[v] += cu[k].count;
            }
        };
        auto clear_candidates = [&]() {
            for (uint32_t v : cand_list) is_cand[v] = 0;
            cand_list.clear();
        };

        for (;;)
        {
            // bounds of the ungrouped vertices (AABox::Encapsulate)
            bool any = false;
            V3 bmin, bmax;
            for (int k = 0; k < 3; ++k)
            {
                while (lo_cursor[k] < nv && group_idx[by_axis[k][lo_cursor[k]]] != -1) ++lo_cursor[k];
                while (hi_cursor[k] > lo_cursor[k] && group_idx[by_axis[k][hi_cursor[k] - 1]] != -1) --hi_cursor[k];
                any = lo_cursor[k] < hi_cursor[k];
                if (!any)
                    break;
                (&bmin.x)[k] = vertices[by_axis[k][lo_cursor[k]]].pos[k];
                (&bmax.x)[k] = vertices[by_axis[k][hi_cursor[k] - 1]].pos[k];
            }
            if (!any)
                break;
            V3 size = bmax - bmin;
            uint32_t max_axis = size.x > size.y ? (size.z > size.x ? 2 : 0) : (size.z > size.y ? 2 : 1);
            uint32_t min_axis = (uint32_t)lowest_index(size);
            if (min_axis == max_axis)
                min_axis = (min_axis + 1) % 3;
            uint32_t mid_axis = 3 - min_axis - max_axis;

            // the ungrouped vertex lowest on (max, mid, min axis), the lowest index among equals
            std::vector<uint32_t> &order = priority_order(max_axis, mid_axis, min_axis);
            size_t &cursor = priority_cursor[3 * max_axis + min_axis];
            while (cursor < order.size() && group_idx[order[cursor]] != -1) ++cursor;
            if (cursor == order.size())
                break;
            join(order[cursor]);

            for (;;)
            {
                uint32_t best_vertex = UINT_MAX;
                size_t live = 0;
                for (uint32_t v : cand_list)
                {
                    if (!is_cand[v])
                        continue;
                    cand_list[live++] = v;
                    if (best_vertex == UINT_MAX || cand_num[v] > cand_num[best_vertex]
                        || (cand_num[v] == cand_num[best_vertex]
                            && (cand_dist[v] < cand_dist[best_vertex]
                                || (cand_dist[v] == cand_dist[best_vertex] && cand_first[v] < cand_first[best_vertex]))))
                        best_vertex = v;
                }
                cand_list.resize(live);
                if (best_vertex != UINT_MAX)
                    join(best_vertex);
                if (current_group.size() >= kBatch || (current_group.size() > kBatch / 2 && best_vertex == UINT_MAX))
                {
                    current_group.clear();
                    clear_candidates();
                    current_group_idx++;
                    break;
                }
                if (best_vertex == UINT_MAX)
                    break;
            }
        }
        if (current_group.size() > kBatch / 2)
            ++current_group_idx;
        current_group.clear();
        conn.clear();

        struct Group
        {
            std::vector<uint32_t> e, l, d, t;
            uint32_t Size() const { return (uint32_t)(e.size() + l.size() + d.size() + t.size()); }
        };
        std::vector<Group> gr(current_group_idx + 1);
        for (uint32_t k = 0; k < edges.size(); ++k)
        {
            int g1 = group_idx[edges[k].v[0]], g2 = group_idx[edges[k].v[1]];
            (g1 == g2 ? gr[g1] : gr.back()).e.push_back(k);
        }
        for (uint32_t k = 0; k < lras.size(); ++k)
        {
            int g1 = group_idx[lras[k].v[0]], g2 = group_idx[lras[k].v[1]];
            (g1 == g2 ? gr[g1] : gr.back()).l.push_back(k);
        }
        for (uint32_t k = 0; k < bends.size(); ++k)
        {
            const uint32_t *v = bends[k].v;
            int g1 = group_idx[v[0]], g2 = group_idx[v[1]], g3 = group_idx[v[2]], g4 = group_idx[v[3]];
            (g1 == g2 && g1 == g3 && g1 == g4 ? gr[g1] : gr.back()).d.push_back(k);
        }
        for (uint32_t k = 0; k < volumes.size(); ++k)
        {
            const uint32_t *v = volumes[k].v;
            int g1 = group_idx[v[0]], g2 = group_idx[v[1]], g3 = group_idx[v[2]], g4 = group_idx[v[3]];
            (g1 == g2 && g1 == g3 && g1 == g4 ? gr[g1] : gr.back()).t.push_back(k);
        }

        // Parallel groups big to small (ties in the reference's quicksort order)
        QuickSortSteps(gr.begin(), gr.end() - 1, [](const Group &a, const Group &b) { return a.Size() > b.Size(); });

        CalculateClosestKinematic();

        for (Group &g : gr)
        {
            QuickSortSteps(g.e.begin(), g.e.end(), [this](uint32_t l, uint32_t r) {
                const EdgeRec &e1 = edges[l], &e2 = edges[r];
                float d1 = MinF(closest[e1.v[0]].distance, closest[e1.v[1]].distance);
                float d2 = MinF(closest[e2.v[0]].distance, closest[e2.v[1]].distance);
                if (d1 != d2) return d1 < d2;
                uint32_t m1 = std::min(e1.v[0], e1.v[1]), m2 = std::min(e2.v[0], e2.v[1]);
                if (m1 != m2) return m1 < m2;
                return l < r;
            });
            QuickSortSteps(g.l.begin(), g.l.end(), [this](uint32_t l, uint32_t r) {
                const LRARec &a = lras[l], &b = lras[r];
                if (a.max_distance != b.max_distance) return a.max_distance > b.max_distance;
                uint32_t m1 = std::min(a.v[0], a.v[1]), m2 = std::min(b.v[0], b.v[1]);
                if (m1 != m2) return m1 < m2;
                return l < r;
            });
            auto min4 = [](const uint32_t *v) { return std::min(std::min(v[0], v[1]), std::min(v[2], v[3])); };
            auto dmin4 = [this](const uint32_t *v) {
                return MinF(MinF(closest[v[0]].distance, closest[v[1]].distance), MinF(closest[v[2]].distance, closest[v[3]].distance));
            };
            QuickSortSteps(g.d.begin(), g.d.end(), [&](uint32_t l, uint32_t r) {
                float d1 = dmin4(bends[l].v), d2 = dmin4(bends[r].v);
                if (d1 != d2) return d1 < d2;
                uint32_t m1 = min4(bends[l].v), m2 = min4(bends[r].v);
                if (m1 != m2) return m1 < m2;
                return l < r;
            });
            QuickSortSteps(g.t.begin(), g.t.end(), [&](uint32_t l, uint32_t r) {
                float d1 = dmin4(volumes[l].v), d2 = dmin4(volumes[r].v);
                if (d1 != d2) return d1 < d2;
                uint32_t m1 = min4(volumes[l].v), m2 = min4(volumes[r].v);
                if (m1 != m2) return m1 < m2;
                return l < r;
            });
        }

        std::vector<EdgeRec> te;
        te.swap(edges);
        std::vector<LRARec> tl;
        tl.swap(lras);
        std::vector<BendRec> td;
        td.swap(bends);
        std::vector<VolumeRec> tt;
        tt.swap(volumes);
        for (const Group &g : gr)
        {
            for (uint32_t k : g.e) edges.push_back(te[k]);
            for (uint32_t k : g.l) lras.push_back(tl[k]);
            for (uint32_t k : g.d) bends.push_back(td[k]);
            for (uint32_t k : g.t) volumes.push_back(tt[k]);
            groups.push_back({ (uint32_t)edges.size(), (uint32_t)lras.size(), 0u, 0u, (uint32_t)bends.size(), (uint32_t)volumes.size(), 0u });
        }
        closest.clear();
    }
};

} // namespace

struct xpbd_built_settings
{
    Builder b;
};

extern "C" {

const char *xpbd_build_last_error(void) { return g_build_error.c_str(); }

int xpbd_settings_build(const xpbd_build_desc *d, xpbd_built_settings **out)
{
    if (!d || !out)
        return BuildFail(XPBD_E_INVALID, "null argument");
    *out = nullptr;
    if (d->num_vertices == 0 || !d->vertices)
        return BuildFail(XPBD_E_INVALID, "no vertices");
    std::unique_ptr<xpbd_built_settings> s(new xpbd_built_settings);
    Builder &b = s->b;
    const uint32_t nv = d->num_vertices;
    const VertexRec *vr = (const VertexRec *)d->vertices;
    b.vertices.assign(vr, vr + nv);
    for (uint32_t f = 0; f < d->num_faces; ++f)
    {
        const uint32_t *fv = d->faces + 3ull * f;
        if (fv[0] >= nv || fv[1] >= nv || fv[2] >= nv)
            return BuildFail(XPBD_E_INVALID, "face vertex out of range");
        if (fv[0] == fv[1] || fv[0] == fv[2] || fv[1] == fv[2])
            return BuildFail(XPBD_E_INVALID, "degenerate face");
        b.faces.push_back({ { fv[0], fv[1], fv[2] }, 0u });
    }
    const EdgeRec *er = (const EdgeRec *)d->edges;
    for (uint32_t k = 0; k < d->num_edges; ++k)
    {
        if (er[k].v[0] >= nv || er[k].v[1] >= nv)
            return BuildFail(XPBD_E_INVALID, "edge vertex out of range");
        b.edges.push_back(er[k]);
    }
    const VolumeRec *tr = (const VolumeRec *)d->volumes;
    for (uint32_t k = 0; k < d->num_volumes; ++k)
    {
        for (int j = 0; j < 4; ++j)
            if (tr[k].v[j] >= nv)
                return BuildFail(XPBD_E_INVALID, "volume vertex out of range");
        b.volumes.push_back(tr[k]);
    }
    const LRARec *lr = (const LRARec *)d->lras;
    if (d->num_lras && !lr)
        return BuildFail(XPBD_E_INVALID, "null LRA array");
    for (uint32_t k = 0; k < d->num_lras; ++k)
    {
        if (lr[k].v[0] >= nv || lr[k].v[1] >= nv || lr[k].v[0] == lr[k].v[1])
            return BuildFail(XPBD_E_INVALID, "LRA vertices out of range or equal");
        b.lras.push_back(lr[k]);
    }
    if (d->flags & XPBD_BUILD_CREATE_CONSTRAINTS)
    {
        if (d->num_attributes == 0 || !d->attributes)
            return BuildFail(XPBD_E_INVALID, "CreateConstraints needs at least one vertex attribute");
        if (d->bend_type < 0 || d->bend_type > 2)
            return BuildFail(XPBD_E_INVALID, "bend type must be 0 (none), 1 (distance) or 2 (dihedral)");
        std::vector<Attr> attrs(d->num_attributes);
        for (uint32_t k = 0; k < d->num_attributes; ++k)
        {
            const xpbd_vertex_attributes &a = d->attributes[k];
            attrs[k] = { a.compliance, a.shear_compliance, a.bend_compliance, a.lra_type, a.lra_max_distance_multiplier };
        }
        b.CreateConstraints(attrs.data(), d->num_attributes, d->bend_type, d->angle_tolerance);
    }
    if (d->flags & XPBD_BUILD_EDGE_LENGTHS) b.CalculateEdgeLengths();
    if (d->flags & XPBD_BUILD_VOLUMES) b.CalculateVolumes();
    if (d->flags & XPBD_BUILD_LRA_LENGTHS) b.CalculateLRALengths(d->lra_max_distance_multiplier);
    if (d->flags & XPBD_BUILD_BEND_CONSTANTS) b.CalculateBendConstants();
    if (d->flags & XPBD_BUILD_OPTIMIZE) b.Optimize();
    *out = s.release();
    return XPBD_OK;
}

int xpbd_built_counts(const xpbd_built_settings *s, uint32_t counts[7])
{
    if (!s || !counts)
        return BuildFail(XPBD_E_INVALID, "null argument");
    const Builder &b = s->b;
    counts[0] = (uint32_t)b.vertices.size(); counts[1] = (uint32_t)b.faces.size(); counts[2] = (uint32_t)b.edges.size();
    counts[3] = (uint32_t)b.bends.size(); counts[4] = (uint32_t)b.volumes.size(); counts[5] = (uint32_t)b.lras.size();
    counts[6] = (uint32_t)b.groups.size();
    return XPBD_OK;
}

int xpbd_built_read(const xpbd_built_settings *s, void *vertices, void *faces, void *edges, void *bends, void *volumes, void *lras, void *groups)
{
    if (!s)
        return BuildFail(XPBD_E_INVALID, "null argument");
    const Builder &b = s->b;
    if (vertices && !b.vertices.empty()) memcpy(vertices, b.vertices.data(), sizeof(VertexRec) * b.vertices.size());
    if (faces && !b.faces.empty()) memcpy(faces, b.faces.data(), sizeof(FaceRec) * b.faces.size());
    if (edges && !b.edges.empty()) memcpy(edges, b.edges.data(), sizeof(EdgeRec) * b.edges.size());
    if (bends && !b.bends.empty()) memcpy(bends, b.bends.data(), sizeof(BendRec) * b.bends.size());
    if (volumes && !b.volumes.empty()) memcpy(volumes, b.volumes.data(), sizeof(VolumeRec) * b.volumes.size());
    if (lras && !b.lras.empty()) memcpy(lras, b.lras.data(), sizeof(LRARec) * b.lras.size());
    if (groups && !b.groups.empty()) memcpy(groups, b.groups.data(), sizeof(GroupRec) * b.groups.size());
    return XPBD_OK;
}

int xpbd_built_free(xpbd_built_settings *s)
{
    delete s;
    return XPBD_OK;
}

} // extern "C"
