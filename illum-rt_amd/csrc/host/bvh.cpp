#include "bvh.h"

#include <algorithm>
#include <cfloat>
#include <utility>

namespace rt {

namespace {

// A primitive is a box (a triangle's box, or with ESC one fragment's box) and
// the triangle it belongs to (rt/cpu/bvh-ctor.h:24-29).
struct prim {
    aabb box;
    uint32_t tri;
    float center(int axis) const { return (comp(box.min, axis) + comp(box.max, axis)) * 0.5f; }
};

int largest_axis(const vec3f &e) {
    float largest = std::max(e.x, std::max(e.y, e.z));
    return largest == e.x ? 0 : (largest == e.y ? 1 : 2);
}

class builder {
public:
    builder(const bvh_params &p, std::vector<prim> &prims, std::vector<uint32_t> &index, bvh &out)
        : p_(p), prims_(prims), index_(index), out_(out) {}

    // Iterative form of the reference's recursive subdivide: a node id is
    // allocated when its range is first visited and the left range is always
    // finished before the right one, so ids come out in the same pre-order.
    void run() {
        struct item {
            uint32_t start, end;
            int32_t parent;
            int side;
        };
        std::vector<item> work;
        work.push_back(item{0, (uint32_t)index_.size(), -1, 0});
        while (!work.empty()) {
            item it = work.back();
            work.pop_back();
            uint32_t id = (uint32_t)out_.nodes.size();
            out_.nodes.emplace_back();
            if (it.parent >= 0) {
                if (it.side == 0) out_.nodes[it.parent].link_l = (int32_t)id;
                else out_.nodes[it.parent].link_r = (int32_t)id;
            }
            uint32_t n = it.end - it.start;
            if (n <= (uint32_t)p_.max_tris) {
                out_.nodes[id].link_l = -(int32_t)it.start;
                out_.nodes[id].link_r = -(int32_t)n;
                out_.nodes[id].box_l = empty_box();
                out_.nodes[id].box_r = empty_box();
                continue;
            }
            aabb bl = empty_box(), br = empty_box();
            uint32_t mid = split(it.start, it.end, bl, br);
            out_.nodes[id].box_l = bl;
            out_.nodes[id].box_r = br;
            out_.nodes[id].link_l = 0;
            out_.nodes[id].link_r = 0;
            work.push_back(item{mid, it.end, (int32_t)id, 1});
            work.push_back(item{it.start, mid, (int32_t)id, 0});
        }
    }

private:
    const bvh_params &p_;
    std::vector<prim> &prims_;
    std::vector<uint32_t> &index_;
    bvh &out_;

    aabb range_box(uint32_t b, uint32_t e) const {
        aabb box = empty_box();
        for (uint32_t i = b; i < e; ++i) grow(box, prims_[index_[i]].box);
        return box;
    }
    void sort_by_center(uint32_t start, uint32_t end, int axis) {
        std::sort(index_.data() + start, index_.data() + end,
                  [&](uint32_t a, uint32_t b) { return prims_[a].center(axis) < prims_[b].center(axis); });
    }
    // The in-place partition of rt/cpu/bvh-ctor.h:331-343 and :427-440: two
    // cursors close in from both ends; the count of elements stepped over by
    // the left cursor is the split position.  The slot where the cursors meet
    // is never counted, so "all but the last element on the left" reports
    // mid == end-1, which the callers treat as degenerate.
    uint32_t partition(uint32_t start, uint32_t end, int axis, float plane) {
        uint32_t mid = start;
        uint32_t *l = index_.data() + start;
        uint32_t *r = index_.data() + end - 1;
        auto c = [&](const uint32_t *x) { return prims_[*x].center(axis); };
        while (l < r) {
            while (c(l) <= plane && l < r) {
                ++l;
                ++mid;
            }
            while (c(r) > plane && l < r) --r;
            if (c(l) > c(r) && l < r) std::swap(*l, *r);
        }
        return mid;
    }

    uint32_t split(uint32_t start, uint32_t end, aabb &box_l, aabb &box_r) {
        switch (p_.split) {
        case split_kind::om: return split_om(start, end, box_l, box_r);
        case split_kind::sm: return split_sm(start, end, box_l, box_r);
        default: return split_sah(start, end, box_l, box_r);
        }
    }

    // rt/cpu/bvh-ctor.h:251-300
    uint32_t split_om(uint32_t start, uint32_t end, aabb &box_l, aabb &box_r) {
        aabb box = range_box(start, end);
        sort_by_center(start, end, largest_axis(box.max - box.min));
        uint32_t mid = start + (end - start) / 2;
        box_l = range_box(start, mid);
        box_r = range_box(mid, end);
        return mid;
    }

    // rt/cpu/bvh-ctor.h:302-369
    uint32_t split_sm(uint32_t start, uint32_t end, aabb &box_l, aabb &box_r) {
        aabb box = range_box(start, end);
        int axis = largest_axis(box.max - box.min);
        float median = comp(box.min + (box.max - box.min) * 0.5f, axis);
        uint32_t mid = partition(start, end, axis, median);
        if (mid == start || mid == end - 1) {
            sort_by_center(start, end, axis);
            mid = start + (end - start) / 2;
        }
        box_l = range_box(start, mid);
        box_r = range_box(mid, end);
        return mid;
    }

    // rt/cpu/bvh-ctor.h:371-503.  Only the largest axis is swept, with
    // P = min(planes, n) planes at min + (i+1)*extent/(P+1), in increasing
    // order, each one re-partitioning the same range.  The first degenerate
    // partition falls back to an object-median sort; later degenerate ones are
    // skipped.  Strict "<" keeps the earliest of equally good candidates.
    uint32_t split_sah(uint32_t start, uint32_t end, aabb &box_l, aabb &box_r) {
        aabb box = range_box(start, end);
        float box_surf = box_area(box);
        vec3f extent = box.max - box.min;
        int axis = largest_axis(extent);
        uint32_t n = end - start;
        if (box_surf == 0) {
            sort_by_center(start, end, axis);
            uint32_t mid = start + n / 2;
            box_l = range_box(start, mid);
            box_r = range_box(mid, end);
            return mid;
        }
        float cost_l = FLT_MAX, cost_r = FLT_MAX;
        bool use_om = true;
        uint32_t mid = start;
        int planes = p_.planes;
        if ((int)n < planes) planes = (int)n;
        float ext_a = comp(extent, axis), min_a = comp(box.min, axis);
        for (int i = 0; i < planes; ++i) {
            float plane = min_a + (float)(i + 1) * (ext_a / (float)(planes + 1));
            uint32_t cur = partition(start, end, axis, plane);
            if (cur == start || cur == end - 1) {
                if (!use_om) continue;
                sort_by_center(start, end, axis);
                cur = start + n / 2;
                use_om = false;
            }
            aabb cl = range_box(start, cur), cr = range_box(cur, end);
            float c_l = (box_area(cl) / box_surf) * (float)(cur - start);
            float c_r = (box_area(cr) / box_surf) * (float)(end - cur);
            if (c_l + c_r < cost_l + cost_r) {
                box_l = cl;
                box_r = cr;
                cost_l = c_l;
                cost_r = c_r;
                mid = cur;
            }
        }
        // The reference's SAH leaf-termination test (bvh-ctor.h:479-487) also
        // requires n <= max_tris, which cannot hold here, so it never fires.
        return mid;
    }
};

float prim_area(const aabb &b) { return box_area(b); }

// rt/cpu/bvh-ctor.h:181-226: the threshold is the box area at rank 9N/10 of
// the ascending order; every triangle is split until its fragments' boxes are
// no larger; fragment 0 replaces the triangle's primitive, the others are
// appended and get fresh index slots.
void early_split_clipping(const scene &s, std::vector<prim> &prims, std::vector<uint32_t> &index) {
    std::vector<float> areas(prims.size());
    for (size_t i = 0; i < prims.size(); ++i) areas[i] = prim_area(prims[i].box);
    size_t rank = 9 * areas.size() / 10;
    std::nth_element(areas.begin(), areas.begin() + rank, areas.end());
    float thres = areas[rank];
    // With most boxes flat (zero area) the threshold is 0 and no fragment of
    // a triangle with area could ever get below it: the reference recurses
    // without end there.  Such scenes keep whole-triangle boxes.
    if (!(thres > 0.0f)) return;
    uint32_t n = (uint32_t)prims.size();
    std::vector<vec3f> poly(3);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t t = index[i];
        poly[0] = s.a(t);
        poly[1] = s.b(t);
        poly[2] = s.c(t);
        std::vector<aabb> frags = esc_split(poly, thres);
        prims[i] = prim{frags[0], i};
        for (size_t j = 1; j < frags.size(); ++j) {
            prims.push_back(prim{frags[j], i});
            index.push_back((uint32_t)prims.size() - 1);
        }
    }
}

} // namespace

namespace {

bool same_box(const aabb &x, const aabb &y) {
    return x.min.x == y.min.x && x.min.y == y.min.y && x.min.z == y.min.z && x.max.x == y.max.x &&
           x.max.y == y.max.y && x.max.z == y.max.z;
}

aabb poly_box(const std::vector<vec3f> &poly) {
    aabb b = empty_box();
    for (const vec3f &v : poly) grow(b, v);
    return b;
}

// Splits one polygon; `budget` counts the splits this triangle may still
// make.  Two guards end recursions the reference would not end: a split
// whose halves are not smaller than the polygon (rounding at large
// coordinates) and a triangle that would need more than kEscMaxSplits
// splits (a threshold far below its area).  The workloads' triangles need
// at most 5 splits, so neither guard changes their fragments.
constexpr int kEscMaxSplits = 1023;

void esc_rec(const std::vector<vec3f> &poly, float threshold, int &budget, std::vector<aabb> &out) {
    aabb b = poly_box(poly);
    if (box_area(b) <= threshold || budget <= 0) {
        out.push_back(b);
        return;
    }
    int axis = largest_axis(b.max - b.min);
    float center = (comp(b.max, axis) + comp(b.min, axis)) * 0.5f;

    // Find the two polygon edges that cross the plane.  The scan does not wrap
    // around, so a crossing on the closing edge is reported as the last vertex.
    int n = (int)poly.size();
    bool last_left = comp(poly[0], axis) < center;
    int cross_first = -1, cross_second = -1;
    for (int i = 0; i < n; ++i) {
        bool left = comp(poly[i], axis) < center;
        if (left != last_left) {
            if (cross_first == -1) cross_first = i - 1;
            else cross_second = i - 1;
        }
        last_left = left;
    }
    if (cross_second == -1) cross_second = n - 1;
    // The reference asserts here; it can only happen when rounding puts the
    // box centre onto the box minimum, and the unsplit box is the safe answer.
    if (cross_first == -1) {
        out.push_back(b);
        return;
    }

    auto crossing = [&](int a) {
        float delta = center - comp(poly[a], axis);
        int nb = a + 1 < n ? a + 1 : 0;
        return poly[a] + (poly[nb] - poly[a]) * (delta / (comp(poly[nb], axis) - comp(poly[a], axis)));
    };
    vec3f c1 = crossing(cross_first), c2 = crossing(cross_second);
    std::vector<vec3f> p1, p2;
    for (int i = 0; i <= cross_first; ++i) p1.push_back(poly[i]);
    p1.push_back(c1);
    p2.push_back(c1);
    for (int i = cross_first + 1; i <= cross_second; ++i) p2.push_back(poly[i]);
    p2.push_back(c2);
    p1.push_back(c2);
    for (int i = cross_second + 1; i < n; ++i) p1.push_back(poly[i]);
    p1.erase(std::unique(p1.begin(), p1.end()), p1.end());
    p2.erase(std::unique(p2.begin(), p2.end()), p2.end());

    if (same_box(poly_box(p1), b) || same_box(poly_box(p2), b)) {
        out.push_back(b);
        return;
    }
    --budget;
    esc_rec(p1, threshold, budget, out);
    esc_rec(p2, threshold, budget, out);
}

} // namespace

std::vector<aabb> esc_split(const std::vector<vec3f> &poly, float threshold) {
    std::vector<aabb> out;
    int budget = kEscMaxSplits;
    esc_rec(poly, threshold, budget, out);
    return out;
}

bvh build_bvh(const scene &s, const bvh_params &p) {
    bvh out;
    uint32_t n = s.n_tris();
    if (n == 0) return out;
    std::vector<prim> prims(n);
    std::vector<uint32_t> index(n);
    for (uint32_t i = 0; i < n; ++i) {
        aabb b = empty_box();
        grow(b, s.a(i));
        grow(b, s.b(i));
        grow(b, s.c(i));
        prims[i] = prim{b, i};
        index[i] = i;
    }
    if (p.esc) early_split_clipping(s, prims, index);
    bvh_params q = p;
    if (q.max_tris < 1) q.max_tris = 1;
    if (q.planes < 1) q.planes = 1;
    builder(q, prims, index, out).run();
    out.root = 0;
    // rt/cpu/bvh-ctor.h:239-248: slots refer to primitives; map them to triangles.
    for (uint32_t &x : index) x = prims[x].tri;
    out.index = std::move(index);
    return out;
}

bvh_stats compute_stats(const bvh &b) {
    bvh_stats st;
    if (b.nodes.empty()) return st;
    std::vector<std::pair<uint32_t, uint32_t>> stack{{b.root, 0}};
    while (!stack.empty()) {
        auto [id, d] = stack.back();
        stack.pop_back();
        const bbvh_node &nd = b.nodes[id];
        if (nd.inner()) {
            ++st.inner;
            st.depth = std::max(st.depth, d + 1);
            stack.push_back({(uint32_t)nd.link_l, d + 1});
            stack.push_back({(uint32_t)nd.link_r, d + 1});
        } else {
            ++st.leaves;
            st.slots += (uint32_t)nd.tri_count();
            st.max_leaf = std::max(st.max_leaf, (uint32_t)nd.tri_count());
        }
    }
    return st;
}

std::string validate_bvh(const scene &s, const bvh &b, int max_tris) {
    uint32_t n = s.n_tris();
    if (n == 0) return b.nodes.empty() ? "" : "nodes for an empty scene";
    if (b.nodes.empty() || b.root >= b.nodes.size()) return "root out of range";
    std::vector<uint8_t> seen_node(b.nodes.size(), 0), seen_tri(n, 0);
    std::vector<uint8_t> seen_slot(b.index.size(), 0);
    struct item {
        uint32_t id;
        aabb bound;   // box the parent assigned to this subtree
    };
    aabb everything{v3(-FLT_MAX), v3(FLT_MAX)};
    std::vector<item> stack{{b.root, everything}};
    uint32_t reached = 0;
    while (!stack.empty()) {
        item it = stack.back();
        stack.pop_back();
        if (it.id >= b.nodes.size()) return "link out of range";
        if (seen_node[it.id]++) return "node reached twice";
        ++reached;
        const bbvh_node &nd = b.nodes[it.id];
        if (nd.inner()) {
            stack.push_back({(uint32_t)nd.link_l, nd.box_l});
            stack.push_back({(uint32_t)nd.link_r, nd.box_r});
            continue;
        }
        int32_t off = nd.tri_offset(), cnt = nd.tri_count();
        if (cnt < 1 || cnt > max_tris) return "leaf size out of range";
        if (off < 0 || (size_t)off + cnt > b.index.size()) return "leaf slots out of range";
        for (int32_t k = 0; k < cnt; ++k) {
            if (seen_slot[off + k]++) return "slot in two leaves";
            uint32_t t = b.index[off + k];
            if (t >= n) return "triangle id out of range";
            seen_tri[t] = 1;
            // The triangle (or, with ESC, at least its fragment) lies in the box.
            aabb tb = empty_box();
            grow(tb, s.a(t));
            grow(tb, s.b(t));
            grow(tb, s.c(t));
            const aabb &bx = it.bound;
            if (tb.max.x < bx.min.x || tb.max.y < bx.min.y || tb.max.z < bx.min.z || tb.min.x > bx.max.x ||
                tb.min.y > bx.max.y || tb.min.z > bx.max.z)
                return "leaf box misses its triangle";
        }
    }
    if (reached != b.nodes.size()) return "unreachable nodes";
    for (uint32_t t = 0; t < n; ++t)
        if (!seen_tri[t]) return "triangle missing from the tree";
    for (size_t k = 0; k < seen_slot.size(); ++k)
        if (!seen_slot[k]) return "slot in no leaf";
    return "";
}

} // namespace rt
