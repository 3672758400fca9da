#include "flatten.h"

#include <cfloat>
#include <cmath>
#include <cstring>
#include <utility>

namespace rt {

namespace {

// Reference into the flattened tree for bbvh node `id`.
struct ref_map {
    const bvh &b;
    const std::vector<int32_t> &inner_id;   // bbvh id -> flattened inner index, -1 for leaves
    int32_t operator()(uint32_t id) const {
        const bbvh_node &n = b.nodes[id];
        if (n.inner()) return inner_id[id];
        return make_leaf_ref((uint32_t)n.tri_offset(), (uint32_t)n.tri_count());
    }
};

aabb child_box(const bbvh_node &n, int side) { return side == 0 ? n.box_l : n.box_r; }

} // namespace

std::string flatten(const scene &s, const bvh &b, int bvh_kind, int tri_kind, flat_scene &out) {
    out = flat_scene{};
    out.bvh_kind = bvh_kind;
    out.tri_kind = tri_kind;
    if (b.nodes.empty()) return "";
    if (b.index.size() > kMaxLeafSlot) return "too many leaf slots for the 26-bit slot field";

    // Triangle records in slot order.
    out.tris.resize(b.index.size());
    out.slot_ids = b.index;
    for (size_t k = 0; k < b.index.size(); ++k) {
        uint32_t t = b.index[k];
        out.tris[k] = tri_kind == TRI_WOOP ? make_woop_rec(s.a(t), s.b(t), s.c(t))
                                           : make_exact_rec(s.a(t), s.b(t), s.c(t), t);
    }
    for (const bbvh_node &n : b.nodes)
        if (!n.inner() && (uint32_t)n.tri_count() > kMaxLeafCount)
            return "leaf holds more than 32 triangles; rebuild with max_tris <= 32";

    const bbvh_node &root = b.nodes[b.root];
    if (bvh_kind == BVH2) {
        // Number the inner nodes in pre-order (the builder's order, minus leaves).
        std::vector<int32_t> inner_id(b.nodes.size(), -1);
        std::vector<uint32_t> order;
        std::vector<std::pair<uint32_t, uint32_t>> stack{{b.root, 1}};
        while (!stack.empty()) {
            auto [id, d] = stack.back();
            stack.pop_back();
            const bbvh_node &n = b.nodes[id];
            if (!n.inner()) continue;
            inner_id[id] = (int32_t)order.size();
            order.push_back(id);
            if (d > out.depth) out.depth = d;
            stack.push_back({(uint32_t)n.link_r, d + 1});
            stack.push_back({(uint32_t)n.link_l, d + 1});
        }
        ref_map ref{b, inner_id};
        out.nodes2.resize(order.size());
        for (size_t k = 0; k < order.size(); ++k) {
            const bbvh_node &n = b.nodes[order[k]];
            node2 &o = out.nodes2[k];
            o = node2{n.box_l.min.x, n.box_l.max.x, n.box_l.min.y, n.box_l.max.y,
                      n.box_r.min.x, n.box_r.max.x, n.box_r.min.y, n.box_r.max.y,
                      n.box_l.min.z, n.box_l.max.z, n.box_r.min.z, n.box_r.max.z,
                      ref((uint32_t)n.link_l), ref((uint32_t)n.link_r), 0, 0};
        }
        out.root_ref = root.inner() ? 0 : make_leaf_ref((uint32_t)root.tri_offset(), (uint32_t)root.tri_count());
        return "";
    }

    // BVH4: each wide node starts from a binary inner node's two children and
    // repeatedly opens the inner child with the largest box surface until it
    // has four children or only leaves remain.
    struct child {
        aabb box;
        uint32_t id;   // bbvh node id
    };
    auto collapse = [&](uint32_t id, std::vector<child> &kids) {
        const bbvh_node &n = b.nodes[id];
        kids = {child{n.box_l, (uint32_t)n.link_l}, child{n.box_r, (uint32_t)n.link_r}};
        while (kids.size() < 4) {
            int best = -1;
            float best_area = -1.0f;
            for (size_t k = 0; k < kids.size(); ++k) {
                if (!b.nodes[kids[k].id].inner()) continue;
                float a = box_area(kids[k].box);
                if (a > best_area) {
                    best_area = a;
                    best = (int)k;
                }
            }
            if (best < 0) break;
            const bbvh_node &c = b.nodes[kids[best].id];
            child l{child_box(c, 0), (uint32_t)c.link_l}, r{child_box(c, 1), (uint32_t)c.link_r};
            kids[best] = l;
            kids.insert(kids.begin() + best + 1, r);
        }
    };
    if (!root.inner()) {
        out.root_ref = make_leaf_ref((uint32_t)root.tri_offset(), (uint32_t)root.tri_count());
        return "";
    }
    // Pre-order numbering of wide nodes; child lists are computed once.
    std::vector<std::vector<child>> kids_of;
    std::vector<uint32_t> wide_src;   // wide node -> bbvh id
    std::vector<std::pair<uint32_t, uint32_t>> stack{{b.root, 1}};
    std::vector<int32_t> wide_id(b.nodes.size(), -1);
    while (!stack.empty()) {
        auto [id, d] = stack.back();
        stack.pop_back();
        wide_id[id] = (int32_t)wide_src.size();
        wide_src.push_back(id);
        if (d > out.depth) out.depth = d;
        std::vector<child> kids;
        collapse(id, kids);
        for (size_t k = kids.size(); k-- > 0;)
            if (b.nodes[kids[k].id].inner()) stack.push_back({kids[k].id, d + 1});
        kids_of.push_back(std::move(kids));
    }
    out.nodes4.resize(wide_src.size());
    for (size_t w = 0; w < wide_src.size(); ++w) {
        node4 &o = out.nodes4[w];
        for (int k = 0; k < 4; ++k) {
            o.lox[k] = o.loy[k] = o.loz[k] = INFINITY;
            o.hix[k] = o.hiy[k] = o.hiz[k] = -INFINITY;
            o.ref[k] = kInvalidRef;
            o.pad[k] = 0;
        }
        const std::vector<child> &kids = kids_of[w];
        for (size_t k = 0; k < kids.size(); ++k) {
            const aabb &bx = kids[k].box;
            o.lox[k] = bx.min.x;
            o.hix[k] = bx.max.x;
            o.loy[k] = bx.min.y;
            o.hiy[k] = bx.max.y;
            o.loz[k] = bx.min.z;
            o.hiz[k] = bx.max.z;
            const bbvh_node &c = b.nodes[kids[k].id];
            o.ref[k] = c.inner() ? wide_id[kids[k].id]
                                 : make_leaf_ref((uint32_t)c.tri_offset(), (uint32_t)c.tri_count());
        }
    }
    out.root_ref = 0;
    return "";
}

} // namespace rt
