// Host wide-BVH builder: binned-SAH binary tree, collapsed to 4- or 8-wide nodes.
//
// Replaces Embree's scene build behind rtcCommitScene (Tracer/Embree/Embree.cpp:648). The hit
// result is defined independently of the tree (closest t, ties to the smaller triangle index), so
// any valid tree gives the same answers; the SAH only decides the traversal cost.
#include "pg_bvh.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <future>
#include <stdexcept>

namespace pg {
namespace {

struct Aabb {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const Aabb& b) {
        for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], b.lo[k]), hi[k] = std::max(hi[k], b.hi[k]);
    }
    void grow(const float* p) {
        for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], p[k]), hi[k] = std::max(hi[k], p[k]);
    }
    float halfArea() const {
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        return dx < 0.0f ? 0.0f : dx * dy + dy * dz + dz * dx;
    }
};

struct Node2 {
    Aabb box;
    uint32_t a = 0, b = 0;  // internal: children; leaf: first, count
    bool leaf = false;
};

struct Builder {
    const float* tri;  // 9 floats per triangle
    std::vector<Aabb> boxes;
    std::vector<float> cen;
    std::vector<uint32_t> idx;
    uint32_t leafSize, bins;
    std::vector<Node2> nodes;
    std::mutex mu;

    uint32_t alloc(uint32_t n) {
        std::lock_guard<std::mutex> g(mu);
        const uint32_t r = static_cast<uint32_t>(nodes.size());
        nodes.resize(nodes.size() + n);
        return r;
    }

    // Builds the subtree over idx[begin, end) into node `self`. Large subtrees fork.
    void build(uint32_t self, uint32_t begin, uint32_t end, int forkDepth) {
        Aabb box, cbox;
        for (uint32_t i = begin; i < end; ++i) {
            box.grow(boxes[idx[i]]);
            cbox.grow(&cen[3 * idx[i]]);
        }
        const uint32_t n = end - begin;
        int axis = -1;
        uint32_t split = 0;
        float bestCost = static_cast<float>(n) * box.halfArea();  // cost of a leaf
        if (n > 1) {
            std::vector<Aabb> bb(bins);
            std::vector<uint32_t> bc(bins);
            std::vector<float> rArea(bins);
            std::vector<uint32_t> rCnt(bins);
            for (int k = 0; k < 3; ++k) {
                const float ext = cbox.hi[k] - cbox.lo[k];
                if (!(ext > 0.0f)) continue;
                const float scale = static_cast<float>(bins) / ext;
                std::fill(bb.begin(), bb.end(), Aabb{});
                std::fill(bc.begin(), bc.end(), 0u);
                for (uint32_t i = begin; i < end; ++i) {
                    const uint32_t t = idx[i];
                    const uint32_t bi = std::min(bins - 1, static_cast<uint32_t>(std::max(0.0f, (cen[3 * t + k] - cbox.lo[k]) * scale)));
                    bb[bi].grow(boxes[t]);
                    ++bc[bi];
                }
                Aabb acc;
                uint32_t ac = 0;
                for (uint32_t i = bins - 1; i > 0; --i) {
                    acc.grow(bb[i]);
                    ac += bc[i];
                    rArea[i] = acc.halfArea();
                    rCnt[i] = ac;
                }
                Aabb lacc;
                uint32_t lc = 0;
                for (uint32_t i = 0; i + 1 < bins; ++i) {
                    lacc.grow(bb[i]);
                    lc += bc[i];
                    if (lc == 0 || rCnt[i + 1] == 0) continue;
                    // traversal step costs about one triangle test
                    const float cost = box.halfArea() + lacc.halfArea() * lc + rArea[i + 1] * rCnt[i + 1];
                    if (cost < bestCost) bestCost = cost, axis = k, split = i + 1;
                }
            }
        }
        uint32_t mid;
        if (axis >= 0) {
            const float scale = static_cast<float>(bins) / (cbox.hi[axis] - cbox.lo[axis]);
            const float lo = cbox.lo[axis];
            auto it = std::partition(idx.begin() + begin, idx.begin() + end, [&](uint32_t t) {
                return std::min(bins - 1, static_cast<uint32_t>(std::max(0.0f, (cen[3 * t + axis] - lo) * scale))) < split;
            });
            mid = static_cast<uint32_t>(it - idx.begin());
        } else if (n > leafSize) {
            // no useful SAH split (coincident centroids) but too many for a leaf: split by index
            mid = begin + n / 2;
        } else {
            nodes[self].box = box;
            nodes[self].leaf = true;
            nodes[self].a = begin;
            nodes[self].b = n;
            return;
        }
        if (axis >= 0 && n <= leafSize && bestCost >= static_cast<float>(n) * box.halfArea()) {
            nodes[self].box = box;
            nodes[self].leaf = true;
            nodes[self].a = begin;
            nodes[self].b = n;
            return;
        }
        const uint32_t c = alloc(2);
        nodes[self].box = box;
        nodes[self].leaf = false;
        nodes[self].a = c;
        nodes[self].b = c + 1;
        if (forkDepth > 0 && n > 65536) {
            auto f = std::async(std::launch::async, [&, c, begin, mid, forkDepth] { build(c, begin, mid, forkDepth - 1); });
            build(c + 1, mid, end, forkDepth - 1);
            f.get();
        } else {
            build(c, begin, mid, 0);
            build(c + 1, mid, end, 0);
        }
    }
};

}  // namespace

void buildWideBvh(const float* triVerts, uint32_t numTris, const BvhParams& params, WideBvh& out) {
    const uint32_t W = params.width;
    if (W != 4 && W != 8) throw std::invalid_argument("BVH width must be 4 or 8");
    if (params.leafSize < 1 || params.leafSize > kMaxLeafSize) throw std::invalid_argument("leaf size must be 1..8");
    if (numTris >= (1u << kLeafCountShift)) throw std::invalid_argument("too many triangles for the leaf encoding");
    out = WideBvh{};
    out.width = W;
    if (numTris == 0) return;

    Builder b;
    b.tri = triVerts;
    b.leafSize = params.leafSize;
    b.bins = std::max(4u, std::min(64u, params.bins));
    b.boxes.resize(numTris);
    b.cen.resize(3 * numTris);
    b.idx.resize(numTris);
    float maxAbs = 1.0f;
    for (uint32_t i = 0; i < numTris; ++i) {
        for (int v = 0; v < 3; ++v) b.boxes[i].grow(triVerts + 9 * i + 3 * v);
        for (int k = 0; k < 3; ++k) {
            b.cen[3 * i + k] = 0.5f * (b.boxes[i].lo[k] + b.boxes[i].hi[k]);
            maxAbs = std::max(maxAbs, std::max(std::fabs(b.boxes[i].lo[k]), std::fabs(b.boxes[i].hi[k])));
        }
        b.idx[i] = i;
    }
    b.nodes.reserve(2 * numTris / std::max(1u, params.leafSize / 2) + 16);
    b.alloc(1);
    b.build(0, 0, numTris, 4);

    // Leaf-order triangles: (v0, prim) (v1, 0) (v2, 0)
    out.tris.resize(12 * static_cast<size_t>(numTris));
    out.order = b.idx;
    for (uint32_t i = 0; i < numTris; ++i) {
        const uint32_t t = b.idx[i];
        float* o = &out.tris[12 * static_cast<size_t>(i)];
        for (int v = 0; v < 3; ++v)
            for (int k = 0; k < 3; ++k) o[4 * v + k] = triVerts[9 * t + 3 * v + k];
        std::memcpy(&o[3], &t, 4);
        o[7] = o[11] = 0.0f;
    }

    // Collapse to W-wide nodes. Boxes are padded so that rounding in the slab test can never
    // cull a box that holds a hit the triangle test accepts.
    const float pad = 1e-5f * maxAbs;
    out.pad = pad;
    const size_t nodeFloats = 7 * W;  // 6W box planes + W child refs
    auto leafRef = [](const Node2& n) {
        return kLeafBit | ((n.b - 1) << kLeafCountShift) | n.a;
    };
    struct Task {
        uint32_t node2, wide, depth;
    };
    std::vector<Task> stack;
    out.nodes.clear();
    out.numNodes = 0;
    auto newWide = [&]() {
        out.nodes.resize(out.nodes.size() + nodeFloats);
        return out.numNodes++;
    };
    const uint32_t root = newWide();
    stack.push_back({0, root, 1});
    out.maxDepth = 1;
    while (!stack.empty()) {
        const Task t = stack.back();
        stack.pop_back();
        out.maxDepth = std::max(out.maxDepth, t.depth);
        std::vector<uint32_t> kids;
        if (b.nodes[t.node2].leaf) {
            kids.push_back(t.node2);
        } else {
            kids.push_back(b.nodes[t.node2].a);
            kids.push_back(b.nodes[t.node2].b);
            while (kids.size() < W) {
                int best = -1;
                float area = -1.0f;
                for (size_t i = 0; i < kids.size(); ++i) {
                    const Node2& k = b.nodes[kids[i]];
                    if (!k.leaf && k.box.halfArea() > area) area = k.box.halfArea(), best = static_cast<int>(i);
                }
                if (best < 0) break;
                const Node2 k = b.nodes[kids[best]];
                kids[best] = k.a;
                kids.push_back(k.b);
            }
        }
        float* lox = &out.nodes[nodeFloats * t.wide];
        float* loy = lox + W;
        float* loz = loy + W;
        float* hix = loz + W;
        float* hiy = hix + W;
        float* hiz = hiy + W;
        uint32_t* child = reinterpret_cast<uint32_t*>(hiz + W);
        for (uint32_t c = 0; c < W; ++c) {
            if (c >= kids.size()) {
                lox[c] = loy[c] = loz[c] = 1.0f;
                hix[c] = hiy[c] = hiz[c] = -1.0f;
                child[c] = kInvalidChild;
                continue;
            }
            const Node2& k = b.nodes[kids[c]];
            lox[c] = k.box.lo[0] - pad, loy[c] = k.box.lo[1] - pad, loz[c] = k.box.lo[2] - pad;
            hix[c] = k.box.hi[0] + pad, hiy[c] = k.box.hi[1] + pad, hiz[c] = k.box.hi[2] + pad;
            if (k.leaf) {
                child[c] = leafRef(k);
                ++out.numLeaves;
            } else {
                const uint32_t w = newWide();
                // newWide may reallocate: recompute pointers next iteration via index
                lox = &out.nodes[nodeFloats * t.wide];
                loy = lox + W, loz = loy + W, hix = loz + W, hiy = hix + W, hiz = hiy + W;
                child = reinterpret_cast<uint32_t*>(hiz + W);
                child[c] = w;
                stack.push_back({kids[c], w, t.depth + 1});
            }
        }
    }
}

}  // namespace pg
