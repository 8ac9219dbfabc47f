// ytrc_kdtree.hip -- the kd-tree of an AMR grid hierarchy (host code): the
// brick set and the brick order of the volume render.
//
// Build: AMRKDTree / Tree.build (yt/utilities/amr_kdtree/amr_kdtree.py:63-118)
// adds the grids level by level, coarsest first, to a root that spans all
// space; Node.add_grids / insert_grids / split_grids / divide
// (yt/utilities/lib/amr_kdtools.pyx:193-532) send them down existing splits
// and split leaves at the median unique grid face of the axis with the most
// faces strictly inside the node (kdtree_get_choices, :760-823).
// Order: step_viewpoint (amr_kdtools.pyx:708-755), farthest from the
// viewpoint first: at a split, the right child first when
// viewpoint[dim] <= pos.  Single process: should_i_build is always true and
// the parallel geometric splits (should_i_split) never happen.
//
// The tree is a few thousand nodes for hierarchies of ~10^3 grids; the work
// is branchy and serial, so it runs on the host, once per dataset (build)
// and once per camera (order).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ytrc.h"
#include "ytrc_host.h"

struct ytrc_kdtree {
    std::vector<ytrc_kd_node> nodes;  // nodes[0] is the root
    std::vector<int32_t> leaf_index;  // per node: depth-first index among leaves with a grid, or -1
    int32_t n_leaves = 0;
};

namespace {

struct Builder {
    ytrc_kdtree &t;
    const double *gle;
    const double *gre;
    const int64_t *gid;

    bool is_leaf(int n) const { return t.nodes[n].left < 0 || t.nodes[n].right < 0; }

    int new_node(int parent, const double le[3], const double re[3], int64_t grid, int64_t id)
    {
        ytrc_kd_node d;
        for (int a = 0; a < 3; a++) {
            d.left_edge[a] = le[a];
            d.right_edge[a] = re[a];
        }
        d.split_pos = NAN;
        d.grid = grid;
        d.node_id = id;
        d.parent = parent;
        d.left = d.right = -1;
        d.split_dim = -1;
        t.nodes.push_back(d);
        return (int)t.nodes.size() - 1;
    }

    // amr_kdtools.pyx:509-532: children inherit the node's grid.
    void divide(int n, int dim, double pos)
    {
        double le[3], re[3];
        const ytrc_kd_node p = t.nodes[n];
        for (int a = 0; a < 3; a++) {
            le[a] = p.left_edge[a];
            re[a] = p.right_edge[a];
        }
        re[dim] = pos;
        const int l = new_node(n, le, re, p.grid, (int64_t)((uint64_t)p.node_id << 1));
        re[dim] = p.right_edge[dim];
        le[dim] = pos;
        const int r = new_node(n, le, re, p.grid, (int64_t)(((uint64_t)p.node_id << 1) + 1));
        ytrc_kd_node &q = t.nodes[n];
        q.split_dim = dim;
        q.split_pos = pos;
        q.left = l;
        q.right = r;
    }

    // kdtree_get_choices (amr_kdtools.pyx:760-823): per axis, the unique grid
    // faces strictly inside the node; the axis with the most wins (the first
    // on ties); the split is the ((n - 1) / 2)-th smallest.
    int choose(int n, const std::vector<int> &ids, double &split) const
    {
        const ytrc_kd_node &nd = t.nodes[n];
        int best = -1;
        size_t my_max = 0;
        std::vector<double> v;
        for (int dim = 0; dim < 3; dim++) {
            v.clear();
            for (int i : ids) {
                const double a = gle[3 * i + dim], b = gre[3 * i + dim];
                if (nd.left_edge[dim] < a && a < nd.right_edge[dim]) v.push_back(a);
                if (nd.left_edge[dim] < b && b < nd.right_edge[dim]) v.push_back(b);
            }
            std::sort(v.begin(), v.end());
            v.erase(std::unique(v.begin(), v.end()), v.end());
            if (v.size() > my_max) {
                best = dim;
                my_max = v.size();
                split = v[(v.size() - 1) / 2];
            }
        }
        return best;
    }

    // amr_kdtools.pyx:369-458
    int split_grids(int n, const std::vector<int> &ids)
    {
        if (ids.empty()) return 0;
        double split = 0.0;
        const int dim = choose(n, ids, split);
        if (dim < 0) return -1;
        divide(n, dim, split);
        std::vector<int> less, greater;
        for (int i : ids) {
            if (gle[3 * i + dim] < split) less.push_back(i);
            if (gre[3 * i + dim] > split) greater.push_back(i);
        }
        const int l = t.nodes[n].left, r = t.nodes[n].right;
        if (!less.empty()) insert_grids(l, less);
        if (!greater.empty()) insert_grids(r, greater);
        return 0;
    }

    // amr_kdtools.pyx:275-313
    void insert_grids(int n, const std::vector<int> &ids)
    {
        if (ids.empty()) return;
        if (ids.size() == 1) {
            const int i = ids[0];
            bool covers = true;
            for (int a = 0; a < 3; a++)
                covers = covers && gle[3 * i + a] <= t.nodes[n].left_edge[a] &&
                         gre[3 * i + a] >= t.nodes[n].right_edge[a];
            if (covers) {
                t.nodes[n].grid = gid[i];
                return;
            }
        }
        if (split_grids(n, ids) == -1) t.nodes[n].grid = -1;
    }

    // amr_kdtools.pyx:193-262
    void add_grids(int n, const std::vector<int> &ids)
    {
        if (is_leaf(n)) {
            insert_grids(n, ids);
            return;
        }
        const int d = t.nodes[n].split_dim;
        const double p = t.nodes[n].split_pos;
        std::vector<int> less, greater;
        for (int i : ids) {
            if (gle[3 * i + d] < p) less.push_back(i);
            if (gre[3 * i + d] > p) greater.push_back(i);
        }
        const int l = t.nodes[n].left, r = t.nodes[n].right;
        if (!less.empty()) add_grids(l, less);
        if (!greater.empty()) add_grids(r, greater);
    }
};

bool leaf_node(const ytrc_kd_node &n) { return n.left < 0 || n.right < 0; }

}  // namespace

extern "C" {

int ytrc_kdtree_build(int32_t n_grids, const double *left_edges, const double *right_edges,
                      const int64_t *grid_ids, const int32_t *levels, ytrc_kdtree **out)
{
    if (out == nullptr) return ytrc::set_error(YTRC_EINVAL, "ytrc_kdtree_build: NULL out");
    *out = nullptr;
    if (n_grids < 0 || (n_grids > 0 && (left_edges == nullptr || right_edges == nullptr ||
                                        grid_ids == nullptr || levels == nullptr)))
        return ytrc::set_error(YTRC_EINVAL, "ytrc_kdtree_build: bad grid arrays");
    for (int i = 0; i < n_grids; i++) {
        if (grid_ids[i] < 0)
            return ytrc::set_error(YTRC_EINVAL, "ytrc_kdtree_build: grid ids must be >= 0 (-1 marks no grid)");
        for (int a = 0; a < 3; a++)
            if (!(right_edges[3 * i + a] > left_edges[3 * i + a]))
                return ytrc::set_error(YTRC_EINVAL, "ytrc_kdtree_build: grid " + std::to_string(i) +
                                                        " has right_edge <= left_edge");
    }
    ytrc_kdtree *t = new ytrc_kdtree();
    Builder B{*t, left_edges, right_edges, grid_ids};
    const double inf = INFINITY;
    const double le[3] = {-inf, -inf, -inf}, re[3] = {inf, inf, inf};
    B.new_node(-1, le, re, -1, 1);
    // Levels coarse to fine (amr_kdtree.py:100-118); within a level, the
    // caller's (blocks) order.
    std::vector<int32_t> lv(levels, levels + n_grids);
    std::sort(lv.begin(), lv.end());
    lv.erase(std::unique(lv.begin(), lv.end()), lv.end());
    for (int32_t L : lv) {
        std::vector<int> ids;
        for (int i = 0; i < n_grids; i++)
            if (levels[i] == L) ids.push_back(i);
        if (!ids.empty()) B.add_grids(0, ids);
    }
    // Depth-first (left first) index of every leaf that carries a grid.
    t->leaf_index.assign(t->nodes.size(), -1);
    std::vector<int> stack{0};
    while (!stack.empty()) {
        const int n = stack.back();
        stack.pop_back();
        const ytrc_kd_node &d = t->nodes[n];
        if (leaf_node(d)) {
            if (d.grid != -1) t->leaf_index[n] = t->n_leaves++;
        } else {
            stack.push_back(d.right);
            stack.push_back(d.left);
        }
    }
    *out = t;
    return YTRC_OK;
}

int ytrc_kdtree_size(const ytrc_kdtree *t, int32_t *n_nodes, int32_t *n_leaves)
{
    if (t == nullptr) return ytrc::set_error(YTRC_EINVAL, "ytrc_kdtree_size: NULL tree");
    if (n_nodes) *n_nodes = (int32_t)t->nodes.size();
    if (n_leaves) *n_leaves = t->n_leaves;
    return YTRC_OK;
}

int ytrc_kdtree_read_nodes(const ytrc_kdtree *t, ytrc_kd_node *nodes)
{
    if (t == nullptr || nodes == nullptr) return ytrc::set_error(YTRC_EINVAL, "ytrc_kdtree_read_nodes: NULL");
    std::copy(t->nodes.begin(), t->nodes.end(), nodes);
    return YTRC_OK;
}

int ytrc_kdtree_order(const ytrc_kdtree *t, const double *viewpoint, int32_t *order)
{
    if (t == nullptr || viewpoint == nullptr || order == nullptr)
        return ytrc::set_error(YTRC_EINVAL, "ytrc_kdtree_order: NULL");
    int32_t k = 0;
    std::vector<int> stack{0};
    while (!stack.empty()) {
        const int n = stack.back();
        stack.pop_back();
        const ytrc_kd_node &d = t->nodes[n];
        if (leaf_node(d)) {
            if (t->leaf_index[n] >= 0) order[k++] = t->leaf_index[n];
            continue;
        }
        // step_viewpoint: the right child first when viewpoint[dim] <= pos.
        const bool right_first = viewpoint[d.split_dim] <= d.split_pos;
        const int first = right_first ? d.right : d.left;
        const int second = right_first ? d.left : d.right;
        stack.push_back(second);
        stack.push_back(first);
    }
    return YTRC_OK;
}

void ytrc_kdtree_destroy(ytrc_kdtree *t) { delete t; }

}  // extern "C"
