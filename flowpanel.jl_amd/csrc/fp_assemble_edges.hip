// fp_assemble_edges.hip -- vortex-ring AIC assembly with node and edge terms
// shared between the panels of a block (gfx950).
//
// G(i, j) = n_i . sum over the edges (a -> b) of panel j of the regularised
// Biot-Savart segment at control point i (src/FLOWPanel_elements.jl:345-467,
// assembled as src/FLOWPanel_liftingbody.jl:455-572).  The segment splits into
//   node terms   r = X - p,  iv = 1 / (|r| + offset)            (per mesh node)
//   edge terms   c = r_a x r_b,  proj = (p_b - p_a) . (r_a iv_a - r_b iv_b),
//                s = proj / (4 pi (|c|^2 + offset^2)) (c . n_i)    (per mesh edge)
// and a panel only adds its edges' s with the sign of its traversal (the
// segment b -> a gives -s exactly in exact arithmetic).  A quad mesh has about
// one node and two edges per panel against the four and four a panel-wise
// kernel evaluates, a triangle mesh a half and one and a half against three
// and three.
//
// Columns are grouped on the host into blocks of at most 256 panels, 512
// nodes and 768 edges: runs of up to 16 consecutive columns that share nodes
// (so each row of a block is written as a few contiguous segments), a block
// seeded by the first unplaced run along a Morton curve of the run centroids
// and grown by the run sharing the most nodes with it (a compact patch: on a
// structured loft, runs stacked across strips), else by the next unplaced runs
// along the curve.  A
// workgroup takes one block and a tile of TM targets; for TT targets at a time
// it evaluates the block's node terms (lane = node slot) into LDS, then the
// edge terms (lane = edge slot) into LDS, then each lane sums its panel's up to
// four signed edge terms and stores G(i, j) (and its FP32 rounding for the
// mixed path).  Panels with more than four vertices use the panel-wise kernel.
#include "fp_internal.h"
#include "fp_elements.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <unordered_map>

namespace {

constexpr int VRE_P = 256;                 // panels (lanes) per block
constexpr int VRE_NS = 2, VRE_ES = 3;      // node and edge slots per lane
constexpr int VRE_NMAX = VRE_NS * VRE_P;   // node slots per block in the table (512)
constexpr int VRE_EMAX = VRE_ES * VRE_P;   // edge slots per block in the table (768)
// blocks are packed to at most VRE_NPACK nodes and VRE_EPACK edges (the LDS the
// kernel holds per target: 4 workgroups fit a CU at two targets per pass)
constexpr int VRE_NPACK = 384, VRE_EPACK = 640;
constexpr int VRE_RUN = 16;                // columns per run
// per block (int32): node refs (panel*8 + vertex), edges (a | b << 16), columns, edge refs (2 per panel)
constexpr int VRE_OFF_NODE = 0, VRE_OFF_EDGE = VRE_NMAX, VRE_OFF_COL = VRE_NMAX + VRE_EMAX,
              VRE_OFF_REF = VRE_NMAX + VRE_EMAX + VRE_P;
constexpr int VRE_BLOCK_INTS = VRE_OFF_REF + 2 * VRE_P;
constexpr unsigned VRE_NOREF = 0xFFFFu;

struct __align__(32) NodeTerm {
    double x, y, z, iv;
};

uint32_t spread10(uint32_t v)
{
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

}  // namespace

// Host: the block tables of a body whose panels have at most four vertices
// (cell_ptr/cell_idx 0-based CSR, nodes 3 x nn).  Appends VRE_BLOCK_INTS int32
// per block to `tab`; returns the number of blocks and, in sums[0..1], the
// distinct nodes and edges summed over the blocks (the terms per target).
static int64_t build_edge_blocks(const double *nodes, int64_t nn, const int64_t *cell_ptr, const int64_t *cell_idx,
                                 int64_t N, std::vector<int32_t> &tab, int64_t sums[2], int npack = VRE_NPACK,
                                 int epack = VRE_EPACK)
{
    auto cell = [&](int64_t j, int k) { return cell_idx[cell_ptr[j] + k]; };
    auto nv = [&](int64_t j) { return (int)(cell_ptr[j + 1] - cell_ptr[j]); };
    auto shares_node = [&](int64_t j0, int64_t j1) {
        for (int k = 0; k < nv(j0); ++k)
            for (int q = 0; q < nv(j1); ++q)
                if (cell(j0, k) == cell(j1, q)) return true;
        return false;
    };
    // runs of up to VRE_RUN consecutive columns, broken where neighbours share no node
    std::vector<int64_t> run0;
    for (int64_t j = 0; j < N;) {
        int64_t e = j + 1;
        while (e < N && e - j < VRE_RUN && shares_node(e - 1, e)) ++e;
        run0.push_back(j);
        j = e;
    }
    const int64_t nr = (int64_t)run0.size();
    run0.push_back(N);
    // Morton order of the run centroids
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    std::vector<double> cen(3 * nr, 0.0);
    for (int64_t r = 0; r < nr; ++r) {
        int cnt = 0;
        for (int64_t j = run0[r]; j < run0[r + 1]; ++j)
            for (int k = 0; k < nv(j); ++k, ++cnt)
                for (int d = 0; d < 3; ++d) cen[3 * r + d] += nodes[3 * cell(j, k) + d];
        for (int d = 0; d < 3; ++d) {
            cen[3 * r + d] /= cnt;
            lo[d] = std::min(lo[d], cen[3 * r + d]);
            hi[d] = std::max(hi[d], cen[3 * r + d]);
        }
    }
    const double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-300});
    std::vector<uint32_t> code(nr);
    for (int64_t r = 0; r < nr; ++r) {
        uint32_t c = 0;
        for (int d = 0; d < 3; ++d) {
            const double u = (cen[3 * r + d] - lo[d]) / ext;
            c |= spread10((uint32_t)std::min(1023.0, std::max(0.0, u * 1023.0))) << d;
        }
        code[r] = c;
    }
    std::vector<int64_t> order(nr);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return code[x] < code[y]; });

    std::vector<int64_t> rank(nr);
    for (int64_t q = 0; q < nr; ++q) rank[order[q]] = q;
    // nodes of each run, runs of each node
    std::vector<std::vector<int64_t>> rnodes(nr);
    std::vector<std::vector<int64_t>> noderuns(nn);
    for (int64_t r = 0; r < nr; ++r) {
        auto &v = rnodes[r];
        for (int64_t j = run0[r]; j < run0[r + 1]; ++j)
            for (int k = 0; k < nv(j); ++k) v.push_back(cell(j, k));
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        for (int64_t u : v) noderuns[u].push_back(r);
    }

    // blocks grown greedily: seeded by the first unplaced run along the curve, then the
    // unplaced run sharing the most nodes with the block (ties: earlier on the curve)
    // while the panel, node and edge limits hold
    std::vector<int64_t> stamp(nn, -1);
    std::vector<int32_t> local(nn, 0);
    std::vector<char> placed(nr, 0);
    std::unordered_map<uint64_t, int> emap, enew;
    std::unordered_map<int64_t, int> score;
    int64_t nb = -1;
    int np = 0, nnod = 0, ned = 0;
    sums[0] = sums[1] = 0;
    auto ekey = [](int64_t u, int64_t w) {
        return u < w ? ((uint64_t)u << 32) | (uint64_t)w : ((uint64_t)w << 32) | (uint64_t)u;
    };
    const size_t base = tab.size();
    auto fits = [&](int64_t r) {
        const int len = (int)(run0[r + 1] - run0[r]);
        if (np + len > VRE_P) return false;
        int add_n = 0, add_e = 0;
        for (int64_t u : rnodes[r]) add_n += stamp[u] != nb;
        if (nnod + add_n > npack) return false;
        enew.clear();
        for (int64_t j = run0[r]; j < run0[r + 1]; ++j)
            for (int k = 0; k < nv(j); ++k) {
                const int64_t u = cell(j, k), w = cell(j, (k + 1) % nv(j));
                if (u != w) {
                    const uint64_t key = ekey(u, w);
                    if (!emap.count(key) && enew.emplace(key, 1).second) ++add_e;
                }
            }
        return ned + add_e <= epack;
    };
    auto place = [&](int64_t r) {
        placed[r] = 1;
        score.erase(r);
        int32_t *T = &tab[base + (size_t)nb * VRE_BLOCK_INTS];
        for (int64_t j = run0[r]; j < run0[r + 1]; ++j) {
            const int lane = np++;
            T[VRE_OFF_COL + lane] = (int32_t)j;
            unsigned ref[4] = {VRE_NOREF, VRE_NOREF, VRE_NOREF, VRE_NOREF};
            for (int k = 0; k < nv(j); ++k) {
                const int64_t u = cell(j, k), w = cell(j, (k + 1) % nv(j));
                for (int64_t v : {u, w})
                    if (stamp[v] != nb) {
                        stamp[v] = nb;
                        local[v] = nnod;
                        int kk = 0;
                        while (cell(j, kk) != v) ++kk;
                        T[VRE_OFF_NODE + nnod] = (int32_t)(j * 8 + kk);
                        ++nnod;
                        for (int64_t r2 : noderuns[v])
                            if (!placed[r2]) ++score[r2];
                    }
                if (u == w) continue;  // repeated vertex: a zero-length edge adds nothing
                const uint64_t key = ekey(u, w);
                auto it = emap.find(key);
                int e;
                unsigned neg = 0;
                if (it == emap.end()) {
                    e = ned++;
                    emap.emplace(key, e);
                    T[VRE_OFF_EDGE + e] = (int32_t)((uint32_t)local[u] | ((uint32_t)local[w] << 16));
                } else {
                    e = it->second;
                    neg = ((uint32_t)T[VRE_OFF_EDGE + e] & 0xFFFFu) == (uint32_t)local[u] ? 0u : 1u;
                }
                ref[k] = ((unsigned)e << 1) | neg;
            }
            T[VRE_OFF_REF + 2 * lane] = (int32_t)(ref[0] | (ref[1] << 16));
            T[VRE_OFF_REF + 2 * lane + 1] = (int32_t)(ref[2] | (ref[3] << 16));
        }
    };
    std::vector<std::pair<int64_t, int64_t>> cand;  // (-score, rank)
    int64_t next = 0;
    for (;;) {
        while (next < nr && placed[order[next]]) ++next;
        if (next == nr) break;
        sums[0] += nnod;
        sums[1] += ned;
        ++nb;
        tab.resize(base + (size_t)(nb + 1) * VRE_BLOCK_INTS, -1);
        np = nnod = ned = 0;
        emap.clear();
        score.clear();
        place(order[next]);
        for (;;) {
            cand.clear();
            for (const auto &kv : score) cand.emplace_back(-(int64_t)kv.second, rank[kv.first]);
            std::sort(cand.begin(), cand.end());
            bool added = false;
            for (const auto &c : cand)
                if (fits(order[c.second])) {
                    place(order[c.second]);
                    added = true;
                    break;
                }
            // nothing adjacent fits (or nothing is adjacent): the next unplaced runs along the curve
            for (int64_t q = next, tries = 0; !added && q < nr && tries < 8; ++q) {
                if (placed[order[q]]) continue;
                ++tries;
                if (fits(order[q])) {
                    place(order[q]);
                    added = true;
                }
            }
            if (!added) break;
        }
    }
    sums[0] += nnod;
    sums[1] += ned;
    return nb + 1;
}

static int64_t build_pair_blocks(const double *nodes, int64_t nn, const int64_t *cell_ptr, const int64_t *cell_idx,
                                 int64_t N, std::vector<int32_t> &out, int64_t sums[2]);

// Build and upload the block tables of a vortex-ring body of triangles and quads.
int fp_body_build_edge_blocks(fp_ctx *ctx, fp_body *b)
{
    if (b->element != FP_ELEM_VORTEXRING || b->nvp != 4 || b->N == 0) return FP_OK;
    std::vector<int32_t> tab;
    int64_t sums[2];
    const int64_t nblocks = build_edge_blocks(b->nodes.data(), (int64_t)b->nodes.size() / 3, b->cell_ptr.data(),
                                              b->cell_idx.data(), b->N, tab, sums);
    FP_HIP(ctx, hipMalloc((void **)&b->d_eb, tab.size() * sizeof(int32_t)));
    FP_HIP(ctx, hipMemcpy(b->d_eb, tab.data(), tab.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    b->eb_nblocks = nblocks;
    b->eb_nodes = sums[0];
    b->eb_edges = sums[1];
    return FP_OK;
}

// The edge-pair tables, built on the first assembly that asks for them.
static int body_build_pair_blocks(fp_ctx *ctx, fp_body *b)
{
    std::vector<int32_t> pt;
    int64_t sums[2];
    const int64_t npb = build_pair_blocks(b->nodes.data(), (int64_t)b->nodes.size() / 3, b->cell_ptr.data(),
                                          b->cell_idx.data(), b->N, pt, sums);
    if (npb <= 0) return fp_fail(ctx, FP_ERR_UNSUPPORTED, "edge-pair tables: a block does not pair into 256 lanes");
    FP_HIP(ctx, hipMalloc((void **)&b->d_pb, pt.size() * sizeof(int32_t)));
    FP_HIP(ctx, hipMemcpy(b->d_pb, pt.data(), pt.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    b->pb_nblocks = npb;
    return FP_OK;
}

extern "C" int64_t fp_assembly_blocks(const double *nodes, int64_t nnodes, const int64_t *cell_ptr,
                                      const int64_t *cell_idx, int64_t ncells, int index_base, int32_t *tab,
                                      int64_t cap_blocks, int64_t *terms)
{
    if (!nodes || !cell_ptr || !cell_idx || ncells <= 0 || nnodes <= 0) return -FP_ERR_ARG;
    std::vector<int64_t> cp(ncells + 1), ci((size_t)(cell_ptr[ncells] - cell_ptr[0]));
    for (int64_t c = 0; c <= ncells; ++c) cp[c] = cell_ptr[c] - cell_ptr[0];
    for (int64_t c = 0; c < ncells; ++c)
        if (cp[c + 1] - cp[c] < 3 || cp[c + 1] - cp[c] > 4) return -FP_ERR_UNSUPPORTED;
    for (size_t k = 0; k < ci.size(); ++k) {
        ci[k] = cell_idx[cell_ptr[0] + (int64_t)k] - index_base;
        if (ci[k] < 0 || ci[k] >= nnodes) return -FP_ERR_ARG;
    }
    std::vector<int32_t> t;
    int64_t sums[2];
    const int64_t nb = build_edge_blocks(nodes, nnodes, cp.data(), ci.data(), ncells, t, sums);
    if (terms) { terms[0] = sums[0]; terms[1] = sums[1]; }
    if (tab && cap_blocks >= nb) std::memcpy(tab, t.data(), t.size() * sizeof(int32_t));
    return nb;
}

// ---------------------------------------------------------------------------
// Edge-pair tables.  The same column blocks (at most 504 distinct edges), each
// block's edges paired at shared nodes: lane l of the workgroup owns a path
// n0 - n1 - n2 of two edges (n0 -> n1 in edge slot l, n1 -> n2 in slot 256 + l)
// or a single edge n0 -> n1 (n2 = -1).  The lane evaluates the three node terms
// in registers, the edge terms go through LDS to the panels' sums.  A depth-first
// pass pairs every edge of a connected block except at most one (each vertex
// pairs the edges left to it by its subtree and its non-tree edges, and hands an
// odd one up with the edge to its parent).
// Per block (int32): columns, column edge refs (2 per panel: slot << 1 | sign, 16
// bits each), node refs (panel * 8 + vertex) of n0, n1, n2 of every lane.
namespace {
constexpr int VRC_OFF_COL = 0, VRC_OFF_REF = VRE_P, VRC_OFF_NODE = 3 * VRE_P;
constexpr int VRC_BLOCK_INTS = 6 * VRE_P;
constexpr int VRC_EPACK = 504;     // first distinct-edge limit per block
constexpr int VRC_NPACK = VRE_NMAX;

// Pairs the edges of one block (local node pairs ea/eb, nl nodes); lanes get
// (n0, n1, n2) local nodes, n2 = -1 for a single edge; slot_of[e] = the slot of
// edge e and fwd[e] = whether that slot runs a -> b.  Returns the lanes used.
int pair_block_edges(int nl, const std::vector<int> &ea, const std::vector<int> &eb, std::vector<int> &l0,
                     std::vector<int> &l1, std::vector<int> &l2, std::vector<int> &slot_of, std::vector<char> &fwd)
{
    const int ne = (int)ea.size();
    std::vector<std::vector<int>> inc(nl);
    for (int e = 0; e < ne; ++e) {
        inc[ea[e]].push_back(e);
        inc[eb[e]].push_back(e);
    }
    std::vector<char> visited(nl, 0), claimed(ne, 0);
    l0.clear(); l1.clear(); l2.clear();
    slot_of.assign(ne, -1);
    fwd.assign(ne, 1);
    auto other = [&](int e, int v) { return ea[e] == v ? eb[e] : ea[e]; };
    auto lane_pair = [&](int e1, int e2, int v) {   // path other(e1) - v - other(e2)
        const int lane = (int)l0.size();
        const int n0 = other(e1, v), n2 = other(e2, v);
        l0.push_back(n0); l1.push_back(v); l2.push_back(n2);
        slot_of[e1] = lane;                    // n0 -> v
        fwd[e1] = ea[e1] == n0;
        slot_of[e2] = VRE_P + lane;            // v -> n2
        fwd[e2] = ea[e2] == v;
    };
    auto lane_single = [&](int e) {
        const int lane = (int)l0.size();
        l0.push_back(ea[e]); l1.push_back(eb[e]); l2.push_back(-1);
        slot_of[e] = lane;
        fwd[e] = 1;
    };
    // iterative depth-first search: frame = (vertex, parent edge, next incidence, pending edges)
    struct Frame { int v, pe; size_t next; std::vector<int> pend; };
    for (int root = 0; root < nl; ++root) {
        if (visited[root] || inc[root].empty()) continue;
        std::vector<Frame> st;
        visited[root] = 1;
        st.push_back({root, -1, 0, {}});
        while (!st.empty()) {
            Frame &f = st.back();
            if (f.next < inc[f.v].size()) {
                const int e = inc[f.v][f.next++];
                if (e == f.pe || claimed[e]) continue;
                claimed[e] = 1;
                const int u = other(e, f.v);
                if (!visited[u]) {
                    visited[u] = 1;
                    st.push_back({u, e, 0, {}});   // tree edge: settled when the child returns
                } else {
                    f.pend.push_back(e);           // non-tree edge, paired at this vertex
                }
                continue;
            }
            // all incidences seen: pair what is pending here, an odd one with the parent edge
            while (f.pend.size() >= 2) {
                const int e1 = f.pend.back(); f.pend.pop_back();
                const int e2 = f.pend.back(); f.pend.pop_back();
                lane_pair(e1, e2, f.v);
            }
            const int v = f.v, pe = f.pe;
            const bool odd = !f.pend.empty();
            const int left = odd ? f.pend[0] : -1;
            st.pop_back();
            if (pe < 0) {
                if (odd) lane_single(left);
            } else if (odd) {
                lane_pair(left, pe, v);            // the parent edge is used here
            } else {
                st.back().pend.push_back(pe);      // the parent pairs it
            }
        }
    }
    return (int)l0.size();
}
}  // namespace

// Host: edge-pair tables of a body (cell lists as build_edge_blocks).  Appends
// VRC_BLOCK_INTS int32 per block; returns the number of blocks (-1 if a block
// cannot be paired into 256 lanes at any edge limit) and, in sums[0..1], the
// lanes and the distinct edges summed over the blocks.
static int64_t build_pair_blocks(const double *nodes, int64_t nn, const int64_t *cell_ptr, const int64_t *cell_idx,
                                 int64_t N, std::vector<int32_t> &out, int64_t sums[2])
{
    for (int epack = VRC_EPACK; epack >= 96; epack -= 48) {
        std::vector<int32_t> tab;
        int64_t s2[2];
        const int64_t nb = build_edge_blocks(nodes, nn, cell_ptr, cell_idx, N, tab, s2, VRC_NPACK, epack);
        out.assign((size_t)nb * VRC_BLOCK_INTS, -1);
        sums[0] = sums[1] = 0;
        bool ok = true;
        std::vector<int> ea, eb, l0, l1, l2, slot_of;
        std::vector<char> fwd;
        for (int64_t b = 0; b < nb && ok; ++b) {
            const int32_t *T = &tab[(size_t)b * VRE_BLOCK_INTS];
            int32_t *O = &out[(size_t)b * VRC_BLOCK_INTS];
            int nl = 0;
            while (nl < VRE_NMAX && T[VRE_OFF_NODE + nl] >= 0) ++nl;
            ea.clear(); eb.clear();
            for (int e = 0; e < VRE_EMAX && T[VRE_OFF_EDGE + e] >= 0; ++e) {
                ea.push_back(T[VRE_OFF_EDGE + e] & 0xFFFF);
                eb.push_back((T[VRE_OFF_EDGE + e] >> 16) & 0xFFFF);
            }
            const int lanes = pair_block_edges(nl, ea, eb, l0, l1, l2, slot_of, fwd);
            if (lanes > VRE_P) { ok = false; break; }
            sums[0] += lanes;
            sums[1] += (int64_t)ea.size();
            for (int l = 0; l < lanes; ++l) {
                O[VRC_OFF_NODE + l] = T[VRE_OFF_NODE + l0[l]];
                O[VRC_OFF_NODE + VRE_P + l] = T[VRE_OFF_NODE + l1[l]];
                O[VRC_OFF_NODE + 2 * VRE_P + l] = l2[l] >= 0 ? T[VRE_OFF_NODE + l2[l]] : -1;
            }
            for (int lane = 0; lane < VRE_P; ++lane) {
                O[VRC_OFF_COL + lane] = T[VRE_OFF_COL + lane];
                uint32_t r[4];
                const uint32_t r01 = (uint32_t)T[VRE_OFF_REF + 2 * lane], r23 = (uint32_t)T[VRE_OFF_REF + 2 * lane + 1];
                const uint32_t in[4] = {r01 & 0xFFFFu, r01 >> 16, r23 & 0xFFFFu, r23 >> 16};
                for (int k = 0; k < 4; ++k) {
                    if (T[VRE_OFF_COL + lane] < 0 || in[k] == VRE_NOREF) { r[k] = VRE_NOREF; continue; }
                    const int e = (int)(in[k] >> 1);
                    const uint32_t neg = (in[k] & 1u) ^ (fwd[e] ? 0u : 1u);
                    r[k] = ((uint32_t)slot_of[e] << 1) | neg;
                }
                O[VRC_OFF_REF + 2 * lane] = (int32_t)(r[0] | (r[1] << 16));
                O[VRC_OFF_REF + 2 * lane + 1] = (int32_t)(r[2] | (r[3] << 16));
            }
        }
        if (ok) return nb;
    }
    return -1;
}

extern "C" int64_t fp_assembly_pairs(const double *nodes, int64_t nnodes, const int64_t *cell_ptr,
                                     const int64_t *cell_idx, int64_t ncells, int index_base, int32_t *tab,
                                     int64_t cap_blocks, int64_t *terms)
{
    if (!nodes || !cell_ptr || !cell_idx || ncells <= 0 || nnodes <= 0) return -FP_ERR_ARG;
    std::vector<int64_t> cp(ncells + 1), ci((size_t)(cell_ptr[ncells] - cell_ptr[0]));
    for (int64_t c = 0; c <= ncells; ++c) cp[c] = cell_ptr[c] - cell_ptr[0];
    for (int64_t c = 0; c < ncells; ++c)
        if (cp[c + 1] - cp[c] < 3 || cp[c + 1] - cp[c] > 4) return -FP_ERR_UNSUPPORTED;
    for (size_t k = 0; k < ci.size(); ++k) {
        ci[k] = cell_idx[cell_ptr[0] + (int64_t)k] - index_base;
        if (ci[k] < 0 || ci[k] >= nnodes) return -FP_ERR_ARG;
    }
    std::vector<int32_t> t;
    int64_t sums[2];
    const int64_t nb = build_pair_blocks(nodes, nnodes, cp.data(), ci.data(), ncells, t, sums);
    if (nb < 0) return -FP_ERR_UNSUPPORTED;
    if (terms) { terms[0] = sums[0]; terms[1] = sums[1]; }
    if (tab && cap_blocks >= nb) std::memcpy(tab, t.data(), t.size() * sizeof(int32_t));
    return nb;
}

template <typename TO, int TM, int TT>
__global__ __launch_bounds__(256) void k_assemble_vre(const int32_t *__restrict__ tab, const double *__restrict__ pv,
                                                      int64_t N, TargetView tv, TO *__restrict__ G, int64_t ld,
                                                      int64_t col0, float *__restrict__ G2, double off, double cut)
{
    __shared__ double st[6][TM];
    __shared__ NodeTerm sn[TT][VRE_NPACK];
    __shared__ double se[TT][VRE_EPACK];
    const int tid = threadIdx.x;
    const int32_t *T = tab + (int64_t)blockIdx.x * VRE_BLOCK_INTS;
    const int64_t i0 = (int64_t)blockIdx.y * TM;
    for (int t = tid; t < TM; t += blockDim.x) {
        const int64_t i = i0 + t;
        const bool ok = i < tv.M;
        st[0][t] = ok ? tv.x[i] : 0.0;
        st[1][t] = ok ? tv.y[i] : 0.0;
        st[2][t] = ok ? tv.z[i] : 0.0;
        st[3][t] = ok ? tv.nx[i] : 0.0;
        st[4][t] = ok ? tv.ny[i] : 0.0;
        st[5][t] = ok ? tv.nz[i] : 0.0;
    }
    auto node_xyz = [&](int32_t ref, double &x, double &y, double &z) {
        const int64_t j = ref >> 3;
        const int k = ref & 7;
        x = pv[(3 * k + 0) * N + j];
        y = pv[(3 * k + 1) * N + j];
        z = pv[(3 * k + 2) * N + j];
    };
    // node slots of this lane (slot s = node s * 256 + tid of the block)
    double px[VRE_NS], py[VRE_NS], pz[VRE_NS];
    bool nok[VRE_NS];
#pragma unroll
    for (int s = 0; s < VRE_NS; ++s) {
        const int n = s * VRE_P + tid;
        const int32_t ref = n < VRE_NPACK ? T[VRE_OFF_NODE + n] : -1;
        nok[s] = ref >= 0;
        px[s] = py[s] = pz[s] = 0.0;
        if (nok[s]) node_xyz(ref, px[s], py[s], pz[s]);
    }
    // edge slots: local end nodes and the edge vector b - a in FP64 coordinates
    int ea[VRE_ES], eb[VRE_ES];
    double ex[VRE_ES], ey[VRE_ES], ez[VRE_ES];
    bool eok[VRE_ES];
#pragma unroll
    for (int s = 0; s < VRE_ES; ++s) {
        const int e = s * VRE_P + tid;
        const int32_t ab = e < VRE_EPACK ? T[VRE_OFF_EDGE + e] : -1;
        eok[s] = ab >= 0;
        ea[s] = eok[s] ? (ab & 0xFFFF) : 0;
        eb[s] = eok[s] ? ((ab >> 16) & 0xFFFF) : 0;
        ex[s] = ey[s] = ez[s] = 0.0;
        if (eok[s]) {
            double ax, ay, az, bx, by, bz;
            node_xyz(T[VRE_OFF_NODE + ea[s]], ax, ay, az);
            node_xyz(T[VRE_OFF_NODE + eb[s]], bx, by, bz);
            ex[s] = bx - ax; ey[s] = by - ay; ez[s] = bz - az;
        }
    }
    const int32_t col = T[VRE_OFF_COL + tid];
    int ri[4];
    double rs[4];
    {
        const uint32_t r01 = (uint32_t)T[VRE_OFF_REF + 2 * tid], r23 = (uint32_t)T[VRE_OFF_REF + 2 * tid + 1];
        const uint32_t r[4] = {r01 & 0xFFFFu, r01 >> 16, r23 & 0xFFFFu, r23 >> 16};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool ok = col >= 0 && r[k] != VRE_NOREF;
            ri[k] = ok ? (int)(r[k] >> 1) : 0;
            rs[k] = ok ? ((r[k] & 1u) ? -1.0 : 1.0) : 0.0;
        }
    }
    const double off2 = off * off, cut2 = cut * cut, k4 = 4.0 * FP_PI;
    const int iend = (int)min((int64_t)TM, tv.M - i0);
    TO *g = G + i0 * ld + col0 + (col >= 0 ? col : 0);
    float *g2 = G2 ? G2 + i0 * ld + col0 + (col >= 0 ? col : 0) : nullptr;
    __syncthreads();
    for (int t0 = 0; t0 < iend; t0 += TT) {
        // node terms (a partial last pass repeats the last target: same values, never stored twice);
        // per slot, the TT targets are independent chains and are stored together
#pragma unroll
        for (int s = 0; s < VRE_NS; ++s) {
            NodeTerm q[TT];
#pragma unroll
            for (int tt = 0; tt < TT; ++tt) {
                const int t = min(t0 + tt, iend - 1);
                const double rx = st[0][t] - px[s], ry = st[1][t] - py[s], rz = st[2][t] - pz[s];
                q[tt].x = rx; q[tt].y = ry; q[tt].z = rz;
                q[tt].iv = fp_recip(fp_sqrt_nb(rx * rx + ry * ry + rz * rz) + off);
            }
            if (nok[s]) {
#pragma unroll
                for (int tt = 0; tt < TT; ++tt) sn[tt][s * VRE_P + tid] = q[tt];
            }
        }
        __syncthreads();
        // edge terms, branch-free (the regularised factor is finite for c2 = 0 when off > 0;
        // the select drops it below the cutoff either way)
#pragma unroll
        for (int s = 0; s < VRE_ES; ++s) {
            double v[TT];
#pragma unroll
            for (int tt = 0; tt < TT; ++tt) {
                const int t = min(t0 + tt, iend - 1);
                const NodeTerm A = sn[tt][ea[s]], B = sn[tt][eb[s]];
                const double cx = A.y * B.z - A.z * B.y;
                const double cy = A.z * B.x - A.x * B.z;
                const double cz = A.x * B.y - A.y * B.x;
                const double c2 = cx * cx + cy * cy + cz * cz;
                const double proj = ex[s] * (A.x * A.iv - B.x * B.iv) + ey[s] * (A.y * A.iv - B.y * B.iv) +
                                    ez[s] * (A.z * A.iv - B.z * B.iv);
                const double w = proj * fp_recip(k4 * (c2 + off2)) * (cx * st[3][t] + cy * st[4][t] + cz * st[5][t]);
                v[tt] = c2 > cut2 ? w : 0.0;
            }
            if (eok[s]) {
#pragma unroll
                for (int tt = 0; tt < TT; ++tt) se[tt][s * VRE_P + tid] = v[tt];
            }
        }
        __syncthreads();
        // panel sums
        if (col >= 0) {
#pragma unroll
            for (int tt = 0; tt < TT; ++tt) {
                const int t = t0 + tt;
                // every stored edge term is finite (the select above), so the unused references
                // (sign 0, slot 0) add exactly zero
                double v = 0.0;
#pragma unroll
                for (int k = 0; k < 4; ++k) v += rs[k] * se[tt][ri[k]];
                if (t < iend) {
                    g[(int64_t)t * ld] = TO(v);
                    if (g2) g2[(int64_t)t * ld] = float(v);
                }
            }
        }
    }
}

// Edge-pair kernel: lane l evaluates the node terms of its path n0 - n1 - n2 for
// TT targets at a time in registers and writes the two edge terms to LDS (double
// buffered, one barrier per pass); lane l then sums the signed edge terms of
// column l.  A missing n2 (single edge) or an idle lane repeats a node, whose
// zero-length edge gives c = 0 and a dropped term.
template <typename TO, int TM, int TT>
__global__ __launch_bounds__(256) void k_assemble_vrc(const int32_t *__restrict__ tab, const double *__restrict__ pv,
                                                      int64_t N, TargetView tv, TO *__restrict__ G, int64_t ld,
                                                      int64_t col0, float *__restrict__ G2, double off, double cut)
{
    __shared__ double st[6][TM];
    __shared__ double se[2][TT][2 * VRE_P];
    const int tid = threadIdx.x;
    const int32_t *T = tab + (int64_t)blockIdx.x * VRC_BLOCK_INTS;
    const int64_t i0 = (int64_t)blockIdx.y * TM;
    for (int t = tid; t < TM; t += blockDim.x) {
        const int64_t i = i0 + t;
        const bool ok = i < tv.M;
        st[0][t] = ok ? tv.x[i] : 0.0;
        st[1][t] = ok ? tv.y[i] : 0.0;
        st[2][t] = ok ? tv.z[i] : 0.0;
        st[3][t] = ok ? tv.nx[i] : 0.0;
        st[4][t] = ok ? tv.ny[i] : 0.0;
        st[5][t] = ok ? tv.nz[i] : 0.0;
    }
    auto node_xyz = [&](int32_t ref, double &x, double &y, double &z) {
        const int64_t j = ref >> 3;
        const int k = ref & 7;
        x = pv[(3 * k + 0) * N + j];
        y = pv[(3 * k + 1) * N + j];
        z = pv[(3 * k + 2) * N + j];
    };
    // the lane's path (idle lanes take lane 0's first node, or vertex 0 of panel 0)
    const int32_t q0 = T[VRC_OFF_NODE + tid], q1 = T[VRC_OFF_NODE + VRE_P + tid],
                  q2 = T[VRC_OFF_NODE + 2 * VRE_P + tid];
    const int32_t f0 = q0 >= 0 ? q0 : (T[VRC_OFF_NODE] >= 0 ? T[VRC_OFF_NODE] : 0);
    const int32_t f1 = q1 >= 0 ? q1 : f0, f2 = q2 >= 0 ? q2 : f1;
    double p0x, p0y, p0z, p1x, p1y, p1z, p2x, p2y, p2z;
    node_xyz(f0, p0x, p0y, p0z);
    node_xyz(f1, p1x, p1y, p1z);
    node_xyz(f2, p2x, p2y, p2z);
    const double e0x = p1x - p0x, e0y = p1y - p0y, e0z = p1z - p0z;
    const double e1x = p2x - p1x, e1y = p2y - p1y, e1z = p2z - p1z;
    const int32_t col = T[VRC_OFF_COL + tid];
    int ri[4];
    double rs[4];
    {
        const uint32_t r01 = (uint32_t)T[VRC_OFF_REF + 2 * tid], r23 = (uint32_t)T[VRC_OFF_REF + 2 * tid + 1];
        const uint32_t r[4] = {r01 & 0xFFFFu, r01 >> 16, r23 & 0xFFFFu, r23 >> 16};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool ok = col >= 0 && r[k] != VRE_NOREF;
            ri[k] = ok ? (int)(r[k] >> 1) : 0;
            rs[k] = ok ? ((r[k] & 1u) ? -1.0 : 1.0) : 0.0;
        }
    }
    const double off2 = off * off, cut2 = cut * cut, k4 = 4.0 * FP_PI;
    const int iend = (int)min((int64_t)TM, tv.M - i0);
    TO *g = G + i0 * ld + col0 + (col >= 0 ? col : 0);
    float *g2 = G2 ? G2 + i0 * ld + col0 + (col >= 0 ? col : 0) : nullptr;
    __syncthreads();
    int buf = 0;
    for (int t0 = 0; t0 < iend; t0 += TT, buf ^= 1) {
        // node and edge terms (a partial last pass repeats the last target: never stored twice)
#pragma unroll
        for (int tt = 0; tt < TT; ++tt) {
            const int t = min(t0 + tt, iend - 1);
            const double X = st[0][t], Y = st[1][t], Z = st[2][t];
            const double nx = st[3][t], ny = st[4][t], nz = st[5][t];
            const double ax = X - p0x, ay = Y - p0y, az = Z - p0z;
            const double bx = X - p1x, by = Y - p1y, bz = Z - p1z;
            const double cx_ = X - p2x, cy_ = Y - p2y, cz_ = Z - p2z;
            const double ia = fp_recip(fp_sqrt_nb(ax * ax + ay * ay + az * az) + off);
            const double ib = fp_recip(fp_sqrt_nb(bx * bx + by * by + bz * bz) + off);
            const double ic = fp_recip(fp_sqrt_nb(cx_ * cx_ + cy_ * cy_ + cz_ * cz_) + off);
            // edge n0 -> n1
            {
                const double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
                const double c2 = cx * cx + cy * cy + cz * cz;
                const double proj = e0x * (ax * ia - bx * ib) + e0y * (ay * ia - by * ib) + e0z * (az * ia - bz * ib);
                const double w = proj * fp_recip(k4 * (c2 + off2)) * (cx * nx + cy * ny + cz * nz);
                se[buf][tt][tid] = c2 > cut2 ? w : 0.0;
            }
            // edge n1 -> n2
            {
                const double cx = by * cz_ - bz * cy_, cy = bz * cx_ - bx * cz_, cz = bx * cy_ - by * cx_;
                const double c2 = cx * cx + cy * cy + cz * cz;
                const double proj = e1x * (bx * ib - cx_ * ic) + e1y * (by * ib - cy_ * ic) + e1z * (bz * ib - cz_ * ic);
                const double w = proj * fp_recip(k4 * (c2 + off2)) * (cx * nx + cy * ny + cz * nz);
                se[buf][tt][VRE_P + tid] = c2 > cut2 ? w : 0.0;
            }
        }
        __syncthreads();
        // panel sums; the next pass writes the other buffer, and the pass after it
        // comes after the next barrier, which every lane reaches after these reads
        if (col >= 0) {
#pragma unroll
            for (int tt = 0; tt < TT; ++tt) {
                const int t = t0 + tt;
                double v = 0.0;
#pragma unroll
                for (int k = 0; k < 4; ++k) v += rs[k] * se[buf][tt][ri[k]];
                if (t < iend) {
                    g[(int64_t)t * ld] = TO(v);
                    if (g2) g2[(int64_t)t * ld] = float(v);
                }
            }
        }
    }
}

// Launch the shared-term kernel for a body with block tables; returns 1 when the
// body has none (the caller then uses the panel-wise kernel).
template <typename T>
int fp_launch_assemble_edges(fp_ctx *ctx, const fp_body *b, const TargetView &tv, T *G, int64_t ld, int64_t col0,
                             float *G32, int *used)
{
    *used = 0;
    if (!b->d_eb || b->eb_nblocks == 0) return FP_OK;
    const char *env = getenv("FLOWPANEL_ASSEMBLY");
    if (env && strcmp(env, "panel") == 0) return FP_OK;
    *used = 1;
    if (tv.M == 0) return FP_OK;
    if (env && strcmp(env, "pairs") == 0) {
        if (!b->d_pb) {
            const int st = body_build_pair_blocks(ctx, const_cast<fp_body *>(b));
            if (st) return st;
        }
        // edge pairs (FLOWPANEL_ASSEMBLY=pairs: 40 % more vector instructions than the node terms in
        // LDS for 59 % fewer LDS instructions, 8.2 against 7.5 ms on the 30,000-panel loft);
        // FLOWPANEL_VRC_SHAPE for measurement
        const char *shape = getenv("FLOWPANEL_VRC_SHAPE");
        const std::string sh = shape ? shape : "64x2";
        ProfScope ps(ctx, sizeof(T) == 8 ? "assemble_vrc_f64" : "assemble_vrc_f32");
#define FP_VRC(TMV, TTV)                                                                                       \
        if (sh == #TMV "x" #TTV) {                                                                               \
            dim3 grid((unsigned)b->pb_nblocks, (unsigned)((tv.M + TMV - 1) / TMV));                            \
            hipLaunchKernelGGL((k_assemble_vrc<T, TMV, TTV>), grid, dim3(VRE_P), 0, ctx->stream, b->d_pb,       \
                               b->d_pv, b->N, tv, G, ld, col0, G32, b->kernel_offset, b->kernel_cutoff);      \
            FP_HIP(ctx, hipGetLastError());                                                                    \
            return FP_OK;                                                                                      \
        }
        FP_VRC(64, 2) FP_VRC(64, 1) FP_VRC(64, 4) FP_VRC(128, 2) FP_VRC(32, 2)
#undef FP_VRC
        return fp_fail(ctx, FP_ERR_ARG, "FLOWPANEL_VRC_SHAPE: one of 64x2, 64x1, 64x4, 128x2, 32x2");
    }
    // targets per workgroup tile x targets per pass (FLOWPANEL_VRE_SHAPE, for measurement)
    const char *shape = getenv("FLOWPANEL_VRE_SHAPE");
    const std::string sh = shape ? shape : "64x2";
    ProfScope ps(ctx, sizeof(T) == 8 ? "assemble_vre_f64" : "assemble_vre_f32");
#define FP_VRE(TMV, TTV)                                                                                       \
    if (sh == #TMV "x" #TTV) {                                                                                   \
        dim3 grid((unsigned)b->eb_nblocks, (unsigned)((tv.M + TMV - 1) / TMV));                              \
        hipLaunchKernelGGL((k_assemble_vre<T, TMV, TTV>), grid, dim3(VRE_P), 0, ctx->stream, b->d_eb, b->d_pv, \
                           b->N, tv, G, ld, col0, G32, b->kernel_offset, b->kernel_cutoff);                    \
        FP_HIP(ctx, hipGetLastError());                                                                        \
        return FP_OK;                                                                                          \
    }
    FP_VRE(64, 2) FP_VRE(128, 2) FP_VRE(32, 2) FP_VRE(64, 4) FP_VRE(128, 4) FP_VRE(64, 1)
#undef FP_VRE
    return fp_fail(ctx, FP_ERR_ARG, "FLOWPANEL_VRE_SHAPE: one of 64x2, 128x2, 32x2, 64x4, 128x4, 64x1");
}
template int fp_launch_assemble_edges<double>(fp_ctx *, const fp_body *, const TargetView &, double *, int64_t,
                                              int64_t, float *, int *);
template int fp_launch_assemble_edges<float>(fp_ctx *, const fp_body *, const TargetView &, float *, int64_t, int64_t,
                                             float *, int *);
