// Mesh::Quad4::FineLayer (include/GooseFEM/MeshQuad4.h:200-1135): a Quad4 mesh whose element layers
// are finest in the middle (around y = Ly / 2) and coarsen towards the bottom and top edges by
// factors of three in x, through transition layers of four elements per coarse element.  Host
// set-up code.  The node and element numbering must equal the reference's (it is the mesh's
// identity), so the layer bookkeeping follows the reference's steps (MeshQuad4.h:900-1071):
//   * the upper half of the layers is planned from the middle outwards: layer iy may coarsen when
//     three times its height fits below it (3 nhy <= ntot), nelx is divisible by 3 nhx, and the
//     half height nmin is not reached; a coarsening layer doubles its own height and triples the
//     height of the layers above it and the width of itself and the layers above it;
//   * the full mesh mirrors that half about the middle layer;
//   * nodes run bottom to top: below the middle each layer contributes its bottom row and, when it
//     refines, its two intermediate nodes per coarse element; from the middle up each layer
//     contributes its intermediate nodes (when it coarsens) and its top row;
//   * elements run bottom to top, left to right, four per coarse element in a transition layer.
#include <string>
#include <vector>

#include "gf_internal.hpp"

using namespace gf;

namespace {

struct FineLayerQ4 {
    double h = 1.0, Lx = 0.0;
    int64_t nelem = 0, nnode = 0;
    std::vector<int64_t> nhx, nhy, lnx, nnd, start_elem, start_node;
    std::vector<int> refine; // 0: transition layer, -1: plain layer
    int64_t layers() const { return static_cast<int64_t>(nhy.size()); }
    int64_t middle() const { return (layers() - 1) / 2; }
};

FineLayerQ4 plan(int64_t nelx, int64_t nely, double h, int64_t nfine)
{
    GF_REQUIRE(nelx >= 1 && nely >= 1, "nelx, nely >= 1");
    FineLayerQ4 m;
    m.h = h;
    m.Lx = h * static_cast<double>(nelx);
    // half of the mesh, from the middle layer (index 0) outwards
    std::vector<int64_t> hx(nely, 1), hy(nely, 1);
    std::vector<int> ref(nely, -1);
    const int64_t nmin = (nely + 1) / 2;                   // nely / 2 for even nely, (nely + 1) / 2 for odd
    int64_t nf = nfine % 2 == 0 ? nfine / 2 + 1 : (nfine + 1) / 2; // fine layers in this half
    nf = std::max<int64_t>(1, std::min(nf, nmin));
    int64_t iy = nf, ntot = nf;
    while (iy < nely && ntot < nmin) {
        if (3 * hy[iy] <= ntot && nelx % (3 * hx[iy]) == 0 && ntot + hy[iy] < nmin) {
            ref[iy] = 0;
            hy[iy] *= 2;
            for (int64_t j = iy + 1; j < nely; ++j) hy[j] *= 3;
            for (int64_t j = iy; j < nely; ++j) hx[j] *= 3;
        }
        ntot += hy[iy];
        ++iy;
    }
    const int64_t half = iy, L = 2 * half - 1;
    // mirror: the lower half reversed, then the upper half without the shared middle layer
    m.nhx.resize(L);
    m.nhy.resize(L);
    m.refine.resize(L);
    for (int64_t k = 0; k < L; ++k) {
        const int64_t src = k < half ? half - 1 - k : k - half + 1;
        m.nhx[k] = hx[src];
        m.nhy[k] = hy[src];
        m.refine[k] = ref[src];
    }
    m.lnx.resize(L);
    for (int64_t k = 0; k < L; ++k) m.lnx[k] = nelx / m.nhx[k];
    // nodes of the main node rows: a layer's bottom row below the middle, its top row above
    const int64_t mid = m.middle();
    m.nnd.assign(L + 1, 0);
    for (int64_t k = 0; k <= mid; ++k) m.nnd[k] = m.lnx[k] + 1;
    for (int64_t k = mid; k < L; ++k) m.nnd[k + 1] = m.lnx[k] + 1;
    // first node and element of each layer
    m.start_elem.assign(L, 0);
    m.start_node.assign(L + 1, 0);
    for (int64_t k = 0; k < L; ++k) {
        m.start_elem[k] = m.nelem;
        const bool tr = m.refine[k] == 0;
        // below the middle: bottom row (+ two intermediate nodes per coarse element); from the
        // middle up: bottom row, which is three times finer in a transition layer (+ intermediates)
        if (k < mid) m.nnode += tr ? 3 * m.lnx[k] + 1 : m.lnx[k] + 1;
        else m.nnode += tr ? 5 * m.lnx[k] + 1 : m.lnx[k] + 1;
        m.nelem += tr ? 4 * m.lnx[k] : m.lnx[k];
        m.start_node[k + 1] = m.nnode;
    }
    m.nnode += m.lnx[L - 1] + 1; // the top row
    return m;
}

// xt::linspace(0, Lx, n + 1)[i] = 0 + step * i
inline double row_x(const FineLayerQ4& m, int64_t n, int64_t i)
{
    const double step = m.Lx / static_cast<double>(n);
    return step * static_cast<double>(i);
}

void coor_of(const FineLayerQ4& m, double* c)
{
    const int64_t L = m.layers(), mid = m.middle();
    std::vector<double> y(L + 1, 0.0);
    for (int64_t k = 1; k <= L; ++k) y[k] = y[k - 1] + static_cast<double>(m.nhy[k - 1]) * m.h;
    int64_t n = 0;
    auto put = [&](double x, double yy) {
        c[2 * n] = x;
        c[2 * n + 1] = yy;
        ++n;
    };
    auto row = [&](int64_t k, double yy) {
        for (int64_t ix = 0; ix <= m.lnx[k]; ++ix) put(row_x(m, m.lnx[k], ix), yy);
    };
    auto intermediate = [&](int64_t k) {
        const double dx = m.h * static_cast<double>(m.nhx[k] / 3), dy = m.h * static_cast<double>(m.nhy[k] / 2);
        for (int64_t ix = 0; ix < m.lnx[k]; ++ix)
            for (int j = 1; j <= 2; ++j) put(row_x(m, m.lnx[k], ix) + dx * static_cast<double>(j), y[k] + dy);
    };
    for (int64_t k = 0; k <= mid; ++k) {
        row(k, y[k]);
        if (k < mid && m.refine[k] == 0) intermediate(k);
    }
    for (int64_t k = mid; k < L; ++k) {
        if (m.refine[k] == 0) intermediate(k);
        row(k, y[k + 1]);
    }
}

void conn_of(const FineLayerQ4& m, int64_t* c)
{
    const int64_t L = m.layers(), mid = m.middle();
    int64_t e = 0;
    auto put = [&](int64_t a, int64_t b, int64_t cc, int64_t d) {
        c[4 * e] = a;
        c[4 * e + 1] = b;
        c[4 * e + 2] = cc;
        c[4 * e + 3] = d;
        ++e;
    };
    for (int64_t k = 0; k < L; ++k) {
        const int64_t bot = m.start_node[k], mi = m.start_node[k] + m.nnd[k], top = m.start_node[k + 1];
        for (int64_t ix = 0; ix < m.lnx[k]; ++ix) {
            if (m.refine[k] != 0) {
                put(bot + ix, bot + ix + 1, top + ix + 1, top + ix);
            }
            else if (k <= mid) {
                // refining upwards: bottom, top-right, top-centre, top-left
                put(bot + ix, bot + ix + 1, mi + 2 * ix + 1, mi + 2 * ix);
                put(bot + ix + 1, top + 3 * ix + 3, top + 3 * ix + 2, mi + 2 * ix + 1);
                put(mi + 2 * ix, mi + 2 * ix + 1, top + 3 * ix + 2, top + 3 * ix + 1);
                put(bot + ix, mi + 2 * ix, top + 3 * ix + 1, top + 3 * ix);
            }
            else {
                // coarsening upwards: lower-left, lower-centre, lower-right, upper
                put(bot + 3 * ix, bot + 3 * ix + 1, mi + 2 * ix, top + ix);
                put(bot + 3 * ix + 1, bot + 3 * ix + 2, mi + 2 * ix + 1, mi + 2 * ix);
                put(bot + 3 * ix + 2, bot + 3 * ix + 3, top + ix + 1, mi + 2 * ix + 1);
                put(mi + 2 * ix, mi + 2 * ix + 1, top + ix + 1, top + ix);
            }
        }
    }
}

// Named sets (MeshQuad4.h:287-292, 747-880; Mesh.h:511-540 for nodesPeriodic)
bool named(const FineLayerQ4& m, const std::string& name, std::vector<int64_t>& r)
{
    r.clear();
    const int64_t L = m.layers(), mid = m.middle();
    const int64_t BL = m.start_node[0], BR = m.start_node[0] + m.lnx[0], TL = m.start_node[L],
                  TR = m.start_node[L] + m.lnx[L - 1];
    auto range = [&](int64_t first, int64_t count) {
        for (int64_t i = 0; i < count; ++i) r.push_back(first + i);
    };
    // left / right main nodes of node row k (rows 0..L), and of the open edge (rows 1..L-1)
    auto left = [&](int64_t k) { return m.start_node[k]; };
    auto right = [&](int64_t k) { return m.start_node[k] + m.lnx[k <= mid ? k : k - 1]; };
    if (name == "nodesBottomEdge") range(BL, m.lnx[0] + 1);
    else if (name == "nodesTopEdge") range(TL, m.lnx[L - 1] + 1);
    else if (name == "nodesBottomOpenEdge") range(BL + 1, m.lnx[0] - 1);
    else if (name == "nodesTopOpenEdge") range(TL + 1, m.lnx[L - 1] - 1);
    else if (name == "nodesLeftEdge") for (int64_t k = 0; k <= L; ++k) r.push_back(left(k));
    else if (name == "nodesRightEdge") for (int64_t k = 0; k <= L; ++k) r.push_back(right(k));
    else if (name == "nodesLeftOpenEdge") for (int64_t k = 1; k < L; ++k) r.push_back(left(k));
    else if (name == "nodesRightOpenEdge") for (int64_t k = 1; k < L; ++k) r.push_back(right(k));
    else if (name == "nodesBottomLeftCorner" || name == "nodesLeftBottomCorner" || name == "nodesOrigin") r.push_back(BL);
    else if (name == "nodesBottomRightCorner" || name == "nodesRightBottomCorner") r.push_back(BR);
    else if (name == "nodesTopLeftCorner" || name == "nodesLeftTopCorner") r.push_back(TL);
    else if (name == "nodesTopRightCorner" || name == "nodesRightTopCorner") r.push_back(TR);
    else if (name == "elementsMiddleLayer") range(m.start_elem[mid], m.lnx[mid]);
    else if (name == "nodesPeriodic") {
        // corners, then bottom / top open edges, then left / right open edges
        auto pair = [&](int64_t a, int64_t b) {
            r.push_back(a);
            r.push_back(b);
        };
        pair(BL, BR);
        pair(BL, TR);
        pair(BL, TL);
        for (int64_t i = 1; i < m.lnx[0]; ++i) pair(BL + i, TL + i);
        for (int64_t k = 1; k < L; ++k) pair(left(k), right(k));
    }
    else return false;
    return true;
}

} // namespace

extern "C" {

int gf_mesh_quad4_finelayer(int64_t nelx, int64_t nely, double h, int64_t nfine, int64_t* sizes, double* coor,
                            int64_t* conn)
{
    return guard([&] {
        const auto m = plan(nelx, nely, h, nfine);
        if (sizes) {
            sizes[0] = m.nelem;
            sizes[1] = m.nnode;
            sizes[2] = *std::max_element(m.lnx.begin(), m.lnx.end());
            int64_t ny = 0;
            for (int64_t v : m.nhy) ny += v;
            sizes[3] = ny;
        }
        if (coor) coor_of(m, coor);
        if (conn) conn_of(m, conn);
    });
}

int gf_mesh_quad4_finelayer_set(int64_t nelx, int64_t nely, double h, int64_t nfine, const char* name, int64_t* out,
                                int64_t* n)
{
    return guard([&] {
        GF_REQUIRE(name != nullptr, "name");
        const auto m = plan(nelx, nely, h, nfine);
        std::vector<int64_t> r;
        GF_REQUIRE(named(m, name, r), std::string("unknown FineLayer set ") + name);
        if (n) *n = static_cast<int64_t>(r.size());
        if (out) std::copy(r.begin(), r.end(), out);
    });
}

// FineLayer::roll (MeshQuad4.h:539-564): the element map that shifts every layer by n elements of
// the coarsest layer in x (periodically); a layer k times finer shifts by k n elements, a
// transition layer (four elements per coarse element) by 4 k n.  out[nelem]: new_elemvar =
// elemvar[out].
int gf_mesh_quad4_finelayer_roll(int64_t nelx, int64_t nely, double h, int64_t nfine, int64_t n, int64_t* out)
{
    return guard([&] {
        GF_REQUIRE(n >= 0 && out != nullptr, "roll: n >= 0 and an output");
        const auto m = plan(nelx, nely, h, nfine);
        for (int64_t k = 0; k < m.layers(); ++k) {
            const int64_t per = m.refine[k] == -1 ? 1 : 4;
            const int64_t nel = per * m.lnx[k], shift = (n * (m.lnx[k] / m.lnx[0]) * per) % nel;
            for (int64_t i = 0; i < nel; ++i) out[m.start_elem[k] + (i + shift) % nel] = m.start_elem[k] + i;
        }
    });
}

} // extern "C"
