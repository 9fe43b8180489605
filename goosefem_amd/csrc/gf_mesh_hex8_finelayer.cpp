// Mesh::Hex8::FineLayer (include/GooseFEM/MeshHex8.h:613-2296): a Hex8 mesh whose element layers
// are finest in the middle (around y = Ly / 2) and coarsen towards the bottom and top faces by
// factors of three, in x or in z, through transition layers of four elements per coarse element.
// Host set-up code.  The node and element numbering must equal the reference's (it is the mesh's
// identity), so the layer bookkeeping follows the reference's steps (MeshHex8.h:640-850):
//   * the upper half of the layers is planned from the middle outwards: layer iy may coarsen in x
//     when three times its height fits below it, nelx is divisible by 3 nhx and the half height
//     nmin is not reached; it may then be followed directly by a layer that coarsens in z (when
//     nelz allows it and the height budget does).  Otherwise it may coarsen in z alone.  A
//     coarsening layer doubles its own height and triples the height of the layers above it and
//     the element size of itself and the layers above it in the coarsened direction;
//   * the full mesh mirrors that half about the middle layer;
//   * nodes run bottom to top; a node row is a grid of (lnx + 1) x (lnz + 1) nodes, x fastest.
//     Below the middle each layer contributes its bottom row and, when it refines, its
//     intermediate nodes (two per coarse element edge); from the middle up each layer contributes
//     its bottom row (three times finer in the refined direction in a transition layer), its
//     intermediate nodes, and the last layer its top row;
//   * elements run bottom to top, z outer, x inner, four per coarse element in a transition layer.
// Node sets are selected on the node coordinates (which lie exactly on the bounding planes) and
// ordered as the reference orders them: by node number, except the bottom and top faces, which
// run x outer, z inner (MeshHex8.h:1531-1573, 1851-1892).  The open side faces leave out the
// intermediate nodes of the two outermost layers, as the reference's layer loops do.
// nodesPeriodic pairs corners, then edges, then faces (Mesh.h:1577-1692).
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "gf_internal.hpp"

using namespace gf;

namespace {

struct FineLayerH8 {
    double h = 1.0, Lx = 0.0, Lz = 0.0;
    int64_t nelem = 0, nnode = 0;
    std::vector<int64_t> nhx, nhy, nhz, lnx, lnz, nnd, start_elem, start_node;
    std::vector<int> refine; // -1: plain layer, 0: transition in x, 2: transition in z
    int64_t layers() const { return static_cast<int64_t>(nhy.size()); }
    int64_t middle() const { return (layers() - 1) / 2; }
};

FineLayerH8 plan(int64_t nelx, int64_t nely, int64_t nelz, double h, int64_t nfine)
{
    GF_REQUIRE(nelx >= 1 && nely >= 1 && nelz >= 1, "nelx, nely, nelz >= 1");
    FineLayerH8 m;
    m.h = h;
    m.Lx = h * static_cast<double>(nelx);
    m.Lz = h * static_cast<double>(nelz);
    // half of the mesh, from the middle layer (index 0) outwards
    std::vector<int64_t> hx(nely, 1), hy(nely, 1), hz(nely, 1);
    std::vector<int> ref(nely, -1);
    auto scale = [&](std::vector<int64_t>& v, int64_t from) {
        for (int64_t j = from; j < nely; ++j) v[j] *= 3;
    };
    const int64_t nmin = (nely + 1) / 2;
    int64_t nf = nfine % 2 == 0 ? nfine / 2 + 1 : (nfine + 1) / 2;
    nf = std::max<int64_t>(1, std::min(nf, nmin));
    int64_t iy = nf, ntot = nf;
    while (iy < nely && ntot < nmin) {
        const bool room = 3 * hy[iy] <= ntot && ntot + hy[iy] < nmin;
        if (room && nelx % (3 * hx[iy]) == 0) {
            ref[iy] = 0;
            hy[iy] *= 2;
            scale(hy, iy + 1);
            scale(hx, iy);
            if (iy + 1 < nely && ntot + 2 * hy[iy] < nmin && nelz % (3 * hz[iy + 1]) == 0) {
                ntot += hy[iy];
                ++iy;
                ref[iy] = 2;
                hy[iy] = hy[iy - 1];
                scale(hz, iy);
            }
        }
        else if (room && nelz % (3 * hz[iy]) == 0) {
            ref[iy] = 2;
            hy[iy] *= 2;
            scale(hy, iy + 1);
            scale(hz, iy);
        }
        ntot += hy[iy];
        ++iy;
    }
    const int64_t half = iy, L = 2 * half - 1;
    m.nhx.resize(L);
    m.nhy.resize(L);
    m.nhz.resize(L);
    m.refine.resize(L);
    for (int64_t k = 0; k < L; ++k) {
        const int64_t src = k < half ? half - 1 - k : k - half + 1;
        m.nhx[k] = hx[src];
        m.nhy[k] = hy[src];
        m.nhz[k] = hz[src];
        m.refine[k] = ref[src];
    }
    m.lnx.resize(L);
    m.lnz.resize(L);
    for (int64_t k = 0; k < L; ++k) {
        m.lnx[k] = nelx / m.nhx[k];
        m.lnz[k] = nelz / m.nhz[k];
    }
    const int64_t mid = m.middle();
    m.nnd.assign(L + 1, 0);
    for (int64_t k = 0; k <= mid; ++k) m.nnd[k] = (m.lnx[k] + 1) * (m.lnz[k] + 1);
    for (int64_t k = mid; k < L; ++k) m.nnd[k + 1] = (m.lnx[k] + 1) * (m.lnz[k] + 1);
    m.start_elem.assign(L, 0);
    m.start_node.assign(L + 1, 0);
    for (int64_t k = 0; k < L; ++k) {
        m.start_elem[k] = m.nelem;
        const int64_t ax = m.lnx[k], az = m.lnz[k];
        // nodes of the bottom row (three times finer in the refined direction from the middle up)
        // and of the intermediate rows (two per coarse element edge in the refined direction)
        const int64_t f = k < mid ? 3 : 5;
        if (m.refine[k] == 0) m.nnode += (f * ax + 1) * (az + 1);
        else if (m.refine[k] == 2) m.nnode += (ax + 1) * (f * az + 1);
        else m.nnode += (ax + 1) * (az + 1);
        m.nelem += (m.refine[k] == -1 ? 1 : 4) * ax * az;
        m.start_node[k + 1] = m.nnode;
    }
    m.nnode += (m.lnx[L - 1] + 1) * (m.lnz[L - 1] + 1);
    return m;
}

// xt::linspace(0, L, n + 1)[i] = 0 + step * i
inline double lin(double Lt, int64_t n, int64_t i) { return (Lt / static_cast<double>(n)) * static_cast<double>(i); }

void coor_of(const FineLayerH8& m, double* c)
{
    const int64_t L = m.layers(), mid = m.middle();
    std::vector<double> y(L + 1, 0.0);
    for (int64_t k = 1; k <= L; ++k) y[k] = y[k - 1] + static_cast<double>(m.nhy[k - 1]) * m.h;
    int64_t n = 0;
    auto put = [&](double px, double py, double pz) {
        c[3 * n] = px;
        c[3 * n + 1] = py;
        c[3 * n + 2] = pz;
        ++n;
    };
    auto row = [&](int64_t k, double py) {
        for (int64_t iz = 0; iz <= m.lnz[k]; ++iz)
            for (int64_t ix = 0; ix <= m.lnx[k]; ++ix) put(lin(m.Lx, m.lnx[k], ix), py, lin(m.Lz, m.lnz[k], iz));
    };
    auto inter = [&](int64_t k) {
        const double dy = m.h * static_cast<double>(m.nhy[k] / 2);
        if (m.refine[k] == 0) {
            const double dx = m.h * static_cast<double>(m.nhx[k] / 3);
            for (int64_t iz = 0; iz <= m.lnz[k]; ++iz)
                for (int64_t ix = 0; ix < m.lnx[k]; ++ix)
                    for (int j = 1; j <= 2; ++j)
                        put(lin(m.Lx, m.lnx[k], ix) + dx * j, y[k] + dy, lin(m.Lz, m.lnz[k], iz));
        }
        else if (m.refine[k] == 2) {
            const double dz = m.h * static_cast<double>(m.nhz[k] / 3);
            for (int64_t iz = 0; iz < m.lnz[k]; ++iz)
                for (int j = 1; j <= 2; ++j)
                    for (int64_t ix = 0; ix <= m.lnx[k]; ++ix)
                        put(lin(m.Lx, m.lnx[k], ix), y[k] + dy, lin(m.Lz, m.lnz[k], iz) + dz * j);
        }
    };
    for (int64_t k = 0; k <= mid; ++k) {
        row(k, y[k]);
        if (k < mid) inter(k);
    }
    for (int64_t k = mid; k < L; ++k) {
        inter(k);
        row(k, y[k + 1]);
    }
}

void conn_of(const FineLayerH8& m, int64_t* conn)
{
    const int64_t L = m.layers(), mid = m.middle();
    int64_t e = 0;
    for (int64_t k = 0; k < L; ++k) {
        const int64_t X = m.lnx[k], Z = m.lnz[k];
        const int64_t bot = m.start_node[k], im = m.start_node[k] + m.nnd[k], top = m.start_node[k + 1];
        const bool lower = k <= mid;
        // node (i, j) of a row grid that starts at s and has w nodes along x
        auto at = [](int64_t s, int64_t w, int64_t i, int64_t j) { return s + i + j * w; };
        for (int64_t iz = 0; iz < Z; ++iz)
            for (int64_t ix = 0; ix < X; ++ix) {
                auto emit = [&](std::initializer_list<int64_t> v) {
                    std::copy(v.begin(), v.end(), conn + 8 * e);
                    ++e;
                };
                if (m.refine[k] == -1) {
                    auto b = [&](int64_t i, int64_t j) { return at(bot, X + 1, i, j); };
                    auto t = [&](int64_t i, int64_t j) { return at(top, X + 1, i, j); };
                    emit({b(ix, iz), b(ix + 1, iz), t(ix + 1, iz), t(ix, iz), b(ix, iz + 1), b(ix + 1, iz + 1),
                          t(ix + 1, iz + 1), t(ix, iz + 1)});
                }
                else if (m.refine[k] == 0) {
                    // x transition: every element is a quad in the x-y plane extruded over [iz, iz+1]
                    const int64_t wb = lower ? X + 1 : 3 * X + 1, wt = lower ? 3 * X + 1 : X + 1;
                    auto b = [&](int64_t i, int64_t j) { return at(bot, wb, i, j); };
                    auto q = [&](int64_t i, int64_t j) { return at(im, 2 * X, i, j); };
                    auto t = [&](int64_t i, int64_t j) { return at(top, wt, i, j); };
                    auto ext = [&](auto f0, auto f1, auto f2, auto f3) {
                        emit({f0(iz), f1(iz), f2(iz), f3(iz), f0(iz + 1), f1(iz + 1), f2(iz + 1), f3(iz + 1)});
                    };
                    const int64_t a = 2 * ix, c3 = 3 * ix;
                    if (lower) {
                        ext([&](int64_t j) { return b(ix, j); }, [&](int64_t j) { return b(ix + 1, j); },
                            [&](int64_t j) { return q(a + 1, j); }, [&](int64_t j) { return q(a, j); });
                        ext([&](int64_t j) { return b(ix + 1, j); }, [&](int64_t j) { return t(c3 + 3, j); },
                            [&](int64_t j) { return t(c3 + 2, j); }, [&](int64_t j) { return q(a + 1, j); });
                        ext([&](int64_t j) { return q(a, j); }, [&](int64_t j) { return q(a + 1, j); },
                            [&](int64_t j) { return t(c3 + 2, j); }, [&](int64_t j) { return t(c3 + 1, j); });
                        ext([&](int64_t j) { return b(ix, j); }, [&](int64_t j) { return q(a, j); },
                            [&](int64_t j) { return t(c3 + 1, j); }, [&](int64_t j) { return t(c3, j); });
                    }
                    else {
                        ext([&](int64_t j) { return b(c3, j); }, [&](int64_t j) { return b(c3 + 1, j); },
                            [&](int64_t j) { return q(a, j); }, [&](int64_t j) { return t(ix, j); });
                        ext([&](int64_t j) { return b(c3 + 1, j); }, [&](int64_t j) { return b(c3 + 2, j); },
                            [&](int64_t j) { return q(a + 1, j); }, [&](int64_t j) { return q(a, j); });
                        ext([&](int64_t j) { return b(c3 + 2, j); }, [&](int64_t j) { return b(c3 + 3, j); },
                            [&](int64_t j) { return t(ix + 1, j); }, [&](int64_t j) { return q(a + 1, j); });
                        ext([&](int64_t j) { return q(a, j); }, [&](int64_t j) { return q(a + 1, j); },
                            [&](int64_t j) { return t(ix + 1, j); }, [&](int64_t j) { return t(ix, j); });
                    }
                }
                else {
                    // z transition: all rows are X + 1 nodes wide; the fine side has 3 rows per
                    // coarse element, the intermediate rows 2
                    auto b = [&](int64_t i, int64_t j) { return at(bot, X + 1, i, j); };
                    auto q = [&](int64_t i, int64_t j) { return at(im, X + 1, i, j); };
                    auto t = [&](int64_t i, int64_t j) { return at(top, X + 1, i, j); };
                    const int64_t x0 = ix, x1 = ix + 1, a = 2 * iz, c3 = 3 * iz;
                    if (lower) {
                        emit({b(x0, iz), b(x0, iz + 1), b(x1, iz + 1), b(x1, iz), q(x0, a), q(x0, a + 1), q(x1, a + 1),
                              q(x1, a)});
                        emit({q(x0, a + 1), q(x1, a + 1), t(x1, c3 + 2), t(x0, c3 + 2), b(x0, iz + 1), b(x1, iz + 1),
                              t(x1, c3 + 3), t(x0, c3 + 3)});
                        emit({q(x0, a), q(x1, a), t(x1, c3 + 1), t(x0, c3 + 1), q(x0, a + 1), q(x1, a + 1),
                              t(x1, c3 + 2), t(x0, c3 + 2)});
                        emit({b(x0, iz), b(x1, iz), t(x1, c3), t(x0, c3), q(x0, a), q(x1, a), t(x1, c3 + 1),
                              t(x0, c3 + 1)});
                    }
                    else {
                        emit({b(x0, c3), b(x1, c3), t(x1, iz), t(x0, iz), b(x0, c3 + 1), b(x1, c3 + 1), q(x1, a),
                              q(x0, a)});
                        emit({b(x0, c3 + 1), b(x1, c3 + 1), q(x1, a), q(x0, a), b(x0, c3 + 2), b(x1, c3 + 2),
                              q(x1, a + 1), q(x0, a + 1)});
                        emit({b(x0, c3 + 2), b(x1, c3 + 2), q(x1, a + 1), q(x0, a + 1), b(x0, c3 + 3), b(x1, c3 + 3),
                              t(x1, iz + 1), t(x0, iz + 1)});
                        emit({q(x0, a), q(x1, a), t(x1, iz), t(x0, iz), q(x0, a + 1), q(x1, a + 1), t(x1, iz + 1),
                              t(x0, iz + 1)});
                    }
                }
            }
    }
}

// Named node set: "nodes" + any order of {Front, Back, Left, Right, Bottom, Top} + one of
// {"" (whole face), "Face" (face interior), "Edge", "OpenEdge" (edge without its corners),
// "Corner"}.  Front / Back: z = 0 / Lz; Left / Right: x = 0 / Lx; Bottom / Top: y = 0 / Ly.
bool named(const FineLayerH8& m, const std::string& name, std::vector<int64_t>& r)
{
    r.clear();
    const int64_t L = m.layers();
    if (name == "elementsMiddleLayer") {
        const int64_t k = m.middle();
        for (int64_t i = 0; i < m.lnx[k] * m.lnz[k]; ++i) r.push_back(m.start_elem[k] + i);
        return true;
    }
    if (name.rfind("nodes", 0) != 0) return false;
    std::string s = name.substr(5), kind;
    for (const char* suf : {"OpenEdge", "Edge", "Face", "Corner"}) {
        const std::string t(suf);
        if (s.size() > t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0) {
            kind = t;
            s = s.substr(0, s.size() - t.size());
            break;
        }
    }
    // fixed[a] = -1 (free), 0 (low plane) or 1 (high plane) for axes x, y, z
    int fixed[3] = {-1, -1, -1};
    int nfix = 0;
    static const struct { const char* w; int axis, side; } words[] = {
        {"Front", 2, 0}, {"Back", 2, 1}, {"Left", 0, 0}, {"Right", 0, 1}, {"Bottom", 1, 0}, {"Top", 1, 1}};
    if (s == "Origin") {
        fixed[0] = fixed[1] = fixed[2] = 0;
        nfix = 3;
        kind = "Corner";
    }
    else {
        while (!s.empty()) {
            bool hit = false;
            for (const auto& w : words) {
                const std::string t(w.w);
                if (s.compare(0, t.size(), t) == 0) {
                    if (fixed[w.axis] != -1) return false;
                    fixed[w.axis] = w.side;
                    ++nfix;
                    s = s.substr(t.size());
                    hit = true;
                    break;
                }
            }
            if (!hit) return false;
        }
    }
    const int need = kind.empty() || kind == "Face" ? 1 : (kind == "Corner" ? 3 : 2);
    if (nfix != need) return false;
    const bool open = kind == "Face" || kind == "OpenEdge";
    std::vector<double> c(static_cast<size_t>(3 * m.nnode));
    coor_of(m, c.data());
    double hi[3] = {m.Lx, 0.0, m.Lz};
    for (int64_t k = 0; k < L; ++k) hi[1] += static_cast<double>(m.nhy[k]) * m.h;
    const double tol = 1e-9 * m.h;
    auto on = [&](double v, int side, int a) { return std::abs(v - (side ? hi[a] : 0.0)) <= tol; };
    auto inside = [&](double v, int a) { return v > tol && v < hi[a] - tol; };
    // The reference's open Front / Back / Left / Right faces run over the node rows 1 .. L - 1 and
    // the intermediate nodes of the transition layers 1 .. L - 2 only (MeshHex8.h:1575-1850): the
    // intermediate nodes of the outermost layers 0 and L - 1, which lie inside those faces, are not
    // part of them, and so are not tied by nodesPeriodic either.
    auto outer_intermediate = [&](int64_t n) {
        for (int64_t k : {int64_t(0), L - 1})
            if (m.refine[k] != -1 && n >= m.start_node[k] + m.nnd[k] && n < m.start_node[k + 1]) return true;
        return false;
    };
    const bool side_face = kind == "Face" && fixed[1] == -1;
    for (int64_t n = 0; n < m.nnode; ++n) {
        bool ok = true;
        for (int a = 0; a < 3 && ok; ++a) {
            const double v = c[3 * n + a];
            if (fixed[a] != -1) ok = on(v, fixed[a], a);
            else if (open) ok = inside(v, a);
        }
        if (ok && side_face && outer_intermediate(n)) ok = false;
        if (ok) r.push_back(n);
    }
    // the bottom and top faces run x outer, z inner over their node row
    if (fixed[1] != -1 && fixed[0] == -1 && fixed[2] == -1) {
        const int64_t s0 = fixed[1] ? m.start_node[L] : m.start_node[0];
        const int64_t w = (fixed[1] ? m.lnx[L - 1] : m.lnx[0]) + 1;
        std::stable_sort(r.begin(), r.end(), [&](int64_t a, int64_t b) {
            const int64_t ia = (a - s0) % w, ib = (b - s0) % w;
            return ia != ib ? ia < ib : a < b;
        });
    }
    return true;
}

std::vector<int64_t> set_of(const FineLayerH8& m, const char* name)
{
    std::vector<int64_t> r;
    GF_REQUIRE(named(m, name, r), std::string("unknown FineLayer set ") + name);
    return r;
}

std::vector<int64_t> periodic(const FineLayerH8& m)
{
    std::vector<int64_t> r;
    auto pair = [&](int64_t a, int64_t b) {
        r.push_back(a);
        r.push_back(b);
    };
    auto zip = [&](const char* a, const char* b) {
        const auto u = set_of(m, a), v = set_of(m, b);
        GF_REQUIRE(u.size() == v.size(), "periodic node sets differ in size");
        for (size_t j = 0; j < u.size(); ++j) pair(u[j], v[j]);
    };
    const int64_t o = set_of(m, "nodesFrontBottomLeftCorner")[0];
    for (const char* c : {"nodesFrontBottomRightCorner", "nodesBackBottomRightCorner", "nodesBackBottomLeftCorner",
                          "nodesFrontTopLeftCorner", "nodesFrontTopRightCorner", "nodesBackTopRightCorner",
                          "nodesBackTopLeftCorner"})
        pair(o, set_of(m, c)[0]);
    for (const char* e : {"nodesBackBottomOpenEdge", "nodesBackTopOpenEdge", "nodesFrontTopOpenEdge"})
        zip("nodesFrontBottomOpenEdge", e);
    for (const char* e : {"nodesBottomRightOpenEdge", "nodesTopRightOpenEdge", "nodesTopLeftOpenEdge"})
        zip("nodesBottomLeftOpenEdge", e);
    for (const char* e : {"nodesFrontRightOpenEdge", "nodesBackRightOpenEdge", "nodesBackLeftOpenEdge"})
        zip("nodesFrontLeftOpenEdge", e);
    zip("nodesFrontFace", "nodesBackFace");
    zip("nodesLeftFace", "nodesRightFace");
    zip("nodesBottomFace", "nodesTopFace");
    return r;
}

} // namespace

extern "C" {

int gf_mesh_hex8_finelayer(int64_t nelx, int64_t nely, int64_t nelz, double h, int64_t nfine, int64_t* sizes,
                           double* coor, int64_t* conn)
{
    return guard([&] {
        const auto m = plan(nelx, nely, nelz, h, nfine);
        if (sizes) {
            sizes[0] = m.nelem;
            sizes[1] = m.nnode;
            sizes[2] = *std::max_element(m.lnx.begin(), m.lnx.end());
            int64_t ny = 0;
            for (int64_t v : m.nhy) ny += v;
            sizes[3] = ny;
            sizes[4] = *std::max_element(m.lnz.begin(), m.lnz.end());
        }
        if (coor) coor_of(m, coor);
        if (conn) conn_of(m, conn);
    });
}

int gf_mesh_hex8_finelayer_set(int64_t nelx, int64_t nely, int64_t nelz, double h, int64_t nfine, const char* name,
                               int64_t* out, int64_t* n)
{
    return guard([&] {
        GF_REQUIRE(name != nullptr, "name");
        const auto m = plan(nelx, nely, nelz, h, nfine);
        std::vector<int64_t> r;
        if (std::string(name) == "nodesPeriodic") r = periodic(m);
        else GF_REQUIRE(named(m, name, r), std::string("unknown FineLayer set ") + name);
        if (n) *n = static_cast<int64_t>(r.size());
        if (out) std::copy(r.begin(), r.end(), out);
    });
}

} // extern "C"
