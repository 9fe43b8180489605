// Mesh bookkeeping and periodic tyings on the host (set-up only).  Integer outputs are bit-exact
// with the reference: Regular meshes (include/GooseFEM/MeshQuad4.h:43-187, MeshHex8.h:41-609),
// dofs / dofsPeriodic / renumber (Mesh.h:93-175, 291-304), nodesPeriodic (Mesh.h:511-540,
// 1577-1692), coordination / elem2node / nodaltyings (Mesh.h:2462-2551, 2839-2861), Reorder
// (Mesh.h:2384-2454) and Tyings::Control / Periodic (TyingsPeriodic.h:82-439).
#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>

#include "gf_internal.hpp"

namespace gf {

thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

std::vector<int32_t> narrow(const int64_t* a, size_t n, const char* what)
{
    std::vector<int32_t> r(n);
    for (size_t i = 0; i < n; ++i) {
        if (a[i] < 0 || a[i] > INT32_MAX)
            throw Error(GF_EINVAL, std::string(what) + ": index out of int32 range");
        r[i] = static_cast<int32_t>(a[i]);
    }
    return r;
}

namespace {

// Regular-grid node coordinate: linspace(0, h*nel, nel+1) as i * (h*nel/nel); exact for h = 1.
inline double grid(int64_t i, int64_t nel, double h)
{
    double stop = h * static_cast<double>(nel);
    return static_cast<double>(i) * (stop / static_cast<double>(nel));
}

struct Grid3 {
    int64_t nx, ny, nz;
    int64_t id(int64_t x, int64_t y, int64_t z) const { return x + y * (nx + 1) + z * (nx + 1) * (ny + 1); }
};

// Named node set of a Regular mesh.  2-D: z ignored (nz <= 0).
bool nodeset(int64_t nx, int64_t ny, int64_t nz, const std::string& name, std::vector<int64_t>& r)
{
    r.clear();
    if (nz <= 0) {
        auto id = [&](int64_t x, int64_t y) { return x + y * (nx + 1); };
        const std::map<std::string, std::pair<int, bool>> edges = {
            {"nodesBottomEdge", {0, false}}, {"nodesTopEdge", {1, false}},
            {"nodesLeftEdge", {2, false}},   {"nodesRightEdge", {3, false}},
            {"nodesBottomOpenEdge", {0, true}}, {"nodesTopOpenEdge", {1, true}},
            {"nodesLeftOpenEdge", {2, true}},   {"nodesRightOpenEdge", {3, true}}};
        auto it = edges.find(name);
        if (it != edges.end()) {
            int w = it->second.first;
            bool open = it->second.second;
            int64_t n = (w < 2) ? nx : ny;
            for (int64_t i = open ? 1 : 0; i < (open ? n : n + 1); ++i) {
                if (w == 0) r.push_back(id(i, 0));
                if (w == 1) r.push_back(id(i, ny));
                if (w == 2) r.push_back(id(0, i));
                if (w == 3) r.push_back(id(nx, i));
            }
            return true;
        }
        const std::map<std::string, int64_t> corners = {
            {"nodesBottomLeftCorner", id(0, 0)},   {"nodesBottomRightCorner", id(nx, 0)},
            {"nodesTopLeftCorner", id(0, ny)},     {"nodesTopRightCorner", id(nx, ny)},
            {"nodesLeftBottomCorner", id(0, 0)},   {"nodesLeftTopCorner", id(0, ny)},
            {"nodesRightBottomCorner", id(nx, 0)}, {"nodesRightTopCorner", id(nx, ny)},
            {"nodesOrigin", id(0, 0)}};
        auto jt = corners.find(name);
        if (jt != corners.end()) {
            r.push_back(jt->second);
            return true;
        }
        return false;
    }
    Grid3 g{nx, ny, nz};
    // faces: x fixed (Left/Right), y fixed (Bottom/Top), z fixed (Front/Back)
    auto full_face = [&](char axis, int64_t v, bool open) {
        // iteration order of MeshHex8.h:133-294: (Front/Back) y outer x inner; (Left/Right) z outer
        // y inner; (Bottom/Top) z outer x inner
        int64_t a0 = open ? 1 : 0;
        if (axis == 'z') {
            for (int64_t y = a0; y < (open ? ny : ny + 1); ++y)
                for (int64_t x = a0; x < (open ? nx : nx + 1); ++x) r.push_back(g.id(x, y, v));
        }
        else if (axis == 'x') {
            for (int64_t z = a0; z < (open ? nz : nz + 1); ++z)
                for (int64_t y = a0; y < (open ? ny : ny + 1); ++y) r.push_back(g.id(v, y, z));
        }
        else {
            for (int64_t z = a0; z < (open ? nz : nz + 1); ++z)
                for (int64_t x = a0; x < (open ? nx : nx + 1); ++x) r.push_back(g.id(x, v, z));
        }
    };
    const std::map<std::string, std::pair<char, int>> faces = {
        {"Front", {'z', 0}}, {"Back", {'z', 1}}, {"Left", {'x', 0}},
        {"Right", {'x', 1}}, {"Bottom", {'y', 0}}, {"Top", {'y', 1}}};
    auto val = [&](char axis, int hi) { return hi ? (axis == 'x' ? nx : axis == 'y' ? ny : nz) : 0; };
    for (auto& f : faces) {
        if (name == "nodes" + f.first) {
            full_face(f.second.first, val(f.second.first, f.second.second), false);
            return true;
        }
        if (name == "nodes" + f.first + "Face") {
            full_face(f.second.first, val(f.second.first, f.second.second), true);
            return true;
        }
    }
    // edges: two faces intersect; the free axis is the one not fixed by either face
    const std::vector<std::string> order = {"Front", "Back", "Bottom", "Top", "Left", "Right"};
    for (auto& a : order)
        for (auto& b : order) {
            auto fa = faces.at(a), fb = faces.at(b);
            if (fa.first == fb.first) continue;
            for (int open = 0; open < 2; ++open) {
                std::string nm = "nodes" + a + b + (open ? "OpenEdge" : "Edge");
                if (nm != name) continue;
                char free = 'x';
                for (char c : {'x', 'y', 'z'})
                    if (c != fa.first && c != fb.first) free = c;
                int64_t fix[3] = {0, 0, 0};
                fix[fa.first - 'x'] = val(fa.first, fa.second);
                fix[fb.first - 'x'] = val(fb.first, fb.second);
                int64_t nf = free == 'x' ? nx : free == 'y' ? ny : nz;
                for (int64_t i = open ? 1 : 0; i < (open ? nf : nf + 1); ++i) {
                    int64_t c[3] = {fix[0], fix[1], fix[2]};
                    c[free - 'x'] = i;
                    r.push_back(g.id(c[0], c[1], c[2]));
                }
                return true;
            }
        }
    // corners: three faces
    for (auto& a : {"Front", "Back"})
        for (auto& b : {"Bottom", "Top"})
            for (auto& c : {"Left", "Right"}) {
                std::string nm = std::string("nodes") + a + b + c + "Corner";
                if (nm != name) continue;
                r.push_back(g.id(std::string(c) == "Right" ? nx : 0, std::string(b) == "Top" ? ny : 0,
                                 std::string(a) == "Back" ? nz : 0));
                return true;
            }
    if (name == "nodesOrigin") {
        r.push_back(0);
        return true;
    }
    return false;
}

std::vector<int64_t> open_edge(int64_t nx, int64_t ny, int64_t nz, const char* name)
{
    std::vector<int64_t> r;
    nodeset(nx, ny, nz, name, r);
    return r;
}

std::vector<int64_t> periodic_pairs(int64_t nx, int64_t ny, int64_t nz)
{
    std::vector<int64_t> r;
    auto pair = [&](int64_t a, int64_t b) {
        r.push_back(a);
        r.push_back(b);
    };
    if (nz <= 0) {
        // Mesh.h:511-540
        int64_t BL = 0, BR = nx, TL = ny * (nx + 1), TR = ny * (nx + 1) + nx;
        pair(BL, BR);
        pair(BL, TR);
        pair(BL, TL);
        for (int64_t i = 1; i < nx; ++i) pair(i, i + ny * (nx + 1));
        for (int64_t i = 1; i < ny; ++i) pair(i * (nx + 1), i * (nx + 1) + nx);
        return r;
    }
    // Mesh.h:1577-1692: corners, then edges, then faces
    Grid3 g{nx, ny, nz};
    int64_t o = g.id(0, 0, 0);
    pair(o, g.id(nx, 0, 0));
    pair(o, g.id(nx, 0, nz));
    pair(o, g.id(0, 0, nz));
    pair(o, g.id(0, ny, 0));
    pair(o, g.id(nx, ny, 0));
    pair(o, g.id(nx, ny, nz));
    pair(o, g.id(0, ny, nz));
    auto E = [&](const char* n) { return open_edge(nx, ny, nz, n); };
    auto zip = [&](const std::vector<int64_t>& a, const std::vector<int64_t>& b) {
        for (size_t j = 0; j < a.size(); ++j) pair(a[j], b[j]);
    };
    auto froBot = E("nodesFrontBottomOpenEdge");
    zip(froBot, E("nodesBackBottomOpenEdge"));
    zip(froBot, E("nodesBackTopOpenEdge"));
    zip(froBot, E("nodesFrontTopOpenEdge"));
    auto botLft = E("nodesBottomLeftOpenEdge");
    zip(botLft, E("nodesBottomRightOpenEdge"));
    zip(botLft, E("nodesTopRightOpenEdge"));
    zip(botLft, E("nodesTopLeftOpenEdge"));
    auto froLft = E("nodesFrontLeftOpenEdge");
    zip(froLft, E("nodesFrontRightOpenEdge"));
    zip(froLft, E("nodesBackRightOpenEdge"));
    zip(froLft, E("nodesBackLeftOpenEdge"));
    zip(E("nodesFrontFace"), E("nodesBackFace"));
    zip(E("nodesLeftFace"), E("nodesRightFace"));
    zip(E("nodesBottomFace"), E("nodesTopFace"));
    return r;
}

std::vector<int64_t> sorted_unique(const int64_t* a, size_t n)
{
    std::vector<int64_t> v(a, a + n);
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    return v;
}

void renumber(const int64_t* in, int64_t n, int64_t* out)
{
    auto u = sorted_unique(in, static_cast<size_t>(n));
    std::vector<int64_t> idx(u.empty() ? 0 : u.back() + 1, 0);
    for (size_t i = 0; i < u.size(); ++i) idx[u[i]] = static_cast<int64_t>(i);
    for (int64_t i = 0; i < n; ++i) out[i] = idx[in[i]];
}

template <class Rows>
void write_csr(const Rows& rows, int64_t* ptr, int64_t* idx, int64_t* nnz)
{
    int64_t k = 0;
    for (auto& r : rows) k += static_cast<int64_t>(r.size());
    if (nnz) *nnz = k;
    if (ptr) {
        ptr[0] = 0;
        for (size_t i = 0; i < rows.size(); ++i) ptr[i + 1] = ptr[i] + static_cast<int64_t>(rows[i].size());
    }
    if (idx) {
        k = 0;
        for (auto& r : rows)
            for (auto v : r) idx[k++] = v;
    }
}

std::vector<std::vector<int64_t>> nodaltyings(const int64_t* dofs, int64_t nnode, int ndim)
{
    int64_t ndof = 0;
    for (int64_t i = 0; i < nnode * ndim; ++i) ndof = std::max(ndof, dofs[i] + 1);
    std::vector<std::vector<int64_t>> per(ndof);
    for (int64_t n = 0; n < nnode; ++n)
        for (int i = 0; i < ndim; ++i) per[dofs[n * ndim + i]].push_back(n);
    std::vector<std::vector<int64_t>> r(nnode);
    for (int64_t n = 0; n < nnode; ++n) {
        r[n] = per[dofs[n * ndim]];
        std::sort(r[n].begin(), r[n].end());
    }
    return r;
}

} // namespace
} // namespace gf

using namespace gf;

extern "C" {

const char* gf_last_error(void) { return g_last_error.c_str(); }

int gf_version(int* major, int* minor, int* patch)
{
    if (major) *major = 0;
    if (minor) *minor = 1;
    if (patch) *patch = 0;
    return GF_OK;
}

int gf_stream_sync(void* stream)
{
    return guard([&] { GF_HIP(hipStreamSynchronize(S(stream))); });
}

int gf_device_malloc(void** ptr, int64_t bytes)
{
    return guard([&] {
        GF_REQUIRE(ptr && bytes >= 0, "invalid argument");
        *ptr = nullptr;
        if (bytes) GF_HIP(hipMalloc(ptr, static_cast<size_t>(bytes)));
    });
}

int gf_device_free(void* ptr)
{
    return guard([&] {
        if (ptr) GF_HIP(hipFree(ptr));
    });
}

int gf_memcpy_h2d(void* dst, const void* src, int64_t bytes)
{
    return guard([&] {
        if (bytes) GF_HIP(hipMemcpy(dst, src, static_cast<size_t>(bytes), hipMemcpyHostToDevice));
    });
}

int gf_memcpy_d2h(void* dst, const void* src, int64_t bytes)
{
    return guard([&] {
        if (bytes) GF_HIP(hipMemcpy(dst, src, static_cast<size_t>(bytes), hipMemcpyDeviceToHost));
    });
}

int gf_memcpy_d2d(void* dst, const void* src, int64_t bytes)
{
    return guard([&] {
        if (bytes) GF_HIP(hipMemcpy(dst, src, static_cast<size_t>(bytes), hipMemcpyDeviceToDevice));
    });
}

int gf_mesh_quad4_regular(int64_t nx, int64_t ny, double h, double* coor, int64_t* conn)
{
    return guard([&] {
        GF_REQUIRE(nx >= 1 && ny >= 1, "nelx, nely >= 1");
        if (coor) {
            int64_t k = 0;
            for (int64_t iy = 0; iy <= ny; ++iy)
                for (int64_t ix = 0; ix <= nx; ++ix, ++k) {
                    coor[2 * k] = grid(ix, nx, h);
                    coor[2 * k + 1] = grid(iy, ny, h);
                }
        }
        if (conn) {
            for (int64_t iy = 0, e = 0; iy < ny; ++iy)
                for (int64_t ix = 0; ix < nx; ++ix, ++e) {
                    int64_t n0 = iy * (nx + 1) + ix;
                    conn[4 * e + 0] = n0;
                    conn[4 * e + 1] = n0 + 1;
                    conn[4 * e + 2] = n0 + nx + 2;
                    conn[4 * e + 3] = n0 + nx + 1;
                }
        }
    });
}

int gf_mesh_hex8_regular(int64_t nx, int64_t ny, int64_t nz, double h, double* coor, int64_t* conn)
{
    return guard([&] {
        GF_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, "nelx, nely, nelz >= 1");
        Grid3 g{nx, ny, nz};
        if (coor) {
            int64_t k = 0;
            for (int64_t iz = 0; iz <= nz; ++iz)
                for (int64_t iy = 0; iy <= ny; ++iy)
                    for (int64_t ix = 0; ix <= nx; ++ix, ++k) {
                        coor[3 * k] = grid(ix, nx, h);
                        coor[3 * k + 1] = grid(iy, ny, h);
                        coor[3 * k + 2] = grid(iz, nz, h);
                    }
        }
        if (conn) {
            int64_t e = 0;
            for (int64_t iz = 0; iz < nz; ++iz)
                for (int64_t iy = 0; iy < ny; ++iy)
                    for (int64_t ix = 0; ix < nx; ++ix, ++e) {
                        int64_t* c = conn + 8 * e;
                        c[0] = g.id(ix, iy, iz);
                        c[1] = g.id(ix + 1, iy, iz);
                        c[2] = g.id(ix + 1, iy + 1, iz);
                        c[3] = g.id(ix, iy + 1, iz);
                        c[4] = g.id(ix, iy, iz + 1);
                        c[5] = g.id(ix + 1, iy, iz + 1);
                        c[6] = g.id(ix + 1, iy + 1, iz + 1);
                        c[7] = g.id(ix, iy + 1, iz + 1);
                    }
        }
    });
}

int gf_mesh_nodeset(int64_t nx, int64_t ny, int64_t nz, const char* name, int64_t* out, int64_t* n)
{
    return guard([&] {
        GF_REQUIRE(name != nullptr, "name");
        std::vector<int64_t> r;
        GF_REQUIRE(nodeset(nx, ny, nz, name, r), std::string("unknown node set ") + name);
        if (n) *n = static_cast<int64_t>(r.size());
        if (out) std::copy(r.begin(), r.end(), out);
    });
}

int gf_mesh_nodes_periodic(int64_t nx, int64_t ny, int64_t nz, int64_t* out, int64_t* nties)
{
    return guard([&] {
        auto r = periodic_pairs(nx, ny, nz);
        if (nties) *nties = static_cast<int64_t>(r.size() / 2);
        if (out) std::copy(r.begin(), r.end(), out);
    });
}

int gf_mesh_dofs(int64_t nnode, int ndim, int64_t* dofs)
{
    return guard([&] {
        for (int64_t i = 0; i < nnode * ndim; ++i) dofs[i] = i;
    });
}

int gf_mesh_dofs_periodic(int64_t nnode, int ndim, const int64_t* ties, int64_t nties, int64_t* dofs)
{
    return guard([&] {
        std::vector<int64_t> d(static_cast<size_t>(nnode * ndim));
        std::iota(d.begin(), d.end(), 0);
        for (int j = 0; j < ndim; ++j)
            for (int64_t t = 0; t < nties; ++t) {
                GF_REQUIRE(ties[2 * t] < nnode && ties[2 * t + 1] < nnode, "tie node out of range");
                d[ties[2 * t + 1] * ndim + j] = d[ties[2 * t] * ndim + j];
            }
        renumber(d.data(), nnode * ndim, dofs);
    });
}

int gf_mesh_renumber(const int64_t* in, int64_t n, int64_t* out)
{
    return guard([&] { renumber(in, n, out); });
}

int gf_mesh_coordination(const int64_t* conn, int64_t n, int64_t nvalues, int64_t* out)
{
    return guard([&] {
        std::fill(out, out + nvalues, 0);
        for (int64_t i = 0; i < n; ++i) {
            GF_REQUIRE(conn[i] >= 0 && conn[i] < nvalues, "value out of range");
            out[conn[i]] += 1;
        }
    });
}

int gf_mesh_elem2node(const int64_t* conn, int64_t nelem, int nne, const int64_t* dofs, int64_t nnode,
                      int ndim, int64_t* ptr, int64_t* idx, int64_t* nnz)
{
    return guard([&] {
        std::vector<std::vector<int64_t>> ret(static_cast<size_t>(nnode));
        for (int64_t e = 0; e < nelem; ++e)
            for (int m = 0; m < nne; ++m) {
                int64_t n = conn[e * nne + m];
                GF_REQUIRE(n >= 0 && n < nnode, "conn out of range");
                ret[n].push_back(e);
            }
        for (auto& r : ret) std::sort(r.begin(), r.end());
        if (dofs) {
            // Mesh.h:2516-2551: merge the lists of tied nodes (first of each group collects)
            auto ties = nodaltyings(dofs, nnode, ndim);
            for (int64_t m = 0; m < nnode; ++m) {
                if (ties[m].size() <= 1 || m > ties[m][0]) continue;
                for (size_t j = 1; j < ties[m].size(); ++j) {
                    auto& o = ret[ties[m][j]];
                    ret[m].insert(ret[m].end(), o.begin(), o.end());
                }
                std::sort(ret[m].begin(), ret[m].end());
                for (size_t j = 1; j < ties[m].size(); ++j) ret[ties[m][j]] = ret[m];
            }
        }
        write_csr(ret, ptr, idx, nnz);
    });
}

int gf_mesh_nodaltyings(const int64_t* dofs, int64_t nnode, int ndim, int64_t* ptr, int64_t* idx,
                        int64_t* nnz)
{
    return guard([&] { write_csr(nodaltyings(dofs, nnode, ndim), ptr, idx, nnz); });
}

int gf_tyings_control(const double* coor, const int64_t* dofs, int64_t nnode, int ndim, double* coor_out,
                      int64_t* dofs_out, int64_t* control_dofs, int64_t* control_nodes)
{
    return guard([&] {
        int64_t mx = -1;
        for (int64_t i = 0; i < nnode * ndim; ++i) mx = std::max(mx, dofs[i]);
        std::copy(coor, coor + nnode * ndim, coor_out);
        std::copy(dofs, dofs + nnode * ndim, dofs_out);
        for (int i = 0; i < ndim * ndim; ++i) {
            control_dofs[i] = mx + 1 + i;
            coor_out[nnode * ndim + i] = 0.0;
            dofs_out[nnode * ndim + i] = mx + 1 + i;
        }
        for (int i = 0; i < ndim; ++i) control_nodes[i] = nnode + i;
    });
}

int gf_tyings_periodic(const int64_t* dofs, int64_t nnode, int ndim, const int64_t* control,
                       const int64_t* ties, int64_t nties, const int64_t* iip, int64_t niip,
                       int64_t* dofs_out, int64_t* control_out, int64_t* sizes)
{
    return guard([&] {
        // iid: dofs of dependent nodes in tie order (row-major); iii = setdiff1d(dofs, iid);
        // iiu = setdiff1d(iii, iip); Reorder({iiu, iip, iid})
        std::vector<int64_t> iid;
        iid.reserve(static_cast<size_t>(nties * ndim));
        for (int64_t t = 0; t < nties; ++t)
            for (int j = 0; j < ndim; ++j) iid.push_back(dofs[ties[2 * t + 1] * ndim + j]);
        auto all = sorted_unique(dofs, static_cast<size_t>(nnode * ndim));
        auto dep = sorted_unique(iid.data(), iid.size());
        std::vector<int64_t> iii, iiu;
        std::set_difference(all.begin(), all.end(), dep.begin(), dep.end(), std::back_inserter(iii));
        auto pre = sorted_unique(iip, static_cast<size_t>(niip));
        std::set_difference(iii.begin(), iii.end(), pre.begin(), pre.end(), std::back_inserter(iiu));
        int64_t n = 0;
        for (auto v : iiu) n = std::max(n, v + 1);
        for (int64_t i = 0; i < niip; ++i) n = std::max(n, iip[i] + 1);
        for (auto v : iid) n = std::max(n, v + 1);
        std::vector<int64_t> idx(static_cast<size_t>(n), -1);
        int64_t k = 0;
        for (auto v : iiu) idx[v] = k++;
        for (int64_t i = 0; i < niip; ++i) idx[iip[i]] = k++;
        for (auto v : iid) idx[v] = k++;
        for (int64_t i = 0; i < nnode * ndim; ++i) {
            GF_REQUIRE(idx[dofs[i]] >= 0, "dof not covered by [iiu, iip, iid]");
            dofs_out[i] = idx[dofs[i]];
        }
        for (int i = 0; i < ndim * ndim; ++i) control_out[i] = idx[control[i]];
        sizes[0] = static_cast<int64_t>(iiu.size());
        sizes[1] = niip;
        sizes[2] = static_cast<int64_t>(iii.size());
        sizes[3] = static_cast<int64_t>(iid.size());
    });
}

int gf_tyings_matrix(int which, const double* coor, const int64_t* dofs, int ndim, const int64_t* control,
                     const int64_t* ties, int64_t nties, int64_t nnu, int64_t* rowptr, int64_t* col,
                     double* val, int64_t* nnz)
{
    return guard([&] {
        GF_REQUIRE(which >= 0 && which <= 2, "which must be 0 (Cdi), 1 (Cdu) or 2 (Cdp)");
        int64_t nrows = nties * ndim;
        std::vector<std::vector<std::pair<int64_t, double>>> rows(static_cast<size_t>(nrows));
        auto keep = [&](int64_t c) { return which == 0 || (which == 1 ? c < nnu : c >= nnu); };
        auto shift = [&](int64_t c) { return which == 2 ? c - nnu : c; };
        for (int64_t i = 0; i < nties; ++i)
            for (int j = 0; j < ndim; ++j) {
                int64_t ni = ties[2 * i], nd = ties[2 * i + 1];
                auto& row = rows[i * ndim + j];
                int64_t c0 = dofs[ni * ndim + j];
                if (keep(c0)) row.push_back({shift(c0), 1.0});
                for (int k = 0; k < ndim; ++k) {
                    int64_t c = control[j * ndim + k];
                    if (keep(c)) row.push_back({shift(c), coor[nd * ndim + k] - coor[ni * ndim + k]});
                }
            }
        int64_t total = 0;
        for (auto& row : rows) {
            std::stable_sort(row.begin(), row.end(),
                             [](const std::pair<int64_t, double>& a, const std::pair<int64_t, double>& b) {
                                 return a.first < b.first;
                             });
            std::vector<std::pair<int64_t, double>> m;
            for (auto& p : row) {
                if (!m.empty() && m.back().first == p.first) m.back().second += p.second;
                else m.push_back(p);
            }
            row.swap(m);
            total += static_cast<int64_t>(row.size());
        }
        if (nnz) *nnz = total;
        if (rowptr) {
            rowptr[0] = 0;
            for (int64_t r = 0; r < nrows; ++r) rowptr[r + 1] = rowptr[r] + static_cast<int64_t>(rows[r].size());
        }
        if (col && val) {
            int64_t k = 0;
            for (auto& row : rows)
                for (auto& p : row) {
                    col[k] = p.first;
                    val[k] = p.second;
                    ++k;
                }
        }
    });
}

int gf_randn(uint64_t seed, int64_t n, double mean, double stddev, double* out)
{
    return guard([&] {
        uint64_t s = seed;
        auto next = [&]() {
            uint64_t z = (s += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            return z ^ (z >> 31);
        };
        const double two_pi = 6.283185307179586476925286766559;
        for (int64_t i = 0; i < n; i += 2) {
            double u1 = (static_cast<double>(next() >> 11) + 1.0) * (1.0 / 9007199254740992.0);
            double u2 = static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0);
            double r = std::sqrt(-2.0 * std::log(u1));
            out[i] = mean + stddev * r * std::cos(two_pi * u2);
            if (i + 1 < n) out[i + 1] = mean + stddev * r * std::sin(two_pi * u2);
        }
    });
}

} // extern "C"
