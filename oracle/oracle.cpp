// FP64 CPU oracle: a loop-for-loop restatement of GooseFEM's element-quadrature and assembly
// path, used ONLY as the reference results in tests/, __graft_entry__.smoke() and the CPU-baseline leg of
// bench.py.  Nothing in the product (goosefem_amd/) links, loads or calls this file.
//
// Every function follows the reference loop order and expression order so that FP64 results are
// reproducible bit for bit.  It is compiled with -ffp-contract=off (no fused multiply-add), like
// the product's FP64 kernels.  Citations are relative to the GooseFEM repository root.
//
// Parity pins (see tests/test_oracle_*.py):
//   mesh / dofs / periodic / elem2node   golden arrays of tests/basic/Mesh*.cpp, tests/test_Mesh.py
//   compute_dN, dV                       closed form (J = h/2 I) + finite-difference isoparametric map
//   gradN / symGradN / int_* kernels     patch tests (affine u), brute-force einsum on tiny meshes
//   int_N_scalar_NT_dV + diag assemble   golden lumped-mass vector tests/basic/MatrixDiagonal.cpp
//   vector conversions                   golden periodic values tests/basic/Vector.cpp + brute force
//   sparse assemble / dot                dense brute force (numpy) on tiny meshes
//   tyings / K' / CG                     scipy spsolve on K' built from dense blocks (tiny meshes)
//   material laws                        parity unpinned by the reference (external GMat); closed
//                                        form identities only (elastic; J2 linear hardening)
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef int64_t i64;

namespace {

// ------------------------------------------------------------------------------------------------
// small tensor inverse, include/GooseFEM/detail.h:17-70 (cofactor inverse, returns determinant)
// ------------------------------------------------------------------------------------------------
inline double inv2(const double A[2][2], double B[2][2])
{
    double det = A[0][0] * A[1][1] - A[0][1] * A[1][0];
    B[0][0] = A[1][1] / det;
    B[0][1] = -1.0 * A[0][1] / det;
    B[1][0] = -1.0 * A[1][0] / det;
    B[1][1] = A[0][0] / det;
    return det;
}

inline double inv3(const double A[3][3], double B[3][3])
{
    double det = (A[0][0] * A[1][1] * A[2][2] + A[0][1] * A[1][2] * A[2][0] +
                  A[0][2] * A[1][0] * A[2][1]) -
                 (A[0][2] * A[1][1] * A[2][0] + A[0][1] * A[1][0] * A[2][2] +
                  A[0][0] * A[1][2] * A[2][1]);
    B[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) / det;
    B[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
    B[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
    B[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) / det;
    B[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
    B[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
    B[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) / det;
    B[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
    B[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
    return det;
}

// element types: 0 = Quad4 (tdim 2), 1 = Quad4 planar (tdim 3), 2 = Hex8 (tdim 3)
inline int et_nne(int et) { return et == 2 ? 8 : 4; }
inline int et_ndim(int et) { return et == 2 ? 3 : 2; }
inline int et_tdim(int et) { return et == 0 ? 2 : 3; }

// sorted unique values
std::vector<i64> unique_sorted(const i64* a, i64 n)
{
    std::vector<i64> v(a, a + n);
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    return v;
}

} // namespace

extern "C" {

// ================================================================================================
// Mesh: include/GooseFEM/MeshQuad4.h:43-187, MeshHex8.h:41-609, Mesh.h:93-175,291-304
// ================================================================================================

// Node coordinates: xt::linspace(0, h*nel, nel+1) evaluated as i * (h*nel / nel); exact for h = 1.
static inline double lin(i64 i, i64 nel, double h)
{
    double stop = h * static_cast<double>(nel);
    double step = stop / static_cast<double>(nel);
    return static_cast<double>(i) * step;
}

// MeshQuad4.h:83-104 (coor_impl) and :106-122 (conn_impl)
void or_quad4_regular(i64 nx, i64 ny, double h, double* coor, i64* conn)
{
    i64 inode = 0;
    for (i64 iy = 0; iy < ny + 1; ++iy) {
        for (i64 ix = 0; ix < nx + 1; ++ix) {
            coor[inode * 2 + 0] = lin(ix, nx, h);
            coor[inode * 2 + 1] = lin(iy, ny, h);
            ++inode;
        }
    }
    i64 e = 0;
    for (i64 iy = 0; iy < ny; ++iy) {
        for (i64 ix = 0; ix < nx; ++ix) {
            conn[e * 4 + 0] = iy * (nx + 1) + ix;
            conn[e * 4 + 1] = iy * (nx + 1) + (ix + 1);
            conn[e * 4 + 3] = (iy + 1) * (nx + 1) + ix;
            conn[e * 4 + 2] = (iy + 1) * (nx + 1) + (ix + 1);
            ++e;
        }
    }
}

// MeshHex8.h:77-103 (coor_impl) and :104-131 (conn_impl)
void or_hex8_regular(i64 nx, i64 ny, i64 nz, double h, double* coor, i64* conn)
{
    i64 inode = 0;
    for (i64 iz = 0; iz < nz + 1; ++iz) {
        for (i64 iy = 0; iy < ny + 1; ++iy) {
            for (i64 ix = 0; ix < nx + 1; ++ix) {
                coor[inode * 3 + 0] = lin(ix, nx, h);
                coor[inode * 3 + 1] = lin(iy, ny, h);
                coor[inode * 3 + 2] = lin(iz, nz, h);
                ++inode;
            }
        }
    }
    i64 e = 0;
    i64 sx = nx + 1, sxy = (nx + 1) * (ny + 1);
    for (i64 iz = 0; iz < nz; ++iz) {
        for (i64 iy = 0; iy < ny; ++iy) {
            for (i64 ix = 0; ix < nx; ++ix) {
                conn[e * 8 + 0] = iy * sx + ix + iz * sxy;
                conn[e * 8 + 1] = iy * sx + (ix + 1) + iz * sxy;
                conn[e * 8 + 3] = (iy + 1) * sx + ix + iz * sxy;
                conn[e * 8 + 2] = (iy + 1) * sx + (ix + 1) + iz * sxy;
                conn[e * 8 + 4] = iy * sx + ix + (iz + 1) * sxy;
                conn[e * 8 + 5] = iy * sx + (ix + 1) + (iz + 1) * sxy;
                conn[e * 8 + 7] = (iy + 1) * sx + ix + (iz + 1) * sxy;
                conn[e * 8 + 6] = (iy + 1) * sx + (ix + 1) + (iz + 1) * sxy;
                ++e;
            }
        }
    }
}

// Quad4 node sets, MeshQuad4.h:128-187.  which: 0 bottom, 1 top, 2 left, 3 right edge,
// 4..7 the same "open" edges (without corners).  Returns the count; writes when out != NULL.
i64 or_quad4_nodeset(i64 nx, i64 ny, int which, i64* out)
{
    std::vector<i64> r;
    switch (which) {
    case 0: for (i64 i = 0; i < nx + 1; ++i) r.push_back(i); break;
    case 1: for (i64 i = 0; i < nx + 1; ++i) r.push_back(i + ny * (nx + 1)); break;
    case 2: for (i64 i = 0; i < ny + 1; ++i) r.push_back(i * (nx + 1)); break;
    case 3: for (i64 i = 0; i < ny + 1; ++i) r.push_back(i * (nx + 1) + nx); break;
    case 4: for (i64 i = 1; i < nx; ++i) r.push_back(i); break;
    case 5: for (i64 i = 1; i < nx; ++i) r.push_back(i + ny * (nx + 1)); break;
    case 6: for (i64 i = 1; i < ny; ++i) r.push_back(i * (nx + 1)); break;
    case 7: for (i64 i = 1; i < ny; ++i) r.push_back(i * (nx + 1) + nx); break;
    default: return -1;
    }
    if (out) std::copy(r.begin(), r.end(), out);
    return static_cast<i64>(r.size());
}

// Mesh.h:511-540: periodic node pairs (independent, dependent) of a 2-D regular mesh.
i64 or_quad4_nodes_periodic(i64 nx, i64 ny, i64* out)
{
    i64 BL = 0, BR = nx, TL = ny * (nx + 1), TR = ny * (nx + 1) + nx;
    std::vector<i64> r = {BL, BR, BL, TR, BL, TL};
    for (i64 i = 1; i < nx; ++i) {
        r.push_back(i);
        r.push_back(i + ny * (nx + 1));
    }
    for (i64 i = 1; i < ny; ++i) {
        r.push_back(i * (nx + 1));
        r.push_back(i * (nx + 1) + nx);
    }
    if (out) std::copy(r.begin(), r.end(), out);
    return static_cast<i64>(r.size() / 2);
}

// Mesh.h:1577-1692 with the face/edge/corner sets of MeshHex8.h:214-597.
i64 or_hex8_nodes_periodic(i64 nx, i64 ny, i64 nz, i64* out)
{
    const i64 sx = nx + 1, sxy = (nx + 1) * (ny + 1);
    auto face = [&](int f) {
        std::vector<i64> r;
        if (f == 0 || f == 1) { // front (z=0) / back (z=nz), interior ix,iy
            for (i64 iy = 1; iy < ny; ++iy)
                for (i64 ix = 1; ix < nx; ++ix) r.push_back(iy * sx + ix + (f ? nz * sxy : 0));
        }
        else if (f == 2 || f == 3) { // left (x=0) / right (x=nx), interior iy,iz
            for (i64 iz = 1; iz < nz; ++iz)
                for (i64 iy = 1; iy < ny; ++iy) r.push_back(iy * sx + iz * sxy + (f == 3 ? nx : 0));
        }
        else { // bottom (y=0) / top (y=ny), interior ix,iz
            for (i64 iz = 1; iz < nz; ++iz)
                for (i64 ix = 1; ix < nx; ++ix) r.push_back(ix + iz * sxy + (f == 5 ? ny * sx : 0));
        }
        return r;
    };
    // open edges along x: (y, z) in {0, ny} x {0, nz}
    auto edge_x = [&](i64 y, i64 z) {
        std::vector<i64> r;
        for (i64 ix = 1; ix < nx; ++ix) r.push_back(ix + y * sx + z * sxy);
        return r;
    };
    auto edge_y = [&](i64 x, i64 z) {
        std::vector<i64> r;
        for (i64 iy = 1; iy < ny; ++iy) r.push_back(iy * sx + x + z * sxy);
        return r;
    };
    auto edge_z = [&](i64 x, i64 y) {
        std::vector<i64> r;
        for (i64 iz = 1; iz < nz; ++iz) r.push_back(iz * sxy + x + y * sx);
        return r;
    };
    auto corner = [&](i64 x, i64 y, i64 z) { return x + y * sx + z * sxy; };

    std::vector<i64> r;
    auto pair = [&](i64 a, i64 b) {
        r.push_back(a);
        r.push_back(b);
    };
    i64 FBL = corner(0, 0, 0);
    pair(FBL, corner(nx, 0, 0));   // front-bottom-right
    pair(FBL, corner(nx, 0, nz));  // back-bottom-right
    pair(FBL, corner(0, 0, nz));   // back-bottom-left
    pair(FBL, corner(0, ny, 0));   // front-top-left
    pair(FBL, corner(nx, ny, 0));  // front-top-right
    pair(FBL, corner(nx, ny, nz)); // back-top-right
    pair(FBL, corner(0, ny, nz));  // back-top-left

    auto froBot = edge_x(0, 0), bckBot = edge_x(0, nz), bckTop = edge_x(ny, nz), froTop = edge_x(ny, 0);
    auto botLft = edge_z(0, 0), botRgt = edge_z(nx, 0), topRgt = edge_z(nx, ny), topLft = edge_z(0, ny);
    auto froLft = edge_y(0, 0), froRgt = edge_y(nx, 0), bckRgt = edge_y(nx, nz), bckLft = edge_y(0, nz);

    for (size_t j = 0; j < froBot.size(); ++j) pair(froBot[j], bckBot[j]);
    for (size_t j = 0; j < froBot.size(); ++j) pair(froBot[j], bckTop[j]);
    for (size_t j = 0; j < froBot.size(); ++j) pair(froBot[j], froTop[j]);
    for (size_t j = 0; j < botLft.size(); ++j) pair(botLft[j], botRgt[j]);
    for (size_t j = 0; j < botLft.size(); ++j) pair(botLft[j], topRgt[j]);
    for (size_t j = 0; j < botLft.size(); ++j) pair(botLft[j], topLft[j]);
    for (size_t j = 0; j < froLft.size(); ++j) pair(froLft[j], froRgt[j]);
    for (size_t j = 0; j < froLft.size(); ++j) pair(froLft[j], bckRgt[j]);
    for (size_t j = 0; j < froLft.size(); ++j) pair(froLft[j], bckLft[j]);

    auto fro = face(0), bck = face(1), lft = face(2), rgt = face(3), bot = face(4), top = face(5);
    for (size_t j = 0; j < fro.size(); ++j) pair(fro[j], bck[j]);
    for (size_t j = 0; j < lft.size(); ++j) pair(lft[j], rgt[j]);
    for (size_t j = 0; j < bot.size(); ++j) pair(bot[j], top[j]);

    if (out) std::copy(r.begin(), r.end(), out);
    return static_cast<i64>(r.size() / 2);
}

// Mesh.h:114-175: renumber to the lowest possible index (rank in the sorted unique set).
void or_renumber(const i64* in, i64 n, i64* out)
{
    std::vector<i64> u = unique_sorted(in, n);
    i64 mx = u.empty() ? 0 : u.back() + 1;
    std::vector<i64> renum(mx, 0);
    for (size_t i = 0; i < u.size(); ++i) renum[u[i]] = static_cast<i64>(i);
    for (i64 i = 0; i < n; ++i) out[i] = renum[in[i]];
}

// Mesh.h:291-304: dofs with dependent nodes copied from independent nodes, then renumbered.
void or_dofs_periodic(i64 nnode, int ndim, const i64* ties, i64 nties, i64* out)
{
    std::vector<i64> d(nnode * ndim);
    for (i64 i = 0; i < nnode * ndim; ++i) d[i] = i;
    for (int j = 0; j < ndim; ++j)
        for (i64 t = 0; t < nties; ++t) d[ties[t * 2 + 1] * ndim + j] = d[ties[t * 2 + 0] * ndim + j];
    or_renumber(d.data(), nnode * ndim, out);
}

// Mesh.h:2462-2477: number of entries equal to each value (conn -> elements per node).
void or_coordination(const i64* conn, i64 n, i64 nvalues, i64* out)
{
    for (i64 i = 0; i < nvalues; ++i) out[i] = 0;
    for (i64 i = 0; i < n; ++i) out[conn[i]] += 1;
}

// Mesh.h:2485-2510 (elem2node, sorted) as CSR.  Returns nnz; ptr[nnode+1], idx[nnz].
i64 or_elem2node(const i64* conn, i64 nelem, int nne, i64 nnode, i64* ptr, i64* idx)
{
    std::vector<std::vector<i64>> ret(nnode);
    for (i64 e = 0; e < nelem; ++e)
        for (int m = 0; m < nne; ++m) ret[conn[e * nne + m]].push_back(e);
    i64 nnz = 0;
    for (auto& row : ret) {
        std::sort(row.begin(), row.end());
        nnz += static_cast<i64>(row.size());
    }
    if (ptr) {
        ptr[0] = 0;
        for (i64 i = 0; i < nnode; ++i) ptr[i + 1] = ptr[i] + static_cast<i64>(ret[i].size());
    }
    if (idx) {
        i64 k = 0;
        for (auto& row : ret)
            for (auto v : row) idx[k++] = v;
    }
    return nnz;
}

// Mesh.h:2839-2861 nodaltyings via node2dof (Mesh.h:2560) = elem2node(dofs).  CSR output.
i64 or_nodaltyings(const i64* dofs, i64 nnode, int ndim, i64* ptr, i64* idx)
{
    i64 ndof = 0;
    for (i64 i = 0; i < nnode * ndim; ++i) ndof = std::max(ndof, dofs[i] + 1);
    std::vector<std::vector<i64>> node2dof(ndof);
    for (i64 n = 0; n < nnode; ++n)
        for (int i = 0; i < ndim; ++i) node2dof[dofs[n * ndim + i]].push_back(n);
    for (auto& r : node2dof) std::sort(r.begin(), r.end());
    i64 nnz = 0;
    for (i64 n = 0; n < nnode; ++n) nnz += static_cast<i64>(node2dof[dofs[n * ndim]].size());
    if (ptr) {
        ptr[0] = 0;
        for (i64 n = 0; n < nnode; ++n)
            ptr[n + 1] = ptr[n] + static_cast<i64>(node2dof[dofs[n * ndim]].size());
    }
    if (idx) {
        i64 k = 0;
        for (i64 n = 0; n < nnode; ++n)
            for (auto v : node2dof[dofs[n * ndim]]) idx[k++] = v;
    }
    return nnz;
}

// Mesh.h:2516-2551: periodic-aware elem2node (lists of tied nodes merged, sorted).  CSR output.
i64 or_elem2node_periodic(const i64* conn, i64 nelem, int nne, const i64* dofs, i64 nnode, int ndim,
                          i64* ptr, i64* idx)
{
    std::vector<std::vector<i64>> ret(nnode);
    for (i64 e = 0; e < nelem; ++e)
        for (int m = 0; m < nne; ++m) ret[conn[e * nne + m]].push_back(e);
    for (auto& row : ret) std::sort(row.begin(), row.end());
    i64 tp = or_nodaltyings(dofs, nnode, ndim, nullptr, nullptr);
    std::vector<i64> tptr(nnode + 1), tidx(tp);
    or_nodaltyings(dofs, nnode, ndim, tptr.data(), tidx.data());
    for (i64 m = 0; m < nnode; ++m) {
        i64 cnt = tptr[m + 1] - tptr[m];
        if (cnt <= 1) continue;
        if (m > tidx[tptr[m]]) continue;
        for (i64 j = 1; j < cnt; ++j) {
            auto& o = ret[tidx[tptr[m] + j]];
            ret[m].insert(ret[m].end(), o.begin(), o.end());
        }
        std::sort(ret[m].begin(), ret[m].end());
        for (i64 j = 1; j < cnt; ++j) ret[tidx[tptr[m] + j]] = ret[m];
    }
    i64 nnz = 0;
    for (auto& r : ret) nnz += static_cast<i64>(r.size());
    if (ptr) {
        ptr[0] = 0;
        for (i64 i = 0; i < nnode; ++i) ptr[i + 1] = ptr[i] + static_cast<i64>(ret[i].size());
    }
    if (idx) {
        i64 k = 0;
        for (auto& row : ret)
            for (auto v : row) idx[k++] = v;
    }
    return nnz;
}

// ================================================================================================
// Quadrature: ElementQuad4.h:36-163,255-318; ElementQuad4Planar.h:80-123; ElementHex8.h:26-317
// ================================================================================================

// Gauss (nodal = 0 -> Gauss, 1 -> Nodal) integration points and weights.
void or_quad_rule(int et, int nodal, double* xi, double* w)
{
    const double a = nodal ? 1.0 : 1.0 / std::sqrt(3.0);
    if (et_ndim(et) == 2) {
        const double s[4][2] = {{-1, -1}, {+1, -1}, {+1, +1}, {-1, +1}};
        for (int q = 0; q < 4; ++q) {
            xi[q * 2 + 0] = s[q][0] < 0 ? -a : +a;
            xi[q * 2 + 1] = s[q][1] < 0 ? -a : +a;
            w[q] = 1.0;
        }
    }
    else {
        const double s[8][3] = {{-1, -1, -1}, {+1, -1, -1}, {+1, +1, -1}, {-1, +1, -1},
                                {-1, -1, +1}, {+1, -1, +1}, {+1, +1, +1}, {-1, +1, +1}};
        for (int q = 0; q < 8; ++q) {
            for (int d = 0; d < 3; ++d) xi[q * 3 + d] = s[q][d] < 0 ? -a : +a;
            w[q] = 1.0;
        }
    }
}

// Shape functions and local gradients: ElementQuad4.h:287-305, ElementHex8.h:265-304.
void or_quad_shape(int et, int nip, const double* xi, double* N, double* dNxi)
{
    if (et_ndim(et) == 2) {
        for (int q = 0; q < nip; ++q) {
            const double x0 = xi[q * 2 + 0], x1 = xi[q * 2 + 1];
            N[q * 4 + 0] = 0.25 * (1.0 - x0) * (1.0 - x1);
            N[q * 4 + 1] = 0.25 * (1.0 + x0) * (1.0 - x1);
            N[q * 4 + 2] = 0.25 * (1.0 + x0) * (1.0 + x1);
            N[q * 4 + 3] = 0.25 * (1.0 - x0) * (1.0 + x1);
            double* d = dNxi + q * 8;
            d[0 * 2 + 0] = -0.25 * (1.0 - x1);
            d[1 * 2 + 0] = +0.25 * (1.0 - x1);
            d[2 * 2 + 0] = +0.25 * (1.0 + x1);
            d[3 * 2 + 0] = -0.25 * (1.0 + x1);
            d[0 * 2 + 1] = -0.25 * (1.0 - x0);
            d[1 * 2 + 1] = -0.25 * (1.0 + x0);
            d[2 * 2 + 1] = +0.25 * (1.0 + x0);
            d[3 * 2 + 1] = +0.25 * (1.0 - x0);
        }
    }
    else {
        for (int q = 0; q < nip; ++q) {
            const double x0 = xi[q * 3 + 0], x1 = xi[q * 3 + 1], x2 = xi[q * 3 + 2];
            double* n = N + q * 8;
            n[0] = 0.125 * (1.0 - x0) * (1.0 - x1) * (1.0 - x2);
            n[1] = 0.125 * (1.0 + x0) * (1.0 - x1) * (1.0 - x2);
            n[2] = 0.125 * (1.0 + x0) * (1.0 + x1) * (1.0 - x2);
            n[3] = 0.125 * (1.0 - x0) * (1.0 + x1) * (1.0 - x2);
            n[4] = 0.125 * (1.0 - x0) * (1.0 - x1) * (1.0 + x2);
            n[5] = 0.125 * (1.0 + x0) * (1.0 - x1) * (1.0 + x2);
            n[6] = 0.125 * (1.0 + x0) * (1.0 + x1) * (1.0 + x2);
            n[7] = 0.125 * (1.0 - x0) * (1.0 + x1) * (1.0 + x2);
            double* d = dNxi + q * 24;
            d[0 * 3 + 0] = -0.125 * (1.0 - x1) * (1.0 - x2);
            d[1 * 3 + 0] = +0.125 * (1.0 - x1) * (1.0 - x2);
            d[2 * 3 + 0] = +0.125 * (1.0 + x1) * (1.0 - x2);
            d[3 * 3 + 0] = -0.125 * (1.0 + x1) * (1.0 - x2);
            d[4 * 3 + 0] = -0.125 * (1.0 - x1) * (1.0 + x2);
            d[5 * 3 + 0] = +0.125 * (1.0 - x1) * (1.0 + x2);
            d[6 * 3 + 0] = +0.125 * (1.0 + x1) * (1.0 + x2);
            d[7 * 3 + 0] = -0.125 * (1.0 + x1) * (1.0 + x2);
            d[0 * 3 + 1] = -0.125 * (1.0 - x0) * (1.0 - x2);
            d[1 * 3 + 1] = -0.125 * (1.0 + x0) * (1.0 - x2);
            d[2 * 3 + 1] = +0.125 * (1.0 + x0) * (1.0 - x2);
            d[3 * 3 + 1] = +0.125 * (1.0 - x0) * (1.0 - x2);
            d[4 * 3 + 1] = -0.125 * (1.0 - x0) * (1.0 + x2);
            d[5 * 3 + 1] = -0.125 * (1.0 + x0) * (1.0 + x2);
            d[6 * 3 + 1] = +0.125 * (1.0 + x0) * (1.0 + x2);
            d[7 * 3 + 1] = +0.125 * (1.0 - x0) * (1.0 + x2);
            d[0 * 3 + 2] = -0.125 * (1.0 - x0) * (1.0 - x1);
            d[1 * 3 + 2] = -0.125 * (1.0 + x0) * (1.0 - x1);
            d[2 * 3 + 2] = -0.125 * (1.0 + x0) * (1.0 + x1);
            d[3 * 3 + 2] = -0.125 * (1.0 - x0) * (1.0 + x1);
            d[4 * 3 + 2] = +0.125 * (1.0 - x0) * (1.0 - x1);
            d[5 * 3 + 2] = +0.125 * (1.0 + x0) * (1.0 - x1);
            d[6 * 3 + 2] = +0.125 * (1.0 + x0) * (1.0 + x1);
            d[7 * 3 + 2] = +0.125 * (1.0 - x0) * (1.0 + x1);
        }
    }
}

// compute_dN: ElementQuad4.h:493-532 (Quad4), ElementQuad4Planar.h:280-319 (x thick),
// Element.h:1178-1229 (generic, Hex8).  The generic reference loop reads dNxi(m,i) where the
// chain rule needs dNxi(m,j) (quirk Q1, Element.h:1220); faithful_q1 = 1 reproduces it, 0 uses
// the correct chain rule.  Both agree whenever J is diagonal (every Regular Hex8 element).
void or_quad_compute_dN(int et, i64 nelem, int nip, const double* x, const double* dNxi,
                        const double* w, double thick, int faithful_q1, double* dNx, double* vol)
{
    const int nne = et_nne(et), nd = et_ndim(et);
    for (i64 e = 0; e < nelem; ++e) {
        const double* xe = x + e * nne * nd;
        for (int q = 0; q < nip; ++q) {
            const double* g = dNxi + q * nne * nd;
            double* o = dNx + (e * nip + q) * nne * nd;
            if (nd == 2) {
                double J[2][2], Ji[2][2];
                J[0][0] = g[0] * xe[0] + g[2] * xe[2] + g[4] * xe[4] + g[6] * xe[6];
                J[0][1] = g[0] * xe[1] + g[2] * xe[3] + g[4] * xe[5] + g[6] * xe[7];
                J[1][0] = g[1] * xe[0] + g[3] * xe[2] + g[5] * xe[4] + g[7] * xe[6];
                J[1][1] = g[1] * xe[1] + g[3] * xe[3] + g[5] * xe[5] + g[7] * xe[7];
                double det = inv2(J, Ji);
                for (int m = 0; m < 4; ++m) {
                    o[m * 2 + 0] = Ji[0][0] * g[m * 2 + 0] + Ji[0][1] * g[m * 2 + 1];
                    o[m * 2 + 1] = Ji[1][0] * g[m * 2 + 0] + Ji[1][1] * g[m * 2 + 1];
                }
                vol[e * nip + q] = (et == 1) ? w[q] * det * thick : w[q] * det;
            }
            else {
                double J[3][3], Ji[3][3];
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) J[i][j] = 0.0;
                for (int m = 0; m < 8; ++m)
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j) J[i][j] += g[m * 3 + i] * xe[m * 3 + j];
                double det = inv3(J, Ji);
                for (int k = 0; k < 24; ++k) o[k] = 0.0;
                for (int m = 0; m < 8; ++m)
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j)
                            o[m * 3 + i] += Ji[i][j] * g[m * 3 + (faithful_q1 ? i : j)];
                vol[e * nip + q] = w[q] * det;
            }
        }
    }
}

// InterpQuad_vector: ElementQuad4.h:324-346 (Quad4 override), Element.h:892-922 (generic).
void or_quad_interp_vector(int et, i64 nelem, int nip, const double* N, const double* elemvec, int n,
                           double* qvector)
{
    const int nne = et_nne(et);
    if (et == 0 && n == 2) {
        for (i64 e = 0; e < nelem; ++e) {
            const double* u = elemvec + e * 8;
            for (int q = 0; q < nip; ++q) {
                const double* Nq = N + q * 4;
                double* ui = qvector + (e * nip + q) * 2;
                ui[0] = Nq[0] * u[0] + Nq[1] * u[2] + Nq[2] * u[4] + Nq[3] * u[6];
                ui[1] = Nq[0] * u[1] + Nq[1] * u[3] + Nq[2] * u[5] + Nq[3] * u[7];
            }
        }
        return;
    }
    for (i64 i = 0; i < nelem * nip * n; ++i) qvector[i] = 0.0;
    for (i64 e = 0; e < nelem; ++e) {
        const double* fq = elemvec + e * nne * n;
        for (int q = 0; q < nip; ++q) {
            const double* Nq = N + q * nne;
            double* tq = qvector + (e * nip + q) * n;
            for (int m = 0; m < nne; ++m)
                for (int i = 0; i < n; ++i) tq[i] += Nq[m] * fq[m * n + i];
        }
    }
}

// GradN_vector (mode 0), GradN_vector_T (mode 1), SymGradN_vector (mode 2):
// Quad4 ElementQuad4.h:348-434, Planar ElementQuad4Planar.h:129-221, Hex8 generic
// Element.h:924-1022.  qtensor [nelem, nip, tdim, tdim].
void or_quad_grad(int et, int mode, i64 nelem, int nip, const double* dNx, const double* elemvec,
                  double* qtensor)
{
    const int nne = et_nne(et), nd = et_ndim(et), td = et_tdim(et);
    if (et == 1 || et == 2) {
        for (i64 i = 0; i < nelem * nip * td * td; ++i) qtensor[i] = 0.0;
    }
    for (i64 e = 0; e < nelem; ++e) {
        const double* u = elemvec + e * nne * nd;
        for (int q = 0; q < nip; ++q) {
            const double* g = dNx + (e * nip + q) * nne * nd;
            double* t = qtensor + (e * nip + q) * td * td;
            if (nd == 2) {
                // u(m,j) = u[m*2+j], dNx(m,i) = g[m*2+i]
                double a00 = g[0] * u[0] + g[2] * u[2] + g[4] * u[4] + g[6] * u[6];
                double a01 = g[0] * u[1] + g[2] * u[3] + g[4] * u[5] + g[6] * u[7];
                double a10 = g[1] * u[0] + g[3] * u[2] + g[5] * u[4] + g[7] * u[6];
                double a11 = g[1] * u[1] + g[3] * u[3] + g[5] * u[5] + g[7] * u[7];
                if (mode == 0) {
                    t[0 * td + 0] = a00;
                    t[0 * td + 1] = a01;
                    t[1 * td + 0] = a10;
                    t[1 * td + 1] = a11;
                }
                else if (mode == 1) {
                    t[0 * td + 0] = a00;
                    t[1 * td + 0] = a01;
                    t[0 * td + 1] = a10;
                    t[1 * td + 1] = a11;
                }
                else {
                    t[0 * td + 0] = a00;
                    t[1 * td + 1] = a11;
                    t[0 * td + 1] = 0.5 * (g[0] * u[1] + g[2] * u[3] + g[4] * u[5] + g[6] * u[7] +
                                           g[1] * u[0] + g[3] * u[2] + g[5] * u[4] + g[7] * u[6]);
                    t[1 * td + 0] = t[0 * td + 1];
                }
            }
            else {
                for (int m = 0; m < nne; ++m)
                    for (int i = 0; i < nd; ++i)
                        for (int j = 0; j < nd; ++j) {
                            if (mode == 0) t[i * td + j] += g[m * 3 + i] * u[m * 3 + j];
                            else if (mode == 1) t[j * td + i] += g[m * 3 + i] * u[m * 3 + j];
                            else {
                                t[i * td + j] += 0.5 * g[m * 3 + i] * u[m * 3 + j];
                                t[j * td + i] += 0.5 * g[m * 3 + i] * u[m * 3 + j];
                            }
                        }
            }
        }
    }
}

// Int_N_vector_dV: Element.h:1024-1055 (generic).
void or_quad_int_N_vector_dV(int et, i64 nelem, int nip, const double* N, const double* vol,
                             const double* qvector, int n, double* elemvec)
{
    const int nne = et_nne(et);
    for (i64 i = 0; i < nelem * nne * n; ++i) elemvec[i] = 0.0;
    for (i64 e = 0; e < nelem; ++e) {
        double* f = elemvec + e * nne * n;
        for (int q = 0; q < nip; ++q) {
            const double* Ne = N + q * nne;
            const double* tq = qvector + (e * nip + q) * n;
            double v = vol[e * nip + q];
            for (int m = 0; m < nne; ++m)
                for (int i = 0; i < n; ++i) f[m * n + i] += Ne[m] * tq[i] * v;
        }
    }
}

// Int_N_scalar_NT_dV: ElementQuad4.h:436-464, ElementQuad4Planar.h:223-251, ElementHex8.h:323-352.
void or_quad_int_N_scalar_NT_dV(int et, i64 nelem, int nip, const double* N, const double* vol,
                                const double* qscalar, double* elemmat)
{
    const int nne = et_nne(et), nd = et_ndim(et), K = nne * nd;
    for (i64 i = 0; i < nelem * K * K; ++i) elemmat[i] = 0.0;
    for (i64 e = 0; e < nelem; ++e) {
        double* M = elemmat + e * K * K;
        for (int q = 0; q < nip; ++q) {
            const double* Nq = N + q * nne;
            double v = vol[e * nip + q];
            double rho = qscalar[e * nip + q];
            for (int m = 0; m < nne; ++m)
                for (int n = 0; n < nne; ++n)
                    for (int i = 0; i < nd; ++i) M[(m * nd + i) * K + n * nd + i] += Nq[m] * rho * Nq[n] * v;
        }
    }
}

// Int_gradN_dot_tensor2_dV: ElementQuad4.h:466-491, ElementQuad4Planar.h:253-278,
// ElementHex8.h:354-383.  f(m,j) += (sum_i dNx(m,i) * sig(i,j)) * dV.
void or_quad_int_gradN_dot_tensor2_dV(int et, i64 nelem, int nip, const double* dNx,
                                      const double* vol, const double* qtensor, double* elemvec)
{
    const int nne = et_nne(et), nd = et_ndim(et), td = et_tdim(et);
    for (i64 i = 0; i < nelem * nne * nd; ++i) elemvec[i] = 0.0;
    for (i64 e = 0; e < nelem; ++e) {
        double* f = elemvec + e * nne * nd;
        for (int q = 0; q < nip; ++q) {
            const double* g = dNx + (e * nip + q) * nne * nd;
            const double* s = qtensor + (e * nip + q) * td * td;
            double v = vol[e * nip + q];
            for (int m = 0; m < nne; ++m) {
                if (nd == 2) {
                    f[m * 2 + 0] += (g[m * 2 + 0] * s[0 * td + 0] + g[m * 2 + 1] * s[1 * td + 0]) * v;
                    f[m * 2 + 1] += (g[m * 2 + 0] * s[0 * td + 1] + g[m * 2 + 1] * s[1 * td + 1]) * v;
                }
                else {
                    for (int j = 0; j < 3; ++j)
                        f[m * 3 + j] += (g[m * 3 + 0] * s[0 * 3 + j] + g[m * 3 + 1] * s[1 * 3 + j] +
                                         g[m * 3 + 2] * s[2 * 3 + j]) *
                                        v;
                }
            }
        }
    }
}

// Int_gradN_dot_tensor4_dot_gradNT_dV: Element.h:1131-1176 (generic, all elements).
// K(m*nd+j, n*nd+k) += dNx(m,i) * C(i,j,k,l) * dNx(n,l) * dV, tensor strides use tdim.
void or_quad_int_gradN_dot_tensor4_dot_gradNT_dV(int et, i64 nelem, int nip, const double* dNx,
                                                 const double* vol, const double* qtensor,
                                                 double* elemmat)
{
    const int nne = et_nne(et), nd = et_ndim(et), td = et_tdim(et), K = nne * nd;
    const int s3 = td * td * td, s2 = td * td, s1 = td;
    for (i64 i = 0; i < nelem * K * K; ++i) elemmat[i] = 0.0;
    for (i64 e = 0; e < nelem; ++e) {
        double* Ke = elemmat + e * K * K;
        for (int q = 0; q < nip; ++q) {
            const double* g = dNx + (e * nip + q) * nne * nd;
            const double* C = qtensor + (e * nip + q) * td * td * td * td;
            double v = vol[e * nip + q];
            for (int m = 0; m < nne; ++m)
                for (int n = 0; n < nne; ++n)
                    for (int i = 0; i < nd; ++i)
                        for (int j = 0; j < nd; ++j)
                            for (int k = 0; k < nd; ++k)
                                for (int l = 0; l < nd; ++l)
                                    Ke[(m * nd + j) * K + n * nd + k] +=
                                        g[m * nd + i] * C[i * s3 + j * s2 + k * s1 + l] * g[n * nd + l] * v;
        }
    }
}

// ================================================================================================
// Vector: include/GooseFEM/Vector.h:568-757 (serial loop order; overwrite = last write wins)
// ================================================================================================

void or_vec_as_element_nodevec(const i64* conn, i64 nelem, int nne, int ndim, const double* nodevec,
                               double* elemvec)
{
    for (i64 e = 0; e < nelem; ++e)
        for (int m = 0; m < nne; ++m)
            for (int i = 0; i < ndim; ++i)
                elemvec[(e * nne + m) * ndim + i] = nodevec[conn[e * nne + m] * ndim + i];
}

void or_vec_as_element_dofval(const i64* conn, const i64* dofs, i64 nelem, int nne, int ndim,
                              const double* dofval, double* elemvec)
{
    for (i64 e = 0; e < nelem; ++e)
        for (int m = 0; m < nne; ++m)
            for (int i = 0; i < ndim; ++i)
                elemvec[(e * nne + m) * ndim + i] = dofval[dofs[conn[e * nne + m] * ndim + i]];
}

void or_vec_as_dofs_nodevec(const i64* dofs, i64 nnode, int ndim, i64 ndof, const double* nodevec,
                            double* dofval)
{
    for (i64 d = 0; d < ndof; ++d) dofval[d] = 0.0;
    for (i64 n = 0; n < nnode; ++n)
        for (int i = 0; i < ndim; ++i) dofval[dofs[n * ndim + i]] = nodevec[n * ndim + i];
}

void or_vec_as_dofs_elemvec(const i64* conn, const i64* dofs, i64 nelem, int nne, int ndim, i64 ndof,
                            const double* elemvec, double* dofval)
{
    for (i64 d = 0; d < ndof; ++d) dofval[d] = 0.0;
    for (i64 e = 0; e < nelem; ++e)
        for (int m = 0; m < nne; ++m)
            for (int i = 0; i < ndim; ++i)
                dofval[dofs[conn[e * nne + m] * ndim + i]] = elemvec[(e * nne + m) * ndim + i];
}

void or_vec_as_node_dofval(const i64* dofs, i64 nnode, int ndim, const double* dofval, double* nodevec)
{
    for (i64 n = 0; n < nnode; ++n)
        for (int i = 0; i < ndim; ++i) nodevec[n * ndim + i] = dofval[dofs[n * ndim + i]];
}

void or_vec_as_node_elemvec(const i64* conn, i64 nelem, int nne, i64 nnode, int ndim,
                            const double* elemvec, double* nodevec)
{
    for (i64 k = 0; k < nnode * ndim; ++k) nodevec[k] = 0.0;
    for (i64 e = 0; e < nelem; ++e)
        for (int m = 0; m < nne; ++m)
            for (int i = 0; i < ndim; ++i)
                nodevec[conn[e * nne + m] * ndim + i] = elemvec[(e * nne + m) * ndim + i];
}

void or_vec_assemble_dofs_nodevec(const i64* dofs, i64 nnode, int ndim, i64 ndof,
                                  const double* nodevec, double* dofval)
{
    for (i64 d = 0; d < ndof; ++d) dofval[d] = 0.0;
    for (i64 n = 0; n < nnode; ++n)
        for (int i = 0; i < ndim; ++i) dofval[dofs[n * ndim + i]] += nodevec[n * ndim + i];
}

void or_vec_assemble_dofs_elemvec(const i64* conn, const i64* dofs, i64 nelem, int nne, int ndim,
                                  i64 ndof, const double* elemvec, double* dofval)
{
    for (i64 d = 0; d < ndof; ++d) dofval[d] = 0.0;
    for (i64 e = 0; e < nelem; ++e)
        for (int m = 0; m < nne; ++m)
            for (int i = 0; i < ndim; ++i)
                dofval[dofs[conn[e * nne + m] * ndim + i]] += elemvec[(e * nne + m) * ndim + i];
}

// Vector.h:746-757: AsNode(AssembleDofs(elemvec)).
void or_vec_assemble_node_elemvec(const i64* conn, const i64* dofs, i64 nelem, int nne, i64 nnode,
                                  int ndim, i64 ndof, const double* elemvec, double* nodevec)
{
    std::vector<double> dv(ndof);
    or_vec_assemble_dofs_elemvec(conn, dofs, nelem, nne, ndim, ndof, elemvec, dv.data());
    or_vec_as_node_dofval(dofs, nnode, ndim, dv.data(), nodevec);
}

// ================================================================================================
// MatrixDiagonal: MatrixDiagonal.h:150-167 (assemble: diagonal entries only, serial),
// :212-259 (dot / solve), :268-280 (inverse)
// ================================================================================================

void or_diag_assemble(const i64* conn, const i64* dofs, i64 nelem, int nne, int ndim, i64 ndof,
                      const double* elemmat, double* A)
{
    const int K = nne * ndim;
    for (i64 d = 0; d < ndof; ++d) A[d] = 0.0;
    for (i64 e = 0; e < nelem; ++e)
        for (int m = 0; m < nne; ++m)
            for (int i = 0; i < ndim; ++i)
                A[dofs[conn[e * nne + m] * ndim + i]] += elemmat[e * K * K + (m * ndim + i) * K + m * ndim + i];
}

void or_diag_inverse(i64 ndof, const double* A, double* inv)
{
    for (i64 d = 0; d < ndof; ++d) inv[d] = 1.0 / A[d];
}

// ================================================================================================
// Sparse Matrix: Matrix.h:704-727 (triplets in e,m,i,n,j order, duplicates summed in triplet
// order by setFromTriplets) and :804-815 (SpMV, ascending column order per row).
// CSR with sorted columns; two-call: pass NULL col/val to get nnz.
// ================================================================================================

i64 or_matrix_assemble(const i64* conn, const i64* dofs, i64 nelem, int nne, int ndim, i64 ndof,
                       const double* elemmat, i64* rowptr, i64* col, double* val)
{
    const int K = nne * ndim;
    std::vector<std::vector<std::pair<i64, double>>> rows(ndof);
    for (i64 e = 0; e < nelem; ++e)
        for (int m = 0; m < nne; ++m)
            for (int i = 0; i < ndim; ++i) {
                i64 r = dofs[conn[e * nne + m] * ndim + i];
                for (int n = 0; n < nne; ++n)
                    for (int j = 0; j < ndim; ++j) {
                        i64 c = dofs[conn[e * nne + n] * ndim + j];
                        double v = elemmat ? elemmat[e * K * K + (m * ndim + i) * K + n * ndim + j] : 0.0;
                        rows[r].push_back({c, v});
                    }
            }
    i64 nnz = 0;
    std::vector<std::vector<std::pair<i64, double>>> merged(ndof);
    for (i64 r = 0; r < ndof; ++r) {
        auto& row = rows[r];
        std::stable_sort(row.begin(), row.end(),
                         [](const std::pair<i64, double>& a, const std::pair<i64, double>& b) {
                             return a.first < b.first;
                         });
        auto& mr = merged[r];
        for (auto& p : row) {
            if (!mr.empty() && mr.back().first == p.first) mr.back().second += p.second;
            else mr.push_back(p);
        }
        nnz += static_cast<i64>(mr.size());
    }
    if (rowptr) {
        rowptr[0] = 0;
        for (i64 r = 0; r < ndof; ++r) rowptr[r + 1] = rowptr[r] + static_cast<i64>(merged[r].size());
    }
    if (col && val) {
        i64 k = 0;
        for (i64 r = 0; r < ndof; ++r)
            for (auto& p : merged[r]) {
                col[k] = p.first;
                val[k] = p.second;
                ++k;
            }
    }
    return nnz;
}

void or_csr_dot(i64 nrows, const i64* rowptr, const i64* col, const double* val, const double* x, double* b)
{
    for (i64 r = 0; r < nrows; ++r) {
        double s = 0.0;
        for (i64 k = rowptr[r]; k < rowptr[r + 1]; ++k) s += val[k] * x[col[k]];
        b[r] = s;
    }
}

// ================================================================================================
// Tyings: TyingsPeriodic.h:82-121 (Periodic ctor), :238-343 (Cdi/Cdu/Cdp), :361-439 (Control);
// Mesh.h:2384-2454 (Reorder).
// ================================================================================================

// Control: appends ndim control nodes (coordinates 0) with dofs max(dofs)+1+arange(ndim*ndim).
void or_tyings_control(const double* coor, const i64* dofs, i64 nnode, int ndim, double* coor_out,
                       i64* dofs_out, i64* control_dofs, i64* control_nodes)
{
    i64 mx = 0;
    for (i64 i = 0; i < nnode * ndim; ++i) mx = std::max(mx, dofs[i]);
    for (i64 i = 0; i < nnode * ndim; ++i) {
        coor_out[i] = coor[i];
        dofs_out[i] = dofs[i];
    }
    for (int i = 0; i < ndim * ndim; ++i) {
        control_dofs[i] = mx + 1 + i;
        coor_out[nnode * ndim + i] = 0.0;
        dofs_out[nnode * ndim + i] = mx + 1 + i;
    }
    for (int i = 0; i < ndim; ++i) control_nodes[i] = nnode + i;
}

// setdiff1d semantics: sorted unique values of a not in b.
static std::vector<i64> setdiff1d(const std::vector<i64>& a, const std::vector<i64>& b)
{
    std::vector<i64> ua = unique_sorted(a.data(), static_cast<i64>(a.size()));
    std::vector<i64> ub = unique_sorted(b.data(), static_cast<i64>(b.size()));
    std::vector<i64> r;
    std::set_difference(ua.begin(), ua.end(), ub.begin(), ub.end(), std::back_inserter(r));
    return r;
}

// Periodic: reorders dofs to [iiu, iip, iid].  sizes_out = {nnu, nnp, nni, nnd}.
void or_tyings_periodic(const i64* dofs, i64 nnode, int ndim, const i64* control, const i64* ties,
                        i64 nties, const i64* iip, i64 niip, i64* dofs_out, i64* control_out,
                        i64* sizes_out)
{
    std::vector<i64> iid;
    for (i64 t = 0; t < nties; ++t)
        for (int j = 0; j < ndim; ++j) iid.push_back(dofs[ties[t * 2 + 1] * ndim + j]);
    std::vector<i64> all(dofs, dofs + nnode * ndim);
    std::vector<i64> iii = setdiff1d(all, iid);
    std::vector<i64> vip(iip, iip + niip);
    std::vector<i64> iiu = setdiff1d(iii, vip);
    i64 n = 0;
    for (auto* v : {&iiu, &vip, &iid})
        for (auto j : *v) n = std::max(n, j + 1);
    std::vector<i64> renum(n, 0);
    i64 k = 0;
    for (auto* v : {&iiu, &vip, &iid})
        for (auto j : *v) renum[j] = k++;
    for (i64 i = 0; i < nnode * ndim; ++i) dofs_out[i] = renum[dofs[i]];
    for (int i = 0; i < ndim * ndim; ++i) control_out[i] = renum[control[i]];
    sizes_out[0] = static_cast<i64>(iiu.size());
    sizes_out[1] = static_cast<i64>(vip.size());
    sizes_out[2] = static_cast<i64>(iii.size());
    sizes_out[3] = static_cast<i64>(iid.size());
}

// Tying matrices as CSR (duplicates summed in triplet order, columns sorted).
// which: 0 = Cdi [nnd x nni], 1 = Cdu [nnd x nnu], 2 = Cdp [nnd x nnp].
// dofs/control are the reordered ones.  Returns nnz (two-call).
i64 or_tyings_C(int which, const double* coor, const i64* dofs, int ndim, const i64* control,
                const i64* ties, i64 nties, i64 nnu, i64* rowptr, i64* col, double* val)
{
    i64 nrows = nties * ndim;
    std::vector<std::vector<std::pair<i64, double>>> rows(nrows);
    for (i64 i = 0; i < nties; ++i)
        for (int j = 0; j < ndim; ++j) {
            i64 ni = ties[i * 2 + 0], nd = ties[i * 2 + 1];
            i64 r = i * ndim + j;
            i64 c0 = dofs[ni * ndim + j];
            bool keep0 = which == 0 || (which == 1 && c0 < nnu) || (which == 2 && c0 >= nnu);
            if (keep0) rows[r].push_back({which == 2 ? c0 - nnu : c0, +1.0});
            for (int k = 0; k < ndim; ++k) {
                i64 c = control[j * ndim + k];
                bool keep = which == 0 || (which == 1 && c < nnu) || (which == 2 && c >= nnu);
                if (keep)
                    rows[r].push_back({which == 2 ? c - nnu : c, coor[nd * ndim + k] - coor[ni * ndim + k]});
            }
        }
    i64 nnz = 0;
    std::vector<std::vector<std::pair<i64, double>>> merged(nrows);
    for (i64 r = 0; r < nrows; ++r) {
        auto& row = rows[r];
        std::stable_sort(row.begin(), row.end(),
                         [](const std::pair<i64, double>& a, const std::pair<i64, double>& b) {
                             return a.first < b.first;
                         });
        for (auto& p : row) {
            if (!merged[r].empty() && merged[r].back().first == p.first) merged[r].back().second += p.second;
            else merged[r].push_back(p);
        }
        nnz += static_cast<i64>(merged[r].size());
    }
    if (rowptr) {
        rowptr[0] = 0;
        for (i64 r = 0; r < nrows; ++r) rowptr[r + 1] = rowptr[r] + static_cast<i64>(merged[r].size());
    }
    if (col && val) {
        i64 k = 0;
        for (i64 r = 0; r < nrows; ++r)
            for (auto& p : merged[r]) {
                col[k] = p.first;
                val[k] = p.second;
                ++k;
            }
    }
    return nnz;
}

// ================================================================================================
// Material (not in the reference; external GMat* packages at the call sites
// docs/examples/dynamics/Elastic-VelocityVerlet/main.cpp:64-74,125).  Contract:
//   sigma = K tr(eps) I + 2 G dev(eps),  dev(eps) = eps - tr(eps)/3 I  (3-D tensors; a tdim = 2
//   strain is the in-plane block of a plane-strain 3-D strain).  Parity unpinned by reference tests.
// ================================================================================================

void or_elastic_stress(i64 nq, int td, const double* eps, const double* K, const double* G, double* sig)
{
    for (i64 p = 0; p < nq; ++p) {
        const double* e = eps + p * td * td;
        double* s = sig + p * td * td;
        double tr = (td == 3) ? e[0] + e[4] + e[8] : e[0] + e[3];
        double k = K[p], g = G[p];
        for (int i = 0; i < td; ++i)
            for (int j = 0; j < td; ++j) {
                if (i == j) s[i * td + j] = k * tr + 2.0 * g * (e[i * td + j] - tr / 3.0);
                else s[i * td + j] = 2.0 * g * e[i * td + j];
            }
    }
}

// C_ijkl = K d_ij d_kl + 2G (1/2 (d_ik d_jl + d_il d_jk) - 1/3 d_ij d_kl), tdim^4 per point.
void or_elastic_tangent(i64 nq, int td, const double* K, const double* G, double* C)
{
    for (i64 p = 0; p < nq; ++p) {
        double* c = C + p * td * td * td * td;
        double k = K[p], g = G[p];
        for (int i = 0; i < td; ++i)
            for (int j = 0; j < td; ++j)
                for (int a = 0; a < td; ++a)
                    for (int b = 0; b < td; ++b) {
                        double dij = i == j, dkl = a == b, dik = i == a, djl = j == b, dil = i == b, djk = j == a;
                        c[((i * td + j) * td + a) * td + b] =
                            k * dij * dkl + 2.0 * g * (0.5 * (dik * djl + dil * djk) - dij * dkl / 3.0);
                    }
    }
}

// ================================================================================================
// Element::Quad4::QuadratureAxisymmetric (include/GooseFEM/ElementQuad4Axisymmetric.h:33-467).
// Nodal coordinates (z, r); tensors in (r, theta, z) order.  Of the B-matrix only five components
// per node are non-zero (:401-438): B(m,r,r,r) = B(m,r,z,z) = dN/dr, B(m,t,t,r) = N / r_q,
// B(m,z,r,r) = B(m,z,z,z) = dN/dz; dV = w J 2 pi r_q.  Geometry is stored as
// G[e, q, m, {dNdr, Nr, dNdz}] and vol[e, q].
// ================================================================================================

void or_axi_geometry(i64 nelem, int nip, const double* x, const double* N, const double* dNxi, const double* w,
                     double* G, double* vol)
{
    for (i64 e = 0; e < nelem; ++e) {
        const double* xe = x + e * 8;
        for (int q = 0; q < nip; ++q) {
            const double* g = dNxi + q * 8;
            const double* n = N + q * 4;
            double J[2][2], Ji[2][2];
            J[0][0] = g[0] * xe[0] + g[2] * xe[2] + g[4] * xe[4] + g[6] * xe[6];
            J[0][1] = g[0] * xe[1] + g[2] * xe[3] + g[4] * xe[5] + g[6] * xe[7];
            J[1][0] = g[1] * xe[0] + g[3] * xe[2] + g[5] * xe[4] + g[7] * xe[6];
            J[1][1] = g[1] * xe[1] + g[3] * xe[3] + g[5] * xe[5] + g[7] * xe[7];
            const double det = inv2(J, Ji);
            const double rq = n[0] * xe[1] + n[1] * xe[3] + n[2] * xe[5] + n[3] * xe[7];
            double* o = G + (e * nip + q) * 12;
            for (int m = 0; m < 4; ++m) {
                o[3 * m + 0] = Ji[1][0] * g[2 * m] + Ji[1][1] * g[2 * m + 1];
                o[3 * m + 1] = 1.0 / rq * n[m];
                o[3 * m + 2] = Ji[0][0] * g[2 * m] + Ji[0][1] * g[2 * m + 1];
            }
            vol[e * nip + q] = w[q] * det * 2.0 * M_PI * rq;
        }
    }
}

// gradN_vector (mode 0, :140-169), gradN_vector_T (1, :171-200), symGradN_vector (2, :202-233)
void or_axi_grad(int mode, i64 nelem, int nip, const double* G, const double* elemvec, double* out)
{
    for (i64 e = 0; e < nelem; ++e) {
        const double* u = elemvec + e * 8; // u[m*2 + 0] = u_z, u[m*2 + 1] = u_r
        for (int q = 0; q < nip; ++q) {
            const double* b = G + (e * nip + q) * 12;
            double* t = out + (e * nip + q) * 9;
            for (int k = 0; k < 9; ++k) t[k] = 0.0;
            // sums over the four nodes in node order
            const double rr = b[0] * u[1] + b[3] * u[3] + b[6] * u[5] + b[9] * u[7];
            const double tt = b[1] * u[1] + b[4] * u[3] + b[7] * u[5] + b[10] * u[7];
            const double zz = b[2] * u[0] + b[5] * u[2] + b[8] * u[4] + b[11] * u[6];
            t[0] = rr;
            t[4] = tt;
            t[8] = zz;
            if (mode == 2) {
                const double s = 0.5 * (b[0] * u[0] + b[3] * u[2] + b[6] * u[4] + b[9] * u[6] + b[2] * u[1] +
                                        b[5] * u[3] + b[8] * u[5] + b[11] * u[7]);
                t[6] = s;
                t[2] = s;
            }
            else {
                const double dzr = b[0] * u[0] + b[3] * u[2] + b[6] * u[4] + b[9] * u[6]; // B(m,r,z,z) u_z
                const double drz = b[2] * u[1] + b[5] * u[3] + b[8] * u[5] + b[11] * u[7]; // B(m,z,r,r) u_r
                t[mode == 0 ? 2 : 6] = dzr;
                t[mode == 0 ? 6 : 2] = drz;
            }
        }
    }
}

// int_N_scalar_NT_dV (:235-266)
void or_axi_int_N_scalar_NT_dV(i64 nelem, int nip, const double* N, const double* vol, const double* qs,
                               double* M)
{
    for (i64 e = 0; e < nelem; ++e) {
        double* Me = M + e * 64;
        for (int k = 0; k < 64; ++k) Me[k] = 0.0;
        for (int q = 0; q < nip; ++q) {
            const double* n = N + q * 4;
            const double v = vol[e * nip + q], rho = qs[e * nip + q];
            for (int m = 0; m < 4; ++m)
                for (int k = 0; k < 4; ++k) {
                    Me[(m * 2) * 8 + k * 2] += n[m] * rho * n[k] * v;
                    Me[(m * 2 + 1) * 8 + k * 2 + 1] += n[m] * rho * n[k] * v;
                }
        }
    }
}

// int_gradN_dot_tensor2_dV (:268-296)
void or_axi_int_tensor2(i64 nelem, int nip, const double* G, const double* vol, const double* sig, double* f)
{
    for (i64 e = 0; e < nelem; ++e) {
        double* fe = f + e * 8;
        for (int k = 0; k < 8; ++k) fe[k] = 0.0;
        for (int q = 0; q < nip; ++q) {
            const double* b = G + (e * nip + q) * 12;
            const double* s = sig + (e * nip + q) * 9;
            const double v = vol[e * nip + q];
            for (int m = 0; m < 4; ++m) {
                fe[2 * m] += v * (b[3 * m + 2] * s[8] + b[3 * m] * s[2]);
                fe[2 * m + 1] += v * (b[3 * m] * s[0] + b[3 * m + 1] * s[4] + b[3 * m + 2] * s[6]);
            }
        }
    }
}

// int_gradN_dot_tensor4_dot_gradNT_dV (:298-398).  The 25 products per node pair run over the
// five non-zero B slots a (outer) and b (inner), slot = (i, j, value, nodal component):
//   (r,r: dNdr, u_r) (t,t: Nr, u_r) (z,z: dNdz, u_z) (r,z: dNdr, u_z) (z,r: dNdz, u_r);
// K(m, ca; n, cb) += B_a(m) C(i_a, j_a, j_b, i_b) B_b(n) dV.
void or_axi_int_tensor4(i64 nelem, int nip, const double* G, const double* vol, const double* C, double* K)
{
    const int si[5] = {0, 1, 2, 0, 2}, sj[5] = {0, 1, 2, 2, 0}, sv[5] = {0, 1, 2, 0, 2}, sc[5] = {1, 1, 0, 0, 1};
    for (i64 e = 0; e < nelem; ++e) {
        double* Ke = K + e * 64;
        for (int k = 0; k < 64; ++k) Ke[k] = 0.0;
        for (int q = 0; q < nip; ++q) {
            const double* b = G + (e * nip + q) * 12;
            const double* c = C + (e * nip + q) * 81;
            const double v = vol[e * nip + q];
            for (int m = 0; m < 4; ++m)
                for (int n = 0; n < 4; ++n)
                    for (int a = 0; a < 5; ++a)
                        for (int bb = 0; bb < 5; ++bb) {
                            const double cab = c[((si[a] * 3 + sj[a]) * 3 + sj[bb]) * 3 + si[bb]];
                            Ke[(m * 2 + sc[a]) * 8 + n * 2 + sc[bb]] += b[3 * m + sv[a]] * cab * b[3 * n + sv[bb]] * v;
                        }
        }
    }
}

// ================================================================================================
// J2 elasto-plasticity with linear isotropic hardening, small strain, 3-d tensors: the law behind
// GMatElastoPlastic.Cartesian3d.LinearHardening2d (python-gmatelastoplastic >= 0.3.1,
// environment.yaml:16; called at docs/examples/statics/periodic_elastoplastic.py:71-104).  GMat is
// not part of the reference, so this restates the published radial-return algorithm and its
// consistent tangent (de Souza Neto, Peric, Owen, "Computational Methods for Plasticity", 2008,
// Box 7.3 and eq. 7.120; Simo and Hughes, "Computational Inelasticity", 1998, Box 3.2):
//   Ee = Eps - Epsp_t;  sigm = K tr(Ee);  s = 2G dev(Ee);  q = sqrt(3/2 s:s)
//   phi = q - (sigy0 + H epsp_t);  elastic if phi <= 0, else
//   dg = phi / (3G + H);  Sig = sigm I + (1 - 3G dg / q) s;  Epsp = Epsp_t + (3/2 dg / q) s;
//   epsp = epsp_t + dg
//   C = K I (x) I + 2G f I_d + c n (x) n,  f = 1 - 3G dg / q,  c = 6G^2 (dg / q - 1 / (3G + H)),
//   n = s / |s|  (f = 1, c = 0 in the elastic branch);  I_d = I_s - I (x) I / 3.
// Parity with GMat's own implementation is unpinned (absent package); the tests pin this
// restatement by the identities of the law (consistency, incompressible plastic flow, exact radial
// return under proportional loading, the tangent as the derivative of the stress, elastic
// unloading, the elastic limit).
// ================================================================================================

void or_plastic_refresh(i64 nq, const double* Eps, const double* K, const double* G, const double* sy0,
                        const double* H, const double* Epsp_t, const double* epsp_t, double* Epsp, double* epsp,
                        double* Sig, double* C)
{
    for (i64 p = 0; p < nq; ++p) {
        const double* e = Eps + 9 * p;
        const double* ept = Epsp_t + 9 * p;
        double ee[9], s[9], n[9];
        for (int k = 0; k < 9; ++k) ee[k] = e[k] - ept[k];
        const double tr = ee[0] + ee[4] + ee[8];
        const double epsm = tr / 3.0;
        const double k_ = K[p], g = G[p];
        double ss = 0.0;
        for (int k = 0; k < 9; ++k) {
            s[k] = 2.0 * g * (k % 4 == 0 ? ee[k] - epsm : ee[k]);
            ss += s[k] * s[k];
        }
        const double q = std::sqrt(1.5 * ss);
        const double phi = q - (sy0[p] + H[p] * epsp_t[p]);
        double f = 1.0, c = 0.0, dg = 0.0;
        for (int k = 0; k < 9; ++k) n[k] = 0.0;
        if (phi > 0.0) {
            dg = phi / (3.0 * g + H[p]);
            f = 1.0 - 3.0 * g * dg / q;
            c = 6.0 * g * g * (dg / q - 1.0 / (3.0 * g + H[p]));
            const double nrm = std::sqrt(ss);
            for (int k = 0; k < 9; ++k) n[k] = s[k] / nrm;
        }
        const double sigm = k_ * tr;
        const double a = 1.5 * dg / q;
        for (int k = 0; k < 9; ++k) {
            Sig[9 * p + k] = f * s[k] + (k % 4 == 0 ? sigm : 0.0);
            Epsp[9 * p + k] = phi > 0.0 ? ept[k] + a * s[k] : ept[k];
        }
        epsp[p] = epsp_t[p] + dg;
        if (!C) continue;
        const double g2f = 2.0 * g * f;
        double* cp = C + 81 * p;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int a_ = 0; a_ < 3; ++a_)
                    for (int b = 0; b < 3; ++b) {
                        const double dij = i == j, dkl = a_ == b, dik = i == a_, djl = j == b, dil = i == b, djk = j == a_;
                        cp[((i * 3 + j) * 3 + a_) * 3 + b] = k_ * dij * dkl +
                                                             g2f * (0.5 * (dik * djl + dil * djk) - dij * dkl / 3.0) +
                                                             c * (n[i * 3 + j] * n[a_ * 3 + b]);
                    }
    }
}

// ================================================================================================
// Velocity-Verlet step of docs/examples/dynamics/Elastic-VelocityVerlet/main.cpp:110-157 on a
// Quad4 planar mesh with per-element (K, G), state as dofval.  Prescribed DOFs carry inv_mass = 0
// (MatrixDiagonalPartitioned::solve touches only unknown DOFs, MatrixDiagonalPartitioned.h:305-335),
// so their acceleration stays 0 and their velocity stays the prescribed value.
//   u <- (u + dt v) + c a,  c = 0.5 dt^2 (precomputed by the caller)
//   fint = AssembleDofs(Int_gradN_dot_tensor2_dV(sigma(SymGradN_vector(AsElement(u)))))
//   a_new = inv_mass * (fext - fint);  v <- v + (0.5 dt) (a + a_new);  a <- a_new
// Geometry (dNx, vol) is precomputed by or_quad_compute_dN on the Gauss rule.  fint_out optional.
// ================================================================================================

void or_vv_steps(i64 nsteps, int et, const i64* conn, const i64* dofs, i64 nelem, i64 ndof,
                 const double* dNx, const double* vol, const double* Ke, const double* Ge,
                 const double* inv_mass, const double* fext, double dt, double c, double* u, double* v,
                 double* a, double* fint_out)
{
    const int nne = et_nne(et), nd = et_ndim(et), td = et_tdim(et), nip = nne;
    std::vector<double> ue(nelem * nne * nd), eps(nelem * nip * td * td), sig(nelem * nip * td * td),
        fe(nelem * nne * nd), fint(ndof), Kq(nelem * nip), Gq(nelem * nip);
    for (i64 e = 0; e < nelem; ++e)
        for (int q = 0; q < nip; ++q) {
            Kq[e * nip + q] = Ke[e];
            Gq[e * nip + q] = Ge[e];
        }
    for (i64 s = 0; s < nsteps; ++s) {
        for (i64 d = 0; d < ndof; ++d) u[d] = u[d] + dt * v[d] + c * a[d];
        or_vec_as_element_dofval(conn, dofs, nelem, nne, nd, u, ue.data());
        or_quad_grad(et, 2, nelem, nip, dNx, ue.data(), eps.data());
        or_elastic_stress(nelem * nip, td, eps.data(), Kq.data(), Gq.data(), sig.data());
        or_quad_int_gradN_dot_tensor2_dV(et, nelem, nip, dNx, vol, sig.data(), fe.data());
        or_vec_assemble_dofs_elemvec(conn, dofs, nelem, nne, nd, ndof, fe.data(), fint.data());
        for (i64 d = 0; d < ndof; ++d) {
            double fres = (fext ? fext[d] : 0.0) - fint[d];
            double an = inv_mass[d] * fres;
            v[d] = v[d] + 0.5 * dt * (a[d] + an);
            a[d] = an;
        }
    }
    if (fint_out)
        for (i64 d = 0; d < ndof; ++d) fint_out[d] = fint[d];
}

// The same steps on all host threads (OpenMP): the element passes run the functions above on
// contiguous element chunks, and the assembly is a per-DOF gather of the element-vector entries in
// ascending (e, m, i) order, the order of the serial scatter, so the results are bit-identical to
// or_vv_steps.  Returns the number of threads used.  This is the "all host cores" CPU baseline; the
// reference itself cannot run its OpenMP loops (Element.h:122-145 returns from inside a parallel
// loop, which GCC rejects with -fopenmp).
i64 or_vv_steps_parallel(i64 nsteps, int et, const i64* conn, const i64* dofs, i64 nelem, i64 ndof,
                         const double* dNx, const double* vol, const double* Ke, const double* Ge,
                         const double* inv_mass, const double* fext, double dt, double c, double* u, double* v,
                         double* a, double* fint_out)
{
    const int nne = et_nne(et), nd = et_ndim(et), td = et_tdim(et), nip = nne;
    std::vector<double> ue(nelem * nne * nd), eps(nelem * nip * td * td), sig(nelem * nip * td * td),
        fe(nelem * nne * nd), fint(ndof), Kq(nelem * nip), Gq(nelem * nip);
    for (i64 e = 0; e < nelem; ++e)
        for (int q = 0; q < nip; ++q) {
            Kq[e * nip + q] = Ke[e];
            Gq[e * nip + q] = Ge[e];
        }
    // DOF -> element-vector entries k = (e nne + m) nd + i, ascending (counting sort)
    std::vector<i64> ptr(ndof + 1, 0), idx(nelem * nne * nd);
    for (i64 k = 0; k < nelem * nne * nd; ++k) ++ptr[dofs[conn[k / nd] * nd + k % nd] + 1];
    for (i64 d = 0; d < ndof; ++d) ptr[d + 1] += ptr[d];
    {
        std::vector<i64> fill(ptr.begin(), ptr.end() - 1);
        for (i64 k = 0; k < nelem * nne * nd; ++k) idx[fill[dofs[conn[k / nd] * nd + k % nd]]++] = k;
    }
    const i64 chunk = 4096;
    const i64 nchunk = (nelem + chunk - 1) / chunk;
    int nthreads = 1;
#pragma omp parallel
    {
#pragma omp single
        {
#ifdef _OPENMP
            nthreads = omp_get_num_threads();
#endif
        }
    }
    for (i64 s = 0; s < nsteps; ++s) {
#pragma omp parallel for schedule(static)
        for (i64 d = 0; d < ndof; ++d) u[d] = u[d] + dt * v[d] + c * a[d];
#pragma omp parallel for schedule(dynamic, 1)
        for (i64 b = 0; b < nchunk; ++b) {
            const i64 e0 = b * chunk, n = std::min(chunk, nelem - e0);
            or_vec_as_element_dofval(conn + e0 * nne, dofs, n, nne, nd, u, ue.data() + e0 * nne * nd);
            or_quad_grad(et, 2, n, nip, dNx + e0 * nip * nne * nd, ue.data() + e0 * nne * nd,
                         eps.data() + e0 * nip * td * td);
            or_elastic_stress(n * nip, td, eps.data() + e0 * nip * td * td, Kq.data() + e0 * nip,
                              Gq.data() + e0 * nip, sig.data() + e0 * nip * td * td);
            or_quad_int_gradN_dot_tensor2_dV(et, n, nip, dNx + e0 * nip * nne * nd, vol + e0 * nip,
                                             sig.data() + e0 * nip * td * td, fe.data() + e0 * nne * nd);
        }
#pragma omp parallel for schedule(static)
        for (i64 d = 0; d < ndof; ++d) {
            double f = 0.0;
            for (i64 j = ptr[d]; j < ptr[d + 1]; ++j) f += fe[idx[j]];
            fint[d] = f;
            double fres = (fext ? fext[d] : 0.0) - f;
            double an = inv_mass[d] * fres;
            v[d] = v[d] + 0.5 * dt * (a[d] + an);
            a[d] = an;
        }
    }
    if (fint_out)
        for (i64 d = 0; d < ndof; ++d) fint_out[d] = fint[d];
    return nthreads;
}

// ================================================================================================
// Plain FP64 Jacobi-preconditioned CG (not in the reference, whose solves are direct:
// Matrix.h:866-909).  Operator is a CSR matrix.  Returns the iteration count; res_out receives the
// true relative residual ||b - A x|| / ||b||.
// ================================================================================================

i64 or_pcg_csr(i64 n, const i64* rowptr, const i64* col, const double* val, const double* diag,
               const double* b, double* x, double tol, i64 maxiter, double* res_out)
{
    std::vector<double> r(n), z(n), p(n), Ap(n);
    or_csr_dot(n, rowptr, col, val, x, Ap.data());
    double bb = 0.0;
    for (i64 i = 0; i < n; ++i) {
        r[i] = b[i] - Ap[i];
        bb += b[i] * b[i];
    }
    double nb = std::sqrt(bb);
    if (nb == 0.0) nb = 1.0;
    double rz = 0.0, rr = 0.0;
    for (i64 i = 0; i < n; ++i) {
        z[i] = r[i] / diag[i];
        p[i] = z[i];
        rz += r[i] * z[i];
        rr += r[i] * r[i];
    }
    i64 it = 0;
    while (std::sqrt(rr) / nb > tol && it < maxiter) {
        or_csr_dot(n, rowptr, col, val, p.data(), Ap.data());
        double pAp = 0.0;
        for (i64 i = 0; i < n; ++i) pAp += p[i] * Ap[i];
        double alpha = rz / pAp;
        rr = 0.0;
        for (i64 i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * Ap[i];
            rr += r[i] * r[i];
        }
        double rz_new = 0.0;
        for (i64 i = 0; i < n; ++i) {
            z[i] = r[i] / diag[i];
            rz_new += r[i] * z[i];
        }
        double beta = rz_new / rz;
        rz = rz_new;
        for (i64 i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        ++it;
    }
    or_csr_dot(n, rowptr, col, val, x, Ap.data());
    double rt = 0.0;
    for (i64 i = 0; i < n; ++i) rt += (b[i] - Ap[i]) * (b[i] - Ap[i]);
    if (res_out) *res_out = std::sqrt(rt) / nb;
    return it;
}

// ================================================================================================
// Portable input generator: SplitMix64 -> 53-bit uniform -> Box-Muller (host FP64).  Shared by the
// product's host code by specification (same constants), so inputs are identical everywhere.
// ================================================================================================

void or_randn(uint64_t seed, i64 n, double mean, double stddev, double* out)
{
    uint64_t s = seed;
    auto next = [&]() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    const double two_pi = 6.283185307179586476925286766559;
    for (i64 i = 0; i < n; i += 2) {
        double u1 = (static_cast<double>(next() >> 11) + 1.0) * (1.0 / 9007199254740992.0);
        double u2 = static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0);
        double r = std::sqrt(-2.0 * std::log(u1));
        out[i] = mean + stddev * r * std::cos(two_pi * u2);
        if (i + 1 < n) out[i + 1] = mean + stddev * r * std::sin(two_pi * u2);
    }
}

} // extern "C"
