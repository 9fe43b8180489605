/*
 * GooseFEM/GooseFEM.h -- header-only C++ face of goosefem_amd: the class names and call forms of
 * the reference's C++ API (the include/GooseFEM headers of tdegeus/GooseFEM) for the element-quadrature and
 * assembly path, forwarded to the C ABI in goosefem_amd.h (link with -lgoosefem_amd).
 *
 * Containers: any array with data(), size(), shape() / shape(i) and a row-major layout is accepted
 * (xt::xtensor, xt::pytensor, GooseFEM::array_type::tensor below).  Results are returned as
 * GooseFEM::array_type::tensor.  Host containers are copied to the GPU for the call and back, as the
 * reference's value semantics require; GooseFEM::amd::device_array arguments stay in HBM and
 * results can be written into them without host traffic (the lower-case write-into forms).
 * Errors: every nonzero status of the C ABI throws std::runtime_error with the library's message
 * (the reference's GOOSEFEM_CHECK behaviour, include/GooseFEM/config.h:65-98).
 */
#ifndef GOOSEFEM_AMD_GOOSEFEM_H
#define GOOSEFEM_AMD_GOOSEFEM_H

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <initializer_list>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../goosefem_amd.h"

namespace GooseFEM {

namespace amd {

inline void check(int rc)
{
    if (rc != GF_OK) throw std::runtime_error(std::string("goosefem_amd: ") + gf_last_error());
}

/* Owning row-major host array with xtensor-like accessors. */
template <class T>
class array {
public:
    array() = default;
    explicit array(std::vector<size_t> shape, T fill = T(0)) : m_shape(std::move(shape)), m_data(count(), fill) {}
    array(std::initializer_list<size_t> shape) : m_shape(shape), m_data(count(), T(0)) {}
    static array from_shape(std::vector<size_t> shape) { return array(std::move(shape)); }
    T* data() { return m_data.data(); }
    const T* data() const { return m_data.data(); }
    size_t size() const { return m_data.size(); }
    size_t dimension() const { return m_shape.size(); }
    const std::vector<size_t>& shape() const { return m_shape; }
    size_t shape(size_t i) const { return m_shape.at(i); }
    template <class... I>
    T& operator()(I... idx) { return m_data[offset({static_cast<size_t>(idx)...})]; }
    template <class... I>
    const T& operator()(I... idx) const { return m_data[offset({static_cast<size_t>(idx)...})]; }
    T& operator[](size_t i) { return m_data[i]; }
    const T& operator[](size_t i) const { return m_data[i]; }
    typename std::vector<T>::iterator begin() { return m_data.begin(); }
    typename std::vector<T>::iterator end() { return m_data.end(); }
    typename std::vector<T>::const_iterator begin() const { return m_data.begin(); }
    typename std::vector<T>::const_iterator end() const { return m_data.end(); }
    void fill(T v) { std::fill(m_data.begin(), m_data.end(), v); }

private:
    size_t count() const
    {
        return std::accumulate(m_shape.begin(), m_shape.end(), size_t(1), std::multiplies<size_t>());
    }
    size_t offset(std::initializer_list<size_t> idx) const
    {
        size_t o = 0, d = 0;
        for (size_t i : idx) o = o * m_shape[d++] + i;
        return o;
    }
    std::vector<size_t> m_shape;
    std::vector<T> m_data;
};

/* Owning device (HBM) array. */
template <class T>
class device_array {
public:
    device_array() = default;
    explicit device_array(std::vector<size_t> shape) : m_shape(std::move(shape)) { allocate(); }
    template <class H>
    static device_array from_host(const H& h)
    {
        std::vector<size_t> s;
        for (size_t i = 0; i < h.dimension(); ++i) s.push_back(h.shape()[i]);
        device_array d(s);
        check(gf_memcpy_h2d(d.m_ptr, h.data(), static_cast<int64_t>(d.bytes())));
        return d;
    }
    device_array(const device_array&) = delete;
    device_array& operator=(const device_array&) = delete;
    device_array(device_array&& o) noexcept : m_shape(std::move(o.m_shape)), m_ptr(o.m_ptr) { o.m_ptr = nullptr; }
    device_array& operator=(device_array&& o) noexcept
    {
        release();
        m_shape = std::move(o.m_shape);
        m_ptr = o.m_ptr;
        o.m_ptr = nullptr;
        return *this;
    }
    ~device_array() { release(); }
    T* data() { return m_ptr; }
    const T* data() const { return m_ptr; }
    size_t size() const
    {
        return std::accumulate(m_shape.begin(), m_shape.end(), size_t(1), std::multiplies<size_t>());
    }
    size_t bytes() const { return size() * sizeof(T); }
    size_t dimension() const { return m_shape.size(); }
    const std::vector<size_t>& shape() const { return m_shape; }
    size_t shape(size_t i) const { return m_shape.at(i); }
    array<T> to_host() const
    {
        array<T> h(m_shape);
        check(gf_memcpy_d2h(h.data(), m_ptr, static_cast<int64_t>(bytes())));
        return h;
    }

private:
    void allocate()
    {
        void* p = nullptr;
        check(gf_device_malloc(&p, static_cast<int64_t>(bytes())));
        m_ptr = static_cast<T*>(p);
    }
    void release()
    {
        if (m_ptr) gf_device_free(m_ptr);
        m_ptr = nullptr;
    }
    std::vector<size_t> m_shape;
    T* m_ptr = nullptr;
};

template <class C>
struct is_device : std::false_type {};
template <class T>
struct is_device<device_array<T>> : std::true_type {};

template <class C>
std::vector<size_t> shape_of(const C& c)
{
    std::vector<size_t> s;
    for (size_t i = 0; i < c.dimension(); ++i) s.push_back(static_cast<size_t>(c.shape()[i]));
    return s;
}

/* Device view of an argument: no copy for device arrays, an upload for host containers. */
template <class C>
struct dev_in {
    device_array<double> tmp;
    const double* ptr = nullptr;
    explicit dev_in(const C& c)
    {
        if constexpr (is_device<C>::value) {
            ptr = c.data();
        }
        else {
            static_assert(std::is_same<typename std::decay<decltype(*c.data())>::type, double>::value,
                          "host containers must hold double");
            tmp = device_array<double>(shape_of(c));
            check(gf_memcpy_h2d(tmp.data(), c.data(), static_cast<int64_t>(tmp.bytes())));
            ptr = tmp.data();
        }
    }
};

/* Output for a write-into call: device arrays are written directly, host containers via a
 * device temporary copied back in finish(). */
template <class R>
struct dev_out {
    device_array<double> tmp;
    double* ptr = nullptr;
    R& ret;
    dev_out(R& r, const std::vector<size_t>& shape) : ret(r)
    {
        if constexpr (is_device<R>::value) {
            ptr = r.data();
        }
        else {
            tmp = device_array<double>(shape);
            ptr = tmp.data();
        }
    }
    void finish()
    {
        if constexpr (!is_device<R>::value) check(gf_memcpy_d2h(ret.data(), tmp.data(), static_cast<int64_t>(tmp.bytes())));
        else check(gf_stream_sync(nullptr));
    }
};

/* In-out argument: host containers are uploaded, written on the device and copied back in
 * finish(); device arrays are used in place. */
template <class R>
struct dev_inout {
    device_array<double> tmp;
    double* ptr = nullptr;
    R& ret;
    explicit dev_inout(R& r) : ret(r)
    {
        if constexpr (is_device<R>::value) {
            ptr = r.data();
        }
        else {
            tmp = device_array<double>(shape_of(r));
            check(gf_memcpy_h2d(tmp.data(), r.data(), static_cast<int64_t>(tmp.bytes())));
            ptr = tmp.data();
        }
    }
    void finish()
    {
        if constexpr (!is_device<R>::value) check(gf_memcpy_d2h(ret.data(), tmp.data(), static_cast<int64_t>(tmp.bytes())));
        else check(gf_stream_sync(nullptr));
    }
};

/* Compressed sparse rows (in place of the reference's Eigen::SparseMatrix). */
struct SparseMatrix {
    size_t rows = 0, cols = 0;
    std::vector<int64_t> rowptr, col;
    std::vector<double> val;
};

template <class C>
std::vector<int64_t> to_i64(const C& c)
{
    std::vector<int64_t> r(c.size());
    const auto* p = c.data();
    for (size_t i = 0; i < r.size(); ++i) r[i] = static_cast<int64_t>(p[i]);
    return r;
}

inline array<size_t> from_i64(const std::vector<int64_t>& v, std::vector<size_t> shape)
{
    array<size_t> r(std::move(shape));
    for (size_t i = 0; i < v.size(); ++i) r[i] = static_cast<size_t>(v[i]);
    return r;
}

} // namespace amd

namespace array_type {
template <class T, size_t N>
using tensor = amd::array<T>;
}

// =================================================================================================
// Mesh (include/GooseFEM/Mesh.h, MeshQuad4.h, MeshHex8.h)
// =================================================================================================
namespace Mesh {

inline amd::array<size_t> dofs(size_t nnode, size_t ndim)
{
    amd::array<size_t> r({nnode, ndim});
    for (size_t i = 0; i < r.size(); ++i) r[i] = i;
    return r;
}

template <class T>
inline amd::array<size_t> renumber(const T& dofs)
{
    auto in = amd::to_i64(dofs);
    std::vector<int64_t> out(in.size());
    amd::check(gf_mesh_renumber(in.data(), static_cast<int64_t>(in.size()), out.data()));
    return amd::from_i64(out, amd::shape_of(dofs));
}

template <class T>
inline amd::array<size_t> coordination(const T& conn)
{
    auto c = amd::to_i64(conn);
    int64_t nv = c.empty() ? 0 : *std::max_element(c.begin(), c.end()) + 1;
    std::vector<int64_t> out(static_cast<size_t>(nv));
    amd::check(gf_mesh_coordination(c.data(), static_cast<int64_t>(c.size()), nv, out.data()));
    return amd::from_i64(out, {static_cast<size_t>(nv)});
}

namespace detail {
inline std::vector<std::vector<size_t>> csr_lists(const std::vector<int64_t>& ptr, const std::vector<int64_t>& idx)
{
    std::vector<std::vector<size_t>> r(ptr.size() - 1);
    for (size_t i = 0; i + 1 < ptr.size(); ++i)
        for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k) r[i].push_back(static_cast<size_t>(idx[k]));
    return r;
}
} // namespace detail

/* Mesh::elem2node(conn[, dofs])  (Mesh.h:2485-2551) */
template <class E>
inline std::vector<std::vector<size_t>> elem2node(const E& conn)
{
    auto c = amd::to_i64(conn);
    int64_t nnode = c.empty() ? 0 : *std::max_element(c.begin(), c.end()) + 1, nnz = 0;
    std::vector<int64_t> ptr(nnode + 1);
    const int nne = static_cast<int>(conn.shape()[1]);
    const int64_t ne = static_cast<int64_t>(conn.shape()[0]);
    amd::check(gf_mesh_elem2node(c.data(), ne, nne, nullptr, nnode, 1, ptr.data(), nullptr, &nnz));
    std::vector<int64_t> idx(nnz);
    amd::check(gf_mesh_elem2node(c.data(), ne, nne, nullptr, nnode, 1, ptr.data(), idx.data(), &nnz));
    return detail::csr_lists(ptr, idx);
}

template <class E, class D>
inline std::vector<std::vector<size_t>> elem2node(const E& conn, const D& dofs)
{
    auto c = amd::to_i64(conn);
    auto d = amd::to_i64(dofs);
    const int64_t nnode = static_cast<int64_t>(dofs.shape()[0]);
    const int ndim = static_cast<int>(dofs.shape()[1]);
    const int nne = static_cast<int>(conn.shape()[1]);
    const int64_t ne = static_cast<int64_t>(conn.shape()[0]);
    int64_t nnz = 0;
    std::vector<int64_t> ptr(nnode + 1);
    amd::check(gf_mesh_elem2node(c.data(), ne, nne, d.data(), nnode, ndim, ptr.data(), nullptr, &nnz));
    std::vector<int64_t> idx(nnz);
    amd::check(gf_mesh_elem2node(c.data(), ne, nne, d.data(), nnode, ndim, ptr.data(), idx.data(), &nnz));
    return detail::csr_lists(ptr, idx);
}

/* Mesh::nodaltyings(dofs)  (Mesh.h:2839-2861) */
template <class D>
inline std::vector<std::vector<size_t>> nodaltyings(const D& dofs)
{
    auto d = amd::to_i64(dofs);
    const int64_t nnode = static_cast<int64_t>(dofs.shape()[0]);
    const int ndim = static_cast<int>(dofs.shape()[1]);
    int64_t nnz = 0;
    std::vector<int64_t> ptr(nnode + 1);
    amd::check(gf_mesh_nodaltyings(d.data(), nnode, ndim, ptr.data(), nullptr, &nnz));
    std::vector<int64_t> idx(nnz);
    amd::check(gf_mesh_nodaltyings(d.data(), nnode, ndim, ptr.data(), idx.data(), &nnz));
    return detail::csr_lists(ptr, idx);
}

namespace detail {

class RegularBase {
public:
    size_t nelem() const { return m_nelem; }
    size_t nnode() const { return m_nnode; }
    size_t nne() const { return m_nne; }
    size_t ndim() const { return m_ndim; }
    size_t nelx() const { return static_cast<size_t>(m_nx); }
    size_t nely() const { return static_cast<size_t>(m_ny); }
    double h() const { return m_h; }
    amd::array<size_t> dofs() const { return Mesh::dofs(m_nnode, m_ndim); }
    amd::array<size_t> nodesPeriodic() const
    {
        int64_t n = 0;
        amd::check(gf_mesh_nodes_periodic(m_nx, m_ny, m_nz, nullptr, &n));
        std::vector<int64_t> r(static_cast<size_t>(2 * n));
        amd::check(gf_mesh_nodes_periodic(m_nx, m_ny, m_nz, r.data(), &n));
        return amd::from_i64(r, {static_cast<size_t>(n), 2});
    }
    amd::array<size_t> dofsPeriodic() const
    {
        auto t = amd::to_i64(nodesPeriodic());
        std::vector<int64_t> r(m_nnode * m_ndim);
        amd::check(gf_mesh_dofs_periodic(static_cast<int64_t>(m_nnode), static_cast<int>(m_ndim), t.data(),
                                         static_cast<int64_t>(t.size() / 2), r.data()));
        return amd::from_i64(r, {m_nnode, m_ndim});
    }
    /* any named node set of the reference (e.g. "nodesBottomEdge", "nodesFrontFace") */
    amd::array<size_t> nodeset(const std::string& name) const
    {
        int64_t n = 0;
        amd::check(gf_mesh_nodeset(m_nx, m_ny, m_nz, name.c_str(), nullptr, &n));
        std::vector<int64_t> r(static_cast<size_t>(n));
        amd::check(gf_mesh_nodeset(m_nx, m_ny, m_nz, name.c_str(), r.data(), &n));
        return amd::from_i64(r, {static_cast<size_t>(n)});
    }
    size_t node(const std::string& name) const { return nodeset(name)[0]; }
    size_t nodesOrigin() const { return node("nodesOrigin"); }

protected:
    int64_t m_nx = 0, m_ny = 0, m_nz = 0;
    double m_h = 1.0;
    size_t m_nelem = 0, m_nnode = 0, m_nne = 0, m_ndim = 0;
};

} // namespace detail

namespace Quad4 {

/* Mesh::Quad4::Regular  (include/GooseFEM/MeshQuad4.h:32-195) */
class Regular : public detail::RegularBase {
public:
    Regular(size_t nelx, size_t nely, double h = 1.0)
    {
        m_nx = static_cast<int64_t>(nelx);
        m_ny = static_cast<int64_t>(nely);
        m_nz = 0;
        m_h = h;
        m_nne = 4;
        m_ndim = 2;
        m_nelem = nelx * nely;
        m_nnode = (nelx + 1) * (nely + 1);
    }
    amd::array<double> coor() const
    {
        amd::array<double> c({m_nnode, 2});
        amd::check(gf_mesh_quad4_regular(m_nx, m_ny, m_h, c.data(), nullptr));
        return c;
    }
    amd::array<size_t> conn() const
    {
        std::vector<int64_t> c(m_nelem * 4);
        amd::check(gf_mesh_quad4_regular(m_nx, m_ny, m_h, nullptr, c.data()));
        return amd::from_i64(c, {m_nelem, 4});
    }
    amd::array<size_t> nodesBottomEdge() const { return nodeset("nodesBottomEdge"); }
    amd::array<size_t> nodesTopEdge() const { return nodeset("nodesTopEdge"); }
    amd::array<size_t> nodesLeftEdge() const { return nodeset("nodesLeftEdge"); }
    amd::array<size_t> nodesRightEdge() const { return nodeset("nodesRightEdge"); }
    amd::array<size_t> nodesBottomOpenEdge() const { return nodeset("nodesBottomOpenEdge"); }
    amd::array<size_t> nodesTopOpenEdge() const { return nodeset("nodesTopOpenEdge"); }
    amd::array<size_t> nodesLeftOpenEdge() const { return nodeset("nodesLeftOpenEdge"); }
    amd::array<size_t> nodesRightOpenEdge() const { return nodeset("nodesRightOpenEdge"); }
    size_t nodesBottomLeftCorner() const { return node("nodesBottomLeftCorner"); }
    size_t nodesBottomRightCorner() const { return node("nodesBottomRightCorner"); }
    size_t nodesTopLeftCorner() const { return node("nodesTopLeftCorner"); }
    size_t nodesTopRightCorner() const { return node("nodesTopRightCorner"); }
};

/* Mesh::Quad4::FineLayer  (include/GooseFEM/MeshQuad4.h:200-1135): layers finest around the
 * middle, coarsened by 3 in x towards the bottom and top; node sets by name via nodeset(name) */
class FineLayer {
public:
    FineLayer(size_t nelx, size_t nely, double h = 1.0, size_t nfine = 1)
        : m_nx(static_cast<int64_t>(nelx)), m_ny(static_cast<int64_t>(nely)), m_nf(static_cast<int64_t>(nfine)), m_h(h)
    {
        int64_t sz[4] = {0, 0, 0, 0};
        amd::check(gf_mesh_quad4_finelayer(m_nx, m_ny, m_h, m_nf, sz, nullptr, nullptr));
        m_nelem = static_cast<size_t>(sz[0]);
        m_nnode = static_cast<size_t>(sz[1]);
        m_nelx = static_cast<size_t>(sz[2]);
        m_nely = static_cast<size_t>(sz[3]);
    }
    size_t nelem() const { return m_nelem; }
    size_t nnode() const { return m_nnode; }
    size_t nne() const { return 4; }
    size_t ndim() const { return 2; }
    size_t nelx() const { return m_nelx; }
    size_t nely() const { return m_nely; }
    double h() const { return m_h; }
    amd::array<double> coor() const
    {
        amd::array<double> c({m_nnode, 2});
        amd::check(gf_mesh_quad4_finelayer(m_nx, m_ny, m_h, m_nf, nullptr, c.data(), nullptr));
        return c;
    }
    amd::array<size_t> conn() const
    {
        std::vector<int64_t> c(m_nelem * 4);
        amd::check(gf_mesh_quad4_finelayer(m_nx, m_ny, m_h, m_nf, nullptr, nullptr, c.data()));
        return amd::from_i64(c, {m_nelem, 4});
    }
    amd::array<size_t> dofs() const { return Mesh::dofs(m_nnode, 2); }
    amd::array<size_t> nodeset(const std::string& name) const
    {
        int64_t n = 0;
        amd::check(gf_mesh_quad4_finelayer_set(m_nx, m_ny, m_h, m_nf, name.c_str(), nullptr, &n));
        std::vector<int64_t> r(static_cast<size_t>(n));
        amd::check(gf_mesh_quad4_finelayer_set(m_nx, m_ny, m_h, m_nf, name.c_str(), r.data(), &n));
        return amd::from_i64(r, {static_cast<size_t>(n)});
    }
    size_t node(const std::string& name) const { return nodeset(name)[0]; }
    amd::array<size_t> elementsMiddleLayer() const { return nodeset("elementsMiddleLayer"); }
    amd::array<size_t> roll(size_t n) const
    {
        std::vector<int64_t> r(m_nelem);
        amd::check(gf_mesh_quad4_finelayer_roll(m_nx, m_ny, m_h, m_nf, static_cast<int64_t>(n), r.data()));
        return amd::from_i64(r, {m_nelem});
    }
    amd::array<size_t> nodesBottomEdge() const { return nodeset("nodesBottomEdge"); }
    amd::array<size_t> nodesTopEdge() const { return nodeset("nodesTopEdge"); }
    amd::array<size_t> nodesLeftEdge() const { return nodeset("nodesLeftEdge"); }
    amd::array<size_t> nodesRightEdge() const { return nodeset("nodesRightEdge"); }
    size_t nodesOrigin() const { return node("nodesOrigin"); }
    amd::array<size_t> nodesPeriodic() const
    {
        auto r = nodeset("nodesPeriodic");
        return amd::from_i64(amd::to_i64(r), {r.size() / 2, 2});
    }
    amd::array<size_t> dofsPeriodic() const
    {
        auto t = amd::to_i64(nodesPeriodic());
        std::vector<int64_t> r(m_nnode * 2);
        amd::check(gf_mesh_dofs_periodic(static_cast<int64_t>(m_nnode), 2, t.data(), static_cast<int64_t>(t.size() / 2),
                                         r.data()));
        return amd::from_i64(r, {m_nnode, 2});
    }

private:
    int64_t m_nx, m_ny, m_nf;
    double m_h;
    size_t m_nelem = 0, m_nnode = 0, m_nelx = 0, m_nely = 0;
};

} // namespace Quad4

namespace Hex8 {

/* Mesh::Hex8::Regular  (include/GooseFEM/MeshHex8.h:26-609); node sets via nodeset(name) */
class Regular : public detail::RegularBase {
public:
    Regular(size_t nelx, size_t nely, size_t nelz, double h = 1.0)
    {
        m_nx = static_cast<int64_t>(nelx);
        m_ny = static_cast<int64_t>(nely);
        m_nz = static_cast<int64_t>(nelz);
        m_h = h;
        m_nne = 8;
        m_ndim = 3;
        m_nelem = nelx * nely * nelz;
        m_nnode = (nelx + 1) * (nely + 1) * (nelz + 1);
    }
    size_t nelz() const { return static_cast<size_t>(m_nz); }
    amd::array<double> coor() const
    {
        amd::array<double> c({m_nnode, 3});
        amd::check(gf_mesh_hex8_regular(m_nx, m_ny, m_nz, m_h, c.data(), nullptr));
        return c;
    }
    amd::array<size_t> conn() const
    {
        std::vector<int64_t> c(m_nelem * 8);
        amd::check(gf_mesh_hex8_regular(m_nx, m_ny, m_nz, m_h, nullptr, c.data()));
        return amd::from_i64(c, {m_nelem, 8});
    }
};

/* Mesh::Hex8::FineLayer  (include/GooseFEM/MeshHex8.h:613-2296): layers finest around the
 * middle, coarsened by 3 in x or z towards the bottom and top; node sets by name via
 * nodeset(name) ("nodesFront", "nodesFrontFace", "nodesFrontBottomEdge", ...) */
class FineLayer {
public:
    FineLayer(size_t nelx, size_t nely, size_t nelz, double h = 1.0, size_t nfine = 1)
        : m_nx(static_cast<int64_t>(nelx)), m_ny(static_cast<int64_t>(nely)), m_nz(static_cast<int64_t>(nelz)),
          m_nf(static_cast<int64_t>(nfine)), m_h(h)
    {
        int64_t sz[5] = {0, 0, 0, 0, 0};
        amd::check(gf_mesh_hex8_finelayer(m_nx, m_ny, m_nz, m_h, m_nf, sz, nullptr, nullptr));
        m_nelem = static_cast<size_t>(sz[0]);
        m_nnode = static_cast<size_t>(sz[1]);
        m_nelx = static_cast<size_t>(sz[2]);
        m_nely = static_cast<size_t>(sz[3]);
        m_nelz = static_cast<size_t>(sz[4]);
    }
    size_t nelem() const { return m_nelem; }
    size_t nnode() const { return m_nnode; }
    size_t nne() const { return 8; }
    size_t ndim() const { return 3; }
    size_t nelx() const { return m_nelx; }
    size_t nely() const { return m_nely; }
    size_t nelz() const { return m_nelz; }
    double h() const { return m_h; }
    amd::array<double> coor() const
    {
        amd::array<double> c({m_nnode, 3});
        amd::check(gf_mesh_hex8_finelayer(m_nx, m_ny, m_nz, m_h, m_nf, nullptr, c.data(), nullptr));
        return c;
    }
    amd::array<size_t> conn() const
    {
        std::vector<int64_t> c(m_nelem * 8);
        amd::check(gf_mesh_hex8_finelayer(m_nx, m_ny, m_nz, m_h, m_nf, nullptr, nullptr, c.data()));
        return amd::from_i64(c, {m_nelem, 8});
    }
    amd::array<size_t> dofs() const { return Mesh::dofs(m_nnode, 3); }
    amd::array<size_t> nodeset(const std::string& name) const
    {
        int64_t n = 0;
        amd::check(gf_mesh_hex8_finelayer_set(m_nx, m_ny, m_nz, m_h, m_nf, name.c_str(), nullptr, &n));
        std::vector<int64_t> r(static_cast<size_t>(n));
        amd::check(gf_mesh_hex8_finelayer_set(m_nx, m_ny, m_nz, m_h, m_nf, name.c_str(), r.data(), &n));
        return amd::from_i64(r, {static_cast<size_t>(n)});
    }
    size_t node(const std::string& name) const { return nodeset(name)[0]; }
    amd::array<size_t> elementsMiddleLayer() const { return nodeset("elementsMiddleLayer"); }
    amd::array<size_t> nodesFront() const { return nodeset("nodesFront"); }
    amd::array<size_t> nodesBack() const { return nodeset("nodesBack"); }
    amd::array<size_t> nodesLeft() const { return nodeset("nodesLeft"); }
    amd::array<size_t> nodesRight() const { return nodeset("nodesRight"); }
    amd::array<size_t> nodesBottom() const { return nodeset("nodesBottom"); }
    amd::array<size_t> nodesTop() const { return nodeset("nodesTop"); }
    size_t nodesOrigin() const { return node("nodesOrigin"); }
    amd::array<size_t> nodesPeriodic() const
    {
        auto r = nodeset("nodesPeriodic");
        return amd::from_i64(amd::to_i64(r), {r.size() / 2, 2});
    }
    amd::array<size_t> dofsPeriodic() const
    {
        auto t = amd::to_i64(nodesPeriodic());
        std::vector<int64_t> r(m_nnode * 3);
        amd::check(gf_mesh_dofs_periodic(static_cast<int64_t>(m_nnode), 3, t.data(), static_cast<int64_t>(t.size() / 2),
                                         r.data()));
        return amd::from_i64(r, {m_nnode, 3});
    }

private:
    int64_t m_nx, m_ny, m_nz, m_nf;
    double m_h;
    size_t m_nelem = 0, m_nnode = 0, m_nelx = 0, m_nely = 0, m_nelz = 0;
};

} // namespace Hex8
} // namespace Mesh

// =================================================================================================
// Vector (include/GooseFEM/Vector.h)
// =================================================================================================

class Vector {
public:
    template <class S, class T>
    Vector(const S& conn, const T& dofs)
    {
        auto c = amd::to_i64(conn);
        auto d = amd::to_i64(dofs);
        m_nelem = conn.shape()[0];
        m_nne = conn.shape()[1];
        m_nnode = dofs.shape()[0];
        m_ndim = dofs.shape()[1];
        amd::check(gf_vector_create(&m_h, c.data(), static_cast<int64_t>(m_nelem), static_cast<int>(m_nne), d.data(),
                                    static_cast<int64_t>(m_nnode), static_cast<int>(m_ndim)));
        int64_t nd = 0;
        amd::check(gf_vector_sizes(m_h, nullptr, nullptr, nullptr, nullptr, &nd));
        m_ndof = static_cast<size_t>(nd);
    }
    Vector(const Vector&) = delete;
    Vector& operator=(const Vector&) = delete;
    ~Vector() { gf_vector_destroy(m_h); }

    size_t nelem() const { return m_nelem; }
    size_t nne() const { return m_nne; }
    size_t nnode() const { return m_nnode; }
    size_t ndim() const { return m_ndim; }
    size_t ndof() const { return m_ndof; }
    std::vector<size_t> shape_elemvec() const { return {m_nelem, m_nne, m_ndim}; }
    std::vector<size_t> shape_nodevec() const { return {m_nnode, m_ndim}; }
    std::vector<size_t> shape_dofval() const { return {m_ndof}; }
    amd::array<double> allocate_elemvec(double v = 0.0) const { return amd::array<double>(shape_elemvec(), v); }
    amd::array<double> allocate_nodevec(double v = 0.0) const { return amd::array<double>(shape_nodevec(), v); }
    amd::array<double> allocate_dofval(double v = 0.0) const { return amd::array<double>(shape_dofval(), v); }
    gf_vector handle() const { return m_h; }

    /* AsElement(dofval | nodevec)  (Vector.h:193-211) */
    template <class T>
    amd::array<double> AsElement(const T& arg) const
    {
        amd::array<double> r(shape_elemvec());
        asElement(arg, r);
        return r;
    }
    template <class T, class R>
    void asElement(const T& arg, R& ret) const
    {
        amd::dev_in<T> in(arg);
        amd::dev_out<R> out(ret, shape_elemvec());
        if (arg.dimension() == 1) amd::check(gf_vector_as_element_dofval(m_h, in.ptr, out.ptr, nullptr));
        else amd::check(gf_vector_as_element_nodevec(m_h, in.ptr, out.ptr, nullptr));
        out.finish();
    }
    /* AsDofs(nodevec | elemvec)  (Vector.h:141-159) */
    template <class T>
    amd::array<double> AsDofs(const T& arg) const
    {
        amd::array<double> r(shape_dofval());
        asDofs(arg, r);
        return r;
    }
    template <class T, class R>
    void asDofs(const T& arg, R& ret) const
    {
        amd::dev_in<T> in(arg);
        amd::dev_out<R> out(ret, shape_dofval());
        if (arg.dimension() == 2) amd::check(gf_vector_as_dofs_nodevec(m_h, in.ptr, out.ptr, nullptr));
        else amd::check(gf_vector_as_dofs_elemvec(m_h, in.ptr, out.ptr, nullptr));
        out.finish();
    }
    /* AsNode(dofval | elemvec)  (Vector.h:167-185) */
    template <class T>
    amd::array<double> AsNode(const T& arg) const
    {
        amd::array<double> r(shape_nodevec());
        asNode(arg, r);
        return r;
    }
    template <class T, class R>
    void asNode(const T& arg, R& ret) const
    {
        amd::dev_in<T> in(arg);
        amd::dev_out<R> out(ret, shape_nodevec());
        if (arg.dimension() == 1) amd::check(gf_vector_as_node_dofval(m_h, in.ptr, out.ptr, nullptr));
        else amd::check(gf_vector_as_node_elemvec(m_h, in.ptr, out.ptr, nullptr));
        out.finish();
    }
    /* AssembleDofs(nodevec | elemvec)  (Vector.h:219-237) */
    template <class T>
    amd::array<double> AssembleDofs(const T& arg) const
    {
        amd::array<double> r(shape_dofval());
        assembleDofs(arg, r);
        return r;
    }
    template <class T, class R>
    void assembleDofs(const T& arg, R& ret) const
    {
        amd::dev_in<T> in(arg);
        amd::dev_out<R> out(ret, shape_dofval());
        if (arg.dimension() == 2) amd::check(gf_vector_assemble_dofs_nodevec(m_h, in.ptr, out.ptr, nullptr));
        else amd::check(gf_vector_assemble_dofs_elemvec(m_h, in.ptr, GF_F64, out.ptr, nullptr));
        out.finish();
    }
    /* AssembleNode(elemvec)  (Vector.h:245-263) */
    template <class T>
    amd::array<double> AssembleNode(const T& arg) const
    {
        amd::array<double> r(shape_nodevec());
        assembleNode(arg, r);
        return r;
    }
    template <class T, class R>
    void assembleNode(const T& arg, R& ret) const
    {
        amd::dev_in<T> in(arg);
        amd::dev_out<R> out(ret, shape_nodevec());
        amd::check(gf_vector_assemble_node_elemvec(m_h, in.ptr, GF_F64, out.ptr, nullptr));
        out.finish();
    }

private:
    gf_vector m_h = nullptr;
    size_t m_nelem = 0, m_nne = 0, m_nnode = 0, m_ndim = 0, m_ndof = 0;
};

// =================================================================================================
// Element quadrature (include/GooseFEM/Element.h, ElementQuad4.h, ElementQuad4Planar.h, ElementHex8.h)
// =================================================================================================
namespace Element {
namespace detail {

inline amd::array<double> rule(size_t ndim, bool nodal, bool weights)
{
    const double a = nodal ? 1.0 : 1.0 / std::sqrt(3.0);
    const size_t nip = ndim == 2 ? 4 : 8;
    static const int s2[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
    static const int s3[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                 {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    if (weights) return amd::array<double>({nip}, 1.0);
    amd::array<double> xi({nip, ndim});
    for (size_t q = 0; q < nip; ++q)
        for (size_t d = 0; d < ndim; ++d) xi(q, d) = (ndim == 2 ? s2[q][d] : s3[q][d]) < 0 ? -a : a;
    return xi;
}

class QuadratureBase {
public:
    QuadratureBase(const QuadratureBase&) = delete;
    QuadratureBase& operator=(const QuadratureBase&) = delete;
    ~QuadratureBase() { gf_quad_destroy(m_h); }
    size_t nelem() const { return m_nelem; }
    size_t nne() const { return m_nne; }
    size_t ndim() const { return m_ndim; }
    size_t tdim() const { return m_tdim; }
    size_t nip() const { return m_nip; }
    gf_quad handle() const { return m_h; }
    std::vector<size_t> shape_elemvec() const { return {m_nelem, m_nne, m_ndim}; }
    std::vector<size_t> shape_elemmat() const { return {m_nelem, m_nne * m_ndim, m_nne * m_ndim}; }
    std::vector<size_t> shape_qscalar() const { return {m_nelem, m_nip}; }
    std::vector<size_t> shape_qtensor2() const { return {m_nelem, m_nip, m_tdim, m_tdim}; }
    std::vector<size_t> shape_qtensor4() const { return {m_nelem, m_nip, m_tdim, m_tdim, m_tdim, m_tdim}; }

    /* update_x(x)  (Element.h:563-569) */
    template <class T>
    void update_x(const T& x)
    {
        amd::dev_in<T> in(x);
        amd::check(gf_quad_update_x(m_h, in.ptr, nullptr));
        m_synced = false;
    }
    /* GradN() / dV()  (Element.h:575-587): const references, refreshed from the device on demand */
    const amd::array<double>& GradN() const
    {
        sync();
        return m_dNx;
    }
    const amd::array<double>& dV() const
    {
        sync();
        return m_vol;
    }
    /* InterpQuad_vector  (Element.h:597-616) */
    template <class T>
    amd::array<double> InterpQuad_vector(const T& elemvec) const
    {
        const size_t n = elemvec.shape()[2];
        amd::array<double> r({m_nelem, m_nip, n});
        amd::dev_in<T> in(elemvec);
        amd::dev_out<amd::array<double>> out(r, r.shape());
        amd::check(gf_quad_interp_quad_vector(m_h, in.ptr, static_cast<int>(n), out.ptr, nullptr));
        out.finish();
        return r;
    }
#define GOOSEFEM_AMD_QT2(Name, name, fn)                                                                    \
    template <class T>                                                                                     \
    amd::array<double> Name(const T& elemvec) const                                                        \
    {                                                                                                      \
        amd::array<double> r(shape_qtensor2());                                                            \
        name(elemvec, r);                                                                                  \
        return r;                                                                                          \
    }                                                                                                      \
    template <class T, class R>                                                                            \
    void name(const T& elemvec, R& qtensor) const                                                          \
    {                                                                                                      \
        amd::dev_in<T> in(elemvec);                                                                        \
        amd::dev_out<R> out(qtensor, shape_qtensor2());                                                    \
        amd::check(fn(m_h, in.ptr, out.ptr, GF_F64, nullptr));                                             \
        out.finish();                                                                                      \
    }
    /* GradN_vector / GradN_vector_T / SymGradN_vector  (Element.h:634-717) */
    GOOSEFEM_AMD_QT2(GradN_vector, gradN_vector, gf_quad_gradN_vector)
    GOOSEFEM_AMD_QT2(GradN_vector_T, gradN_vector_T, gf_quad_gradN_vector_T)
    GOOSEFEM_AMD_QT2(SymGradN_vector, symGradN_vector, gf_quad_symgradN_vector)
#undef GOOSEFEM_AMD_QT2
    /* Int_N_vector_dV  (Element.h:731-750) */
    template <class T>
    amd::array<double> Int_N_vector_dV(const T& qvector) const
    {
        const size_t n = qvector.shape()[2];
        amd::array<double> r({m_nelem, m_nne, n});
        amd::dev_in<T> in(qvector);
        amd::dev_out<amd::array<double>> out(r, r.shape());
        amd::check(gf_quad_int_N_vector_dV(m_h, in.ptr, static_cast<int>(n), out.ptr, nullptr));
        out.finish();
        return r;
    }
    /* Int_N_scalar_NT_dV  (Element.h:770-788) */
    template <class T>
    amd::array<double> Int_N_scalar_NT_dV(const T& qscalar) const
    {
        amd::array<double> r(shape_elemmat());
        int_N_scalar_NT_dV(qscalar, r);
        return r;
    }
    template <class T, class R>
    void int_N_scalar_NT_dV(const T& qscalar, R& elemmat) const
    {
        amd::dev_in<T> in(qscalar);
        amd::dev_out<R> out(elemmat, shape_elemmat());
        amd::check(gf_quad_int_N_scalar_NT_dV(m_h, in.ptr, out.ptr, nullptr));
        out.finish();
    }
    /* Int_gradN_dot_tensor2_dV  (Element.h:807-825) */
    template <class T>
    amd::array<double> Int_gradN_dot_tensor2_dV(const T& qtensor) const
    {
        amd::array<double> r(shape_elemvec());
        int_gradN_dot_tensor2_dV(qtensor, r);
        return r;
    }
    template <class T, class R>
    void int_gradN_dot_tensor2_dV(const T& qtensor, R& elemvec) const
    {
        amd::dev_in<T> in(qtensor);
        amd::dev_out<R> out(elemvec, shape_elemvec());
        amd::check(gf_quad_int_gradN_dot_tensor2_dV(m_h, in.ptr, out.ptr, GF_F64, nullptr));
        out.finish();
    }
    /* Int_gradN_dot_tensor4_dot_gradNT_dV  (Element.h:849-868); precision GF_F64 or GF_BF16X3 */
    template <class T>
    amd::array<double> Int_gradN_dot_tensor4_dot_gradNT_dV(const T& qtensor, int precision = GF_F64) const
    {
        amd::array<double> r(shape_elemmat());
        int_gradN_dot_tensor4_dot_gradNT_dV(qtensor, r, precision);
        return r;
    }
    template <class T, class R>
    void int_gradN_dot_tensor4_dot_gradNT_dV(const T& qtensor, R& elemmat, int precision = GF_F64) const
    {
        amd::dev_in<T> in(qtensor);
        amd::dev_out<R> out(elemmat, shape_elemmat());
        amd::check(gf_quad_int_gradN_dot_tensor4_dot_gradNT_dV(m_h, in.ptr, out.ptr, precision, nullptr));
        out.finish();
    }

protected:
    QuadratureBase() = default;
    template <class T, class X, class W>
    void init(int etype, const T& x, const X& xi, const W& w, double thick)
    {
        amd::dev_in<T> in(x);
        m_nelem = x.shape()[0];
        amd::check(gf_quad_create(&m_h, etype, in.ptr, static_cast<int64_t>(m_nelem), GF_GAUSS, xi.data(), w.data(),
                                  static_cast<int>(w.size()), thick, nullptr));
        int64_t ne = 0;
        int nne = 0, nd = 0, td = 0, ni = 0;
        amd::check(gf_quad_sizes(m_h, &ne, &nne, &nd, &td, &ni));
        m_nne = nne;
        m_ndim = nd;
        m_tdim = td;
        m_nip = ni;
        amd::check(gf_stream_sync(nullptr));
    }
    void sync() const
    {
        if (m_synced) return;
        const double *g = nullptr, *v = nullptr;
        amd::check(gf_quad_gradN(m_h, &g));
        amd::check(gf_quad_dV(m_h, &v));
        m_dNx = amd::array<double>({m_nelem, m_nip, m_nne, m_ndim});
        m_vol = amd::array<double>({m_nelem, m_nip});
        amd::check(gf_memcpy_d2h(m_dNx.data(), g, static_cast<int64_t>(m_dNx.size() * sizeof(double))));
        amd::check(gf_memcpy_d2h(m_vol.data(), v, static_cast<int64_t>(m_vol.size() * sizeof(double))));
        m_synced = true;
    }
    gf_quad m_h = nullptr;
    size_t m_nelem = 0, m_nne = 0, m_ndim = 0, m_tdim = 0, m_nip = 0;
    mutable amd::array<double> m_dNx, m_vol;
    mutable bool m_synced = false;
};

} // namespace detail

namespace Quad4 {
namespace Gauss {
inline size_t nip() { return 4; }
inline amd::array<double> xi() { return detail::rule(2, false, false); }
inline amd::array<double> w() { return detail::rule(2, false, true); }
} // namespace Gauss
namespace Nodal {
inline size_t nip() { return 4; }
inline amd::array<double> xi() { return detail::rule(2, true, false); }
inline amd::array<double> w() { return detail::rule(2, true, true); }
} // namespace Nodal

/* Element::Quad4::Quadrature  (ElementQuad4.h:229-547) */
class Quadrature : public detail::QuadratureBase {
public:
    template <class T>
    explicit Quadrature(const T& x) { init(GF_QUAD4, x, Gauss::xi(), Gauss::w(), 1.0); }
    template <class T, class X, class W>
    Quadrature(const T& x, const X& xi, const W& w) { init(GF_QUAD4, x, xi, w, 1.0); }
};

/* Element::Quad4::QuadraturePlanar  (ElementQuad4Planar.h:37-335) */
class QuadraturePlanar : public detail::QuadratureBase {
public:
    template <class T>
    explicit QuadraturePlanar(const T& x, double thick = 1.0) { init(GF_QUAD4_PLANAR, x, Gauss::xi(), Gauss::w(), thick); }
    template <class T, class X, class W>
    QuadraturePlanar(const T& x, const X& xi, const W& w, double thick = 1.0)
    {
        init(GF_QUAD4_PLANAR, x, xi, w, thick);
    }
};
} // namespace Quad4

namespace Hex8 {
namespace Gauss {
inline size_t nip() { return 8; }
inline amd::array<double> xi() { return detail::rule(3, false, false); }
inline amd::array<double> w() { return detail::rule(3, false, true); }
} // namespace Gauss
namespace Nodal {
inline size_t nip() { return 8; }
inline amd::array<double> xi() { return detail::rule(3, true, false); }
inline amd::array<double> w() { return detail::rule(3, true, true); }
} // namespace Nodal

/* Element::Hex8::Quadrature  (ElementHex8.h:206-400) */
class Quadrature : public detail::QuadratureBase {
public:
    template <class T>
    explicit Quadrature(const T& x) { init(GF_HEX8, x, Gauss::xi(), Gauss::w(), 1.0); }
    template <class T, class X, class W>
    Quadrature(const T& x, const X& xi, const W& w) { init(GF_HEX8, x, xi, w, 1.0); }
};
} // namespace Hex8
} // namespace Element

namespace Element {
namespace Quad4 {

/* Element::Quad4::QuadratureAxisymmetric  (include/GooseFEM/ElementQuad4Axisymmetric.h:33-467):
 * nodal coordinates (z, r), tensors 3 x 3 in (r, theta, z) order */
class QuadratureAxisymmetric {
public:
    template <class T>
    explicit QuadratureAxisymmetric(const T& x) : QuadratureAxisymmetric(x, Gauss::xi(), Gauss::w())
    {
    }
    template <class T, class X, class W>
    QuadratureAxisymmetric(const T& x, const X& xi, const W& w)
    {
        m_nelem = x.shape()[0];
        m_nip = w.size();
        std::vector<double> xih(xi.data(), xi.data() + xi.size()), wh(w.data(), w.data() + w.size());
        amd::dev_in<T> in(x);
        amd::check(gf_axi_create(&m_h, static_cast<int64_t>(m_nelem), static_cast<int>(m_nip), in.ptr, xih.data(),
                                 wh.data(), nullptr));
        amd::check(gf_stream_sync(nullptr));
    }
    QuadratureAxisymmetric(const QuadratureAxisymmetric&) = delete;
    QuadratureAxisymmetric& operator=(const QuadratureAxisymmetric&) = delete;
    ~QuadratureAxisymmetric() { gf_axi_destroy(m_h); }
    size_t nelem() const { return m_nelem; }
    size_t nne() const { return 4; }
    size_t ndim() const { return 2; }
    size_t tdim() const { return 3; }
    size_t nip() const { return m_nip; }
    std::vector<size_t> shape_qtensor2() const { return {m_nelem, m_nip, 3, 3}; }
    std::vector<size_t> shape_qtensor4() const { return {m_nelem, m_nip, 3, 3, 3, 3}; }
    template <class T>
    void update_x(const T& x)
    {
        amd::dev_in<T> in(x);
        amd::check(gf_axi_update_x(m_h, in.ptr, nullptr));
        amd::check(gf_stream_sync(nullptr));
    }
    amd::array<double> dV() const
    {
        const double* p = nullptr;
        amd::check(gf_axi_geometry(m_h, nullptr, &p));
        amd::array<double> r({m_nelem, m_nip});
        amd::check(gf_memcpy_d2h(r.data(), p, static_cast<int64_t>(r.size() * sizeof(double))));
        return r;
    }
    template <class T>
    amd::array<double> GradN_vector(const T& ue) const { return grad(0, ue); }
    template <class T>
    amd::array<double> GradN_vector_T(const T& ue) const { return grad(1, ue); }
    template <class T>
    amd::array<double> SymGradN_vector(const T& ue) const { return grad(2, ue); }
    template <class T, class R>
    void symGradN_vector(const T& ue, R& eps) const { grad_into(2, ue, eps); }
    template <class T>
    amd::array<double> Int_N_scalar_NT_dV(const T& qs) const { return integrate(0, qs, {m_nelem, 8, 8}); }
    template <class T>
    amd::array<double> Int_gradN_dot_tensor2_dV(const T& qt) const { return integrate(1, qt, {m_nelem, 4, 2}); }
    template <class T>
    amd::array<double> Int_gradN_dot_tensor4_dot_gradNT_dV(const T& qt) const { return integrate(2, qt, {m_nelem, 8, 8}); }

private:
    template <class T>
    amd::array<double> grad(int mode, const T& ue) const
    {
        amd::array<double> r(shape_qtensor2());
        grad_into(mode, ue, r);
        return r;
    }
    template <class T, class R>
    void grad_into(int mode, const T& ue, R& ret) const
    {
        amd::dev_in<T> in(ue);
        amd::dev_out<R> out(ret, shape_qtensor2());
        amd::check(gf_axi_grad(m_h, mode, in.ptr, out.ptr, nullptr));
        out.finish();
    }
    template <class T>
    amd::array<double> integrate(int which, const T& q, std::vector<size_t> shape) const
    {
        amd::array<double> r(std::move(shape));
        amd::dev_in<T> in(q);
        amd::dev_out<amd::array<double>> out(r, r.shape());
        if (which == 0) amd::check(gf_axi_int_N_scalar_NT_dV(m_h, in.ptr, out.ptr, nullptr));
        else if (which == 1) amd::check(gf_axi_int_gradN_dot_tensor2_dV(m_h, in.ptr, out.ptr, nullptr));
        else amd::check(gf_axi_int_gradN_dot_tensor4_dot_gradNT_dV(m_h, in.ptr, out.ptr, nullptr));
        out.finish();
        return r;
    }
    gf_axi m_h = nullptr;
    size_t m_nelem = 0, m_nip = 0;
};

} // namespace Quad4
} // namespace Element

// =================================================================================================
// Material laws the examples take from GMat (not part of GooseFEM), with GMat's interface:
// Eps(), Sig(), C() as device arrays, refresh(), increment()
// =================================================================================================
namespace material {

/* J2 elasto-plasticity with linear isotropic hardening, as
 * GMatElastoPlastic::Cartesian3d::LinearHardening (docs/examples/statics/periodic_elastoplastic.py:71-104) */
class LinearHardening {
public:
    template <class K, class G, class S, class H>
    LinearHardening(const K& kappa, const G& mu, const S& sigy0, const H& hard)
    {
        m_shape = amd::shape_of(kappa);
        m_nq = kappa.size();
        amd::dev_in<K> k(kappa);
        amd::dev_in<G> g(mu);
        amd::dev_in<S> sy(sigy0);
        amd::dev_in<H> hh(hard);
        amd::check(gf_plastic_create(&m_h, static_cast<int64_t>(m_nq), k.ptr, g.ptr, sy.ptr, hh.ptr, nullptr));
        auto s2 = m_shape, s4 = m_shape;
        s2.insert(s2.end(), {3, 3});
        s4.insert(s4.end(), {3, 3, 3, 3});
        m_Eps = amd::device_array<double>(s2);
        m_Sig = amd::device_array<double>(s2);
        m_C = amd::device_array<double>(s4);
        amd::array<double> z(s2);
        amd::check(gf_memcpy_h2d(m_Eps.data(), z.data(), static_cast<int64_t>(m_Eps.bytes())));
    }
    LinearHardening(const LinearHardening&) = delete;
    LinearHardening& operator=(const LinearHardening&) = delete;
    ~LinearHardening() { gf_plastic_destroy(m_h); }
    amd::device_array<double>& Eps() { return m_Eps; }
    const amd::device_array<double>& Sig() const { return m_Sig; }
    const amd::device_array<double>& C() const { return m_C; }
    void refresh(bool compute_tangent = true)
    {
        amd::check(gf_plastic_refresh(m_h, m_Eps.data(), m_Sig.data(), compute_tangent ? m_C.data() : nullptr, nullptr));
        amd::check(gf_stream_sync(nullptr));
    }
    void increment()
    {
        amd::check(gf_plastic_increment(m_h, nullptr));
        amd::check(gf_stream_sync(nullptr));
    }
    /* equivalent plastic strain (host copy) */
    amd::array<double> epsp() const
    {
        const double* p = nullptr;
        amd::check(gf_plastic_state(m_h, nullptr, &p, nullptr, nullptr));
        amd::array<double> r(m_shape);
        amd::check(gf_memcpy_d2h(r.data(), p, static_cast<int64_t>(r.size() * sizeof(double))));
        return r;
    }

private:
    gf_plastic m_h = nullptr;
    std::vector<size_t> m_shape;
    size_t m_nq = 0;
    amd::device_array<double> m_Eps, m_Sig, m_C;
};

} // namespace material

// =================================================================================================
// MatrixDiagonal (include/GooseFEM/MatrixDiagonal.h) and Matrix (include/GooseFEM/Matrix.h)
// =================================================================================================

class MatrixDiagonal {
public:
    template <class C, class D>
    MatrixDiagonal(const C& conn, const D& dofs) : m_vec(conn, dofs)
    {
        amd::check(gf_diag_create(&m_h, m_vec.handle()));
    }
    ~MatrixDiagonal() { gf_diag_destroy(m_h); }
    size_t ndof() const { return m_vec.ndof(); }
    /* assemble(elemmat): diagonal entries only (MatrixDiagonal.h:150-167) */
    template <class T>
    void assemble(const T& elemmat)
    {
        amd::dev_in<T> in(elemmat);
        amd::check(gf_diag_assemble(m_h, in.ptr, nullptr));
        amd::check(gf_stream_sync(nullptr));
    }
    template <class T>
    void set(const T& A)
    {
        amd::dev_in<T> in(A);
        amd::check(gf_diag_set(m_h, in.ptr, nullptr));
        amd::check(gf_stream_sync(nullptr));
    }
    /* data()  (MatrixDiagonal.h:206-209) */
    amd::array<double> data() const
    {
        const double* p = nullptr;
        amd::check(gf_diag_data(m_h, &p));
        amd::array<double> r({ndof()});
        amd::check(gf_memcpy_d2h(r.data(), p, static_cast<int64_t>(r.size() * sizeof(double))));
        return r;
    }
    /* Dot(x) / Solve(b) on dofval or nodevec  (MatrixDiagonal.h:212-259) */
    template <class T>
    amd::array<double> Dot(const T& x) const
    {
        amd::array<double> r(amd::shape_of(x));
        amd::dev_in<T> in(x);
        amd::dev_out<amd::array<double>> out(r, r.shape());
        if (x.dimension() == 1) amd::check(gf_diag_dot_dofval(m_h, in.ptr, out.ptr, nullptr));
        else amd::check(gf_diag_dot_nodevec(m_h, in.ptr, out.ptr, nullptr));
        out.finish();
        return r;
    }
    template <class T>
    amd::array<double> Solve(const T& b)
    {
        amd::array<double> r(amd::shape_of(b));
        solve(b, r);
        return r;
    }
    template <class T, class R>
    void solve(const T& b, R& x)
    {
        amd::dev_in<T> in(b);
        amd::dev_out<R> out(x, amd::shape_of(b));
        if (b.dimension() == 1) amd::check(gf_diag_solve_dofval(m_h, in.ptr, out.ptr, nullptr));
        else amd::check(gf_diag_solve_nodevec(m_h, in.ptr, out.ptr, nullptr));
        out.finish();
    }

private:
    Vector m_vec;
    gf_diag m_h = nullptr;
};

class Matrix {
public:
    template <class C, class D>
    Matrix(const C& conn, const D& dofs) : m_vec(conn, dofs)
    {
        amd::check(gf_matrix_create(&m_h, m_vec.handle()));
        int64_t nr = 0, nz = 0;
        amd::check(gf_matrix_sizes(m_h, &nr, &nz));
        m_nrows = static_cast<size_t>(nr);
        m_nnz = static_cast<size_t>(nz);
    }
    ~Matrix() { gf_matrix_destroy(m_h); }
    size_t ndof() const { return m_nrows; }
    size_t nnz() const { return m_nnz; }
    /* assemble(elemmat)  (Matrix.h:704-727) */
    template <class T>
    void assemble(const T& elemmat)
    {
        amd::dev_in<T> in(elemmat);
        amd::check(gf_matrix_assemble(m_h, in.ptr, nullptr));
        amd::check(gf_stream_sync(nullptr));
    }
    /* Dot(x)  (Matrix.h:804-815) */
    template <class T>
    amd::array<double> Dot(const T& x) const
    {
        amd::array<double> r(amd::shape_of(x));
        amd::dev_in<T> in(x);
        amd::dev_out<amd::array<double>> out(r, r.shape());
        if (x.dimension() == 1) amd::check(gf_matrix_dot_dofval(m_h, in.ptr, out.ptr, nullptr));
        else amd::check(gf_matrix_dot_nodevec(m_h, in.ptr, out.ptr, nullptr));
        out.finish();
        return r;
    }
    /* Todense()  (Matrix.h:338-355) */
    amd::array<double> Todense() const
    {
        amd::array<double> r({m_nrows, m_nrows});
        amd::dev_out<amd::array<double>> out(r, r.shape());
        amd::check(gf_matrix_todense(m_h, out.ptr, nullptr));
        out.finish();
        return r;
    }

private:
    Vector m_vec;
    gf_matrix m_h = nullptr;
    size_t m_nrows = 0, m_nnz = 0;
};

// =================================================================================================
// Dirichlet partitioning: VectorPartitioned (include/GooseFEM/VectorPartitioned.h),
// MatrixDiagonalPartitioned (MatrixDiagonalPartitioned.h), MatrixPartitioned and
// MatrixPartitionedSolver (MatrixPartitioned.h)
// =================================================================================================

/* VectorPartitioned(conn, dofs, iip)  (VectorPartitioned.h:72-661) */
class VectorPartitioned : public Vector {
public:
    template <class C, class D, class P>
    VectorPartitioned(const C& conn, const D& dofs, const P& iip) : Vector(conn, dofs)
    {
        auto p = amd::to_i64(iip);
        amd::check(gf_vecpart_create(&m_ph, handle(), p.data(), static_cast<int64_t>(p.size())));
        int64_t nu = 0, np = 0;
        amd::check(gf_vecpart_sizes(m_ph, &nu, &np));
        m_nnu = static_cast<size_t>(nu);
        m_nnp = static_cast<size_t>(np);
        std::vector<int64_t> iiu(m_nnu), iipv(m_nnp), part(nnode() * ndim());
        amd::check(gf_vecpart_index(m_ph, iiu.data(), iipv.data(), part.data()));
        m_iiu = amd::from_i64(iiu, {m_nnu});
        m_iip = amd::from_i64(iipv, {m_nnp});
        m_part = amd::from_i64(part, {nnode(), ndim()});
    }
    ~VectorPartitioned() { gf_vecpart_destroy(m_ph); }
    size_t nnu() const { return m_nnu; }
    size_t nnp() const { return m_nnp; }
    const amd::array<size_t>& iiu() const { return m_iiu; }
    const amd::array<size_t>& iip() const { return m_iip; }
    /* the reordered DOFs: unknowns numbered [0, nnu), prescribed [nnu, nnu + nnp) */
    const amd::array<size_t>& part() const { return m_part; }
    gf_vecpart part_handle() const { return m_ph; }

    /* AsDofs_u / AsDofs_p of a dofval, nodevec or elemvec  (VectorPartitioned.h:428-661) */
    template <class T>
    amd::array<double> AsDofs_u(const T& arg) const { return as_dofs(arg, true); }
    template <class T>
    amd::array<double> AsDofs_p(const T& arg) const { return as_dofs(arg, false); }
    /* DofsFromParitioned / NodeFromPartitioned / ElementFromPartitioned  (VectorPartitioned.h:276-424) */
    template <class U, class P>
    amd::array<double> DofsFromParitioned(const U& u, const P& p) const { return from_partitioned(0, u, p, shape_dofval()); }
    template <class U, class P>
    amd::array<double> NodeFromPartitioned(const U& u, const P& p) const { return from_partitioned(1, u, p, shape_nodevec()); }
    template <class U, class P>
    amd::array<double> ElementFromPartitioned(const U& u, const P& p) const
    {
        return from_partitioned(2, u, p, shape_elemvec());
    }
    /* Copy_u / Copy_p (VectorPartitioned.h:198-263): dest with its unknown / prescribed entries from src */
    template <class S, class T>
    amd::array<double> Copy_u(const S& src, const T& dest) const { return copy(0, src, dest); }
    template <class S, class T>
    amd::array<double> Copy_p(const S& src, const T& dest) const { return copy(1, src, dest); }

private:
    template <class T>
    static int kind_of(const T& arg) { return static_cast<int>(arg.dimension()) - 1; }
    template <class T>
    amd::array<double> as_dofs(const T& arg, bool u) const
    {
        amd::array<double> r({u ? m_nnu : m_nnp});
        amd::dev_in<T> in(arg);
        amd::dev_out<amd::array<double>> out(r, r.shape());
        amd::check(gf_vecpart_as_dofs(m_ph, kind_of(arg), in.ptr, u ? out.ptr : nullptr, u ? nullptr : out.ptr, nullptr));
        out.finish();
        return r;
    }
    template <class U, class P>
    amd::array<double> from_partitioned(int kind, const U& u, const P& p, std::vector<size_t> shape) const
    {
        amd::array<double> r(std::move(shape));
        amd::dev_in<U> du(u);
        amd::dev_in<P> dp(p);
        amd::dev_out<amd::array<double>> out(r, r.shape());
        amd::check(gf_vecpart_from_partitioned(m_ph, kind, du.ptr, dp.ptr, out.ptr, nullptr));
        out.finish();
        return r;
    }
    template <class S, class T>
    amd::array<double> copy(int which, const S& src, const T& dest) const
    {
        amd::array<double> r(amd::shape_of(dest));
        std::copy(dest.data(), dest.data() + r.size(), r.data());
        amd::dev_in<S> in(src);
        amd::dev_inout<amd::array<double>> io(r);
        amd::check(gf_vecpart_copy(m_ph, which, in.ptr, io.ptr, nullptr));
        io.finish();
        return r;
    }
    gf_vecpart m_ph = nullptr;
    size_t m_nnu = 0, m_nnp = 0;
    amd::array<size_t> m_iiu, m_iip, m_part;
};

/* MatrixDiagonalPartitioned(conn, dofs, iip)  (MatrixDiagonalPartitioned.h:42-452) */
class MatrixDiagonalPartitioned {
public:
    template <class C, class D, class P>
    MatrixDiagonalPartitioned(const C& conn, const D& dofs, const P& iip) : m_vec(conn, dofs, iip)
    {
        amd::check(gf_diagpart_create(&m_h, m_vec.part_handle()));
        amd::check(gf_diagpart_diag(m_h, &m_d));
    }
    ~MatrixDiagonalPartitioned() { gf_diagpart_destroy(m_h); }
    size_t ndof() const { return m_vec.ndof(); }
    size_t nnu() const { return m_vec.nnu(); }
    size_t nnp() const { return m_vec.nnp(); }
    const VectorPartitioned& vector() const { return m_vec; }
    /* assemble(elemmat)  (MatrixDiagonalPartitioned.h:73-98) */
    template <class T>
    void assemble(const T& elemmat)
    {
        amd::dev_in<T> in(elemmat);
        amd::check(gf_diag_assemble(m_d, in.ptr, nullptr));
        amd::check(gf_stream_sync(nullptr));
    }
    /* set(A)  (:121-135) */
    template <class T>
    void set(const T& A)
    {
        amd::dev_in<T> in(A);
        amd::check(gf_diagpart_set(m_h, in.ptr, nullptr));
        amd::check(gf_stream_sync(nullptr));
    }
    /* data(), data_uu(), data_pp()  (:142-175) */
    amd::array<double> data() const
    {
        const double* p = nullptr;
        amd::check(gf_diag_data(m_d, &p));
        amd::array<double> r({ndof()});
        amd::check(gf_memcpy_d2h(r.data(), p, static_cast<int64_t>(r.size() * sizeof(double))));
        return r;
    }
    amd::array<double> data_uu() const { return data_part(true); }
    amd::array<double> data_pp() const { return data_part(false); }
    /* Dot(x)  (:189-224) */
    template <class T>
    amd::array<double> Dot(const T& x) const
    {
        amd::array<double> r(amd::shape_of(x));
        amd::dev_in<T> in(x);
        amd::dev_out<amd::array<double>> out(r, r.shape());
        if (x.dimension() == 1) amd::check(gf_diag_dot_dofval(m_d, in.ptr, out.ptr, nullptr));
        else amd::check(gf_diag_dot_nodevec(m_d, in.ptr, out.ptr, nullptr));
        out.finish();
        return r;
    }
    /* Dot_u = A_uu x_u, Dot_p = A_pp x_p  (:248-300); Reaction_p (:405-420) is Dot_p */
    template <class U, class P>
    amd::array<double> Dot_u(const U& x_u, const P&) const { return dot_part(0, x_u, nnu()); }
    template <class U, class P>
    amd::array<double> Dot_p(const U&, const P& x_p) const { return dot_part(1, x_p, nnp()); }
    template <class U, class P>
    amd::array<double> Reaction_p(const U& x_u, const P& x_p) const { return Dot_p(x_u, x_p); }
    /* Solve(b, x) / solve(b, x)  (:306-337): the unknown entries A_uu^-1 b_u */
    template <class B, class X>
    amd::array<double> Solve(const B& b, const X& x)
    {
        amd::array<double> r(amd::shape_of(x));
        std::copy(x.data(), x.data() + r.size(), r.data());
        solve(b, r);
        return r;
    }
    template <class B, class X>
    void solve(const B& b, X& x)
    {
        apply_sel(0, b, x);
    }
    /* Solve_u(b_u, x_p)  (:356-372) */
    template <class B, class P>
    amd::array<double> Solve_u(const B& b_u, const P&)
    {
        amd::array<double> r({nnu()});
        amd::dev_in<B> in(b_u);
        amd::dev_out<amd::array<double>> out(r, r.shape());
        amd::check(gf_diagpart_solve_u(m_h, in.ptr, out.ptr, nullptr));
        out.finish();
        return r;
    }
    /* Reaction(x, b) / reaction(x, b)  (Matrix.h:490-514, :378-403) */
    template <class X, class B>
    amd::array<double> Reaction(const X& x, const B& b) const
    {
        amd::array<double> r(amd::shape_of(b));
        std::copy(b.data(), b.data() + r.size(), r.data());
        reaction(x, r);
        return r;
    }
    template <class X, class B>
    void reaction(const X& x, B& b) const
    {
        apply_sel(1, x, b);
    }

private:
    amd::array<double> data_part(bool u) const
    {
        amd::array<double> r({u ? nnu() : nnp()});
        amd::dev_out<amd::array<double>> out(r, r.shape());
        amd::check(gf_diagpart_data_part(m_h, u ? out.ptr : nullptr, u ? nullptr : out.ptr, nullptr));
        out.finish();
        return r;
    }
    template <class T>
    amd::array<double> dot_part(int which, const T& x, size_t n) const
    {
        amd::array<double> r({n});
        amd::dev_in<T> in(x);
        amd::dev_out<amd::array<double>> out(r, r.shape());
        amd::check(gf_diagpart_dot_part(m_h, which, in.ptr, out.ptr, nullptr));
        out.finish();
        return r;
    }
    template <class T, class R>
    void apply_sel(int op, const T& in_arg, R& out_arg) const
    {
        amd::dev_in<T> in(in_arg);
        amd::dev_inout<R> io(out_arg);
        amd::check(gf_diagpart_apply_sel(m_h, op, in_arg.dimension() == 1 ? 0 : 1, in.ptr, io.ptr, nullptr));
        io.finish();
    }
    VectorPartitioned m_vec;
    gf_diagpart m_h = nullptr;
    gf_diag m_d = nullptr;
};

/* MatrixPartitioned(conn, dofs, iip)  (MatrixPartitioned.h:34-523): CSR blocks uu, up, pu, pp */
class MatrixPartitioned {
public:
    template <class C, class D, class P>
    MatrixPartitioned(const C& conn, const D& dofs, const P& iip) : m_vec(conn, dofs, iip)
    {
        amd::check(gf_matpart_create(&m_h, m_vec.part_handle()));
        amd::check(gf_matpart_sizes(m_h, m_nnz));
    }
    ~MatrixPartitioned() { gf_matpart_destroy(m_h); }
    size_t ndof() const { return m_vec.ndof(); }
    size_t nnu() const { return m_vec.nnu(); }
    size_t nnp() const { return m_vec.nnp(); }
    const amd::array<size_t>& iiu() const { return m_vec.iiu(); }
    const amd::array<size_t>& iip() const { return m_vec.iip(); }
    const VectorPartitioned& vector() const { return m_vec; }
    gf_matpart handle() const { return m_h; }
    /* assemble(elemmat)  (MatrixPartitioned.h:156-210) */
    template <class T>
    void assemble(const T& elemmat)
    {
        amd::dev_in<T> in(elemmat);
        amd::check(gf_matpart_assemble(m_h, in.ptr, nullptr));
        amd::check(gf_stream_sync(nullptr));
    }
    /* data_uu() ... data_pp()  (:126-153) as host CSR copies */
    amd::SparseMatrix data_uu() const { return block(0); }
    amd::SparseMatrix data_up() const { return block(1); }
    amd::SparseMatrix data_pu() const { return block(2); }
    amd::SparseMatrix data_pp() const { return block(3); }
    /* Todense()  (:324-350) */
    amd::array<double> Todense() const
    {
        gf_matrix full = nullptr;
        amd::check(gf_matpart_matrix(m_h, &full));
        amd::array<double> r({ndof(), ndof()});
        amd::dev_out<amd::array<double>> out(r, r.shape());
        amd::check(gf_matrix_todense(full, out.ptr, nullptr));
        out.finish();
        return r;
    }
    /* Dot(x)  (:354-396) */
    template <class T>
    amd::array<double> Dot(const T& x) const
    {
        amd::array<double> r(amd::shape_of(x));
        amd::dev_in<T> in(x);
        amd::dev_out<amd::array<double>> out(r, r.shape());
        amd::check(gf_matpart_apply(m_h, 0, x.dimension() == 1 ? 0 : 1, in.ptr, out.ptr, nullptr));
        out.finish();
        return r;
    }
    /* Reaction(x, b) / reaction(x, b)  (Matrix.h:490-514, :401-433) */
    template <class X, class B>
    amd::array<double> Reaction(const X& x, const B& b) const
    {
        amd::array<double> r(amd::shape_of(b));
        std::copy(b.data(), b.data() + r.size(), r.data());
        reaction(x, r);
        return r;
    }
    template <class X, class B>
    void reaction(const X& x, B& b) const
    {
        amd::dev_in<X> in(x);
        amd::dev_inout<B> io(b);
        amd::check(gf_matpart_apply(m_h, 1, x.dimension() == 1 ? 0 : 1, in.ptr, io.ptr, nullptr));
        io.finish();
    }
    /* Reaction_p(x_u, x_p)  (:436-447) */
    template <class U, class P>
    amd::array<double> Reaction_p(const U& x_u, const P& x_p) const
    {
        amd::array<double> r({nnp()});
        amd::dev_in<U> du(x_u);
        amd::dev_in<P> dp(x_p);
        amd::dev_out<amd::array<double>> out(r, r.shape());
        amd::check(gf_matpart_dot_part(m_h, 1, du.ptr, dp.ptr, out.ptr, nullptr));
        out.finish();
        return r;
    }

private:
    amd::SparseMatrix block(int b) const
    {
        const int32_t *rp = nullptr, *cl = nullptr;
        const double* val = nullptr;
        amd::check(gf_matpart_csr(m_h, b, &rp, &cl, &val));
        amd::SparseMatrix M;
        M.rows = b < 2 ? nnu() : nnp();
        M.cols = (b % 2 == 0) ? nnu() : nnp();
        const size_t nnz = static_cast<size_t>(m_nnz[b]);
        std::vector<int32_t> r32(M.rows + 1), c32(nnz);
        M.val.resize(nnz);
        amd::check(gf_memcpy_d2h(r32.data(), rp, static_cast<int64_t>(r32.size() * sizeof(int32_t))));
        if (nnz) {
            amd::check(gf_memcpy_d2h(c32.data(), cl, static_cast<int64_t>(nnz * sizeof(int32_t))));
            amd::check(gf_memcpy_d2h(M.val.data(), val, static_cast<int64_t>(nnz * sizeof(double))));
        }
        M.rowptr.assign(r32.begin(), r32.end());
        M.col.assign(c32.begin(), c32.end());
        return M;
    }
    VectorPartitioned m_vec;
    gf_matpart m_h = nullptr;
    int64_t m_nnz[4] = {0, 0, 0, 0};
};

/* MatrixPartitionedSolver  (MatrixPartitioned.h:526-640): x_u = A_uu^-1 (b_u - A_up x_p) by an FP64
 * Jacobi-preconditioned CG on the device (A_uu symmetric positive definite) in place of the
 * reference's SimplicialLDLT.  Throws if the true relative residual exceeds 10 tol. */
class MatrixPartitionedSolver {
public:
    explicit MatrixPartitionedSolver(double tol = 1e-12, int64_t maxiter = 0) : m_tol(tol), m_maxiter(maxiter) {}
    int64_t iterations() const { return m_it; }
    double relres() const { return m_res; }
    /* Solve(A, b, x) / solve(A, b, x)  (:583-613) */
    template <class B, class X>
    amd::array<double> Solve(MatrixPartitioned& A, const B& b, const X& x)
    {
        amd::array<double> r(amd::shape_of(x));
        std::copy(x.data(), x.data() + r.size(), r.data());
        solve(A, b, r);
        return r;
    }
    template <class B, class X>
    void solve(MatrixPartitioned& A, const B& b, X& x)
    {
        amd::dev_in<B> in(b);
        amd::dev_inout<X> io(x);
        amd::check(gf_matpart_solve(A.handle(), b.dimension() == 1 ? 0 : 1, in.ptr, io.ptr, 0, m_tol, cap(A), &m_it,
                                    &m_res, nullptr));
        io.finish();
    }
    /* Solve_u(A, b_u, x_p) / solve_u(A, b_u, x_p, x_u)  (:545-580, :615-631) */
    template <class B, class P>
    amd::array<double> Solve_u(MatrixPartitioned& A, const B& b_u, const P& x_p)
    {
        amd::array<double> r({A.nnu()});
        solve_u(A, b_u, x_p, r);
        return r;
    }
    template <class B, class P, class U>
    void solve_u(MatrixPartitioned& A, const B& b_u, const P& x_p, U& x_u)
    {
        amd::dev_in<B> db(b_u);
        amd::dev_in<P> dp(x_p);
        amd::dev_out<U> out(x_u, {A.nnu()});
        amd::check(gf_matpart_solve_u(A.handle(), db.ptr, dp.ptr, out.ptr, 0, m_tol, cap(A), &m_it, &m_res, nullptr));
        out.finish();
    }

private:
    int64_t cap(const MatrixPartitioned& A) const
    {
        return m_maxiter > 0 ? m_maxiter : std::max<int64_t>(100, 10 * static_cast<int64_t>(A.nnu()));
    }
    double m_tol;
    int64_t m_maxiter;
    int64_t m_it = 0;
    double m_res = 0.0;
};

// =================================================================================================
// Tyings (include/GooseFEM/TyingsPeriodic.h)
// =================================================================================================
namespace Tyings {

/* Compressed sparse rows of a tying matrix (in place of the reference's Eigen::SparseMatrix). */
using SparseMatrix = amd::SparseMatrix;

/* Tyings::Control  (TyingsPeriodic.h:361-439) */
class Control {
public:
    template <class C, class N>
    Control(const C& coor, const N& dofs)
    {
        const size_t nnode = coor.shape()[0], ndim = coor.shape()[1];
        std::vector<double> c(coor.data(), coor.data() + coor.size());
        auto d = amd::to_i64(dofs);
        std::vector<double> co((nnode + ndim) * ndim);
        std::vector<int64_t> dout((nnode + ndim) * ndim), cd(ndim * ndim), cn(ndim);
        amd::check(gf_tyings_control(c.data(), d.data(), static_cast<int64_t>(nnode), static_cast<int>(ndim), co.data(),
                                     dout.data(), cd.data(), cn.data()));
        m_coor = amd::array<double>({nnode + ndim, ndim});
        std::copy(co.begin(), co.end(), m_coor.data());
        m_dofs = amd::from_i64(dout, {nnode + ndim, ndim});
        m_control_dofs = amd::from_i64(cd, {ndim, ndim});
        m_control_nodes = amd::from_i64(cn, {ndim});
    }
    const amd::array<double>& coor() const { return m_coor; }
    const amd::array<size_t>& dofs() const { return m_dofs; }
    const amd::array<size_t>& controlDofs() const { return m_control_dofs; }
    const amd::array<size_t>& controlNodes() const { return m_control_nodes; }

private:
    amd::array<double> m_coor;
    amd::array<size_t> m_dofs, m_control_dofs, m_control_nodes;
};

/* Tyings::Periodic  (TyingsPeriodic.h:82-343) */
class Periodic {
public:
    template <class C, class D, class S, class T, class U>
    Periodic(const C& coor, const D& dofs, const S& control_dofs, const T& nodal_tyings, const U& iip)
    {
        m_ndim = coor.shape()[1];
        m_coor.assign(coor.data(), coor.data() + coor.size());
        auto d = amd::to_i64(dofs);
        auto c = amd::to_i64(control_dofs);
        m_ties = amd::to_i64(nodal_tyings);
        auto p = amd::to_i64(iip);
        std::vector<int64_t> dout(d.size()), cout(c.size()), sizes(4);
        amd::check(gf_tyings_periodic(d.data(), static_cast<int64_t>(dofs.shape()[0]), static_cast<int>(m_ndim), c.data(),
                                      m_ties.data(), static_cast<int64_t>(m_ties.size() / 2), p.data(),
                                      static_cast<int64_t>(p.size()), dout.data(), cout.data(), sizes.data()));
        m_dofs = amd::from_i64(dout, amd::shape_of(dofs));
        m_control = amd::from_i64(cout, {m_ndim, m_ndim});
        m_nnu = static_cast<size_t>(sizes[0]);
        m_nnp = static_cast<size_t>(sizes[1]);
        m_nni = static_cast<size_t>(sizes[2]);
        m_nnd = static_cast<size_t>(sizes[3]);
    }
    size_t nnu() const { return m_nnu; }
    size_t nnp() const { return m_nnp; }
    size_t nni() const { return m_nni; }
    size_t nnd() const { return m_nnd; }
    const amd::array<size_t>& dofs() const { return m_dofs; }
    const amd::array<size_t>& control() const { return m_control; }
    SparseMatrix Cdi() const { return matrix(0, m_nni); }
    SparseMatrix Cdu() const { return matrix(1, m_nnu); }
    SparseMatrix Cdp() const { return matrix(2, m_nnp); }

private:
    SparseMatrix matrix(int which, size_t ncols) const
    {
        auto d = amd::to_i64(m_dofs);
        auto c = amd::to_i64(m_control);
        const int64_t nties = static_cast<int64_t>(m_ties.size() / 2);
        SparseMatrix M;
        M.rows = m_nnd;
        M.cols = ncols;
        M.rowptr.resize(m_nnd + 1);
        int64_t nnz = 0;
        amd::check(gf_tyings_matrix(which, m_coor.data(), d.data(), static_cast<int>(m_ndim), c.data(), m_ties.data(),
                                    nties, static_cast<int64_t>(m_nnu), M.rowptr.data(), nullptr, nullptr, &nnz));
        M.col.resize(static_cast<size_t>(nnz));
        M.val.resize(static_cast<size_t>(nnz));
        amd::check(gf_tyings_matrix(which, m_coor.data(), d.data(), static_cast<int>(m_ndim), c.data(), m_ties.data(),
                                    nties, static_cast<int64_t>(m_nnu), M.rowptr.data(), M.col.data(), M.val.data(),
                                    &nnz));
        return M;
    }
    size_t m_ndim = 0, m_nnu = 0, m_nnp = 0, m_nni = 0, m_nnd = 0;
    std::vector<double> m_coor;
    std::vector<int64_t> m_ties;
    amd::array<size_t> m_dofs, m_control;
};

} // namespace Tyings

// =================================================================================================
// MatrixPartitionedTyings + MatrixPartitionedTyingsSolver (include/GooseFEM/MatrixPartitionedTyings.h)
// for any element matrices: K as one CSR over the tied numbering [u | p | d], the condensed system
// solved by a matrix-free FP64 Jacobi-PCG (K' symmetric positive definite) instead of LDLT
// =================================================================================================

class MatrixPartitionedTyings {
public:
    template <class C, class D>
    MatrixPartitionedTyings(const C& conn, const D& dofs, const Tyings::SparseMatrix& Cdu, const Tyings::SparseMatrix& Cdp)
        : m_vec(conn, dofs)
    {
        m_nnu = Cdu.cols;
        m_nnp = Cdp.cols;
        m_nnd = Cdu.rows;
        amd::check(gf_mattyings_create(&m_h, m_vec.handle(), static_cast<int64_t>(m_nnu), static_cast<int64_t>(m_nnp),
                                       static_cast<int64_t>(m_nnd), Cdu.rowptr.data(), Cdu.col.data(), Cdu.val.data(),
                                       Cdp.rowptr.data(), Cdp.col.data(), Cdp.val.data()));
    }
    MatrixPartitionedTyings(const MatrixPartitionedTyings&) = delete;
    MatrixPartitionedTyings& operator=(const MatrixPartitionedTyings&) = delete;
    ~MatrixPartitionedTyings() { gf_mattyings_destroy(m_h); }
    size_t nnu() const { return m_nnu; }
    size_t nnp() const { return m_nnp; }
    size_t nnd() const { return m_nnd; }
    size_t ndof() const { return m_vec.ndof(); }
    gf_mattyings handle() const { return m_h; }
    /* assemble(elemmat)  (MatrixPartitionedTyings.h:210-305) */
    template <class T>
    void assemble(const T& elemmat)
    {
        amd::dev_in<T> in(elemmat);
        amd::check(gf_mattyings_assemble(m_h, in.ptr, nullptr));
        amd::check(gf_stream_sync(nullptr));
    }
    amd::array<double> Todense() const
    {
        gf_matrix K = nullptr;
        amd::check(gf_mattyings_matrix(m_h, &K));
        amd::array<double> r({ndof(), ndof()});
        amd::dev_out<amd::array<double>> out(r, r.shape());
        amd::check(gf_matrix_todense(K, out.ptr, nullptr));
        out.finish();
        return r;
    }

private:
    Vector m_vec;
    gf_mattyings m_h = nullptr;
    size_t m_nnu = 0, m_nnp = 0, m_nnd = 0;
};

class MatrixPartitionedTyingsSolver {
public:
    explicit MatrixPartitionedTyingsSolver(double tol = 1e-12, int64_t maxiter = 0) : m_tol(tol), m_maxiter(maxiter) {}
    int64_t iterations() const { return m_it; }
    double relres() const { return m_res; }
    /* solve(A, b, x) / Solve(A, b, x)  (MatrixPartitionedTyings.h:540-564) on nodevecs */
    template <class B, class X>
    void solve(MatrixPartitionedTyings& A, const B& b, X& x)
    {
        amd::dev_in<B> in(b);
        amd::dev_inout<X> io(x);
        const int64_t cap = m_maxiter > 0 ? m_maxiter : std::max<int64_t>(100, 10 * static_cast<int64_t>(A.nnu()));
        amd::check(gf_mattyings_solve(A.handle(), in.ptr, io.ptr, m_tol, cap, &m_it, &m_res, nullptr));
        io.finish();
    }
    template <class B, class X>
    amd::array<double> Solve(MatrixPartitionedTyings& A, const B& b, const X& x)
    {
        amd::array<double> r(amd::shape_of(x));
        std::copy(x.data(), x.data() + r.size(), r.data());
        solve(A, b, r);
        return r;
    }

private:
    double m_tol;
    int64_t m_maxiter;
    int64_t m_it = 0;
    double m_res = 0.0;
};

// =================================================================================================
// Iterate::StopList  (include/GooseFEM/Iterate.h): the last n residuals of an iteration
// =================================================================================================
namespace Iterate {

class StopList {
public:
    explicit StopList(size_t n = 1) { reset(n); }
    void reset() { std::fill(m_res.begin(), m_res.end(), std::numeric_limits<double>::infinity()); }
    void reset(size_t n)
    {
        m_res.assign(n, std::numeric_limits<double>::infinity());
    }
    void roll_insert(double res)
    {
        if (m_res.empty()) return;
        for (size_t i = 1; i < m_res.size(); ++i) m_res[i - 1] = m_res[i];
        m_res.back() = res;
    }
    bool descending() const
    {
        for (size_t i = 1; i < m_res.size(); ++i)
            if (m_res[i] > m_res[i - 1]) return false;
        return true;
    }
    bool all_less(double tol) const
    {
        for (double r : m_res)
            if (r >= tol) return false;
        return true;
    }
    const std::vector<double>& data() const { return m_res; }

private:
    std::vector<double> m_res;
};

} // namespace Iterate

} // namespace GooseFEM

#endif
