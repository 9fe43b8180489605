/*
 * GooseFEM/H5Easy.h -- header-only C++ face of libgoosefem_amd_h5.so with the names of the H5Easy
 * calls the reference's examples make (HighFive's H5Easy::File, dump, load, exist, getShape;
 * docs/examples/dynamics/Elastic-VelocityVerlet/main.cpp:161-176), so that their output code
 * compiles unchanged:
 *
 *     H5Easy::File file("example.hdf5", H5Easy::File::Overwrite);
 *     H5Easy::dump(file, "/global/Epot", Epot);
 *     H5Easy::dump(file, "/mesh/conn", conn);
 *
 * Containers: any row-major array with data(), size() and shape() (xt::xtensor, GooseFEM::amd::array);
 * load<C> needs C::from_shape.  Element types: floating point is stored as double, unsigned integers
 * (size_t) as uint64, signed integers as int64.  Errors throw std::runtime_error.
 * Link with -lgoosefem_amd_h5.
 */
#ifndef GOOSEFEM_AMD_H5EASY_H
#define GOOSEFEM_AMD_H5EASY_H

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../goosefem_amd_h5.h"

namespace H5Easy {

enum class DumpMode { Create, Overwrite };

namespace detail {

inline void check(int rc)
{
    if (rc != GF_OK) throw std::runtime_error(std::string("H5Easy: ") + gf_h5_last_error());
}

// storage type of an element type, and the library's code for it
template <class T>
using stored_t = std::conditional_t<std::is_floating_point<T>::value, double,
                                    std::conditional_t<std::is_unsigned<T>::value, uint64_t, int64_t>>;

template <class S>
constexpr int code()
{
    return std::is_same<S, double>::value ? GF_H5_F64 : (std::is_same<S, uint64_t>::value ? GF_H5_U64 : GF_H5_I64);
}

} // namespace detail

/* H5Easy::File(path, mode) */
class File {
public:
    enum Mode { Overwrite = GF_H5_OVERWRITE, ReadOnly = GF_H5_READONLY, ReadWrite = GF_H5_READWRITE };

    explicit File(const std::string& path, int mode = ReadOnly) { detail::check(gf_h5_open(&m_h, path.c_str(), mode)); }
    ~File()
    {
        if (m_h) gf_h5_close(m_h);
    }
    File(const File&) = delete;
    File& operator=(const File&) = delete;
    gf_h5 handle() const { return m_h; }

private:
    gf_h5 m_h = nullptr;
};

inline bool exist(const File& file, const std::string& path)
{
    int e = 0;
    detail::check(gf_h5_exist(file.handle(), path.c_str(), &e));
    return e != 0;
}

inline std::vector<size_t> getShape(const File& file, const std::string& path)
{
    int rank = 0;
    int64_t shape[8] = {};
    detail::check(gf_h5_shape(file.handle(), path.c_str(), &rank, shape));
    return std::vector<size_t>(shape, shape + rank);
}

/* H5Easy::dump(file, path, array[, mode]) */
template <class C, std::enable_if_t<!std::is_arithmetic<C>::value, int> = 0>
void dump(File& file, const std::string& path, const C& data, DumpMode mode = DumpMode::Create)
{
    using T = std::decay_t<decltype(*data.data())>;
    using S = detail::stored_t<T>;
    std::vector<int64_t> shape;
    for (auto s : data.shape()) shape.push_back(static_cast<int64_t>(s));
    const size_t n = static_cast<size_t>(data.size());
    const void* ptr = data.data();
    std::vector<S> buf;
    if (!std::is_same<T, S>::value) {
        buf.assign(data.data(), data.data() + n);
        ptr = buf.data();
    }
    detail::check(gf_h5_dump(file.handle(), path.c_str(), ptr, detail::code<S>(), static_cast<int>(shape.size()),
                             shape.data(), mode == DumpMode::Overwrite));
}

/* H5Easy::dump(file, path, scalar[, mode]) */
template <class T, std::enable_if_t<std::is_arithmetic<T>::value, int> = 0>
void dump(File& file, const std::string& path, T value, DumpMode mode = DumpMode::Create)
{
    using S = detail::stored_t<T>;
    const S v = static_cast<S>(value);
    detail::check(gf_h5_dump(file.handle(), path.c_str(), &v, detail::code<S>(), 0, nullptr, mode == DumpMode::Overwrite));
}

/* H5Easy::dump(file, path, value, {index}): entry of an extendible 1-D dataset */
template <class T, std::enable_if_t<std::is_arithmetic<T>::value, int> = 0>
void dump(File& file, const std::string& path, T value, const std::vector<size_t>& index)
{
    if (index.size() != 1) throw std::runtime_error("H5Easy: dump at an index of a 1-D dataset only");
    using S = detail::stored_t<T>;
    const S v = static_cast<S>(value);
    detail::check(gf_h5_dump_at(file.handle(), path.c_str(), &v, detail::code<S>(), static_cast<int64_t>(index[0])));
}

/* H5Easy::load<C>(file, path) */
template <class C>
C load(const File& file, const std::string& path)
{
    using T = std::decay_t<decltype(*std::declval<C&>().data())>;
    using S = detail::stored_t<T>;
    C out = C::from_shape(getShape(file, path));
    if (std::is_same<T, S>::value) {
        detail::check(gf_h5_load(file.handle(), path.c_str(), out.data(), detail::code<S>()));
        return out;
    }
    std::vector<S> buf(static_cast<size_t>(out.size()));
    detail::check(gf_h5_load(file.handle(), path.c_str(), buf.data(), detail::code<S>()));
    for (size_t i = 0; i < buf.size(); ++i) out.data()[i] = static_cast<T>(buf[i]);
    return out;
}

} // namespace H5Easy

#endif
