// HDF5 field output of the examples (docs/examples/dynamics/Elastic-VelocityVerlet/main.cpp:161-176:
// "/global/{t,Ekin,Epot}", "/mesh/{conn,coor,disp}"), with the semantics of the H5Easy calls the
// reference's examples use (HighFive's H5Easy::File, dump, load, exist, getShape), over the HDF5 C
// library.  Host code; built as its own library (libgoosefem_amd_h5.so) so that the GPU library
// does not depend on HDF5.
//   dump(file, path, array)          a dataset of the array's rank and shape; intermediate groups
//                                    are created; an existing dataset is an error unless overwrite
//                                    is asked, in which case its shape must match
//   dump(file, path, value, index)   entry `index` of an extendible 1-D dataset (created chunked,
//                                    grown as needed): the per-increment scalars of an example
//   load / shape / exist             read back (with HDF5's numeric conversion to the asked type)
#include <hdf5.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/goosefem_amd_h5.h"

namespace {

thread_local std::string g_error;

struct Fail {
    std::string msg;
};

template <class F>
int guard(F&& f)
{
    try {
        f();
        return GF_OK;
    }
    catch (const Fail& e) {
        g_error = e.msg;
        return GF_EINVAL;
    }
    catch (const std::exception& e) {
        g_error = e.what();
        return GF_EINVAL;
    }
}

void require(bool ok, const std::string& msg)
{
    if (!ok) throw Fail{msg};
}

// HDF5 object handle closed at scope exit
struct Id {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~Id()
    {
        if (id >= 0 && close) close(id);
    }
    Id(const Id&) = delete;
    Id& operator=(const Id&) = delete;
    operator hid_t() const { return id; }
};

hid_t native(int dtype)
{
    switch (dtype) {
        case GF_H5_F64: return H5T_NATIVE_DOUBLE;
        case GF_H5_I64: return H5T_NATIVE_INT64;
        case GF_H5_U64: return H5T_NATIVE_UINT64;
        default: throw Fail{"dtype must be GF_H5_F64, GF_H5_I64 or GF_H5_U64"};
    }
}

// does every component of the absolute path exist (as a link)?
bool exists(hid_t file, const std::string& path)
{
    require(!path.empty() && path[0] == '/', "path must be absolute (start with '/')");
    size_t pos = 1;
    while (pos <= path.size()) {
        size_t next = path.find('/', pos);
        if (next == std::string::npos) next = path.size();
        const std::string prefix = path.substr(0, next);
        if (prefix.size() > 1) {
            htri_t e = H5Lexists(file, prefix.c_str(), H5P_DEFAULT);
            if (e <= 0) return false;
        }
        pos = next + 1;
    }
    return true;
}

struct gf_h5_file {
    hid_t id;
};

gf_h5_file* F(gf_h5 f)
{
    require(f != nullptr, "null file");
    return reinterpret_cast<gf_h5_file*>(f);
}

std::vector<hsize_t> dims_of(hid_t dset)
{
    Id space(H5Dget_space(dset), H5Sclose);
    require(space >= 0, "cannot read the dataset's dataspace");
    int r = H5Sget_simple_extent_ndims(space);
    require(r >= 0, "cannot read the dataset's rank");
    std::vector<hsize_t> d(static_cast<size_t>(r));
    if (r) H5Sget_simple_extent_dims(space, d.data(), nullptr);
    return d;
}

} // namespace

extern "C" {

const char* gf_h5_last_error(void) { return g_error.c_str(); }

int gf_h5_open(gf_h5* out, const char* path, int mode)
{
    return guard([&] {
        require(out && path, "null argument");
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); // errors come back as status codes
        hid_t id = -1;
        if (mode == GF_H5_OVERWRITE) id = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
        else if (mode == GF_H5_READONLY) id = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
        else if (mode == GF_H5_READWRITE) {
            htri_t is = H5Fis_hdf5(path);
            id = is > 0 ? H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT) : H5Fcreate(path, H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT);
        }
        else throw Fail{"mode must be GF_H5_OVERWRITE, GF_H5_READONLY or GF_H5_READWRITE"};
        require(id >= 0, std::string("cannot open ") + path);
        *out = reinterpret_cast<gf_h5>(new gf_h5_file{id});
    });
}

int gf_h5_close(gf_h5 f)
{
    if (!f) return GF_OK;
    auto* p = reinterpret_cast<gf_h5_file*>(f);
    herr_t e = H5Fclose(p->id);
    delete p;
    if (e < 0) {
        g_error = "H5Fclose failed";
        return GF_EINVAL;
    }
    return GF_OK;
}

int gf_h5_exist(gf_h5 f, const char* path, int* exist)
{
    return guard([&] {
        require(path && exist, "null argument");
        *exist = exists(F(f)->id, path) ? 1 : 0;
    });
}

int gf_h5_dump(gf_h5 f, const char* path, const void* data_h, int dtype, int rank, const int64_t* shape, int overwrite)
{
    return guard([&] {
        require(path && (data_h || rank > 0) && rank >= 0 && rank <= 8, "path, data; 0 <= rank <= 8");
        hid_t file = F(f)->id;
        std::vector<hsize_t> dims(static_cast<size_t>(rank));
        for (int i = 0; i < rank; ++i) {
            require(shape && shape[i] >= 0, "shape entries >= 0");
            dims[i] = static_cast<hsize_t>(shape[i]);
        }
        hid_t mem = native(dtype);
        if (exists(file, path)) {
            require(overwrite != 0, std::string("dataset exists: ") + path + " (dump with overwrite)");
            Id dset(H5Dopen2(file, path, H5P_DEFAULT), H5Dclose);
            require(dset >= 0, std::string("cannot open ") + path);
            require(dims_of(dset) == dims, std::string("overwrite of ") + path + " with another shape");
            require(H5Dwrite(dset, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, data_h) >= 0, std::string("write failed: ") + path);
            return;
        }
        Id space(rank ? H5Screate_simple(rank, dims.data(), nullptr) : H5Screate(H5S_SCALAR), H5Sclose);
        Id lcpl(H5Pcreate(H5P_LINK_CREATE), H5Pclose);
        H5Pset_create_intermediate_group(lcpl, 1);
        Id dset(H5Dcreate2(file, path, mem, space, lcpl, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
        require(dset >= 0, std::string("cannot create ") + path);
        size_t n = 1;
        for (auto d : dims) n *= d;
        if (n) require(H5Dwrite(dset, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, data_h) >= 0, std::string("write failed: ") + path);
    });
}

int gf_h5_dump_at(gf_h5 f, const char* path, const void* value_h, int dtype, int64_t index)
{
    return guard([&] {
        require(path && value_h && index >= 0, "path, value; index >= 0");
        hid_t file = F(f)->id;
        hid_t mem = native(dtype);
        const hsize_t at = static_cast<hsize_t>(index), one = 1;
        if (!exists(file, path)) {
            const hsize_t init = at + 1, maxd = H5S_UNLIMITED, chunk = 1024;
            Id space(H5Screate_simple(1, &init, &maxd), H5Sclose);
            Id dcpl(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
            H5Pset_chunk(dcpl, 1, &chunk);
            Id lcpl(H5Pcreate(H5P_LINK_CREATE), H5Pclose);
            H5Pset_create_intermediate_group(lcpl, 1);
            Id dset(H5Dcreate2(file, path, mem, space, lcpl, dcpl, H5P_DEFAULT), H5Dclose);
            require(dset >= 0, std::string("cannot create ") + path);
        }
        Id dset(H5Dopen2(file, path, H5P_DEFAULT), H5Dclose);
        require(dset >= 0, std::string("cannot open ") + path);
        auto dims = dims_of(dset);
        require(dims.size() == 1, std::string(path) + " is not 1-D");
        if (at >= dims[0]) {
            const hsize_t grow = at + 1;
            require(H5Dset_extent(dset, &grow) >= 0, std::string(path) + " is not extendible");
        }
        Id fspace(H5Dget_space(dset), H5Sclose);
        H5Sselect_hyperslab(fspace, H5S_SELECT_SET, &at, nullptr, &one, nullptr);
        Id mspace(H5Screate_simple(1, &one, nullptr), H5Sclose);
        require(H5Dwrite(dset, mem, mspace, fspace, H5P_DEFAULT, value_h) >= 0, std::string("write failed: ") + path);
    });
}

int gf_h5_shape(gf_h5 f, const char* path, int* rank, int64_t* shape)
{
    return guard([&] {
        require(path && rank, "null argument");
        hid_t file = F(f)->id;
        require(exists(file, path), std::string("no dataset ") + path);
        Id dset(H5Dopen2(file, path, H5P_DEFAULT), H5Dclose);
        require(dset >= 0, std::string("cannot open ") + path);
        auto d = dims_of(dset);
        require(d.size() <= 8, "rank above 8");
        *rank = static_cast<int>(d.size());
        if (shape)
            for (size_t i = 0; i < d.size(); ++i) shape[i] = static_cast<int64_t>(d[i]);
    });
}

int gf_h5_load(gf_h5 f, const char* path, void* data_h, int dtype)
{
    return guard([&] {
        require(path && data_h, "null argument");
        hid_t file = F(f)->id;
        require(exists(file, path), std::string("no dataset ") + path);
        Id dset(H5Dopen2(file, path, H5P_DEFAULT), H5Dclose);
        require(dset >= 0, std::string("cannot open ") + path);
        require(H5Dread(dset, native(dtype), H5S_ALL, H5S_ALL, H5P_DEFAULT, data_h) >= 0,
                std::string("read failed: ") + path);
    });
}

} // extern "C"
