// The output code of docs/examples/dynamics/Elastic-VelocityVerlet/main.cpp:161-176 (and the
// per-increment series of its loop) compiled against include/GooseFEM/H5Easy.h, written and read
// back through libgoosefem_amd_h5.so.  Host only.
#include <GooseFEM/GooseFEM.h>
#include <GooseFEM/H5Easy.h>

#include <cstdio>
#include <string>

static int g_fail = 0;
#define REQUIRE(c)                                                                                  \
    do {                                                                                            \
        if (!(c)) {                                                                                 \
            std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #c);                              \
            ++g_fail;                                                                               \
        }                                                                                           \
    } while (0)

int main(int argc, char** argv)
{
    namespace GF = GooseFEM;
    namespace H5 = H5Easy;
    const std::string path = argc > 1 ? argv[1] : "h5easy_test.hdf5";
    GF::Mesh::Quad4::Regular mesh(6, 4, 1.0);
    auto coor = mesh.coor();
    auto conn = mesh.conn();
    GF::amd::array<double> u({mesh.nnode(), 2});
    for (size_t i = 0; i < u.size(); ++i) u[i] = 0.001 * static_cast<double>(i);
    GF::amd::array<double> Epot({5}), Ekin({5}), t({5});
    {
        H5::File file(path, H5::File::Overwrite);
        for (size_t inc = 0; inc < 5; ++inc) {
            t[inc] = static_cast<double>(inc) * 0.01;
            Ekin[inc] = 0.5 * static_cast<double>(inc);
            Epot[inc] = 2.0 * static_cast<double>(inc);
            H5::dump(file, "/series/Ekin", Ekin[inc], {inc});
        }
        H5::dump(file, "/global/Epot", Epot);
        H5::dump(file, "/global/Ekin", Ekin);
        H5::dump(file, "/global/t", t);
        H5::dump(file, "/mesh/conn", conn);
        H5::dump(file, "/mesh/coor", coor);
        H5::dump(file, "/mesh/disp", u);
        H5::dump(file, "/meta/nelem", mesh.nelem());
        bool threw = false;
        try {
            H5::dump(file, "/mesh/disp", u);
        }
        catch (const std::runtime_error&) {
            threw = true;
        }
        REQUIRE(threw);
        H5::dump(file, "/mesh/disp", u, H5::DumpMode::Overwrite);
    }
    {
        H5::File file(path, H5::File::ReadOnly);
        REQUIRE(H5::exist(file, "/mesh/conn") && !H5::exist(file, "/mesh/none"));
        auto c = H5::load<GF::amd::array<size_t>>(file, "/mesh/conn");
        REQUIRE(c.shape() == conn.shape());
        bool same = true;
        for (size_t i = 0; i < c.size(); ++i) same &= c[i] == conn[i];
        REQUIRE(same);
        auto x = H5::load<GF::amd::array<double>>(file, "/mesh/coor");
        same = x.shape() == coor.shape();
        for (size_t i = 0; same && i < x.size(); ++i) same &= x[i] == coor[i];
        REQUIRE(same);
        auto e = H5::load<GF::amd::array<double>>(file, "/series/Ekin");
        REQUIRE(e.size() == 5 && e[4] == 2.0);
        REQUIRE(H5::getShape(file, "/global/t") == std::vector<size_t>{5});
        REQUIRE(H5::getShape(file, "/meta/nelem").empty());
    }
    if (g_fail) {
        std::printf("%d check(s) failed\n", g_fail);
        return 1;
    }
    std::printf("ALL OK\n");
    return 0;
}
