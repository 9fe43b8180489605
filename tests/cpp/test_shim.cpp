// Drop-in check of the C++ face (include/GooseFEM/GooseFEM.h): the same calls and checks as the
// reference's C++ unit tests (tests/basic/ElementQuad4Planar.cpp, ElementHex8.cpp,
// MatrixDiagonal.cpp, Matrix.cpp, Vector.cpp, MeshQuad4.cpp, VectorPartitioned.cpp,
// MatrixDiagonalPartitioned.cpp, MatrixParitioned.cpp), the FineLayer meshes, the axisymmetric
// element and the examples' J2 material law, against the GPU path.
#include <GooseFEM/GooseFEM.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

static int g_fail = 0;
#define REQUIRE(cond)                                                                               \
    do {                                                                                            \
        if (!(cond)) {                                                                              \
            std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);                          \
            ++g_fail;                                                                               \
        }                                                                                           \
    } while (0)

template <class A>
static double maxabs(const A& a)
{
    double m = 0.0;
    for (size_t i = 0; i < a.size(); ++i) m = std::max(m, std::abs(a.data()[i]));
    return m;
}

int main()
{
    namespace GF = GooseFEM;
    // MeshQuad4.cpp: Regular(5, 3)
    {
        GF::Mesh::Quad4::Regular mesh(5, 3);
        auto conn = mesh.conn();
        REQUIRE(mesh.nelem() == 15 && mesh.nnode() == 24);
        REQUIRE(conn(0, 0) == 0 && conn(0, 1) == 1 && conn(0, 2) == 7 && conn(0, 3) == 6);
        REQUIRE(conn(14, 0) == 16 && conn(14, 2) == 23);
        auto per = mesh.nodesPeriodic();
        REQUIRE(per.shape(0) == 9 && per(1, 1) == 23 && per(8, 0) == 12 && per(8, 1) == 17);
        auto dp = mesh.dofsPeriodic();
        REQUIRE(dp(5, 0) == 0 && dp(11, 0) == 10 && dp(23, 1) == 1);
        REQUIRE(mesh.nodesTopRightCorner() == 23 && mesh.nodesBottomEdge().size() == 6);
    }
    // ElementQuad4Planar.cpp: dV, patch tests, zero periodic force
    {
        GF::Mesh::Quad4::Regular mesh(3, 3);
        GF::Vector vec(mesh.conn(), mesh.dofs());
        GF::Element::Quad4::QuadraturePlanar quad(vec.AsElement(mesh.coor()));
        REQUIRE(std::abs(maxabs(quad.dV()) - 0.25) < 1e-14);
        auto coor = mesh.coor();
        GF::array_type::tensor<double, 2> disp(coor.shape());
        for (size_t n = 0; n < mesh.nnode(); ++n) disp(n, 0) = 0.1 * coor(n, 1);
        auto ue = vec.AsElement(disp);
        auto gu = quad.GradN_vector(ue);
        auto eps = quad.SymGradN_vector(ue);
        bool ok = true;
        for (size_t e = 0; e < mesh.nelem(); ++e)
            for (size_t q = 0; q < quad.nip(); ++q) {
                ok &= std::abs(gu(e, q, 1, 0) - 0.1) < 1e-12 && std::abs(gu(e, q, 0, 1)) < 1e-12;
                ok &= std::abs(eps(e, q, 0, 1) - 0.05) < 1e-12 && std::abs(eps(e, q, 2, 2)) < 1e-14;
            }
        REQUIRE(ok);
        GF::Vector vp(mesh.conn(), mesh.dofsPeriodic());
        auto Fi = vp.AssembleDofs(quad.Int_gradN_dot_tensor2_dV(eps));
        REQUIRE(Fi.size() == vp.ndof() && maxabs(Fi) < 1e-12);
    }
    // MatrixDiagonal.cpp: lumped mass known answer and Dot / Solve round trip
    {
        GF::Mesh::Quad4::Regular mesh(2, 2);
        GF::Vector vector(mesh.conn(), mesh.dofs());
        GF::MatrixDiagonal A(mesh.conn(), mesh.dofs());
        GF::Element::Quad4::Quadrature quad(vector.AsElement(mesh.coor()), GF::Element::Quad4::Nodal::xi(),
                                            GF::Element::Quad4::Nodal::w());
        GF::array_type::tensor<double, 2> rho({mesh.nelem(), quad.nip()}, 1.0);
        A.assemble(quad.Int_N_scalar_NT_dV(rho));
        const double a[18] = {0.25, 0.25, 0.5, 0.5, 0.25, 0.25, 0.5, 0.5, 1.0,
                              1.0,  0.5,  0.5, 0.25, 0.25, 0.5, 0.5, 0.25, 0.25};
        auto d = A.data();
        bool ok = true;
        for (int i = 0; i < 18; ++i) ok &= std::abs(d[i] - a[i]) < 1e-12;
        REQUIRE(ok);
        std::mt19937 rng(1);
        std::uniform_real_distribution<double> U(0.5, 1.5);
        GF::array_type::tensor<double, 1> x({A.ndof()});
        for (auto& v : x) v = U(rng);
        auto b = A.Dot(x);
        auto X = A.Solve(b);
        ok = true;
        for (size_t i = 0; i < x.size(); ++i) ok &= std::abs(X[i] - x[i]) < 1e-12;
        REQUIRE(ok);
    }
    // Matrix.cpp: assemble random symmetric element matrices, Dot against the dense matrix
    {
        GF::Mesh::Quad4::Regular mesh(2, 2);
        GF::Matrix A(mesh.conn(), mesh.dofs());
        std::mt19937 rng(2);
        std::uniform_real_distribution<double> U(0.0, 1.0);
        GF::array_type::tensor<double, 3> a({mesh.nelem(), 8, 8});
        for (size_t e = 0; e < mesh.nelem(); ++e)
            for (size_t i = 0; i < 8; ++i)
                for (size_t j = 0; j <= i; ++j) a(e, i, j) = a(e, j, i) = U(rng);
        A.assemble(a);
        auto K = A.Todense();
        GF::array_type::tensor<double, 1> x({A.ndof()});
        for (auto& v : x) v = U(rng);
        auto b = A.Dot(x);
        bool ok = true;
        for (size_t i = 0; i < A.ndof(); ++i) {
            double s = 0.0;
            for (size_t j = 0; j < A.ndof(); ++j) s += K(i, j) * x[j];
            ok &= std::abs(s - b[i]) < 1e-12 && std::abs(K(i, (i + 3) % A.ndof()) - K((i + 3) % A.ndof(), i)) < 1e-14;
        }
        REQUIRE(ok);
    }
    // VectorPartitioned.cpp:40-95: tied Dirichlet mesh, nodevec -> partitioned -> nodevec round trip
    {
        GF::Mesh::Quad4::Regular mesh(9, 9);
        auto dofs = mesh.dofs();
        auto left = mesh.nodesLeftOpenEdge();
        auto right = mesh.nodesRightOpenEdge();
        for (size_t k = 0; k < left.size(); ++k)
            for (size_t i = 0; i < 2; ++i) dofs(right[k], i) = dofs(left[k], i);
        auto top = mesh.nodesTopEdge();
        auto bottom = mesh.nodesBottomEdge();
        GF::array_type::tensor<size_t, 1> iip({2 * (top.size() + bottom.size())});
        size_t k = 0;
        for (size_t i = 0; i < 2; ++i)
            for (size_t n = 0; n < top.size(); ++n) iip[k++] = dofs(top[n], i);
        for (size_t i = 0; i < 2; ++i)
            for (size_t n = 0; n < bottom.size(); ++n) iip[k++] = dofs(bottom[n], i);
        GF::VectorPartitioned vec(mesh.conn(), dofs, iip);
        GF::array_type::tensor<double, 2> u(vec.shape_nodevec());
        std::mt19937 rng(5);
        std::uniform_real_distribution<double> U(0.0, 1.0);
        for (auto& v : u) v = U(rng);
        for (size_t n = 0; n < left.size(); ++n)
            for (size_t i = 0; i < 2; ++i) u(right[n], i) = u(left[n], i);
        auto u_u = vec.AsDofs_u(u);
        auto u_p = vec.AsDofs_p(u);
        REQUIRE(u_u.size() == vec.nnu() && u_p.size() == vec.nnp() && vec.nnp() == iip.size());
        auto back = vec.NodeFromPartitioned(u_u, u_p);
        bool ok = true;
        for (size_t i = 0; i < u.size(); ++i) ok &= back[i] == u[i];
        REQUIRE(ok);
        GF::array_type::tensor<double, 2> zero(vec.shape_nodevec());
        auto c = vec.Copy_p(u, vec.Copy_u(u, zero));
        ok = true;
        for (size_t i = 0; i < u.size(); ++i) ok &= c[i] == u[i];
        REQUIRE(ok);
        // MatrixDiagonalPartitioned.cpp:55-94: set, Dot, Solve, Reaction
        GF::MatrixDiagonalPartitioned D(mesh.conn(), dofs, iip);
        GF::array_type::tensor<double, 1> a({vec.ndof()}), x({vec.ndof()});
        for (auto& v : a) v = 0.5 + U(rng);
        for (auto& v : x) v = U(rng);
        D.set(a);
        auto b = D.Dot(x);
        auto X = D.Solve(b, x);
        ok = true;
        for (size_t d = 0; d < vec.nnu(); ++d) ok &= std::abs(X[vec.iiu()[d]] - x[vec.iiu()[d]]) < 1e-14;
        REQUIRE(ok);
        auto R = D.Reaction(x, b);
        ok = true;
        for (size_t d = 0; d < vec.nnp(); ++d) ok &= R[vec.iip()[d]] == a[vec.iip()[d]] * x[vec.iip()[d]];
        REQUIRE(ok);
    }
    // MatrixParitioned.cpp:11-44 and :46-108: assemble, Todense, Dot, Solve (SPD matrices: the device
    // solve is a CG)
    {
        GF::Mesh::Quad4::Regular mesh(2, 2);
        GF::array_type::tensor<size_t, 1> iip({4});
        iip[0] = 0;
        iip[1] = 5;
        iip[2] = 7;
        iip[3] = 13;
        GF::MatrixPartitioned A(mesh.conn(), mesh.dofs(), iip);
        std::mt19937 rng(6);
        std::uniform_real_distribution<double> U(0.0, 1.0);
        GF::array_type::tensor<double, 3> a({mesh.nelem(), 8, 8});
        for (size_t e = 0; e < mesh.nelem(); ++e) {
            double r[8][8];
            for (auto& row : r)
                for (auto& v : row) v = U(rng);
            for (size_t i = 0; i < 8; ++i)
                for (size_t j = 0; j < 8; ++j) {
                    double s = i == j ? 8.0 : 0.0;
                    for (size_t k = 0; k < 8; ++k) s += r[i][k] * r[j][k];
                    a(e, i, j) = s;
                }
        }
        A.assemble(a);
        GF::array_type::tensor<double, 1> x({A.ndof()});
        for (auto& v : x) v = U(rng);
        auto b = A.Dot(x);
        auto K = A.Todense();
        bool ok = true;
        for (size_t i = 0; i < A.ndof(); ++i) {
            double s = 0.0;
            for (size_t j = 0; j < A.ndof(); ++j) s += K(i, j) * x[j];
            ok &= std::abs(s - b[i]) < 1e-12;
        }
        REQUIRE(ok);
        GF::MatrixPartitionedSolver solver;
        auto xc = solver.Solve(A, b, x);
        ok = true;
        for (size_t i = 0; i < x.size(); ++i) ok &= std::abs(xc[i] - x[i]) < 1e-9;
        REQUIRE(ok && solver.iterations() > 0 && solver.relres() < 1e-11);
        auto uu = A.data_uu();
        REQUIRE(uu.rows == A.nnu() && uu.cols == A.nnu() && uu.rowptr.back() == static_cast<int64_t>(uu.val.size()));
        auto R = A.Reaction(x, b);
        ok = true;
        for (size_t d = 0; d < A.nnp(); ++d) ok &= std::abs(R[A.iip()[d]] - b[A.iip()[d]]) < 1e-12;
        REQUIRE(ok);
    }
    // MeshQuad4.cpp:103-256: FineLayer(9, 18) sizes, middle layer, periodic pairs
    {
        GF::Mesh::Quad4::FineLayer mesh(9, 18);
        REQUIRE(mesh.nelem() == 81 && mesh.nnode() == 96 && mesh.nelx() == 9 && mesh.nely() == 21);
        auto mid = mesh.elementsMiddleLayer();
        REQUIRE(mid.size() == 9 && mid[0] == 36 && mid[8] == 44);
        auto per = mesh.nodesPeriodic();
        REQUIRE(per.shape(0) == 15 && per(1, 1) == 95 && per(14, 0) == 88 && per(14, 1) == 91);
        auto conn = mesh.conn();
        REQUIRE(conn(7, 1) == 21 && conn(63, 3) == 84);
    }
    // MeshHex8.cpp:192-377: FineLayer(3, 3, 3) sizes, connectivity, node sets and periodic pairs;
    // a mesh with transition layers keeps nelx, nelz and the middle layer's size
    {
        GF::Mesh::Hex8::FineLayer mesh(3, 3, 3);
        REQUIRE(mesh.nelem() == 27 && mesh.nnode() == 64 && mesh.nely() == 3);
        auto conn = mesh.conn();
        REQUIRE(conn(26, 0) == 42 && conn(26, 3) == 58 && conn(26, 7) == 62);
        auto bot = mesh.nodesBottom();
        REQUIRE(bot.size() == 16 && bot[1] == 4 && bot[4] == 1);
        auto face = mesh.nodeset("nodesBottomFace");
        REQUIRE(face.size() == 4 && face[0] == 5 && face[1] == 9 && face[2] == 6);
        auto per = mesh.nodesPeriodic();
        REQUIRE(per.shape(0) == 37 && per(0, 1) == 3 && per(36, 0) == 10 && per(36, 1) == 58);
        REQUIRE(mesh.dofsPeriodic()(63, 2) == 2 && mesh.nodesOrigin() == 0);
        GF::Mesh::Hex8::FineLayer fine(9, 17, 27);
        REQUIRE(fine.nelx() == 9 && fine.nelz() == 27 && fine.elementsMiddleLayer().size() == 243);
    }
    // QuadratureAxisymmetric: the volume of a hollow cylinder, and the material law's consistency
    {
        GF::Mesh::Quad4::Regular mesh(4, 3);
        auto coor = mesh.coor();
        for (size_t n = 0; n < mesh.nnode(); ++n) {
            coor(n, 0) = 0.5 * coor(n, 0);       // z in [0, 2]
            coor(n, 1) = 0.5 + coor(n, 1) / 3.0; // r in [0.5, 1.5]
        }
        GF::Vector vec(mesh.conn(), mesh.dofs());
        GF::Element::Quad4::QuadratureAxisymmetric quad(vec.AsElement(coor));
        auto dV = quad.dV();
        double V = 0.0;
        for (size_t i = 0; i < dV.size(); ++i) V += dV[i];
        REQUIRE(std::abs(V - M_PI * (1.5 * 1.5 - 0.25) * 2.0) < 1e-12);
        GF::array_type::tensor<double, 2> K({quad.nelem(), quad.nip()}, 10.0), G({quad.nelem(), quad.nip()}, 1.0),
            sy({quad.nelem(), quad.nip()}, 0.01), H({quad.nelem(), quad.nip()}, 0.5);
        GF::material::LinearHardening mat(K, G, sy, H);
        GF::array_type::tensor<double, 2> u(vec.shape_nodevec());
        for (size_t n = 0; n < mesh.nnode(); ++n) u(n, 1) = 0.02 * coor(n, 1); // radial expansion
        quad.symGradN_vector(vec.AsElement(u), mat.Eps());
        mat.refresh();
        auto ep = mat.epsp();
        auto sig = mat.Sig().to_host();
        bool ok = true;
        for (size_t i = 0; i < ep.size(); ++i) {
            const double* s9 = sig.data() + 9 * i;
            const double m = (s9[0] + s9[4] + s9[8]) / 3.0;
            double ss = 0.0;
            for (int k = 0; k < 9; ++k) {
                const double d = s9[k] - (k % 4 == 0 ? m : 0.0);
                ss += d * d;
            }
            ok &= ep[i] > 0.0 && std::abs(std::sqrt(1.5 * ss) - (0.01 + 0.5 * ep[i])) < 1e-12;
        }
        REQUIRE(ok);
    }
    // ElementHex8.cpp: dV and periodic lumped mass of ones; BF16x3 stiffness close to FP64
    {
        GF::Mesh::Hex8::Regular mesh(3, 3, 3);
        GF::Vector vec(mesh.conn(), mesh.dofsPeriodic());
        GF::MatrixDiagonal M(mesh.conn(), mesh.dofsPeriodic());
        GF::Element::Hex8::Quadrature nodal(vec.AsElement(mesh.coor()), GF::Element::Hex8::Nodal::xi(),
                                            GF::Element::Hex8::Nodal::w());
        GF::array_type::tensor<double, 2> rho({mesh.nelem(), nodal.nip()}, 1.0);
        M.assemble(nodal.Int_N_scalar_NT_dV(rho));
        auto m = M.data();
        bool ok = m.size() == vec.ndof();
        for (size_t i = 0; i < m.size(); ++i) ok &= std::abs(m[i] - 1.0) < 1e-12;
        REQUIRE(ok);
        GF::Element::Hex8::Quadrature quad(vec.AsElement(mesh.coor()));
        REQUIRE(std::abs(maxabs(quad.dV()) - 0.125) < 1e-14);
        GF::array_type::tensor<double, 6> C(quad.shape_qtensor4());
        for (size_t e = 0; e < mesh.nelem(); ++e)
            for (size_t q = 0; q < 8; ++q)
                for (size_t i = 0; i < 3; ++i)
                    for (size_t j = 0; j < 3; ++j)
                        for (size_t k = 0; k < 3; ++k)
                            for (size_t l = 0; l < 3; ++l)
                                C(e, q, i, j, k, l) = 10.0 * (i == j) * (k == l) +
                                                      2.0 * ((0.5 * ((i == k) * (j == l) + (i == l) * (j == k))) -
                                                             (i == j) * (k == l) / 3.0);
        auto K64 = quad.Int_gradN_dot_tensor4_dot_gradNT_dV(C, GF_F64);
        auto Kbf = quad.Int_gradN_dot_tensor4_dot_gradNT_dV(C, GF_BF16X3);
        double num = 0.0, den = 0.0;
        for (size_t i = 0; i < K64.size(); ++i) {
            num += (Kbf[i] - K64[i]) * (Kbf[i] - K64[i]);
            den += K64[i] * K64[i];
        }
        REQUIRE(std::sqrt(num / den) < 2e-5);
    }
    // device-resident path: no host round trips between calls
    {
        GF::Mesh::Quad4::Regular mesh(16, 16);
        GF::Vector vec(mesh.conn(), mesh.dofs());
        GF::Element::Quad4::QuadraturePlanar quad(vec.AsElement(mesh.coor()));
        GF::array_type::tensor<double, 2> u(vec.shape_nodevec());
        std::mt19937 rng(3);
        std::normal_distribution<double> N(0.0, 1e-3);
        for (auto& v : u) v = N(rng);
        auto du = GF::amd::device_array<double>::from_host(u);
        GF::amd::device_array<double> ue(vec.shape_elemvec()), eps(quad.shape_qtensor2()), fe(vec.shape_elemvec()),
            f(vec.shape_nodevec());
        vec.asElement(du, ue);
        quad.symGradN_vector(ue, eps);
        quad.int_gradN_dot_tensor2_dV(eps, fe);
        vec.assembleNode(fe, f);
        auto fh = f.to_host();
        auto ref = vec.AssembleNode(quad.Int_gradN_dot_tensor2_dV(quad.SymGradN_vector(vec.AsElement(u))));
        bool ok = true;
        for (size_t i = 0; i < ref.size(); ++i) ok &= fh[i] == ref[i];
        REQUIRE(ok);
    }
    // Tyings::Control + Tyings::Periodic as in docs/examples/statics/periodic_elastoplastic.py
    {
        GF::Mesh::Quad4::Regular mesh(5, 5);
        GF::Tyings::Control control(mesh.coor(), mesh.dofs());
        auto cd = control.controlDofs();
        GF::array_type::tensor<size_t, 1> iip({6});
        for (size_t i = 0; i < 4; ++i) iip[i] = cd[i];
        iip[4] = control.dofs()(mesh.nodesOrigin(), 0);
        iip[5] = control.dofs()(mesh.nodesOrigin(), 1);
        GF::Tyings::Periodic tyings(control.coor(), control.dofs(), cd, mesh.nodesPeriodic(), iip);
        REQUIRE(tyings.nnp() == 6 && tyings.nnd() == 2 * mesh.nodesPeriodic().shape(0));
        REQUIRE(tyings.nnu() + tyings.nnp() == tyings.nni());
        auto Cdu = tyings.Cdu();
        REQUIRE(Cdu.rows == tyings.nnd() && Cdu.cols == tyings.nnu());
    }
    // MatrixPartitionedTyings: a homogeneous elastic periodic cell sheared through a control DOF
    // deforms affinely (zero fluctuation), so the tied solve must return u = A x + c at every node
    {
        GF::Mesh::Quad4::Regular mesh(6, 6);
        GF::Tyings::Control control(mesh.coor(), mesh.dofs());
        auto cd = control.controlDofs();
        auto cn = control.controlNodes();
        auto dofs0 = control.dofs();
        GF::array_type::tensor<size_t, 1> iip({6});
        for (size_t i = 0; i < 4; ++i) iip[i] = cd[i];
        iip[4] = dofs0(mesh.nodesOrigin(), 0);
        iip[5] = dofs0(mesh.nodesOrigin(), 1);
        GF::Tyings::Periodic tyings(control.coor(), dofs0, cd, mesh.nodesPeriodic(), iip);
        auto dofs = tyings.dofs();
        auto coor = control.coor();
        GF::Vector vec(mesh.conn(), dofs);
        GF::Element::Quad4::QuadraturePlanar quad(vec.AsElement(coor));
        GF::array_type::tensor<double, 6> C(quad.shape_qtensor4());
        for (size_t e = 0; e < quad.nelem(); ++e)
            for (size_t q = 0; q < quad.nip(); ++q)
                for (size_t i = 0; i < 3; ++i)
                    for (size_t j = 0; j < 3; ++j)
                        for (size_t k = 0; k < 3; ++k)
                            for (size_t l = 0; l < 3; ++l)
                                C(e, q, i, j, k, l) = 10.0 * (i == j) * (k == l) +
                                                      2.0 * ((0.5 * ((i == k) * (j == l) + (i == l) * (j == k))) -
                                                             (i == j) * (k == l) / 3.0);
        GF::MatrixPartitionedTyings K(mesh.conn(), dofs, tyings.Cdu(), tyings.Cdp());
        K.assemble(quad.Int_gradN_dot_tensor4_dot_gradNT_dV(C, GF_F64));
        GF::array_type::tensor<double, 2> b(vec.shape_nodevec()), u(vec.shape_nodevec());
        u(cn[0], 1) = 0.01;
        GF::MatrixPartitionedTyingsSolver solver;
        solver.solve(K, b, u);
        REQUIRE(solver.relres() < 1e-11);
        // least-squares affine fit per component over the mesh nodes (not the control nodes)
        bool ok = true;
        for (size_t c = 0; c < 2; ++c) {
            double S[3][3] = {{0}}, R[3] = {0};
            for (size_t n = 0; n < mesh.nnode(); ++n) {
                const double f[3] = {coor(n, 0), coor(n, 1), 1.0};
                for (int a = 0; a < 3; ++a) {
                    R[a] += f[a] * u(n, c);
                    for (int bb = 0; bb < 3; ++bb) S[a][bb] += f[a] * f[bb];
                }
            }
            // solve the 3 x 3 normal equations by Cramer's rule
            auto det3 = [](double M[3][3]) {
                return M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                       M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
            };
            double coef[3];
            const double D = det3(S);
            for (int a = 0; a < 3; ++a) {
                double M[3][3];
                for (int r = 0; r < 3; ++r)
                    for (int cc = 0; cc < 3; ++cc) M[r][cc] = cc == a ? R[r] : S[r][cc];
                coef[a] = det3(M) / D;
            }
            for (size_t n = 0; n < mesh.nnode(); ++n)
                ok &= std::abs(coef[0] * coor(n, 0) + coef[1] * coor(n, 1) + coef[2] - u(n, c)) < 1e-10;
        }
        REQUIRE(ok);
        double umax = 0.0;
        for (size_t n = 0; n < mesh.nnode(); ++n) umax = std::max(umax, std::abs(u(n, 0)) + std::abs(u(n, 1)));
        REQUIRE(umax > 1e-3); // the cell did deform
    }
    if (g_fail) {
        std::printf("%d check(s) failed\n", g_fail);
        return 1;
    }
    std::printf("ALL OK\n");
    return 0;
}
