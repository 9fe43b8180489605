// MatrixDiagonal (include/GooseFEM/MatrixDiagonal.h) and the sparse Matrix
// (include/GooseFEM/Matrix.h:653-860) on the device.
//
// Matrix assembly: the reference pushes triplets in (e, m, i, n, j) order and lets
// Eigen::SparseMatrix::setFromTriplets sum duplicates in that order (Matrix.h:704-727).  Here each
// CSR row d is owned by a group of lanes: the row's contributions (e, m, i) are visited in
// ascending order (the DOF-centric list of the Vector), and for each contribution the lanes add
// the element-matrix row Ke[e, m*nd+i, :] into the row's accumulators (in LDS) at the positions of
// their columns.  Every CSR value therefore receives its element contributions in triplet order:
// FP64 bit-identical to the reference's duplicate sum, with no atomics.  The column positions
// are found once, when the pattern is built (binary search of each element column in the row's
// sorted column list), and kept as one byte per element-matrix entry (rows of at most 256
// entries), so an assembly only streams the element matrices, the byte map and the values.
// SpMV: lanes form the products val*x in parallel, one lane sums them in ascending column order
// (bit-identical to a sequential row loop).
#include <algorithm>
#include <mutex>

#include "gf_internal.hpp"

using namespace gf;

// Staging of gf_matrix_dot_nodevec_noalloc: the dofval copies of x and of A x (2 x at most 2^20
// doubles, 16 MiB).  Every device the library runs on holds its own copy for the life of the
// process.  In host code `scratch` is a host-side stand-in, not a device address: the device
// address comes from hipGetSymbolAddress (scratch_ptr below).
constexpr int64_t NOALLOC_MAX_DOFS = int64_t(1) << 20;
static __device__ __align__(256) double scratch[2 * NOALLOC_MAX_DOFS];

struct gf_diag_s {
    gf_vector v = nullptr;
    DBuf<double> A, inv;
    bool changed = true;
};

struct gf_matrix_s {
    gf_vector v = nullptr;
    int64_t nrows = 0, nnz = 0;
    int maxrow = 0;
    bool degenerate = false; // an element holds the same DOF twice
    bool assembled = false;
    DBuf<int32_t> rowptr, col;
    DBuf<double> val;
    DBuf<uint8_t> posmap; // [contribution c][column s]: position of the column in row c's CSR row
};

namespace {

constexpr int BS = 256;

__global__ void k_diag_assemble(int64_t ndof, int K, const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                                const double* __restrict__ elemmat, double* __restrict__ A)
{
    int64_t d = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (d >= ndof) return;
    double s = 0.0;
    for (int32_t j = ptr[d]; j < ptr[d + 1]; ++j) {
        int64_t k = idx[j];
        int64_t e = k / K;
        int64_t r = k - e * K;
        s += elemmat[e * K * K + r * K + r];
    }
    A[d] = s;
}

__global__ void k_inverse(int64_t n, const double* __restrict__ A, double* __restrict__ inv)
{
    int64_t d = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (d < n) inv[d] = 1.0 / A[d];
}

__global__ void k_mul(int64_t n, const double* __restrict__ a, const double* __restrict__ x, double* __restrict__ b)
{
    int64_t d = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (d < n) b[d] = a[d] * x[d];
}

__global__ void k_mul_nodevec(int64_t n, const int32_t* __restrict__ dofs, const double* __restrict__ a,
                              const double* __restrict__ x, double* __restrict__ b)
{
    int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k < n) b[k] = a[dofs[k]] * x[k];
}

// binary search of c in sorted cols[0..n)
__device__ inline int find_col(const int32_t* cols, int n, int32_t c)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (cols[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Column positions, once per pattern: posmap[c * K + s] is the position, in CSR row d, of column
// s of contribution c (the c-th entry of the DOF-element list, c in [cptr[d], cptr[d + 1])).
// L lanes per row, CAP >= the longest row (<= 256, so a position fits one byte).
template <int L, int CAP>
__global__ void __launch_bounds__(BS) k_matrix_posmap(int64_t nrows, int nne, int nd, const int32_t* __restrict__ conn,
                                                      const int32_t* __restrict__ dofs,
                                                      const int32_t* __restrict__ cptr,
                                                      const int32_t* __restrict__ cidx,
                                                      const int32_t* __restrict__ rowptr,
                                                      const int32_t* __restrict__ col, uint8_t* __restrict__ posmap)
{
    constexpr int R = BS / L;
    __shared__ int32_t s_col[R][CAP];
    const int g = threadIdx.x / L, lane = threadIdx.x % L;
    const int64_t row = static_cast<int64_t>(blockIdx.x) * R + g;
    const bool active = row < nrows;
    int n = 0;
    if (active) {
        const int b = rowptr[row];
        n = rowptr[row + 1] - b;
        for (int p = lane; p < n; p += L) s_col[g][p] = col[b + p];
    }
    __syncthreads();
    if (!active) return;
    const int K = nne * nd;
    for (int32_t c = cptr[row]; c < cptr[row + 1]; ++c) {
        const int64_t e = cidx[c] / K;
        for (int s = lane; s < K; s += L) {
            const int32_t cc = dofs[static_cast<int64_t>(conn[e * nne + s / nd]) * nd + s % nd];
            posmap[static_cast<int64_t>(c) * K + s] = static_cast<uint8_t>(find_col(s_col[g], n, cc));
        }
    }
}

// Assembly from the position map.  L lanes per row (K <= L), CAP >= the longest row.  Lane s adds
// column s of each contribution; the values and positions of up to 8 contributions are loaded
// together, then added to the row's LDS accumulators strictly in contribution order.  Loads and
// adds are unconditional: indices past the row's end are clamped to its last contribution and
// their values go to a spare accumulator (slot CAP) that is never stored.  Index products are
// 32 x 32 -> 64-bit: contribution indices and elemvec positions are below 2^31.
template <int L, int CAP, class TE>
__global__ void __launch_bounds__(BS) k_matrix_assemble(int64_t nrows, int K, const int32_t* __restrict__ cptr,
                                                        const int32_t* __restrict__ cidx,
                                                        const uint8_t* __restrict__ posmap,
                                                        const int32_t* __restrict__ rowptr,
                                                        const TE* __restrict__ elemmat, double* __restrict__ val)
{
    constexpr int R = BS / L;
    __shared__ double s_acc[R][CAP + 1];
    const int g = threadIdx.x / L, lane = threadIdx.x % L;
    const int64_t row = static_cast<int64_t>(blockIdx.x) * R + g;
    const bool active = row < nrows;
    int n = 0, b = 0;
    if (active) {
        b = rowptr[row];
        n = rowptr[row + 1] - b;
        for (int p = lane; p < n; p += L) s_acc[g][p] = 0.0;
    }
    __syncthreads();
    if (active && lane < K) {
        const uint32_t uK = static_cast<uint32_t>(K);
        const int32_t cend = cptr[row + 1];
        for (int32_t c0 = cptr[row]; c0 < cend; c0 += 8) {
            uint32_t cc[8], kk[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) cc[j] = static_cast<uint32_t>(min(c0 + j, cend - 1));
#pragma unroll
            for (int j = 0; j < 8; ++j) kk[j] = static_cast<uint32_t>(cidx[cc[j]]);
            int pos[8];
            double pv[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                pos[j] = c0 + j < cend ? static_cast<int>(posmap[static_cast<uint64_t>(cc[j]) * uK + lane]) : CAP;
                pv[j] = static_cast<double>(elemmat[static_cast<uint64_t>(kk[j]) * uK + lane]);
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                s_acc[g][pos[j]] += pv[j];
                // contributions of one row are added strictly in order: make this contribution's
                // LDS writes visible to the next one's reads across the row's lanes
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            }
        }
    }
    __syncthreads();
    if (active)
        for (int p = lane; p < n; p += L) val[b + p] = s_acc[g][p];
}

// Fallback: one thread per row, sequential (handles degenerate elements and long rows).
template <class TE>
__global__ void k_matrix_assemble_serial(int64_t nrows, int nne, int nd, const int32_t* __restrict__ conn,
                                         const int32_t* __restrict__ dofs, const int32_t* __restrict__ cptr,
                                         const int32_t* __restrict__ cidx, const int32_t* __restrict__ rowptr,
                                         const int32_t* __restrict__ col, const TE* __restrict__ elemmat,
                                         double* __restrict__ val)
{
    int64_t row = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (row >= nrows) return;
    const int b = rowptr[row], n = rowptr[row + 1] - b;
    for (int p = 0; p < n; ++p) val[b + p] = 0.0;
    const int K = nne * nd;
    for (int32_t c = cptr[row]; c < cptr[row + 1]; ++c) {
        const int64_t k = cidx[c];
        const int64_t e = k / K;
        const int r = static_cast<int>(k - e * K);
        for (int s = 0; s < K; ++s) {
            const int32_t cc = dofs[static_cast<int64_t>(conn[e * nne + s / nd]) * nd + s % nd];
            const int p = find_col(col + b, n, cc);
            val[b + p] += static_cast<double>(elemmat[(e * K + r) * K + s]);
        }
    }
}

// SpMV, L lanes per row: products in parallel, ordered sum by lane 0.
template <int L, int CAP>
__global__ void __launch_bounds__(BS) k_spmv(int64_t nrows, const int32_t* __restrict__ rowptr,
                                             const int32_t* __restrict__ col, const double* __restrict__ val,
                                             const double* __restrict__ x, double* __restrict__ y)
{
    constexpr int R = BS / L;
    __shared__ double s_p[R][CAP];
    const int g = threadIdx.x / L, lane = threadIdx.x % L;
    const int64_t row = static_cast<int64_t>(blockIdx.x) * R + g;
    int b = 0, n = 0;
    if (row < nrows) {
        b = rowptr[row];
        n = rowptr[row + 1] - b;
        if (n > 0) {
            // all of the lane's loads are issued together (indices past the row's end are clamped
            // to its last entry and not stored): the columns first, then the values and x
            constexpr int NI = CAP / L;
            int32_t cc[NI];
#pragma unroll
            for (int i = 0; i < NI; ++i) cc[i] = col[b + min(lane + i * L, n - 1)];
            double vv[NI], xv[NI];
#pragma unroll
            for (int i = 0; i < NI; ++i) {
                vv[i] = val[b + min(lane + i * L, n - 1)];
                xv[i] = x[cc[i]];
            }
#pragma unroll
            for (int i = 0; i < NI; ++i)
                if (lane + i * L < n) s_p[g][lane + i * L] = vv[i] * xv[i];
        }
    }
    __syncthreads();
    if (row < nrows && lane == 0) {
        double s = 0.0;
        for (int p = 0; p < n; ++p) s += s_p[g][p];
        y[row] = s;
    }
}

__global__ void k_spmv_serial(int64_t nrows, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                              const double* __restrict__ val, const double* __restrict__ x, double* __restrict__ y)
{
    int64_t row = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (row >= nrows) return;
    double s = 0.0;
    for (int32_t k = rowptr[row]; k < rowptr[row + 1]; ++k) s += val[k] * x[col[k]];
    y[row] = s;
}

__global__ void k_todense(int64_t nrows, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                          const double* __restrict__ val, double* __restrict__ out)
{
    int64_t row = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (row >= nrows) return;
    for (int32_t k = rowptr[row]; k < rowptr[row + 1]; ++k) out[row * nrows + col[k]] = val[k];
}

__global__ void k_fill(int64_t n, double v, double* __restrict__ a)
{
    int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i < n) a[i] = v;
}

void spmv(gf_matrix_s* A, const double* x, double* y, hipStream_t s)
{
    if (A->nrows == 0) return;
    if (A->maxrow <= 32)
        k_spmv<8, 32><<<nblocks(A->nrows, BS / 8), BS, 0, s>>>(A->nrows, A->rowptr.p, A->col.p, A->val.p, x, y);
    else if (A->maxrow <= 88) // Hex8 rows of up to 81 entries: 11 per lane, 8 rows per wave
        k_spmv<8, 88><<<nblocks(A->nrows, BS / 8), BS, 0, s>>>(A->nrows, A->rowptr.p, A->col.p, A->val.p, x, y);
    else if (A->maxrow <= 96)
        k_spmv<16, 96><<<nblocks(A->nrows, BS / 16), BS, 0, s>>>(A->nrows, A->rowptr.p, A->col.p, A->val.p, x, y);
    else if (A->maxrow <= 128)
        k_spmv<16, 128><<<nblocks(A->nrows, BS / 16), BS, 0, s>>>(A->nrows, A->rowptr.p, A->col.p, A->val.p, x, y);
    else k_spmv_serial<<<nblocks(A->nrows, BS), BS, 0, s>>>(A->nrows, A->rowptr.p, A->col.p, A->val.p, x, y);
    check_launch();
}

} // namespace

// exported for the solver translation units
namespace gf {
void matrix_spmv(gf_matrix A, const double* x, double* y, hipStream_t s) { spmv(A, x, y, s); }
} // namespace gf

namespace {

template <class TE>
void matrix_assemble(gf_matrix A, const TE* elemmat, hipStream_t s)
{
    auto v = A->v;
    const int K = v->nne * v->ndim;
    if (A->nrows) {
        if (A->posmap.n && K <= 8 && A->maxrow <= 64)
            k_matrix_assemble<8, 64, TE><<<nblocks(A->nrows, BS / 8), BS, 0, s>>>(
                A->nrows, K, v->dof_eptr.p, v->dof_eidx.p, A->posmap.p, A->rowptr.p, elemmat, A->val.p);
        else if (A->posmap.n && A->maxrow <= 128)
            k_matrix_assemble<32, 128, TE><<<nblocks(A->nrows, BS / 32), BS, 0, s>>>(
                A->nrows, K, v->dof_eptr.p, v->dof_eidx.p, A->posmap.p, A->rowptr.p, elemmat, A->val.p);
        else if (A->posmap.n)
            k_matrix_assemble<32, 256, TE><<<nblocks(A->nrows, BS / 32), BS, 0, s>>>(
                A->nrows, K, v->dof_eptr.p, v->dof_eidx.p, A->posmap.p, A->rowptr.p, elemmat, A->val.p);
        else
            k_matrix_assemble_serial<TE><<<nblocks(A->nrows, BS), BS, 0, s>>>(
                A->nrows, v->nne, v->ndim, v->conn.p, v->dofs.p, v->dof_eptr.p, v->dof_eidx.p, A->rowptr.p,
                A->col.p, elemmat, A->val.p);
        check_launch();
    }
    A->assembled = true;
}

} // namespace

extern "C" {

// ---- MatrixDiagonal ----------------------------------------------------------------------------

int gf_diag_create(gf_diag* out, gf_vector v)
{
    return guard([&] {
        GF_REQUIRE(out && v, "null argument");
        auto d = new gf_diag_s();
        try {
            d->v = v;
            d->A.alloc(v->ndof);
            d->inv.alloc(v->ndof);
        }
        catch (...) {
            delete d;
            throw;
        }
        *out = d;
    });
}

int gf_diag_destroy(gf_diag d)
{
    delete d;
    return GF_OK;
}

int gf_diag_assemble(gf_diag d, const double* elemmat, void* stream)
{
    return guard([&] {
        GF_REQUIRE(d && elemmat, "null argument");
        auto v = d->v;
        if (v->ndof)
            k_diag_assemble<<<nblocks(v->ndof, BS), BS, 0, S(stream)>>>(v->ndof, v->nne * v->ndim, v->dof_eptr.p,
                                                                      v->dof_eidx.p, elemmat, d->A.p);
        check_launch();
        d->changed = true;
    });
}

int gf_diag_set(gf_diag d, const double* A, void* stream)
{
    return guard([&] {
        GF_REQUIRE(d && A, "null argument");
        if (d->A.n)
            GF_HIP(hipMemcpyAsync(d->A.p, A, d->A.n * sizeof(double), hipMemcpyDeviceToDevice, S(stream)));
        d->changed = true;
    });
}

int gf_diag_data(gf_diag d, const double** A)
{
    return guard([&] {
        GF_REQUIRE(d && A, "null argument");
        *A = d->A.p;
    });
}

static void factorize(gf_diag d, hipStream_t s)
{
    if (!d->changed) return;
    int64_t n = d->v->ndof;
    if (n) k_inverse<<<nblocks(n, BS), BS, 0, s>>>(n, d->A.p, d->inv.p);
    check_launch();
    d->changed = false;
}

int gf_diag_inverse(gf_diag d, const double** inv, void* stream)
{
    return guard([&] {
        GF_REQUIRE(d && inv, "null argument");
        factorize(d, S(stream));
        *inv = d->inv.p;
    });
}

int gf_diag_dot_dofval(gf_diag d, const double* x, double* b, void* stream)
{
    return guard([&] {
        GF_REQUIRE(d && x && b, "null argument");
        int64_t n = d->v->ndof;
        if (n) k_mul<<<nblocks(n, BS), BS, 0, S(stream)>>>(n, d->A.p, x, b);
        check_launch();
    });
}

int gf_diag_dot_nodevec(gf_diag d, const double* x, double* b, void* stream)
{
    return guard([&] {
        GF_REQUIRE(d && x && b, "null argument");
        int64_t n = d->v->nnode * d->v->ndim;
        if (n) k_mul_nodevec<<<nblocks(n, BS), BS, 0, S(stream)>>>(n, d->v->dofs.p, d->A.p, x, b);
        check_launch();
    });
}

int gf_diag_solve_dofval(gf_diag d, const double* b, double* x, void* stream)
{
    return guard([&] {
        GF_REQUIRE(d && x && b, "null argument");
        factorize(d, S(stream));
        int64_t n = d->v->ndof;
        if (n) k_mul<<<nblocks(n, BS), BS, 0, S(stream)>>>(n, d->inv.p, b, x);
        check_launch();
    });
}

int gf_diag_solve_nodevec(gf_diag d, const double* b, double* x, void* stream)
{
    return guard([&] {
        GF_REQUIRE(d && x && b, "null argument");
        factorize(d, S(stream));
        int64_t n = d->v->nnode * d->v->ndim;
        if (n) k_mul_nodevec<<<nblocks(n, BS), BS, 0, S(stream)>>>(n, d->v->dofs.p, d->inv.p, b, x);
        check_launch();
    });
}

// ---- sparse Matrix -------------------------------------------------------------------------------

int gf_matrix_create(gf_matrix* out, gf_vector v)
{
    return guard([&] {
        GF_REQUIRE(out && v, "null argument");
        auto A = new gf_matrix_s();
        try {
            A->v = v;
            const int nne = v->nne, nd = v->ndim, K = nne * nd;
            const int64_t ndof = v->ndof, nelem = v->nelem;
            const auto& conn = v->conn_h;
            const auto& dofs = v->dofs_h;
            // elements touching each DOF (ascending, unique)
            std::vector<int64_t> eptr(ndof + 1, 0);
            for (int64_t e = 0; e < nelem; ++e)
                for (int s = 0; s < K; ++s) eptr[dofs[conn[e * nne + s / nd] * nd + s % nd] + 1]++;
            for (int64_t d = 0; d < ndof; ++d) eptr[d + 1] += eptr[d];
            std::vector<int64_t> eidx(eptr[ndof]), fill(eptr.begin(), eptr.end() - 1);
            for (int64_t e = 0; e < nelem; ++e) {
                int64_t dd[64];
                for (int s = 0; s < K; ++s) dd[s] = dofs[conn[e * nne + s / nd] * nd + s % nd];
                for (int s = 0; s < K; ++s) {
                    bool dup = false;
                    for (int t = 0; t < s; ++t) dup |= dd[t] == dd[s];
                    if (dup) A->degenerate = true;
                    eidx[fill[dd[s]]++] = e;
                }
            }
            // CSR pattern: per row, sorted unique columns of all elements touching the row DOF
            std::vector<int32_t> rowptr(ndof + 1, 0);
            std::vector<std::vector<int32_t>> cols(ndof);
            parallel_for(ndof, [&](int64_t d) {
                auto& c = cols[d];
                for (int64_t j = eptr[d]; j < eptr[d + 1]; ++j) {
                    int64_t e = eidx[j];
                    for (int s = 0; s < K; ++s) c.push_back(static_cast<int32_t>(dofs[conn[e * nne + s / nd] * nd + s % nd]));
                }
                std::sort(c.begin(), c.end());
                c.erase(std::unique(c.begin(), c.end()), c.end());
            });
            int64_t nnz = 0;
            int maxrow = 0;
            for (int64_t d = 0; d < ndof; ++d) {
                nnz += static_cast<int64_t>(cols[d].size());
                maxrow = std::max(maxrow, static_cast<int>(cols[d].size()));
                GF_REQUIRE(nnz < INT32_MAX, "nnz exceeds int32");
                rowptr[d + 1] = static_cast<int32_t>(nnz);
            }
            std::vector<int32_t> colv(nnz);
            parallel_for(ndof, [&](int64_t d) { std::copy(cols[d].begin(), cols[d].end(), colv.begin() + rowptr[d]); });
            A->nrows = ndof;
            A->nnz = nnz;
            A->maxrow = maxrow;
            A->rowptr.upload(rowptr);
            A->col.upload(colv);
            A->val.alloc(nnz);
            if (!A->degenerate && K <= 32 && maxrow <= 256 && ndof > 0) {
                A->posmap.alloc(v->dof_eidx.n * static_cast<size_t>(K));
                k_matrix_posmap<32, 256><<<nblocks(ndof, BS / 32), BS, 0, nullptr>>>(
                    ndof, nne, nd, v->conn.p, v->dofs.p, v->dof_eptr.p, v->dof_eidx.p, A->rowptr.p, A->col.p,
                    A->posmap.p);
                check_launch();
                GF_HIP(hipStreamSynchronize(nullptr));
            }
        }
        catch (...) {
            delete A;
            throw;
        }
        *out = A;
    });
}

int gf_matrix_destroy(gf_matrix A)
{
    delete A;
    return GF_OK;
}

int gf_matrix_sizes(gf_matrix A, int64_t* nrows, int64_t* nnz)
{
    return guard([&] {
        GF_REQUIRE(A, "null matrix");
        if (nrows) *nrows = A->nrows;
        if (nnz) *nnz = A->nnz;
    });
}

int gf_matrix_assemble(gf_matrix A, const double* elemmat, void* stream)
{
    return guard([&] {
        GF_REQUIRE(A && elemmat, "null argument");
        matrix_assemble(A, elemmat, S(stream));
    });
}

int gf_matrix_assemble_f32(gf_matrix A, const float* elemmat, void* stream)
{
    return guard([&] {
        GF_REQUIRE(A && elemmat, "null argument");
        matrix_assemble(A, elemmat, S(stream));
    });
}

int gf_matrix_csr(gf_matrix A, const int32_t** rowptr, const int32_t** col, const double** val)
{
    return guard([&] {
        GF_REQUIRE(A, "null matrix");
        if (rowptr) *rowptr = A->rowptr.p;
        if (col) *col = A->col.p;
        if (val) *val = A->val.p;
    });
}

int gf_matrix_dot_dofval(gf_matrix A, const double* x, double* b, void* stream)
{
    return guard([&] {
        GF_REQUIRE(A && x && b, "null argument");
        if (!A->assembled) throw Error(GF_ESTATE, "matrix not assembled");
        spmv(A, x, b, S(stream));
    });
}

int gf_matrix_dot_nodevec(gf_matrix A, const double* x, double* b, void* stream)
{
    return guard([&] {
        GF_REQUIRE(A && x && b, "null argument");
        if (!A->assembled) throw Error(GF_ESTATE, "matrix not assembled");
        // Matrix.h:804-809: b = AsNode(A * AsDofs(x)) with AsDofs overwriting (last node wins)
        DBuf<double> xd(A->nrows), bd(A->nrows);
        int rc = gf_vector_as_dofs_nodevec(A->v, x, xd.p, stream);
        if (rc) throw Error(rc, gf_last_error());
        spmv(A, xd.p, bd.p, S(stream));
        rc = gf_vector_as_node_dofval(A->v, bd.p, b, stream);
        if (rc) throw Error(rc, gf_last_error());
        GF_HIP(hipStreamSynchronize(S(stream)));
    });
}

} // extern "C"

namespace {

std::mutex g_scratch_mu;
constexpr int MAX_DEVICES = 64;
double* g_scratch_ptr[MAX_DEVICES] = {};

// Device address of `scratch` on the current device, looked up once per device (with that device
// current) and kept.
double* scratch_ptr()
{
    int dev = 0;
    GF_HIP(hipGetDevice(&dev));
    GF_REQUIRE(dev >= 0 && dev < MAX_DEVICES, "device index below 64");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        GF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    return g_scratch_ptr[dev];
}

} // namespace

extern "C" {

int gf_matrix_dot_nodevec_noalloc(gf_matrix A, const double* x, double* b, void* stream)
{
    return guard([&] {
        GF_REQUIRE(A && x && b, "null argument");
        if (!A->assembled) throw Error(GF_ESTATE, "matrix not assembled");
        if (A->nrows > NOALLOC_MAX_DOFS)
            throw Error(GF_EINVAL, "gf_matrix_dot_nodevec_noalloc: " + std::to_string(A->nrows) +
                                       " DOFs exceed the static staging of " + std::to_string(NOALLOC_MAX_DOFS) +
                                       " (use gf_matrix_dot_nodevec)");
        double* xd = scratch_ptr();
        double* bd = xd + NOALLOC_MAX_DOFS;
        // the same three kernels as gf_matrix_dot_nodevec; k_row_last writes every DOF of xd, and
        // the product every row of bd, so nothing left by an earlier call is read
        int rc = gf_vector_as_dofs_nodevec(A->v, x, xd, stream);
        if (rc) throw Error(rc, gf_last_error());
        spmv(A, xd, bd, S(stream));
        rc = gf_vector_as_node_dofval(A->v, bd, b, stream);
        if (rc) throw Error(rc, gf_last_error());
    });
}

int gf_matrix_todense(gf_matrix A, double* out, void* stream)
{
    return guard([&] {
        GF_REQUIRE(A && out, "null argument");
        int64_t n = A->nrows * A->nrows;
        if (n) k_fill<<<nblocks(n, BS), BS, 0, S(stream)>>>(n, 0.0, out);
        if (A->nrows)
            k_todense<<<nblocks(A->nrows, BS), BS, 0, S(stream)>>>(A->nrows, A->rowptr.p, A->col.p, A->val.p, out);
        check_launch();
    });
}

} // extern "C"
