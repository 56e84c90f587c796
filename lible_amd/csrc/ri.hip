// Density-fitted (RI) Coulomb and exchange matrices from the dense 3-index and 2-index
// integrals of the library (the natural consumer of eri3()/eri2(), eri3.cpp:7-46,
// eri2.cpp:27-72; the reference has no J/K build).  With B_P(mn) = (mn|P), V = (P|Q) = L L^T:
//   J = sum_P c_P B_P,                c = V^-1 g,   g_P = sum_mn B_P(mn) D_mn
//   K = sum_P W_P D W_P,              W_P = sum_Q (L^-1)_PQ B_Q
// The 3-index tensor is held P-major (one nbf x nbf block per auxiliary function) so that the
// exchange is one batched GEMM W_P D and one GEMM over k = nbf * naux; the Cholesky
// factorisation and solves are rocSOLVER's, the products rocBLAS's (FP64 matrix cores).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "lbx_internal.hpp"

#include <string>

namespace lbx
{
    namespace
    {
        // out (cols, rows) = in (rows, cols)^T, both row-major, through 32 x 32 LDS tiles
        __global__ __launch_bounds__(256) void transposeKernel(const double *__restrict__ in,
                                                               double *__restrict__ out, long long rows,
                                                               long long cols)
        {
            __shared__ double tile[32][33];
            const long long tr = (rows + 31) / 32, tc = (cols + 31) / 32;
            for (long long t = blockIdx.x; t < tr * tc; t += gridDim.x)
            {
                const long long r0 = (t / tc) * 32, c0 = (t % tc) * 32;
                const int x = threadIdx.x & 31, y = threadIdx.x >> 5;
                for (int k = y; k < 32; k += 8)
                {
                    const long long r = r0 + k, c = c0 + x;
                    tile[k][x] = (r < rows && c < cols) ? in[r * cols + c] : 0.0;
                }
                __syncthreads();
                for (int k = y; k < 32; k += 8)
                {
                    const long long c = c0 + k, r = r0 + x;
                    if (c < cols && r < rows)
                        out[c * rows + r] = tile[x][k];
                }
                __syncthreads();
            }
        }

        void rb(rocblas_status s, const char *what)
        {
            if (s != rocblas_status_success)
                throw Error(s == rocblas_status_memory_error ? LBX_ERR_NOMEM : LBX_ERR_HIP,
                            std::string(what) + " failed (rocblas status " + std::to_string(static_cast<int>(s)) + ")");
        }

        struct Handle
        {
            rocblas_handle h = nullptr;
            explicit Handle(hipStream_t s)
            {
                rb(rocblas_create_handle(&h), "rocblas_create_handle");
                rb(rocblas_set_stream(h, s), "rocblas_set_stream");
            }
            ~Handle()
            {
                if (h)
                    (void)rocblas_destroy_handle(h);
            }
        };

        struct Buf
        {
            void *p = nullptr;
            explicit Buf(size_t bytes) { LBX_HIP(hipMalloc(&p, bytes ? bytes : 1)); }
            ~Buf()
            {
                if (p)
                    (void)hipFree(p);
            }
            double *d() const { return static_cast<double *>(p); }
        };
    }

    void riJK(lbx_context *ctx, hipStream_t stream, long long n, long long naux, const double *D, double *J,
              double *K, double *gamma_out)
    {
        const long long n2 = n * n;
        // dense (mn|P), mn-major as the eri3 entry writes it, then P-major
        Buf raw(sizeof(double) * n2 * naux);
        {
            const int rc = lbx_eri3(ctx, raw.d());
            if (rc != LBX_OK)
                throw Error(rc, std::string("lbx_ri_jk: ") + lbx_last_error());
        }
        Buf xp(sizeof(double) * n2 * naux);
        {
            const long long tiles = ((n2 + 31) / 32) * ((naux + 31) / 32);
            const int grid = static_cast<int>(tiles < 65536 ? tiles : 65536);
            hipLaunchKernelGGL(transposeKernel, dim3(grid), dim3(256), 0, stream, raw.d(), xp.d(), n2, naux);
            LBX_HIP(hipGetLastError());
        }
        Buf V(sizeof(double) * naux * naux);
        {
            const int rc = lbx_eri2(ctx, V.d());
            if (rc != LBX_OK)
                throw Error(rc, std::string("lbx_ri_jk: ") + lbx_last_error());
        }
        Handle h(stream);
        // the 3-index products run through rocBLAS's 64-bit-index entries (nbf^2 naux exceeds
        // 2^31 elements at a few hundred functions); the metric's factorisation needs 32-bit
        const int NA = static_cast<int>(naux);
        // V = L L^T (V symmetric: its row-major storage is also its column-major storage)
        Buf info(sizeof(rocblas_int));
        rb(rocsolver_dpotrf(h.h, rocblas_fill_lower, NA, V.d(), NA, static_cast<rocblas_int *>(info.p)),
           "rocsolver_dpotrf");
        rocblas_int hinfo = 0;
        LBX_HIP(hipMemcpyAsync(&hinfo, info.p, sizeof(hinfo), hipMemcpyDeviceToHost, stream));
        LBX_HIP(hipStreamSynchronize(stream));
        if (hinfo != 0)
            throw Error(LBX_ERR_STATE, "lbx_ri_jk: the auxiliary metric (P|Q) is not positive definite (pivot " +
                                               std::to_string(hinfo) + ")");
        const double one = 1.0, zero = 0.0;
        // X (column-major N2 x NA, ld N2): column P = the P-major block B_P
        Buf g(sizeof(double) * naux);
        rb(rocblas_dgemv_64(h.h, rocblas_operation_transpose, n2, naux, &one, xp.d(), n2, D, 1, &zero, g.d(), 1),
           "rocblas_dgemv (g = X^T d)");
        if (gamma_out)
            LBX_HIP(hipMemcpyAsync(gamma_out, g.d(), sizeof(double) * naux, hipMemcpyDeviceToDevice, stream));
        rb(rocsolver_dpotrs(h.h, rocblas_fill_lower, NA, 1, V.d(), NA, g.d(), NA), "rocsolver_dpotrs");
        rb(rocblas_dgemv_64(h.h, rocblas_operation_none, n2, naux, &one, xp.d(), n2, g.d(), 1, &zero, J, 1),
           "rocblas_dgemv (J = X c)");
        if (K)
        {
            // W = X L^-T: column P of W is W_P = sum_Q (L^-1)_PQ B_Q
            rb(rocblas_dtrsm_64(h.h, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                                rocblas_diagonal_non_unit, n2, naux, &one, V.d(), naux, xp.d(), n2),
               "rocblas_dtrsm (W = X L^-T)");
            // T_P = W_P D (W_P and D symmetric, so row- and column-major readings agree), into
            // the buffer of the mn-major tensor
            double *T = raw.d();
            rb(rocblas_dgemm_strided_batched_64(h.h, rocblas_operation_none, rocblas_operation_none, n, n, n, &one,
                                                xp.d(), n, n2, D, n, 0, &zero, T, n, n2, naux),
               "rocblas_dgemm_strided_batched (T_P = W_P D)");
            // K = sum_P T_P W_P^T = [T_1 .. T_naux] [W_1 .. W_naux]^T, k = nbf * naux
            rb(rocblas_dgemm_64(h.h, rocblas_operation_none, rocblas_operation_transpose, n, n, n * naux, &one, T,
                                n, xp.d(), n, &zero, K, n),
               "rocblas_dgemm (K)");
        }
        LBX_HIP(hipStreamSynchronize(stream));
    }
}
