// Density-fitted (RI) Coulomb and exchange matrices from the dense 3-index and 2-index
// integrals of the library (the natural consumer of eri3()/eri2(), eri3.cpp:7-46,
// eri2.cpp:27-72; the reference has no J/K build).  With B_P(mn) = (mn|P), V = (P|Q) = L L^T
// and the fitted tensor W_P = sum_Q (L^-1)_PQ B_Q (so that sum_P W_P W_P^T = B V^-1 B^T):
//   J = sum_P W_P h_P,   h_P = sum_mn W_P(mn) D_mn
//   K = sum_P W_P D W_P
// W depends on the geometry only: the first call builds it (eri3, a transpose to P-major, eri2,
// the Cholesky factorisation, one triangular solve) and the context keeps it, so every later
// call (an SCF iteration) costs two matrix-vector products for J and, for K, one batched GEMM
// W_P D and one GEMM over k = nbf * naux.  Factorisation by rocSOLVER, products by rocBLAS.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "lbx_internal.hpp"

#include <memory>
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

    void riJK(lbx_context *ctx, Context &c, hipStream_t stream, long long n, long long naux, const double *D,
              double *J, double *K)
    {
        std::lock_guard<std::mutex> lk(c.ri_mu);
        const long long n2 = n * n;
        Handle h(stream);
        if (K && !c.ri_t)
            c.ri_t = c.alloc<double>(static_cast<size_t>(n2 * naux));
        if (!c.ri_ready)
        {
            if (!c.ri_w)
                c.ri_w = c.alloc<double>(static_cast<size_t>(n2 * naux));
            // dense (mn|P), mn-major as the eri3 entry writes it, then P-major into W's buffer
            std::unique_ptr<Buf> tmp;
            double *raw = c.ri_t;
            if (!raw)
            {
                tmp = std::make_unique<Buf>(sizeof(double) * n2 * naux);
                raw = tmp->d();
            }
            {
                const int rc = lbx_eri3(ctx, raw);
                if (rc != LBX_OK)
                    throw Error(rc, std::string("lbx_ri_jk: ") + lbx_last_error());
            }
            {
                const long long tiles = ((n2 + 31) / 32) * ((naux + 31) / 32);
                const int grid = static_cast<int>(tiles < 65536 ? tiles : 65536);
                hipLaunchKernelGGL(transposeKernel, dim3(grid), dim3(256), 0, stream, raw, c.ri_w, n2, naux);
                LBX_HIP(hipGetLastError());
            }
            Buf V(sizeof(double) * naux * naux);
            {
                const int rc = lbx_eri2(ctx, V.d());
                if (rc != LBX_OK)
                    throw Error(rc, std::string("lbx_ri_jk: ") + lbx_last_error());
            }
            // V = L L^T (V symmetric: its row-major storage is also its column-major storage);
            // the factorisation takes 32-bit sizes, the 3-index products rocBLAS's 64-bit
            // entries (nbf^2 naux exceeds 2^31 elements at a few hundred functions)
            const int NA = static_cast<int>(naux);
            Buf info(sizeof(rocblas_int));
            rb(rocsolver_dpotrf(h.h, rocblas_fill_lower, NA, V.d(), NA, static_cast<rocblas_int *>(info.p)),
               "rocsolver_dpotrf");
            rocblas_int hinfo = 0;
            LBX_HIP(hipMemcpyAsync(&hinfo, info.p, sizeof(hinfo), hipMemcpyDeviceToHost, stream));
            LBX_HIP(hipStreamSynchronize(stream));
            if (hinfo != 0)
                throw Error(LBX_ERR_STATE, "lbx_ri_jk: the auxiliary metric (P|Q) is not positive definite (pivot " +
                                                   std::to_string(hinfo) + ")");
            // W = X L^-T with X column-major n2 x naux (column P = the P-major block B_P)
            const double one = 1.0;
            rb(rocblas_dtrsm_64(h.h, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                                rocblas_diagonal_non_unit, n2, naux, &one, V.d(), naux, c.ri_w, n2),
               "rocblas_dtrsm (W = X L^-T)");
            LBX_HIP(hipStreamSynchronize(stream));
            c.ri_ready = true;
        }
        const double one = 1.0, zero = 0.0;
        Buf g(sizeof(double) * naux);
        rb(rocblas_dgemv_64(h.h, rocblas_operation_transpose, n2, naux, &one, c.ri_w, n2, D, 1, &zero, g.d(), 1),
           "rocblas_dgemv (h = W^T d)");
        rb(rocblas_dgemv_64(h.h, rocblas_operation_none, n2, naux, &one, c.ri_w, n2, g.d(), 1, &zero, J, 1),
           "rocblas_dgemv (J = W h)");
        if (K)
        {
            // T_P = W_P D (W_P and D symmetric, so row- and column-major readings agree)
            rb(rocblas_dgemm_strided_batched_64(h.h, rocblas_operation_none, rocblas_operation_none, n, n, n, &one,
                                                c.ri_w, n, n2, D, n, 0, &zero, c.ri_t, n, n2, naux),
               "rocblas_dgemm_strided_batched (T_P = W_P D)");
            // K = sum_P T_P W_P^T = [T_1 .. T_naux] [W_1 .. W_naux]^T, k = nbf * naux
            rb(rocblas_dgemm_64(h.h, rocblas_operation_none, rocblas_operation_transpose, n, n, n * naux, &one,
                                c.ri_t, n, c.ri_w, n, &zero, K, n),
               "rocblas_dgemm (K)");
        }
        LBX_HIP(hipStreamSynchronize(stream));
    }
}
