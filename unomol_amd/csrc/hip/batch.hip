// Dense kernels of the batched SCF: many small molecules (N <= 256 functions each) whose
// matrices live back to back in one store, row-major N_m x N_m at offset off[m].
//
//  * grouped GEMM: C_m = alpha op(A_m) op(B_m) + beta C_m for every molecule, one 32 x 32
//    output tile per 256-thread block (2 x 2 outputs per thread, K through LDS in steps
//    of 32); the host lists the (molecule, tile) pairs;
//  * batched Jacobi eigensolver: one block per molecule, cyclic sweeps in parallel
//    (round-robin) order, each round applying N/2 disjoint rotations to rows, then
//    columns, then the eigenvector columns; eigenvalues ascending with their vectors;
//  * per-molecule reductions (trace products, difference norms) and helpers.
#include <hip/hip_runtime.h>
#include "batch_kernels.hpp"

namespace uno {

// ---------------------------------------------------------------------------
// grouped GEMM
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_grouped_gemm(const GemmTile* __restrict__ tiles, const GemmGroup* __restrict__ groups,
                                                      double alpha, double beta) {
    __shared__ double sA[32][33];
    __shared__ double sB[32][33];
    const GemmTile t = tiles[blockIdx.x];
    const GemmGroup g = groups[t.group];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;   // 16 x 16 threads, 2 x 2 outputs each
    const int i0 = t.ti * 32, j0 = t.tj * 32;
    double acc[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (int k0 = 0; k0 < g.k; k0 += 32) {
        for (int e = threadIdx.x; e < 32 * 32; e += 256) {
            const int r = e >> 5, c = e & 31;
            // op(A)[i][k], i = i0 + r, k = k0 + c
            const int i = i0 + r, k = k0 + c;
            double av = 0.0;
            if (i < g.m && k < g.k) av = g.ta ? g.A[(long)k * g.lda + i] : g.A[(long)i * g.lda + k];
            sA[r][c] = av;
            // op(B)[k][j], k = k0 + r, j = j0 + c
            const int kk = k0 + r, j = j0 + c;
            double bv = 0.0;
            if (kk < g.k && j < g.n) bv = g.tb ? g.B[(long)j * g.ldb + kk] : g.B[(long)kk * g.ldb + j];
            sB[r][c] = bv;
        }
        __syncthreads();
#pragma unroll 8
        for (int k = 0; k < 32; ++k) {
            const double a0 = sA[ty * 2][k], a1 = sA[ty * 2 + 1][k];
            const double b0 = sB[k][tx * 2], b1 = sB[k][tx * 2 + 1];
            acc[0][0] = fma(a0, b0, acc[0][0]);
            acc[0][1] = fma(a0, b1, acc[0][1]);
            acc[1][0] = fma(a1, b0, acc[1][0]);
            acc[1][1] = fma(a1, b1, acc[1][1]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int i = i0 + ty * 2 + a, j = j0 + tx * 2 + b;
            if (i < g.m && j < g.n) {
                double* c = g.C + (long)i * g.ldc + j;
                *c = alpha * acc[a][b] + (beta != 0.0 ? beta * *c : 0.0);
            }
        }
}

hipError_t grouped_gemm(int ntiles, const GemmTile* tiles, const GemmGroup* groups, double alpha, double beta,
                        hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_grouped_gemm, dim3(ntiles), dim3(256), 0, st, tiles, groups, alpha, beta);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// batched Jacobi: A (destroyed) -> eigenvalues ascending, eigenvectors in the columns of Z
// ---------------------------------------------------------------------------
// Cyclic Jacobi in round-robin order: with m = N rounded up to even (a dummy index N when N
// is odd), round r pairs (m-1, r) and ((r+k) mod (m-1), (r-k) mod (m-1)), k = 1..m/2-1, so
// that each round is a perfect matching and m-1 rounds visit every pair once.  A rotation
// is skipped when |a_pq| <= max(eps/2 sqrt(|a_pp a_qq|), 1e-15 ||A||_F) (it would not change
// the diagonal, or the pair is converged to working accuracy of the whole matrix: the
// relative test alone never settles next to zero diagonal entries, where rounding keeps
// a_pq above it); sweeps stop when none is applied.
constexpr int kJacobiThreads = 1024;
__global__ __launch_bounds__(kJacobiThreads) void k_batched_jacobi(const BatchMat* __restrict__ mats, int max_sweeps,
                                                                   int* __restrict__ sweeps_out) {
    __shared__ double cs[128], sn[128];
    __shared__ int pp[128], qq[128];
    __shared__ int nrot;
    __shared__ double red[kJacobiThreads / 64];
    const BatchMat M = mats[blockIdx.x];
    const int n = M.n;
    double* A = M.A;
    double* V = M.V;
    const int tid = threadIdx.x;
    double f2 = 0.0;
    for (long e = tid; e < (long)n * n; e += blockDim.x) {
        V[e] = (e / n == e % n) ? 1.0 : 0.0;
        f2 = fma(A[e], A[e], f2);
    }
    for (int o = 32; o >= 1; o >>= 1) f2 += __shfl_xor(f2, o, 64);
    if ((tid & 63) == 0) red[tid >> 6] = f2;
    __syncthreads();
    double fro2 = 0.0;
    for (int w = 0; w < kJacobiThreads / 64; ++w) fro2 += red[w];
    const double abs_tol = 1e-15 * sqrt(fro2);
    const int m = n + (n & 1);
    int sweep = 0;
    while (sweep < max_sweeps && n > 1) {
        ++sweep;
        if (tid == 0) nrot = 0;
        __syncthreads();
        for (int r = 0; r < m - 1; ++r) {
            for (int k = tid; k < m / 2; k += blockDim.x) {
                int a, b;
                if (k == 0) { a = m - 1; b = r; }
                else { a = (r + k) % (m - 1); b = (r - k + (m - 1)) % (m - 1); }
                const int p = a < b ? a : b, q = a < b ? b : a;
                double c = 1.0, s = 0.0;
                if (q < n) {
                    const double apq = A[(long)p * n + q];
                    const double app = A[(long)p * n + p], aqq = A[(long)q * n + q];
                    if (fabs(apq) > fmax(1.1102230246251565e-16 * sqrt(fabs(app * aqq)), abs_tol)) {
                        const double theta = (aqq - app) / (2.0 * apq);
                        const double t = fabs(theta) < 1e150
                                             ? (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0))
                                             : 0.5 / theta;
                        c = 1.0 / sqrt(t * t + 1.0);
                        s = t * c;
                        atomicAdd(&nrot, 1);
                    }
                }
                pp[k] = p;
                qq[k] = q;
                cs[k] = c;
                sn[k] = s;
            }
            __syncthreads();
            // rows p, q of A: A <- J^T A
            for (long e = tid; e < (long)(m / 2) * n; e += blockDim.x) {
                const int k = (int)(e / n), j = (int)(e % n);
                const int p = pp[k], q = qq[k];
                if (q >= n || sn[k] == 0.0) continue;
                const double c = cs[k], s = sn[k];
                const double ap = A[(long)p * n + j], aq = A[(long)q * n + j];
                A[(long)p * n + j] = c * ap - s * aq;
                A[(long)q * n + j] = s * ap + c * aq;
            }
            __syncthreads();
            // columns p, q of A and of V: A <- A J, V <- V J
            for (long e = tid; e < (long)(m / 2) * n; e += blockDim.x) {
                const int k = (int)(e / n), i = (int)(e % n);
                const int p = pp[k], q = qq[k];
                if (q >= n || sn[k] == 0.0) continue;
                const double c = cs[k], s = sn[k];
                const double ap = A[(long)i * n + p], aq = A[(long)i * n + q];
                A[(long)i * n + p] = c * ap - s * aq;
                A[(long)i * n + q] = s * ap + c * aq;
                const double vp = V[(long)i * n + p], vq = V[(long)i * n + q];
                V[(long)i * n + p] = c * vp - s * vq;
                V[(long)i * n + q] = s * vp + c * vq;
            }
            __syncthreads();
        }
        const int done = nrot == 0;
        __syncthreads();
        if (done) break;
    }
    __syncthreads();
    // eigenvalues ascending (ties by index) with their vectors as columns of Z
    for (int i = tid; i < n; i += blockDim.x) {
        const double di = A[(long)i * n + i];
        int rank = 0;
        for (int j = 0; j < n; ++j) {
            const double dj = A[(long)j * n + j];
            if (dj < di || (dj == di && j < i)) ++rank;
        }
        M.evals[rank] = di;
        for (int r = 0; r < n; ++r) M.Z[(long)r * n + rank] = V[(long)r * n + i];
    }
    if (tid == 0 && sweeps_out) sweeps_out[blockIdx.x] = sweep;
}

hipError_t batched_jacobi(int nmat, const BatchMat* mats, int max_sweeps, int* sweeps, hipStream_t st) {
    if (nmat <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_batched_jacobi, dim3(nmat), dim3(kJacobiThreads), 0, st, mats, max_sweeps, sweeps);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// per-molecule helpers: one 256-thread block per slot
// ---------------------------------------------------------------------------
__device__ __forceinline__ double block_reduce_sum(double v, double* red) {
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += red[k];
    __syncthreads();
    return s;
}

// out[k] (+)= sum_ij A_k,ij B_k,ij over slot k's n x n matrices
__global__ __launch_bounds__(256) void k_batch_dot(const MatSlot* __restrict__ slots, const double* __restrict__ A,
                                                   const double* __restrict__ B, double* __restrict__ out, int acc) {
    __shared__ double red[4];
    const MatSlot m = slots[blockIdx.x];
    double s = 0.0;
    for (long e = threadIdx.x; e < (long)m.n * m.n; e += blockDim.x) s += A[m.off + e] * B[m.off + e];
    s = block_reduce_sum(s, red);
    if (threadIdx.x == 0) out[blockIdx.x] = acc ? out[blockIdx.x] + s : s;
}

// out[k] = sqrt(sum (A - B)^2) / n over slot k
__global__ __launch_bounds__(256) void k_batch_diffnorm(const MatSlot* __restrict__ slots, const double* __restrict__ A,
                                                        const double* __restrict__ B, double* __restrict__ out) {
    __shared__ double red[4];
    const MatSlot m = slots[blockIdx.x];
    double s = 0.0;
    for (long e = threadIdx.x; e < (long)m.n * m.n; e += blockDim.x) {
        const double d = A[m.off + e] - B[m.off + e];
        s += d * d;
    }
    s = block_reduce_sum(s, red);
    if (threadIdx.x == 0) out[blockIdx.x] = sqrt(s) / m.n;
}

// G = W + W^T (or G += ...) per slot
__global__ __launch_bounds__(256) void k_batch_sym(const MatSlot* __restrict__ slots, const double* __restrict__ W,
                                                   double* __restrict__ G, int acc) {
    const MatSlot m = slots[blockIdx.x];
    for (long e = threadIdx.x; e < (long)m.n * m.n; e += blockDim.x) {
        const int i = (int)(e / m.n), j = (int)(e % m.n);
        const double v = W[m.off + e] + W[m.off + (long)j * m.n + i];
        G[m.off + e] = acc ? G[m.off + e] + v : v;
    }
}

// Dsh_k[s][t] = max |D| over the function block of shells s, t of slot k's molecule (D1, D2
// optional further densities of the same layout)
__global__ __launch_bounds__(256) void k_batch_shell_max(const MatSlot* __restrict__ slots, const int* __restrict__ sh_off,
                                                         const int* __restrict__ sh_n, const double* __restrict__ D,
                                                         const double* __restrict__ D1, const double* __restrict__ D2,
                                                         double* __restrict__ Dsh) {
    const MatSlot m = slots[blockIdx.x];
    const int ns = m.nsh;
    for (int st = threadIdx.x; st < ns * ns; st += blockDim.x) {
        const int s = st / ns, t = st % ns;
        const int o1 = sh_off[m.shoff + s], n1 = sh_n[m.shoff + s];
        const int o2 = sh_off[m.shoff + t], n2 = sh_n[m.shoff + t];
        double v = 0.0;
        for (int a = 0; a < n1; ++a)
            for (int b = 0; b < n2; ++b) {
                const long ij = m.off + (long)(o1 + a) * m.n + o2 + b;
                v = fmax(v, fabs(D[ij]));
                if (D1) v = fmax(v, fabs(D1[ij]));
                if (D2) v = fmax(v, fabs(D2[ij]));
            }
        Dsh[m.dshoff + st] = v;
    }
}

// F_k = sum_j c[k * ncoef + j] H_j,k: a linear combination of ncoef stores per slot
__global__ __launch_bounds__(256) void k_batch_lincomb(const MatSlot* __restrict__ slots, int ncoef,
                                                       const double* __restrict__ coef, const double* const* __restrict__ src,
                                                       double* __restrict__ F) {
    const MatSlot m = slots[blockIdx.x];
    for (long e = threadIdx.x; e < (long)m.n * m.n; e += blockDim.x) {
        double v = 0.0;
        for (int j = 0; j < ncoef; ++j) {
            const double cj = coef[(long)blockIdx.x * ncoef + j];
            if (cj != 0.0) v = fma(cj, src[j][m.off + e], v);
        }
        F[m.off + e] = v;
    }
}

hipError_t batch_dot(int nslot, const MatSlot* slots, const double* A, const double* B, double* out, bool acc,
                     hipStream_t st) {
    if (nslot <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch_dot, dim3(nslot), dim3(256), 0, st, slots, A, B, out, acc ? 1 : 0);
    return hipGetLastError();
}
hipError_t batch_diffnorm(int nslot, const MatSlot* slots, const double* A, const double* B, double* out,
                          hipStream_t st) {
    if (nslot <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch_diffnorm, dim3(nslot), dim3(256), 0, st, slots, A, B, out);
    return hipGetLastError();
}
hipError_t batch_sym(int nslot, const MatSlot* slots, const double* W, double* G, bool acc, hipStream_t st) {
    if (nslot <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch_sym, dim3(nslot), dim3(256), 0, st, slots, W, G, acc ? 1 : 0);
    return hipGetLastError();
}
hipError_t batch_shell_max(int nslot, const MatSlot* slots, const int* sh_off, const int* sh_n, const double* D,
                           const double* D1, const double* D2, double* Dsh, hipStream_t st) {
    if (nslot <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch_shell_max, dim3(nslot), dim3(256), 0, st, slots, sh_off, sh_n, D, D1, D2, Dsh);
    return hipGetLastError();
}
hipError_t batch_lincomb(int nslot, const MatSlot* slots, int ncoef, const double* coef, const double* const* src,
                         double* F, hipStream_t st) {
    if (nslot <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch_lincomb, dim3(nslot), dim3(256), 0, st, slots, ncoef, coef, src, F);
    return hipGetLastError();
}

}  // namespace uno
