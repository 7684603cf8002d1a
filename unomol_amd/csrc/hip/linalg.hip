// Dense FP64 kernels: MFMA GEMM, Householder tridiagonalisation with reflector
// accumulation, row-parallel application of QL rotation sweeps, and the small
// element-wise / reduction kernels of the SCF loop.
#include <hip/hip_runtime.h>
#include "kernels.hpp"

namespace uno {

typedef double double4_t __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// element-wise
// ---------------------------------------------------------------------------
__global__ void k_unpack_(int n, const double* __restrict__ pk, double* __restrict__ D) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n * n) return;
    const int i = (int)(t / n), j = (int)(t % n);
    D[t] = i >= j ? pk[(long)i * (i + 1) / 2 + j] : pk[(long)j * (j + 1) / 2 + i];
}
hipError_t k_unpack(int n, const double* packed, double* dense, hipStream_t st) {
    const long len = (long)n * n;
    hipLaunchKernelGGL(k_unpack_, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, n, packed, dense);
    return hipGetLastError();
}

__global__ void k_pack_(int n, const double* __restrict__ D, double* __restrict__ pk, int mode) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n2 = (long)n * (n + 1) / 2;
    if (t >= n2) return;
    // invert t = i(i+1)/2 + j
    int i = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((long)i * (i + 1) / 2 > t) --i;
    while ((long)(i + 1) * (i + 2) / 2 <= t) ++i;
    const int j = (int)(t - (long)i * (i + 1) / 2);
    if (mode == 0) pk[t] = D[(long)i * n + j];
    else pk[t] += D[(long)i * n + j] + D[(long)j * n + i];
}
hipError_t k_pack(int n, const double* dense, double* packed, hipStream_t st) {
    const long n2 = (long)n * (n + 1) / 2;
    hipLaunchKernelGGL(k_pack_, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, st, n, dense, packed, 0);
    return hipGetLastError();
}
hipError_t k_pack_sym_add(int n, const double* W, double* packed, hipStream_t st) {
    const long n2 = (long)n * (n + 1) / 2;
    hipLaunchKernelGGL(k_pack_, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, st, n, W, packed, 1);
    return hipGetLastError();
}

__global__ void k_sym_w_(int n, const double* __restrict__ W, double* __restrict__ G) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n * n) return;
    const int i = (int)(t / n), j = (int)(t % n);
    G[t] = W[t] + W[(long)j * n + i];
}
hipError_t k_sym_w(int n, const double* W, double* G, hipStream_t st) {
    const long len = (long)n * n;
    hipLaunchKernelGGL(k_sym_w_, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, n, W, G);
    return hipGetLastError();
}

__global__ void k_axpby_(long len, double a, const double* x, double b, const double* y, double* z) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < len) z[t] = a * x[t] + b * y[t];
}
hipError_t k_axpby(long len, double a, const double* x, double b, const double* y, double* z, hipStream_t st) {
    hipLaunchKernelGGL(k_axpby_, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, len, a, x, b, y, z);
    return hipGetLastError();
}

__global__ void k_fill_(long len, double v, double* x) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < len) x[t] = v;
}
hipError_t k_fill(long len, double v, double* x, hipStream_t st) {
    if (len <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, len, v, x);
    return hipGetLastError();
}

// deterministic single-block reductions (1024 threads, fixed order)
__device__ double block_sum(double v, double* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] += sh[t + s];
        __syncthreads();
    }
    return sh[0];
}

__global__ __launch_bounds__(1024) void k_dot_(long len, const double* A, const double* B, double* out) {
    __shared__ double sh[1024];
    double s = 0.0;
    for (long i = threadIdx.x; i < len; i += 1024) s = fma(A[i], B[i], s);
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[0] += s;
}
hipError_t k_dot(long len, const double* A, const double* B, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_dot_, dim3(1), dim3(1024), 0, st, len, A, B, out);
    return hipGetLastError();
}

__global__ __launch_bounds__(1024) void k_diffnorm_(long len, int n, const double* A, const double* B, double* out) {
    __shared__ double sh[1024];
    double s = 0.0;
    for (long i = threadIdx.x; i < len; i += 1024) {
        const double d = A[i] - B[i];
        s = fma(d, d, s);
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[0] = sqrt(s) / n;
}
hipError_t k_diffnorm(long len, int n, const double* A, const double* B, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_diffnorm_, dim3(1), dim3(1024), 0, st, len, n, A, B, out);
    return hipGetLastError();
}

// one block per shell pair (s, t): max |D| over its function block of up to three matrices
__global__ void k_shell_max_(int n, int nsh, const int* off, const int* nf, const double* D, const double* D1,
                             const double* D2, double* Dsh) {
    const int s = blockIdx.x / nsh, t = blockIdx.x % nsh;
    double m = 0.0;
    const int ns = nf[s], nt = nf[t];
    for (int q = threadIdx.x; q < ns * nt; q += 64) {
        const long ij = (long)(off[s] + q / nt) * n + off[t] + q % nt;
        m = fmax(m, fabs(D[ij]));
        if (D1) m = fmax(m, fabs(D1[ij]));
        if (D2) m = fmax(m, fabs(D2[ij]));
    }
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o, 64));
    if (threadIdx.x == 0) Dsh[blockIdx.x] = m;
}
hipError_t k_shell_max(int n, int nsh, const int* sh_off, const int* sh_n, const double* D, const double* D1,
                       const double* D2, double* Dsh, hipStream_t st) {
    hipLaunchKernelGGL(k_shell_max_, dim3((unsigned)(nsh * nsh)), dim3(64), 0, st, n, nsh, sh_off, sh_n, D, D1, D2,
                       Dsh);
    return hipGetLastError();
}

__global__ __launch_bounds__(1024) void k_absmax_(long len, const double* x, double* out) {
    __shared__ double sh[1024];
    double m = 0.0;
    for (long i = threadIdx.x; i < len; i += 1024) m = fmax(m, fabs(x[i]));
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}
hipError_t k_absmax(long len, const double* x, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_absmax_, dim3(1), dim3(1024), 0, st, len, x, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// FP64 MFMA GEMM: 64x64 block tile, 4 waves each owning a 32x32 quarter made of
// 2x2 tiles of v_mfma_f64_16x16x4_f64.  Operand layout of that instruction:
// A(16x4): lane l holds A[l%16][l/16]; B(4x16): lane l holds B[l/16][l%16];
// D(16x16): lane l holds D[4*r + l/16][l%16] in register r (pinned by an on-device probe).
// ---------------------------------------------------------------------------
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, double alpha, const double* __restrict__ A,
                                              int lda, const double* __restrict__ B, int ldb, double beta,
                                              double* __restrict__ C, int ldc) {
    __shared__ double sA[16][64 + 2];
    __shared__ double sB[16][64 + 2];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int bi = blockIdx.y * 64, bj = blockIdx.x * 64;
    const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
    double4_t acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = (double4_t){0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < K; k0 += 16) {
        for (int e = tid; e < 1024; e += 256) {
            int i, kk;
            if (TA) { kk = e / 64; i = e % 64; } else { i = e / 16; kk = e % 16; }
            const int gi = bi + i, gk = k0 + kk;
            double v = 0.0;
            if (gi < M && gk < K) v = TA ? A[(long)gk * lda + gi] : A[(long)gi * lda + gk];
            sA[kk][i] = v;
        }
        for (int e = tid; e < 1024; e += 256) {
            int j, kk;
            if (TB) { j = e / 16; kk = e % 16; } else { kk = e / 64; j = e % 64; }
            const int gj = bj + j, gk = k0 + kk;
            double v = 0.0;
            if (gj < N && gk < K) v = TB ? B[(long)gj * ldb + gk] : B[(long)gk * ldb + gj];
            sB[kk][j] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; kk += 4) {
            double av[2], bv[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                av[t] = sA[kk + (lane >> 4)][wi + t * 16 + (lane & 15)];
                bv[t] = sB[kk + (lane >> 4)][wj + t * 16 + (lane & 15)];
            }
#pragma unroll
            for (int ta = 0; ta < 2; ++ta)
#pragma unroll
                for (int tb = 0; tb < 2; ++tb)
                    acc[ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[ta], bv[tb], acc[ta][tb], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int ta = 0; ta < 2; ++ta)
#pragma unroll
        for (int tb = 0; tb < 2; ++tb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gi = bi + wi + ta * 16 + 4 * r + (lane >> 4);
                const int gj = bj + wj + tb * 16 + (lane & 15);
                if (gi < M && gj < N) {
                    double* c = C + (long)gi * ldc + gj;
                    *c = (beta == 0.0 ? 0.0 : beta * *c) + alpha * acc[ta][tb][r];
                }
            }
}

hipError_t gemm_f64(bool ta, bool tb, int m, int n, int k, double alpha, const double* A, int lda, const double* B,
                    int ldb, double beta, double* C, int ldc, hipStream_t st) {
    const dim3 grid((unsigned)((n + 63) / 64), (unsigned)((m + 63) / 64)), blk(256);
    if (!ta && !tb) hipLaunchKernelGGL((k_gemm<false, false>), grid, blk, 0, st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    else if (ta && !tb) hipLaunchKernelGGL((k_gemm<true, false>), grid, blk, 0, st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    else if (!ta && tb) hipLaunchKernelGGL((k_gemm<false, true>), grid, blk, 0, st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    else hipLaunchKernelGGL((k_gemm<true, true>), grid, blk, 0, st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Householder tridiagonalisation A = Q T Q^T (one reflector per column k):
//   x = A[k+1:,k], alpha = -sign(x0)|x|, v = x - alpha e1, tau = 2/|v|^2
//   p = tau A22 v, K = tau/2 p.v, w = p - K v, A22 -= v w^T + w v^T
// d[k] = A[k][k], e[k] = alpha.  Reflector k is stored in row k of V.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_hh_vec(int n, int k, double* A, double* V, double* tau, double* d,
                                                 double* e) {
    __shared__ double sh[1024];
    const int m = n - k - 1;
    double s = 0.0;
    for (int i = threadIdx.x; i < m; i += 1024) {
        const double x = A[(long)(k + 1 + i) * n + k];
        s = fma(x, x, s);
    }
    const double nx2 = block_sum(s, sh);
    const double x0 = A[(long)(k + 1) * n + k];
    const double nx = sqrt(nx2);
    const double alpha = x0 > 0.0 ? -nx : nx;
    const double v0 = x0 - alpha;
    const double vn2 = nx2 - x0 * x0 + v0 * v0;
    for (int i = threadIdx.x; i < m; i += 1024) {
        const double x = A[(long)(k + 1 + i) * n + k];
        V[(long)k * n + k + 1 + i] = i == 0 ? v0 : x;
    }
    if (threadIdx.x == 0) {
        d[k] = A[(long)k * n + k];
        if (nx2 == 0.0 || vn2 == 0.0) {
            tau[k] = 0.0;
            e[k] = x0;
        } else {
            tau[k] = 2.0 / vn2;
            e[k] = alpha;
        }
    }
}

// p_i = tau sum_j A22[i][j] v_j, one wave per row
__global__ __launch_bounds__(256) void k_hh_symv(int n, int k, const double* A, const double* V, const double* tau,
                                                 double* p) {
    const int m = n - k - 1;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= m) return;
    const double* a = A + (long)(k + 1 + row) * n + k + 1;
    const double* v = V + (long)k * n + k + 1;
    double s = 0.0;
    for (int j = lane; j < m; j += 64) s = fma(a[j], v[j], s);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if (lane == 0) p[row] = tau[k] * s;
}

// scal[0] = K = tau/2 p.v ; p <- w = p - K v
__global__ __launch_bounds__(1024) void k_hh_w(int n, int k, const double* V, const double* tau, double* p,
                                               double* scal) {
    __shared__ double sh[1024];
    const int m = n - k - 1;
    const double* v = V + (long)k * n + k + 1;
    double s = 0.0;
    for (int i = threadIdx.x; i < m; i += 1024) s = fma(p[i], v[i], s);
    const double K = 0.5 * tau[k] * block_sum(s, sh);
    for (int i = threadIdx.x; i < m; i += 1024) p[i] -= K * v[i];
}

__global__ __launch_bounds__(256) void k_hh_rank2(int n, int k, double* A, const double* V, const double* w) {
    const int m = n - k - 1;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)m * m) return;
    const int i = (int)(t / m), j = (int)(t % m);
    const double* v = V + (long)k * n + k + 1;
    A[(long)(k + 1 + i) * n + k + 1 + j] -= v[i] * w[j] + w[i] * v[j];
}

__global__ void k_hh_last(int n, double* A, double* d, double* e) {
    if (n >= 2) {
        d[n - 2] = A[(long)(n - 2) * n + n - 2];
        e[n - 2] = A[(long)(n - 1) * n + n - 2];
    }
    d[n - 1] = A[(long)(n - 1) * n + n - 1];
    e[n - 1] = 0.0;
}

hipError_t householder_tridiag(int n, EigWork& w, hipStream_t st) {
    for (int k = 0; k + 2 < n; ++k) {
        const int m = n - k - 1;
        hipLaunchKernelGGL(k_hh_vec, dim3(1), dim3(1024), 0, st, n, k, w.A, w.V, w.tau, w.d, w.e);
        hipLaunchKernelGGL(k_hh_symv, dim3((unsigned)((m + 3) / 4)), dim3(256), 0, st, n, k, w.A, w.V, w.tau, w.p);
        hipLaunchKernelGGL(k_hh_w, dim3(1), dim3(1024), 0, st, n, k, w.V, w.tau, w.p, w.scal);
        const long mm = (long)m * m;
        hipLaunchKernelGGL(k_hh_rank2, dim3((unsigned)((mm + 255) / 256)), dim3(256), 0, st, n, k, w.A, w.V, w.p);
    }
    hipLaunchKernelGGL(k_hh_last, dim3(1), dim3(1), 0, st, n, w.A, w.d, w.e);
    return hipGetLastError();
}

// Q = H_0 H_1 ... H_{n-3}: starting from I apply H_k for k descending; thread per column
__global__ void k_eye(int n, double* Q) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (long)n * n) Q[t] = (t / n == t % n) ? 1.0 : 0.0;
}
__global__ __launch_bounds__(256) void k_q_apply(int n, int k, double* Q, const double* V, const double* tau) {
    const int m = n - k - 1;
    const int j = k + 1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double tk = tau[k];
    if (tk == 0.0) return;
    const double* v = V + (long)k * n + k + 1;
    double s = 0.0;
    for (int i = 0; i < m; ++i) s = fma(v[i], Q[(long)(k + 1 + i) * n + j], s);
    s *= tk;
    for (int i = 0; i < m; ++i) Q[(long)(k + 1 + i) * n + j] -= s * v[i];
}
hipError_t accumulate_q(int n, EigWork& w, hipStream_t st) {
    const long nn = (long)n * n;
    hipLaunchKernelGGL(k_eye, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, n, w.Q);
    for (int k = n - 3; k >= 0; --k) {
        const int m = n - k - 1;
        hipLaunchKernelGGL(k_q_apply, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, n, k, w.Q, w.V, w.tau);
    }
    return hipGetLastError();
}

// Row-parallel application of QL sweeps.  Sweep s rotates columns (i, i+1) for
// i = hi-1 down to lo with (c, s) pairs stored consecutively; each row of Q is
// independent, so every thread owns one row, staged in LDS.
__global__ void k_apply_sweeps(int n, double* Q, int nsweep, const int* __restrict__ sweeps,
                               const double* __restrict__ cs, int rows_per_block) {
    extern __shared__ double srow[];
    const int r0 = blockIdx.x * rows_per_block;
    const int nr = min(rows_per_block, n - r0);
    for (int q = threadIdx.x; q < nr * n; q += blockDim.x) srow[q] = Q[(long)r0 * n + q];
    __syncthreads();
    if ((int)threadIdx.x < nr) {
        double* z = srow + threadIdx.x * n;
        for (int s = 0; s < nsweep; ++s) {
            const int lo = sweeps[3 * s], hi = sweeps[3 * s + 1];
            const double* r = cs + 2 * (long)sweeps[3 * s + 2];
            double carry = z[hi];
            for (int i = hi - 1; i >= lo; --i) {
                const double c = r[0], sn = r[1];
                r += 2;
                const double lo_v = z[i];
                z[i + 1] = sn * lo_v + c * carry;
                carry = c * lo_v - sn * carry;
            }
            z[lo] = carry;
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < nr * n; q += blockDim.x) Q[(long)r0 * n + q] = srow[q];
}
hipError_t apply_sweeps(int n, double* Q, int nsweep, const int* sweeps, const double* cs, hipStream_t st) {
    int rpb = (int)((60 * 1024) / (8 * (long)n));
    if (rpb > 64) rpb = 64;
    if (rpb < 1) rpb = 1;
    const size_t lds = (size_t)rpb * n * sizeof(double);
    const unsigned nb = (unsigned)((n + rpb - 1) / rpb);
    hipLaunchKernelGGL(k_apply_sweeps, dim3(nb), dim3(64), lds, st, n, Q, nsweep, sweeps, cs, rpb);
    return hipGetLastError();
}

__global__ void k_perm(int n, const double* Q, const int* perm, double* Z) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n * n) return;
    const int i = (int)(t / n), j = (int)(t % n);
    Z[t] = Q[(long)i * n + perm[j]];
}
hipError_t permute_cols(int n, const double* Q, const int* perm, double* Z, hipStream_t st) {
    const long nn = (long)n * n;
    hipLaunchKernelGGL(k_perm, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, n, Q, perm, Z);
    return hipGetLastError();
}

}  // namespace uno
