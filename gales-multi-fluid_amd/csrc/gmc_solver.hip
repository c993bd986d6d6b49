// gmc_solver.hip -- device kernels of the Krylov consumer and Newton update:
// BSR SpMV, point-Jacobi, the batched dot / axpy of classical Gram-Schmidt
// (applied twice) for GMRES(k), and the corrector + Y trim.
// Reductions are two-pass with a fixed summation tree: results do not depend
// on scheduling (run-to-run bitwise reproducible).
#include <hip/hip_runtime.h>

#include "gmc_kernels.hpp"
#include "gmc_physics.hpp"

namespace gmc {

static constexpr int kThreads = 256;

static inline unsigned grid_for(long long n, int threads = kThreads) {
  long long g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  return (unsigned)g;
}

__global__ void k_fill(double* x, long long n, double v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

hipError_t launch_fill(double* x, long long n, double v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(kThreads), 0, s, x, n, v);
  return hipGetLastError();
}

// y(row) = sum over blocks of the node row; one thread per scalar row,
// ndof consecutive threads share the node's block row (contiguous reads).
template <int N, bool RES>
__global__ __launch_bounds__(256) void k_spmv(int n_nodes, const int* __restrict__ rowptr,
                                              const int* __restrict__ colidx,
                                              const double* __restrict__ vals,
                                              const double* __restrict__ x,
                                              const double* __restrict__ b,
                                              double* __restrict__ y) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= (long long)n_nodes * N) return;
  const int i = (int)(r / N), k = (int)(r % N);
  double acc = 0.0;
  for (int t = rowptr[i]; t < rowptr[i + 1]; t++) {
    const double* V = vals + (long long)t * N * N + k * N;
    const double* xx = x + (long long)colidx[t] * N;
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < N; j++) s += V[j] * xx[j];
    acc += s;
  }
  y[r] = RES ? b[r] - acc : acc;
}

template <bool RES>
static hipError_t spmv_any(int ndof, int n_nodes, const int* rowptr, const int* colidx,
                           const double* vals, const double* x, const double* b, double* y,
                           hipStream_t s) {
  const long long n = (long long)n_nodes * ndof;
  const dim3 g(grid_for(n)), t(kThreads);
  switch (ndof) {
    case 5: hipLaunchKernelGGL((k_spmv<5, RES>), g, t, 0, s, n_nodes, rowptr, colidx, vals, x, b, y); break;
    case 6: hipLaunchKernelGGL((k_spmv<6, RES>), g, t, 0, s, n_nodes, rowptr, colidx, vals, x, b, y); break;
    case 7: hipLaunchKernelGGL((k_spmv<7, RES>), g, t, 0, s, n_nodes, rowptr, colidx, vals, x, b, y); break;
    case 8: hipLaunchKernelGGL((k_spmv<8, RES>), g, t, 0, s, n_nodes, rowptr, colidx, vals, x, b, y); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_spmv(int ndof, int n_nodes, const int* rowptr, const int* colidx,
                       const double* vals, const double* x, double* y, hipStream_t s) {
  return spmv_any<false>(ndof, n_nodes, rowptr, colidx, vals, x, nullptr, y, s);
}

hipError_t launch_residual_vec(int ndof, int n_nodes, const int* rowptr, const int* colidx,
                               const double* vals, const double* x, const double* b, double* y,
                               hipStream_t s) {
  return spmv_any<true>(ndof, n_nodes, rowptr, colidx, vals, x, b, y, s);
}

__global__ void k_jacobi_inv(int ndof, int n_nodes, const int* __restrict__ diag_blk,
                             const double* __restrict__ vals, double* __restrict__ dinv) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= (long long)n_nodes * ndof) return;
  const int i = (int)(r / ndof), k = (int)(r % ndof);
  const double d = vals[(long long)diag_blk[i] * ndof * ndof + k * ndof + k];
  dinv[r] = 1.0 / d;
}

hipError_t launch_jacobi_inv(int ndof, int n_nodes, const int* diag_blk, const double* vals,
                             double* dinv, hipStream_t s) {
  const long long n = (long long)n_nodes * ndof;
  hipLaunchKernelGGL(k_jacobi_inv, dim3(grid_for(n)), dim3(kThreads), 0, s, ndof, n_nodes,
                     diag_blk, vals, dinv);
  return hipGetLastError();
}

// Block Jacobi: inverse of every node's ndof x ndof diagonal block by LU with
// partial pivoting (one thread per node).
template <int N>
__global__ void k_block_jacobi_inv(int n_nodes, const int* __restrict__ diag_blk,
                                   const double* __restrict__ vals, double* __restrict__ binv,
                                   int* __restrict__ status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  double A[N][N], Inv[N][N];
  const double* V = vals + (long long)diag_blk[i] * N * N;
#pragma unroll
  for (int r = 0; r < N; r++)
#pragma unroll
    for (int c = 0; c < N; c++) A[r][c] = V[r * N + c];
  if (!lu_inverse<double, N>(A, Inv)) atomicOr(status, 4);
#pragma unroll
  for (int r = 0; r < N; r++)
#pragma unroll
    for (int c = 0; c < N; c++) binv[(long long)i * N * N + r * N + c] = Inv[r][c];
}

hipError_t launch_block_jacobi_inv(int ndof, int n_nodes, const int* diag_blk, const double* vals,
                                   double* binv, int* status, hipStream_t s) {
  const dim3 g(grid_for(n_nodes)), t(kThreads);
  switch (ndof) {
    case 5: hipLaunchKernelGGL(k_block_jacobi_inv<5>, g, t, 0, s, n_nodes, diag_blk, vals, binv, status); break;
    case 6: hipLaunchKernelGGL(k_block_jacobi_inv<6>, g, t, 0, s, n_nodes, diag_blk, vals, binv, status); break;
    case 7: hipLaunchKernelGGL(k_block_jacobi_inv<7>, g, t, 0, s, n_nodes, diag_blk, vals, binv, status); break;
    case 8: hipLaunchKernelGGL(k_block_jacobi_inv<8>, g, t, 0, s, n_nodes, diag_blk, vals, binv, status); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// y = Binv x (ACC: y += Binv x), one thread per scalar row
template <int N, bool ACC>
__global__ void k_block_apply(int n_nodes, const double* __restrict__ binv,
                              const double* __restrict__ x, double* __restrict__ y) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= (long long)n_nodes * N) return;
  const long long i = r / N;
  const int k = (int)(r - i * N);
  const double* B = binv + i * N * N + k * N;
  const double* xx = x + i * N;
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < N; c++) s += B[c] * xx[c];
  y[r] = ACC ? y[r] + s : s;
}

hipError_t launch_block_apply(int ndof, int n_nodes, const double* binv, const double* x, double* y,
                              bool accumulate, hipStream_t s) {
  const dim3 g(grid_for((long long)n_nodes * ndof)), t(kThreads);
#define GMC_BA(N_)                                                                           \
  if (accumulate) hipLaunchKernelGGL((k_block_apply<N_, true>), g, t, 0, s, n_nodes, binv, x, y); \
  else hipLaunchKernelGGL((k_block_apply<N_, false>), g, t, 0, s, n_nodes, binv, x, y);
  switch (ndof) {
    case 5: GMC_BA(5) break;
    case 6: GMC_BA(6) break;
    case 7: GMC_BA(7) break;
    case 8: GMC_BA(8) break;
    default: return hipErrorInvalidValue;
  }
#undef GMC_BA
  return hipGetLastError();
}

int dot_blocks(long long n) {
  long long b = (n + 4 * kThreads - 1) / (4 * kThreads);
  if (b > 1024) b = 1024;
  if (b < 1) b = 1;
  return (int)b;
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (t < w) sh[t] += sh[t + w];
    __syncthreads();
  }
  return sh[0];
}

// blockIdx.y selects the basis vector; blockIdx.x a fixed strided chunk
__global__ __launch_bounds__(256) void k_multi_dot(const double* __restrict__ V, long long ld,
                                                   const double* __restrict__ w, long long n,
                                                   double* __restrict__ part) {
  __shared__ double sh[kThreads];
  const int i = blockIdx.y;
  const double* v = V + ld * i;
  double s = 0.0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x)
    s += v[t] * w[t];
  const double tot = block_sum(s, sh);
  if (threadIdx.x == 0) part[(long long)i * gridDim.x + blockIdx.x] = tot;
}

hipError_t launch_multi_dot(const double* V, long long ld, int m, const double* w, long long n,
                            double* part, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_multi_dot, dim3(dot_blocks(n), m), dim3(kThreads), 0, s, V, ld, w, n,
                     part);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_reduce_parts(const double* __restrict__ part, int nblk,
                                                      double* __restrict__ out) {
  __shared__ double sh[kThreads];
  const int i = blockIdx.x;
  double s = 0.0;
  for (int t = threadIdx.x; t < nblk; t += blockDim.x) s += part[(long long)i * nblk + t];
  const double tot = block_sum(s, sh);
  if (threadIdx.x == 0) out[i] = tot;
}

hipError_t launch_reduce_parts(const double* part, int m, int nblk, double* out, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_reduce_parts, dim3(m), dim3(kThreads), 0, s, part, nblk, out);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_multi_axpy(const double* __restrict__ V, long long ld,
                                                    int m, const double* __restrict__ h,
                                                    double* __restrict__ w, long long n) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double s = w[t];
  for (int i = 0; i < m; i++) s -= h[i] * V[ld * i + t];
  w[t] = s;
}

hipError_t launch_multi_axpy(const double* V, long long ld, int m, const double* h, double* w,
                             long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_multi_axpy, dim3(grid_for(n)), dim3(kThreads), 0, s, V, ld, m, h, w, n);
  return hipGetLastError();
}

__global__ void k_scale_div(const double* __restrict__ x, const double* __restrict__ dn,
                            double* __restrict__ y, long long n) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double d = dn[0];
  y[t] = (d != 0.0) ? x[t] / d : 0.0;
}

hipError_t launch_scale_div(const double* x, const double* dnorm, double* y, long long n,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_scale_div, dim3(grid_for(n)), dim3(kThreads), 0, s, x, dnorm, y, n);
  return hipGetLastError();
}

__global__ void k_scale_div_h(const double* __restrict__ x, double d, double* __restrict__ y,
                              long long n) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  y[t] = (d != 0.0) ? x[t] / d : 0.0;
}

hipError_t launch_scale_div_h(const double* x, double d, double* y, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_scale_div_h, dim3(grid_for(n)), dim3(kThreads), 0, s, x, d, y, n);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_update_x(const double* __restrict__ V, long long ld,
                                                  int m, const double* __restrict__ y,
                                                  const double* __restrict__ dinv,
                                                  double* __restrict__ x, long long n) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double s = 0.0;
  for (int i = 0; i < m; i++) s += V[ld * i + t] * y[i];
  x[t] += (dinv ? dinv[t] : 1.0) * s;
}

hipError_t launch_update_x(const double* V, long long ld, int m, const double* y,
                           const double* dinv, double* x, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_update_x, dim3(grid_for(n)), dim3(kThreads), 0, s, V, ld, m, y, dinv, x,
                     n);
  return hipGetLastError();
}

__global__ void k_mul(const double* __restrict__ a, const double* __restrict__ b,
                      double* __restrict__ y, long long n) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] = a[t] * b[t];
}

hipError_t launch_mul(const double* a, const double* b, double* y, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_mul, dim3(grid_for(n)), dim3(kThreads), 0, s, a, b, y, n);
  return hipGetLastError();
}

// model::dof_trimmer (model.hpp:176-184): std::max(lo, std::min(x, hi)),
// spelled with the comparisons of std::min/std::max so a NaN maps as there.
__device__ __forceinline__ double trim_value(double x, double lo, double hi) {
  const double m = (hi < x) ? hi : x;
  return (lo < m) ? m : lo;
}

__global__ void k_newton_update(double* __restrict__ I, const double* __restrict__ d,
                                long long n_nodes, int ndof, int trim) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_nodes * ndof) return;
  double v = I[t] + d[t];
  if ((int)(t % ndof) == trim) v = trim_value(v, 0.0, 1.0);
  I[t] = v;
}

hipError_t launch_newton_update(double* I, const double* d, long long n_nodes, int ndof, int trim,
                                hipStream_t s) {
  hipLaunchKernelGGL(k_newton_update, dim3(grid_for(n_nodes * ndof)), dim3(kThreads), 0, s, I, d,
                     n_nodes, ndof, trim);
  return hipGetLastError();
}

__global__ void k_trim(double* __restrict__ I, long long n_nodes, int ndof, int k, double lo,
                       double hi) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  I[i * ndof + k] = trim_value(I[i * ndof + k], lo, hi);
}

hipError_t launch_trim(double* I, long long n_nodes, int ndof, int k, double lo, double hi,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_trim, dim3(grid_for(n_nodes)), dim3(kThreads), 0, s, I, n_nodes, ndof, k, lo,
                     hi);
  return hipGetLastError();
}

__global__ void k_axpy(double alpha, const double* __restrict__ x, double* __restrict__ y,
                       long long n) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] += alpha * x[t];
}

hipError_t launch_axpy(double alpha, const double* x, double* y, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy, dim3(grid_for(n)), dim3(kThreads), 0, s, alpha, x, y, n);
  return hipGetLastError();
}

__global__ void k_f64_to_f32(const double* __restrict__ x, float* __restrict__ y, long long n) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] = (float)x[t];
}

hipError_t launch_copy_f64_to_f32(const double* x, float* y, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_f64_to_f32, dim3(grid_for(n)), dim3(kThreads), 0, s, x, y, n);
  return hipGetLastError();
}

}  // namespace gmc
