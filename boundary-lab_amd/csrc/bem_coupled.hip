// Dense blocks of the coupled FEM-BEM system on the device.
//
// Replaces the reference's ROCm assembly of the coupled matrix
//   src/blab/solvers/julia_local/src/BeatEngineRocmOperators.jl:40-108  (sparse -> dense scatter kernel)
//   src/blab/solvers/julia_local/src/BeatEngineCoupled.jl:1264-1406     (BEM interface / motion blocks as
//                                                                        dense operator-times-flux products)
//   src/blab/solvers/julia_local/src/BeatEngineCoupled.jl:1917-2080     (block placement into the matrix)
// The reference forms all four dense Galerkin operators and multiplies the RHS operator by the dense flux
// matrices with rocBLAS; here the RHS operator is never formed: k_rhs_cols (bem_assembly.hip) integrates the
// (test element, source face) blocks of the faces that carry a column, and k_cols_gather combines them per P1
// row and per sparse column, writing straight into the coupled matrix. Each output entry is written by one
// thread in a fixed order: deterministic.

#include "bem_kernels.hpp"

template <typename T>
__global__ __launch_bounds__(256) void k_cols_gather(const cplx<double>* part, int F, int n_src,
                                                     const int* vstar_ptr, const int* vstar_row, const double* roww,
                                                     int N, int n_cols, const int* col_ptr, const int* col_idx,
                                                     const cplx<double>* col_w, const cplx<double>* alpha,
                                                     cplx<T>* out, long long ldo, long long strideO) {
  const int i = blockIdx.y;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int f = blockIdx.z;
  if (c >= n_cols) return;
  const cplx<double>* pf = part + (size_t)f * 3 * F * n_src;
  const int s0 = vstar_ptr[i], s1 = vstar_ptr[i + 1];
  const int e0 = col_ptr[c], e1 = col_ptr[c + 1];
  double re = 0.0, im = 0.0;
  for (int r = s0; r < s1; ++r) {
    const cplx<double>* row = pf + (size_t)vstar_row[r] * n_src;
    for (int e = e0; e < e1; ++e) {
      const cplx<double> v = row[col_idx[e]];
      const cplx<double> w = col_w[e];
      re += v.re * w.re - v.im * w.im;
      im += v.re * w.im + v.im * w.re;
    }
  }
  const double wi = roww ? roww[i] : 1.0;
  const cplx<double> a = alpha[f];
  re *= wi;
  im *= wi;
  out[(size_t)f * strideO + (size_t)i * ldo + c] = mk<T>(T(re * a.re - im * a.im), T(re * a.im + im * a.re));
}

template <typename T>
__global__ __launch_bounds__(256) void k_scatter_coo(long long nnz, const int* rows, const int* cols,
                                                     const cplx<double>* vals, cplx<double> alpha, int add,
                                                     cplx<T>* out, long long ldo) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  const cplx<double> v = vals[e];
  cplx<T>* o = out + (size_t)rows[e] * ldo + cols[e];
  const T vr = T(v.re * alpha.re - v.im * alpha.im), vi = T(v.re * alpha.im + v.im * alpha.re);
  if (add)
    *o = mk<T>(o->re + vr, o->im + vi);
  else
    *o = mk<T>(vr, vi);
}

template <typename T>
__global__ __launch_bounds__(256) void k_dense_block(int m, int n, const cplx<double>* src, long long lds,
                                                     cplx<double> alpha, int add, cplx<T>* out, long long ldo) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= n || i >= m) return;
  cplx<double> v = alpha;
  if (src) {
    const cplx<double> s = src[(size_t)i * lds + j];
    v = mk<double>(s.re * alpha.re - s.im * alpha.im, s.re * alpha.im + s.im * alpha.re);
  }
  cplx<T>* o = out + (size_t)i * ldo + j;
  if (add)
    *o = mk<T>(o->re + T(v.re), o->im + T(v.im));
  else
    *o = mk<T>(T(v.re), T(v.im));
}

// Row equilibration by powers of two: row i is multiplied by s_i = 2^-e_i with max_j (|re| + |im|) of the row in
// [2^(e_i - 1), 2^e_i), so the largest entry lands in [1/2, 1) and no entry is rounded. scale[i] = s_i (1 for an all-
// zero row). One workgroup per row; the maximum is reduced in a fixed order.
template <typename T>
__global__ __launch_bounds__(256) void k_row_equilibrate(cplx<T>* A, long long lda, int ncols, double* scale) {
  const int i = blockIdx.x;
  const int tid = threadIdx.x;
  __shared__ double red[256];
  cplx<T>* row = A + (size_t)i * lda;
  double m = 0.0;
  for (int j = tid; j < ncols; j += 256) m = fmax(m, fabs((double)row[j].re) + fabs((double)row[j].im));
  red[tid] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (tid < o) red[tid] = fmax(red[tid], red[tid + o]);
    __syncthreads();
  }
  const double mx = red[0];
  int e = 0;
  if (mx > 0.0 && isfinite(mx)) (void)frexp(mx, &e);
  const double s = ldexp(1.0, -e);
  const T st = T(s);
  for (int j = tid; j < ncols; j += 256) row[j] = mk<T>(row[j].re * st, row[j].im * st);
  if (tid == 0) scale[i] = s;
}

template <typename T>
hipError_t row_equilibrate(cplx<T>* A, long long lda, int n, int ncols, double* scale, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_row_equilibrate<T>, dim3(n), dim3(256), 0, st, A, lda, ncols, scale);
  return hipGetLastError();
}

template <typename T>
hipError_t cols_gather(const cplx<double>* part, int F, int n_src, const int* vstar_ptr, const int* vstar_row,
                       const double* roww, int N, int n_cols, const int* col_ptr, const int* col_idx,
                       const cplx<double>* col_w, const cplx<double>* alpha, cplx<T>* out, long long ldo,
                       long long strideO, int nf, hipStream_t st) {
  if (N <= 0 || n_cols <= 0 || nf <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_cols_gather<T>, dim3((n_cols + 255) / 256, N, nf), dim3(256), 0, st, part, F, n_src,
                     vstar_ptr, vstar_row, roww, N, n_cols, col_ptr, col_idx, col_w, alpha, out, ldo, strideO);
  return hipGetLastError();
}

template <typename T>
hipError_t scatter_coo(long long nnz, const int* rows, const int* cols, const cplx<double>* vals,
                       cplx<double> alpha, int add, cplx<T>* out, long long ldo, hipStream_t st) {
  if (nnz <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scatter_coo<T>, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, st, nnz, rows, cols, vals,
                     alpha, add, out, ldo);
  return hipGetLastError();
}

template <typename T>
hipError_t dense_block(int m, int n, const cplx<double>* src, long long lds, cplx<double> alpha, int add,
                       cplx<T>* out, long long ldo, hipStream_t st) {
  if (m <= 0 || n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_dense_block<T>, dim3((n + 255) / 256, m), dim3(256), 0, st, m, n, src, lds, alpha, add, out,
                     ldo);
  return hipGetLastError();
}

template hipError_t cols_gather<float>(const cplx<double>*, int, int, const int*, const int*, const double*, int, int,
                                       const int*, const int*, const cplx<double>*, const cplx<double>*,
                                       cplx<float>*, long long, long long, int, hipStream_t);
template hipError_t cols_gather<double>(const cplx<double>*, int, int, const int*, const int*, const double*, int,
                                        int, const int*, const int*, const cplx<double>*, const cplx<double>*,
                                        cplx<double>*, long long, long long, int, hipStream_t);
template hipError_t scatter_coo<float>(long long, const int*, const int*, const cplx<double>*, cplx<double>, int,
                                       cplx<float>*, long long, hipStream_t);
template hipError_t scatter_coo<double>(long long, const int*, const int*, const cplx<double>*, cplx<double>, int,
                                        cplx<double>*, long long, hipStream_t);
template hipError_t row_equilibrate<float>(cplx<float>*, long long, int, int, double*, hipStream_t);
template hipError_t row_equilibrate<double>(cplx<double>*, long long, int, int, double*, hipStream_t);
template hipError_t dense_block<float>(int, int, const cplx<double>*, long long, cplx<double>, int, cplx<float>*,
                                       long long, hipStream_t);
template hipError_t dense_block<double>(int, int, const cplx<double>*, long long, cplx<double>, int, cplx<double>*,
                                        long long, hipStream_t);
