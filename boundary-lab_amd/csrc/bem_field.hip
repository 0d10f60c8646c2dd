// Field evaluation at observation points (polar arcs and the Fibonacci sphere).
//
// Replaces src/blab/solvers/julia_local/src/BeatEngineRocmField.jl:70-202 (weighted-source and
// thread-per-point field kernels) and BeatEngineCpuField.jl:1-45:
//   u(x) = sum_m [ dG/dn_y(x, y_m) p_w(m) - G(x, y_m) q_w(m) ],
//   p_w(m) = w_m sum_l phi_l(q) p[face_l], q_w(m) = w_m q[face], y_m the regular rule points (nq per face).
// With x / xy plane symmetry the sources of every image transform follow those of the fundamental domain
// (transform-major) and share its weighted values (BeatEngineCore.jl:1245-1280).
// Design: the transfer matrix is generated on the fly (VALU bound, never stored). A block holds 64 observation
// points (one per lane) and its four waves sweep disjoint slices of sources whose coordinates and weights are
// wave-uniform (scalar loads). Slices are reduced in a fixed order by k_field_reduce: deterministic.

#include "bem_kernels.hpp"

#define MAX_CH BEM_MAX_CH


template <typename T>
struct FieldParams {
  const SrcRec<T>* src;       // [M]
  const int4* src_face;       // [F] vertex ids per face (shared with assembly)
  const T* src_w;             // [M] weights w_q * 2A
  const cplx<T>* wsrc;        // [nf][n_ch][M][2] (p_w, q_w)
  const double* obs;          // [n_obs][3]
  const double* k;            // [nf]
  cplx<double>* partial;      // [nf][n_slices][n_ch][n_obs]
  int M, n_obs, n_ch, n_slices, slice_len;
  int M0;  // weighted sources per transform: source m uses the weights of source m % M0 (its original)
};

// weighted sources: p_w and q_w per (f, c, m); p is [nf][n_ch][ldp] FP64 boundary pressure, q likewise per face
template <typename T>
__global__ void k_field_sources(const int4* fv, const T* src_w, const cplx<double>* p, long long ldp,
                                long long strideP, const cplx<double>* q, long long ldq, long long strideQ,
                                cplx<T>* wsrc, int M, int n_ch, int nq) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y, f = blockIdx.z;
  if (m >= M) return;
  const int face = m / nq, qp = m % nq;  // nq rule points per face
  const int4 v = fv[face];
  const cplx<double>* pp = p + (size_t)f * strideP + (size_t)c * ldp;
  const double b0 = qrule_phi(nq, 0, qp), b1 = qrule_phi(nq, 1, qp), b2 = qrule_phi(nq, 2, qp);
  const double w = double(src_w[m]);
  const cplx<double> p0 = pp[v.x], p1 = pp[v.y], p2 = pp[v.z];
  const double pr = (b0 * p0.re + b1 * p1.re + b2 * p2.re) * w;
  const double pi = (b0 * p0.im + b1 * p1.im + b2 * p2.im) * w;
  const cplx<double> qv = q[(size_t)f * strideQ + (size_t)c * ldq + face];
  cplx<T>* o = wsrc + (((size_t)f * n_ch + c) * M + m) * 2;
  o[0] = mk<T>(T(pr), T(pi));
  o[1] = mk<T>(T(qv.re * w), T(qv.im * w));
}

template <typename T>
__global__ __launch_bounds__(256) void k_field(FieldParams<T> p) {
  // w through readfirstlane: the wave's source range is then provably uniform and the source records and
  // weights come in through scalar loads
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int o = blockIdx.x * 64 + lane;
  const int slice = blockIdx.y, f = blockIdx.z;
  const double kd = p.k[f];
  const T k = T(kd);
  const T kappa = T(kd / (2.0 * BEM_PI));
  T ox = 0, oy = 0, oz = 0;
  if (o < p.n_obs) {
    ox = T(p.obs[3 * o]);
    oy = T(p.obs[3 * o + 1]);
    oz = T(p.obs[3 * o + 2]);
  }
  T ur[MAX_CH], ui[MAX_CH];
#pragma unroll
  for (int c = 0; c < MAX_CH; ++c) ur[c] = ui[c] = T(0);
  const int m0 = slice * p.slice_len;
  const int m1 = min(p.M, m0 + p.slice_len);
  const int sub = (m1 - m0 + 3) / 4;
  const int s0 = m0 + w * sub, s1 = min(m1, s0 + sub);
  for (int m = s0; m < s1; ++m) {
    const SrcRec<T> s = p.src[m];
    const T dx = s.y[0] - ox, dy = s.y[1] - oy, dz = s.y[2] - oz;
    const T r2 = dx * dx + dy * dy + dz * dz;
    const T rinv = Phase<T>::rsqrt_(r2);
    const T r = r2 * rinv;
    T sn, cs;
    Phase<T>::sincos_kr(kappa, k, r, &sn, &cs);
    const T sc = rinv * T(BEM_INV_FOUR_PI);
    const T gr = cs * sc, gi = sn * sc;
    const T proj = (dx * s.n[0] + dy * s.n[1] + dz * s.n[2]) * rinv;
    // dl = G (ik - 1/r) proj
    const T dr = (-gr * rinv - gi * k) * proj;
    const T di = (gr * k - gi * rinv) * proj;
#pragma unroll
    for (int c = 0; c < MAX_CH; ++c) {
      if (c < p.n_ch) {
        const cplx<T>* ws = p.wsrc + (((size_t)f * p.n_ch + c) * p.M0 + (m % p.M0)) * 2;
        const cplx<T> pw = ws[0], qw = ws[1];
        ur[c] += dr * pw.re - di * pw.im - (gr * qw.re - gi * qw.im);
        ui[c] += dr * pw.im + di * pw.re - (gr * qw.im + gi * qw.re);
      }
    }
  }
  __shared__ T red[4][MAX_CH][2][64];
#pragma unroll
  for (int c = 0; c < MAX_CH; ++c) {
    red[w][c][0][lane] = ur[c];
    red[w][c][1][lane] = ui[c];
  }
  __syncthreads();
  if (w == 0 && o < p.n_obs) {
    for (int c = 0; c < p.n_ch; ++c) {
      double sr = 0.0, si = 0.0;
      for (int q = 0; q < 4; ++q) {
        sr += double(red[q][c][0][lane]);
        si += double(red[q][c][1][lane]);
      }
      p.partial[(((size_t)f * p.n_slices + slice) * p.n_ch + c) * p.n_obs + o] = mk<double>(sr, si);
    }
  }
}

__global__ void k_field_reduce(const cplx<double>* partial, cplx<double>* out, long long ldo, long long strideO,
                               int n_obs, int n_ch, int n_slices) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y, f = blockIdx.z;
  if (o >= n_obs) return;
  double sr = 0.0, si = 0.0;
  for (int s = 0; s < n_slices; ++s) {
    const cplx<double> v = partial[(((size_t)f * n_slices + s) * n_ch + c) * n_obs + o];
    sr += v.re;
    si += v.im;
  }
  out[(size_t)f * strideO + (size_t)c * ldo + o] = mk<double>(sr, si);
}

template <typename T>
hipError_t field_batch(const SrcRec<T>* src, const int4* fv, const T* src_w, const double* obs, int M, int M0,
                       int n_obs, const double* k, const cplx<double>* p, long long ldp, long long strideP,
                       const cplx<double>* q, long long ldq, long long strideQ, int n_ch, int nf, cplx<T>* wsrc,
                       cplx<double>* partial, int n_slices, cplx<double>* out, long long ldo, long long strideO,
                       int nq, hipStream_t st) {
  // weighted sources of the fundamental domain (M0 = nq F); image sources (M = R * M0) reuse them
  hipLaunchKernelGGL(k_field_sources<T>, dim3((M0 + 255) / 256, n_ch, nf), dim3(256), 0, st, fv, src_w, p, ldp,
                     strideP, q, ldq, strideQ, wsrc, M0, n_ch, nq);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  FieldParams<T> fp;
  fp.src = src;
  fp.src_face = fv;
  fp.src_w = src_w;
  fp.wsrc = wsrc;
  fp.obs = obs;
  fp.k = k;
  fp.partial = partial;
  fp.M = M;
  fp.M0 = M0;
  fp.n_obs = n_obs;
  fp.n_ch = n_ch;
  fp.n_slices = n_slices;
  fp.slice_len = (M + n_slices - 1) / n_slices;
  hipLaunchKernelGGL(k_field<T>, dim3((n_obs + 63) / 64, n_slices, nf), dim3(256), 0, st, fp);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_field_reduce, dim3((n_obs + 255) / 256, n_ch, nf), dim3(256), 0, st, partial, out, ldo,
                     strideO, n_obs, n_ch, n_slices);
  return hipGetLastError();
}

// Neumann data of every channel: q[f][c][j] = scale[f] * qunit[c][j] (scale = i rho omega per frequency).
__global__ void k_neumann(const cplx<double>* scale, const cplx<double>* qunit, cplx<double>* q, int n_ch, int F) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y, f = blockIdx.z;
  if (j >= F) return;
  q[((size_t)f * n_ch + c) * F + j] = cmul(scale[f], qunit[(size_t)c * F + j]);
}

hipError_t neumann_batch(const cplx<double>* scale, const cplx<double>* qunit, cplx<double>* q, int n_ch, int F,
                         int nf, hipStream_t st) {
  hipLaunchKernelGGL(k_neumann, dim3((F + 255) / 256, n_ch, nf), dim3(256), 0, st, scale, qunit, q, n_ch, F);
  return hipGetLastError();
}

template hipError_t field_batch<float>(const SrcRec<float>*, const int4*, const float*, const double*, int, int, int,
                                       const double*, const cplx<double>*, long long, long long,
                                       const cplx<double>*, long long, long long, int, int, cplx<float>*,
                                       cplx<double>*, int, cplx<double>*, long long, long long, int, hipStream_t);
template hipError_t field_batch<double>(const SrcRec<double>*, const int4*, const double*, const double*, int, int, int,
                                        const double*, const cplx<double>*, long long, long long,
                                        const cplx<double>*, long long, long long, int, int, cplx<double>*,
                                        cplx<double>*, int, cplx<double>*, long long, long long, int, hipStream_t);
