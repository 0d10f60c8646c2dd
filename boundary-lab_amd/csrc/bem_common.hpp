// Shared device/host helpers for the MI355X (gfx950) Helmholtz BEM kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define BEM_PI 3.14159265358979323846
#define BEM_INV_FOUR_PI 0.07957747154594767  // 1 / (4 pi)

template <typename T>
struct __align__(2 * sizeof(T)) cplx {
  T re, im;
};

template <typename T>
__host__ __device__ __forceinline__ cplx<T> mk(T r, T i) {
  cplx<T> c;
  c.re = r;
  c.im = i;
  return c;
}
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return mk<T>(a.re + b.re, a.im + b.im); }
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return mk<T>(a.re - b.re, a.im - b.im); }
template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return mk<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
// a - b*c
template <typename T>
__device__ __forceinline__ cplx<T> cfms(cplx<T> a, cplx<T> b, cplx<T> c) {
  return mk<T>(a.re - b.re * c.re + b.im * c.im, a.im - b.re * c.im - b.im * c.re);
}
// a + b*c
template <typename T>
__device__ __forceinline__ cplx<T> cfma(cplx<T> a, cplx<T> b, cplx<T> c) {
  return mk<T>(a.re + b.re * c.re - b.im * c.im, a.im + b.re * c.im + b.im * c.re);
}
template <typename T>
__device__ __forceinline__ cplx<T> cdiv(cplx<T> a, cplx<T> b) {
  // Smith-free scaled division; pivots are O(1) after partial pivoting so plain form is adequate.
  T d = b.re * b.re + b.im * b.im;
  T inv = T(1) / d;
  return mk<T>((a.re * b.re + a.im * b.im) * inv, (a.im * b.re - a.re * b.im) * inv);
}
template <typename T>
__device__ __forceinline__ T cabs1(cplx<T> a) {  // |re| + |im|, the LAPACK i?amax measure
  return fabs(a.re) + fabs(a.im);
}

// Trigonometry of the Helmholtz phase k*r.
//  float : hardware v_sin_f32 / v_cos_f32 take revolutions; the phase is formed as (k/2pi)*r and
//          reduced to [0,1) with v_fract_f32 before the call.
//  double: library sincos in radians.
// sin and cos of a non-negative FP64 phase of moderate size (k r below 2^20 rad: element-scale distances):
// quadrant reduction q = rint(x 2/pi) with pi/2 split in two doubles (fused multiply-adds keep it exact at these
// magnitudes), then the Taylor series of sin and cos on |y| <= pi/4 to degree 15 / 16 (truncation below 2e-17).
__device__ __forceinline__ void sincos_phase(double x, double* s, double* c) {
  constexpr double kTwoOverPi = 0.6366197723675814;
  constexpr double kPio2Hi = 1.5707963267948966, kPio2Lo = 6.123233995736766e-17;
  const double q = rint(x * kTwoOverPi);
  double y = fma(-q, kPio2Hi, x);
  y = fma(-q, kPio2Lo, y);
  const double y2 = y * y;
  // 1/n! coefficients, Horner form
  double ps = -1.0 / 1307674368000.0;                 // -1/15!
  ps = fma(ps, y2, 1.0 / 6227020800.0);               // +1/13!
  ps = fma(ps, y2, -1.0 / 39916800.0);                // -1/11!
  ps = fma(ps, y2, 1.0 / 362880.0);                   // +1/9!
  ps = fma(ps, y2, -1.0 / 5040.0);                    // -1/7!
  ps = fma(ps, y2, 1.0 / 120.0);                      // +1/5!
  ps = fma(ps, y2, -1.0 / 6.0);                       // -1/3!
  const double sy = fma(ps * y2, y, y);
  double pc = 1.0 / 20922789888000.0;                 // +1/16!
  pc = fma(pc, y2, -1.0 / 87178291200.0);             // -1/14!
  pc = fma(pc, y2, 1.0 / 479001600.0);                // +1/12!
  pc = fma(pc, y2, -1.0 / 3628800.0);                 // -1/10!
  pc = fma(pc, y2, 1.0 / 40320.0);                    // +1/8!
  pc = fma(pc, y2, -1.0 / 720.0);                     // -1/6!
  pc = fma(pc, y2, 1.0 / 24.0);                       // +1/4!
  pc = fma(pc, y2, -0.5);                             // -1/2!
  const double cy = fma(pc, y2, 1.0);
  const int iq = (int)q & 3;
  const double ss = (iq & 1) ? cy : sy, cc = (iq & 1) ? sy : cy;
  *s = (iq & 2) ? -ss : ss;
  *c = ((iq + 1) & 2) ? -cc : cc;
}

template <typename T>
struct Phase;
template <>
struct Phase<float> {
  __device__ __forceinline__ static void sincos_kr(float kappa_rev, float k, float r, float* s, float* c) {
    (void)k;
    float ph = __builtin_amdgcn_fractf(kappa_rev * r);
    *s = __builtin_amdgcn_sinf(ph);
    *c = __builtin_amdgcn_cosf(ph);
  }
  __device__ __forceinline__ static float rsqrt_(float x) { return x > 0.f ? __builtin_amdgcn_rsqf(x) : 0.f; }
};
template <>
struct Phase<double> {
  __device__ __forceinline__ static void sincos_kr(double kappa_rev, double k, double r, double* s, double* c) {
    (void)kappa_rev;
    sincos_phase(k * r, s, c);
  }
  __device__ __forceinline__ static double rsqrt_(double x) {
    if (!(x > 0.0)) return 0.0;
    double y = __builtin_amdgcn_rsq(x);  // hardware estimate, two Newton steps to full FP64 accuracy
    y = y * fma(-0.5 * x * y, y, 1.5);
    return y * fma(-0.5 * x * y, y, 1.5);
  }
};

// Wave-level reductions (wave64).
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Order-4 triangle rule (7-digit-asymmetric constants as tabulated by the reference,
// BeatEngineCore.jl:517-549). Index q = 0..5.
__host__ __device__ constexpr double rule_xi(int q) {
  return q == 0 ? 0.4459484909159651
         : q == 1 ? 0.0915762135097710
         : q == 2 ? 0.1081030181680700
         : q == 3 ? 0.4459484909159651
         : q == 4 ? 0.8168475729804590
                  : 0.0915762135097710;
}
__host__ __device__ constexpr double rule_eta(int q) {
  return q == 0 ? 0.4459484909159651
         : q == 1 ? 0.0915762135097700
         : q == 2 ? 0.4459484909159651
         : q == 3 ? 0.1081030181680700
         : q == 4 ? 0.0915762135097700
                  : 0.8168475729804580;
}
__host__ __device__ constexpr double rule_w(int q) {
  return 0.5 * (q == 0 || q == 2 || q == 3 ? 0.2233815896780110 : 0.1099517436553220);
}
// P1 basis value phi_l at rule point q.
__host__ __device__ constexpr double rule_phi(int l, int q) {
  return l == 0 ? 1.0 - rule_xi(q) - rule_eta(q) : (l == 1 ? rule_xi(q) : rule_eta(q));
}

// The regular triangle rules a mesh can be integrated with, by point count (BeatEngineCore.jl:517-549):
// NQ = 6 the order-4 rule above, NQ = 3 the order-2 rule (points (1/6, 1/6), (2/3, 1/6), (1/6, 2/3), weights 1/6),
// NQ = 1 the centroid rule (weight 1/2). Weights sum to 1/2, the reference triangle's area.
template <int NQ>
struct QRule;
template <>
struct QRule<6> {
  __host__ __device__ static constexpr double xi(int q) { return rule_xi(q); }
  __host__ __device__ static constexpr double eta(int q) { return rule_eta(q); }
  __host__ __device__ static constexpr double w(int q) { return rule_w(q); }
};
template <>
struct QRule<3> {
  __host__ __device__ static constexpr double xi(int q) { return q == 1 ? 2.0 / 3.0 : 1.0 / 6.0; }
  __host__ __device__ static constexpr double eta(int q) { return q == 2 ? 2.0 / 3.0 : 1.0 / 6.0; }
  __host__ __device__ static constexpr double w(int q) { return (void)q, 1.0 / 6.0; }
};
template <>
struct QRule<1> {
  __host__ __device__ static constexpr double xi(int q) { return (void)q, 1.0 / 3.0; }
  __host__ __device__ static constexpr double eta(int q) { return (void)q, 1.0 / 3.0; }
  __host__ __device__ static constexpr double w(int q) { return (void)q, 0.5; }
};
template <int NQ>
__host__ __device__ constexpr double qphi(int l, int q) {
  return l == 0 ? 1.0 - QRule<NQ>::xi(q) - QRule<NQ>::eta(q) : (l == 1 ? QRule<NQ>::xi(q) : QRule<NQ>::eta(q));
}
// w_q and phi_l(q) of the NQ-point rule at run time (q < NQ).
__host__ __device__ inline double qrule_w(int nq, int q) {
  return nq == 6 ? QRule<6>::w(q) : nq == 3 ? QRule<3>::w(q) : QRule<1>::w(q);
}
__host__ __device__ inline double qrule_phi(int nq, int l, int q) {
  return nq == 6 ? qphi<6>(l, q) : nq == 3 ? qphi<3>(l, q) : qphi<1>(l, q);
}
