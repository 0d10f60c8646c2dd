// Burton-Miller Galerkin assembly for gfx950.
//
// Replaces the reference's colour-pair regular kernels
//   src/blab/solvers/julia_local/src/BeatEngineRocmPairKernels.jl:28-341 (2*C^2 launches, 4 dense operators)
// singular block + gather kernels
//   src/blab/solvers/julia_local/src/BeatEngineRocmSingular.jl:84-374
// and the LHS broadcast / RHS gemv formation
//   src/blab/solvers/julia_local/src/BeatEngineRocmOperators.jl:110-121,239-240
// with four kernels that produce only what the solve needs: A = 1/2 M11 - D + (i/k) H and
// b = -(S + (i/k)(D* + 1/2 M10)) q for every channel, for a batch of wavenumbers at once.
//
// Data flow per frequency f:
//   k_singular    (T)      one lane per vertex-sharing element pair and group of wavenumbers: Duffy-rule blocks
//                          of the LHS (9 complex) and of the RHS operator (3 complex), written as FP64, identity
//                          terms folded into coincident pairs (FP32 sums for mixed batches unless FP64 is asked).
//   k_regular_lhs (T)      one wave per test element t; lanes sweep trial elements in locality-ordered chunks of
//                          64. Each lane integrates its (t,s) pair with the 6x6 order-4 rule into a 3x3 block
//                          (or picks up the FP64 singular block for adjacent pairs). Blocks go to LDS, then lanes
//                          switch to "slot" ownership: each slot (a trial vertex of the chunk) sums the block
//                          columns that touch it, in a fixed order, into the vertex's column of the three
//                          test-element strip rows (stored by its first chunk, accumulated by later ones).
//                          Only this wave touches its strip rows: no atomics, no colouring, deterministic.
//   k_rhs_part    (T)      one wave per test element: sum over radiating faces of the RHS-operator blocks times
//                          the unit channel drives.
//   k_gather      (T)      one block per P1 row i: A[i, j] = sum over (t,l) in star(i) of strip[3t+l][j].
//
// Regular-pair arithmetic follows BeatEngineCpuAssembly.jl:278-373 / BeatEngineRocmPairKernels.jl:100-150:
//   G = e^{ikr}/(4 pi r), grad = G (ik - 1/r) r_hat, D uses the trial normal, D* uses -r.n_test,
//   H = (curl_t.curl_s - k^2 phi psi n_t.n_s) G, weight w_a w_b 4 A_t A_s, pairs sharing a vertex are singular.
// The 3x3 block is contracted in two stages using the tensor structure of the rule:
//   inner over trial points b with constants psi_m(b) w_b, outer over test points a with phi_l(a) w_a.

#include "bem_kernels.hpp"

#define SLOT_STRIDE BEM_SLOT_STRIDE
#define MAX_CH BEM_MAX_CH



// phi_l(a) * w_a for the NQ points of a regular rule (index l * NQ + a), read from constant memory by the rolled
// test-point loops.
template <int NQ>
struct RuleWeights {
  double v[3 * NQ];
  constexpr RuleWeights() : v() {
    for (int l = 0; l < 3; ++l)
      for (int a = 0; a < NQ; ++a) v[l * NQ + a] = qphi<NQ>(l, a) * QRule<NQ>::w(a);
  }
};
template <int NQ>
struct RuleWeightsF {
  float v[3 * NQ];
  constexpr RuleWeightsF() : v() {
    for (int i = 0; i < 3 * NQ; ++i) v[i] = float(RuleWeights<NQ>().v[i]);
  }
};
__constant__ RuleWeights<6> c_phiw6 = RuleWeights<6>();
__constant__ RuleWeights<3> c_phiw3 = RuleWeights<3>();
__constant__ RuleWeights<1> c_phiw1 = RuleWeights<1>();
__constant__ RuleWeightsF<6> c_phiw6_f = RuleWeightsF<6>();
__constant__ RuleWeightsF<3> c_phiw3_f = RuleWeightsF<3>();
__constant__ RuleWeightsF<1> c_phiw1_f = RuleWeightsF<1>();
template <int NQ>
__device__ __forceinline__ double phiw(int i) {
  if constexpr (NQ == 6) return c_phiw6.v[i];
  else if constexpr (NQ == 3) return c_phiw3.v[i];
  else return c_phiw1.v[i];
}
template <int NQ>
__device__ __forceinline__ float phiw_f(int i) {
  if constexpr (NQ == 6) return c_phiw6_f.v[i];
  else if constexpr (NQ == 3) return c_phiw3_f.v[i];
  else return c_phiw1_f.v[i];
}

// 3x3 LHS block of a regular (t,s) pair, scaled and combined: C += J/(4pi) (-D + (i/k) H).
// UNROLL_A = NQ fully unrolls the test-point loop (FP32: test points live in SGPRs); UNROLL_A = 1 keeps it
// rolled with wave-uniform scalar loads of the test point (FP64: halves register pressure, no spills).
// NQ: points of the regular rule (6, 3 or 1).
template <typename T, int UNROLL_A, int NQ>
__device__ __forceinline__ void regular_block(const T* __restrict__ xq, const T (&nt)[3], const T (&ct)[9], T a2t,
                                              const ElemRec<T>& se, T k, T kappa, cplx<T> (&C)[9]) {
  T y[3 * NQ];
#pragma unroll
  for (int i = 0; i < 3 * NQ; ++i) y[i] = se.q[i];
  const T nsx = se.n[0], nsy = se.n[1], nsz = se.n[2];
  T DMr[9], DMi[9], GMr[9], GMi[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) DMr[i] = DMi[i] = GMr[i] = GMi[i] = T(0);
#pragma unroll UNROLL_A
  for (int a = 0; a < NQ; ++a) {
    const T xa = xq[3 * a + 0], ya = xq[3 * a + 1], za = xq[3 * a + 2];
    T tDr[3] = {0, 0, 0}, tDi[3] = {0, 0, 0}, tGr[3] = {0, 0, 0}, tGi[3] = {0, 0, 0};
#pragma unroll
    for (int b = 0; b < NQ; ++b) {
      const T dx = y[3 * b + 0] - xa;
      const T dy = y[3 * b + 1] - ya;
      const T dz = y[3 * b + 2] - za;
      const T r2 = dx * dx + dy * dy + dz * dz;
      const T rinv = Phase<T>::rsqrt_(r2);
      const T r = r2 * rinv;
      T sn, cs;
      Phase<T>::sincos_kr(kappa, k, r, &sn, &cs);
      const T gr = cs * rinv, gi = sn * rinv;  // 4 pi G
      const T hr = -gr * rinv - gi * k;        // 4 pi G (ik - 1/r)
      const T hi = gr * k - gi * rinv;
      const T td = (dx * nsx + dy * nsy + dz * nsz) * rinv;
      const T dgr = hr * td, dgi = hi * td;
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        const T cw = T(qphi<NQ>(m, b) * QRule<NQ>::w(b));
        tDr[m] += cw * dgr;
        tDi[m] += cw * dgi;
        tGr[m] += cw * gr;
        tGi[m] += cw * gi;
      }
    }
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      const T wa = T(phiw<NQ>(l * NQ + a));
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        DMr[3 * l + m] += wa * tDr[m];
        DMi[3 * l + m] += wa * tDi[m];
        GMr[3 * l + m] += wa * tGr[m];
        GMi[3 * l + m] += wa * tGi[m];
      }
    }
  }
  T gsr = 0, gsi = 0;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    gsr += GMr[i];
    gsi += GMi[i];
  }
  const T nn = nt[0] * nsx + nt[1] * nsy + nt[2] * nsz;
  const T k2nn = k * k * nn;
  const T scale = a2t * se.a2 * T(BEM_INV_FOUR_PI);
  const T kinv = T(1) / k;
#pragma unroll
  for (int l = 0; l < 3; ++l) {
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const T cp = ct[3 * l + 0] * se.curl[3 * m + 0] + ct[3 * l + 1] * se.curl[3 * m + 1] +
                   ct[3 * l + 2] * se.curl[3 * m + 2];
      const T Hr = cp * gsr - k2nn * GMr[3 * l + m];
      const T Hi = cp * gsi - k2nn * GMi[3 * l + m];
      // -D + (i/k) H, accumulated (image transforms add into the same block)
      C[3 * l + m] = cadd(C[3 * l + m], mk<T>(scale * (-DMr[3 * l + m] - Hi * kinv), scale * (-DMi[3 * l + m] + Hr * kinv)));
    }
  }
}

// Orbit structure of the order-4 rule: points {0,2,3} share one weight, points {1,4,5} the other, and within an
// orbit basis function m takes one value at its own "special" point ({0,2,3}[m], {1,4,5}[m]) and a common value
// at the other two. Hence psi_m(b) w_b = base(orbit) + [b special for m] * sp(orbit) (to the last digit of the
// tabulated rule), and the trial contraction of six values costs 4 adds + 8 FMAs instead of 18 FMAs.
constexpr double kBaseA = rule_phi(1, 0) * rule_w(0);
constexpr double kSpA = rule_phi(0, 0) * rule_w(0) - kBaseA;
constexpr double kBaseB = rule_phi(1, 1) * rule_w(1);
constexpr double kSpB = rule_phi(0, 1) * rule_w(1) - kBaseB;

__device__ __forceinline__ void trial_contract(const float (&x)[6], float (&t)[3]) {
  const float base = float(kBaseA) * (x[0] + x[2] + x[3]) + float(kBaseB) * (x[1] + x[4] + x[5]);
  t[0] = base + float(kSpA) * x[0] + float(kSpB) * x[1];
  t[1] = base + float(kSpA) * x[2] + float(kSpB) * x[4];
  t[2] = base + float(kSpA) * x[3] + float(kSpB) * x[5];
}
// Trial contraction of the NQ-point rule: the orbit form above for NQ = 6, plain weighted sums otherwise.
template <int NQ>
__device__ __forceinline__ void trial_contract_n(const float (&x)[NQ], float (&t)[3]) {
  if constexpr (NQ == 6) {
    trial_contract(x, t);
  } else {
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      float acc = 0.f;
#pragma unroll
      for (int b = 0; b < NQ; ++b) acc += float(qphi<NQ>(m, b) * QRule<NQ>::w(b)) * x[b];
      t[m] = acc;
    }
  }
}

// FP32 block for NFW wavenumbers at once: the pair geometry (r, 1/r) is shared by the frequencies. The trial
// element is flat, so (y_b - x_a).n_s is the same for the six trial points of a test point: the double-layer
// kernel is u (y_b - x_a).n_s with u = 4 pi G (ik - 1/r) / r, and the n_s offset scales the trial-contracted sum
// once per test point. The points of a regular pair (no shared vertex, images included) are distinct.
template <int NFW, int NQ>
__device__ __forceinline__ void regular_block_f32(const float* __restrict__ xq, const float (&nt)[3],
                                                  const float (&ct)[9], float a2t, const ElemRec<float>& se,
                                                  const float (&k)[NFW], const float (&kappa)[NFW],
                                                  cplx<float> (&C)[NFW][9]) {
  float y[3 * NQ];
#pragma unroll
  for (int i = 0; i < 3 * NQ; ++i) y[i] = se.q[i];
  const float nsx = se.n[0], nsy = se.n[1], nsz = se.n[2];
  const float cs_plane = y[0] * nsx + y[1] * nsy + y[2] * nsz;
  float DMr[NFW][9], DMi[NFW][9], GMr[NFW][9], GMi[NFW][9];
#pragma unroll
  for (int j = 0; j < NFW; ++j)
#pragma unroll
    for (int i = 0; i < 9; ++i) DMr[j][i] = DMi[j][i] = GMr[j][i] = GMi[j][i] = 0.f;
#pragma unroll 1
  for (int a = 0; a < NQ; ++a) {
    const float xa = xq[3 * a + 0], ya = xq[3 * a + 1], za = xq[3 * a + 2];
    const float ps = cs_plane - (xa * nsx + ya * nsy + za * nsz);  // (y_b - x_a).n_s for every b
    float xDr[NFW][NQ], xDi[NFW][NQ], xGr[NFW][NQ], xGi[NFW][NQ];
#pragma unroll
    for (int b = 0; b < NQ; ++b) {
      const float dx = y[3 * b + 0] - xa;
      const float dy = y[3 * b + 1] - ya;
      const float dz = y[3 * b + 2] - za;
      const float r2 = dx * dx + dy * dy + dz * dz;
      const float rinv = __builtin_amdgcn_rsqf(r2);
      const float r = r2 * rinv;
      const float rinv2 = rinv * rinv;
#pragma unroll
      for (int j = 0; j < NFW; ++j) {
        float sn, cs;
        Phase<float>::sincos_kr(kappa[j], k[j], r, &sn, &cs);
        const float gr = cs * rinv, gi = sn * rinv;  // 4 pi G
        xDr[j][b] = -rinv2 * fmaf(sn, k[j], gr);      // 4 pi G (ik - 1/r) / r
        xDi[j][b] = rinv2 * fmaf(cs, k[j], -gi);
        xGr[j][b] = gr;
        xGi[j][b] = gi;
      }
    }
#pragma unroll
    for (int j = 0; j < NFW; ++j) {
      float tDr[3], tDi[3], tGr[3], tGi[3];
      trial_contract_n<NQ>(xDr[j], tDr);
      trial_contract_n<NQ>(xDi[j], tDi);
      trial_contract_n<NQ>(xGr[j], tGr);
      trial_contract_n<NQ>(xGi[j], tGi);
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        const float wa = phiw_f<NQ>(l * NQ + a);
        const float wd = wa * ps;
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          DMr[j][3 * l + m] += wd * tDr[m];
          DMi[j][3 * l + m] += wd * tDi[m];
          GMr[j][3 * l + m] += wa * tGr[m];
          GMi[j][3 * l + m] += wa * tGi[m];
        }
      }
    }
  }
  const float nn = nt[0] * nsx + nt[1] * nsy + nt[2] * nsz;
  const float scale = a2t * se.a2 * float(BEM_INV_FOUR_PI);
  float cp[9];
#pragma unroll
  for (int l = 0; l < 3; ++l)
#pragma unroll
    for (int m = 0; m < 3; ++m)
      cp[3 * l + m] = ct[3 * l + 0] * se.curl[3 * m + 0] + ct[3 * l + 1] * se.curl[3 * m + 1] +
                      ct[3 * l + 2] * se.curl[3 * m + 2];
#pragma unroll
  for (int j = 0; j < NFW; ++j) {
    float gsr = 0.f, gsi = 0.f;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      gsr += GMr[j][i];
      gsi += GMi[j][i];
    }
    const float k2nn = k[j] * k[j] * nn;
    const float kinv = 1.f / k[j];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const float Hr = cp[i] * gsr - k2nn * GMr[j][i];
      const float Hi = cp[i] * gsi - k2nn * GMi[j][i];
      // -D + (i/k) H, accumulated (image transforms add into the same block)
      C[j][i] = cadd(C[j][i], mk<float>(scale * (-DMr[j][i] - Hi * kinv), scale * (-DMi[j][i] + Hr * kinv)));
    }
  }
}

// Both blocks of an unordered pair {t, s} from one evaluation of the kernel at the 36 point pairs: G is symmetric,
// so the (s, t) block differs from the (t, s) block only in its double-layer part, which takes the derivative
// along the other element's normal. C receives the (t, s) block, Cp[l][m] the (s, t) block's entry (m, l).
template <int NFW>
__device__ __forceinline__ void regular_pair_f32(const float* __restrict__ xq, const float (&nt)[3],
                                                 const float (&ct)[9], float a2t, const ElemRec<float>& se,
                                                 const float (&k)[NFW], const float (&kappa)[NFW],
                                                 cplx<float> (&C)[NFW][9], cplx<float> (&Cp)[NFW][9]) {
  float y[18];
#pragma unroll
  for (int i = 0; i < 18; ++i) y[i] = se.q[i];
  const float nsx = se.n[0], nsy = se.n[1], nsz = se.n[2];
  const float ntx = nt[0], nty = nt[1], ntz = nt[2];
  // Both elements are flat: (y_b - x_a).n_s = (y_0 - x_a).n_s depends on the test point only and
  // -(y_b - x_a).n_t = (x_0 - y_b).n_t on the trial point only. The double-layer kernels of the two blocks are
  // u (y_b - x_a).n / r with u = 4 pi G (ik - 1/r) / r; the n_s factor of a test point multiplies its trial
  // contraction once, the n_t factors of the six trial points are formed once per pair.
  const float cs_plane = y[0] * nsx + y[1] * nsy + y[2] * nsz;
  const float ct_plane = xq[0] * ntx + xq[1] * nty + xq[2] * ntz;
  float pt[6];
#pragma unroll
  for (int b = 0; b < 6; ++b) pt[b] = ct_plane - (y[3 * b] * ntx + y[3 * b + 1] * nty + y[3 * b + 2] * ntz);
  float DMr[NFW][9], DMi[NFW][9], GMr[NFW][9], GMi[NFW][9], EMr[NFW][9], EMi[NFW][9];
#pragma unroll
  for (int j = 0; j < NFW; ++j)
#pragma unroll
    for (int i = 0; i < 9; ++i) DMr[j][i] = DMi[j][i] = GMr[j][i] = GMi[j][i] = EMr[j][i] = EMi[j][i] = 0.f;
#pragma unroll 1
  for (int a = 0; a < 6; ++a) {
    const float xa = xq[3 * a + 0], ya = xq[3 * a + 1], za = xq[3 * a + 2];
    const float ps = cs_plane - (xa * nsx + ya * nsy + za * nsz);  // (y_b - x_a).n_s for every b
    float xDr[NFW][6], xDi[NFW][6], xGr[NFW][6], xGi[NFW][6], xEr[NFW][6], xEi[NFW][6];
#pragma unroll
    for (int b = 0; b < 6; ++b) {
      const float dx = y[3 * b + 0] - xa;
      const float dy = y[3 * b + 1] - ya;
      const float dz = y[3 * b + 2] - za;
      const float r2 = dx * dx + dy * dy + dz * dz;
      // no guard against r = 0: the points of two elements without a shared vertex are distinct
      const float rinv = __builtin_amdgcn_rsqf(r2);
      const float r = r2 * rinv;
      const float rinv2 = rinv * rinv;
#pragma unroll
      for (int j = 0; j < NFW; ++j) {
        float sn, cs;
        Phase<float>::sincos_kr(kappa[j], k[j], r, &sn, &cs);
        const float gr = cs * rinv, gi = sn * rinv;       // 4 pi G
        const float ur = -rinv2 * fmaf(sn, k[j], gr);     // 4 pi G (ik - 1/r) / r
        const float ui = rinv2 * fmaf(cs, k[j], -gi);
        xDr[j][b] = ur;
        xDi[j][b] = ui;
        xEr[j][b] = ur * pt[b];
        xEi[j][b] = ui * pt[b];
        xGr[j][b] = gr;
        xGi[j][b] = gi;
      }
    }
#pragma unroll
    for (int j = 0; j < NFW; ++j) {
      float tDr[3], tDi[3], tGr[3], tGi[3], tEr[3], tEi[3];
      trial_contract(xDr[j], tDr);
      trial_contract(xDi[j], tDi);
      trial_contract(xGr[j], tGr);
      trial_contract(xGi[j], tGi);
      trial_contract(xEr[j], tEr);
      trial_contract(xEi[j], tEi);
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        const float wa = phiw_f<6>(l * 6 + a);
        const float wd = wa * ps;
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          DMr[j][3 * l + m] += wd * tDr[m];
          DMi[j][3 * l + m] += wd * tDi[m];
          GMr[j][3 * l + m] += wa * tGr[m];
          GMi[j][3 * l + m] += wa * tGi[m];
          EMr[j][3 * l + m] += wa * tEr[m];
          EMi[j][3 * l + m] += wa * tEi[m];
        }
      }
    }
  }
  const float nn = ntx * nsx + nty * nsy + ntz * nsz;
  const float scale = a2t * se.a2 * float(BEM_INV_FOUR_PI);
  float cp[9];
#pragma unroll
  for (int l = 0; l < 3; ++l)
#pragma unroll
    for (int m = 0; m < 3; ++m)
      cp[3 * l + m] = ct[3 * l + 0] * se.curl[3 * m + 0] + ct[3 * l + 1] * se.curl[3 * m + 1] +
                      ct[3 * l + 2] * se.curl[3 * m + 2];
#pragma unroll
  for (int j = 0; j < NFW; ++j) {
    float gsr = 0.f, gsi = 0.f;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      gsr += GMr[j][i];
      gsi += GMi[j][i];
    }
    const float k2nn = k[j] * k[j] * nn;
    const float kinv = 1.f / k[j];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const float Hr = cp[i] * gsr - k2nn * GMr[j][i];
      const float Hi = cp[i] * gsi - k2nn * GMi[j][i];
      // -D + (i/k) H for both orientations; H is the same (transposed) for both
      C[j][i] = mk<float>(scale * (-DMr[j][i] - Hi * kinv), scale * (-DMi[j][i] + Hr * kinv));
      Cp[j][i] = mk<float>(scale * (-EMr[j][i] - Hi * kinv), scale * (-EMi[j][i] + Hr * kinv));
    }
  }
}

// One wave per (test element t, group of NFW wavenumbers that share the pair geometry); both precisions launch
// NFW = 1: FP32 then needs at most 128 VGPRs (149 with the image loop) and 7 KB of LDS, four (three) waves per SIMD.
template <typename T, int NFW, bool SYM, int NQ>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? (SYM ? 3 : 4) : 1))) void
k_regular_lhs(RegParams<T> p, int nf) {
  const int t = blockIdx.x;
  const int f0 = blockIdx.y * NFW;
  const int nfw = min(NFW, nf - f0);  // the last group may be partial: its spare slot repeats a wavenumber
  const int lane = threadIdx.x;
  __shared__ cplx<T> blk[NFW][64][9];
  __shared__ cplx<T> pend[BEM_PEND_SLOTS][NFW][3];  // partial strip sums of vertices shared with later chunks

  T k[NFW], kappa[NFW];
#pragma unroll
  for (int j = 0; j < NFW; ++j) {
    const double kd = p.k[f0 + min(j, nfw - 1)];
    k[j] = T(kd);
    kappa[j] = T(kd / (2.0 * BEM_PI));
  }
  const ElemRec<T>* te = p.rec + t;
  T nt[3], ct[9];
#pragma unroll
  for (int i = 0; i < 3; ++i) nt[i] = te->n[i];
#pragma unroll
  for (int i = 0; i < 9; ++i) ct[i] = te->curl[i];
  const T a2t = te->a2;
  const int4 tv = p.fv[t];
  const int sp0 = p.sing_ptr[t], sp1 = p.sing_ptr[t + 1];
  const int mt = p.pmask ? p.pmask[t] : 0;

  for (int c = 0; c < p.n_chunks; ++c) {
    const int s = p.order[c * 64 + lane];
    // Strip columns of vertices that an earlier chunk already stored are read-modify-written below; the
    // first slot of each lane is loaded here, so the load overlaps the pair integrals of this chunk.
    const int nsl = p.slot_count[c];
    const int* sv = p.slot_vertex + (size_t)c * 192;
    const int vc0 = lane < nsl ? sv[lane] : 0;
    cplx<T> pre[NFW][3];
#pragma unroll
    for (int j = 0; j < NFW; ++j) {
      pre[j][0] = pre[j][1] = pre[j][2] = mk<T>(0, 0);
      if ((vc0 & 1) && j < nfw) {
        const cplx<T>* rows = p.strip + (size_t)(f0 + j) * p.strip_fstride + (size_t)(3 * t) * p.L + (vc0 >> 1);
        pre[j][0] = rows[0];
        pre[j][1] = rows[p.L];
        pre[j][2] = rows[2 * (size_t)p.L];
      }
    }
    cplx<T> C[NFW][9];
#pragma unroll
    for (int j = 0; j < NFW; ++j)
#pragma unroll
      for (int i = 0; i < 9; ++i) C[j][i] = mk<T>(0, 0);
    if (s >= 0) {
      const int4 sv = p.fv[s];
      const bool adj = sv.x == tv.x || sv.x == tv.y || sv.x == tv.z || sv.y == tv.x || sv.y == tv.y ||
                       sv.y == tv.z || sv.z == tv.x || sv.z == tv.y || sv.z == tv.z;
      const int ms = p.pmask ? (mt & p.pmask[s]) : 0;
      // the trial element and its images under the symmetry transforms (r = 0: itself); a pair listed among
      // the singular pairs (topologically adjacent, or an image touching t) takes its FP64 Duffy block.
      // Without symmetry (SYM = false) the loop has one trip and the block is assigned, not accumulated.
      const int R = SYM ? p.R : 1;
      for (int r = 0; r < R; ++r) {
        int qs = -1;
        if (r == 0 ? adj : ((ms >> r) & 1)) {
          const int key = s + r * p.F;
          for (int q = sp0; q < sp1; ++q)
            if (p.sing_trial[q] == key) {
              qs = q;
              break;
            }
        }
        if (qs >= 0) {
#pragma unroll
          for (int j = 0; j < NFW; ++j) {
            const cplx<double>* sl = p.sing_lhs + ((size_t)(f0 + min(j, nfw - 1)) * p.P + qs) * 9;
#pragma unroll
            for (int i = 0; i < 9; ++i) C[j][i] = cadd(C[j][i], mk<T>(T(sl[i].re), T(sl[i].im)));
          }
        } else if (r != 0 || !adj) {
          const ElemRec<T>& se = p.rec[(size_t)r * p.F + s];
          if constexpr (sizeof(T) == 4) {
            regular_block_f32<NFW, NQ>(te->q, nt, ct, a2t, se, k, kappa, C);
          } else {
            static_assert(NFW == 1, "FP64 runs one wavenumber per wave");
            regular_block<T, 1, NQ>(te->q, nt, ct, a2t, se, k[0], kappa[0], C[0]);
          }
        }
      }
    }
#pragma unroll
    for (int j = 0; j < NFW; ++j)
#pragma unroll
      for (int i = 0; i < 9; ++i) blk[j][lane][i] = C[j][i];
    __syncthreads();
    const uint16_t* st = p.slot_start + (size_t)c * SLOT_STRIDE;
    const uint16_t* inc = p.slot_inc + (size_t)c * 192;
    const int* sp = p.slot_pend + (size_t)c * 192;
    for (int kk = lane; kk < nsl; kk += 64) {
      const int e0 = st[kk], e1 = st[kk + 1];
      const bool first_slot = kk == lane;
      const int vcode = first_slot ? vc0 : sv[kk];
      const int v = vcode >> 1;
      const bool again = vcode & 1;  // an earlier chunk of this wave already stored this vertex column
      const int pc = sp[kk];
      const int pidx = pc & 0xFFFF;  // LDS slot of a vertex shared with other chunks (or none)
#pragma unroll
      for (int j = 0; j < NFW; ++j) {
        if (j < nfw) {
          cplx<T> a0 = mk<T>(0, 0), a1 = mk<T>(0, 0), a2 = mk<T>(0, 0);
          for (int e = e0; e < e1; ++e) {
            const int code = inc[e];
            const int src = code >> 2, m = code & 3;
            a0 = cadd(a0, blk[j][src][m]);
            a1 = cadd(a1, blk[j][src][3 + m]);
            a2 = cadd(a2, blk[j][src][6 + m]);
          }
          cplx<T>* rows = p.strip + (size_t)(f0 + j) * p.strip_fstride + (size_t)(3 * t) * p.L + v;
          if (pidx != BEM_PEND_NONE) {
            // partial sums travel in LDS between the vertex's chunks; its last chunk stores them once
            if (pc & BEM_PEND_FROM_LDS) {
              a0 = cadd(a0, pend[pidx][j][0]);
              a1 = cadd(a1, pend[pidx][j][1]);
              a2 = cadd(a2, pend[pidx][j][2]);
            }
            if (!(pc & BEM_PEND_TO_STRIP)) {
              pend[pidx][j][0] = a0;
              pend[pidx][j][1] = a1;
              pend[pidx][j][2] = a2;
              continue;
            }
          } else if (again && first_slot) {
            a0 = cadd(a0, pre[j][0]);
            a1 = cadd(a1, pre[j][1]);
            a2 = cadd(a2, pre[j][2]);
          } else if (again) {
            a0 = cadd(a0, rows[0]);
            a1 = cadd(a1, rows[p.L]);
            a2 = cadd(a2, rows[2 * (size_t)p.L]);
          }
          rows[0] = a0;
          rows[p.L] = a1;
          rows[2 * (size_t)p.L] = a2;
        }
      }
    }
    // the barrier orders this chunk's strip and LDS partial-sum stores before the next chunk reads them
    __syncthreads();
  }
}

// Paired FP32 assembly (no symmetry images): the wave of test element t evaluates each unordered pair once.
// It visits the trial chunks from chunk_start[t] (its own chunk, or the first chunk holding an element that
// shares a vertex with t); a trial element s after t in the trial order that shares no vertex with t gives the
// (t, s) block to t's row strip and the transposed (s, t) block to t's column strip (one kernel evaluation for
// both, regular_pair_f32); an element sharing a vertex gives its FP64 singular block to the row strip; earlier
// non-adjacent elements are left to their own waves. Row-strip columns are stored for vertices whose last chunk
// is at or after chunk_start[t], column-strip columns for those at or after t's own chunk; the gather masks the
// others. Partial sums of shared vertices travel in LDS (or are read-modify-written) exactly as in k_regular_lhs,
// with "an earlier chunk of this wave" judged from the vertex's previous chunk.
// Four waves per SIMD (128 VGPRs, a few spilled outside the pair loop; 9 KB of LDS per wave) measured 6 % faster
// than the three the unconstrained allocation (145 VGPRs) allows.
template <int NFW>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_regular_pair(RegParams<float> p,
                                                                                              int nf) {
  const int t = blockIdx.x;
  const int f0 = blockIdx.y * NFW;
  const int nfw = min(NFW, nf - f0);
  const int lane = threadIdx.x;
  __shared__ cplx<float> blk[NFW][64][9];
  __shared__ cplx<float> pend[BEM_PEND_SLOTS][NFW][6];  // row-strip sums [0..2], column-strip sums [3..5]
  __shared__ uint32_t tab[128];                          // the chunk's slot byte table

  float k[NFW], kappa[NFW];
#pragma unroll
  for (int j = 0; j < NFW; ++j) {
    const double kd = p.k[f0 + min(j, nfw - 1)];
    k[j] = float(kd);
    kappa[j] = float(kd / (2.0 * BEM_PI));
  }
  const ElemRec<float>* te = p.rec + t;
  float nt[3], ct[9];
#pragma unroll
  for (int i = 0; i < 3; ++i) nt[i] = te->n[i];
#pragma unroll
  for (int i = 0; i < 9; ++i) ct[i] = te->curl[i];
  const float a2t = te->a2;
  const int4 tv = p.fv[t];
  const int sp0 = p.sing_ptr[t], sp1 = p.sing_ptr[t + 1];
  const int tpos = p.elem_pos[t];
  const int c_row = p.chunk_start[t], c_col = tpos >> 6;

  // per-slot sums of the chunk's blocks in blk into one strip (w = 0 row strip, 1 column strip); the slot
  // incidence ranges and codes come from the chunk's byte table in LDS (tb), staged before the chunk's pairs
  const uint8_t* tb = reinterpret_cast<const uint8_t*>(tab);
  auto store_slots = [&](int c, int w, int c_first) {
    const int nsl = p.slot_count[c];
    const int* sv = p.slot_vertex + (size_t)c * 192;
    const int* sp = p.slot_pend + (size_t)c * 192;
    const int* spv = p.slot_prev + (size_t)c * 192;
    cplx<float>* strip = w == 0 ? p.strip : p.strip2;
    for (int kk = lane; kk < nsl; kk += 64) {
      const int e0 = tb[kk], e1 = tb[kk + 1];
      const int vcode = sv[kk];
      const int v = vcode >> 1;
      const bool seen = spv[kk] >= c_first;  // an earlier chunk of this wave held the vertex
      const bool again = (vcode & 1) && seen;
      const int pc = sp[kk];
      const int pidx = pc & 0xFFFF;
#pragma unroll
      for (int j = 0; j < NFW; ++j) {
        if (j < nfw) {
          cplx<float> a0 = mk<float>(0, 0), a1 = mk<float>(0, 0), a2 = mk<float>(0, 0);
          for (int e = e0; e < e1; ++e) {
            const int code = tb[256 + e];
            const int src = code >> 2, m = code & 3;
            a0 = cadd(a0, blk[j][src][m]);
            a1 = cadd(a1, blk[j][src][3 + m]);
            a2 = cadd(a2, blk[j][src][6 + m]);
          }
          cplx<float>* rows = strip + (size_t)(f0 + j) * p.strip_fstride + (size_t)(3 * t) * p.L + v;
          if (pidx != BEM_PEND_NONE) {
            if ((pc & BEM_PEND_FROM_LDS) && seen) {
              a0 = cadd(a0, pend[pidx][j][3 * w + 0]);
              a1 = cadd(a1, pend[pidx][j][3 * w + 1]);
              a2 = cadd(a2, pend[pidx][j][3 * w + 2]);
            }
            if (!(pc & BEM_PEND_TO_STRIP)) {
              pend[pidx][j][3 * w + 0] = a0;
              pend[pidx][j][3 * w + 1] = a1;
              pend[pidx][j][3 * w + 2] = a2;
              continue;
            }
          } else if (again) {
            a0 = cadd(a0, rows[0]);
            a1 = cadd(a1, rows[p.L]);
            a2 = cadd(a2, rows[2 * (size_t)p.L]);
          }
          rows[0] = a0;
          rows[p.L] = a1;
          rows[2 * (size_t)p.L] = a2;
        }
      }
    }
  };

  for (int c = c_row; c < p.n_chunks; ++c) {
    const int spos = c * 64 + lane;
    const int s = p.order[spos];
    // the chunk's slot byte table, loaded now and stored to LDS after the pairs: its latency hides behind them
    const uint2 tw = reinterpret_cast<const uint2*>(p.slot_bytes + (size_t)c * 128)[lane];
    cplx<float> C[NFW][9], Cp[NFW][9];
#pragma unroll
    for (int j = 0; j < NFW; ++j)
#pragma unroll
      for (int i = 0; i < 9; ++i) C[j][i] = Cp[j][i] = mk<float>(0, 0);
    if (s >= 0) {
      const int4 sv = p.fv[s];
      const bool adj = sv.x == tv.x || sv.x == tv.y || sv.x == tv.z || sv.y == tv.x || sv.y == tv.y ||
                       sv.y == tv.z || sv.z == tv.x || sv.z == tv.y || sv.z == tv.z;
      if (adj) {
        int qs = -1;
        for (int q = sp0; q < sp1; ++q)
          if (p.sing_trial[q] == s) {
            qs = q;
            break;
          }
        if (qs >= 0) {
#pragma unroll
          for (int j = 0; j < NFW; ++j) {
            const cplx<double>* sl = p.sing_lhs + ((size_t)(f0 + min(j, nfw - 1)) * p.P + qs) * 9;
#pragma unroll
            for (int i = 0; i < 9; ++i) C[j][i] = mk<float>(float(sl[i].re), float(sl[i].im));
          }
        }
      } else if (spos > tpos) {
        regular_pair_f32<NFW>(te->q, nt, ct, a2t, p.rec[s], k, kappa, C, Cp);
      }
    }
    tab[2 * lane] = tw.x;
    tab[2 * lane + 1] = tw.y;
#pragma unroll
    for (int j = 0; j < NFW; ++j)
#pragma unroll
      for (int i = 0; i < 9; ++i) blk[j][lane][i] = C[j][i];
    __syncthreads();
    store_slots(c, 0, c_row);
    __syncthreads();  // orders blk reads before the column blocks overwrite it
    if (c >= c_col) {
#pragma unroll
      for (int j = 0; j < NFW; ++j)
#pragma unroll
        for (int i = 0; i < 9; ++i) blk[j][lane][i] = Cp[j][i];
      __syncthreads();
      store_slots(c, 1, c_col);
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------------------------------------------
// Singular (vertex-sharing) pairs, Duffy rules: BeatEngineCpuAssembly.jl:151-218,
// BeatEngineRocmSingular.jl:84-289. Identity terms 1/2 M11 and 1/2 (i/k) M10 are added to the
// coincident pair of each element (BeatEngineCore.jl:731-749 with the regular rule).
// ------------------------------------------------------------------------------------------------


// One lane per (pair, group of NF frequencies) item; items run pair-major over pairs sorted by Duffy rule,
// frequency group fastest, so the lanes of a wave share one rule (uniform trip count, broadcast rule-point loads)
// and need no cross-lane reduction: each lane integrates its pair at its NF wavenumbers over all Duffy points in
// T arithmetic (the element geometry is formed in FP64 and rounded to T once; the block epilogue runs in FP64).
// The point geometry (r, 1/r, the normal projections, the basis products) is formed once for the NF
// wavenumbers. T = double: 1/r from the hardware estimate with two Newton steps and the FP64 phase series;
// T = float: hardware reciprocal square root and sin/cos of the phase in revolutions, as the regular kernels.
template <typename T, int NF>
__global__ __launch_bounds__(64) void k_singular(SingParams p, int nf) {
  const int ngr = (nf + NF - 1) / NF;  // frequency groups per pair
  const long long item = (long long)blockIdx.x * 64 + threadIdx.x;
  if (item >= (long long)p.P * ngr) return;
  const int pi = p.by_rule[item / ngr];
  const int f0 = (int)(item % ngr) * NF;
  const int nfv = min(NF, nf - f0);  // valid frequencies of this lane
  const int t = p.test[pi], key = p.trial[pi], ru = p.rule[pi];
  // trial key = s + r F: element s reflected by transform r (r = 0: itself)
  const int s = key % p.F, tr = key / p.F;
  const double sg[3] = {double(p.sgn[3 * tr]), double(p.sgn[3 * tr + 1]), double(p.sgn[3 * tr + 2])};
  const double det = sg[0] * sg[1] * sg[2];
  double k[NF];
  T kt[NF], kap[NF];  // wavenumber and wavenumber / 2 pi (revolutions per metre) in T
#pragma unroll
  for (int j = 0; j < NF; ++j) {
    k[j] = p.k[f0 + min(j, nfv - 1)];
    kt[j] = T(k[j]);
    kap[j] = T(k[j] / (2.0 * BEM_PI));
  }
  // y - x = d0 + sxi se1 + set se2 - txi te1 - tet te2; both elements are flat, so (y - x).n_s = pa0 + txi pa1
  // + tet pa2 and (y - x).n_t = pb0 + sxi pb1 + set pb2 (the in-plane edges of an element are normal to its n)
  T d0[3], te1[3], te2[3], se1[3], se2[3];
  T pa0, pa1, pa2, pb0, pb1, pb2;
  double nn = 0;
  {
    double a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double t0 = p.fvtx[9 * t + c], s0 = sg[c] * p.fvtx[9 * s + c];
      const double e1 = p.fvtx[9 * t + 3 + c] - t0, e2 = p.fvtx[9 * t + 6 + c] - t0;
      const double f1 = sg[c] * p.fvtx[9 * s + 3 + c] - s0, f2 = sg[c] * p.fvtx[9 * s + 6 + c] - s0;
      const double dd = s0 - t0;
      const double ntc = p.fnrm[3 * t + c], nsc = sg[c] * p.fnrm[3 * s + c];
      a0 += dd * nsc;
      a1 -= e1 * nsc;
      a2 -= e2 * nsc;
      b0 += dd * ntc;
      b1 += f1 * ntc;
      b2 += f2 * ntc;
      nn += ntc * nsc;
      d0[c] = T(dd);
      te1[c] = T(e1);
      te2[c] = T(e2);
      se1[c] = T(f1);
      se2[c] = T(f2);
    }
    pa0 = T(a0), pa1 = T(a1), pa2 = T(a2), pb0 = T(b0), pb1 = T(b1), pb2 = T(b2);
  }
  T Ar[NF][3], Ai[NF][3];
  T Dr[NF][9], Di[NF][9], Gr[NF][9], Gi[NF][9];
#pragma unroll
  for (int j = 0; j < NF; ++j) {
#pragma unroll
    for (int l = 0; l < 3; ++l) Ar[j][l] = Ai[j][l] = T(0);
#pragma unroll
    for (int i = 0; i < 9; ++i) Dr[j][i] = Di[j][i] = Gr[j][i] = Gi[j][i] = T(0);
  }
  const int off = p.rule_off[ru], len = p.rule_len[ru];
  const double* rp = p.rpts + (size_t)off * 5;
  // the next rule point is loaded while the current one is integrated
  double nxt[5] = {0, 0, 0, 0, 0};
  if (len > 0)
#pragma unroll
    for (int i = 0; i < 5; ++i) nxt[i] = rp[i];
  for (int q = 0; q < len; ++q) {
    const T txi = T(nxt[0]), tet = T(nxt[1]), sxi = T(nxt[2]), set = T(nxt[3]), w = T(nxt[4]);
    if (q + 1 < len) {
      rp += 5;
#pragma unroll
      for (int i = 0; i < 5; ++i) nxt[i] = rp[i];
    }
    T d[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) d[c] = d0[c] + sxi * se1[c] + set * se2[c] - txi * te1[c] - tet * te2[c];
    const T r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (r2 == T(0)) continue;
    const T rinv = Phase<T>::rsqrt_(r2);
    const T r = r2 * rinv;
    const T td = (pa0 + txi * pa1 + tet * pa2) * rinv;
    const T ad = -(pb0 + sxi * pb1 + set * pb2) * rinv;
    const T bt[3] = {T(1) - txi - tet, txi, tet};
    const T bs[3] = {T(1) - sxi - set, sxi, set};
#pragma unroll
    for (int j = 0; j < NF; ++j) {
      T sn, cs;
      Phase<T>::sincos_kr(kap[j], kt[j], r, &sn, &cs);
      const T gr = cs * rinv * w, gi = sn * rinv * w;  // 4 pi G w
      const T hr = -gr * rinv - gi * kt[j], hi = gr * kt[j] - gi * rinv;
      const T dgr = hr * td, dgi = hi * td;
      const T har = hr * ad, hai = hi * ad;
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        Ar[j][l] += bt[l] * har;
        Ai[j][l] += bt[l] * hai;
      }
#pragma unroll
      for (int l = 0; l < 3; ++l)
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          const T bb = bt[l] * bs[m];
          Dr[j][3 * l + m] += bb * dgr;
          Di[j][3 * l + m] += bb * dgi;
          Gr[j][3 * l + m] += bb * gr;
          Gi[j][3 * l + m] += bb * gi;
        }
    }
  }
  const double scale = p.fa2[t] * p.fa2[s] * BEM_INV_FOUR_PI;
  const double* ctp = p.fcurl + 9 * t;
  const double* csp = p.fcurl + 9 * s;
  const bool self = (t == key);  // the coincident pair of an element with itself (not with an image)
  const double a2 = p.fa2[t];
  const double cs0 = det * sg[0], cs1 = det * sg[1], cs2 = det * sg[2];  // reflected curls: det * signs
#pragma unroll
  for (int j = 0; j < NF; ++j) {
    if (j >= nfv) break;
    const int f = f0 + j;
    const double kinv = 1.0 / k[j];
    double gsr = 0, gsi = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      gsr += double(Gr[j][i]);
      gsi += double(Gi[j][i]);
    }
    cplx<double>* out = p.lhs + ((size_t)f * p.P + pi) * 9;
#pragma unroll
    for (int l = 0; l < 3; ++l) {
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        const double cp = ctp[3 * l] * cs0 * csp[3 * m] + ctp[3 * l + 1] * cs1 * csp[3 * m + 1] +
                          ctp[3 * l + 2] * cs2 * csp[3 * m + 2];
        const double Hr = cp * gsr - k[j] * k[j] * nn * double(Gr[j][3 * l + m]);
        const double Hi = cp * gsi - k[j] * k[j] * nn * double(Gi[j][3 * l + m]);
        double re = scale * (-double(Dr[j][3 * l + m]) - Hi * kinv);
        const double im = scale * (-double(Di[j][3 * l + m]) + Hr * kinv);
        if (self) re += 0.5 * p.m11[3 * l + m] * a2;
        out[3 * l + m] = mk<double>(re, im);
      }
    }
    cplx<double>* ro = p.rhs + ((size_t)f * p.P + pi) * 3;
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      // -(S + (i/k) D*) ; then -(1/2)(i/k) M10 on the coincident pair. S_l = sum_m G_lm: the trial
      // barycentric weights sum to one at every point.
      const double Sr = double(Gr[j][3 * l]) + double(Gr[j][3 * l + 1]) + double(Gr[j][3 * l + 2]);
      const double Si = double(Gi[j][3 * l]) + double(Gi[j][3 * l + 1]) + double(Gi[j][3 * l + 2]);
      const double re = scale * (-Sr + double(Ai[j][l]) * kinv);
      double im = scale * (-Si - double(Ar[j][l]) * kinv);
      if (self) im -= 0.5 * kinv * p.m10[l] * a2;
      ro[l] = mk<double>(re, im);
    }
  }
}

// ------------------------------------------------------------------------------------------------
// RHS partial sums per test element: rhs_part[f][3t+l][c] = sum_s B_ts[l] * q_unit[c][s]
// ------------------------------------------------------------------------------------------------


// RHS-operator block of the pair (t, s) with the images of s: br + i bi [l] = sum over r of B_{t, s_r}[l], each
// regular pair from the NQ-point rule in T arithmetic, each singular pair (adjacent, or an image touching t) from
// its FP64 Duffy block of frequency f.
template <typename T, bool SYM, int NQ>
__device__ __forceinline__ void rhs_pair_block(const RhsParams<T>& p, const ElemRec<T>* te, int4 tv, int sp0,
                                               int sp1, int mt, int s, int f, double kd, T k, T kappa,
                                               double (&br)[3], double (&bi)[3]) {
  constexpr int RHS_UNROLL_A = sizeof(T) == 4 ? NQ : 1;  // FP64: rolled test-point loop, no spills
  const int4 sv = p.fv[s];
  const bool adj = sv.x == tv.x || sv.x == tv.y || sv.x == tv.z || sv.y == tv.x || sv.y == tv.y ||
                   sv.y == tv.z || sv.z == tv.x || sv.z == tv.y || sv.z == tv.z;
  const int ms = p.pmask ? (mt & p.pmask[s]) : 0;
#pragma unroll
  for (int l = 0; l < 3; ++l) br[l] = bi[l] = 0.0;
  // the radiating face and its images (r = 0: itself); singular pairs take their FP64 Duffy block
  const int R = SYM ? p.R : 1;
  for (int r = 0; r < R; ++r) {
    int qs = -1;
    if (r == 0 ? adj : ((ms >> r) & 1)) {
      const int key = s + r * p.F;
      for (int q = sp0; q < sp1; ++q)
        if (p.sing_trial[q] == key) {
          qs = q;
          break;
        }
    }
    if (qs < 0 && (r != 0 || !adj)) {
      const ElemRec<T>& se = p.rec[(size_t)r * p.F + s];
      T Sr[3] = {0, 0, 0}, Si[3] = {0, 0, 0}, Ar[3] = {0, 0, 0}, Ai[3] = {0, 0, 0};
      const T ntx = te->n[0], nty = te->n[1], ntz = te->n[2];
#pragma unroll RHS_UNROLL_A
      for (int a = 0; a < NQ; ++a) {
        T sgr = 0, sgi = 0, sar = 0, sai = 0;
#pragma unroll
        for (int b = 0; b < NQ; ++b) {
          const T dx = se.q[3 * b] - te->q[3 * a];
          const T dy = se.q[3 * b + 1] - te->q[3 * a + 1];
          const T dz = se.q[3 * b + 2] - te->q[3 * a + 2];
          const T r2 = dx * dx + dy * dy + dz * dz;
          const T rinv = Phase<T>::rsqrt_(r2);
          const T rr = r2 * rinv;
          T sn, cs;
          Phase<T>::sincos_kr(kappa, k, rr, &sn, &cs);
          const T gr = cs * rinv, gi = sn * rinv;
          const T hr = -gr * rinv - gi * k, hi = gr * k - gi * rinv;
          const T ad = -(dx * ntx + dy * nty + dz * ntz) * rinv;
          const T wb = T(QRule<NQ>::w(b));
          sgr += wb * gr;
          sgi += wb * gi;
          sar += wb * hr * ad;
          sai += wb * hi * ad;
        }
#pragma unroll
        for (int l = 0; l < 3; ++l) {
          const T wa = T(qphi<NQ>(l, a) * QRule<NQ>::w(a));
          Sr[l] += wa * sgr;
          Si[l] += wa * sgi;
          Ar[l] += wa * sar;
          Ai[l] += wa * sai;
        }
      }
      const double scale = double(te->a2) * double(se.a2) * BEM_INV_FOUR_PI;
      const double kinv = 1.0 / kd;
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        br[l] += scale * (-double(Sr[l]) + double(Ai[l]) * kinv);
        bi[l] += scale * (-double(Si[l]) - double(Ar[l]) * kinv);
      }
    } else if (qs >= 0) {
      const cplx<double>* sr = p.sing_rhs + (size_t)f * p.P * 3;
      for (int l = 0; l < 3; ++l) {
        br[l] += sr[(size_t)qs * 3 + l].re;
        bi[l] += sr[(size_t)qs * 3 + l].im;
      }
    }
  }
}

template <typename T, bool SYM, int NQ>
__global__ __launch_bounds__(64) void k_rhs_part(RhsParams<T> p) {
  const int t = blockIdx.x;
  const int f = blockIdx.y;
  const int lane = threadIdx.x;
  const double kd = p.k[f];
  const T k = T(kd);
  const T kappa = T(kd / (2.0 * BEM_PI));
  const ElemRec<T>* te = p.rec + t;
  const int4 tv = p.fv[t];
  const int sp0 = p.sing_ptr[t], sp1 = p.sing_ptr[t + 1];
  double accr[3][MAX_CH], acci[3][MAX_CH];
#pragma unroll
  for (int l = 0; l < 3; ++l)
#pragma unroll
    for (int c = 0; c < MAX_CH; ++c) accr[l][c] = acci[l][c] = 0.0;

  const int mt = p.pmask ? p.pmask[t] : 0;
  for (int j = lane; j < p.n_src; j += 64) {
    const int s = p.src[j];
    double br[3], bi[3];
    rhs_pair_block<T, SYM, NQ>(p, te, tv, sp0, sp1, mt, s, f, kd, k, kappa, br, bi);
#pragma unroll
    for (int c = 0; c < MAX_CH; ++c) {
      if (c < p.n_ch) {
        const cplx<double> qv = p.qunit[(size_t)c * p.F + s];
#pragma unroll
        for (int l = 0; l < 3; ++l) {
          accr[l][c] += br[l] * qv.re - bi[l] * qv.im;
          acci[l][c] += br[l] * qv.im + bi[l] * qv.re;
        }
      }
    }
  }
#pragma unroll
  for (int l = 0; l < 3; ++l)
#pragma unroll
    for (int c = 0; c < MAX_CH; ++c) {
      accr[l][c] = wave_sum(accr[l][c]);
      acci[l][c] = wave_sum(acci[l][c]);
    }
  if (lane == 0) {
    cplx<double>* o = p.out + ((size_t)f * 3 * p.F + 3 * t) * MAX_CH;
    for (int l = 0; l < 3; ++l)
      for (int c = 0; c < MAX_CH; ++c) o[l * MAX_CH + c] = mk<double>(accr[l][c], acci[l][c]);
  }
}

// RHS-operator columns of listed source faces, one wave per (test element t, frequency):
// out[f][3t+l][j] = B_{t, src[j]}[l] for j < n_src (lanes along j, coalesced stores). The coupled FEM-BEM
// system needs the operator applied to many sparse Neumann columns (interface fluxes, piston motions); the
// per-element blocks are formed once and combined per column by k_cols_gather.
template <typename T, bool SYM, int NQ>
__global__ __launch_bounds__(64) void k_rhs_cols(RhsParams<T> p) {
  const int t = blockIdx.x;
  const int f = blockIdx.y;
  const int lane = threadIdx.x;
  const double kd = p.k[f];
  const T k = T(kd);
  const T kappa = T(kd / (2.0 * BEM_PI));
  const ElemRec<T>* te = p.rec + t;
  const int4 tv = p.fv[t];
  const int sp0 = p.sing_ptr[t], sp1 = p.sing_ptr[t + 1];
  const int mt = p.pmask ? p.pmask[t] : 0;
  cplx<double>* o = p.out + ((size_t)f * 3 * p.F + 3 * (size_t)t) * p.n_src;
  for (int j = lane; j < p.n_src; j += 64) {
    double br[3], bi[3];
    rhs_pair_block<T, SYM, NQ>(p, te, tv, sp0, sp1, mt, p.src[j], f, kd, k, kappa, br, bi);
#pragma unroll
    for (int l = 0; l < 3; ++l) o[(size_t)l * p.n_src + j] = mk<double>(br[l], bi[l]);
  }
}

// ------------------------------------------------------------------------------------------------
// Row gather: A[i][j] = sum_{(t,l) in star(i)} strip[3t+l][j]
//             b[c][i] = scale_f * sum_{(t,l) in star(i)} rhs_part[3t+l][c]
// ------------------------------------------------------------------------------------------------


template <typename T>
__global__ __launch_bounds__(256) void k_gather(GatherParams<T> g) {
  const int i = blockIdx.x;
  const int f = blockIdx.y;
  const int tid = threadIdx.x;
  __shared__ int rows[64];  // the first 64 strip rows of the star; the rest (valence above 64) come from memory
  __shared__ double red[256];
  const int s0 = g.vstar_ptr[i], s1 = g.vstar_ptr[i + 1];
  const int ns = s1 - s0;
  if (tid < min(ns, 64)) rows[tid] = g.vstar_row[s0 + tid];
  __syncthreads();
  auto row = [&](int r) { return r < 64 ? rows[r] : g.vstar_row[s0 + r]; };
  const cplx<T>* strip = g.strip + (size_t)f * g.strip_fstride;
  cplx<T>* Ar = g.A + (size_t)f * g.strideA + (size_t)i * g.lda;
  cplx<T>* A0r = g.A0 ? g.A0 + (size_t)f * g.strideA + (size_t)i * g.lda : nullptr;
  const double wi = g.roww ? g.roww[i] : 1.0;  // P1 orbit row weight (plane symmetry)
  double rmax = 0.0;
  // strip rows are vertex-indexed: row i of A is the plain sum of the strip rows of its star (coalesced)
  for (int j = tid; j < g.N; j += 256) {
    cplx<T> acc = mk<T>(0, 0);
    for (int r = 0; r < ns; ++r) acc = cadd(acc, strip[(size_t)row(r) * g.L + j]);
    if (g.roww) acc = mk<T>(acc.re * T(wi), acc.im * T(wi));
    Ar[j] = acc;
    if (A0r) A0r[j] = acc;
    rmax = fmax(rmax, (double)(fabs(acc.re) + fabs(acc.im)));
  }
  red[tid] = rmax;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (tid < o) red[tid] = fmax(red[tid], red[tid + o]);
    __syncthreads();
  }
  if (tid == 0) g.rowmax[(size_t)f * g.N + i] = red[0];
  if (tid < g.n_ch) {
    const cplx<double>* rp = g.rhs_part + (size_t)f * 3 * g.F * MAX_CH;
    cplx<double> acc = mk<double>(0, 0);
    for (int r = 0; r < ns; ++r) acc = cadd(acc, rp[(size_t)row(r) * MAX_CH + tid]);
    acc = mk<double>(acc.re * wi, acc.im * wi);
    g.b[(size_t)f * g.strideB + (size_t)tid * g.ldb + i] = cmul(acc, g.rhs_scale[f]);
  }
}

// Paired gather, one 64 x 64 tile of A per workgroup: A[i][j] = sum over the row strips of star(i) at column j
// (stored when j's last chunk is at or after the element's chunk_start) + sum over the column strips of star(j)
// at column i (stored when i's last chunk is at or after the element's own chunk). Wave w owns tile rows
// [16w, 16w+16) for the row strips and tile columns [16w, 16w+16) for the column strips, walks the flat star
// entries of those 16 vertices eight at a time and adds each 64-column strip segment into the LDS tile. The
// segments are buffer loads with one descriptor per strip row: a lane whose column was not stored gets an
// offset past the descriptor's range, which reads zero without touching memory, so the eight loads of a batch
// are issued back to back with no branch between them. Row maxima of |A| (|re|+|im|) meet across tiles by an
// atomic max on the FP64 bit pattern (non-negative, order-independent).
typedef unsigned int gather_u2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ gather_u2 strip_segment(const cplx<float>* row, int len, int off) {
  __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void*)row, (short)0, len * 8, 0x00020000);
  return __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0);
}

// Adds the strip segments of star entries [qa, qb) into the LDS tile: entry q's strip row (column) is read at
// byte offset `off` per lane when the lane's vertex_last `vl` reaches the entry's first stored chunk (field
// `thr_field` of the entry), and added to tile element `(key - base, lane)` (TRANSPOSED: `(lane, key - base)`).
// One vector load brings 64 entries into the lanes; each is then broadcast by readlane, so the descriptors are
// scalar and NB segment loads are in flight back to back.
template <int NB, bool TRANSPOSED>
__device__ __forceinline__ void add_star_segments(cplx<float> (*tile)[65], const cplx<float>* strip, int len,
                                                  const int4* ent, int qa, int qb, int base, int lane, int vl,
                                                  int off, int thr_field) {
  constexpr int OOB = 0x40000000;  // byte offset past every strip row (at most 2^24 vertices per mesh)
  for (int q0 = qa; q0 < qb; q0 += 64) {
    const int4 my = ent[min(q0 + lane, qb - 1)];
    const int mthr = thr_field == 1 ? my.y : my.z;
    const int n = min(64, qb - q0);
    for (int u0 = 0; u0 < n; u0 += NB) {
      gather_u2 v[NB];
      int key[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int src = min(u0 + u, n - 1);
        const int r = __builtin_amdgcn_readlane(my.x, src);
        const int thr = __builtin_amdgcn_readlane(mthr, src);
        key[u] = u0 + u < n ? __builtin_amdgcn_readlane(my.w, src) - base : -1;
        v[u] = strip_segment(strip + (size_t)r * len, len, (u0 + u < n && vl >= thr) ? off : OOB);
      }
#pragma unroll
      for (int u = 0; u < NB; ++u)
        if (key[u] >= 0) {
          cplx<float>& a = TRANSPOSED ? tile[lane][key[u]] : tile[key[u]][lane];
          a = mk<float>(a.re + __uint_as_float(v[u].x), a.im + __uint_as_float(v[u].y));
        }
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_gather_pair(GatherParams<T> g) {
  static_assert(sizeof(T) == 4, "the paired strips are single precision");
  const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 64, f = blockIdx.z;
  // the wave index is uniform: keeping it scalar keeps the entry walk and the descriptors in SGPRs
  const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  __shared__ cplx<T> tile[64][65];
  const cplx<T>* rs = g.strip + (size_t)f * g.strip_fstride;
  const cplx<T>* cs = g.strip2 + (size_t)f * g.strip_fstride;
#pragma unroll
  for (int r = 0; r < 16; ++r) tile[16 * w + r][lane] = mk<T>(0, 0);
  // part 1: row strips of the star of rows [i0 + 16w, i0 + 16w + 16), columns j0 + lane
  {
    const int j = j0 + lane;
    const int jl = j < g.N ? g.vertex_last[j] : -1;
    const int qa = g.vstar_ptr[min(i0 + 16 * w, g.N)], qb = g.vstar_ptr[min(i0 + 16 * w + 16, g.N)];
    add_star_segments<8, false>(tile, rs, g.L, g.vstar_ent, qa, qb, i0, lane, jl, j * 8, 1);
  }
  __syncthreads();
  // part 2: column strips of the star of columns [j0 + 16w, j0 + 16w + 16), rows i0 + lane, added transposed
  {
    const int i = i0 + lane;
    const int il = i < g.N ? g.vertex_last[i] : -1;
    const int qa = g.vstar_ptr[min(j0 + 16 * w, g.N)], qb = g.vstar_ptr[min(j0 + 16 * w + 16, g.N)];
    add_star_segments<8, true>(tile, cs, g.L, g.vstar_ent, qa, qb, j0, lane, il, i * 8, 2);
  }
  __syncthreads();
  // store rows (one wave per row at a time) and fold the row maxima
  for (int ii = w; ii < 64; ii += 4) {
    const int i = i0 + ii, j = j0 + lane;
    double v = 0.0;
    if (i < g.N && j < g.N) {
      const cplx<T> a = tile[ii][lane];
      g.A[(size_t)f * g.strideA + (size_t)i * g.lda + j] = a;
      if (g.A0) g.A0[(size_t)f * g.strideA + (size_t)i * g.lda + j] = a;
      v = (double)(fabs(a.re) + fabs(a.im));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    if (lane == 0 && i < g.N)
      atomicMax(reinterpret_cast<unsigned long long*>(g.rowmax + (size_t)f * g.N + i),
                (unsigned long long)__double_as_longlong(v));
  }
}

// b[c][i] = scale_f * w_i * sum over (t, l) in star(i) of rhs_part[3t+l][c]; rowmax[f][i] = 0 (before the paired
// gather's atomic maxima)
template <typename T>
__global__ __launch_bounds__(256) void k_gather_rhs(GatherParams<T> g) {
  const int i = blockIdx.x * 256 + threadIdx.x, f = blockIdx.y;
  if (i >= g.N) return;
  if (g.rowmax) g.rowmax[(size_t)f * g.N + i] = 0.0;
  const double wi = g.roww ? g.roww[i] : 1.0;
  const cplx<double>* rp = g.rhs_part + (size_t)f * 3 * g.F * MAX_CH;
  for (int c = 0; c < g.n_ch; ++c) {
    cplx<double> acc = mk<double>(0, 0);
    for (int q = g.vstar_ptr[i]; q < g.vstar_ptr[i + 1]; ++q) acc = cadd(acc, rp[(size_t)g.vstar_row[q] * MAX_CH + c]);
    acc = mk<double>(acc.re * wi, acc.im * wi);
    g.b[(size_t)f * g.strideB + (size_t)c * g.ldb + i] = cmul(acc, g.rhs_scale[f]);
  }
}

template <typename T>
hipError_t launch_regular(const RegParams<T>& p, int nf, hipStream_t st) {
  constexpr int NFW = 1;  // one wavenumber per wave: four FP32 waves per SIMD (see k_regular_pair)
  if constexpr (sizeof(T) == 4) {
    if (p.strip2 && p.nq == 6) {
      // one wavenumber per wave: 9 KB of LDS and at most 128 VGPRs give four waves per SIMD, against two
      // (255 VGPRs, 18 KB) when a wave shares the pair geometry between two wavenumbers: 13 % faster on the
      // 6,000-triangle horn job
      hipLaunchKernelGGL((k_regular_pair<1>), dim3(p.F, nf), dim3(64), 0, st, p, nf);
      return hipGetLastError();
    }
  }
  const dim3 grid(p.F, (nf + NFW - 1) / NFW);
  if (p.nq == 6) {
    if (p.R > 1)
      hipLaunchKernelGGL((k_regular_lhs<T, NFW, true, 6>), grid, dim3(64), 0, st, p, nf);
    else
      hipLaunchKernelGGL((k_regular_lhs<T, NFW, false, 6>), grid, dim3(64), 0, st, p, nf);
  } else if (p.nq == 3) {
    if (p.R > 1)
      hipLaunchKernelGGL((k_regular_lhs<T, NFW, true, 3>), grid, dim3(64), 0, st, p, nf);
    else
      hipLaunchKernelGGL((k_regular_lhs<T, NFW, false, 3>), grid, dim3(64), 0, st, p, nf);
  } else {
    if (p.R > 1)
      hipLaunchKernelGGL((k_regular_lhs<T, NFW, true, 1>), grid, dim3(64), 0, st, p, nf);
    else
      hipLaunchKernelGGL((k_regular_lhs<T, NFW, false, 1>), grid, dim3(64), 0, st, p, nf);
  }
  return hipGetLastError();
}
hipError_t launch_singular(const SingParams& p, int nf, bool fp32, hipStream_t st) {
  if (p.P == 0) return hipSuccess;
  const int NF = nf >= 2 ? 2 : 1;
  const long long items = (long long)p.P * ((nf + NF - 1) / NF);
  const dim3 grid((unsigned)((items + 63) / 64));
  if (fp32) {
    if (NF == 2)
      hipLaunchKernelGGL((k_singular<float, 2>), grid, dim3(64), 0, st, p, nf);
    else
      hipLaunchKernelGGL((k_singular<float, 1>), grid, dim3(64), 0, st, p, nf);
  } else {
    if (NF == 2)
      hipLaunchKernelGGL((k_singular<double, 2>), grid, dim3(64), 0, st, p, nf);
    else
      hipLaunchKernelGGL((k_singular<double, 1>), grid, dim3(64), 0, st, p, nf);
  }
  return hipGetLastError();
}
template <typename T>
hipError_t launch_rhs(const RhsParams<T>& p, int nf, hipStream_t st) {
  const dim3 grid(p.F, nf);
  if (p.nq == 6) {
    if (p.R > 1)
      hipLaunchKernelGGL((k_rhs_part<T, true, 6>), grid, dim3(64), 0, st, p);
    else
      hipLaunchKernelGGL((k_rhs_part<T, false, 6>), grid, dim3(64), 0, st, p);
  } else if (p.nq == 3) {
    if (p.R > 1)
      hipLaunchKernelGGL((k_rhs_part<T, true, 3>), grid, dim3(64), 0, st, p);
    else
      hipLaunchKernelGGL((k_rhs_part<T, false, 3>), grid, dim3(64), 0, st, p);
  } else {
    if (p.R > 1)
      hipLaunchKernelGGL((k_rhs_part<T, true, 1>), grid, dim3(64), 0, st, p);
    else
      hipLaunchKernelGGL((k_rhs_part<T, false, 1>), grid, dim3(64), 0, st, p);
  }
  return hipGetLastError();
}
template <typename T>
hipError_t launch_rhs_cols(const RhsParams<T>& p, int nf, hipStream_t st) {
  const dim3 grid(p.F, nf);
  if (p.nq == 6) {
    if (p.R > 1)
      hipLaunchKernelGGL((k_rhs_cols<T, true, 6>), grid, dim3(64), 0, st, p);
    else
      hipLaunchKernelGGL((k_rhs_cols<T, false, 6>), grid, dim3(64), 0, st, p);
  } else if (p.nq == 3) {
    if (p.R > 1)
      hipLaunchKernelGGL((k_rhs_cols<T, true, 3>), grid, dim3(64), 0, st, p);
    else
      hipLaunchKernelGGL((k_rhs_cols<T, false, 3>), grid, dim3(64), 0, st, p);
  } else {
    if (p.R > 1)
      hipLaunchKernelGGL((k_rhs_cols<T, true, 1>), grid, dim3(64), 0, st, p);
    else
      hipLaunchKernelGGL((k_rhs_cols<T, false, 1>), grid, dim3(64), 0, st, p);
  }
  return hipGetLastError();
}
template <typename T>
hipError_t launch_gather(const GatherParams<T>& g, int nf, hipStream_t st) {
  if constexpr (sizeof(T) == 4) {
    if (g.strip2) {
      hipLaunchKernelGGL(k_gather_rhs<T>, dim3((g.N + 255) / 256, nf), dim3(256), 0, st, g);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
      const int nt = (g.N + 63) / 64;
      hipLaunchKernelGGL(k_gather_pair<T>, dim3(nt, nt, nf), dim3(256), 0, st, g);
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL(k_gather<T>, dim3(g.N, nf), dim3(256), 0, st, g);
  return hipGetLastError();
}

// b only (the RHS of a further channel group; A, A0 and rowmax untouched when g.rowmax is null)
template <typename T>
hipError_t launch_gather_rhs(const GatherParams<T>& g, int nf, hipStream_t st) {
  hipLaunchKernelGGL(k_gather_rhs<T>, dim3((g.N + 255) / 256, nf), dim3(256), 0, st, g);
  return hipGetLastError();
}

template hipError_t launch_gather_rhs<float>(const GatherParams<float>&, int, hipStream_t);
template hipError_t launch_gather_rhs<double>(const GatherParams<double>&, int, hipStream_t);
template hipError_t launch_regular<float>(const RegParams<float>&, int, hipStream_t);
template hipError_t launch_regular<double>(const RegParams<double>&, int, hipStream_t);
template hipError_t launch_rhs<float>(const RhsParams<float>&, int, hipStream_t);
template hipError_t launch_rhs_cols<float>(const RhsParams<float>&, int, hipStream_t);
template hipError_t launch_rhs_cols<double>(const RhsParams<double>&, int, hipStream_t);
template hipError_t launch_rhs<double>(const RhsParams<double>&, int, hipStream_t);
template hipError_t launch_gather<float>(const GatherParams<float>&, int, hipStream_t);
template hipError_t launch_gather<double>(const GatherParams<double>&, int, hipStream_t);
