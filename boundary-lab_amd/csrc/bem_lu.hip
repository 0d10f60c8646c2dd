// Blocked dense complex LU with partial pivoting, triangular solves and FP64 iterative refinement.
//
// Replaces the reference's rocSOLVER getrf/getrs calls
//   src/blab/solvers/julia_local/src/BeatEngineRocmOperators.jl:132-222 (rocm_dense_lu!,
//   solve_rocm_dense_factorization, d_lhs \ d_rhs) and LAPACK cgetrf/cgetrs on the CPU path
//   (BeatEngineCpuSolve.jl:55-71).
//
// Storage: each matrix is row-major complex (re,im interleaved), leading dimension lda, matrices of a batch
// at a fixed stride. LAPACK semantics: P A = L U, unit lower L below the diagonal, U on and above it, pivot
// chosen with |re|+|im| (i?amax), ties to the lowest row; ipiv[c] is the row swapped with row c. A running
// permutation perm (perm[i] = original row now at position i) is kept for O(N) RHS permutation.
//
// Three levels. Right-looking over blocks of NBT = 128 columns whose trailing updates C -= L21 U12 are applied
// to the matrix beyond a pair of blocks at once (K = 256 on MFMA). Each block is two halves of NBO = 64: the
// second half is brought up to date with the first by an MFMA product (K = 64) before it is factored.
// Left-looking inside a half: sub-panels of NBI columns are factored by one 1024-thread workgroup that holds
// the sub-panel in registers (k_lu_panel), after an update with the sub-panels of the half already done; each
// sub-panel kernel also extends the half's inverted L block by its rows and writes the next sub-panel's U rows
// (one product with that inverse). Row interchanges are applied inside the trailing block by the panel kernel
// and outside it by k_laswp once the block is factored. The 64 x 64 diagonal blocks of L (built by the sub-panels)
// and U (inverted after the factorization) live in a workspace; the U12 solves and the diagonal steps of the
// triangular sweeps multiply by those inverses.
//
// Complex GEMM on real MFMA: C (M x 2Nc reals, interleaved) = A (M x 2K reals) * B'' where B'' is the 2x2 real
// expansion [[br, bi], [-bi, br]] of each complex entry of B, formed on the fly from B in LDS. FP32 uses
// v_mfma_f32_32x32x2f32 (k_gemm_tile32), FP64 v_mfma_f64_16x16x4f64 (k_gemm_sub).

#include "bem_kernels.hpp"
#include "bemhip.h"

#define NBO BEM_NBO
#define MAX_CH BEM_MAX_CH
constexpr int NBT = 2 * NBO;  // trailing block: two halves

static_assert(NBO == 64, "the inverted diagonal blocks of the LU workspace are one half block wide");
long long lu_workspace_elems(int N) { return (long long)((N + 63) / 64) * 2 * 4096; }

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<float> {
  typedef f32x4 acc_t;
  __device__ __forceinline__ static acc_t step(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // output register v of lane l holds tile row 4*(l>>4)+v, column l&15
  __device__ __forceinline__ static int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};
template <>
struct Mfma<double> {
  typedef f64x4 acc_t;
  __device__ __forceinline__ static acc_t step(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // output register v of lane l holds tile row 4*v+(l>>4), column l&15
  __device__ __forceinline__ static int row(int lane, int v) { return 4 * v + (lane >> 4); }
};


// DPP lane moves for 32-bit and 64-bit values (two 32-bit halves with the same control); lanes without a
// source keep their own value.
template <int CTRL, int RM>
__device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, RM, 0xf, false));
}
template <int CTRL, int RM>
__device__ __forceinline__ double dpp_mov(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = (int)b, hi = (int)(b >> 32);
  const int nlo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, RM, 0xf, false);
  const int nhi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, RM, 0xf, false);
  return __longlong_as_double(((long long)nhi << 32) | (unsigned)nlo);
}
template <int CTRL, int RM>
__device__ __forceinline__ int dpp_mov(int v) {
  return __builtin_amdgcn_update_dpp(v, v, CTRL, RM, 0xf, false);
}

// ------------------------------------------------------------------------------------------------
// GEMM: C[c_r0+i][c_c0+j] -= sum_k A[a_r0+i][a_c0+k] * B[b_r0+k][b_c0+j], all blocks of one matrix.
// ------------------------------------------------------------------------------------------------

struct GemmShape {
  int a_r0, a_c0, b_r0, b_c0, c_r0, c_c0, M, Nc, K;
};

template <typename T>
__global__ __launch_bounds__(256) void k_gemm_sub(MatBatch<T> mb, GemmShape g, const int* skip) {
  const int mat = blockIdx.z;
  if (skip && skip[mat]) return;
  cplx<T>* A = mb.A + (size_t)mat * mb.strideA;
  const long long lda = mb.lda;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wy = w >> 1, wx = w & 1;
  const int bm = blockIdx.y * 64, bn = blockIdx.x * 64;
  __shared__ cplx<T> As[64][17];
  __shared__ cplx<T> Bs[16][65];
  typedef typename Mfma<T>::acc_t acc_t;
  acc_t acc[2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[i][j][v] = T(0);

  for (int k0 = 0; k0 < g.K; k0 += 16) {
    // stage A tile: 64 rows x 16 complex
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q;
      const int r = e >> 4, kk = e & 15;
      cplx<T> v = mk<T>(0, 0);
      if (bm + r < g.M && k0 + kk < g.K) v = A[(size_t)(g.a_r0 + bm + r) * lda + g.a_c0 + k0 + kk];
      As[r][kk] = v;
    }
    // stage B tile: 16 rows x 64 complex
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q;
      const int kk = e >> 6, j = e & 63;
      cplx<T> v = mk<T>(0, 0);
      if (k0 + kk < g.K && bn + j < g.Nc) v = A[(size_t)(g.b_r0 + k0 + kk) * lda + g.b_c0 + bn + j];
      Bs[kk][j] = v;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
      const int kr = 4 * ks + (lane >> 4);
      const int kc = kr >> 1, pk = kr & 1;
      T av[2];
#pragma unroll
      for (int ti = 0; ti < 2; ++ti) {
        const cplx<T> a = As[32 * wy + 16 * ti + (lane & 15)][kc];
        av[ti] = pk ? a.im : a.re;
      }
#pragma unroll
      for (int tj = 0; tj < 4; ++tj) {
        const int cr = 64 * wx + 16 * tj + (lane & 15);
        const int jc = cr >> 1, pj = cr & 1;
        const cplx<T> b = Bs[kc][jc];
        const T bv = (pk == pj) ? b.re : (pk ? -b.im : b.im);
#pragma unroll
        for (int ti = 0; ti < 2; ++ti) acc[ti][tj] = Mfma<T>::step(av[ti], bv, acc[ti][tj]);
      }
    }
    __syncthreads();
  }
  // epilogue: C -= acc
  T* Cbase = reinterpret_cast<T*>(A);
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 4; ++tj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = bm + 32 * wy + 16 * ti + Mfma<T>::row(lane, v);
        const int cr = 2 * bn + 64 * wx + 16 * tj + (lane & 15);
        const int jc = cr >> 1;
        if (r < g.M && jc < g.Nc) {
          T* cp = Cbase + 2 * ((size_t)(g.c_r0 + r) * lda + g.c_c0 + jc) + (cr & 1);
          *cp = *cp - acc[ti][tj][v];
        }
      }
}

// Trailing update on v_mfma_f32_32x32x2f32 (FP32): a workgroup owns 128 rows x 64 complex columns of C, its four
// waves 64 x 32 complex each as 2 x 2 tiles of 32 x 32 reals. Complex arithmetic without expanding B: for each
// complex k every lane issues two MFMAs on the same loaded values, A.re against the B row (br, bi) and A.im
// against (-bi, br), with the real/imaginary column parity of the lane selecting the entries. K is staged in two
// stages of 16 complex; the next stage is fetched into registers while the current one is multiplied.
// Output layout of the 32x32 MFMA: register v of lane l holds row (v/4)*8 + (l/32)*4 + v%4, column l%32.
typedef float f32x16 __attribute__((ext_vector_type(16)));

// ASSIGN = false: C -= A B (C starts in the accumulators, A staged negated). ASSIGN = true: C = A B with A read
// from a separate operand array (Aop, leading dimension ldaop, stride strideAop per matrix): the triangular solve
// with an inverted diagonal block, in place on B (a workgroup reads its whole B block before writing C).
template <bool ASSIGN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_gemm_tile32(
    MatBatch<float> mb, GemmShape g, const int* skip, const cplx<float>* Aop, long long ldaop, long long strideAop) {
  constexpr int BM = 128, BN = 64, KC = 16;  // rows, complex columns, complex k per stage
  const int mat = blockIdx.z;
  if (skip && skip[mat]) return;
  cplx<float>* A = mb.A + (size_t)mat * mb.strideA;
  const long long lda = mb.lda;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wy = w >> 1, wx = w & 1;
  const int bm = blockIdx.y * BM, bn = blockIdx.x * BN;
  const int half = lane >> 5, l32 = lane & 31;
  __shared__ cplx<float> As[BM][KC + 1];
  __shared__ cplx<float> Bs[KC][BN + 1];
  // The accumulators start as the C tile itself and the MFMAs add (-A) B, so the epilogue is a plain store and
  // the C reads travel together with the first stage of A and B. Tile rows: (v/4)*8 + half*4 + v%4, real
  // column l32; 32-bit offsets from the matrix base.
  float* Cb = reinterpret_cast<float*>(A);
  const int lda2 = (int)(2 * lda);
  const int rbase = bm + 64 * wy + half * 4;  // + 32 ti + (v/4)*8 + v%4
  const int cbase = 2 * bn + 64 * wx + l32;   // + 32 tj (real column)
  const int obase = (g.c_r0 + rbase) * lda2 + 2 * g.c_c0 + cbase;
  f32x16 acc[2][2];
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int dr = 32 * ti + (v >> 2) * 8 + (v & 3);
        const bool ok = !ASSIGN && rbase + dr < g.M && ((cbase + 32 * tj) >> 1) < g.Nc;
        acc[ti][tj][v] = ok ? Cb[obase + dr * lda2 + 32 * tj] : 0.f;
      }
  const cplx<float>* Asrc = ASSIGN ? Aop + (size_t)mat * strideAop : A;
  const long long ldas = ASSIGN ? ldaop : lda;

  // staging per stage: A = 128 x 16 complex (8 per thread, stored negated), B = 16 x 64 complex (4 per thread)
  cplx<float> ra[8], rb[4];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = tid + 256 * q, r = e >> 4, kk = e & 15;
      ra[q] = (bm + r < g.M && k0 + kk < g.K) ? Asrc[(size_t)(g.a_r0 + bm + r) * ldas + g.a_c0 + k0 + kk]
                                              : mk<float>(0.f, 0.f);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q, kk = e >> 6, j = e & 63;
      rb[q] = (k0 + kk < g.K && bn + j < g.Nc) ? A[(size_t)(g.b_r0 + k0 + kk) * lda + g.b_c0 + bn + j]
                                               : mk<float>(0.f, 0.f);
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = tid + 256 * q;
      As[e >> 4][e & 15] = ASSIGN ? ra[q] : mk<float>(-ra[q].re, -ra[q].im);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q;
      Bs[e >> 6][e & 63] = rb[q];
    }
  };
  const int pj = l32 & 1;  // real/imaginary column parity of this lane's B entries
  auto multiply = [&]() {
#pragma unroll 4
    for (int st = 0; st < KC / 2; ++st) {
      const int kc = 2 * st + half;
      cplx<float> a[2];
      float bre[2], bim[2];
#pragma unroll
      for (int ti = 0; ti < 2; ++ti) a[ti] = As[64 * wy + 32 * ti + l32][kc];
#pragma unroll
      for (int tj = 0; tj < 2; ++tj) {
        const cplx<float> b = Bs[kc][(64 * wx + 32 * tj + l32) >> 1];
        bre[tj] = pj ? b.im : b.re;   // row (kc, re) of the real expansion: (br, bi)
        bim[tj] = pj ? b.re : -b.im;  // row (kc, im): (-bi, br)
      }
      // four independent accumulators between dependent MFMAs
#pragma unroll
      for (int tj = 0; tj < 2; ++tj)
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
          acc[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[ti].re, bre[tj], acc[ti][tj], 0, 0, 0);
#pragma unroll
      for (int tj = 0; tj < 2; ++tj)
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
          acc[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[ti].im, bim[tj], acc[ti][tj], 0, 0, 0);
    }
  };
  for (int k0 = 0; k0 < g.K; k0 += KC) {
    if (k0 == 0) fetch(0);
    stash();
    __syncthreads();
    if (k0 + KC < g.K) fetch(k0 + KC);  // next stage in flight during the multiply
    multiply();
    __syncthreads();
  }
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int dr = 32 * ti + (v >> 2) * 8 + (v & 3);
        if (rbase + dr < g.M && ((cbase + 32 * tj) >> 1) < g.Nc) Cb[obase + dr * lda2 + 32 * tj] = acc[ti][tj][v];
      }
}

// Trailing update C -= A B with every FP32 product taken as three bf16 products on v_mfma_f32_32x32x16_bf16:
// x = hi + lo with hi = bf16(x), lo = bf16(x - hi) (round to nearest even, v_cvt_pk_bf16_f32), and
// a b ~ a_hi b_hi + a_hi b_lo + a_lo b_hi accumulated in FP32. The dropped a_lo b_lo term and the rounding of
// the lo parts leave a relative error near 2^-16 per product: more than FP32's, but the factors only precondition
// the FP64 refinement that follows, which converges in the same number of steps (the LU of a Burton-Miller
// matrix has a condition number of tens to hundreds). Each wave owns 64 rows x 32 TN complex columns of C as
// 32 x 32 blocks with separate real and imaginary accumulators. Complex arithmetic as a real product with K
// doubled: the A operand is A itself ((re, im) interleaved along k), the B operand for the real part is
// (br, -bi) per k and for the imaginary part (bi, br), both formed from the staged (br, bi) pairs by a sign flip
// or a 16-bit rotation of each dword. A is staged negated, split into hi and lo rows in LDS; B transposed (one
// row per column of C) and split the same way.
// Operand layout of the 32x32x16 MFMA: lane l holds row (or column) l%32 and k = 8(l/32) .. 8(l/32)+7.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int gather_u2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void split_bf16(float re, float im, unsigned& hi, unsigned& lo) {
  const f32x2 v = {re, im};
  const bf16x2 h = __builtin_convertvector(v, bf16x2);
  const f32x2 hf = __builtin_convertvector(h, f32x2);
  const bf16x2 l = __builtin_convertvector(v - hf, bf16x2);
  hi = __builtin_bit_cast(unsigned, h);
  lo = __builtin_bit_cast(unsigned, l);
}

// WY x WX waves of 64 rows x 32 TN complex columns each: a tile of 64 WY rows x 32 WX TN complex columns. The large
// updates use 4 x 4 (256 x 128, 1024 threads, one workgroup per CU): every staged A and B element feeds twice
// as many products as with 2 x 2 (128 x 64), which cut a 2,746-row K = 256 update over 40 matrices from 3.31
// to 2.56 ms; updates of one 128-column block keep 2 x 2 (more workgroups). NP: bf16 products per FP32 product,
// 3 (the hi / lo split above) or 1 (the hi parts only: relative error near 2^-8 per product, no lo rows).
template <int WY, int WX, int TN, int NP>
__device__ __forceinline__ void gemm_bf16x3_tile(const MatBatch<float>& mb, const GemmShape& g, int mat, int tx,
                                                 int ty) {
  constexpr int BM = 64 * WY, BN = 32 * WX * TN, KC = 16;  // rows, complex columns, complex k per stage
  constexpr int NT = 64 * WY * WX;                     // threads
  constexpr int QA = BM * KC / NT, QB = KC * BN / NT;  // staged A and B elements per thread and stage
  constexpr int RW = KC + 4;  // LDS row: KC (re, im) bf16 pairs + 16 B of padding, in dwords
  static_assert(QA * NT == BM * KC && QB * NT == KC * BN, "tile and thread count do not match");
  cplx<float>* A = mb.A + (size_t)mat * mb.strideA;
  const long long lda = mb.lda;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wy = w / WX, wx = w % WX;
  const int bm = ty * BM, bn = tx * BN;
  const int half = lane >> 5, l32 = lane & 31;
  // two stage buffers: the stage after the current one is written while the current one is multiplied, so a
  // stage needs one barrier (the 4 x 4 tile: 2 x 60 KB)
  // (NP = 1: no lo rows; their arrays shrink to one row)
  constexpr int LM = NP == 3 ? BM : 1, LN = NP == 3 ? BN : 1;
  __shared__ __attribute__((aligned(16))) unsigned Ah[2][BM][RW], Al[2][LM][RW], Bh[2][BN][RW], Bl[2][LN][RW];

  // accumulators: block (ti, tj), register v holds row 32 ti + (v/4)*8 + half*4 + v%4 of the wave's 64 rows and
  // complex column 32 tj + l32 of its 32 TN. They start at zero and C is added in the epilogue: a C tile read
  // before the first product stalls the workgroup on HBM latency (measured 10 % slower on the 2,874-row updates).
  f32x16 acr[2][TN], aci[2][TN];
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < TN; ++tj)
#pragma unroll
      for (int v = 0; v < 16; ++v) acr[ti][tj][v] = aci[ti][tj][v] = 0.f;

  // staging per stage: A = BM x 16 complex, B = 16 x BN complex, through buffer descriptors with scalar bases at
  // the tile's first A row and first B row: 32-bit byte offsets (at most 256 rows x lda x 8 B) instead of 64-bit
  // addresses, and an element outside the update gets an offset past the descriptor's range, which loads zero
  // (and drops a store) without a branch
  constexpr int OOB = 0x7fffffff;
  const unsigned a_bytes = (unsigned)min((long long)BM * lda * 8, (long long)0x7ffffff0);
  const unsigned b_bytes = (unsigned)min((long long)g.K * lda * 8, (long long)0x7ffffff0);
  const __amdgpu_buffer_rsrc_t rsa = __builtin_amdgcn_make_buffer_rsrc(
      (void*)(A + (size_t)(g.a_r0 + bm) * lda + g.a_c0), (short)0, (int)a_bytes, 0x00020000);
  const __amdgpu_buffer_rsrc_t rsb = __builtin_amdgcn_make_buffer_rsrc(
      (void*)(A + (size_t)g.b_r0 * lda + g.b_c0 + bn), (short)0, (int)b_bytes, 0x00020000);
  const int ldab = (int)lda * 8;
  // element e = tid + NT q of a stage: A row e / 16 and k e % 16, B k e / BN and column e % BN; the lane's part of
  // the byte offset is the vector operand, the part of q and of the stage (the same for all lanes) the scalar one
  static_assert(NT % BN == 0 && NT % 16 == 0, "staging assumes whole rows per q");
  const int a_r = tid >> 4, a_k = tid & 15, b_k = tid / BN, b_j = tid % BN;
  const int a_loff = a_r * ldab + a_k * 8, b_loff = b_k * ldab + b_j * 8;
  const int a_rlim = g.M - bm, b_jlim = g.Nc - bn;
  cplx<float> ra[QA], rb[QB];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int q = 0; q < QA; ++q) {
      const int dr = (NT / 16) * q;
      const int off = (a_r + dr < a_rlim && k0 + a_k < g.K) ? a_loff : OOB;
      const gather_u2 v = __builtin_amdgcn_raw_buffer_load_b64(rsa, off, dr * ldab + k0 * 8, 0);
      ra[q] = mk<float>(__uint_as_float(v.x), __uint_as_float(v.y));
    }
#pragma unroll
    for (int q = 0; q < QB; ++q) {
      const int dk = (NT / BN) * q;
      const int off = (k0 + b_k + dk < g.K && b_j < b_jlim) ? b_loff : OOB;
      const gather_u2 v = __builtin_amdgcn_raw_buffer_load_b64(rsb, off, (k0 + dk) * ldab, 0);
      rb[q] = mk<float>(__uint_as_float(v.x), __uint_as_float(v.y));
    }
  };
  auto stash = [&](int sb) {
#pragma unroll
    for (int q = 0; q < QA; ++q) {
      const int e = tid + NT * q;
      unsigned hi, lo;
      split_bf16(-ra[q].re, -ra[q].im, hi, lo);
      Ah[sb][e >> 4][e & 15] = hi;
      if constexpr (NP == 3) Al[sb][e >> 4][e & 15] = lo;
    }
#pragma unroll
    for (int q = 0; q < QB; ++q) {
      const int e = tid + NT * q;
      unsigned hi, lo;
      split_bf16(rb[q].re, rb[q].im, hi, lo);
      Bh[sb][e % BN][e / BN] = hi;
      if constexpr (NP == 3) Bl[sb][e % BN][e / BN] = lo;
    }
  };
  // the products of one group of eight complex k (the MFMA's sixteen bf16) of stage buffer sb
  auto multiply_kg = [&](int sb, int kg) {
    const int d0 = 8 * kg + 4 * half;  // first complex k of this lane's eight bf16
    bf16x8 Brh[TN], Brl[TN], Bih[TN], Bil[TN];  // (br, -bi) and (bi, br) per complex k
#pragma unroll
    for (int tj = 0; tj < TN; ++tj) {
      const int cl = 32 * (TN * wx + tj) + l32;
      const u32x4 bh = *reinterpret_cast<const u32x4*>(&Bh[sb][cl][d0]);
      const u32x4 bl = NP == 3 ? *reinterpret_cast<const u32x4*>(&Bl[sb][NP == 3 ? cl : 0][d0]) : bh;
      u32x4 brh, brl, bih, bil;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        brh[i] = bh[i] ^ 0x80000000u;
        brl[i] = bl[i] ^ 0x80000000u;
        bih[i] = (bh[i] >> 16) | (bh[i] << 16);
        bil[i] = (bl[i] >> 16) | (bl[i] << 16);
      }
      Brh[tj] = __builtin_bit_cast(bf16x8, brh);
      Brl[tj] = __builtin_bit_cast(bf16x8, brl);
      Bih[tj] = __builtin_bit_cast(bf16x8, bih);
      Bil[tj] = __builtin_bit_cast(bf16x8, bil);
    }
    // one 32-row block's A operands at a time (8 registers live instead of 16)
#pragma unroll
    for (int ti = 0; ti < 2; ++ti) {
      const int r = 64 * wy + 32 * ti + l32;
      const bf16x8 Ahh = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(&Ah[sb][r][d0]));
#pragma unroll
      for (int tj = 0; tj < TN; ++tj) {
        acr[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ahh, Brh[tj], acr[ti][tj], 0, 0, 0);
        aci[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ahh, Bih[tj], aci[ti][tj], 0, 0, 0);
      }
      if constexpr (NP == 3) {
        const bf16x8 All = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(&Al[sb][r][d0]));
#pragma unroll
        for (int tj = 0; tj < TN; ++tj) {
          acr[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ahh, Brl[tj], acr[ti][tj], 0, 0, 0);
          aci[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ahh, Bil[tj], aci[ti][tj], 0, 0, 0);
          acr[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(All, Brh[tj], acr[ti][tj], 0, 0, 0);
          aci[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(All, Bih[tj], aci[ti][tj], 0, 0, 0);
        }
      }
    }
  };
  static_assert(KC == 16, "a stage is two groups of eight complex k");
  // Stage s multiplies buffer s & 1 in two groups; between them the staged registers (stage s + 1, loaded during
  // the previous stage) are split into buffer (s + 1) & 1, whose last reader was stage s - 1 before the previous
  // barrier, and then refilled with stage s + 2: the loads have the second group of this stage and the first of
  // the next (two groups of products) to arrive, with one set of staging registers. A persistent workgroup's
  // previous tile ended with a barrier after its last products.
  const int nst = (g.K + KC - 1) / KC;
  fetch(0);
  stash(0);
  if (nst > 1) fetch(KC);
  __syncthreads();
  for (int s = 0; s < nst; ++s) {
    multiply_kg(s & 1, 0);
    if (s + 1 < nst) stash((s + 1) & 1);
    if (s + 2 < nst) fetch((s + 2) * KC);
    multiply_kg(s & 1, 1);
    __syncthreads();
  }
  // epilogue: C += acc (the accumulators hold -A B). The C loads of eight rows are issued together through a
  // descriptor at the tile's first C row (eight loads in flight per thread, then the stores), not as one dependent
  // load-store round trip per element. The lane's offset is the vector operand (past the range for an element
  // outside the update); the row's part, the same for all lanes, is the scalar offset.
  const unsigned c_bytes = (unsigned)min((long long)BM * lda * 8, (long long)0x7ffffff0);
  const __amdgpu_buffer_rsrc_t rsc = __builtin_amdgcn_make_buffer_rsrc(
      (void*)(A + (size_t)(g.c_r0 + bm) * lda + g.c_c0 + bn), (short)0, (int)c_bytes, 0x00020000);
  const int lrow = 64 * wy + half * 4;  // the lane's first row in the tile
  constexpr int EG = 8;
#pragma unroll
  for (int tj = 0; tj < TN; ++tj) {
    const int cl = 32 * (TN * wx + tj) + l32;
    const int loff = lrow * ldab + cl * 8;
    const int rlim = bn + cl < g.Nc ? g.M - bm : 0;  // tile rows below rlim are in the update
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
      for (int v0 = 0; v0 < 16; v0 += EG) {
        gather_u2 c[EG];
#pragma unroll
        for (int i = 0; i < EG; ++i) {
          const int v = v0 + i, dr = 32 * ti + (v >> 2) * 8 + (v & 3);
          c[i] = __builtin_amdgcn_raw_buffer_load_b64(rsc, lrow + dr < rlim ? loff : OOB, dr * ldab, 0);
        }
#pragma unroll
        for (int i = 0; i < EG; ++i) {
          const int v = v0 + i, dr = 32 * ti + (v >> 2) * 8 + (v & 3);
          gather_u2 o;
          o.x = __float_as_uint(__uint_as_float(c[i].x) + acr[ti][tj][v]);
          o.y = __float_as_uint(__uint_as_float(c[i].y) + aci[ti][tj][v]);
          __builtin_amdgcn_raw_buffer_store_b64(o, rsc, lrow + dr < rlim ? loff : OOB, dr * ldab, 0);
        }
      }
  }
}

// Tile order. A launch's workgroups go to the eight XCDs in turn (workgroup b to XCD b % 8), each XCD with its own
// L2, so the tiles are split into eight contiguous ranges, one per XCD, walked by that XCD's workgroups in turn (a
// launch with fewer workgroups than tiles walks them persistently, so a large update can leave CUs free for a
// concurrent panel chain). Within a matrix the tiles go column-major over groups of GR row bands: the A rows of a
// group stay in the XCD's L2 while its column bands stream past. Against the plain order (columns fastest, tile
// b + k gridDim.x) this reads 36 % fewer bytes past L2 on the 3,002-unknown factorizations, at the same speed.
template <int WY, int WX, int TN, int NP>
__global__ __launch_bounds__(64 * WY * WX) __attribute__((amdgpu_waves_per_eu(2))) void k_gemm_bf16x3(
    MatBatch<float> mb, GemmShape g, const int* skip, int nx, int ny, int ntiles) {
  constexpr int NXCD = 8, GR = 4;
  const int nwg = gridDim.x, b = blockIdx.x;
  int first = b, step = nwg, end = ntiles;
  if (nwg >= NXCD) {
    const int x = b % NXCD, per = ntiles / NXCD, rem = ntiles % NXCD;
    first = x * per + min(x, rem) + b / NXCD;
    step = nwg / NXCD + (x < nwg % NXCD ? 1 : 0);  // workgroups on XCD x
    end = (x + 1) * per + min(x + 1, rem);
  }
  const int per_mat = nx * ny;
  for (int tile = first; tile < end; tile += step) {
    const int mat = tile / per_mat, tm = tile % per_mat;
    const int grp = tm / (GR * nx), tg = tm % (GR * nx), rows = min(GR, ny - grp * GR);
    const int ty = grp * GR + tg % rows, tx = tg / rows;
    if (skip && skip[mat]) continue;
    gemm_bf16x3_tile<WY, WX, TN, NP>(mb, g, mat, tx, ty);
  }
}

// Skinny variant for the sub-panel updates (Nc <= 16 columns, K <= 64): the K x Nc block of B sits in LDS,
// four threads share a row of C (each a quarter of K, contiguous A loads), partial sums meet by DPP quad
// swaps, and each thread writes a quarter of the row. Plain FMA: the update is bound by reading A once.
template <typename T, int NCMAX>
__global__ __launch_bounds__(256) void k_gemm_skinny(MatBatch<T> mb, GemmShape g, const int* skip) {
  constexpr int KMAX = BEM_NBO;
  const int mat = blockIdx.y;
  if (skip && skip[mat]) return;
  cplx<T>* A = mb.A + (size_t)mat * mb.strideA;
  const long long lda = mb.lda;
  const int tid = threadIdx.x;
  __shared__ cplx<T> Bs[KMAX][NCMAX + 1];
  for (int e = tid; e < g.K * NCMAX; e += 256) {
    const int kk = e / NCMAX, j = e % NCMAX;
    Bs[kk][j] = j < g.Nc ? A[(size_t)(g.b_r0 + kk) * lda + g.b_c0 + j] : mk<T>(0, 0);
  }
  __syncthreads();
  const int row = blockIdx.x * 64 + (tid >> 2), part = tid & 3;
  const int kq = (g.K + 3) / 4, ka = part * kq, kb = min(g.K, ka + kq);
  T accr[NCMAX], acci[NCMAX];
#pragma unroll
  for (int j = 0; j < NCMAX; ++j) accr[j] = acci[j] = T(0);
  // this thread's quarter of the row of A, loaded up front (independent loads, one latency)
  constexpr int KQ = KMAX / 4;
  T ar_[KQ], ai_[KQ];
  const cplx<T>* Ar = A + (size_t)(g.a_r0 + min(row, g.M - 1)) * lda + g.a_c0;
#pragma unroll
  for (int i = 0; i < KQ; ++i) {
    const cplx<T> a = (row < g.M && ka + i < kb) ? Ar[ka + i] : mk<T>(0, 0);
    ar_[i] = a.re;
    ai_[i] = a.im;
  }
#pragma unroll
  for (int i = 0; i < KQ; ++i) {
    if (ka + i < kb) {
#pragma unroll
      for (int j = 0; j < NCMAX; ++j) {
        const cplx<T> b = Bs[ka + i][j];
        accr[j] += ar_[i] * b.re - ai_[i] * b.im;
        acci[j] += ar_[i] * b.im + ai_[i] * b.re;
      }
    }
  }
  // sum the four quarter-K partials of a row (lanes 4i..4i+3: quad_perm swaps 1 and 2)
#pragma unroll
  for (int j = 0; j < NCMAX; ++j) {
    accr[j] += dpp_mov<0xb1, 0xf>(accr[j]);  // quad_perm [1,0,3,2]
    acci[j] += dpp_mov<0xb1, 0xf>(acci[j]);
    accr[j] += dpp_mov<0x4e, 0xf>(accr[j]);  // quad_perm [2,3,0,1]
    acci[j] += dpp_mov<0x4e, 0xf>(acci[j]);
  }
  if (row < g.M) {
    cplx<T>* Cr = A + (size_t)(g.c_r0 + row) * lda + g.c_c0;
#pragma unroll
    for (int j = 0; j < NCMAX; ++j)
      if ((j >> 2) == part && j < g.Nc) Cr[j] = mk<T>(Cr[j].re - accr[j], Cr[j].im - acci[j]);
  }
}

// ------------------------------------------------------------------------------------------------
// TRSM with unit lower L: X[r0:r0+nb, col0:col0+ncols] = L[r0:r0+nb, r0:r0+nb]^{-1} X
// ------------------------------------------------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(256) void k_trsm_lower(MatBatch<T> mb, int r0, int nb, int col0, int ncols,
                                                   const int* skip) {
  const int mat = blockIdx.y;
  if (skip && skip[mat]) return;
  constexpr int TW = sizeof(T) == 4 ? 64 : 32;
  cplx<T>* A = mb.A + (size_t)mat * mb.strideA;
  const long long lda = mb.lda;
  const int tid = threadIdx.x;
  const int cb = col0 + blockIdx.x * TW;
  __shared__ cplx<T> X[64][TW + 1];
  __shared__ cplx<T> Ls[64][65];  // strictly lower part of the diagonal block (the serial sweep reads only LDS)
  for (int e = tid; e < 64 * TW; e += 256) {
    const int r = e / TW, j = e % TW;
    cplx<T> v = mk<T>(0, 0);
    if (r < nb && cb + j < col0 + ncols) v = A[(size_t)(r0 + r) * lda + cb + j];
    X[r][j] = v;
  }
  for (int e = tid; e < 64 * 64; e += 256) {
    const int r = e >> 6, c = e & 63;
    Ls[r][c] = (r < nb && c < r) ? A[(size_t)(r0 + r) * lda + r0 + c] : mk<T>(0, 0);
  }
  __syncthreads();
  constexpr int BS = sizeof(T) == 4 ? 16 : 8;  // diagonal block rows held in registers
  for (int b0 = 0; b0 < nb; b0 += BS) {
    const int be = min(nb, b0 + BS);
    // diagonal block: one thread per column, the 16 unknowns in registers (L reads are independent of the
    // substitution chain and pipeline; rows past nb are zero in Ls and never stored)
    if (tid < TW) {
      cplx<T> x[BS];
#pragma unroll
      for (int i = 0; i < BS; ++i) x[i] = b0 + i < be ? X[b0 + i][tid] : mk<T>(0, 0);
#pragma unroll
      for (int i = 1; i < BS; ++i)
#pragma unroll
        for (int c = 0; c < i; ++c) x[i] = cfms(x[i], Ls[b0 + i][b0 + c], x[c]);
#pragma unroll
      for (int i = 0; i < BS; ++i)
        if (b0 + i < be) X[b0 + i][tid] = x[i];
    }
    __syncthreads();
    // update rows below the block
    const int rows_below = nb - be;
    for (int e = tid; e < rows_below * TW; e += 256) {
      const int r = be + e / TW, j = e % TW;
      cplx<T> acc = X[r][j];
#pragma unroll
      for (int c = 0; c < BS; ++c)
        if (b0 + c < be) acc = cfms(acc, Ls[r][b0 + c], X[b0 + c][j]);
      X[r][j] = acc;
    }
    __syncthreads();
  }
  for (int e = tid; e < nb * TW; e += 256) {
    const int r = e / TW, j = e % TW;
    if (cb + j < col0 + ncols) A[(size_t)(r0 + r) * lda + cb + j] = X[r][j];
  }
}

// ------------------------------------------------------------------------------------------------
// Pivot search: exact |re|+|im| argmax (LAPACK i?amax: first index among equal maxima) as a max over
// integer keys {bits of the non-negative magnitude, ~row}; rows with no candidate carry key 0. A wave reduces
// its keys with DPP row shifts and row broadcasts (no LDS round trips); lane 63 holds the result.
// ------------------------------------------------------------------------------------------------

template <typename T>
struct PivKey;
template <>
struct PivKey<float> {
  static constexpr int W = 2;
  __device__ __forceinline__ static void make(float v, int row, unsigned (&k)[2]) {
    k[0] = __float_as_uint(v);
    k[1] = ~(unsigned)row;
  }
};
template <>
struct PivKey<double> {
  static constexpr int W = 3;
  __device__ __forceinline__ static void make(double v, int row, unsigned (&k)[3]) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    k[0] = (unsigned)(b >> 32);
    k[1] = (unsigned)b;
    k[2] = ~(unsigned)row;
  }
};

template <int W>
__device__ __forceinline__ bool key_less(const unsigned (&a)[W], const unsigned (&b)[W]) {
  bool lt = false, eq = true;
#pragma unroll
  for (int i = 0; i < W; ++i) {
    lt = lt || (eq && a[i] < b[i]);
    eq = eq && a[i] == b[i];
  }
  return lt;
}

template <int CTRL, int ROWMASK, int W>
__device__ __forceinline__ void dpp_max_step(unsigned (&k)[W]) {
  unsigned o[W];
#pragma unroll
  for (int i = 0; i < W; ++i) o[i] = (unsigned)__builtin_amdgcn_update_dpp((int)k[i], (int)k[i], CTRL, ROWMASK, 0xf, false);
  if (key_less<W>(k, o))
#pragma unroll
    for (int i = 0; i < W; ++i) k[i] = o[i];
}

// max over the 64 lanes, returned in every lane
template <int W>
__device__ __forceinline__ void wave_max_key(unsigned (&k)[W]) {
  dpp_max_step<0x111, 0xf, W>(k);  // row_shr:1
  dpp_max_step<0x112, 0xf, W>(k);  // row_shr:2
  dpp_max_step<0x114, 0xf, W>(k);  // row_shr:4
  dpp_max_step<0x118, 0xf, W>(k);  // row_shr:8  -> lane 15 of each row holds the row maximum
  dpp_max_step<0x142, 0xa, W>(k);  // row_bcast:15 into rows 1 and 3
  dpp_max_step<0x143, 0xc, W>(k);  // row_bcast:31 into rows 2 and 3 -> lane 63 holds the maximum
#pragma unroll
  for (int i = 0; i < W; ++i) k[i] = (unsigned)__builtin_amdgcn_readlane((int)k[i], 63);
}

// maximum over the 64 lanes (non-NaN inputs), in every lane
template <typename T>
__device__ __forceinline__ T wave_max_dpp(T v) {
  v = fmax(v, dpp_mov<0x111, 0xf>(v));
  v = fmax(v, dpp_mov<0x112, 0xf>(v));
  v = fmax(v, dpp_mov<0x114, 0xf>(v));
  v = fmax(v, dpp_mov<0x118, 0xf>(v));
  v = fmax(v, dpp_mov<0x142, 0xa>(v));
  v = fmax(v, dpp_mov<0x143, 0xc>(v));
  if constexpr (sizeof(T) == 4) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
  } else {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, 63), hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
  }
}

// maximum / minimum over each group of 16 lanes, returned from lane 15 (every lane gets group 0's result)
template <typename T>
__device__ __forceinline__ T row16_max_dpp(T v) {
  v = fmax(v, dpp_mov<0x111, 0xf>(v));
  v = fmax(v, dpp_mov<0x112, 0xf>(v));
  v = fmax(v, dpp_mov<0x114, 0xf>(v));
  v = fmax(v, dpp_mov<0x118, 0xf>(v));
  if constexpr (sizeof(T) == 4) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 15));
  } else {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, 15), hi = __builtin_amdgcn_readlane((int)(b >> 32), 15);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
  }
}
__device__ __forceinline__ int row16_min_dpp(int v) {
  v = min(v, dpp_mov<0x111, 0xf>(v));
  v = min(v, dpp_mov<0x112, 0xf>(v));
  v = min(v, dpp_mov<0x114, 0xf>(v));
  v = min(v, dpp_mov<0x118, 0xf>(v));
  return __builtin_amdgcn_readlane(v, 15);
}

// Composite of the row interchanges (c0+jj, piv[jj]), jj < nb <= 64, applied in order (LAPACK laswp order),
// restricted to the rows it touches: afterwards row rows[q] holds the former row rows[src[q]]. Run by one
// wave; rows [c0, c0+nb) come first, pivot rows beyond are appended in first-seen order. The slots (at most
// 2 nb) live in registers, slot q in lane q % 64 of the low or high half, so each step is a few scalar, cross-lane
// and select operations; rows and src are written to LDS at the end. Returns the count.
__device__ int compose_swaps(const int* piv, int c0, int nb, int* rows, int* src, int lane) {
  int rlo = lane < nb ? c0 + lane : -1, rhi = -1;  // rows of slots lane and 64 + lane (-1: not in use)
  int slo = lane, shi = 64 + lane;                  // src of slots lane and 64 + lane
  const int pl = lane < nb ? piv[lane] : 0;
  int na = nb;
  for (int jj = 0; jj < nb; ++jj) {
    const int pv = __builtin_amdgcn_readlane(pl, jj);
    int b;
    if (pv < c0 + nb) {
      b = pv - c0;
    } else {
      const unsigned long long hlo = __ballot(rlo == pv), hhi = __ballot(rhi == pv);
      if (hlo) {
        b = (int)__builtin_ctzll(hlo);
      } else if (hhi) {
        b = 64 + (int)__builtin_ctzll(hhi);
      } else {
        b = na++;
        if (b < 64) rlo = lane == b ? pv : rlo;
        else rhi = lane == b - 64 ? pv : rhi;
      }
    }
    if (b != jj) {
      const int sj = __builtin_amdgcn_readlane(slo, jj);
      const int sb = b < 64 ? __builtin_amdgcn_readlane(slo, b) : __builtin_amdgcn_readlane(shi, b - 64);
      slo = lane == jj ? sb : slo;
      if (b < 64) slo = lane == b ? sj : slo;
      else shi = lane == b - 64 ? sj : shi;
    }
  }
  if (lane < na) {
    rows[lane] = rlo;
    src[lane] = slo;
  }
  if (64 + lane < na) {
    rows[64 + lane] = rhi;
    src[64 + lane] = shi;
  }
  return na;
}

// ------------------------------------------------------------------------------------------------
// Sub-panel factorization in registers: rows [c0, N), columns [c0, c0+nb), nb <= NBI. The row interchanges
// are applied to the other columns of the trailing block [k0, k1) here; columns outside the block are
// permuted once per block by k_laswp. Then the sub-panel's rows of the half's inverted L block (half [h0, h0+64),
// in W) and the U rows [h0, c0+nb) of the next sub-panel's nbn columns.
// ------------------------------------------------------------------------------------------------

// A row's NBI sub-panel entries as 16-byte granules of CPG complex values.
template <typename T>
struct alignas(16) Granule {
  static constexpr int CPG = 16 / (int)sizeof(cplx<T>);
  cplx<T> c[CPG];
};

template <typename T, int NBI, int MR, int MRL = 0>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_lu_panel(
    MatBatch<T> mb, int c0, int nb, int k0, int k1, int* ipiv, int* perm, long long strideP, int* info,
    const int* skip, cplx<T>* W, long long strideW, int h0, int nbn) {
  const int mat = blockIdx.x;
  if (skip && skip[mat]) return;
  cplx<T>* A = mb.A + (size_t)mat * mb.strideA;
  int* piv_out = ipiv + (size_t)mat * strideP;
  int* prm = perm + (size_t)mat * strideP;
  const long long lda = mb.lda;
  const int N = mb.N;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  // per-column exchange area: every wave publishes its largest magnitude and the first row holding it; after
  // the choice, the owner of the pivot row publishes that row and the owner of row c publishes row c
  __shared__ T s_val[16];
  __shared__ int s_row[16];
  __shared__ int s_pivs[NBI];
  __shared__ int s_rows[2 * NBI];
  __shared__ int s_src[2 * NBI];
  __shared__ int s_na;

  // the sub-panel: separate real and imaginary arrays (scalar arrays stay in registers at every size used); the
  // last MRL row sets live in LDS instead (one column of a thread's row at [j][tid]: conflict-free), which lets a
  // 16-column sub-panel cover 3,072 rows without spilling
  constexpr int MRR = MR - MRL;
  T Pr[MRR > 0 ? MRR : 1][NBI], Pi[MRR > 0 ? MRR : 1][NBI];
  // Rows move between HBM and the threads as whole 16-byte granules: the LPR lanes of a row chunk load (store)
  // its consecutive granules, so one wave instruction covers RPI whole rows (one cache line touched per granule
  // group instead of one per lane), and the chunks change hands between the lanes of their granules and the
  // thread of their row through a wave-private LDS slice of 64 rows (granule g of row q at q LPR + (g ^ q % LPR):
  // conflict-free in both directions). The slices share the LDS of the row sets Xr / Xi when there are any:
  // the register row sets pass through them before the LDS row sets are written and after they are read.
  using G = Granule<T>;
  constexpr int CPG = G::CPG;
  constexpr int LPR = NBI / CPG;  // granules per row chunk: 8, 4, 2 or 1
  constexpr int RPI = 64 / LPR;   // rows per wave instruction
  static_assert(NBI % CPG == 0 && (LPR & (LPR - 1)) == 0, "row chunks of whole granules");
  constexpr int XN = MRL > 0 ? MRL * NBI * 1024 : 0;  // entries of each of Xr, Xi
  constexpr int XBYTES = 2 * XN * (int)sizeof(T);
  constexpr int SBYTES = LPR > 1 ? 16 * 64 * LPR * 16 : 0;
  __shared__ G lds_g[(XBYTES > SBYTES ? XBYTES : SBYTES) / 16 + 1];
  T (*Xr)[1024] = reinterpret_cast<T (*)[1024]>(lds_g);
  T (*Xi)[1024] = reinterpret_cast<T (*)[1024]>(reinterpret_cast<T*>(lds_g) + XN);
  G* stg = lds_g + w * 64 * LPR;
  __shared__ G s_prow_g[LPR], s_crow_g[LPR];
  cplx<T>* const s_prow = reinterpret_cast<cplx<T>*>(s_prow_g);
  cplx<T>* const s_crow = reinterpret_cast<cplx<T>*>(s_crow_g);
  // granule transfers need 16-byte aligned row chunks
  const bool gran = LPR > 1 && ((reinterpret_cast<uintptr_t>(A) | (uintptr_t)lda * sizeof(cplx<T>) |
                                 (uintptr_t)c0 * sizeof(cplx<T>)) & 15) == 0;
  auto getP = [&](int m, int j) {
    if (m < MRR) return mk<T>(Pr[m][j], Pi[m][j]);
    return mk<T>(Xr[(m - MRR) * NBI + j][tid], Xi[(m - MRR) * NBI + j][tid]);
  };
  auto setP = [&](int m, int j, cplx<T> v) {
    if (m < MRR) {
      Pr[m][j] = v.re;
      Pi[m][j] = v.im;
    } else {
      Xr[(m - MRR) * NBI + j][tid] = v.re;
      Xi[(m - MRR) * NBI + j][tid] = v.im;
    }
  };
  // granule i of this lane in a wave instruction: row RPI i + lane / LPR of the wave's 64, granule lane % LPR
  const int gq = lane / LPR, gg = lane % LPR;
  auto load_granules = [&](int m, G (&gv)[LPR]) {
    const int rb = c0 + 1024 * m + 64 * w;
#pragma unroll
    for (int i = 0; i < LPR; ++i) {
      const int r = rb + RPI * i + gq;
      // a granule that starts before column nb lies within the row (a partial one ends at column N < lda)
      if (r < N && gg * CPG < nb) {
        gv[i] = *reinterpret_cast<const G*>(A + (size_t)r * lda + c0 + gg * CPG);
      } else {
#pragma unroll
        for (int c = 0; c < CPG; ++c) gv[i].c[c] = mk<T>(0, 0);
      }
#pragma unroll
      for (int c = 0; c < CPG; ++c)
        if (gg * CPG + c >= nb) gv[i].c[c] = mk<T>(0, 0);
    }
  };
  if (gran) {
#pragma unroll
    for (int m = 0; m < MRR; ++m) {
      G gv[LPR];
      load_granules(m, gv);
#pragma unroll
      for (int i = 0; i < LPR; ++i) {
        const int q = RPI * i + gq;
        stg[q * LPR + (gg ^ (q & (LPR - 1)))] = gv[i];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
      for (int g = 0; g < LPR; ++g) {
        const G v = stg[lane * LPR + (g ^ (lane & (LPR - 1)))];
#pragma unroll
        for (int c = 0; c < CPG; ++c) setP(m, g * CPG + c, v.c[c]);
      }
      __builtin_amdgcn_wave_barrier();
    }
    if constexpr (MRL > 0) {
      __syncthreads();  // every slice read before the LDS row sets overwrite them
#pragma unroll
      for (int m = MRR; m < MR; ++m) {
        G gv[LPR];
        load_granules(m, gv);
#pragma unroll
        for (int i = 0; i < LPR; ++i) {
          const int t = 64 * w + RPI * i + gq;
#pragma unroll
          for (int c = 0; c < CPG; ++c) {
            Xr[(m - MRR) * NBI + gg * CPG + c][t] = gv[i].c[c].re;
            Xi[(m - MRR) * NBI + gg * CPG + c][t] = gv[i].c[c].im;
          }
        }
      }
      __syncthreads();
    }
  } else {
#pragma unroll
    for (int m = 0; m < MR; ++m) {
      const int r = c0 + tid + 1024 * m;
#pragma unroll
      for (int j = 0; j < NBI; ++j)
        setP(m, j, (r < N && j < nb) ? A[(size_t)r * lda + c0 + j] : mk<T>(0, 0));
    }
  }
  int zero_col = 0;  // first breakdown column (1-based), tracked by thread 0
  // one row of this thread to an exchange row, as whole granules
  auto publish = [&](int m, G* dst) {
#pragma unroll
    for (int g = 0; g < LPR; ++g) {
      G v;
#pragma unroll
      for (int c = 0; c < CPG; ++c) v.c[c] = getP(m, g * CPG + c);
      dst[g] = v;
    }
  };

#pragma unroll
  for (int jj = 0; jj < NBI; ++jj) {
    if (jj < nb) {
      const int c = c0 + jj;
      // this thread's best candidate: strict > keeps the lowest of its rows among equal magnitudes
      T bv = T(-1);
      int bm = -1;
#pragma unroll
      for (int m = 0; m < MR; ++m) {
        const int r = c0 + tid + 1024 * m;
        if (r >= c && r < N) {
          const T v = cabs1(getP(m, jj));
          if (v > bv) {
            bv = v;
            bm = m;
          }
        }
      }
      // wave maximum (DPP), then the first row holding it: a wave's rows are ordered by (m, lane)
      const T wmax = wave_max_dpp(bv);
      int wrow = 0x7fffffff;
      if (wmax >= T(0)) {
#pragma unroll
        for (int m = 0; m < MR; ++m) {
          const unsigned long long ball = __ballot(bm == m && bv == wmax);
          if (ball) {
            wrow = c0 + 64 * w + (int)__builtin_ctzll(ball) + 1024 * m;
            break;
          }
        }
      }
      if (lane == 0) {
        s_val[w] = wmax;
        s_row[w] = wrow;
      }
      __syncthreads();
      // across the 16 waves: largest magnitude, then the lowest row holding it (LAPACK i?amax order)
      const T vq = s_val[lane & 15];
      const int rq = s_row[lane & 15];
      const T gmax = row16_max_dpp(vq);
      const int pick = row16_min_dpp(vq == gmax && gmax >= T(0) ? rq : 0x7fffffff);
      // no candidate at all cannot happen for c < N; keep row c then (the pivot test records a breakdown)
      const int pv = gmax >= T(0) ? pick : c;
#pragma unroll
      for (int m = 0; m < MR; ++m) {
        const int r = c0 + tid + 1024 * m;
        if (r == pv) publish(m, s_prow_g);
        if (r == c && pv != c) publish(m, s_crow_g);
      }
      __syncthreads();
      if (pv != c) {
#pragma unroll
        for (int m = 0; m < MR; ++m) {
          const int r = c0 + tid + 1024 * m;
          if (r == c)
#pragma unroll
            for (int j = 0; j < NBI; ++j) setP(m, j, s_prow[j]);
          if (r == pv)
#pragma unroll
            for (int j = 0; j < NBI; ++j) setP(m, j, s_crow[j]);
        }
      }
      if (tid == 0) s_pivs[jj] = pv;
      const cplx<T> pivot = s_prow[jj];
      if (!(cabs1(pivot) > T(0))) {  // zero or NaN pivot: breakdown, no elimination
        if (tid == 0 && zero_col == 0) zero_col = c + 1;
      } else {
        const cplx<T> inv = cdiv(mk<T>(1, 0), pivot);
        // multipliers of the rows below c; rows c0 .. c (U rows of this panel: thread jj and below of the
        // first row set) take multiplier 0 and keep column jj. Rows at or beyond N compute on zeros and are
        // neither searched nor stored.
        auto multiplier = [&](int m) {
          cplx<T> l = cmul(getP(m, jj), inv);
          if (m == 0 && c0 + tid <= c) l = mk<T>(0, 0);
          else setP(m, jj, l);
          return l;
        };
        if constexpr (NBI >= 8) {
          // few row sets: all multipliers first, then each pivot-row entry read once for all of them
          cplx<T> l[MR];
#pragma unroll
          for (int m = 0; m < MR; ++m) l[m] = multiplier(m);
#pragma unroll
          for (int j2 = jj + 1; j2 < NBI; ++j2) {
            const cplx<T> pj = s_prow[j2];
#pragma unroll
            for (int m = 0; m < MR; ++m) setP(m, j2, cfms(getP(m, j2), l[m], pj));
          }
        } else {
          // many row sets, few columns: the pivot row in registers, one multiplier at a time
          cplx<T> pr[NBI];
#pragma unroll
          for (int j2 = jj + 1; j2 < NBI; ++j2) pr[j2] = s_prow[j2];
#pragma unroll
          for (int m = 0; m < MR; ++m) {
            const cplx<T> l = multiplier(m);
#pragma unroll
            for (int j2 = jj + 1; j2 < NBI; ++j2) setP(m, j2, cfms(getP(m, j2), l, pr[j2]));
          }
        }
      }
    }
  }
  // the sub-panel's unit-lower diagonal block (rows c0 .. c0 + nb are threads 0 .. nb - 1 of the first row set)
  __shared__ cplx<T> s_ld[NBI][NBI + 1];
  if (tid < nb)
#pragma unroll
    for (int j = 0; j < NBI; ++j) s_ld[tid][j] = getP(0, j);
  if (gran) {
    // granule i of this lane to HBM: whole granules as one store, the partial last one by entries
    auto store_granule = [&](int r, const G& v) {
      cplx<T>* dst = A + (size_t)r * lda + c0 + gg * CPG;
      if (gg * CPG + CPG <= nb) {
        *reinterpret_cast<G*>(dst) = v;
      } else {
#pragma unroll
        for (int c = 0; c < CPG; ++c)
          if (gg * CPG + c < nb) dst[c] = v.c[c];
      }
    };
    if constexpr (MRL > 0) {
#pragma unroll
      for (int m = MRR; m < MR; ++m) {
        const int rb = c0 + 1024 * m + 64 * w;
#pragma unroll
        for (int i = 0; i < LPR; ++i) {
          const int q = RPI * i + gq;
          if (rb + q < N && gg * CPG < nb) {
            G v;
#pragma unroll
            for (int c = 0; c < CPG; ++c)
              v.c[c] = mk<T>(Xr[(m - MRR) * NBI + gg * CPG + c][64 * w + q], Xi[(m - MRR) * NBI + gg * CPG + c][64 * w + q]);
            store_granule(rb + q, v);
          }
        }
      }
      __syncthreads();  // every LDS row set read before the slices overwrite them
    }
#pragma unroll
    for (int m = 0; m < MRR; ++m) {
#pragma unroll
      for (int g = 0; g < LPR; ++g) {
        G v;
#pragma unroll
        for (int c = 0; c < CPG; ++c) v.c[c] = getP(m, g * CPG + c);
        stg[lane * LPR + (g ^ (lane & (LPR - 1)))] = v;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const int rb = c0 + 1024 * m + 64 * w;
#pragma unroll
      for (int i = 0; i < LPR; ++i) {
        const int q = RPI * i + gq;
        const G v = stg[q * LPR + (gg ^ (q & (LPR - 1)))];
        if (rb + q < N && gg * CPG < nb) store_granule(rb + q, v);
      }
      __builtin_amdgcn_wave_barrier();
    }
  } else {
#pragma unroll
    for (int m = 0; m < MR; ++m) {
      const int r = c0 + tid + 1024 * m;
      if (r < N)
#pragma unroll
        for (int j = 0; j < NBI; ++j)
          if (j < nb) A[(size_t)r * lda + c0 + j] = getP(m, j);
    }
  }
  __syncthreads();
  if (w == 0) {
    const int na = compose_swaps(s_pivs, c0, nb, s_rows, s_src, lane);
    if (lane == 0) s_na = na;
    // LAPACK-style ipiv and the row permutation (loads of the wave complete before its stores)
    if (lane < nb) piv_out[c0 + lane] = s_pivs[lane];
    const int pq = lane < na ? prm[s_rows[s_src[lane]]] : 0;
    if (lane < na) prm[s_rows[lane]] = pq;
    if (lane == 0 && zero_col != 0 && info[mat] == 0) info[mat] = zero_col;
  }
  __syncthreads();
  // interchanges on the other columns of the trailing block [k0, k1): one (row, column) pair per thread,
  // consecutive threads on consecutive columns of a row, all loads first
  const int na = s_na;
  const int ncols = (k1 - k0) - nb;
  constexpr int PER = (2 * NBI * NBT + 1023) / 1024;
  T vr[PER], vi[PER];  // scalar arrays stay in registers
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e = tid + 1024 * i;
    const int q = e / NBT, cc = e % NBT;
    vr[i] = vi[i] = T(0);
    if (cc < ncols && q < na) {
      const int col = k0 + cc + (k0 + cc >= c0 ? nb : 0);
      const cplx<T> x = A[(size_t)s_rows[s_src[q]] * lda + col];
      vr[i] = x.re;
      vi[i] = x.im;
    }
  }
  // the barrier does not wait for outstanding global loads: consume them first, so that no load of this
  // thread can observe a store issued by another thread after the barrier
#pragma unroll
  for (int i = 0; i < PER; ++i) asm volatile("" ::"v"(vr[i]), "v"(vi[i]));
  __syncthreads();
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e = tid + 1024 * i;
    const int q = e / NBT, cc = e % NBT;
    if (cc < ncols && q < na && s_src[q] != q) {
      const int col = k0 + cc + (k0 + cc >= c0 ? nb : 0);
      A[(size_t)s_rows[q] * lda + col] = mk<T>(vr[i], vi[i]);
    }
  }

  // The half's unit-lower diagonal block L[h0:h1, h0:h1] inverted one sub-panel at a time (the lower W block of
  // the half): with I = L[h0:c0, h0:c0]^-1 from the earlier sub-panels, this sub-panel's rows of the inverse are
  // [-Ic Lb I, Ic] (Ic: its diagonal block inverted, Lb: its rows in the earlier columns, both final after the
  // interchanges above). The U rows of the next sub-panel [c0 + nb, c0 + nb + nbn) of the half then follow as one
  // product with the extended inverse: U' = I' A[h0:c0+nb, c0+nb : c0+nb+nbn]. The first sub-panel of a half
  // writes the whole block (identity beyond its rows, as for a partial last half).
  __syncthreads();  // interchanges stored; the LDS of the row sets is free
  {
    const int K = c0 - h0, Kn = K + nb;
    cplx<T>* Wb = W + (size_t)mat * strideW + (size_t)(h0 / 64) * 2 * 4096;
    cplx<T>* sm = reinterpret_cast<cplx<T>*>(lds_g);
    cplx<T>(*Wi)[65] = reinterpret_cast<cplx<T>(*)[65]>(sm);               // 64 x 64 inverse rows so far
    cplx<T>(*Lb)[65] = reinterpret_cast<cplx<T>(*)[65]>(sm + 64 * 65);     // nb x K
    cplx<T>(*T1)[65] = reinterpret_cast<cplx<T>(*)[65]>(sm + 80 * 65);     // Lb I, nb x K
    cplx<T>(*X)[17] = reinterpret_cast<cplx<T>(*)[17]>(sm + 96 * 65);      // next sub-panel's columns, Kn x nbn
    for (int e = tid; e < K * K; e += 1024) Wi[e / K][e % K] = Wb[(size_t)(e / K) * 64 + e % K];
    for (int e = tid; e < nb * K; e += 1024) Lb[e / K][e % K] = A[(size_t)(c0 + e / K) * lda + h0 + e % K];
    for (int e = tid; e < Kn * nbn; e += 1024) X[e / nbn][e % nbn] = A[(size_t)(h0 + e / nbn) * lda + c0 + nb + e % nbn];
    __syncthreads();
    if (tid < nb) {  // Ic column tid by substitution, and the zero upper part of the column
      const int j = tid;
      for (int i = 0; i < j; ++i) Wi[K + i][K + j] = mk<T>(0, 0);
      Wi[K + j][K + j] = mk<T>(1, 0);
      for (int i = j + 1; i < nb; ++i) {
        cplx<T> acc = mk<T>(0, 0);
        for (int k = j; k < i; ++k) acc = cfma(acc, s_ld[i][k], Wi[K + k][K + j]);
        Wi[K + i][K + j] = mk<T>(-acc.re, -acc.im);
      }
    }
    for (int e = tid; e < nb * K; e += 1024) {  // T1 = Lb I (I lower triangular)
      const int i = e / K, j = e % K;
      cplx<T> acc = mk<T>(0, 0);
      for (int k = j; k < K; ++k) acc = cfma(acc, Lb[i][k], Wi[k][j]);
      T1[i][j] = acc;
    }
    __syncthreads();
    for (int e = tid; e < nb * K; e += 1024) {  // -Ic T1 (Ic lower triangular)
      const int i = e / K, j = e % K;
      cplx<T> acc = mk<T>(0, 0);
      for (int k = 0; k <= i; ++k) acc = cfma(acc, Wi[K + i][K + k], T1[k][j]);
      Wi[K + i][j] = mk<T>(-acc.re, -acc.im);
    }
    __syncthreads();
    if (K == 0) {
      for (int e = tid; e < 4096; e += 1024) {
        const int i = e >> 6, j = e & 63;
        Wb[e] = (i < nb && j < nb) ? Wi[i][j] : mk<T>(i == j ? T(1) : T(0), T(0));
      }
    } else {
      for (int e = tid; e < nb * Kn; e += 1024) Wb[(size_t)(K + e / Kn) * 64 + e % Kn] = Wi[K + e / Kn][e % Kn];
    }
    if (tid < Kn * nbn) {  // U' = I' X, I' lower triangular
      const int i = tid / nbn, j = tid % nbn;
      cplx<T> acc = mk<T>(0, 0);
      for (int k = 0; k <= i; ++k) acc = cfma(acc, Wi[i][k], X[k][j]);
      A[(size_t)(h0 + i) * lda + c0 + nb + j] = acc;
    }
  }
}

// Row interchanges of the pivots of rows [p0, p1) (at most NBO of them) on the columns [c_lo, c_hi) outside the
// trailing block [k0, k1). Each workgroup resolves the composite permutation in LDS, stages the touched rows of
// its column block in LDS and writes them back permuted.
template <typename T>
__global__ __launch_bounds__(256) void k_laswp(MatBatch<T> mb, const int* ipiv, long long strideP, int p0, int p1,
                                               int k0, int k1, int c_lo, int c_hi, const int* skip) {
  constexpr int CB = sizeof(T) == 4 ? 64 : 32;  // columns per workgroup
  const int mat = blockIdx.y;
  if (skip && skip[mat]) return;
  cplx<T>* A = mb.A + (size_t)mat * mb.strideA;
  const int* piv = ipiv + (size_t)mat * strideP;
  const long long lda = mb.lda;
  const int N = mb.N;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  __shared__ int s_piv[BEM_NBO];
  __shared__ int s_rows[2 * BEM_NBO];
  __shared__ int s_src[2 * BEM_NBO];
  __shared__ int s_na;
  __shared__ cplx<T> buf[2 * BEM_NBO][CB];
  const int np = p1 - p0, nb = k1 - k0;
  if (tid < np) s_piv[tid] = piv[p0 + tid];
  __syncthreads();
  if (w == 0) {
    const int na = compose_swaps(s_piv, p0, np, s_rows, s_src, lane);
    if (lane == 0) s_na = na;
  }
  __syncthreads();
  const int na = s_na;
  // column block: columns [c_lo, k0) then [k1, c_hi) (the block clipped to the window)
  const int b0 = min(max(k0, c_lo), c_hi), b1 = min(max(k1, c_lo), c_hi);
  const int outside = (c_hi - c_lo) - (b1 - b0);
  const int cb0 = blockIdx.x * CB;
  (void)nb;
  (void)N;
  for (int e = tid; e < na * CB; e += 256) {
    const int q = e / CB, jc = e % CB;
    const int cc = cb0 + jc;
    if (cc < outside) {
      const int col = c_lo + cc < b0 ? c_lo + cc : c_lo + cc + (b1 - b0);
      buf[q][jc] = A[(size_t)s_rows[s_src[q]] * lda + col];
    }
  }
  __syncthreads();
  for (int e = tid; e < na * CB; e += 256) {
    const int q = e / CB, jc = e % CB;
    const int cc = cb0 + jc;
    if (cc < outside && s_src[q] != q) {
      const int col = c_lo + cc < b0 ? c_lo + cc : c_lo + cc + (b1 - b0);
      A[(size_t)s_rows[q] * lda + col] = buf[q][jc];
    }
  }
}

// ------------------------------------------------------------------------------------------------
// Inverses of the 64 x 64 diagonal blocks of L (unit lower) and U (upper), stored per matrix in the LU
// workspace W[mat][block][0 = L, 1 = U][64][64] (identity-padded past N). They turn the U12 solves into MFMA
// products and the diagonal steps of the triangular sweeps into matrix-vector products. The L blocks are built
// by the sub-panel kernel during the factorization, one row block per sub-panel; the U blocks here, after it:
// one workgroup per (block, matrix), block-recursive: the four 16 x 16 diagonal sub-blocks are inverted by
// substitution (one thread per column), then pairs are joined at sizes 16 and 32 with
//   [A C; 0 B]^-1 = [A^-1  -A^-1 C B^-1; 0 B^-1].
// ------------------------------------------------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(256) void k_tri_inv_upper(MatBatch<T> mb, cplx<T>* W, long long strideW,
                                                       const int* skip) {
  const int mat = blockIdx.y;
  if (skip && skip[mat]) return;
  const int kb = blockIdx.x;
  const int N = mb.N;
  const int r0 = kb * 64, nr = min(64, N - r0);
  if (nr <= 0) return;
  const cplx<T>* A = mb.A + (size_t)mat * mb.strideA;
  const long long lda = mb.lda;
  const int tid = threadIdx.x;
  __shared__ cplx<T> Bk[64][65];  // the triangle (identity past nr); joined blocks' off-diagonal parts are
                                  // replaced by the intermediate products once read
  __shared__ cplx<T> X[64][65];
  for (int e = tid; e < 64 * 64; e += 256) {
    const int i = e >> 6, j = e & 63;
    cplx<T> v = mk<T>(i == j ? T(1) : T(0), T(0));
    if (i < nr && j < nr) v = j >= i ? A[(size_t)(r0 + i) * lda + r0 + j] : mk<T>(T(0), T(0));
    Bk[i][j] = v;
    X[i][j] = mk<T>(T(0), T(0));
  }
  __syncthreads();
  // 16 x 16 diagonal sub-blocks by substitution, one thread per column
  if (tid < 64) {
    const int o = tid & 48, j = tid & 15;
    X[o + j][o + j] = cdiv(mk<T>(T(1), T(0)), Bk[o + j][o + j]);
    for (int i = j - 1; i >= 0; --i) {
      cplx<T> acc = mk<T>(T(0), T(0));
      for (int k = i + 1; k <= j; ++k) acc = cfma(acc, Bk[o + i][o + k], X[o + k][o + j]);
      const cplx<T> q = cdiv(acc, Bk[o + i][o + i]);
      X[o + i][o + j] = mk<T>(-q.re, -q.im);
    }
  }
  __syncthreads();
  // join pairs of h x h inverses into 2h x 2h ones (h = 16: two pairs, h = 32: one pair)
  for (int h = 16; h < 64; h <<= 1) {
    const int E = 64 * h / 2;  // output entries over all pairs: (64 / 2h) pairs x h x h
    cplx<T> t[4];
    // T = C B^-1 (C = U12), h x h per pair
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = tid + 256 * u;
      t[u] = mk<T>(T(0), T(0));
      if (e < E) {
        const int p = e / (h * h), r = (e / h) % h, c = e % h, o = 2 * h * p;
        cplx<T> acc = mk<T>(T(0), T(0));
        for (int k = 0; k < h; ++k) acc = cfma(acc, Bk[o + r][o + h + k], X[o + h + k][o + h + c]);
        t[u] = acc;
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = tid + 256 * u;
      if (e < E) {
        const int p = e / (h * h), r = (e / h) % h, c = e % h, o = 2 * h * p;
        Bk[o + r][o + h + c] = t[u];
      }
    }
    __syncthreads();
    // X12 = -A^-1 T
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = tid + 256 * u;
      if (e < E) {
        const int p = e / (h * h), r = (e / h) % h, c = e % h, o = 2 * h * p;
        cplx<T> acc = mk<T>(T(0), T(0));
        for (int k = 0; k < h; ++k) acc = cfma(acc, X[o + r][o + k], Bk[o + k][o + h + c]);
        X[o + r][o + h + c] = mk<T>(-acc.re, -acc.im);
      }
    }
    __syncthreads();
  }
  cplx<T>* Wb = W + (size_t)mat * strideW + ((size_t)kb * 2 + 1) * 4096;
  for (int e = tid; e < 64 * 64; e += 256) Wb[e] = X[e >> 6][e & 63];
}

// ------------------------------------------------------------------------------------------------
// Misc kernels: identity permutation, residual, norms, update.
// ------------------------------------------------------------------------------------------------

__global__ void k_fill_int(int* p, int n, int value) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}

hipError_t fill_int(int* p, int n, int value, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_int, dim3((n + 255) / 256), dim3(256), 0, st, p, n, value);
  return hipGetLastError();
}

__global__ void k_iota(int* perm, long long strideP, int N, int* info) {
  const int mat = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) perm[(size_t)mat * strideP + i] = i;
  if (i == 0) info[mat] = 0;
}

// Whole solve of one matrix per workgroup (1024 threads): y = P b, then the forward sweep with unit L and the
// backward sweep with U, by 64-row blocks. A block's rows take their dot products with every entry solved so far
// (16 threads per row, 4 contiguous columns per thread and group of 64 columns, several groups of loads in flight),
// then the inverted diagonal block finishes the block as a 64 x 64 matrix-vector product. The solution lives in
// LDS when it fits (N <= CAP columns per channel), otherwise solved entries are staged through LDS in chunks.
// One launch per solve: the sweeps are chains of dependent blocks, and one workgroup walks its chain with
// barriers instead of launches, while the matrices of a batch stream their factors on separate CUs.
template <typename T>
__device__ __forceinline__ void load4(const cplx<T>* p, cplx<T> (&a)[4]) {
  if constexpr (sizeof(T) == 4) {
    typedef float v4 __attribute__((ext_vector_type(4)));
    const v4 x = reinterpret_cast<const v4*>(p)[0], y = reinterpret_cast<const v4*>(p)[1];
    a[0] = mk<T>(x[0], x[1]);
    a[1] = mk<T>(x[2], x[3]);
    a[2] = mk<T>(y[0], y[1]);
    a[3] = mk<T>(y[2], y[3]);
  } else {
    typedef double v2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const v2 x = reinterpret_cast<const v2*>(p)[q];
      a[q] = mk<T>(x[0], x[1]);
    }
  }
}

template <typename T, int NR>
__global__ __launch_bounds__(1024) void k_trsv_wg(MatBatch<T> mb, const cplx<double>* b, long long ldb,
                                                  long long strideB, cplx<T>* y, long long ldy, long long strideY,
                                                  const int* perm, long long strideP, const int* skip,
                                                  const cplx<T>* W, long long strideW) {
  constexpr int CAP = 32768 / (NR * (int)sizeof(cplx<T>));  // columns per channel in 32 KB of LDS
  const int mat = blockIdx.x;
  if (skip && skip[mat]) return;
  const int N = mb.N;
  const int nblk = (N + 63) / 64;
  const bool resident = N <= CAP;
  const cplx<T>* A = mb.A + (size_t)mat * mb.strideA;
  const long long lda = mb.lda;
  cplx<T>* Y = y + (size_t)mat * strideY;
  const int tid = threadIdx.x, row = tid >> 4, part = tid & 15;
  __shared__ cplx<T> ys[NR][CAP];
  __shared__ cplx<T> sb[NR][64];

  // y = P b (FP64 right-hand sides rounded to T), into LDS when the solution fits there
  {
    const cplx<double>* B = b + (size_t)mat * strideB;
    const int* pm = perm + (size_t)mat * strideP;
    for (int i = tid; i < N; i += 1024) {
      const int pi = pm[i];
#pragma unroll
      for (int c = 0; c < NR; ++c) {
        const cplx<double> v = B[(size_t)c * ldb + pi];
        const cplx<T> t = mk<T>(T(v.re), T(v.im));
        if (resident)
          ys[c][i] = t;
        else
          Y[(size_t)c * ldy + i] = t;
      }
    }
  }
  __syncthreads();

  for (int pass = 0; pass < 2; ++pass) {
    const bool lower = pass == 0;
    for (int s = 0; s < nblk; ++s) {
      const int kb = lower ? s : nblk - 1 - s;
      const int r0 = kb * 64, nr = min(64, N - r0);
      // inverted diagonal block entries of this thread, loaded before the dot products
      const cplx<T>* Wr = W + (size_t)mat * strideW + ((size_t)kb * 2 + (lower ? 0 : 1)) * 4096 + row * 64 + 4 * part;
      cplx<T> w[4];
      load4(Wr, w);
      // solved entries this block depends on: [0, r0) going forward, [r0 + nr, N) going backward
      const int cb = lower ? 0 : r0 + nr, ce = lower ? r0 : N;
      T accr[NR], acci[NR];
#pragma unroll
      for (int c = 0; c < NR; ++c) accr[c] = acci[c] = T(0);
      const cplx<T>* Ar = A + (size_t)(r0 + min(row, nr - 1)) * lda;
      for (int c0 = cb; c0 < ce; c0 += resident ? (ce - cb) : CAP) {
        const int len = resident ? ce - cb : min(CAP, ce - c0);
        const int yo = resident ? c0 : 0;  // LDS column of global column c0
        if (!resident) {
          for (int e = tid; e < len; e += 1024)
#pragma unroll
            for (int c = 0; c < NR; ++c) ys[c][e] = Y[(size_t)c * ldy + c0 + e];
          __syncthreads();
        }
        if (row < nr) {
          const cplx<T>* Ap = Ar + c0 + 4 * part;
          const int ng = len >> 6;  // full groups of 64 columns
          auto fma4 = [&](const cplx<T> (&a)[4], int col) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
              for (int c = 0; c < NR; ++c) {
                const cplx<T> v = ys[c][col + q];
                accr[c] += a[q].re * v.re - a[q].im * v.im;
                acci[c] += a[q].re * v.im + a[q].im * v.re;
              }
          };
          constexpr int UG = sizeof(T) == 4 ? 4 : 2;  // groups of loads in flight (FP64: register budget)
          int g = 0;
          for (; g + UG <= ng; g += UG) {
            cplx<T> a[UG][4];
#pragma unroll
            for (int u = 0; u < UG; ++u) load4(Ap + 64 * (g + u), a[u]);
#pragma unroll
            for (int u = 0; u < UG; ++u) fma4(a[u], yo + 4 * part + 64 * (g + u));
          }
          for (; g < ng; ++g) {
            cplx<T> a[4];
            load4(Ap + 64 * g, a);
            fma4(a, yo + 4 * part + 64 * g);
          }
          // tail of fewer than 64 columns
          for (int jj = 64 * ng + part; jj < len; jj += 16) {
            const cplx<T> a = Ar[c0 + jj];
#pragma unroll
            for (int c = 0; c < NR; ++c) {
              const cplx<T> v = ys[c][yo + jj];
              accr[c] += a.re * v.re - a.im * v.im;
              acci[c] += a.re * v.im + a.im * v.re;
            }
          }
        }
        if (!resident) __syncthreads();
      }
      // sum the 16 partial dot products of a row (lanes 16r .. 16r+15 of one wave)
#pragma unroll
      for (int c = 0; c < NR; ++c)
#pragma unroll
        for (int m = 1; m < 16; m <<= 1) {
          accr[c] += __shfl_xor(accr[c], m, 64);
          acci[c] += __shfl_xor(acci[c], m, 64);
        }
      if (part == 0)
#pragma unroll
        for (int c = 0; c < NR; ++c) {
          if (row < nr) {
            const cplx<T> v = resident ? ys[c][r0 + row] : Y[(size_t)c * ldy + r0 + row];
            sb[c][row] = mk<T>(v.re - accr[c], v.im - acci[c]);
          } else {
            sb[c][row] = mk<T>(0, 0);  // rows past N meet zero columns of the padded inverse
          }
        }
      __syncthreads();
      // the block's entries: inverted diagonal block times the updated right-hand side
      T dr[NR], di[NR];
#pragma unroll
      for (int c = 0; c < NR; ++c) dr[c] = di[c] = T(0);
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < NR; ++c) {
          const cplx<T> v = sb[c][4 * part + q];
          dr[c] += w[q].re * v.re - w[q].im * v.im;
          di[c] += w[q].re * v.im + w[q].im * v.re;
        }
#pragma unroll
      for (int c = 0; c < NR; ++c)
#pragma unroll
        for (int m = 1; m < 16; m <<= 1) {
          dr[c] += __shfl_xor(dr[c], m, 64);
          di[c] += __shfl_xor(di[c], m, 64);
        }
      if (part == 0 && row < nr)
#pragma unroll
        for (int c = 0; c < NR; ++c) {
          if (resident)
            ys[c][r0 + row] = mk<T>(dr[c], di[c]);
          else
            Y[(size_t)c * ldy + r0 + row] = mk<T>(dr[c], di[c]);
        }
      __syncthreads();
    }
  }
  if (resident)
    for (int i = tid; i < N; i += 1024)
#pragma unroll
      for (int c = 0; c < NR; ++c) Y[(size_t)c * ldy + i] = ys[c][i];
}

// Blocked sweeps over many workgroups, one launch per 64-row block (for large N, where one workgroup per matrix
// leaves most CUs idle and is bound by one CU's bandwidth). y = P b first (k_perm_rhs), then, for block k of
// the forward sweep (unit L) or the backward sweep (U), every workgroup of the launch forms the block's
// solution x_k = W_k y_k (W_k the inverted diagonal block; y_k is final, since the launches of the earlier
// blocks have subtracted their parts) and subtracts A[i][block k] x_k from its tile of TSW_ROWS rows still to be
// solved (after block k going forward, before it going backward). The launch's workgroups all read y_k, so
// none may overwrite it: workgroup 0 of the next launch writes the previous block's solution, recomputed from
// its y (which no later launch of the sweep changes), and the sweep's last launch, a single workgroup with
// no rows left to update, writes its own block too.
constexpr int TSW_ROWS = 256;
// one workgroup per matrix solves when N <= TSW_MIN_N and at least TSW_MIN_BATCH matrices share the launch
constexpr int TSW_MIN_N = 4096;

template <typename T, int NR>
__global__ __launch_bounds__(256) void k_perm_rhs(const cplx<double>* b, long long ldb, long long strideB,
                                                  cplx<T>* y, long long ldy, long long strideY, const int* perm,
                                                  long long strideP, int N, const int* skip) {
  const int mat = blockIdx.y;
  if (skip && skip[mat]) return;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  const int pi = perm[(size_t)mat * strideP + i];
#pragma unroll
  for (int c = 0; c < NR; ++c) {
    const cplx<double> v = b[(size_t)mat * strideB + (size_t)c * ldb + pi];
    y[(size_t)mat * strideY + (size_t)c * ldy + i] = mk<T>(T(v.re), T(v.im));
  }
}

// x = W_kb y_kb for the block's rows (rows past N read as zero) into xs; 4 threads per row of W, 16 columns each.
template <typename T, int NR>
__device__ __forceinline__ void sweep_block_solution(const cplx<T>* Wb, const cplx<T>* Y, long long ldy, int r0,
                                                     int nr, cplx<T> (*sb)[64], cplx<T> (*xs)[64]) {
  const int tid = threadIdx.x;
  if (tid < 64)
#pragma unroll
    for (int c = 0; c < NR; ++c) sb[c][tid] = tid < nr ? Y[(size_t)c * ldy + r0 + tid] : mk<T>(0, 0);
  __syncthreads();
  const int row = tid >> 2, part = tid & 3;
  T dr[NR], di[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) dr[c] = di[c] = T(0);
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    cplx<T> w[4];
    load4(Wb + row * 64 + 16 * part + 4 * u, w);
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int c = 0; c < NR; ++c) {
        const cplx<T> v = sb[c][16 * part + 4 * u + q];
        dr[c] += w[q].re * v.re - w[q].im * v.im;
        di[c] += w[q].re * v.im + w[q].im * v.re;
      }
  }
#pragma unroll
  for (int c = 0; c < NR; ++c)
#pragma unroll
    for (int m = 1; m < 4; m <<= 1) {
      dr[c] += __shfl_xor(dr[c], m, 64);
      di[c] += __shfl_xor(di[c], m, 64);
    }
  if (part == 0)
#pragma unroll
    for (int c = 0; c < NR; ++c) xs[c][row] = row < nr ? mk<T>(dr[c], di[c]) : mk<T>(0, 0);
  __syncthreads();
}

template <typename T, int NR>
__global__ __launch_bounds__(256) void k_trsv_step(MatBatch<T> mb, cplx<T>* y, long long ldy, long long strideY,
                                                   const int* skip, const cplx<T>* W, long long strideW, int kb,
                                                   int kb_prev, int lower) {
  const int mat = blockIdx.y;
  if (skip && skip[mat]) return;
  const int N = mb.N, tid = threadIdx.x;
  const int r0 = kb * 64, nr = min(64, N - r0);
  const int half = lower ? 0 : 1;  // which inverted diagonal block: L (0) or U (1)
  const cplx<T>* Wm = W + (size_t)mat * strideW;
  cplx<T>* Y = y + (size_t)mat * strideY;
  __shared__ cplx<T> sb[NR][64];
  __shared__ cplx<T> xs[NR][64];
  __shared__ cplx<T> xp[NR][64];
  // rows this launch updates: [ua, ub)
  const int ua = lower ? r0 + nr : 0, ub = lower ? N : r0;
  if (blockIdx.x == 0 && kb_prev >= 0) {
    // the previous block's solution, from its right-hand side that no workgroup of this launch reads
    const int p0 = kb_prev * 64, np = min(64, N - p0);
    sweep_block_solution<T, NR>(Wm + ((size_t)kb_prev * 2 + half) * 4096, Y, ldy, p0, np, sb, xp);
    if (tid < np)
#pragma unroll
      for (int c = 0; c < NR; ++c) Y[(size_t)c * ldy + p0 + tid] = xp[c][tid];
    __syncthreads();  // sb is reused below
  }
  sweep_block_solution<T, NR>(Wm + ((size_t)kb * 2 + half) * 4096, Y, ldy, r0, nr, sb, xs);
  if (ua >= ub) {
    // last block of the sweep: nothing left to update, the single workgroup writes the block's solution
    if (tid < nr)
#pragma unroll
      for (int c = 0; c < NR; ++c) Y[(size_t)c * ldy + r0 + tid] = xs[c][tid];
    return;
  }
  const int ta = ua + blockIdx.x * TSW_ROWS, tb = min(ub, ta + TSW_ROWS);
  const int row = tid >> 4, part = tid & 15;
  const bool full = nr == 64;  // a partial last block reads only its columns below N
  for (int i0 = ta; i0 < tb; i0 += 64) {
    cplx<T> a[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = min(i0 + row + 16 * u, tb - 1);
      const cplx<T>* Ap = mb.A + (size_t)mat * mb.strideA + (size_t)i * mb.lda + r0 + 4 * part;
      if (full) {
        load4(Ap, a[u]);
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) a[u][q] = 4 * part + q < nr ? Ap[q] : mk<T>(0, 0);
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      T accr[NR], acci[NR];
#pragma unroll
      for (int c = 0; c < NR; ++c) accr[c] = acci[c] = T(0);
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < NR; ++c) {
          const cplx<T> v = xs[c][4 * part + q];
          accr[c] += a[u][q].re * v.re - a[u][q].im * v.im;
          acci[c] += a[u][q].re * v.im + a[u][q].im * v.re;
        }
#pragma unroll
      for (int c = 0; c < NR; ++c)
#pragma unroll
        for (int m = 1; m < 16; m <<= 1) {
          accr[c] += __shfl_xor(accr[c], m, 64);
          acci[c] += __shfl_xor(acci[c], m, 64);
        }
      const int i = i0 + row + 16 * u;
      if (part == 0 && i < tb)
#pragma unroll
        for (int c = 0; c < NR; ++c) {
          cplx<T>* yi = Y + (size_t)c * ldy + i;
          const cplx<T> v = *yi;
          *yi = mk<T>(v.re - accr[c], v.im - acci[c]);
        }
    }
  }
}

// r = b - A0 x in FP64 (A0 in T, x and b in FP64); one wave per row, all RHS together.
template <typename T>
__global__ __launch_bounds__(256) void k_residual(MatBatch<T> mb, const cplx<double>* b, const cplx<double>* x,
                                                  cplx<double>* r, long long ldv, long long strideV, int nrhs,
                                                  const int* skip) {
  const int mat = blockIdx.y;
  if (skip && skip[mat]) return;
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int N = mb.N;
  if (i >= N) return;
  const cplx<T>* Ar = mb.A + (size_t)mat * mb.strideA + (size_t)i * mb.lda;
  const cplx<double>* X = x + (size_t)mat * strideV;
  double sr[MAX_CH], si[MAX_CH];
#pragma unroll
  for (int c = 0; c < MAX_CH; ++c) sr[c] = si[c] = 0.0;
  for (int j = lane; j < N; j += 64) {
    const cplx<T> a = Ar[j];
    const double ar = a.re, ai = a.im;
#pragma unroll
    for (int c = 0; c < MAX_CH; ++c) {
      if (c < nrhs) {
        const cplx<double> xv = X[(size_t)c * ldv + j];
        sr[c] += ar * xv.re - ai * xv.im;
        si[c] += ar * xv.im + ai * xv.re;
      }
    }
  }
#pragma unroll
  for (int c = 0; c < MAX_CH; ++c) {
    if (c < nrhs) {
      const double tr = wave_sum(sr[c]), ti = wave_sum(si[c]);
      if (lane == 0) {
        const size_t o = (size_t)mat * strideV + (size_t)c * ldv + i;
        r[o] = mk<double>(b[o].re - tr, b[o].im - ti);
      }
    }
  }
}

// Relative residual per matrix: max over RHS of ||r_c|| / ||b_c||; sets done[mat] when <= tol.
__global__ __launch_bounds__(256) void k_resnorm(const cplx<double>* r, const cplx<double>* b, long long ldv,
                                                 long long strideV, int N, int nrhs, double tol, double* rel,
                                                 int* done, int iter, int* iters) {
  const int mat = blockIdx.x;
  if (done[mat]) return;
  __shared__ double sr[256], sb[256];
  const int tid = threadIdx.x;
  double worst = 0.0;
  for (int c = 0; c < nrhs; ++c) {
    double ar = 0.0, ab = 0.0;
    const size_t base = (size_t)mat * strideV + (size_t)c * ldv;
    for (int i = tid; i < N; i += 256) {
      const cplx<double> rv = r[base + i], bv = b[base + i];
      ar += rv.re * rv.re + rv.im * rv.im;
      ab += bv.re * bv.re + bv.im * bv.im;
    }
    sr[tid] = ar;
    sb[tid] = ab;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if (tid < o) {
        sr[tid] += sr[tid + o];
        sb[tid] += sb[tid + o];
      }
      __syncthreads();
    }
    if (tid == 0) {
      const double q = sb[0] > 0.0 ? sqrt(sr[0] / sb[0]) : sqrt(sr[0]);
      worst = fmax(worst, q);
    }
    __syncthreads();
  }
  if (tid == 0) {
    rel[mat] = worst;
    iters[mat] = iter;
    if (worst <= tol) done[mat] = 1;
  }
}

// x64 += d (T), or x64 = d when init
template <typename T>
__global__ void k_update(cplx<double>* x, const cplx<T>* d, long long ldv, long long ldd, long long strideV,
                         long long strideD, int N, int nrhs, int init, const int* skip) {
  const int mat = blockIdx.z;
  if (skip && skip[mat]) return;
  const int c = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N || c >= nrhs) return;
  cplx<double>* xp = x + (size_t)mat * strideV + (size_t)c * ldv + i;
  const cplx<T> dv = d[(size_t)mat * strideD + (size_t)c * ldd + i];
  if (init)
    *xp = mk<double>(dv.re, dv.im);
  else
    *xp = mk<double>(xp->re + double(dv.re), xp->im + double(dv.im));
}

// Row maxima of U (upper triangle incl. diagonal) in |re|+|im|; growth = max|U| / max|A| per matrix.
template <typename T>
__global__ __launch_bounds__(256) void k_umax(MatBatch<T> mb, double* umax) {
  const int mat = blockIdx.y, i = blockIdx.x;
  const cplx<T>* Ar = mb.A + (size_t)mat * mb.strideA + (size_t)i * mb.lda;
  __shared__ double red[256];
  double m = 0.0;
  for (int j = i + threadIdx.x; j < mb.N; j += 256) m = fmax(m, double(cabs1(Ar[j])));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) umax[(size_t)mat * mb.N + i] = red[0];
}

__global__ __launch_bounds__(256) void k_growth(const double* umax, const double* amax, int N, double* growth) {
  const int mat = blockIdx.x;
  __shared__ double ru[256], ra[256];
  double u = 0.0, a = 0.0;
  for (int i = threadIdx.x; i < N; i += 256) {
    u = fmax(u, umax[(size_t)mat * N + i]);
    a = fmax(a, amax[(size_t)mat * N + i]);
  }
  ru[threadIdx.x] = u;
  ra[threadIdx.x] = a;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      ru[threadIdx.x] = fmax(ru[threadIdx.x], ru[threadIdx.x + o]);
      ra[threadIdx.x] = fmax(ra[threadIdx.x], ra[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) growth[mat] = ra[0] > 0.0 ? ru[0] / ra[0] : 0.0;
}

// ------------------------------------------------------------------------------------------------
// Host-side launch sequences
// ------------------------------------------------------------------------------------------------

template <typename T>
static hipError_t launch_panel(const MatBatch<T>& mb, int batch, int c0, int nb, int k0, int k1, int* ipiv,
                               int* perm, long long strideP, int* info, const int* skip, hipStream_t st,
                               cplx<T>* W, long long strideW, int h0, int nbn) {
  const int H = mb.N - c0;
  const int mr = (H + 1023) / 1024;
#define PANEL(NBI_, MR_, ...)                                                                                 \
  hipLaunchKernelGGL((k_lu_panel<T, NBI_, MR_, ##__VA_ARGS__>), dim3(batch), dim3(1024), 0, st, mb, c0, nb, k0, \
                     k1, ipiv, perm, strideP, info, skip, W, strideW, h0, nbn)
  // (NBI, MR, MRL): the sub-panel's register rows stay within 128 VGPRs/lane without spilling; the last MRL row
  // sets per thread live in 128 KB of LDS, which doubles the width wherever more than two row sets are needed
  if constexpr (sizeof(T) == 4) {
    if (mr <= 1) PANEL(16, 1);
    else if (mr <= 2) PANEL(16, 2);
    else if (mr <= 3) PANEL(16, 3, 1);
    else if (mr <= 6) PANEL(8, 6, 2);
    else if (mr <= 12) PANEL(4, 12, 4);
    else if (mr <= 24) PANEL(2, 24, 8);
    else return hipErrorInvalidValue;
  } else {
    if (mr <= 1) PANEL(8, 1);
    else if (mr <= 2) PANEL(8, 2);
    else if (mr <= 3) PANEL(8, 3, 1);
    else if (mr <= 6) PANEL(4, 6, 2);
    else if (mr <= 12) PANEL(2, 12, 4);
    else if (mr <= 24) PANEL(1, 24, 8);
    else return hipErrorInvalidValue;
  }
#undef PANEL
  return hipGetLastError();
}

template <typename T>
int panel_width(int N, int c0) {
  const int mr = (N - c0 + 1023) / 1024;
  if (sizeof(T) == 4) return mr <= 3 ? 16 : (mr <= 6 ? 8 : (mr <= 12 ? 4 : 2));
  return mr <= 3 ? 8 : (mr <= 6 ? 4 : (mr <= 12 ? 2 : 1));
}

// split: the large updates of FP32 factorizations on bf16 MFMA (k_gemm_bf16x3), BEM_LU_SPLIT_BF16 with three bf16
// products per FP32 product, BEM_LU_BF16 with one (BEM_LU_FP32: FP32 MFMA); max_wgs > 0 caps the workgroups of a
// large update (persistent tiles)
template <typename T>
static hipError_t launch_gemm(const MatBatch<T>& mb, int batch, const GemmShape& g, const int* skip,
                              hipStream_t st, int split = BEM_LU_FP32, int max_wgs = 0, int split_tile = 0) {
  if (g.M <= 0 || g.Nc <= 0 || g.K <= 0) return hipSuccess;
  if (g.Nc <= 16 && g.K <= BEM_NBO) {
    // as many accumulator columns as the update has (the narrow panels of large N update 4 or 8 columns)
    const dim3 grid((g.M + 63) / 64, batch);
    if (g.Nc <= 4)
      hipLaunchKernelGGL((k_gemm_skinny<T, 4>), grid, dim3(256), 0, st, mb, g, skip);
    else if (g.Nc <= 8)
      hipLaunchKernelGGL((k_gemm_skinny<T, 8>), grid, dim3(256), 0, st, mb, g, skip);
    else
      hipLaunchKernelGGL((k_gemm_skinny<T, 16>), grid, dim3(256), 0, st, mb, g, skip);
    return hipGetLastError();
  }
  if constexpr (sizeof(T) == 4) {
    if (split) {
      const bool one = split == BEM_LU_BF16;
      if (split_tile == 0 && g.M >= 512 && g.Nc >= 256) {
        const int nx = (g.Nc + 127) / 128, ny = (g.M + 255) / 256, nt = nx * ny * batch;
        const int nwg = max_wgs > 0 ? min(nt, max_wgs) : nt;
        if (one)
          hipLaunchKernelGGL((k_gemm_bf16x3<4, 4, 1, 1>), dim3(nwg), dim3(1024), 0, st, mb, g, skip, nx, ny, nt);
        else
          hipLaunchKernelGGL((k_gemm_bf16x3<4, 4, 1, 3>), dim3(nwg), dim3(1024), 0, st, mb, g, skip, nx, ny, nt);
      } else {
        const int nx = (g.Nc + 63) / 64, ny = (g.M + 127) / 128, nt = nx * ny * batch;
        if (one)
          hipLaunchKernelGGL((k_gemm_bf16x3<2, 2, 1, 1>), dim3(nt), dim3(256), 0, st, mb, g, skip, nx, ny, nt);
        else
          hipLaunchKernelGGL((k_gemm_bf16x3<2, 2, 1, 3>), dim3(nt), dim3(256), 0, st, mb, g, skip, nx, ny, nt);
      }
      return hipGetLastError();
    }
    hipLaunchKernelGGL(k_gemm_tile32<false>, dim3((g.Nc + 63) / 64, (g.M + 127) / 128, batch), dim3(256), 0, st,
                       mb, g, skip, nullptr, 0LL, 0LL);
    return hipGetLastError();
  }
  dim3 grid((g.Nc + 63) / 64, (g.M + 63) / 64, batch);
  hipLaunchKernelGGL(k_gemm_sub<T>, grid, dim3(256), 0, st, mb, g, skip);
  return hipGetLastError();
}

template <typename T>
static hipError_t launch_trsm(const MatBatch<T>& mb, int batch, int r0, int nb, int col0, int ncols,
                              const int* skip, hipStream_t st) {
  if (nb <= 0 || ncols <= 0) return hipSuccess;
  constexpr int TW = sizeof(T) == 4 ? 64 : 32;
  dim3 grid((ncols + TW - 1) / TW, batch);
  hipLaunchKernelGGL(k_trsm_lower<T>, grid, dim3(256), 0, st, mb, r0, nb, col0, ncols, skip);
  return hipGetLastError();
}

#define CHK(x)                            \
  do {                                    \
    hipError_t e_ = (x);                  \
    if (e_ != hipSuccess) return e_;      \
  } while (0)

// X = L^-1 X for the rows [r0, r0 + 64) of half block r0 / 64 (L its unit-lower diagonal block) and the ncols
// columns starting at col0. FP32: one MFMA product with the inverted diagonal block (k_gemm_tile32 in assign
// mode, in place); FP64: substitution (k_trsm_lower).
template <typename T>
static hipError_t solve_upper_rows(const MatBatch<T>& mb, int batch, int r0, int col0, int ncols, cplx<T>* W,
                                   long long strideW, const int* skip, hipStream_t st) {
  const int nr = min(NBO, mb.N - r0);
  if (ncols <= 0 || nr <= 0) return hipSuccess;
  if constexpr (sizeof(T) == 4) {
    GemmShape t = {0, 0, r0, col0, r0, col0, nr, ncols, nr};
    hipLaunchKernelGGL(k_gemm_tile32<true>, dim3((ncols + 63) / 64, 1, batch), dim3(256), 0, st, mb, t, skip,
                       (const cplx<float>*)W + (size_t)(r0 / NBO) * 2 * 4096, 64LL, strideW);
    return hipGetLastError();
  } else {
    (void)W;
    (void)strideW;
    return launch_trsm<T>(mb, batch, r0, nr, col0, ncols, skip, st);
  }
}

template <typename T>
hipError_t lu_factor_batch(const MatBatch<T>& mb, int batch, int* ipiv, int* perm, long long strideP, int* info,
                           const int* skip, cplx<T>* W, hipStream_t st, const LuLookahead* la, int products) {
  const int tile = la ? la->split_tile : 0;
  const int N = mb.N;
  const int split = sizeof(T) == 4 ? products : BEM_LU_FP32;  // bf16 products in the large updates
  const long long strideW = lu_workspace_elems(N);
  const bool look = la && la->stream;
  hipLaunchKernelGGL(k_iota, dim3((N + 255) / 256, batch), dim3(256), 0, st, perm, strideP, N, info);
  CHK(hipGetLastError());

  // factor the trailing block [k0, k1) whose columns are up to date: two halves, sub-panels in registers
  auto factor_block = [&](int k0, int k1) -> hipError_t {
    for (int h0 = k0; h0 < k1; h0 += NBO) {
      const int h1 = min(k1, h0 + NBO);
      if (h0 > k0) {
        // second half: rows [k0, h0) become U with the first half's L, the rows below take the K = 64 update
        CHK(solve_upper_rows(mb, batch, k0, h0, h1 - h0, W, strideW, skip, st));
        GemmShape g = {h0, k0, k0, h0, h0, h0, N - h0, h1 - h0, h0 - k0};
        CHK(launch_gemm<T>(mb, batch, g, skip, st, split, 0, tile));
      }
      // sub-panels: each one's U rows [h0, c0) come from the panel before it in the half, which also extends
      // the half's inverted L block (final once the half is done: later interchanges touch rows >= h1 only)
      int c0 = h0;
      while (c0 < h1) {
        const int nb = min(panel_width<T>(N, c0), h1 - c0);
        if (c0 > h0) {
          GemmShape g = {c0, h0, h0, c0, c0, c0, N - c0, nb, c0 - h0};
          CHK(launch_gemm<T>(mb, batch, g, skip, st));
        }
        const int cn = c0 + nb;
        const int nbn = cn < h1 ? min(panel_width<T>(N, cn), h1 - cn) : 0;
        CHK(launch_panel<T>(mb, batch, c0, nb, k0, k1, ipiv, perm, strideP, info, skip, st, W, strideW, h0, nbn));
        c0 = cn;
      }
    }
    return hipSuccess;
  };

  // interchanges of a block's pivots on the columns [c_lo, c_hi) outside it, one half at a time (in pivot order)
  auto laswp_cols = [&](int b0, int b1, int c_lo, int c_hi, hipStream_t s) -> hipError_t {
    const int bl = min(max(b0, c_lo), c_hi), bh = min(max(b1, c_lo), c_hi);
    const int ncols = (c_hi - c_lo) - (bh - bl);
    if (ncols > 0) {
      constexpr int CB = sizeof(T) == 4 ? 64 : 32;
      for (int h0 = b0; h0 < b1; h0 += NBO) {
        hipLaunchKernelGGL(k_laswp<T>, dim3((ncols + CB - 1) / CB, batch), dim3(256), 0, s, mb, ipiv, strideP, h0,
                           min(b1, h0 + NBO), b0, b1, c_lo, c_hi, skip);
        CHK(hipGetLastError());
      }
    }
    return hipSuccess;
  };
  // U12 = L11^-1 A12 for the rows of block [b0, b1) and the columns [c, ce), by halves
  auto solve_u12 = [&](int b0, int b1, int c, int ce, hipStream_t s) -> hipError_t {
    if (ce <= c) return hipSuccess;
    CHK(solve_upper_rows(mb, batch, b0, c, ce - c, W, strideW, skip, s));
    if (b1 - b0 > NBO) {
      const int h0 = b0 + NBO;
      GemmShape g = {h0, b0, b0, c, h0, c, b1 - h0, ce - c, NBO};
      CHK(launch_gemm<T>(mb, batch, g, skip, s, split, 0, tile));
      CHK(solve_upper_rows(mb, batch, h0, c, ce - c, W, strideW, skip, s));
    }
    return hipSuccess;
  };

  // Blocks go in pairs P = [k0, k1), Q = [k1, k2) whose trailing updates are applied together (K = 256), so the
  // trailing matrix is read and written once per pair instead of once per block. P updates only Q's columns
  // before Q is factored; P's update of Q's rows beyond Q waits until Q's interchanges have been applied to
  // them (every row of the columns beyond Q is then still without P's update, so the interchanges commute with
  // the deferred update, which uses P's equally interchanged L rows). Look-ahead (a second stream): everything
  // on the columns beyond the next block R = [k2, k3) -- the pair's interchanges there, its U12 rows there and
  // the K = 256 update there -- runs on the look-ahead stream, on at most la->gemm_wgs workgroups for the
  // update, beside the chain on the main stream (Q, then R, factored with their columns). The main stream waits
  // for it before it touches those columns again, at the next pair.
  bool pending = false;  // look-ahead work on the columns beyond the next block is in flight
  CHK(factor_block(0, min(N, NBT)));
  int k0 = 0;
  while (true) {
    const int k1 = min(N, k0 + NBT), k2 = min(N, k1 + NBT), k3 = min(N, k2 + NBT);  // P factored
    if (pending) {
      CHK(hipStreamWaitEvent(st, la->done, 0));
      pending = false;
    }
    const bool far = look && k3 < N;  // the columns [k3, N) go to the look-ahead stream
    const int ce = far ? k3 : N;
    hipStream_t as = far ? la->stream : st;
    if (far) {  // P factored on the main stream: its interchanges and U12 rows on the far columns, and its
                // interchanges on the L columns of the blocks before it (read by nothing until the solves)
      CHK(hipEventRecord(la->ready, st));
      CHK(hipStreamWaitEvent(as, la->ready, 0));
      CHK(laswp_cols(k0, k1, k3, N, as));
      CHK(solve_u12(k0, k1, k3, N, as));
      CHK(laswp_cols(k0, k1, 0, k0, as));
    }
    CHK(laswp_cols(k0, k1, far ? k1 : 0, ce, st));
    if (k1 >= N) break;
    CHK(solve_u12(k0, k1, k1, ce, st));
    {
      GemmShape g = {k1, k0, k0, k1, k1, k1, N - k1, k2 - k1, k1 - k0};  // P's update of Q's columns
      CHK(launch_gemm<T>(mb, batch, g, skip, st, split, 0, tile));
    }
    CHK(factor_block(k1, k2));
    if (k2 >= N) {  // Q is the last block: its interchanges at the top of the loop, then done
      k0 = k1;
      continue;
    }
    CHK(laswp_cols(k1, k2, far ? k0 : 0, ce, st));
    if (far) {  // Q's interchanges, P's update of Q's rows and Q's U12 rows on the far columns; Q's
                // interchanges on the L columns before P
      CHK(hipEventRecord(la->ready, st));
      CHK(hipStreamWaitEvent(as, la->ready, 0));
      CHK(laswp_cols(k1, k2, k3, N, as));
      GemmShape g = {k1, k0, k0, k3, k1, k3, k2 - k1, N - k3, k1 - k0};
      CHK(launch_gemm<T>(mb, batch, g, skip, as, split, 0, tile));
      CHK(solve_u12(k1, k2, k3, N, as));
      CHK(laswp_cols(k1, k2, 0, k0, as));
    }
    {
      GemmShape g = {k1, k0, k0, k2, k1, k2, k2 - k1, ce - k2, k1 - k0};  // P's update of Q's rows beyond Q
      CHK(launch_gemm<T>(mb, batch, g, skip, st, split, 0, tile));
    }
    CHK(solve_u12(k1, k2, k2, ce, st));
    // P's and Q's updates of the rows and columns beyond Q in one pass, K = k2 - k0
    {
      GemmShape g = {k2, k0, k0, k2, k2, k2, N - k2, ce - k2, k2 - k0};
      CHK(launch_gemm<T>(mb, batch, g, skip, st, split, 0, tile));
    }
    if (far) {
      GemmShape gr = {k2, k0, k0, k3, k2, k3, N - k2, N - k3, k2 - k0};
      CHK(launch_gemm<T>(mb, batch, gr, skip, as, split, la->gemm_wgs, tile));
      CHK(hipEventRecord(la->done, as));
      pending = true;
    }
    CHK(factor_block(k2, k3));
    k0 = k2;
  }
  if (pending) CHK(hipStreamWaitEvent(st, la->done, 0));
  // inverses of the diagonal blocks of U for the backward sweeps
  hipLaunchKernelGGL(k_tri_inv_upper<T>, dim3((N + 63) / 64, batch), dim3(256), 0, st, mb, W, strideW, skip);
  CHK(hipGetLastError());
  return hipSuccess;
}

// One workgroup per matrix walks the whole chain while the matrices fill the CUs (small N or large batches);
// otherwise the blocked sweeps spread each block's update over TSW_ROWS-row tiles, one launch per block.
template <typename T, int NR>
static hipError_t solve_sweeps(const MatBatch<T>& mb, int batch, const int* perm, long long strideP,
                               const cplx<double>* b, long long ldb, long long strideB, cplx<T>* y, long long ldy,
                               long long strideY, const int* skip, const cplx<T>* W, long long strideW,
                               hipStream_t st, int wg_min_batch) {
  const int N = mb.N, nblk = (N + 63) / 64;
  if (N <= TSW_MIN_N && wg_min_batch > 0 && batch >= wg_min_batch) {
    hipLaunchKernelGGL((k_trsv_wg<T, NR>), dim3(batch), dim3(1024), 0, st, mb, b, ldb, strideB, y, ldy, strideY, perm,
                       strideP, skip, W, strideW);
    return hipGetLastError();
  }
  hipLaunchKernelGGL((k_perm_rhs<T, NR>), dim3((N + 255) / 256, batch), dim3(256), 0, st, b, ldb, strideB, y, ldy,
                     strideY, perm, strideP, N, skip);
  for (int pass = 0; pass < 2; ++pass) {
    const int lower = pass == 0;
    for (int s = 0; s < nblk; ++s) {
      const int kb = lower ? s : nblk - 1 - s;
      const int prev = s == 0 ? -1 : (lower ? kb - 1 : kb + 1);
      const int nupd = lower ? max(0, N - (kb + 1) * 64) : kb * 64;
      const int tiles = max(1, (nupd + TSW_ROWS - 1) / TSW_ROWS);
      hipLaunchKernelGGL((k_trsv_step<T, NR>), dim3(tiles, batch), dim3(256), 0, st, mb, y, ldy, strideY, skip, W,
                         strideW, kb, prev, lower);
    }
  }
  return hipGetLastError();
}

template <typename T>
hipError_t lu_solve_batch(const MatBatch<T>& mb, int batch, const int* perm, long long strideP,
                          const cplx<double>* b, long long ldb, long long strideB, cplx<T>* y, long long ldy,
                          long long strideY, int nrhs, const int* skip, const cplx<T>* W, hipStream_t st,
                          int wg_min_batch) {
  const long long strideW = lu_workspace_elems(mb.N);
  switch (nrhs) {
    case 1: return solve_sweeps<T, 1>(mb, batch, perm, strideP, b, ldb, strideB, y, ldy, strideY, skip, W, strideW, st, wg_min_batch);
    case 2: return solve_sweeps<T, 2>(mb, batch, perm, strideP, b, ldb, strideB, y, ldy, strideY, skip, W, strideW, st, wg_min_batch);
    case 3: return solve_sweeps<T, 3>(mb, batch, perm, strideP, b, ldb, strideB, y, ldy, strideY, skip, W, strideW, st, wg_min_batch);
    case 4: return solve_sweeps<T, 4>(mb, batch, perm, strideP, b, ldb, strideB, y, ldy, strideY, skip, W, strideW, st, wg_min_batch);
    default: return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t residual_batch(const MatBatch<T>& mb, int batch, const cplx<double>* b, const cplx<double>* x,
                          cplx<double>* r, long long ldv, long long strideV, int nrhs, const int* skip,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_residual<T>, dim3((mb.N + 3) / 4, batch), dim3(256), 0, st, mb, b, x, r, ldv, strideV, nrhs,
                     skip);
  return hipGetLastError();
}

hipError_t resnorm_batch(const cplx<double>* r, const cplx<double>* b, long long ldv, long long strideV, int N,
                         int nrhs, double tol, double* rel, int* done, int iter, int* iters, int batch,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_resnorm, dim3(batch), dim3(256), 0, st, r, b, ldv, strideV, N, nrhs, tol, rel, done, iter,
                     iters);
  return hipGetLastError();
}

template <typename T>
hipError_t update_batch(cplx<double>* x, const cplx<T>* d, long long ldv, long long ldd, long long strideV,
                        long long strideD, int N, int nrhs, int init, const int* skip, int batch, hipStream_t st) {
  hipLaunchKernelGGL(k_update<T>, dim3((N + 255) / 256, nrhs, batch), dim3(256), 0, st, x, d, ldv, ldd, strideV,
                     strideD, N, nrhs, init, skip);
  return hipGetLastError();
}

template <typename T>
hipError_t growth_batch(const MatBatch<T>& mb, int batch, double* umax, const double* amax, double* growth,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_umax<T>, dim3(mb.N, batch), dim3(256), 0, st, mb, umax);
  CHK(hipGetLastError());
  hipLaunchKernelGGL(k_growth, dim3(batch), dim3(256), 0, st, umax, amax, mb.N, growth);
  return hipGetLastError();
}

// Outcome of a batch of solves, one workgroup per matrix: flags[m] bit 0 = LU breakdown (info[m] != 0), bit 1 =
// a non-finite solution entry, bit 2 = relative residual above tol (or NaN), bit 3 = pivot growth above gmax (or
// NaN). rel and growth may be null (not checked).
__global__ __launch_bounds__(256) void k_solve_status(const int* info, const cplx<double>* x, long long ldx,
                                                      long long strideX, int N, int nrhs, const double* rel,
                                                      double tol, const double* growth, double gmax, int* flags) {
  const int mat = blockIdx.x, tid = threadIdx.x;
  __shared__ int bad;
  if (tid == 0) bad = 0;
  __syncthreads();
  int local = 0;
  for (int c = 0; c < nrhs; ++c) {
    const cplx<double>* xc = x + (size_t)mat * strideX + (size_t)c * ldx;
    for (int i = tid; i < N; i += 256)
      if (!isfinite(xc[i].re) || !isfinite(xc[i].im)) local = 1;
  }
  if (local) atomicOr(&bad, 1);
  __syncthreads();
  if (tid == 0) {
    int f = 0;
    if (info[mat] != 0) f |= 1;
    if (bad) f |= 2;
    if (rel && !(rel[mat] <= tol)) f |= 4;
    if (growth && !(growth[mat] <= gmax)) f |= 8;
    flags[mat] = f;
  }
}

hipError_t solve_status_batch(const int* info, const cplx<double>* x, long long ldx, long long strideX, int N,
                              int nrhs, const double* rel, double tol, const double* growth, double gmax, int* flags,
                              int batch, hipStream_t st) {
  hipLaunchKernelGGL(k_solve_status, dim3(batch), dim3(256), 0, st, info, x, ldx, strideX, N, nrhs, rel, tol, growth,
                     gmax, flags);
  return hipGetLastError();
}

template hipError_t lu_factor_batch<float>(const MatBatch<float>&, int, int*, int*, long long, int*, const int*,
                                           cplx<float>*, hipStream_t, const LuLookahead*, int);
template hipError_t lu_factor_batch<double>(const MatBatch<double>&, int, int*, int*, long long, int*, const int*,
                                            cplx<double>*, hipStream_t, const LuLookahead*, int);
template hipError_t lu_solve_batch<float>(const MatBatch<float>&, int, const int*, long long, const cplx<double>*,
                                          long long, long long, cplx<float>*, long long, long long, int, const int*,
                                          const cplx<float>*, hipStream_t, int);
template hipError_t lu_solve_batch<double>(const MatBatch<double>&, int, const int*, long long,
                                           const cplx<double>*, long long, long long, cplx<double>*, long long,
                                           long long, int, const int*, const cplx<double>*, hipStream_t, int);
template hipError_t residual_batch<float>(const MatBatch<float>&, int, const cplx<double>*, const cplx<double>*,
                                          cplx<double>*, long long, long long, int, const int*, hipStream_t);
template hipError_t residual_batch<double>(const MatBatch<double>&, int, const cplx<double>*,
                                           const cplx<double>*, cplx<double>*, long long, long long, int,
                                           const int*, hipStream_t);
template hipError_t update_batch<float>(cplx<double>*, const cplx<float>*, long long, long long, long long,
                                        long long, int, int, int, const int*, int, hipStream_t);
template hipError_t update_batch<double>(cplx<double>*, const cplx<double>*, long long, long long, long long,
                                         long long, int, int, int, const int*, int, hipStream_t);
template hipError_t growth_batch<float>(const MatBatch<float>&, int, double*, const double*, double*,
                                        hipStream_t);
template hipError_t growth_batch<double>(const MatBatch<double>&, int, double*, const double*, double*,
                                         hipStream_t);
