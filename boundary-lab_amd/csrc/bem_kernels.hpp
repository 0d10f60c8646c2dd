// Kernel parameter blocks and launcher declarations shared by the translation units of libbemhip.
#pragma once

#include "bem_common.hpp"

#define BEM_SLOT_STRIDE 196  // per-chunk slot_start entries (<= 192 slots + 1, padded)
#define BEM_SLOT_INC 192     // per-chunk incidence entries (64 elements x 3 columns)
#define BEM_PEND_SLOTS 96    // LDS partial strip sums per regular-assembly wave
#define BEM_PEND_NONE 0xFFFF
#define BEM_PEND_FROM_LDS (1 << 16)
#define BEM_PEND_TO_STRIP (1 << 17)
#define BEM_MAX_CH 4
#define BEM_NBO 64           // outer LU panel width (trailing-update K)

// Per-element record used by the regular kernels (32 T values, 128 B in FP32).
template <typename T>
struct ElemRec {
  T q[18];    // the regular rule's points (x,y,z): six (order 4), three (order 2) or one (order 1)
  T n[3];     // unit normal
  T curl[9];  // P1 surface curls, 3 vectors
  T a2;       // 2 * area
  T pad;
};

// Field source point (regular rule point of a face) with its face normal.
template <typename T>
struct SrcRec {
  T y[3];
  T n[3];
};

template <typename T>
struct RegParams {
  const ElemRec<T>* rec;
  const int4* fv;
  const int* order;             // [n_chunks*64] trial element per (chunk, lane), -1 = none
  const int* slot_count;        // [n_chunks]
  const uint16_t* slot_start;   // [n_chunks][BEM_SLOT_STRIDE]
  const uint16_t* slot_inc;     // [n_chunks][BEM_SLOT_INC]  (lane << 2) | m
  const uint32_t* slot_bytes;   // [n_chunks][128] bytes: slot starts at 0.., incidence codes at 256..
  const int* slot_vertex;       // [n_chunks][BEM_SLOT_INC]  2 * vertex + (earlier chunk stored the column)
  const int* slot_pend;         // [n_chunks][BEM_SLOT_INC]  LDS slot | BEM_PEND_FROM_LDS | BEM_PEND_TO_STRIP
  const int* slot_prev;         // [n_chunks][BEM_SLOT_INC]  previous chunk holding the vertex (-1: none)
  const int* elem_pos;          // [F] position in the trial order
  const int* chunk_start;       // [F] first chunk with the element or an element sharing a vertex
  cplx<T>* strip2;              // paired assembly: column strip [nf][3F][L] (transposed (s, t) blocks)
  const int* sing_ptr;          // [F+1]
  const int* sing_trial;        // [P]
  const cplx<double>* sing_lhs; // [nf][P][9]
  const double* k;              // [nf]
  cplx<T>* strip;               // [nf][3F][L], L = V padded to 64: one column per vertex
  long long strip_fstride;
  int L, F, n_chunks, P;
  int R;                        // transforms (identity + images); rec holds R*F records, image r at r*F
  const int* pmask;             // [F] bit r: face in an image pair of transform r (null: no image pairs)
  int nq;                       // points of the regular rule (6, 3 or 1); the paired kernel needs 6
};

struct SingParams {
  const double* fvtx;   // [F][9]
  const double* fnrm;   // [F][3]
  const double* fcurl;  // [F][9]
  const double* fa2;    // [F]
  const int* test;      // [P]
  const int* trial;     // [P]
  const int* rule;      // [P]
  const int* by_rule;   // [P] pair indices sorted by rule
  const int* rule_off;  // [n_rules]
  const int* rule_len;  // [n_rules]
  const double* rpts;   // [total][5]: test xi, eta, trial xi, eta, weight
  const double* k;      // [nf]
  cplx<double>* lhs;    // [nf][P][9]
  cplx<double>* rhs;    // [nf][P][3]
  int P;
  double m11[9];        // reference-element mass block / (2A)
  double m10[3];
  int F;                // trial key = s + r*F: element s reflected by transform r
  int sgn[12];          // transform sign vectors
};

template <typename T>
struct RhsParams {
  const ElemRec<T>* rec;
  const int4* fv;
  const int* src;                // [n_src] radiating faces
  const cplx<double>* qunit;     // [n_ch][F]
  const int* sing_ptr;
  const int* sing_trial;
  const cplx<double>* sing_rhs;  // [nf][P][3]
  const double* k;
  cplx<double>* out;             // [nf][3F][BEM_MAX_CH]
  int F, P, n_src, n_ch;
  int R;                         // transforms; rec holds R*F records
  const int* pmask;              // [F] image-pair participation bits (null: none)
  int nq;                        // points of the regular rule (6, 3 or 1)
};

template <typename T>
struct GatherParams {
  const cplx<T>* strip;
  long long strip_fstride;
  int L;
  const int* vstar_ptr;           // [N+1]
  const int* vstar_row;           // strip rows 3t+l
  const cplx<double>* rhs_part;   // [nf][3F][BEM_MAX_CH]
  const cplx<double>* rhs_scale;  // [nf]  (i rho omega)
  cplx<T>* A;                     // [nf] row-major N x lda
  cplx<T>* A0;                    // optional copy (may be null)
  long long lda, strideA;
  cplx<double>* b;                // [nf][n_ch][ldb]
  long long ldb, strideB;
  double* rowmax;                 // [nf][N]
  int N, F, n_ch;
  const double* roww;             // [N] P1 orbit row weights (null: 1)
  // paired assembly (strip2 non-null): the column strip and the validity of strip columns
  const cplx<T>* strip2;          // [nf][3F][L]
  const int* elem_pos;            // [F]
  const int* chunk_start;         // [F]
  const int* vertex_last;         // [N]
  const int4* vstar_ent;          // per star entry: {strip row, first chunk of its row strip, of its column
                                  // strip, the star's vertex}
};

template <typename T>
struct MatBatch {
  cplx<T>* A;
  long long lda, strideA;
  int N;
};

template <typename T>
hipError_t launch_regular(const RegParams<T>&, int nf, hipStream_t);
// fp32: Duffy sums in FP32 arithmetic (blocks still written as FP64); else FP64 throughout
hipError_t launch_singular(const SingParams&, int nf, bool fp32, hipStream_t);
template <typename T>
hipError_t launch_rhs(const RhsParams<T>&, int nf, hipStream_t);
// p.out is [nf][3F][p.n_src]: the RHS-operator block of every (test element, listed source face)
template <typename T>
hipError_t launch_rhs_cols(const RhsParams<T>&, int nf, hipStream_t);

// Coupled-system blocks (bem_coupled.hip).
// out[f] (cT, row-major, ldo) row i < N, column c < n_cols: alpha[f] * w_i * sum over strip rows r of star(i)
// and entries e of column c of col_w[e] * part[f][r][col_idx[e]] (part: [nf][3F][n_src] from launch_rhs_cols).
template <typename T>
hipError_t cols_gather(const cplx<double>* part, int F, int n_src, const int* vstar_ptr, const int* vstar_row,
                       const double* roww, int N, int n_cols, const int* col_ptr, const int* col_idx,
                       const cplx<double>* col_w, const cplx<double>* alpha, cplx<T>* out, long long ldo,
                       long long strideO, int nf, hipStream_t st);
// out[rows[e] * ldo + cols[e]] (=, or += when add) alpha * vals[e] for e < nnz (entries unique per call).
template <typename T>
hipError_t scatter_coo(long long nnz, const int* rows, const int* cols, const cplx<double>* vals,
                       cplx<double> alpha, int add, cplx<T>* out, long long ldo, hipStream_t st);
// Rows i < n of A (ncols columns) scaled by powers of two so that max |re|+|im| lies in [1/2, 1); scale[i] receives
// the factor.
template <typename T>
hipError_t row_equilibrate(cplx<T>* A, long long lda, int n, int ncols, double* scale, hipStream_t st);
// out[i * ldo + j] (=, or += when add) alpha * src[i * lds + j] (src null: alpha) for i < m, j < n.
template <typename T>
hipError_t dense_block(int m, int n, const cplx<double>* src, long long lds, cplx<double> alpha, int add,
                       cplx<T>* out, long long ldo, hipStream_t st);
template <typename T>
hipError_t launch_gather(const GatherParams<T>&, int nf, hipStream_t);
template <typename T>
hipError_t launch_gather_rhs(const GatherParams<T>&, int nf, hipStream_t);
hipError_t neumann_batch(const cplx<double>* scale, const cplx<double>* qunit, cplx<double>* q, int n_ch, int F,
                         int nf, hipStream_t);
hipError_t solve_status_batch(const int* info, const cplx<double>* x, long long ldx, long long strideX, int N,
                              int nrhs, const double* rel, double tol, const double* growth, double gmax, int* flags,
                              int batch, hipStream_t);

// elements (cT) of LU workspace per matrix: inverted 64 x 64 diagonal blocks of L and U
long long lu_workspace_elems(int N);
// Second stream for the look-ahead factorization (stream null: everything on the main stream). `ready` and
// `done` are events private to this factorization (recorded on the main / look-ahead stream).
struct LuLookahead {
  hipStream_t stream;
  hipEvent_t ready, done;
  int gemm_wgs;     // cap on the workgroups of the look-ahead update (0: none)
  int split_tile;   // 0: 256 x 128 tiles of 16 waves for the large split-bf16 updates, 1: 128 x 64 tiles everywhere
};
template <typename T>
hipError_t lu_factor_batch(const MatBatch<T>&, int batch, int* ipiv, int* perm, long long strideP, int* info,
                           const int* skip, cplx<T>* W, hipStream_t, const LuLookahead* la, int products);
// Triangular solves: one workgroup per matrix when N <= 4096 and batch >= wg_min_batch (0: never), else the
// blocked multi-workgroup sweeps.
constexpr int BEM_TSW_MIN_BATCH_DEFAULT = 8;
template <typename T>
hipError_t lu_solve_batch(const MatBatch<T>&, int batch, const int* perm, long long strideP, const cplx<double>* b,
                          long long ldb, long long strideB, cplx<T>* y, long long ldy, long long strideY, int nrhs,
                          const int* skip, const cplx<T>* W, hipStream_t, int wg_min_batch = BEM_TSW_MIN_BATCH_DEFAULT);
template <typename T>
hipError_t residual_batch(const MatBatch<T>&, int batch, const cplx<double>* b, const cplx<double>* x,
                          cplx<double>* r, long long ldv, long long strideV, int nrhs, const int* skip, hipStream_t);
hipError_t resnorm_batch(const cplx<double>* r, const cplx<double>* b, long long ldv, long long strideV, int N,
                         int nrhs, double tol, double* rel, int* done, int iter, int* iters, int batch, hipStream_t);
template <typename T>
hipError_t update_batch(cplx<double>* x, const cplx<T>* d, long long ldv, long long ldd, long long strideV,
                        long long strideD, int N, int nrhs, int init, const int* skip, int batch, hipStream_t);
template <typename T>
hipError_t growth_batch(const MatBatch<T>&, int batch, double* umax, const double* amax, double* growth,
                        hipStream_t);
hipError_t fill_int(int* p, int n, int value, hipStream_t);

template <typename T>
hipError_t field_batch(const SrcRec<T>* src, const int4* fv, const T* src_w, const double* obs, int M, int M0, int n_obs,
                       const double* k, const cplx<double>* p, long long ldp, long long strideP,
                       const cplx<double>* q, long long ldq, long long strideQ, int n_ch, int nf, cplx<T>* wsrc,
                       cplx<double>* partial, int n_slices, cplx<double>* out, long long ldo, long long strideO,
                       int nq, hipStream_t st);
