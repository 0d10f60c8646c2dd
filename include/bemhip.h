/*
 * bemhip.h — C ABI of the MI355X (gfx950) exterior Helmholtz BEM library (libbemhip.so).
 *
 * This is the drop-in seam for the reference's `beat_rocm` device path. The reference binds no C library:
 * its Julia worker calls AMDGPU.jl kernels and rocBLAS/rocSOLVER directly. Each entry point below replaces
 * one of those Julia entry points and cites it (paths relative to the reference repository root;
 * JS = src/blab/solvers/julia_local/src).
 *
 * Conventions
 *  - Every function returns an int status (BEM_OK == 0) and never throws; the message of the last failure on
 *    the calling thread is available through bem_last_error().
 *  - Device buffers are caller-owned and passed as raw device pointers plus leading dimensions and batch
 *    strides in ELEMENTS. Complex numbers are interleaved (re, im). "c128" = complex double, "cT" = complex
 *    float for BEM_PREC_MIXED and complex double for BEM_PREC_FP64. Matrices are row-major.
 *  - Handles (bem_ctx, bem_mesh) are opaque and owned by the library; no ownership crosses the ABI.
 *  - One HIP stream per context; calls on one context are not reentrant. All work is enqueued on that stream
 *    and is asynchronous unless the function says otherwise.
 */
#ifndef BEMHIP_H
#define BEMHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BEM_OK 0
#define BEM_E_ARG 1      /* invalid argument */
#define BEM_E_HIP 2      /* HIP runtime error */
#define BEM_E_NOMEM 3    /* device allocation failed */
#define BEM_E_STATE 4    /* operation not allowed in the current state (e.g. allocation during capture) */
#define BEM_E_SINGULAR 5 /* LU breakdown (zero pivot) or a non-finite solution (bem_solve_status) */
#define BEM_E_ACCURACY 6 /* refinement short of its tolerance or pivot growth above the limit (bem_solve_status) */

#define BEM_PREC_FP64 0  /* FP64 assembly, FP64 LU, FP64 field */
#define BEM_PREC_MIXED 1 /* FP32 regular assembly + FP64 singular, FP32 LU (FP32 MFMA), FP64 refinement */

#define BEM_MAX_CHANNELS 4
#define BEM_LU_MAX_N 24576 /* largest N bem_lu_factor accepts (24 row sets of 1024 per panel thread) */

typedef struct bem_ctx bem_ctx;
typedef struct bem_mesh bem_mesh;
typedef struct bem_colset bem_colset;
typedef struct bem_sparse bem_sparse;

/* Library version (major*10000 + minor*100 + patch). */
int bem_version(void);

/* Copy the calling thread's last error message into buf (always NUL-terminated). Returns its length. */
int bem_last_error(char* buf, int buflen);

/* Create a context on `device`. `stream` is a hipStream_t to enqueue on, or NULL to create one
 * (default priority). Replaces the device selection / AMDGPU.functional() checks of
 * JS/BeatEngineRocmCommon.jl:1-85. */
int bem_ctx_create(int device, void* stream, bem_ctx** out);
int bem_ctx_destroy(bem_ctx* ctx);
/* Block until all work on the context's stream is done (AMDGPU.synchronize(),
 * JS/BeatEngineRocmAssembly.jl:85). */
int bem_ctx_sync(bem_ctx* ctx);
/* The hipStream_t used by the context. */
int bem_ctx_stream(bem_ctx* ctx, void** stream);

/* Stream fan-out (no reference counterpart; tuning), on default-priority streams forked from and joined back
 * into the context stream. Refinement splits a batch of independent matrices into up to `streams` contiguous
 * groups (at least 4 matrices each); LU factorization uses a second stream for look-ahead when streams >= 2
 * (same results bit for bit). Default 4; 1 runs everything on the context stream. */
int bem_ctx_set_fanout(bem_ctx* ctx, int streams);

/* Products of the large LU updates of BEM_PREC_MIXED factorizations on this context (no reference counterpart:
 * the reference factors in its working precision). BEM_LU_FP32 (default): FP32 MFMA. BEM_LU_SPLIT_BF16: each
 * FP32 product as three bf16 MFMA products (x = hi + lo, a b ~ a_hi b_hi + a_hi b_lo + a_lo b_hi, FP32
 * accumulation), relative error near 2^-16 per product, for factors that precondition the FP64 refinement of
 * bem_refine. BEM_LU_BF16: one bf16 MFMA product per FP32 product (relative error near 2^-8 per product, about
 * one more refinement step, a third of the MFMA work). The sub-panel updates, panels and triangular solves stay
 * FP32. Ignored for BEM_PREC_FP64. */
#define BEM_LU_FP32 0
#define BEM_LU_SPLIT_BF16 1
#define BEM_LU_BF16 2
int bem_ctx_set_lu_products(bem_ctx* ctx, int mode);

/* Arithmetic of the Duffy-rule sums of the vertex-sharing element pairs in BEM_PREC_MIXED assemblies, per context.
 * BEM_SING_WORKING (default): FP32, the working precision, as the reference's ROCm singular-correction kernels
 * (JS/BeatEngineRocmSingular.jl:1-289, instantiated with the assembly's element type); the element geometry is
 * formed in FP64 and rounded once, the blocks are written as FP64. BEM_SING_FP64: FP64 sums. FP64 assemblies always
 * sum in FP64. Refused during graph capture. */
#define BEM_SING_WORKING 0
#define BEM_SING_FP64 1
int bem_ctx_set_singular_precision(bem_ctx* ctx, int mode);

/* Launch-shape choices (no reference counterpart: tuning), per context; the project's tuner (beat_hip/autotune.py)
 * times candidates on the device and the engine applies its table per (precision, N, batch). Captured graphs keep
 * the shapes of the tuning in force when they were captured.
 *  split_tile:         0: the large split-bf16 LU updates (M >= 512, N >= 256) on 256 x 128 tiles of 16 waves,
 *                      smaller ones on 128 x 64 tiles of 4 waves; 1: 128 x 64 tiles everywhere.
 *  trsv_wg_min_batch:  triangular solves of N <= 4096 run one workgroup per matrix when at least this many matrices
 *                      share the launch (0: never; default 8), else as blocked multi-workgroup sweeps.
 *  lookahead_wgs:      workgroups of the look-ahead trailing update: 0 = max(CUs / 2, CUs - batch) (default),
 *                      > 0 that many, -1 uncapped.
 *  field_slice_budget: source slices x observation blocks x frequencies per field launch (0: 2048). */
typedef struct bem_tuning {
  int split_tile;
  int trsv_wg_min_batch;
  int lookahead_wgs;
  int field_slice_budget;
} bem_tuning;
int bem_ctx_set_tuning(bem_ctx* ctx, const bem_tuning* tuning);
int bem_ctx_get_tuning(bem_ctx* ctx, bem_tuning* tuning);

/* Kernel timers (no reference counterpart; diagnostics for roofline measurement). While enabled, each timed
 * launch group is bracketed by HIP events on the context's stream and waited for (not during capture).
 * bem_ctx_timer_read returns accumulated milliseconds, accumulated work units (frequencies or matrices) and the
 * number of timed regions (one per kernel launch for the assembly slots; one per call for LU, solve, field). */
#define BEM_TIMER_REGULAR 0  /* regular-pair assembly kernel */
#define BEM_TIMER_SINGULAR 1 /* Duffy singular-pair kernel */
#define BEM_TIMER_GATHER 2   /* row gather */
#define BEM_TIMER_RHS 3      /* RHS partial sums */
#define BEM_TIMER_LU 4       /* whole LU factorization sequence */
#define BEM_TIMER_FIELD 5    /* field evaluation sequence */
#define BEM_TIMER_SOLVE 6    /* triangular solves + refinement (bem_lu_solve, bem_refine) */
int bem_ctx_timers(bem_ctx* ctx, int enable);
int bem_ctx_timer_read(bem_ctx* ctx, int slot, double* ms, long long* units, long long* regions);

/* Host-side description of a prepared boundary mesh (all pointers are HOST memory, copied by
 * bem_mesh_create). Indices are 0-based. */
typedef struct bem_mesh_desc {
  int n_faces;               /* F  (DP0 dofs) */
  int n_vertices;            /* V  (P1 dofs, N of the dense system) */
  const int* faces;          /* [F][3] vertex ids */
  const double* face_vertices; /* [F][9] */
  const double* normals;     /* [F][3] unit normals, cross(v2-v1, v3-v1) */
  const double* curls;       /* [F][9] P1 surface curls */
  const double* area2;       /* [F] 2*area */
  const double* qpoints;     /* [F][18] regular rule points (n_qp of them) */
  /* regular-assembly tables: trial elements in locality order, chunks of 64, slot tables */
  int n_chunks;
  const int* order;          /* [n_chunks*64], -1 padding */
  const int* slot_count;     /* [n_chunks] */
  const uint16_t* slot_start;/* [n_chunks][196] */
  const uint16_t* slot_inc;  /* [n_chunks][192], (lane << 2) | local column */
  const int* slot_vertex;    /* [n_chunks][192]: 2 * vertex + 1 if an earlier chunk stored the vertex column */
  const int* slot_pend;      /* [n_chunks][192]: LDS partial-sum slot (0xFFFF: none) | 1 << 16 (an earlier chunk
                                left partial sums there) | 1 << 17 (store the sums to the strip) */
  const int* slot_prev;      /* [n_chunks][192]: previous chunk holding the slot's vertex, -1 if none */
  const int* elem_pos;       /* [F] position of each element in the trial order */
  const int* chunk_start;    /* [F] first chunk holding the element or an element sharing a vertex with it */
  const int* vertex_last;    /* [V] last chunk holding each vertex, -1 if unused */
  const int* vstar_ptr;      /* [V+1] vertex -> strip rows (3t+l) CSR */
  const int* vstar_row;
  /* singular (vertex-sharing) pairs, grouped by test element */
  int n_pairs;
  const int* sing_test;      /* [P] */
  const int* sing_trial;     /* [P] */
  const int* sing_rule;      /* [P] */
  const int* sing_ptr;       /* [F+1] */
  int n_rules;
  const int* rule_off;       /* [n_rules] */
  const int* rule_len;       /* [n_rules] */
  const double* rule_pts;    /* [sum len][5]: test xi, eta, trial xi, eta, weight */
  double m11[9];             /* reference-element mass block, sum_q phi_l phi_m w_q */
  double m10[3];             /* sum_q phi_l w_q */
  /* radiating faces (faces with a non-zero Neumann drive in any channel) */
  int n_src;
  const int* src_faces;
  /* x / xy plane symmetry (JS/BeatEngineCore.jl:317-439,917-945; JS/BeatEngineRocmRegular.jl:103-132;
   * JS/BeatEngineRocmPairKernels.jl:442-466; JS/BeatEngineRocmOperators.jl:11-21). The mesh is the
   * fundamental domain. n_transforms = 1 (off), 2 (x) or 4 (xy); transform r reflects by
   * transform_signs[3r .. 3r+2] (+-1; transform 0 is the identity). Every trial element is integrated together
   * with its images; singular pairs whose trial key sing_trial >= n_faces are image pairs (trial element
   * key % n_faces reflected by transform key / n_faces) integrated with their Duffy rule. */
  int n_transforms;
  int transform_signs[12];
  const double* row_weights; /* [V] P1 orbit row weights applied to A and b (NULL: 1) */
  const int* plane_mask;     /* [F] bit r: the face takes part in an image pair of transform r (NULL: none) */
  /* Points of the regular triangle rule (triangle_rule, JS/BeatEngineCore.jl:517-549): 6 (order 4; 0 means 6),
   * 3 (orders 2 and 3) or 1 (order 1). qpoints holds that many points per face (the rest of its 18 values
   * unused) and m11 / m10 must be the identity blocks under the same rule. */
  int n_qp;
} bem_mesh_desc;

/* Check every index table of a mesh description on the host (no device needed): face vertex ids < V, the trial
 * order a permutation of the faces padded with -1 and inverted by elem_pos, slot tables, vertex_last, the
 * vertex -> strip-row CSR (each strip row 3t+l exactly once, under the vertex of corner l of face t), singular
 * pairs grouped by test element with trial keys < n_transforms * F and rule indices < n_rules, rule ranges and
 * radiating faces. BEM_E_ARG with the reason when anything is out of range; bem_mesh_create runs it first, so a
 * malformed description never reaches a kernel. Part of build_rocm_regular_assembly_cache's input contract
 * (JS/BeatEngineRocmRegular.jl:134-205), which the reference trusts unchecked. */
int bem_mesh_validate(const bem_mesh_desc* desc);

/* Host-side tables (no device needed; HOST arrays). bem_host_vertex_pairs lists every (test, trial) element pair
 * sharing at least one vertex id, self pairs included, grouped by test element with trials ascending (at most
 * `capacity`), and classifies it by its Duffy orientation code: 100000 coincident; 200000 + 1000 ia + 100 ib +
 * 10 ja + jb for a shared edge (test corners ia, ib and trial corners ja < jb of the two matches, taken in
 * test-corner-major scan order); 300000 + 1000 ia + 10 ja for a shared vertex. codes (int[64]) receives the
 * distinct codes ascending, rule[p] the rank of pair p's code among them. Replaces adjacency_info /
 * singular_orientation_key / build_singular_correction_cache (JS/BeatEngineCore.jl:441-507,784-915). */
int bem_host_vertex_pairs(int n_faces, int n_vertices, const int* faces, long long capacity, int* test, int* trial,
                          int* rule, int* codes, int* n_codes, long long* n_pairs);
/* The regular assembly's chunk tables for a trial order (`order`: n_chunks*64 face ids padded with -1) and
 * `pend_slots` LDS partial-sum slots per wave: the bem_mesh_desc fields slot_count .. vstar_row (sizes as there).
 * No reference counterpart: it replaces the greedy colouring of the reference's pair kernels
 * (_beat_cpu_element_color_groups, JS/BeatEngineCpuAssembly.jl:102-130; JS/BeatEngineRocmRegular.jl:134-205). */
int bem_host_slot_tables(int n_faces, int n_vertices, const int* faces, const int* order, int pend_slots,
                         int* slot_count, uint16_t* slot_start, uint16_t* slot_inc, int* slot_vertex, int* slot_pend,
                         int* slot_prev, int* elem_pos, int* chunk_start, int* vertex_last, int* vstar_ptr,
                         int* vstar_row);
/* n float32 values as a JSON array (HOST arrays) in `out` (cap >= 16 n + 2 bytes, no terminator), each with nine
 * significant digits, so that every value reads back as the same float32. Returns the length written, -1 for bad
 * arguments or too small a buffer, -2 if a value is not finite (the caller writes those rows itself). The worker's
 * result rows; replaces the JSON writer of the reference's worker for its Float32 arrays (emit_event,
 * julia_local/solver.jl:12-17). */
long long bem_host_format_f32_json(const float* v, long long n, char* out, long long cap);

/* Upload a prepared mesh: SoA geometry, locality/slot tables, singular pairs and Duffy rules, field sources.
 * Replaces build_rocm_regular_assembly_cache (JS/BeatEngineRocmRegular.jl:134-205),
 * build_rocm_singular_correction_cache (JS/BeatEngineRocmSingular.jl:1-55),
 * build_rocm_field_evaluation_cache (JS/BeatEngineRocmField.jl:1-39) and the identity cache
 * (JS/BeatEngineRocmOperators.jl:23-32). Synchronous. */
int bem_mesh_create(bem_ctx* ctx, const bem_mesh_desc* desc, bem_mesh** out);
/* Release the mesh. While captured graphs still use it (see bem_graph_end) the release is deferred until the
 * last of them is destroyed. */
int bem_mesh_destroy(bem_mesh* mesh);
/* Number of live captured graphs whose nodes use this mesh's tables or workspace. While it is non-zero the mesh's
 * workspace cannot grow: bem_mesh_reserve, bem_assemble or bem_field asking for more returns BEM_E_STATE instead
 * of moving buffers the graphs address. */
int bem_mesh_live_graphs(bem_mesh* mesh, int* count);
/* Reserve library workspace for up to `nf` frequencies per call at `precision` (device allocations happen
 * here, never inside a captured graph). `strip_nf` bounds how many frequencies share one strip buffer; when nf
 * exceeds it and a strip set stays within 4 GiB (counted with one column per vertex), two sets are reserved and
 * bem_assemble gathers one frequency group on an auxiliary stream while the next group's strips are assembled
 * (not when the context's fan-out is 1 or its timers are on). */
int bem_mesh_reserve(bem_mesh* mesh, int precision, int nf, int strip_nf, int n_obs, int n_ch);
/* Bytes of device memory held by the mesh (geometry + workspace). */
int bem_mesh_device_bytes(bem_mesh* mesh, long long* bytes);

/* Assemble, for nf wavenumbers k[f] (device FP64 array), the Burton-Miller LHS
 *   A_f = 1/2 M11 - D + (i/k) H            (cT, [nf] x N x lda, stride strideA; optional copy A0, same layout)
 * and channel RHS
 *   b_f[c] = rhs_scale[f] * (-(S + (i/k)(D* + 1/2 M10)) qunit[c])   (c128, [nf][n_ch][ldb], stride strideB)
 * with qunit a device c128 [n_ch][F] array of unit channel drives and rhs_scale a device c128 [nf] array
 * (i rho omega per frequency). rowmax receives max_j |A_ij| (|re|+|im|) as FP64 [nf][N].
 * Replaces assemble_regular_galerkin_operators_rocm_regular (JS/BeatEngineRocmAssembly.jl:16-193),
 * _rocm_burton_miller_rhs (JS/BeatEngineRocmOperators.jl:110-121) and the LHS broadcast (:239-240). */
int bem_assemble(bem_ctx* ctx, bem_mesh* mesh, int precision, int nf, const double* k, const void* qunit,
                 int n_ch, const void* rhs_scale, void* A, long long lda, long long strideA, void* A0, void* b,
                 long long ldb, long long strideB, double* rowmax);

/* Right-hand sides only, for a further group of up to BEM_MAX_CHANNELS channels of the same frequencies: b as in
 * bem_assemble for these unit drives, A untouched. Reuses the singular RHS blocks of the last bem_assemble on this
 * mesh, which must have been called with the same device wavenumber array k (BEM_E_STATE otherwise). The
 * reference re-forms and re-factors the system per channel (julia_local/solver.jl:889-909 calling
 * JS/BeatEngineRocmOperators.jl:224-251); here the factors of one bem_assemble serve every channel group and only
 * _rocm_burton_miller_rhs (JS/BeatEngineRocmOperators.jl:110-121) runs per group. */
int bem_assemble_rhs(bem_ctx* ctx, bem_mesh* mesh, int precision, int nf, const double* k, const void* qunit,
                     int n_ch, const void* rhs_scale, void* b, long long ldb, long long strideB);

/* Neumann data q[f][c][j] = rhs_scale[f] * qunit[c][j] (c128; rhs_scale = i rho omega per frequency, qunit the
 * unit channel drives [n_ch][n_faces]) on the device. Replaces the q = i rho omega v_n of pressure_for_drives
 * (julia_local/solver.jl:434-470, :456). */
int bem_neumann(bem_ctx* ctx, int nf, int n_ch, int n_faces, const void* rhs_scale, const void* qunit, void* q);

/* ---- Coupled FEM-BEM system blocks (the coupled matrix is row-major cT; callers offset `out` to a block) ----
 *
 * A column set: sparse Neumann columns on the mesh's faces (HOST arrays, validated and uploaded once). Column
 * c has entries e in [col_ptr[c], col_ptr[c+1]) with face src[col_idx[e]] and weight col_w[e] (c128). The
 * interface block uses the interface flux basis (orientation / 3 on the faces around each interface vertex,
 * assemble_interface_operators, JS/BeatEngineCoupled.jl:721-773), the motion block the piston normal
 * velocities (assemble_transducer_operators, :775-914). Destroy it before its mesh. */
int bem_colset_create(bem_ctx* ctx, bem_mesh* mesh, int n_src, const int* src, int n_cols, const int* col_ptr,
                      const int* col_idx, const double* col_w, bem_colset** out);
int bem_colset_destroy(bem_colset* cs);
/* out[f][i][c] = alpha[f] * (B W)[i][c] for the N = n_vertices P1 rows and the set's columns, where B is the
 * Burton-Miller RHS operator -S - (i/k)(D* + 1/2 M10) at wavenumber k[f] (device double [nf]; the same pointer
 * and frequencies as the last bem_assemble on this mesh, whose singular blocks it reuses) and W the column set
 * (alpha: device c128 [nf]; out: device cT, row-major with ldo >= n_cols, strideO between frequencies).
 * Replaces bem_rhs_operator * bem_flux / * bem_motion_flux of _accelerator_coupled_bem_blocks
 * (JS/BeatEngineCoupled.jl:1264-1406) without forming the operator. */
int bem_rhs_columns(bem_ctx* ctx, bem_mesh* mesh, const bem_colset* cs, int precision, int nf, const double* k,
                    const void* alpha, void* out, long long ldo, long long strideO);
/* A sparse block (HOST COO arrays of n_rows x n_cols entries, c128 values; entries must be unique), uploaded once.
 * Replaces build_rocm_sparse_scatter_cache (JS/BeatEngineRocmOperators.jl:40-54). */
int bem_sparse_create(bem_ctx* ctx, int n_rows, int n_cols, long long nnz, const int* rows, const int* cols,
                      const double* vals, bem_sparse** out);
int bem_sparse_destroy(bem_sparse* sp);
/* out[rows[e] * ldo + cols[e]] = alpha * vals[e] (add != 0: +=) into the block at `out` (device cT, at least
 * n_rows x n_cols with leading dimension ldo). Replaces scatter_rocm_sparse_to_dense!
 * (JS/BeatEngineRocmOperators.jl:56-100). */
int bem_sparse_scatter(bem_ctx* ctx, const bem_sparse* sp, int precision, double alpha_re, double alpha_im, int add,
                       void* out, long long ldo);
/* out[i * ldo + j] = alpha * src[i * lds + j] (add != 0: +=) for i < m, j < n; src (device c128) may be NULL,
 * which writes alpha (a fill). Places the dense Schur complement and the diagonal transducer blocks, and clears
 * the coupled matrix, on the library's stream (no memset nodes). */
int bem_dense_block(bem_ctx* ctx, int precision, int m, int n, const void* src, long long lds, double alpha_re,
                    double alpha_im, int add, void* out, long long ldo);
/* Scale rows i < n of A (ncols columns, device cT, leading dimension lda) by powers of two so that each row's
 * largest |re| + |im| lies in [1/2, 1) -- exact, no entry is rounded; scale (device double [n]) receives the
 * factors (1 for a zero row). The solution of the scaled system with right-hand-side rows multiplied by the same
 * factors is unchanged. No reference counterpart: the reference's single-precision dense LU runs unscaled;
 * here it lets the FP32 factors of the coupled matrix, whose FEM, BEM, trace and lumped rows differ in scale by
 * orders of magnitude, precondition the FP64 refinement in a few iterations. */
int bem_row_equilibrate(bem_ctx* ctx, int precision, int n, int ncols, void* A, long long lda, double* scale);

/* Bytes of the LU workspace for `batch` N x N matrices: the inverted 64 x 64 diagonal blocks of L and U
 * that bem_lu_factor writes and bem_lu_solve / bem_refine read (cT [batch][ceil(N/64)][2][64][64]). */
int bem_lu_workspace_bytes(int precision, int n, int batch, long long* bytes);

/* In-place LU with partial pivoting of `batch` N x N cT matrices (P A = L U, LAPACK getrf semantics;
 * ipiv[c] = row exchanged with row c, perm = resulting row permutation; info[m] = 0 or first zero pivot + 1;
 * int arrays of stride strideP). `work` (bem_lu_workspace_bytes) receives the inverted diagonal blocks and
 * travels with the factors. N <= BEM_LU_MAX_N. Replaces rocm_dense_lu! (JS/BeatEngineRocmOperators.jl:161-194). */
int bem_lu_factor(bem_ctx* ctx, int precision, int n, int batch, void* A, long long lda, long long strideA,
                  int* ipiv, int* perm, long long strideP, int* info, void* work);

/* y = U^-1 L^-1 P b for nrhs right-hand sides (b c128 [batch][nrhs][ldb]; y cT [batch][nrhs][ldy]).
 * `work` is the workspace bem_lu_factor filled for these factors. `skip` (device int [batch], may be
 * NULL) skips matrices whose flag is set.
 * Replaces solve_rocm_dense_factorization (JS/BeatEngineRocmOperators.jl:196-222). */
int bem_lu_solve(bem_ctx* ctx, int precision, int n, int batch, const void* LU, long long lda, long long strideA,
                 const int* perm, long long strideP, const void* work, const void* b, long long ldb,
                 long long strideB, void* y, long long ldy, long long strideY, int nrhs, const int* skip);

/* Mixed-precision solve with FP64 iterative refinement (no reference counterpart; the reference solves
 * in FP32 once, JS/BeatEngineRocmOperators.jl:132-159): x (c128) = A^-1 b where A0 is the c64 matrix as
 * assembled and LU its c64 factorization. Iterates r = b - A0 x (FP64), d = LU^-1 r, x += d until
 * ||r_c|| / ||b_c|| <= tol for every channel or max_iter corrections. rel[m] receives the final relative
 * residual, iters[m] the corrections applied. lu_work: the factorization's workspace (bem_lu_factor);
 * work_r: c128 like b; work_y: c64 [batch][nrhs][ldb];
 * done: device int [batch] scratch. Works under graph capture (fixed launch count, device-side exit). */
int bem_refine(bem_ctx* ctx, int n, int batch, const void* LU, long long lda, long long strideA, const int* perm,
               long long strideP, const void* lu_work, const void* A0, const void* b, long long ldb, long long strideB, void* x,
               void* work_r, void* work_y, int nrhs, int max_iter, double tol, double* rel, int* iters, int* done);

/* bem_refine without workspace arguments, for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). No counterpart in the reference. The residuals,
 * corrections and done flags live in a static per-device scratch of BEM_NOALLOC_SCRATCH_BYTES (64 MiB, held by
 * each device the library runs on until the process ends): batch * strideB * 24 bytes (+ 4 bytes per matrix,
 * 256-byte aligned parts) must fit, else BEM_E_ARG. Calls no allocating routine. Every call on a device shares
 * that device's scratch, so calls that overlap on one device (other streams, other threads) race; the scratch
 * keeps the last call's contents and each call initialises what it reads. Same results as bem_refine. */
#define BEM_NOALLOC_SCRATCH_BYTES (64LL << 20)
int bem_refine_noalloc(bem_ctx* ctx, int n, int batch, const void* LU, long long lda, long long strideA,
                       const int* perm, long long strideP, const void* lu_work, const void* A0, const void* b,
                       long long ldb, long long strideB, void* x, int nrhs, int max_iter, double tol, double* rel,
                       int* iters);

/* FP64 relative residual ||b_c - A x_c|| / ||b_c|| (max over channels) of a cT matrix and c128 x, b.
 * rel: device double [batch]. work_r: c128 like b. scratch: device int [2*batch]. */
int bem_residual(bem_ctx* ctx, int precision, int n, int batch, const void* A, long long lda, long long strideA,
                 const void* b, const void* x, long long ldb, long long strideB, void* work_r, int nrhs,
                 double* rel, int* scratch);

/* Pivot growth max|U| / max|A| in FP64 (|re|+|im|); amax = rowmax from bem_assemble, umax scratch
 * double [batch][N], growth double [batch]. */
int bem_growth(bem_ctx* ctx, int precision, int n, int batch, const void* LU, long long lda, long long strideA,
               const double* amax, double* umax, double* growth);

/* Outcome of a batch of solves, made visible to the host (synchronizes the context stream; not capturable).
 * flags (HOST int [batch]) receives per matrix bit 0: info[m] != 0 (LU breakdown), bit 1: a non-finite entry of
 * x (c128 [batch][nrhs][ldx], stride strideX), bit 2: rel[m] > tol or NaN, bit 3: growth[m] > growth_max or NaN
 * (rel, growth: device double [batch], may be NULL to skip that check). Returns BEM_E_SINGULAR if any matrix has
 * bit 0 or 1, else BEM_E_ACCURACY if any has bit 2 or 3, else BEM_OK. The reference's counterpart is the
 * rocSOLVER info check (_is_corrupt_rocsolver_info, JS/BeatEngineRocmOperators.jl:123-130) and lu!'s
 * SingularException (JS/BeatEngineCpuSolve.jl:63), which its worker turns into a `failed` event
 * (julia_local/solver.jl:1033). */
int bem_solve_status(bem_ctx* ctx, int n, int batch, const int* info, const void* x, long long ldx,
                     long long strideX, int nrhs, const double* rel, double tol, const double* growth,
                     double growth_max, int* flags);

/* Field at n_obs points (obs: device FP64 [n_obs][3]) for nf frequencies and n_ch channels:
 *   u = sum_m dG/dn p_w - G q_w     (p: c128 [nf][n_ch][ldp] boundary pressure; q: c128 [nf][n_ch][F]
 *   Neumann data including i rho omega; out: c128 [nf][n_ch][ldo]).
 * Replaces evaluate_galerkin_field_rocm (JS/BeatEngineRocmField.jl:150-202). */
int bem_field(bem_ctx* ctx, bem_mesh* mesh, int precision, int nf, const double* k, int n_obs, const double* obs,
              const void* p, long long ldp, long long strideP, const void* q, long long strideQ, int n_ch,
              void* out, long long ldo, long long strideO);

/* Stream capture into a HIP graph on the context's stream (frequency batching under one graph launch).
 * bem_graph_end returns an executable graph handle usable with bem_graph_launch. The handle records the meshes
 * the captured sequence used: until bem_graph_destroy, their workspace cannot grow (bem_mesh_live_graphs) and
 * bem_mesh_destroy defers their release. Replaces the per-frequency launch loop of julia_local/solver.jl:805-997. */
int bem_graph_begin(bem_ctx* ctx);
int bem_graph_end(bem_ctx* ctx, void** graph_exec);
int bem_graph_launch(bem_ctx* ctx, void* graph_exec);
int bem_graph_destroy(void* graph_exec);

#ifdef __cplusplus
}
#endif
#endif /* BEMHIP_H */
