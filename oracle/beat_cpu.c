/*
 * C/OpenMP restatement of the BEAT CPU engine's operator assembly and field evaluation.
 *
 * TEST INFRASTRUCTURE ONLY (oracle and reported CPU baseline), never linked into the product.
 *
 * Follows src/blab/solvers/julia_local/src/BeatEngineCpuAssembly.jl:
 *   - test elements are coloured so that no two elements of a colour share a vertex (:102-130); each colour
 *     group is processed by all threads, one test element per iteration, so rows never collide (:683-698);
 *   - regular pairs (no shared vertex id) accumulate 6x6 q-pairs into register blocks, then one scatter
 *     (:278-373, skip rule :391);
 *   - singular pairs walk their remapped Duffy rule and add each point directly (:151-218, :447-486);
 * and BeatEngineCpuField.jl:1-45 for the field (threads over observation points).
 * Compiled twice: REAL=float (the reference's production precision) and REAL=double (fast FP64 oracle).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef REAL
#define REAL double
#define SUF _f64
#endif

#ifdef BEAT_CPU_SINGLE
#define SQRT sqrtf
#define COS cosf
#define SIN sinf
#else
#define SQRT sqrt
#define COS cos
#define SIN sin
#endif

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SUF)

static const double RXI[6] = {0.4459484909159651, 0.0915762135097710, 0.1081030181680700,
                              0.4459484909159651, 0.8168475729804590, 0.0915762135097710};
static const double RETA[6] = {0.4459484909159651, 0.0915762135097700, 0.4459484909159651,
                               0.1081030181680700, 0.0915762135097700, 0.8168475729804580};
static const double RW[6] = {0.5 * 0.2233815896780110, 0.5 * 0.1099517436553220, 0.5 * 0.2233815896780110,
                             0.5 * 0.2233815896780110, 0.5 * 0.1099517436553220, 0.5 * 0.1099517436553220};

typedef struct {
  REAL re, im;
} cr;

/* Dense column-major complex matrices as in the reference: A[row + col * ld]. */
static inline void cadd(cr* a, REAL re, REAL im) {
  a->re += re;
  a->im += im;
}

static inline int adjacent(const int* fa, const int* fb) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      if (fa[i] == fb[j]) return 1;
  return 0;
}

/*
 * fv[F*9] vertex coordinates, nrm[F*3], area[F], faces[F*3] (0-based), curls[F*9].
 * color_ptr[n_colors+1], color_elem[F]: colour groups of test elements.
 * Singular pairs grouped by test element: sp_ptr[F+1], sp_trial[P], sp_rule[P]; rules: rule_off, rule_len,
 * rule_pts[n][5] (test xi, eta, trial xi, eta, w).
 * Outputs (zero-initialised by the caller): S[V x F], D[V x V], Ds[V x F], H[V x V] column-major complex.
 */
int FN(beat_cpu_assemble)(int F, int V, const REAL* fv, const REAL* nrm, const REAL* area, const int* faces,
                          const REAL* curls, REAL k, int n_colors, const int* color_ptr, const int* color_elem,
                          const int* sp_ptr, const int* sp_trial, const int* sp_rule, const int* rule_off,
                          const int* rule_len, const REAL* rule_pts, int do_regular, int do_singular, cr* S, cr* D,
                          cr* Ds, cr* H) {
  const REAL inv4pi = (REAL)(1.0 / (4.0 * M_PI));
  const REAL k2 = k * k;
  /* quadrature points of every element */
  REAL* qp = (REAL*)malloc(sizeof(REAL) * (size_t)F * 18);
  if (!qp) return 1;
  for (int e = 0; e < F; ++e)
    for (int q = 0; q < 6; ++q) {
      const REAL b0 = (REAL)(1.0 - RXI[q] - RETA[q]), b1 = (REAL)RXI[q], b2 = (REAL)RETA[q];
      for (int c = 0; c < 3; ++c)
        qp[e * 18 + 3 * q + c] = b0 * fv[e * 9 + c] + b1 * fv[e * 9 + 3 + c] + b2 * fv[e * 9 + 6 + c];
    }
  for (int g = 0; g < n_colors; ++g) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int gi = color_ptr[g]; gi < color_ptr[g + 1]; ++gi) {
      const int t = color_elem[gi];
      const int* ft = faces + 3 * t;
      const REAL* nt = nrm + 3 * t;
      const REAL* ct = curls + 9 * t;
      if (do_regular) {
        for (int s = 0; s < F; ++s) {
          const int* fs = faces + 3 * s;
          if (adjacent(ft, fs)) continue;
          const REAL* ns = nrm + 3 * s;
          const REAL* cs = curls + 9 * s;
          const REAL nn = nt[0] * ns[0] + nt[1] * ns[1] + nt[2] * ns[2];
          const REAL jac = (REAL)4.0 * area[t] * area[s];
          REAL sb[3][2] = {{0}}, ab[3][2] = {{0}}, db[3][3][2] = {{{0}}}, hb[3][3][2] = {{{0}}};
          REAL cp[3][3];
          for (int l = 0; l < 3; ++l)
            for (int m = 0; m < 3; ++m)
              cp[l][m] = ct[3 * l] * cs[3 * m] + ct[3 * l + 1] * cs[3 * m + 1] + ct[3 * l + 2] * cs[3 * m + 2];
          for (int a = 0; a < 6; ++a) {
            const REAL* x = qp + t * 18 + 3 * a;
            const REAL tb[3] = {(REAL)(1.0 - RXI[a] - RETA[a]), (REAL)RXI[a], (REAL)RETA[a]};
            for (int b = 0; b < 6; ++b) {
              const REAL* y = qp + s * 18 + 3 * b;
              const REAL rb[3] = {(REAL)(1.0 - RXI[b] - RETA[b]), (REAL)RXI[b], (REAL)RETA[b]};
              const REAL dx = y[0] - x[0], dy = y[1] - x[1], dz = y[2] - x[2];
              const REAL r = SQRT(dx * dx + dy * dy + dz * dz);
              if (r == (REAL)0) continue;
              const REAL ir = (REAL)1 / r;
              const REAL gsc = inv4pi * ir;
              const REAL gr = COS(k * r) * gsc, gi = SIN(k * r) * gsc;
              /* grad_scale = G * (-1/r + i k) */
              const REAL hr = -gr * ir - gi * k, hi = gr * k - gi * ir;
              const REAL trial_dot = (dx * ns[0] + dy * ns[1] + dz * ns[2]) * ir;
              const REAL test_dot = -(dx * nt[0] + dy * nt[1] + dz * nt[2]) * ir;
              const REAL w = (REAL)RW[a] * (REAL)RW[b] * jac;
              const REAL wgr = gr * w, wgi = gi * w;
              const REAL wdr = hr * trial_dot * w, wdi = hi * trial_dot * w;
              const REAL war = hr * test_dot * w, wai = hi * test_dot * w;
              for (int l = 0; l < 3; ++l) {
                sb[l][0] += tb[l] * wgr;
                sb[l][1] += tb[l] * wgi;
                ab[l][0] += tb[l] * war;
                ab[l][1] += tb[l] * wai;
                for (int m = 0; m < 3; ++m) {
                  const REAL bp = tb[l] * rb[m];
                  db[l][m][0] += bp * wdr;
                  db[l][m][1] += bp * wdi;
                  const REAL hf = cp[l][m] - k2 * bp * nn;
                  hb[l][m][0] += hf * wgr;
                  hb[l][m][1] += hf * wgi;
                }
              }
            }
          }
          for (int l = 0; l < 3; ++l) {
            const int row = ft[l];
            cadd(&S[row + (size_t)s * V], sb[l][0], sb[l][1]);
            cadd(&Ds[row + (size_t)s * V], ab[l][0], ab[l][1]);
            for (int m = 0; m < 3; ++m) {
              const int col = fs[m];
              cadd(&D[row + (size_t)col * V], db[l][m][0], db[l][m][1]);
              cadd(&H[row + (size_t)col * V], hb[l][m][0], hb[l][m][1]);
            }
          }
        }
      }
      if (do_singular) {
        for (int p = sp_ptr[t]; p < sp_ptr[t + 1]; ++p) {
          const int s = sp_trial[p];
          const int* fs = faces + 3 * s;
          const REAL* ns = nrm + 3 * s;
          const REAL* cs = curls + 9 * s;
          const REAL nn = nt[0] * ns[0] + nt[1] * ns[1] + nt[2] * ns[2];
          const REAL jac = ((REAL)2 * area[t]) * ((REAL)2 * area[s]);
          REAL cp[3][3];
          for (int l = 0; l < 3; ++l)
            for (int m = 0; m < 3; ++m)
              cp[l][m] = ct[3 * l] * cs[3 * m] + ct[3 * l + 1] * cs[3 * m + 1] + ct[3 * l + 2] * cs[3 * m + 2];
          const int ru = sp_rule[p];
          for (int q = 0; q < rule_len[ru]; ++q) {
            const REAL* rp = rule_pts + (size_t)(rule_off[ru] + q) * 5;
            const REAL tb[3] = {(REAL)1 - rp[0] - rp[1], rp[0], rp[1]};
            const REAL rb[3] = {(REAL)1 - rp[2] - rp[3], rp[2], rp[3]};
            REAL x[3], y[3];
            for (int c = 0; c < 3; ++c) {
              x[c] = tb[0] * fv[t * 9 + c] + tb[1] * fv[t * 9 + 3 + c] + tb[2] * fv[t * 9 + 6 + c];
              y[c] = rb[0] * fv[s * 9 + c] + rb[1] * fv[s * 9 + 3 + c] + rb[2] * fv[s * 9 + 6 + c];
            }
            const REAL dx = y[0] - x[0], dy = y[1] - x[1], dz = y[2] - x[2];
            const REAL r = SQRT(dx * dx + dy * dy + dz * dz);
            if (r == (REAL)0) continue;
            const REAL ir = (REAL)1 / r;
            const REAL gsc = inv4pi * ir;
            const REAL gr = COS(k * r) * gsc, gi = SIN(k * r) * gsc;
            const REAL hr = -gr * ir - gi * k, hi = gr * k - gi * ir;
            const REAL trial_dot = (dx * ns[0] + dy * ns[1] + dz * ns[2]) * ir;
            const REAL test_dot = -(dx * nt[0] + dy * nt[1] + dz * nt[2]) * ir;
            const REAL w = rp[4] * jac;
            for (int l = 0; l < 3; ++l) {
              const int row = ft[l];
              cadd(&S[row + (size_t)s * V], tb[l] * gr * w, tb[l] * gi * w);
              cadd(&Ds[row + (size_t)s * V], tb[l] * hr * test_dot * w, tb[l] * hi * test_dot * w);
              for (int m = 0; m < 3; ++m) {
                const int col = fs[m];
                const REAL bp = tb[l] * rb[m];
                cadd(&D[row + (size_t)col * V], bp * hr * trial_dot * w, bp * hi * trial_dot * w);
                const REAL hf = cp[l][m] - k2 * bp * nn;
                cadd(&H[row + (size_t)col * V], hf * gr * w, hf * gi * w);
              }
            }
          }
        }
      }
    }
  }
  free(qp);
  return 0;
}

/* u(x_o) = sum_m dG/dn p_w - G q_w over all rule points of all faces (threads over observation points).
 * p[V] and q[F] complex (one channel), out[n_obs]. */
int FN(beat_cpu_field)(int F, const REAL* fv, const REAL* nrm, const REAL* area, const int* faces, REAL k,
                       const cr* p, const cr* q, int n_obs, const REAL* obs, cr* out) {
  const REAL inv4pi = (REAL)(1.0 / (4.0 * M_PI));
#pragma omp parallel for schedule(static)
  for (int o = 0; o < n_obs; ++o) {
    const REAL* xo = obs + 3 * o;
    REAL ur = 0, ui = 0;
    for (int e = 0; e < F; ++e) {
      const int* f = faces + 3 * e;
      const REAL* n = nrm + 3 * e;
      for (int a = 0; a < 6; ++a) {
        const REAL b[3] = {(REAL)(1.0 - RXI[a] - RETA[a]), (REAL)RXI[a], (REAL)RETA[a]};
        REAL y[3];
        for (int c = 0; c < 3; ++c) y[c] = b[0] * fv[e * 9 + c] + b[1] * fv[e * 9 + 3 + c] + b[2] * fv[e * 9 + 6 + c];
        const REAL dx = y[0] - xo[0], dy = y[1] - xo[1], dz = y[2] - xo[2];
        const REAL r = SQRT(dx * dx + dy * dy + dz * dz);
        if (r == (REAL)0) continue;
        const REAL w = (REAL)RW[a] * (REAL)2 * area[e];
        const REAL gsc = inv4pi / r;
        const REAL gr = COS(k * r) * gsc, gi = SIN(k * r) * gsc;
        const REAL proj = (dx * n[0] + dy * n[1] + dz * n[2]) / r;
        const REAL dr = (-gr / r - gi * k) * proj, di = (gr * k - gi / r) * proj;
        const REAL pwr = (b[0] * p[f[0]].re + b[1] * p[f[1]].re + b[2] * p[f[2]].re) * w;
        const REAL pwi = (b[0] * p[f[0]].im + b[1] * p[f[1]].im + b[2] * p[f[2]].im) * w;
        const REAL qwr = q[e].re * w, qwi = q[e].im * w;
        ur += dr * pwr - di * pwi - (gr * qwr - gi * qwi);
        ui += dr * pwi + di * pwr - (gr * qwi + gi * qwr);
      }
    }
    out[o].re = ur;
    out[o].im = ui;
  }
  return 0;
}
