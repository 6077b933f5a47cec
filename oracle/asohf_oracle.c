/*
 * asohf_oracle.c -- CPU oracle for ASOHF's particle-to-mesh density build.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  The product
 * (libasohf_hip) never links, loads or calls it.
 *
 * It restates, in plain C, the reference algorithm of
 *   - COMPUTE_KERNELS          grids.f:997-1067   (orc_kernels32 / orc_kernels64)
 *   - MALLA                    grids.f:2509-2573  (orc_malla)
 *   - DENSITY  ("mode A")      grids.f:710-994    (orc_density_a32 / _a64)
 *   - INTERPOLATE_DENSITY_TSC  ("mode B") grids.f:1070-1398 (orc_density_b32 / _b64)
 *   - RENORM_DENSITY           grids.f:2668-2722  (the final "/ fdm" of every cell)
 * plus an FP64 adjoint of mode B (orc_adjoint_b64), which has no reference
 * counterpart (the reference has no derivatives), and two routines beside
 * the path:
 *   - SORT_DM_PARTICLES_LOCALDENSITY reader.f:140-336 (orc_species32)
 *   - the peak candidates of HALOFIND_GRID, haloes_grids.f:316-404 (base)
 *     and 700-790 (AMR levels), sorted by SORT_CELLS / SORT_CELLS_AMR
 *     (haloes_grids.f:1341-1425) with ARGSORT (num.f:150-200)
 *     (orc_argsort32, orc_peaks32);
 *   - the grid masks written with the density dump (asohf.f:515-545):
 *     COMPUTE_CR0AMR grids.f:2076-2226 (orc_cr0amr) and VEINSGRID
 *     grids.f:2229-2506 (orc_veinsgrid).
 *
 * Two arithmetic flavours:
 *   *32  -- the reference's own FP32 expressions, operand order and loop
 *           order.  With nthreads == 1 the result is bit-identical to the
 *           reference compiled without OpenMP (checked against oracle/_ref).
 *           With nthreads > 1 it is the CPU baseline: the base grid is reduced
 *           from per-thread private copies as OpenMP REDUCTION(+:U1) does
 *           (grids.f:787-791) and patches are scheduled dynamically
 *           (grids.f:857).
 *   *64  -- the same control flow and the same FP32 cell/patch index
 *           expressions (indices are integers and must match bit for bit),
 *           but weights, products, sums and the normalisation in FP64,
 *           evaluated on the reference's FP32 operands (positions, masses,
 *           patch origins and the FP32 cell centres RADX / MINIRADX).
 *
 * Build: gcc -O2 -ffp-contract=off -fopenmp -shared -fPIC (oracle/Makefile).
 * -ffp-contract=off is required: the reference's FP32 expressions have no
 * fused multiply-adds.
 *
 * Layouts: every grid is column-major, x fastest, 1-based in the comments,
 * 0-based in memory.  Patch fields are packed one after another in patch
 * order 1..P (no NAMRX padding).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_N_ESP 9    /* asohf_parameters.dat: N_ESP   */
#define ORC_NLEVELS 9  /* asohf_parameters.dat: NLEVELS */

typedef struct {
  int32_t nx, ny, nz;     /* base cells per axis                          */
  int32_t nl;             /* number of AMR levels                         */
  const int32_t *npatch;  /* [nl+1]  NPATCH(0:NL)                          */
  const int32_t *pnx, *pny, *pnz;  /* [P] PATCHNX/NY/NZ                     */
  const float *prx, *pry, *prz;    /* [P] PATCHRX/RY/RZ                     */
  float dx, dy, dz;       /* COMMON /ESPACIADO/                            */
  float lado0;            /* box side (internal length units)              */
  float rote, rete;       /* COMMON /BACK/                                 */
  float fdm;              /* RENORM_DENSITY divisor                        */
  int32_t interp_degree;  /* COMPUTE_KERNELS: 1 (CIC-like) or 2 (TSC-like) */
  const float *radx, *rady, *radz; /* [n+2] cell centres incl. ghosts 0, n+1 */
} orc_geom;

/* ------------------------------------------------------------------ */
/* helpers                                                             */
/* ------------------------------------------------------------------ */

/* Fortran REAL**INTEGER for small positive exponents: binary powering as
 * compiler runtimes do it (x**3 == x*(x*x)). */
static float powi_f(float a, int b) {
  float r = 1.0f;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return r;
}

static int geom_npatch_total(const orc_geom *g) {
  int s = 0;
  for (int l = 0; l <= g->nl; ++l) s += g->npatch[l];
  return s;
}

/* first global patch index (0-based) of level ir, and one past the last:
 * LOW1 = SUM(NPATCH(0:IR-1))+1, LOW2 = SUM(NPATCH(0:IR)) (grids.f:846-847) */
static void level_range(const orc_geom *g, int ir, int *lo, int *hi) {
  int s = 0;
  for (int l = 0; l < ir; ++l) s += g->npatch[l];
  *lo = s;
  *hi = s + g->npatch[ir];
}

/* packed offsets of every patch field, in cells */
static int64_t *patch_offsets(const orc_geom *g) {
  int P = geom_npatch_total(g);
  int64_t *off = (int64_t *)malloc(sizeof(int64_t) * (P + 1));
  off[0] = 0;
  for (int p = 0; p < P; ++p)
    off[p + 1] = off[p] + (int64_t)g->pnx[p] * g->pny[p] * g->pnz[p];
  return off;
}

/* ------------------------------------------------------------------ */
/* MALLA  grids.f:2509-2573                                             */
/* ------------------------------------------------------------------ */
/* rad[0..n+1]; returns the cell size RAD(2)-RAD(1).  lado is the MALLA
 * argument, i.e. LADO0 - LADO0/NX (asohf.f:323). */
float orc_malla(int n, float lado, float *rad) {
  float a = -lado / 2.0f;
  float b = lado / 2.0f;
  float c = (b - a) / (float)(n - 1);
  rad[1] = a;
  for (int i = 2; i <= n; ++i) rad[i] = rad[1] + (float)(i - 1) * c;
  rad[0] = rad[1] - c;
  rad[n + 1] = rad[n] + c;
  return rad[2] - rad[1];
}

/* ------------------------------------------------------------------ */
/* COMPUTE_KERNELS  grids.f:997-1067                                     */
/* ------------------------------------------------------------------ */
/* Table row ir holds KERNELS(-KMAX:KMAX, ir) at K[ir*(2*KMAX+1) + i+KMAX]
 * with KMAX = 2^nl; only |i| <= 2^ir is filled, the rest is 0. */
void orc_kernels32(int nl, int degree, float *K) {
  const int kmax = 1 << nl;
  const int w = 2 * kmax + 1;
  memset(K, 0, sizeof(float) * (size_t)w * (nl + 1));
  for (int ir = 0; ir <= nl; ++ir) {
    float *row = K + (size_t)ir * w + kmax;
    int n = 1 << ir;
    float h32 = (float)n + 0.5f;
    float h = 2.0f * h32 / 3.0f;
    for (int i = -n; i <= n; ++i) {
      if (degree == 2) {
        float bas = (float)abs(i) / h;
        if (bas <= 0.5f)
          row[i] = 0.75f - bas * bas;
        else {
          float t = 1.5f - bas;
          row[i] = 0.5f * (t * t);
        }
      } else if (degree == 1) {
        float bas = (float)abs(i) / h32;
        row[i] = 1.0f - bas;
      }
    }
    /* normalisation: sequential triple sum, I outer, K inner (grids.f:1042-1051) */
    float s = 0.0f;
    for (int i = -n; i <= n; ++i)
      for (int j = -n; j <= n; ++j)
        for (int k = -n; k <= n; ++k) s = s + row[i] * row[j] * row[k];
    s = 1.0f / powf(s, 1.0f / 3.0f);
    for (int i = -n; i <= n; ++i) row[i] = s * row[i];
  }
}

/* exact normalised table: k_i / sum_j k_j, everything in FP64 */
void orc_kernels64(int nl, int degree, double *K) {
  const int kmax = 1 << nl;
  const int w = 2 * kmax + 1;
  memset(K, 0, sizeof(double) * (size_t)w * (nl + 1));
  for (int ir = 0; ir <= nl; ++ir) {
    double *row = K + (size_t)ir * w + kmax;
    int n = 1 << ir;
    double h32 = (double)n + 0.5;
    double h = 2.0 * h32 / 3.0;
    double s = 0.0;
    for (int i = -n; i <= n; ++i) {
      double v = 0.0;
      if (degree == 2) {
        double bas = fabs((double)i) / h;
        v = bas <= 0.5 ? 0.75 - bas * bas : 0.5 * (1.5 - bas) * (1.5 - bas);
      } else if (degree == 1) {
        v = 1.0 - fabs((double)i) / h32;
      }
      row[i] = v;
      s += v;
    }
    for (int i = -n; i <= n; ++i) row[i] /= s;
  }
}

/* ------------------------------------------------------------------ */
/* shared FP32 index expressions (bit-exact with the reference)         */
/* ------------------------------------------------------------------ */

/* mode A base cell: IX = INT((XP-XL)/DX)+1, XL = -LADO0/2 (grids.f:780,798) */
static inline int a_base_index(float xp, float xl, float dx) {
  return (int)((xp - xl) / dx) + 1;
}
/* single periodic wrap of grids.f:808-813 */
static inline int a_wrap(int i, int n) {
  if (i > n) i -= n;
  if (i < 1) i += n;
  return i;
}

/* mode A patch box (grids.f:871-886) */
typedef struct { float xlfix, ylfix, zlfix, xl, xr, yl, yr, zl, zr; } a_box;
static inline a_box a_patch_box(const orc_geom *g, int p, float dxpa, float dypa, float dzpa) {
  a_box b;
  b.xlfix = g->prx[p] - dxpa;
  b.ylfix = g->pry[p] - dypa;
  b.zlfix = g->prz[p] - dzpa;
  b.xl = b.xlfix; b.yl = b.ylfix; b.zl = b.zlfix;
  b.xr = b.xl + (float)g->pnx[p] * dxpa;
  b.yr = b.yl + (float)g->pny[p] * dypa;
  b.zr = b.zl + (float)g->pnz[p] * dzpa;
  b.xl = b.xl - g->dx; b.xr = b.xr + g->dx;   /* MAXKERNELEXT = DX (grids.f:837) */
  b.yl = b.yl - g->dx; b.yr = b.yr + g->dx;
  b.zl = b.zl - g->dx; b.zr = b.zr + g->dx;
  return b;
}
static inline int a_in_box(const a_box *b, float xp, float yp, float zp) {
  return b->xl < xp && xp < b->xr && b->yl < yp && yp < b->yr && b->zl < zp && zp < b->zr;
}
/* IX = FLOOR((XP-XLFIX)/DXPA)+1 (grids.f:912-914) */
static inline int a_patch_index(float xp, float xlfix, float dxpa) {
  return (int)floorf((xp - xlfix) / dxpa) + 1;
}

/* mode B base index: INT((XP-XL)/DX)+1 clamped to [1,N] (grids.f:1180-1192) */
static inline int b_base_index(float xp, float xl, float dx, int n) {
  int i = (int)((xp - xl) / dx) + 1;
  if (i < 1) i = 1;
  if (i > n) i = n;
  return i;
}
/* TSC weights of grids.f:1194-1213; r = centres RAD(I-1), RAD(I), RAD(I+1) */
static inline void b_weights32(float xp, const float r[3], float d, float v[3]) {
  float bas, t;
  bas = fabsf(xp - r[0]) / d; t = 1.5f - bas; v[0] = 0.5f * (t * t);
  bas = fabsf(xp - r[1]) / d; v[1] = 0.75f - bas * bas;
  bas = fabsf(xp - r[2]) / d; t = 1.5f - bas; v[2] = 0.5f * (t * t);
}
static inline void b_weights64(float xp, const float r[3], float d, double v[3]) {
  double bas;
  bas = fabs((double)xp - (double)r[0]) / (double)d; v[0] = 0.5 * (1.5 - bas) * (1.5 - bas);
  bas = fabs((double)xp - (double)r[1]) / (double)d; v[1] = 0.75 - bas * bas;
  bas = fabs((double)xp - (double)r[2]) / (double)d; v[2] = 0.5 * (1.5 - bas) * (1.5 - bas);
}
/* derivative of the three weights w.r.t. xp (FP64) */
static inline void b_dweights64(float xp, const float r[3], float d, double dv[3]) {
  double dd = (double)d, t;
  t = (double)xp - (double)r[0];
  dv[0] = -(1.5 - fabs(t) / dd) * (t > 0 ? 1.0 : (t < 0 ? -1.0 : 0.0)) / dd;
  t = (double)xp - (double)r[1];
  dv[1] = -2.0 * t / (dd * dd);
  t = (double)xp - (double)r[2];
  dv[2] = -(1.5 - fabs(t) / dd) * (t > 0 ? 1.0 : (t < 0 ? -1.0 : 0.0)) / dd;
}

/* mode B patch box, BUF = 1 (grids.f:1294-1302) */
typedef struct { float xl, yl, zl, xr, yr, zr; } b_box;
static inline b_box b_patch_box(const orc_geom *g, int p, float dxpa, float dypa, float dzpa) {
  b_box b;
  b.xl = g->prx[p] - 2.0f * dxpa;
  b.yl = g->pry[p] - 2.0f * dypa;
  b.zl = g->prz[p] - 2.0f * dzpa;
  b.xr = b.xl + (float)(g->pnx[p] + 2) * dxpa;
  b.yr = b.yl + (float)(g->pny[p] + 2) * dypa;
  b.zr = b.zl + (float)(g->pnz[p] + 2) * dzpa;
  return b;
}
static inline int b_in_box(const b_box *b, float xp, float yp, float zp) {
  return b->xl < xp && xp < b->xr && b->yl < yp && yp < b->yr && b->zl < zp && zp < b->zr;
}
/* MINIRADX(II) = PATCHRX + (FLOAT(II)-1.5)*DXPA (grids.f:1304-1312) */
static inline float b_minirad(float pr, int ii, float dxpa) {
  return pr + ((float)ii - 1.5f) * dxpa;
}

static inline size_t cidx(int i, int j, int k, int n1, int n2) {
  return (size_t)(i - 1) + (size_t)n1 * ((size_t)(j - 1) + (size_t)n2 * (size_t)(k - 1));
}

static int nthreads_eff(int nthreads) {
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  return nthreads;
#else
  (void)nthreads;
  return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* Mode A: DENSITY, FP32 restatement                                    */
/* ------------------------------------------------------------------ */
int orc_density_a32(const orc_geom *g, int64_t n, const float *x, const float *y,
                    const float *z, const float *m, const int32_t *npart_esp,
                    float *u1, float *u11, int nthreads) {
  const int nx = g->nx, ny = g->ny, nz = g->nz, nl = g->nl;
  const int kmax = 1 << nl, kw = 2 * kmax + 1;
  float *K = (float *)malloc(sizeof(float) * kw * (nl + 1));
  orc_kernels32(nl, g->interp_degree, K);
  const size_t ncell = (size_t)nx * ny * nz;
  int nt = nthreads_eff(nthreads);
  int64_t ntot = 0;
  for (int s = 0; s < ORC_N_ESP; ++s) ntot += npart_esp[s];
  if (ntot > n) { free(K); return -1; }

  /* base grid, grids.f:764-833 */
  memset(u1, 0, sizeof(float) * ncell);
  const float *k0 = K + kmax; /* KERNELS(-1:1,0) */
  const float xl = -g->lado0 / 2.0f, yl = -g->lado0 / 2.0f, zl = -g->lado0 / 2.0f;
  if (nt == 1) {
    for (int64_t ip = 0; ip < ntot; ++ip) {
      float bas = m[ip];
      int ix = a_base_index(x[ip], xl, g->dx);
      int jy = a_base_index(y[ip], yl, g->dy);
      int kz = a_base_index(z[ip], zl, g->dz);
      for (int kk = -1; kk <= 1; ++kk)
        for (int jj = -1; jj <= 1; ++jj)
          for (int ii = -1; ii <= 1; ++ii) {
            int i1 = a_wrap(ix + ii, nx), j1 = a_wrap(jy + jj, ny), k1 = a_wrap(kz + kk, nz);
            u1[cidx(i1, j1, k1, nx, ny)] += bas * k0[ii] * k0[jj] * k0[kk];
          }
    }
  } else {
#ifdef _OPENMP
    float *priv = (float *)calloc(ncell * (size_t)nt, sizeof(float));
#pragma omp parallel num_threads(nt)
    {
      float *mine = priv + ncell * (size_t)omp_get_thread_num();
#pragma omp for schedule(static)
      for (int64_t ip = 0; ip < ntot; ++ip) {
        float bas = m[ip];
        int ix = a_base_index(x[ip], xl, g->dx);
        int jy = a_base_index(y[ip], yl, g->dy);
        int kz = a_base_index(z[ip], zl, g->dz);
        for (int kk = -1; kk <= 1; ++kk)
          for (int jj = -1; jj <= 1; ++jj)
            for (int ii = -1; ii <= 1; ++ii) {
              int i1 = a_wrap(ix + ii, nx), j1 = a_wrap(jy + jj, ny), k1 = a_wrap(kz + kk, nz);
              mine[cidx(i1, j1, k1, nx, ny)] += bas * k0[ii] * k0[jj] * k0[kk];
            }
      }
#pragma omp for schedule(static)
      for (size_t c = 0; c < ncell; ++c) {
        float s = u1[c];
        for (int t = 0; t < nt; ++t) s += priv[ncell * (size_t)t + c];
        u1[c] = s;
      }
    }
    free(priv);
#endif
  }
  float denbas = g->dx * g->dy * g->dz * g->rote * powi_f(g->rete, 3);
  for (size_t c = 0; c < ncell; ++c) u1[c] = (u1[c] / denbas) / g->fdm;

  /* patches, grids.f:839-989 */
  int64_t *off = patch_offsets(g);
  int64_t sp[ORC_N_ESP + 1];
  sp[0] = 0;
  for (int s = 0; s < ORC_N_ESP; ++s) sp[s + 1] = sp[s] + npart_esp[s];
  for (int ir = 1; ir <= nl; ++ir) {
    float p2 = powi_f(2.0f, ir);
    float dxpa = g->dx / p2, dypa = g->dy / p2, dzpa = g->dz / p2;
    float den = dxpa * dypa * dzpa * g->rote * powi_f(g->rete, 3);
    int lo, hi;
    level_range(g, ir, &lo, &hi);
#pragma omp parallel for schedule(dynamic) num_threads(nt) if (nt > 1)
    for (int p = lo; p < hi; ++p) {
      const int n1 = g->pnx[p], n2 = g->pny[p], n3 = g->pnz[p];
      float *u = u11 + off[p];
      memset(u, 0, sizeof(float) * (size_t)n1 * n2 * n3);
      a_box b = a_patch_box(g, p, dxpa, dypa, dzpa);
      for (int s = 0; s < ORC_N_ESP; ++s) {
        int irk = ir - s;
        if (irk < 0) irk = 0;
        int nk = 1 << irk;
        const float *kr = K + (size_t)irk * kw + kmax;
        for (int64_t ip = sp[s]; ip < sp[s + 1]; ++ip) {
          float xp = x[ip], yp = y[ip], zp = z[ip];
          if (!a_in_box(&b, xp, yp, zp)) continue;
          float bas = m[ip];
          int ix = a_patch_index(xp, b.xlfix, dxpa);
          int jy = a_patch_index(yp, b.ylfix, dypa);
          int kz = a_patch_index(zp, b.zlfix, dzpa);
          int i1 = ix - nk, i2 = ix + nk, j1 = jy - nk, j2 = jy + nk, k1 = kz - nk, k2 = kz + nk;
          if (i1 > n1 || i2 < 1 || j1 > n2 || j2 < 1 || k1 > n3 || j2 < 1) continue;
          if (i1 < 1) i1 = 1;
          if (i2 > n1) i2 = n1;
          if (j1 < 1) j1 = 1;
          if (j2 > n2) j2 = n2;
          if (k1 < 1) k1 = 1;
          if (k2 > n3) k2 = n3;
          for (int kk = k1; kk <= k2; ++kk)
            for (int jj = j1; jj <= j2; ++jj)
              for (int ii = i1; ii <= i2; ++ii)
                u[cidx(ii, jj, kk, n1, n2)] += bas * kr[ii - ix] * kr[jj - jy] * kr[kk - kz];
        }
      }
      size_t nc = (size_t)n1 * n2 * n3;
      for (size_t c = 0; c < nc; ++c) u[c] = (u[c] / den) / g->fdm;
    }
  }
  free(off);
  free(K);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Mode A: FP64                                                          */
/* grid_mass[0] = deposited mass on the base grid, grid_mass[1+p] = on   */
/* patch p (sum of the accumulated terms before normalisation).          */
/* ------------------------------------------------------------------ */
int orc_density_a64(const orc_geom *g, int64_t n, const float *x, const float *y,
                    const float *z, const float *m, const int32_t *npart_esp,
                    double *u1, double *u11, double *grid_mass, int nthreads) {
  const int nx = g->nx, ny = g->ny, nz = g->nz, nl = g->nl;
  const int kmax = 1 << nl, kw = 2 * kmax + 1;
  double *K = (double *)malloc(sizeof(double) * kw * (nl + 1));
  orc_kernels64(nl, g->interp_degree, K);
  const size_t ncell = (size_t)nx * ny * nz;
  int nt = nthreads_eff(nthreads);
  int64_t ntot = 0;
  for (int s = 0; s < ORC_N_ESP; ++s) ntot += npart_esp[s];
  if (ntot > n) { free(K); return -1; }
  memset(u1, 0, sizeof(double) * ncell);
  const double *k0 = K + kmax;
  const float xl = -g->lado0 / 2.0f, yl = -g->lado0 / 2.0f, zl = -g->lado0 / 2.0f;
  for (int64_t ip = 0; ip < ntot; ++ip) {
    double bas = m[ip];
    int ix = a_base_index(x[ip], xl, g->dx);
    int jy = a_base_index(y[ip], yl, g->dy);
    int kz = a_base_index(z[ip], zl, g->dz);
    for (int kk = -1; kk <= 1; ++kk)
      for (int jj = -1; jj <= 1; ++jj)
        for (int ii = -1; ii <= 1; ++ii) {
          int i1 = a_wrap(ix + ii, nx), j1 = a_wrap(jy + jj, ny), k1 = a_wrap(kz + kk, nz);
          u1[cidx(i1, j1, k1, nx, ny)] += bas * k0[ii] * k0[jj] * k0[kk];
        }
  }
  double den0 = (double)g->dx * g->dy * g->dz * g->rote * (double)g->rete * g->rete * g->rete;
  double gm = 0.0;
  for (size_t c = 0; c < ncell; ++c) {
    gm += u1[c];
    u1[c] = u1[c] / den0 / (double)g->fdm;
  }
  if (grid_mass) grid_mass[0] = gm;

  int64_t *off = patch_offsets(g);
  int64_t sp[ORC_N_ESP + 1];
  sp[0] = 0;
  for (int s = 0; s < ORC_N_ESP; ++s) sp[s + 1] = sp[s] + npart_esp[s];
  for (int ir = 1; ir <= nl; ++ir) {
    float p2 = powi_f(2.0f, ir);
    float dxpa = g->dx / p2, dypa = g->dy / p2, dzpa = g->dz / p2;
    double den = (double)dxpa * dypa * dzpa * g->rote * (double)g->rete * g->rete * g->rete;
    int lo, hi;
    level_range(g, ir, &lo, &hi);
#pragma omp parallel for schedule(dynamic) num_threads(nt) if (nt > 1)
    for (int p = lo; p < hi; ++p) {
      const int n1 = g->pnx[p], n2 = g->pny[p], n3 = g->pnz[p];
      double *u = u11 + off[p];
      memset(u, 0, sizeof(double) * (size_t)n1 * n2 * n3);
      a_box b = a_patch_box(g, p, dxpa, dypa, dzpa);
      for (int s = 0; s < ORC_N_ESP; ++s) {
        int irk = ir - s;
        if (irk < 0) irk = 0;
        int nk = 1 << irk;
        const double *kr = K + (size_t)irk * kw + kmax;
        for (int64_t ip = sp[s]; ip < sp[s + 1]; ++ip) {
          float xp = x[ip], yp = y[ip], zp = z[ip];
          if (!a_in_box(&b, xp, yp, zp)) continue;
          double bas = m[ip];
          int ix = a_patch_index(xp, b.xlfix, dxpa);
          int jy = a_patch_index(yp, b.ylfix, dypa);
          int kz = a_patch_index(zp, b.zlfix, dzpa);
          int i1 = ix - nk, i2 = ix + nk, j1 = jy - nk, j2 = jy + nk, k1 = kz - nk, k2 = kz + nk;
          if (i1 > n1 || i2 < 1 || j1 > n2 || j2 < 1 || k1 > n3 || j2 < 1) continue;
          if (i1 < 1) i1 = 1;
          if (i2 > n1) i2 = n1;
          if (j1 < 1) j1 = 1;
          if (j2 > n2) j2 = n2;
          if (k1 < 1) k1 = 1;
          if (k2 > n3) k2 = n3;
          for (int kk = k1; kk <= k2; ++kk)
            for (int jj = j1; jj <= j2; ++jj)
              for (int ii = i1; ii <= i2; ++ii)
                u[cidx(ii, jj, kk, n1, n2)] += bas * kr[ii - ix] * kr[jj - jy] * kr[kk - kz];
        }
      }
      size_t nc = (size_t)n1 * n2 * n3;
      double pm = 0.0;
      for (size_t c = 0; c < nc; ++c) {
        pm += u[c];
        u[c] = u[c] / den / (double)g->fdm;
      }
      if (grid_mass) grid_mass[1 + p] = pm;
    }
  }
  free(off);
  free(K);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Mode B: INTERPOLATE_DENSITY_TSC                                       */
/* ------------------------------------------------------------------ */
typedef struct {
  const float *radx, *rady, *radz;
  float *own;
} rad_set;

static rad_set rad_get(const orc_geom *g) {
  rad_set r = {g->radx, g->rady, g->radz, NULL};
  if (r.radx && r.rady && r.radz) return r;
  /* MALLA with LADO = LADO0 - LADO0/NX (asohf.f:323) */
  r.own = (float *)malloc(sizeof(float) * (g->nx + g->ny + g->nz + 6));
  float *ax = r.own, *ay = ax + g->nx + 2, *az = ay + g->ny + 2;
  orc_malla(g->nx, g->lado0 - g->lado0 / (float)g->nx, ax);
  orc_malla(g->ny, g->lado0 - g->lado0 / (float)g->nx, ay);
  orc_malla(g->nz, g->lado0 - g->lado0 / (float)g->nx, az);
  r.radx = ax; r.rady = ay; r.radz = az;
  return r;
}

int orc_density_b32(const orc_geom *g, int64_t n, const float *x, const float *y,
                    const float *z, const float *m, float *u1, float *u11, int nthreads) {
  const int nx = g->nx, ny = g->ny, nz = g->nz, nl = g->nl;
  const size_t ncell = (size_t)nx * ny * nz;
  int nt = nthreads_eff(nthreads);
  rad_set rs = rad_get(g);
  const float xl = -g->lado0 / 2.0f, yl = -g->lado0 / 2.0f, zl = -g->lado0 / 2.0f;
  memset(u1, 0, sizeof(float) * ncell);
#define B_BASE_BODY(ACC)                                                          \
  {                                                                               \
    float xp = x[i], yp = y[i], zp = z[i];                                        \
    int ix = b_base_index(xp, xl, g->dx, nx);                                     \
    int jy = b_base_index(yp, yl, g->dy, ny);                                     \
    int kz = b_base_index(zp, zl, g->dz, nz);                                     \
    float vx[3], vy[3], vz[3];                                                    \
    b_weights32(xp, rs.radx + ix - 1, g->dx, vx);                                 \
    b_weights32(yp, rs.rady + jy - 1, g->dy, vy);                                 \
    b_weights32(zp, rs.radz + kz - 1, g->dz, vz);                                 \
    for (int kk = -1; kk <= 1; ++kk)                                              \
      for (int jj = -1; jj <= 1; ++jj)                                            \
        for (int ii = -1; ii <= 1; ++ii) {                                        \
          int i1 = ix + ii, j1 = jy + jj, k1 = kz + kk;                           \
          if (i1 == nx + 1) i1 = 1;                                               \
          if (i1 == 0) i1 = nx;                                                   \
          if (j1 == ny + 1) j1 = 1;                                               \
          if (j1 == 0) j1 = ny;                                                   \
          if (k1 == nz + 1) k1 = 1;                                               \
          if (k1 == 0) k1 = nz;                                                   \
          ACC[cidx(i1, j1, k1, nx, ny)] += m[i] * vx[ii + 1] * vy[jj + 1] * vz[kk + 1]; \
        }                                                                         \
  }
  if (nt == 1) {
    for (int64_t i = 0; i < n; ++i) B_BASE_BODY(u1)
  } else {
#ifdef _OPENMP
    float *priv = (float *)calloc(ncell * (size_t)nt, sizeof(float));
#pragma omp parallel num_threads(nt)
    {
      float *mine = priv + ncell * (size_t)omp_get_thread_num();
#pragma omp for schedule(static)
      for (int64_t i = 0; i < n; ++i) B_BASE_BODY(mine)
#pragma omp for schedule(static)
      for (size_t c = 0; c < ncell; ++c) {
        float s = u1[c];
        for (int t = 0; t < nt; ++t) s += priv[ncell * (size_t)t + c];
        u1[c] = s;
      }
    }
    free(priv);
#endif
  }
#undef B_BASE_BODY
  float denbas = g->dx * g->dy * g->dz * g->rote * powi_f(g->rete, 3);
  for (size_t c = 0; c < ncell; ++c) u1[c] = (u1[c] / denbas) / g->fdm;

  /* PLEV (grids.f:1114-1157): the finest level whose BUF=1 box holds the
   * particle; level IR then scans only particles with PLEV >= IR, in their
   * original order (grids.f:1249-1261).  Membership in the patch's own box
   * implies PLEV >= IR, so this is a pure filter of the per-patch scan. */
  int64_t *off = patch_offsets(g);
  int *plev = (int *)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
  for (int ir = nl; ir >= 1; --ir) {
    float p2 = powi_f(2.0f, ir);
    float dxpa = g->dx / p2, dypa = g->dy / p2, dzpa = g->dz / p2;
    int lo, hi;
    level_range(g, ir, &lo, &hi);
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
    for (int64_t i = 0; i < n; ++i) {
      if (plev[i] != 0) continue;
      for (int p = lo; p < hi; ++p) {
        b_box b = b_patch_box(g, p, dxpa, dypa, dzpa);
        if (b_in_box(&b, x[i], y[i], z[i])) { plev[i] = ir; break; }
      }
    }
  }
  int64_t *lst = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  for (int ir = 1; ir <= nl; ++ir) {
    float p2 = powi_f(2.0f, ir);
    float dxpa = g->dx / p2, dypa = g->dy / p2, dzpa = g->dz / p2;
    float den = dxpa * dypa * dzpa * g->rote * powi_f(g->rete, 3);
    int lo, hi;
    level_range(g, ir, &lo, &hi);
    int64_t nbas = 0;
    for (int64_t i = 0; i < n; ++i)
      if (plev[i] >= ir) lst[nbas++] = i;
#pragma omp parallel for schedule(dynamic) num_threads(nt) if (nt > 1)
    for (int p = lo; p < hi; ++p) {
      const int n1 = g->pnx[p], n2 = g->pny[p], n3 = g->pnz[p];
      float *u = u11 + off[p];
      size_t nc = (size_t)n1 * n2 * n3;
      memset(u, 0, sizeof(float) * nc);
      b_box b = b_patch_box(g, p, dxpa, dypa, dzpa);
      for (int64_t q = 0; q < nbas; ++q) {
        const int64_t i = lst[q];
        float xp = x[i], yp = y[i], zp = z[i];
        if (!b_in_box(&b, xp, yp, zp)) continue;
        int ix = (int)((xp - b.xl) / dxpa) + 1 - 1;
        int jy = (int)((yp - b.yl) / dypa) + 1 - 1;
        int kz = (int)((zp - b.zl) / dzpa) + 1 - 1;
        float rx[3], ry[3], rz[3];
        for (int t = 0; t < 3; ++t) {
          rx[t] = b_minirad(g->prx[p], ix - 1 + t, dxpa);
          ry[t] = b_minirad(g->pry[p], jy - 1 + t, dypa);
          rz[t] = b_minirad(g->prz[p], kz - 1 + t, dzpa);
        }
        float vx[3], vy[3], vz[3];
        b_weights32(xp, rx, dxpa, vx);
        b_weights32(yp, ry, dypa, vy);
        b_weights32(zp, rz, dzpa, vz);
        for (int kk = -1; kk <= 1; ++kk)
          for (int jj = -1; jj <= 1; ++jj)
            for (int ii = -1; ii <= 1; ++ii) {
              int i1 = ix + ii, j1 = jy + jj, k1 = kz + kk;
              if (i1 > 0 && i1 <= n1 && j1 > 0 && j1 <= n2 && k1 > 0 && k1 <= n3)
                u[cidx(i1, j1, k1, n1, n2)] += m[i] * vx[ii + 1] * vy[jj + 1] * vz[kk + 1];
            }
      }
      for (size_t c = 0; c < nc; ++c) u[c] = (u[c] / den) / g->fdm;
    }
  }
  free(lst);
  free(plev);
  free(off);
  free(rs.own);
  return 0;
}

int orc_density_b64(const orc_geom *g, int64_t n, const float *x, const float *y,
                    const float *z, const float *m, double *u1, double *u11,
                    double *grid_mass, int nthreads) {
  const int nx = g->nx, ny = g->ny, nz = g->nz, nl = g->nl;
  const size_t ncell = (size_t)nx * ny * nz;
  int nt = nthreads_eff(nthreads);
  rad_set rs = rad_get(g);
  const float xl = -g->lado0 / 2.0f, yl = -g->lado0 / 2.0f, zl = -g->lado0 / 2.0f;
  memset(u1, 0, sizeof(double) * ncell);
  for (int64_t i = 0; i < n; ++i) {
    float xp = x[i], yp = y[i], zp = z[i];
    int ix = b_base_index(xp, xl, g->dx, nx);
    int jy = b_base_index(yp, yl, g->dy, ny);
    int kz = b_base_index(zp, zl, g->dz, nz);
    double vx[3], vy[3], vz[3];
    b_weights64(xp, rs.radx + ix - 1, g->dx, vx);
    b_weights64(yp, rs.rady + jy - 1, g->dy, vy);
    b_weights64(zp, rs.radz + kz - 1, g->dz, vz);
    for (int kk = -1; kk <= 1; ++kk)
      for (int jj = -1; jj <= 1; ++jj)
        for (int ii = -1; ii <= 1; ++ii) {
          int i1 = ix + ii, j1 = jy + jj, k1 = kz + kk;
          if (i1 == nx + 1) i1 = 1;
          if (i1 == 0) i1 = nx;
          if (j1 == ny + 1) j1 = 1;
          if (j1 == 0) j1 = ny;
          if (k1 == nz + 1) k1 = 1;
          if (k1 == 0) k1 = nz;
          u1[cidx(i1, j1, k1, nx, ny)] += (double)m[i] * vx[ii + 1] * vy[jj + 1] * vz[kk + 1];
        }
  }
  double den0 = (double)g->dx * g->dy * g->dz * g->rote * (double)g->rete * g->rete * g->rete;
  double gm = 0.0;
  for (size_t c = 0; c < ncell; ++c) {
    gm += u1[c];
    u1[c] = u1[c] / den0 / (double)g->fdm;
  }
  if (grid_mass) grid_mass[0] = gm;
  int64_t *off = patch_offsets(g);
  for (int ir = 1; ir <= nl; ++ir) {
    float p2 = powi_f(2.0f, ir);
    float dxpa = g->dx / p2, dypa = g->dy / p2, dzpa = g->dz / p2;
    double den = (double)dxpa * dypa * dzpa * g->rote * (double)g->rete * g->rete * g->rete;
    int lo, hi;
    level_range(g, ir, &lo, &hi);
#pragma omp parallel for schedule(dynamic) num_threads(nt) if (nt > 1)
    for (int p = lo; p < hi; ++p) {
      const int n1 = g->pnx[p], n2 = g->pny[p], n3 = g->pnz[p];
      double *u = u11 + off[p];
      size_t nc = (size_t)n1 * n2 * n3;
      memset(u, 0, sizeof(double) * nc);
      b_box b = b_patch_box(g, p, dxpa, dypa, dzpa);
      for (int64_t i = 0; i < n; ++i) {
        float xp = x[i], yp = y[i], zp = z[i];
        if (!b_in_box(&b, xp, yp, zp)) continue;
        int ix = (int)((xp - b.xl) / dxpa) + 1 - 1;
        int jy = (int)((yp - b.yl) / dypa) + 1 - 1;
        int kz = (int)((zp - b.zl) / dzpa) + 1 - 1;
        float rx[3], ry[3], rz[3];
        for (int t = 0; t < 3; ++t) {
          rx[t] = b_minirad(g->prx[p], ix - 1 + t, dxpa);
          ry[t] = b_minirad(g->pry[p], jy - 1 + t, dypa);
          rz[t] = b_minirad(g->prz[p], kz - 1 + t, dzpa);
        }
        double vx[3], vy[3], vz[3];
        b_weights64(xp, rx, dxpa, vx);
        b_weights64(yp, ry, dypa, vy);
        b_weights64(zp, rz, dzpa, vz);
        for (int kk = -1; kk <= 1; ++kk)
          for (int jj = -1; jj <= 1; ++jj)
            for (int ii = -1; ii <= 1; ++ii) {
              int i1 = ix + ii, j1 = jy + jj, k1 = kz + kk;
              if (i1 > 0 && i1 <= n1 && j1 > 0 && j1 <= n2 && k1 > 0 && k1 <= n3)
                u[cidx(i1, j1, k1, n1, n2)] += (double)m[i] * vx[ii + 1] * vy[jj + 1] * vz[kk + 1];
            }
      }
      double pm = 0.0;
      for (size_t c = 0; c < nc; ++c) {
        pm += u[c];
        u[c] = u[c] / den / (double)g->fdm;
      }
      if (grid_mass) grid_mass[1 + p] = pm;
    }
  }
  free(off);
  free(rs.own);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Mode B adjoint, FP64.  L = sum_c g_c rho_c over the base grid and all */
/* patch cells; returns dL/dm, dL/dx, dL/dy, dL/dz per particle.         */
/* No reference counterpart: derived from the mode B forward above.      */
/* ------------------------------------------------------------------ */
int orc_adjoint_b64(const orc_geom *g, int64_t n, const float *x, const float *y,
                    const float *z, const float *m, const double *g1, const double *g11,
                    double *gm, double *gx, double *gy, double *gz, int nthreads) {
  const int nx = g->nx, ny = g->ny, nz = g->nz, nl = g->nl;
  int nt = nthreads_eff(nthreads);
  rad_set rs = rad_get(g);
  int64_t *off = patch_offsets(g);
  const float xl = -g->lado0 / 2.0f, yl = -g->lado0 / 2.0f, zl = -g->lado0 / 2.0f;
  double den0 = (double)g->dx * g->dy * g->dz * g->rote * (double)g->rete * g->rete * g->rete;
  double sc0 = 1.0 / (den0 * (double)g->fdm);
  double lev_sc[ORC_NLEVELS + 1];
  float lev_d[ORC_NLEVELS + 1][3];
  for (int ir = 1; ir <= nl; ++ir) {
    float p2 = powi_f(2.0f, ir);
    lev_d[ir][0] = g->dx / p2; lev_d[ir][1] = g->dy / p2; lev_d[ir][2] = g->dz / p2;
    double den = (double)lev_d[ir][0] * lev_d[ir][1] * lev_d[ir][2] * g->rote *
                 (double)g->rete * g->rete * g->rete;
    lev_sc[ir] = 1.0 / (den * (double)g->fdm);
  }
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
  for (int64_t i = 0; i < n; ++i) {
    float xp = x[i], yp = y[i], zp = z[i];
    double am = 0, ax = 0, ay = 0, az = 0;
    {
      int ix = b_base_index(xp, xl, g->dx, nx);
      int jy = b_base_index(yp, yl, g->dy, ny);
      int kz = b_base_index(zp, zl, g->dz, nz);
      double vx[3], vy[3], vz[3], dvx[3], dvy[3], dvz[3];
      b_weights64(xp, rs.radx + ix - 1, g->dx, vx);
      b_weights64(yp, rs.rady + jy - 1, g->dy, vy);
      b_weights64(zp, rs.radz + kz - 1, g->dz, vz);
      b_dweights64(xp, rs.radx + ix - 1, g->dx, dvx);
      b_dweights64(yp, rs.rady + jy - 1, g->dy, dvy);
      b_dweights64(zp, rs.radz + kz - 1, g->dz, dvz);
      for (int kk = -1; kk <= 1; ++kk)
        for (int jj = -1; jj <= 1; ++jj)
          for (int ii = -1; ii <= 1; ++ii) {
            int i1 = ix + ii, j1 = jy + jj, k1 = kz + kk;
            if (i1 == nx + 1) i1 = 1;
            if (i1 == 0) i1 = nx;
            if (j1 == ny + 1) j1 = 1;
            if (j1 == 0) j1 = ny;
            if (k1 == nz + 1) k1 = 1;
            if (k1 == 0) k1 = nz;
            double gc = g1[cidx(i1, j1, k1, nx, ny)] * sc0;
            am += gc * vx[ii + 1] * vy[jj + 1] * vz[kk + 1];
            ax += gc * dvx[ii + 1] * vy[jj + 1] * vz[kk + 1];
            ay += gc * vx[ii + 1] * dvy[jj + 1] * vz[kk + 1];
            az += gc * vx[ii + 1] * vy[jj + 1] * dvz[kk + 1];
          }
    }
    for (int ir = 1; ir <= nl; ++ir) {
      float dxpa = lev_d[ir][0], dypa = lev_d[ir][1], dzpa = lev_d[ir][2];
      int lo, hi;
      level_range(g, ir, &lo, &hi);
      for (int p = lo; p < hi; ++p) {
        b_box b = b_patch_box(g, p, dxpa, dypa, dzpa);
        if (!b_in_box(&b, xp, yp, zp)) continue;
        const int n1 = g->pnx[p], n2 = g->pny[p], n3 = g->pnz[p];
        int ix = (int)((xp - b.xl) / dxpa);
        int jy = (int)((yp - b.yl) / dypa);
        int kz = (int)((zp - b.zl) / dzpa);
        float rx[3], ry[3], rz[3];
        for (int t = 0; t < 3; ++t) {
          rx[t] = b_minirad(g->prx[p], ix - 1 + t, dxpa);
          ry[t] = b_minirad(g->pry[p], jy - 1 + t, dypa);
          rz[t] = b_minirad(g->prz[p], kz - 1 + t, dzpa);
        }
        double vx[3], vy[3], vz[3], dvx[3], dvy[3], dvz[3];
        b_weights64(xp, rx, dxpa, vx);
        b_weights64(yp, ry, dypa, vy);
        b_weights64(zp, rz, dzpa, vz);
        b_dweights64(xp, rx, dxpa, dvx);
        b_dweights64(yp, ry, dypa, dvy);
        b_dweights64(zp, rz, dzpa, dvz);
        const double *gp = g11 + off[p];
        for (int kk = -1; kk <= 1; ++kk)
          for (int jj = -1; jj <= 1; ++jj)
            for (int ii = -1; ii <= 1; ++ii) {
              int i1 = ix + ii, j1 = jy + jj, k1 = kz + kk;
              if (!(i1 > 0 && i1 <= n1 && j1 > 0 && j1 <= n2 && k1 > 0 && k1 <= n3)) continue;
              double gc = gp[cidx(i1, j1, k1, n1, n2)] * lev_sc[ir];
              am += gc * vx[ii + 1] * vy[jj + 1] * vz[kk + 1];
              ax += gc * dvx[ii + 1] * vy[jj + 1] * vz[kk + 1];
              ay += gc * vx[ii + 1] * dvy[jj + 1] * vz[kk + 1];
              az += gc * vx[ii + 1] * vy[jj + 1] * dvz[kk + 1];
            }
      }
    }
    gm[i] = am;
    gx[i] = (double)m[i] * ax;
    gy[i] = (double)m[i] * ay;
    gz[i] = (double)m[i] * az;
  }
  free(off);
  free(rs.own);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Membership pairs with their reference cell indices.                   */
/* Each record: {grid, particle, ix, jy, kz} with grid 0 = base and      */
/* grid 1+p = patch p (0-based p).  Order: base grid first (particle     */
/* ascending), then patches in order, particles ascending -- the order   */
/* the reference visits them.                                            */
/*   mode 1 (A): base IX=INT((XP-XL)/DX)+1 before wrap; patch membership */
/*               grids.f:902-909, IX=FLOOR(..)+1.  Only pairs that reach */
/*               the stencil loop (the "cloud misses the patch" CYCLE of  */
/*               grids.f:925-927 removes the rest) are listed when       */
/*               mode == 1; mode == 3 lists every box member.            */
/*   mode 2 (B): base index clamped; patch membership grids.f:1318-1320, */
/*               IX = INT((XP-XL)/DXPA)+1-1.                             */
/* Returns the number of records; writes at most cap of them.            */
/* ------------------------------------------------------------------ */
int64_t orc_pairs(const orc_geom *g, int mode, int64_t n, const float *x, const float *y,
                  const float *z, const int32_t *npart_esp, int32_t *out, int64_t cap) {
  const int nx = g->nx, ny = g->ny, nz = g->nz, nl = g->nl;
  const float xl = -g->lado0 / 2.0f, yl = -g->lado0 / 2.0f, zl = -g->lado0 / 2.0f;
  int64_t cnt = 0;
#define EMIT(G, P, I, J, K)            \
  do {                                 \
    if (out && cnt < cap) {            \
      int32_t *r = out + 5 * cnt;      \
      r[0] = (G); r[1] = (int32_t)(P); \
      r[2] = (I); r[3] = (J); r[4] = (K); \
    }                                  \
    ++cnt;                             \
  } while (0)
  for (int64_t i = 0; i < n; ++i) {
    if (mode == 2)
      EMIT(0, i, b_base_index(x[i], xl, g->dx, nx), b_base_index(y[i], yl, g->dy, ny),
           b_base_index(z[i], zl, g->dz, nz));
    else
      EMIT(0, i, a_base_index(x[i], xl, g->dx), a_base_index(y[i], yl, g->dy),
           a_base_index(z[i], zl, g->dz));
  }
  int64_t sp[ORC_N_ESP + 1];
  sp[0] = 0;
  for (int s = 0; s < ORC_N_ESP; ++s) sp[s + 1] = sp[s] + (npart_esp ? npart_esp[s] : 0);
  for (int ir = 1; ir <= nl; ++ir) {
    float p2 = powi_f(2.0f, ir);
    float dxpa = g->dx / p2, dypa = g->dy / p2, dzpa = g->dz / p2;
    int lo, hi;
    level_range(g, ir, &lo, &hi);
    for (int p = lo; p < hi; ++p) {
      const int n1 = g->pnx[p], n2 = g->pny[p], n3 = g->pnz[p];
      if (mode == 2) {
        b_box b = b_patch_box(g, p, dxpa, dypa, dzpa);
        for (int64_t i = 0; i < n; ++i) {
          if (!b_in_box(&b, x[i], y[i], z[i])) continue;
          EMIT(1 + p, i, (int)((x[i] - b.xl) / dxpa), (int)((y[i] - b.yl) / dypa),
               (int)((z[i] - b.zl) / dzpa));
        }
      } else {
        a_box b = a_patch_box(g, p, dxpa, dypa, dzpa);
        for (int s = 0; s < ORC_N_ESP; ++s) {
          int irk = ir - s;
          if (irk < 0) irk = 0;
          int nk = 1 << irk;
          for (int64_t i = sp[s]; i < sp[s + 1]; ++i) {
            if (!a_in_box(&b, x[i], y[i], z[i])) continue;
            int ix = a_patch_index(x[i], b.xlfix, dxpa);
            int jy = a_patch_index(y[i], b.ylfix, dypa);
            int kz = a_patch_index(z[i], b.zlfix, dzpa);
            if (mode == 1 && (ix - nk > n1 || ix + nk < 1 || jy - nk > n2 || jy + nk < 1 ||
                              kz - nk > n3 || kz + nk < 1))
              continue;
            EMIT(1 + p, i, ix, jy, kz);
          }
        }
      }
    }
  }
#undef EMIT
  return cnt;
}

/* Stencil work of mode A: number of (particle, cell) updates per level,
 * out[ir] for ir = 0..nl (the stencil-size statistic of the density build). */
void orc_stencil_work_a(const orc_geom *g, int64_t n, const float *x, const float *y,
                        const float *z, const int32_t *npart_esp, int64_t *out) {
  const int nl = g->nl;
  int64_t ntot = 0;
  for (int s = 0; s < ORC_N_ESP; ++s) ntot += npart_esp[s];
  (void)n;
  out[0] = 27 * ntot;
  int64_t sp[ORC_N_ESP + 1];
  sp[0] = 0;
  for (int s = 0; s < ORC_N_ESP; ++s) sp[s + 1] = sp[s] + npart_esp[s];
  for (int ir = 1; ir <= nl; ++ir) {
    int64_t w = 0;
    float p2 = powi_f(2.0f, ir);
    float dxpa = g->dx / p2, dypa = g->dy / p2, dzpa = g->dz / p2;
    int lo, hi;
    level_range(g, ir, &lo, &hi);
    for (int p = lo; p < hi; ++p) {
      const int n1 = g->pnx[p], n2 = g->pny[p], n3 = g->pnz[p];
      a_box b = a_patch_box(g, p, dxpa, dypa, dzpa);
      for (int s = 0; s < ORC_N_ESP; ++s) {
        int irk = ir - s;
        if (irk < 0) irk = 0;
        int nk = 1 << irk;
        for (int64_t i = sp[s]; i < sp[s + 1]; ++i) {
          if (!a_in_box(&b, x[i], y[i], z[i])) continue;
          int ix = a_patch_index(x[i], b.xlfix, dxpa);
          int jy = a_patch_index(y[i], b.ylfix, dypa);
          int kz = a_patch_index(z[i], b.zlfix, dzpa);
          int i1 = ix - nk < 1 ? 1 : ix - nk, i2 = ix + nk > n1 ? n1 : ix + nk;
          int j1 = jy - nk < 1 ? 1 : jy - nk, j2 = jy + nk > n2 ? n2 : jy + nk;
          int k1 = kz - nk < 1 ? 1 : kz - nk, k2 = kz + nk > n3 ? n3 : kz + nk;
          if (i2 >= i1 && j2 >= j1 && k2 >= k1) w += (int64_t)(i2 - i1 + 1) * (j2 - j1 + 1) * (k2 - k1 + 1);
        }
      }
    }
    out[ir] = w;
  }
}

/* ------------------------------------------------------------------ */
/* Mode B loss of ONE particle, FP64: sum_c g_c rho_c over the cells its */
/* own deposit touches (base grid and every patch whose box holds it).   */
/* By linearity of the deposit, L(x_i + h) - L(x_i - h) for the full     */
/* snapshot equals this quantity's difference, which makes central       */
/* differences cheap at any grid size.  Same formulas as orc_density_b64.*/
/* ------------------------------------------------------------------ */
double orc_loss_one_b64(const orc_geom *g, float xp, float yp, float zp, float m, const double *g1,
                        const double *g11) {
  const int nx = g->nx, ny = g->ny, nz = g->nz, nl = g->nl;
  rad_set rs = rad_get(g);
  const float xl = -g->lado0 / 2.0f, yl = -g->lado0 / 2.0f, zl = -g->lado0 / 2.0f;
  double L = 0.0;
  {
    int ix = b_base_index(xp, xl, g->dx, nx), jy = b_base_index(yp, yl, g->dy, ny),
        kz = b_base_index(zp, zl, g->dz, nz);
    double vx[3], vy[3], vz[3];
    b_weights64(xp, rs.radx + ix - 1, g->dx, vx);
    b_weights64(yp, rs.rady + jy - 1, g->dy, vy);
    b_weights64(zp, rs.radz + kz - 1, g->dz, vz);
    double den0 = (double)g->dx * g->dy * g->dz * g->rote * (double)g->rete * g->rete * g->rete;
    for (int kk = -1; kk <= 1; ++kk)
      for (int jj = -1; jj <= 1; ++jj)
        for (int ii = -1; ii <= 1; ++ii) {
          int i1 = ix + ii, j1 = jy + jj, k1 = kz + kk;
          if (i1 == nx + 1) i1 = 1;
          if (i1 == 0) i1 = nx;
          if (j1 == ny + 1) j1 = 1;
          if (j1 == 0) j1 = ny;
          if (k1 == nz + 1) k1 = 1;
          if (k1 == 0) k1 = nz;
          L += g1[cidx(i1, j1, k1, nx, ny)] * ((double)m * vx[ii + 1] * vy[jj + 1] * vz[kk + 1]) / den0 /
               (double)g->fdm;
        }
  }
  int64_t *off = patch_offsets(g);
  for (int ir = 1; ir <= nl; ++ir) {
    float p2 = powi_f(2.0f, ir);
    float dxpa = g->dx / p2, dypa = g->dy / p2, dzpa = g->dz / p2;
    double den = (double)dxpa * dypa * dzpa * g->rote * (double)g->rete * g->rete * g->rete;
    int lo, hi;
    level_range(g, ir, &lo, &hi);
    for (int p = lo; p < hi; ++p) {
      b_box b = b_patch_box(g, p, dxpa, dypa, dzpa);
      if (!b_in_box(&b, xp, yp, zp)) continue;
      const int n1 = g->pnx[p], n2 = g->pny[p], n3 = g->pnz[p];
      int ix = (int)((xp - b.xl) / dxpa), jy = (int)((yp - b.yl) / dypa), kz = (int)((zp - b.zl) / dzpa);
      float rx[3], ry[3], rz[3];
      for (int t = 0; t < 3; ++t) {
        rx[t] = b_minirad(g->prx[p], ix - 1 + t, dxpa);
        ry[t] = b_minirad(g->pry[p], jy - 1 + t, dypa);
        rz[t] = b_minirad(g->prz[p], kz - 1 + t, dzpa);
      }
      double vx[3], vy[3], vz[3];
      b_weights64(xp, rx, dxpa, vx);
      b_weights64(yp, ry, dypa, vy);
      b_weights64(zp, rz, dzpa, vz);
      for (int kk = -1; kk <= 1; ++kk)
        for (int jj = -1; jj <= 1; ++jj)
          for (int ii = -1; ii <= 1; ++ii) {
            int i1 = ix + ii, j1 = jy + jj, k1 = kz + kk;
            if (i1 > 0 && i1 <= n1 && j1 > 0 && j1 <= n2 && k1 > 0 && k1 <= n3)
              L += g11[off[p] + cidx(i1, j1, k1, n1, n2)] * ((double)m * vx[ii + 1] * vy[jj + 1] * vz[kk + 1]) /
                   den / (double)g->fdm;
          }
    }
  }
  free(off);
  free(rs.own);
  return L;
}

/* ------------------------------------------------------------------ */
/* Species by local density: SORT_DM_PARTICLES_LOCALDENSITY            */
/* (reader.f:140-336), FP32, sequential (the reference without OpenMP). */
/*   grid: the reference's NMAX x NMAY x NMAZ cells of DX, centred on 0  */
/*   (XMIN = -NMAX*DX/2, reader.f:183-185); NGP cell INT((X-XMIN)/DX)+1  */
/*   clamped to [1, NMAX] (reader.f:190-200); mass sums in particle      */
/*   order; divided by BAS = DX*DY*DZ*RODO*RE0**3 (reader.f:203-212);     */
/*   species = MAX(MIN(INT(LOG(D)/LOG(8.0)), N_ESP-1), 0) for D > 0,     */
/*   else 0 (reader.f:219-236); then the stable partition by species     */
/*   (reader.f:244-262).  order[k] = original index (0-based) of the     */
/*   particle at position k.                                             */
/* ------------------------------------------------------------------ */
static int species_cell(float v, float vmin, float d, int n) {
  int i = (int)((v - vmin) / d) + 1;
  if (i < 1) i = 1;
  if (i > n) i = n;
  return i - 1;
}

int orc_species32(int nx, int ny, int nz, float dx, float dy, float dz, float rodo, float re0, int64_t n,
                  const float *x, const float *y, const float *z, const float *m, int32_t *species,
                  int32_t *npart_esp, int64_t *order) {
  const size_t nc = (size_t)nx * ny * nz;
  float *dens = (float *)calloc(nc > 0 ? nc : 1, sizeof(float));
  if (!dens) return -1;
  const float xmin = -(((float)nx * dx) / 2.0f);
  const float ymin = -(((float)ny * dy) / 2.0f);
  const float zmin = -(((float)nz * dz) / 2.0f);
  for (int64_t i = 0; i < n; ++i) {
    const int ix = species_cell(x[i], xmin, dx, nx), jy = species_cell(y[i], ymin, dy, ny),
              kz = species_cell(z[i], zmin, dz, nz);
    dens[(size_t)ix + (size_t)nx * ((size_t)jy + (size_t)ny * kz)] += m[i];
  }
  const float bas = (((dx * dy) * dz) * rodo) * powi_f(re0, 3);
  for (size_t c = 0; c < nc; ++c) dens[c] = dens[c] / bas;
  const float l8 = logf(8.0f);
  for (int64_t i = 0; i < n; ++i) {
    const int ix = species_cell(x[i], xmin, dx, nx), jy = species_cell(y[i], ymin, dy, ny),
              kz = species_cell(z[i], zmin, dz, nz);
    const float b = dens[(size_t)ix + (size_t)nx * ((size_t)jy + (size_t)ny * kz)];
    int s = 0;
    if (b > 0.0f) {
      s = (int)(logf(b) / l8);
      if (s > ORC_N_ESP - 1) s = ORC_N_ESP - 1;
      if (s < 0) s = 0;
    }
    species[i] = s;
  }
  free(dens);
  int64_t k = 0;
  for (int s = 0; s < ORC_N_ESP; ++s) {
    int64_t c0 = k;
    for (int64_t i = 0; i < n; ++i)
      if (species[i] == s) order[k++] = i;
    npart_esp[s] = (int32_t)(k - c0);
  }
  return 0;
}

/* ARGSORT (num.f:150-200): bottom-up merge sort of indices 1..n by a[] in
 * increasing order.  Runs of width DELTA start at 1, 1+2*DELTA, ...; a merge
 * takes the left head only when it is strictly smaller, so of two equal keys
 * the one from the right run goes first.  indx is 1-based as in Fortran. */
void orc_argsort32(int64_t n, const float* a, int32_t* indx) {
  for (int64_t i = 0; i < n; ++i) indx[i] = (int32_t)(i + 1);
  if (n <= 1) return;
  int32_t* tmp = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  for (int64_t delta = 1; delta < n; delta *= 2) {
    for (int64_t iii = 1; iii <= n - delta; iii += 2 * delta) {
      int64_t i1 = iii, i2 = iii + delta;
      const int64_t kkk = (2 * delta < n - iii + 1) ? 2 * delta : n - iii + 1;
      int64_t k = 1;
      while (i1 < iii + delta && i2 < iii + kkk) {
        if (a[indx[i1 - 1] - 1] < a[indx[i2 - 1] - 1]) tmp[(k++) - 1] = indx[(i1++) - 1];
        else tmp[(k++) - 1] = indx[(i2++) - 1];
      }
      if (i1 < iii + delta) {
        for (; k <= kkk; ++k) tmp[k - 1] = indx[(i1++) - 1];
      } else {
        for (; k <= kkk; ++k) tmp[k - 1] = indx[(i2++) - 1];
      }
      for (int64_t q = 0; q < kkk; ++q) indx[iii - 1 + q] = tmp[q];
    }
  }
  free(tmp);
}

/* Peak candidates of HALOFIND_GRID on a set of grids g = 0..ng-1 (the base
 * grid alone, or the patches of one level in patch order): field f, x
 * fastest, grid g at off[g] with n1[g] x n2[g] x n3[g] cells; mask (same
 * layout, or NULL for all ones) plays CONTA1.  A cell outside the one-cell
 * border (BOR = BORAMR = 1) is a candidate when f >= CONTRASTEC/6.0, its
 * mask is 1, and it is a strict local maximum along the six face
 * neighbours, tested as the reference does (haloes_grids.f:325-347 and
 * 712-734: f(x+1)-f(x) < 0 and f(x)-f(x-1) > 0, likewise in y and z).
 * Candidates are collected in grid, z, y, x order and sorted as SORT_CELLS_AMR
 * does: keys 1/f, ARGSORT, and f returned as 1/(1/f).  Returns the number of
 * candidates; the first min(K, cap) are written (ix, iy, iz 1-based). */
int64_t orc_peaks32(int ng, const int32_t* n1, const int32_t* n2, const int32_t* n3, const int64_t* off,
                    const float* f, const int32_t* mask, float contrastec, int64_t cap, float* dens,
                    int32_t* ix, int32_t* iy, int32_t* iz, int32_t* grid) {
  const float peak = contrastec / 6.0f;
  int64_t k = 0, kcap = 1024;
  float* d = (float*)malloc(sizeof(float) * (size_t)kcap);
  int32_t* cx = (int32_t*)malloc(sizeof(int32_t) * (size_t)kcap * 4);
  for (int g = 0; g < ng; ++g) {
    const int a = n1[g], b = n2[g], c = n3[g];
    const float* u = f + off[g];
    const int32_t* mk = mask ? mask + off[g] : NULL;
#define U(i, j, l) u[(size_t)((i) - 1) + (size_t)a * ((size_t)((j) - 1) + (size_t)b * (size_t)((l) - 1))]
    for (int kz = 2; kz <= c - 1; ++kz)
      for (int jy = 2; jy <= b - 1; ++jy)
        for (int ii = 2; ii <= a - 1; ++ii) {
          const float bas = U(ii, jy, kz);
          if (!(bas >= peak)) continue;
          if (mk && mk[(size_t)(ii - 1) + (size_t)a * ((size_t)(jy - 1) + (size_t)b * (size_t)(kz - 1))] != 1)
            continue;
          if (!(U(ii + 1, jy, kz) - bas < 0.0f)) continue;
          if (!(U(ii, jy + 1, kz) - bas < 0.0f)) continue;
          if (!(U(ii, jy, kz + 1) - bas < 0.0f)) continue;
          if (!(bas - U(ii - 1, jy, kz) > 0.0f)) continue;
          if (!(bas - U(ii, jy - 1, kz) > 0.0f)) continue;
          if (!(bas - U(ii, jy, kz - 1) > 0.0f)) continue;
          if (k == kcap) {
            kcap *= 2;
            d = (float*)realloc(d, sizeof(float) * (size_t)kcap);
            cx = (int32_t*)realloc(cx, sizeof(int32_t) * (size_t)kcap * 4);
          }
          d[k] = bas;
          cx[4 * k] = ii;
          cx[4 * k + 1] = jy;
          cx[4 * k + 2] = kz;
          cx[4 * k + 3] = g;
          ++k;
        }
#undef U
  }
  float* d2 = (float*)malloc(sizeof(float) * (size_t)(k > 0 ? k : 1));
  int32_t* ind = (int32_t*)malloc(sizeof(int32_t) * (size_t)(k > 0 ? k : 1));
  for (int64_t i = 0; i < k; ++i) d2[i] = 1.0f / d[i];
  orc_argsort32(k, d2, ind);
  for (int64_t i = 0; i < k && i < cap; ++i) {
    const int64_t j = ind[i] - 1;
    dens[i] = 1.0f / d2[j];
    ix[i] = cx[4 * j];
    iy[i] = cx[4 * j + 1];
    iz[i] = cx[4 * j + 2];
    grid[i] = cx[4 * j + 3];
  }
  free(d);
  free(cx);
  free(d2);
  free(ind);
  return k;
}

/* COMPUTE_CR0AMR (grids.f:2076-2226) on packed patch arrays (patch p at
 * off[p], x fastest).  cr0amr[nx*ny*nz] and cr0amr11 receive 1 for cells not
 * refined, 0 for the mother cells of a finer patch: the level-1 patches'
 * mothers on the base grid (PATCHX..Z = the first mother cell), level >= 2
 * mothers in the parent patch PARE, and cells of level IR whose centre lies
 * between the first and last cell centres of a level IR+1 patch that is not
 * their child.  DX here is LADO0/NX, as the routine computes it.  Returns -1
 * if a mother index falls outside its grid (the reference would write out
 * of bounds). */
int orc_cr0amr(int nx, int ny, int nz, int nl, const int32_t* npatch, const int32_t* pnx, const int32_t* pny,
               const int32_t* pnz, const int32_t* px, const int32_t* py, const int32_t* pz, const int32_t* pare,
               const float* prx, const float* pry, const float* prz, float lado0, int32_t* cr0amr,
               int32_t* cr0amr11) {
  int64_t ptot = 0;
  for (int l = 1; l <= nl; ++l) ptot += npatch[l];
  int64_t* off = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ptot + 1));
  off[0] = 0;
  for (int64_t p = 0; p < ptot; ++p) off[p + 1] = off[p] + (int64_t)pnx[p] * pny[p] * pnz[p];
  for (int64_t c = 0; c < (int64_t)nx * ny * nz; ++c) cr0amr[c] = 1;
  for (int64_t c = 0; c < off[ptot]; ++c) cr0amr11[c] = 1;
  int rc = 0;
  /* level 1: mothers on the base grid (the loop steps by 2, grids.f:2123-2135) */
  for (int64_t i = 0; i < (nl >= 1 ? npatch[1] : 0); ++i)
    for (int kz = 1; kz <= pnz[i]; kz += 2)
      for (int jy = 1; jy <= pny[i]; jy += 2)
        for (int ix = 1; ix <= pnx[i]; ix += 2) {
          const int c1 = px[i] - 1 + (ix + 1) / 2, c2 = py[i] - 1 + (jy + 1) / 2, c3 = pz[i] - 1 + (kz + 1) / 2;
          if (c1 < 1 || c1 > nx || c2 < 1 || c2 > ny || c3 < 1 || c3 > nz) { rc = -1; continue; }
          cr0amr[(size_t)(c1 - 1) + (size_t)nx * ((size_t)(c2 - 1) + (size_t)ny * (size_t)(c3 - 1))] = 0;
        }
  /* levels >= 2: every cell marks its mother in the parent (grids.f:2138-2160) */
  int64_t low1 = nl >= 1 ? npatch[1] : 0;
  for (int ir = 2; ir <= nl; ++ir) {
    for (int64_t i = low1; i < low1 + npatch[ir]; ++i) {
      const int64_t q = pare[i] - 1;
      if (q < 0 || q >= ptot) { rc = -1; continue; }
      for (int kz = 1; kz <= pnz[i]; ++kz)
        for (int jy = 1; jy <= pny[i]; ++jy)
          for (int ix = 1; ix <= pnx[i]; ++ix) {
            const int c1 = px[i] - 1 + (ix + 1) / 2, c2 = py[i] - 1 + (jy + 1) / 2, c3 = pz[i] - 1 + (kz + 1) / 2;
            if (c1 < 1 || c1 > pnx[q] || c2 < 1 || c2 > pny[q] || c3 < 1 || c3 > pnz[q]) { rc = -1; continue; }
            cr0amr11[off[q] + (c1 - 1) + (int64_t)pnx[q] * ((c2 - 1) + (int64_t)pny[q] * (c3 - 1))] = 0;
          }
    }
    low1 += npatch[ir];
  }
  /* cells inside a non-child finer patch (grids.f:2162-2224) */
  const float dx = lado0 / (float)nx, dy = lado0 / (float)ny, dz = lado0 / (float)nz;
  int64_t lo = 0;
  for (int ir = 1; ir <= nl - 1; ++ir) {
    float p2 = 1.0f;
    for (int k = 0; k < ir; ++k) p2 *= 2.0f;
    const float dxpa = dx / p2, dypa = dy / p2, dzpa = dz / p2;
    const int64_t lo2 = lo + npatch[ir], hi2 = lo2 + npatch[ir + 1];
    for (int64_t i = lo; i < lo + npatch[ir]; ++i)
      for (int kz = 1; kz <= pnz[i]; ++kz)
        for (int jy = 1; jy <= pny[i]; ++jy)
          for (int ix = 1; ix <= pnx[i]; ++ix) {
            int32_t* cell = &cr0amr11[off[i] + (ix - 1) + (int64_t)pnx[i] * ((jy - 1) + (int64_t)pny[i] * (kz - 1))];
            if (*cell != 1) continue;
            const float xx = (prx[i] - 0.5f * dxpa) + (float)(ix - 1) * dxpa;
            const float yy = (pry[i] - 0.5f * dypa) + (float)(jy - 1) * dypa;
            const float zz = (prz[i] - 0.5f * dzpa) + (float)(kz - 1) * dzpa;
            for (int64_t ii = lo2; ii < hi2; ++ii) {
              if (pare[ii] == (int32_t)(i + 1)) continue;
              const float xx1 = prx[ii] - 0.5f * (dxpa * 0.5f), yy1 = pry[ii] - 0.5f * (dypa * 0.5f),
                          zz1 = prz[ii] - 0.5f * (dzpa * 0.5f);
              const float xx2 = xx1 + ((float)(pnx[ii] - 1) * dxpa) * 0.5f;
              const float yy2 = yy1 + ((float)(pny[ii] - 1) * dypa) * 0.5f;
              const float zz2 = zz1 + ((float)(pnz[ii] - 1) * dzpa) * 0.5f;
              const float b1 = (xx - xx1) * (xx2 - xx), b2 = (yy - yy1) * (yy2 - yy), b3 = (zz - zz1) * (zz2 - zz);
              if (b1 >= 0.0f && b2 >= 0.0f && b3 >= 0.0f) {
                *cell = 0;
                break;
              }
            }
          }
    lo += npatch[ir];
  }
  free(off);
  return rc;
}

/* VEINSGRID (grids.f:2229-2506) for every level 1..nl, on packed patch
 * arrays: SOLAP = 1, except where patches of one level overlap, resolved in
 * the reference's sequential order (patch I ascending, its neighbours J
 * ascending, cells KK, JJ, II): a cell kept by both I and J is dropped from
 * I when it lies on I's border, else from J.  radx1..3 = RADX(1), RADY(1),
 * RADZ(1) and dx..dz = COMMON /ESPACIADO/ (MALLA's).  Returns -1 if a mapped
 * cell index leaves its patch. */
static int vg_int(float v) { return (int)v; } /* Fortran INT: truncation */
int orc_veinsgrid(int nl, const int32_t* npatch, const int32_t* pnx, const int32_t* pny, const int32_t* pnz,
                  const float* prx, const float* pry, const float* prz, float dx, float dy, float dz, float radx1,
                  float rady1, float radz1, int32_t* solap) {
  int64_t ptot = 0;
  for (int l = 1; l <= nl; ++l) ptot += npatch[l];
  int64_t* off = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ptot + 1));
  off[0] = 0;
  for (int64_t p = 0; p < ptot; ++p) off[p + 1] = off[p] + (int64_t)pnx[p] * pny[p] * pnz[p];
  int rc = 0;
  int64_t low1 = 0;
  for (int ir = 1; ir <= nl; ++ir) {
    const int64_t np = npatch[ir], lo = low1;
    float p2 = 1.0f;
    for (int k = 0; k < ir; ++k) p2 *= 2.0f;
    const float dxpa = dx / p2, dypa = dy / p2, dzpa = dz / p2;
    const float rxfix = (radx1 - dx * 0.5f) + 0.5f * dxpa;
    const float ryfix = (rady1 - dy * 0.5f) + 0.5f * dypa;
    const float rzfix = (radz1 - dz * 0.5f) + 0.5f * dzpa;
    int* g = (int*)malloc(sizeof(int) * 6 * (size_t)(np > 0 ? np : 1));
    for (int64_t k = 0; k < np; ++k) {
      const int64_t i = lo + k;
      const float rx1 = prx[i] - 0.5f * dxpa, ry1 = pry[i] - 0.5f * dypa, rz1 = prz[i] - 0.5f * dzpa;
      g[6 * k] = vg_int(((rx1 - rxfix) / dxpa) + 0.5f) + 1;
      g[6 * k + 1] = vg_int(((ry1 - ryfix) / dypa) + 0.5f) + 1;
      g[6 * k + 2] = vg_int(((rz1 - rzfix) / dzpa) + 0.5f) + 1;
      g[6 * k + 3] = g[6 * k] + pnx[i] - 1;
      g[6 * k + 4] = g[6 * k + 1] + pny[i] - 1;
      g[6 * k + 5] = g[6 * k + 2] + pnz[i] - 1;
      for (int64_t c = off[i]; c < off[i + 1]; ++c) solap[c] = 1;
    }
    for (int64_t k = 0; k < np; ++k) {
      const int64_t i = lo + k;
      const int n1 = pnx[i], n2 = pny[i], n3 = pnz[i];
      const float rx1 = prx[i] - 0.5f * dxpa, ry1 = pry[i] - 0.5f * dypa, rz1 = prz[i] - 0.5f * dzpa;
      const float rx2 = (prx[i] - 0.5f * dxpa) + (float)(n1 - 1) * dxpa;
      const float ry2 = (pry[i] - 0.5f * dypa) + (float)(n2 - 1) * dypa;
      const float rz2 = (prz[i] - 0.5f * dzpa) + (float)(n3 - 1) * dzpa;
      for (int64_t kk2 = 0; kk2 < np; ++kk2) {
        if (kk2 == k) continue;
        if (!(g[6 * k] <= g[6 * kk2 + 3] && g[6 * kk2] <= g[6 * k + 3] && g[6 * k + 1] <= g[6 * kk2 + 4] &&
              g[6 * kk2 + 1] <= g[6 * k + 4] && g[6 * k + 2] <= g[6 * kk2 + 5] && g[6 * kk2 + 2] <= g[6 * k + 5]))
          continue;
        const int64_t j = lo + kk2;
        const int nn1 = pnx[j], nn2 = pny[j], nn3 = pnz[j];
        const float r1[3] = {rx1, ry1, rz1}, r2[3] = {rx2, ry2, rz2};
        const float d[3] = {dxpa, dypa, dzpa};
        const int nme[3] = {n1, n2, n3}, noth[3] = {nn1, nn2, nn3};
        const float pr[3] = {prx[j], pry[j], prz[j]};
        int c1[3], c2[3], cc1[3], cc2[3];
        for (int a = 0; a < 3; ++a) {
          const float q1 = pr[a] - 0.5f * d[a];
          const float q2 = (pr[a] - 0.5f * d[a]) + (float)(noth[a] - 1) * d[a];
          if (q1 >= r1[a] && q2 <= r2[a]) {
            c1[a] = vg_int(((q1 - r1[a]) / d[a]) + 0.5f) + 1;
            c2[a] = vg_int(((q2 - r1[a]) / d[a]) + 0.5f) + 1;
            cc1[a] = 1;
            cc2[a] = noth[a];
          }
          if (q1 >= r1[a] && q2 > r2[a]) {
            c1[a] = vg_int(((q1 - r1[a]) / d[a]) + 0.5f) + 1;
            c2[a] = nme[a];
            cc1[a] = 1;
            cc2[a] = vg_int(((r2[a] - q1) / d[a]) + 0.5f) + 1;
          }
          if (q2 <= r2[a] && q1 < r1[a]) {
            c1[a] = 1;
            c2[a] = vg_int(((q2 - r1[a]) / d[a]) + 0.5f) + 1;
            cc1[a] = vg_int(((r1[a] - q1) / d[a]) + 0.5f) + 1;
            cc2[a] = noth[a];
          }
          if (q1 < r1[a] && q2 > r2[a]) {
            c1[a] = 1;
            c2[a] = nme[a];
            cc1[a] = vg_int(((r1[a] - q1) / d[a]) + 0.5f) + 1;
            cc2[a] = vg_int(((r2[a] - q1) / d[a]) + 0.5f) + 1;
          }
        }
        (void)c2;
        for (int kz2 = cc1[2]; kz2 <= cc2[2]; ++kz2)
          for (int jy2 = cc1[1]; jy2 <= cc2[1]; ++jy2)
            for (int ix2 = cc1[0]; ix2 <= cc2[0]; ++ix2) {
              const int ix = ix2 - cc1[0] + c1[0], jy = jy2 - cc1[1] + c1[1], kz = kz2 - cc1[2] + c1[2];
              if (ix < 1 || ix > n1 || jy < 1 || jy > n2 || kz < 1 || kz > n3 || ix2 < 1 || ix2 > nn1 ||
                  jy2 < 1 || jy2 > nn2 || kz2 < 1 || kz2 > nn3) {
                rc = -1;
                continue;
              }
              int32_t* si = &solap[off[i] + (ix - 1) + (int64_t)n1 * ((jy - 1) + (int64_t)n2 * (kz - 1))];
              int32_t* sj = &solap[off[j] + (ix2 - 1) + (int64_t)nn1 * ((jy2 - 1) + (int64_t)nn2 * (kz2 - 1))];
              if (*si == 1 && *sj == 1) {
                if (ix == 1 || ix == n1 || jy == 1 || jy == n2 || kz == 1 || kz == n3) *si = 0;
                else *sj = 0;
              }
            }
      }
    }
    free(g);
    low1 += np;
  }
  free(off);
  return rc;
}
