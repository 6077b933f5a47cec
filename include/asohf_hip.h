/*
 * asohf_hip.h -- C ABI of libasohf_hip, the MI355X (gfx950) density build of
 * ASOHF: particle-to-mesh assignment onto the base grid and every AMR patch.
 *
 * All entry points are extern "C", take plain pointers and sizes, return an
 * asohf_status, and never abort.  A context is single-threaded: one host
 * thread at a time, one HIP stream.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   SUBROUTINE DENSITY(...)              grids.f:710-994  (+ COMPUTE_KERNELS
 *                                        grids.f:997-1067, RENORM_DENSITY
 *                                        grids.f:2668-2722), called once per
 *                                        snapshot at asohf.f:467-470
 *   SUBROUTINE INTERPOLATE_DENSITY_TSC   grids.f:1070-1398
 */
#ifndef ASOHF_HIP_H
#define ASOHF_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct asohf_ctx asohf_ctx;

typedef enum {
  ASOHF_OK = 0,
  ASOHF_ERR_INVALID = -1,   /* bad argument, geometry or particle outside the wrap range */
  ASOHF_ERR_CAPACITY = -2,  /* a compile-time-style bound of the reference exceeded      */
  ASOHF_ERR_HIP = -3,       /* a HIP runtime call failed                                 */
  ASOHF_ERR_OOM = -4,       /* device allocation failed                                  */
  ASOHF_ERR_NODEVICE = -5   /* no HIP device                                             */
} asohf_status;

/* Deposit semantics. */
enum {
  ASOHF_MODE_TABLE = 1, /* DENSITY: NGP cell + normalised kernel table of width
                           2*2^max(IR-species,0)+1 (grids.f:710-994)           */
  ASOHF_MODE_TSC = 2    /* INTERPOLATE_DENSITY_TSC: position-dependent 27-cell
                           TSC on every level (grids.f:1070-1398)              */
};
/* Reduction of the cell sums. */
enum {
  ASOHF_REDUCE_DETERMINISTIC = 0, /* exact int64 fixed-point sums: bit-identical
                                     on every run, mass-conserving on the base grid */
  ASOHF_REDUCE_ATOMIC = 1         /* FP32 LDS atomics: run-to-run order varies      */
};
/* Execution form. */
enum {
  ASOHF_PATH_BATCHED = 0, /* all grids of all levels in one persistent launch  */
  ASOHF_PATH_PLAIN = 1    /* one launch per grid, global FP32 atomics           */
};

/* Geometry: the DENSITY arguments and COMMON blocks (grids.f:710-742).  Host
 * pointers.  Patches are numbered level-major as in the reference: level IR
 * holds SUM(NPATCH(0:IR-1)) .. SUM(NPATCH(0:IR))-1 (0-based). */
typedef struct asohf_geom {
  int32_t nx, ny, nz;      /* NX, NY, NZ                                   */
  int32_t nl;              /* NL (0..9)                                    */
  const int32_t *npatch;   /* NPATCH(0:NL), nl+1 entries                   */
  const int32_t *patchnx, *patchny, *patchnz; /* PATCHNX/NY/NZ [P]         */
  const float *patchrx, *patchry, *patchrz;   /* PATCHRX/RY/RZ [P]         */
  float dx, dy, dz;        /* COMMON /ESPACIADO/                           */
  float lado0;             /* LADO0                                        */
  float rote, rete;        /* COMMON /BACK/                                */
  float fdm;               /* RENORM_DENSITY divisor (1 = none)            */
  int32_t interp_degree;   /* COMPUTE_KERNELS: 2 TSC-like, 1 CIC-like      */
  const float *radx, *rady, *radz; /* COMMON /GRID/ RADX(0:NX+1) etc., or
                                      NULL: rebuilt as MALLA does with
                                      LADO = LADO0 - LADO0/NX             */
} asohf_geom;

typedef struct asohf_opts {
  int32_t mode;        /* ASOHF_MODE_TABLE or ASOHF_MODE_TSC                 */
  int32_t reduction;   /* ASOHF_REDUCE_*                                     */
  int32_t path;        /* ASOHF_PATH_*                                       */
  int32_t use_graph;   /* 1: replay the batched pipeline as a hipGraph       */
  int32_t u1_ld[2];    /* leading dims of u1 (NMAX, NMAY); 0 = (nx, ny)       */
  int32_t u11_ld[3];   /* NAMRX, NAMRY, NAMRZ of a padded U11; 0 = packed:
                          patch p at offset sum_{q<p} n1*n2*n3, x fastest    */
  void *stream;        /* hipStream_t to run on; NULL = the default (null)
                          stream.  asohf_hip_density_host uses the
                          context's own stream when NULL.                    */
} asohf_opts;

/* Tuning of the batched deposit (chosen by asohf_hip_autotune). */
typedef struct asohf_tune {
  int32_t tile[3];          /* cells per tile along x, y, z (powers of two)   */
  int32_t block;            /* threads per workgroup (64..1024)               */
  int32_t chunk;            /* max entries per work item ("particles per block") */
  int32_t waves_per_cu;     /* persistent grid = waves_per_cu * CUs workgroups */
} asohf_tune;

/* Counters and timings of the last density call. */
typedef struct asohf_stats {
  int64_t n_particles;
  int64_t n_pairs;          /* (particle, grid) members under the mode's box rule */
  int64_t n_entries;        /* binned (particle, tile) entries                     */
  int64_t n_tiles;
  int64_t n_items;          /* work items of the persistent deposit                */
  int64_t n_cells;          /* output cells written                                */
  double ms_total;          /* device time of the whole call (hipEvents)           */
  double ms_bin;            /* count + scan + item planning (before the host read) */
  double ms_scatter;        /* scatter of the entries                              */
  double ms_deposit;        /* the persistent deposit kernel alone                 */
  double ms_sums;           /* item writing and grid sums                          */
  double ms_split;          /* summation of split tiles' partials (split_reduce)   */
} asohf_stats;

/* Context. device < 0 selects the current device. */
asohf_status asohf_hip_create(asohf_ctx **ctx, int32_t device);
void asohf_hip_destroy(asohf_ctx *ctx);
const char *asohf_hip_last_error(const asohf_ctx *ctx);

/* The density build on device arrays.
 *   rx, ry, rz, m : device FP32 [n], particles in reference order (species
 *                   contiguous, ascending), replacing MODULE PARTICLES'
 *                   RXPA, RYPA, RZPA, MASAP (particles.f:9-11)
 *   npart_esp     : host int32 [9], NPART_ESP(0:8); mode TABLE deposits
 *                   SUM(NPART_ESP) particles, mode TSC all n
 *   u1            : device FP32 base field, U1 (interior written only)
 *   u11           : device FP32 patch fields, U11 (interiors written only)
 *   grid_mass     : device FP64 [1+P] or NULL: deposited mass per grid
 *                   (base first), summed in fixed order
 * Replaces DENSITY (grids.f:710) + RENORM_DENSITY (grids.f:2668) in mode
 * TABLE, INTERPOLATE_DENSITY_TSC (grids.f:1070) + RENORM_DENSITY in TSC. */
asohf_status asohf_hip_density(asohf_ctx *ctx, const asohf_geom *geom, int64_t n,
                               const float *rx, const float *ry, const float *rz,
                               const float *m, const int32_t *npart_esp, float *u1,
                               float *u11, double *grid_mass, const asohf_opts *opts);

/* Same on host arrays (copies in and out): the form the Fortran DENSITY
 * shim calls with MODULE PARTICLES and COMMON /VARIA/ U1, U11. */
asohf_status asohf_hip_density_host(asohf_ctx *ctx, const asohf_geom *geom, int64_t n,
                                    const float *rx, const float *ry, const float *rz,
                                    const float *m, const int32_t *npart_esp, float *u1,
                                    float *u11, double *grid_mass, const asohf_opts *opts);

/* Adjoint of mode TSC (no reference counterpart: the reference has no
 * derivatives).  Given dL/du1 and dL/du11 (device FP32, same layouts as the
 * forward outputs), writes dL/dm, dL/dx, dL/dy, dL/dz (device FP32 [n]).
 * Weights are recomputed from the positions. */
asohf_status asohf_hip_density_adjoint(asohf_ctx *ctx, const asohf_geom *geom, int64_t n,
                                       const float *rx, const float *ry, const float *rz,
                                       const float *m, const float *g1, const float *g11,
                                       float *dm, float *dxp, float *dyp, float *dzp,
                                       const asohf_opts *opts);

/* Membership records of the binning stage, for bit-exact index checks:
 * out (device int32 [cap][5]) = {grid, particle, ix, jy, kz} with grid 0 the
 * base grid, 1+p patch p, and the reference's own cell index expressions
 * (mode TABLE: grids.f:798-800, 912-914; mode TSC: grids.f:1180-1192,
 * 1321-1323).  Order is unspecified; *count receives the total. */
asohf_status asohf_hip_membership(asohf_ctx *ctx, const asohf_geom *geom, int64_t n,
                                  const float *rx, const float *ry, const float *rz,
                                  const int32_t *npart_esp, int32_t mode, int32_t *out,
                                  int64_t cap, int64_t *count);

/* Tuning. */
asohf_status asohf_hip_get_tune(const asohf_ctx *ctx, asohf_tune *out);
asohf_status asohf_hip_set_tune(asohf_ctx *ctx, const asohf_tune *tune);
/* Times the batched pipeline over a fixed candidate set on these inputs and
 * keeps the fastest; *out receives it. */
asohf_status asohf_hip_autotune(asohf_ctx *ctx, const asohf_geom *geom, int64_t n,
                                const float *rx, const float *ry, const float *rz,
                                const float *m, const int32_t *npart_esp, float *u1,
                                float *u11, const asohf_opts *opts, asohf_tune *out);

asohf_status asohf_hip_get_stats(const asohf_ctx *ctx, asohf_stats *out);

/* Kernel classes of the particles by local density and their stable
 * partition: SORT_DM_PARTICLES_LOCALDENSITY (reader.f:140-336), the reader's
 * default species rule that feeds NPART_ESP of DENSITY (asohf.f:422-444).
 * nmax[3] = the reference's NMAX, NMAY, NMAZ (its grid for this rule, of
 * cells d[3] = DX, DY, DZ, centred on 0); rodo, re0 = RODO, RE0 (asohf.f:
 * 351-358).  On device arrays: species[n] (may be NULL) receives each
 * particle's class, order[n] the original index (0-based) of the particle at
 * each position of the partition (apply it to every particle array, as the
 * reference reorders RXPA..ORIPA); npart_esp[9] (host) the class counts.
 * Masses are summed per cell in FP32 in particle order, as the reference's
 * loop without OpenMP does.  Synchronises before returning. */
asohf_status asohf_hip_species(asohf_ctx *ctx, const int32_t nmax[3], const float d[3], float rodo, float re0,
                               int64_t n, const float *rx, const float *ry, const float *rz,
                               const float *m, int32_t *species, int64_t *order, int32_t *npart_esp,
                               void *stream);

/* asohf_hip_species without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as asohf_hip_species, for at most 2^20 particles and
 * nmax[0]*nmax[1]*nmax[2] <= 2^21 cells (ASOHF_ERR_CAPACITY beyond).  Its
 * scratch is a 64 MiB array in device memory that every GPU the library runs
 * on holds until the process ends; every call on a GPU shares that GPU's
 * array, so calls that overlap on one GPU (other threads, other contexts,
 * other streams) race.  Synchronises before returning. */
asohf_status asohf_hip_species_noalloc(asohf_ctx *ctx, const int32_t nmax[3], const float d[3], float rodo,
                                       float re0, int64_t n, const float *rx, const float *ry, const float *rz,
                                       const float *m, int32_t *species, int64_t *order, int32_t *npart_esp,
                                       void *stream);

/* Peak candidates of the grid halo finder: the cells HALOFIND_GRID starts
 * its haloes from (haloes_grids.f:316-404 on the base grid, 700-790 on AMR
 * level IR), in the order SORT_CELLS / SORT_CELLS_AMR (haloes_grids.f:
 * 1341-1425, with ARGSORT num.f:150-200) gives them.  A cell outside the
 * one-cell border of its grid is a candidate when its density is at least
 * CONTRASTEC/6 (FP32), its CONTA1 mask is 1 (AMR levels) and it exceeds its
 * six face neighbours.  Order: 1/density ascending; equal keys by
 * decreasing position in the reference's scan order (grid, z, y, x).
 *   level 0  : u1 (device FP32, leading dims opts->u1_ld as asohf_hip_density)
 *   level IR : u11 (device FP32, packed or padded as opts->u11_ld) for the
 *              patches of level IR; conta (device int32, same layout) or NULL
 *              (all ones)
 * Outputs (device, cap entries): dens = 1/(1/density) as the reference
 * writes DDD back, ix/iy/iz 1-based cell indices (DDDX..Z), patch (DDDP: the
 * 1-based global patch number, 0 on the base grid; may be NULL).  *count
 * receives the number of candidates; ASOHF_ERR_CAPACITY if it exceeds cap
 * (nothing written).  Runs on opts->stream and synchronises before
 * returning. */
asohf_status asohf_hip_peak_candidates(asohf_ctx *ctx, const asohf_geom *geom, int32_t level, const float *u1,
                                       const float *u11, const int32_t *conta, float contrastec, float *dens,
                                       int32_t *ix, int32_t *iy, int32_t *iz, int32_t *patch, int64_t cap,
                                       int64_t *count, const asohf_opts *opts);

/* The grid masks of the density dump, as asohf.f:515-528 computes them
 * before writing it.
 * asohf_hip_refinement_masks: COMPUTE_CR0AMR (grids.f:2076-2226).  cr0amr
 * (device int32, u1's layout: opts->u1_ld) and cr0amr11 (device int32,
 * u11's layout: opts->u11_ld, every patch of levels 1..NL; the whole slab in
 * the padded layout) receive 1 for cells that are not refined and 0 for
 * the mother cells of finer patches and for cells whose centre lies within a
 * finer patch that is not their child.  patchx/y/z (host int32 [P]):
 * PATCHX..Z, the parent cell holding a patch's first cells; pare (host int32
 * [P]): PARE, 1-based, 0 for level 1.  DX is LADO0/NX, as the routine
 * computes it. */
asohf_status asohf_hip_refinement_masks(asohf_ctx *ctx, const asohf_geom *geom, const int32_t *patchx,
                                        const int32_t *patchy, const int32_t *patchz, const int32_t *pare,
                                        int32_t *cr0amr, int32_t *cr0amr11, const asohf_opts *opts);

/* asohf_hip_overlap_mask: VEINSGRID (grids.f:2229-2506) for IR = 1..NL.
 * solap (device int32, u11's layout) receives 1, or 0 for a cell that
 * another patch of its level also covers and keeps: overlaps resolved in
 * the reference's sequential order (patches ascending, each with its
 * neighbours ascending; a shared cell kept by both is dropped by the first
 * when it lies on that patch's border, else by the second).  Uses RADX(1)
 * (geom->radx, or MALLA's) and DX..DZ from geom. */
asohf_status asohf_hip_overlap_mask(asohf_ctx *ctx, const asohf_geom *geom, int32_t *solap, const asohf_opts *opts);

/* asohf_hip_write_density_dump: the FW2 density dump of asohf.f:530-545 at
 * `path`, an unformatted sequential file (4-byte record markers): U1 and
 * CR0AMR (interiors 1..NX x 1..NY x 1..NZ, x fastest), then for every patch
 * U11, CR0AMR11 and SOLAP (interiors).  Inputs are device arrays in the
 * layouts of opts (u1_ld for U1 and CR0AMR, u11_ld for the patch arrays).
 * Records longer than 2,147,483,639 bytes are written as gfortran-style
 * subrecords. */
asohf_status asohf_hip_write_density_dump(asohf_ctx *ctx, const char *path, const asohf_geom *geom, const float *u1,
                                          const int32_t *cr0amr, const float *u11, const int32_t *cr0amr11,
                                          const int32_t *solap, const asohf_opts *opts);

/* Parameters of the mesh builder: asohf.dat's mesh block (asohf.f:196-212)
 * and the reference build's compile-time bounds (asohf_parameters.dat). */
typedef struct asohf_mesh_params {
  int32_t nl_mesh;        /* NL_MESH: levels to build (1..9)                     */
  int32_t refine_thr;     /* REFINE_THR                                          */
  int32_t min_patchsize;  /* MIN_PATCHSIZE (child cells)                         */
  float frac_refinable;   /* MINFRAC_REFINABLE                                   */
  int32_t bor, boramr;    /* BOR, BORAMR                                         */
  int32_t bor_ovlp;       /* BOR_OVLP                                            */
  int32_t namr[3];        /* NAMRX, NAMRY, NAMRZ: largest patch, in cells        */
  int32_t npalev;         /* NPALEV: most patches in all                         */
  int32_t mpapolev[9];    /* MPAPOLEV(1:NLEVELS): most patches per level         */
} asohf_mesh_params;

/* The mesh builder CREATE_MESH (grids.f:2-707), called at asohf.f:459-463.
 * base: the base grid (nx..nz, lado0, dx..dz = COMMON /ESPACIADO/, radx..z
 * or NULL for MALLA's); its patch fields are ignored.  rx, ry, rz: device
 * FP32 [n], particles in reference order with npart_esp (host int32 [9])
 * giving the species blocks (PLEV).  Outputs (host): *nl_mesh (levels that
 * received patches), npatch[10] (NPATCH(0:9)), and for the P patches
 * (P <= cap; ASOHF_ERR_CAPACITY otherwise, with npatch set) PATCHNX..Z,
 * PATCHX..Z, PARE and PATCHRX..Z, numbered as the reference numbers them.
 * The particle counts run on the GPU; the patch growth, which each new
 * patch changes for the next, runs on the host as the reference's
 * sequential loops.  Uses opts->stream; synchronises. */
asohf_status asohf_hip_create_mesh(asohf_ctx *ctx, const asohf_geom *base, int64_t n, const float *rx,
                                   const float *ry, const float *rz, const int32_t *npart_esp,
                                   const asohf_mesh_params *params, int32_t *nl_mesh, int32_t *npatch,
                                   int32_t *patchnx, int32_t *patchny, int32_t *patchnz, int32_t *patchx,
                                   int32_t *patchy, int32_t *patchz, int32_t *pare, float *patchrx, float *patchry,
                                   float *patchrz, int64_t cap, const asohf_opts *opts);

/* Frees the device workspaces the context keeps between calls (binned
 * entries, work items, block tables, emission lists, split tiles, the
 * adjoint's sort buffers, the captured graph); the geometry plan stays, and
 * the next call allocates what it needs again.  No counterpart in the
 * reference, which holds no device memory: for callers that alternate the
 * density build with other large device work. */
asohf_status asohf_hip_release_workspaces(asohf_ctx *ctx);

/* Version string of the library. */
const char *asohf_hip_version(void);

#ifdef __cplusplus
}
#endif
#endif
