// api.hip -- host side of libasohf_hip: geometry planning, the batched and
// plain pipelines, the adjoint, and the extern "C" entry points of
// include/asohf_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/asohf_hip.h"
#include "dev.h"
#include "launch.h"

using namespace asohf;

namespace {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
struct DBuf {
  void* p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    size_t want = std::max<size_t>(bytes + bytes / 8, 256);
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) { p = nullptr; return e; }
    cap = want;
    return hipSuccess;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  template <class T> T* as() const { return reinterpret_cast<T*>(p); }
};

// FP32 powers as the reference's REAL**INTEGER (binary powering)
float powi_f(float a, int b) {
  float r = 1.0f;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return r;
}

// MALLA (grids.f:2509-2573): RAD(0:N+1) for a side LADO
void malla(int n, float lado, std::vector<float>& rad) {
  rad.assign(n + 2, 0.f);
  float a = -lado / 2.0f, b = lado / 2.0f;
  float c = (b - a) / (float)(n - 1);
  rad[1] = a;
  for (int i = 2; i <= n; ++i) rad[i] = rad[1] + (float)(i - 1) * c;
  rad[0] = rad[1] - c;
  rad[n + 1] = rad[n] + c;
}

// COMPUTE_KERNELS (grids.f:997-1067) in FP32, the reference's operation
// order, including the sequential normalisation sum.
void compute_kernels(int nl, int degree, std::vector<float>& K) {
  const int kmax = 1 << nl, kw = 2 * kmax + 1;
  K.assign((size_t)kw * (nl + 1), 0.f);
  for (int ir = 0; ir <= nl; ++ir) {
    float* row = K.data() + (size_t)ir * kw + kmax;
    const int n = 1 << ir;
    const float h32 = (float)n + 0.5f;
    const float h = 2.0f * h32 / 3.0f;
    for (int i = -n; i <= n; ++i) {
      if (degree == 2) {
        float bas = (float)std::abs(i) / h;
        if (bas <= 0.5f) row[i] = 0.75f - bas * bas;
        else { float t = 1.5f - bas; row[i] = 0.5f * (t * t); }
      } else {
        float bas = (float)std::abs(i) / h32;
        row[i] = 1.0f - bas;
      }
    }
    float s = 0.0f;
    for (int i = -n; i <= n; ++i)
      for (int j = -n; j <= n; ++j)
        for (int k = -n; k <= n; ++k) s = s + row[i] * row[j] * row[k];
    s = 1.0f / powf(s, 1.0f / 3.0f);
    for (int i = -n; i <= n; ++i) row[i] = s * row[i];
  }
}

// content hash of the geometry arrays (8 bytes per step)
uint64_t fnv(uint64_t h, const void* p, size_t n) {
  const unsigned char* c = (const unsigned char*)p;
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    uint64_t w;
    memcpy(&w, c + i, 8);
    h = (h ^ w) * 0x100000001b3ull;
    h ^= h >> 29;
  }
  for (; i < n; ++i) { h ^= c[i]; h *= 1099511628211ull; }
  return h;
}

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

}  // namespace

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct HostPlan {
  uint64_t key = 0;
  bool valid = false;
  std::vector<GridDesc> grids;
  std::vector<int> tile_grid;
  std::vector<HashSlot> hslots;
  std::vector<BoxRec> hrec;
  std::vector<int> hlist;    // grid id of every candidate (host copy, for the nesting check)
  std::vector<float> radx, rady, radz, ktab;
  // TSC binning along the hierarchy (tgt_off / tgt / grec of Plan)
  std::vector<unsigned> tgt_off;
  std::vector<int> tgt;
  std::vector<BoxRec> grec;
  std::vector<int> level_tile0;   // first tile of every level, [nl + 2]
  unsigned emit_levels = 0;       // bit l: some tile of level l has targets
  Plan dev{};
  long long ncells = 0;
  long long u11_cells = 0;   // cells of the patch output array
  long long u1_cells = 0;
};

// Every value the kernels of one batched call receive.  Zero-initialised
// before it is filled, so two calls with equal arguments compare equal byte
// for byte: a captured graph is replayed only for exactly these values.
struct PipeArgs {
  Plan P;
  Parts A;
  BlockTables BT;
  DepArgs D;
  Counters* C;
  unsigned* cnt;
  unsigned long long* start;
  unsigned long long* cursor;
  unsigned long long* bsum;
  unsigned* nitems;
  unsigned* nsplit;
  unsigned long long* item_off;
  unsigned long long* split_off;
  int* split;
  unsigned* is_split;
  unsigned long long* split_rank;
  int* split_tiles;
  double* grid_mass;
  long long nblk;
  long long nkeys;
  unsigned long long cap_entries;   // entries, items and split-tile slots the buffers hold
  unsigned long long cap_items;
  unsigned long long cap_split;
  int ntiles, chunk, chunk_wide, red, grid, block, cus, elem;
  // TSC binning along the hierarchy
  unsigned* item_pairs;
  long long lvl_items[kNLevels + 2];   // items of level l: [lvl_items[l], lvl_items[l + 1])
  long long n_items;
  unsigned emit_levels;
  int nl;
};

struct asohf_ctx {
  int device = 0;
  int cus = 0;
  hipStream_t stream = nullptr;
  std::string err;
  asohf_tune tune{{16, 8, 8}, 256, 4096, 16};
  HostPlan plan;
  DBuf plan_blob;
  DBuf cnt, start, cursor, bsum, counters, stiles, isplit, srank, nitems, split, tmass, entries, items, scratch;
  DBuf flags, idx, pool, pool_off, pool_nk, blk_pairs, blk_mmax, nsplit, item_off, split_off;
  DBuf elist, eoff, ecount;
  DBuf ipairs;                             // members found per work item (TSC hierarchy binning)
  void* noalloc = nullptr;                 // this device's scratch array of order.hip (device address)
  DBuf sp_k[4], sp_dens, sp_cnt, sp_tmp;   // species partition
  DBuf pk_r[2], pk_k[2], pk_tmp, pk_cnt, pk_geo;   // peak candidates
  DBuf mk_geo, mk_err, vg_cnt, vg_start, vg_pairs;  // grid masks
  long long pool_want = 0;   // pairs the last count pass needed
  long long epool_want = 0;  // emission-list words the last count pass needed
  bool sized = false;        // the entry/item/scratch buffers were sized by an exact call
  // TSC binning along the hierarchy: per-tile capacities and the work items
  // planned on them by an exact count, valid for one (plan, particle count,
  // tuning); cleared when a call overflows a capacity
  bool hier_sized = false;
  bool hier_off = false;     // ASOHF_TSC_BINNING=hash at creation: TSC uses the patch-hash binning
  uint64_t hier_key = 0;
  long long hier_items[kNLevels + 2] = {};
  long long hier_nitems = 0;
  unsigned long long hier_nsplit = 0;
  hipGraphExec_t gexec = nullptr;   // the captured batched build and the arguments it holds
  PipeArgs gkey;
  DBuf h_in[4], h_out[2], h_mass;    // device copies for the host entry point
  DBuf ord[4], ord_tmp;              // adjoint gather order: keys and indices (in, out), sort scratch
  Counters* hcount = nullptr;          // pinned
  unsigned long long* hscalar = nullptr;
  hipEvent_t ev[9] = {};   // 0 start, 1 count pass, 2 scatter start, 3 scatter end, 4 deposit start,
                           // 8 deposit end, 5 split tiles summed, 6 end, 7 the caller's stream
                           // before a graph launch
  asohf_stats stats{};
  bool stats_pending = false;
  bool stats_phases = false;
  std::map<std::pair<int, int>, std::vector<float>> ktab_cache;
};

// particles from which the adjoint gathers in base-cell order
constexpr long long kAdjointSortMin = 1ll << 20;

static asohf_status fail(asohf_ctx* c, asohf_status st, const std::string& msg) {
  if (c) c->err = msg;
  return st;
}
static asohf_status hip_fail(asohf_ctx* c, hipError_t e, const char* what) {
  return fail(c, e == hipErrorOutOfMemory ? ASOHF_ERR_OOM : ASOHF_ERR_HIP,
              std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPCHK(expr, what)                         \
  do {                                             \
    hipError_t _e = (expr);                        \
    if (_e != hipSuccess) return hip_fail(ctx, _e, what); \
  } while (0)

// ---------------------------------------------------------------------------
// geometry validation and planning
// ---------------------------------------------------------------------------
static asohf_status check_geom(asohf_ctx* ctx, const asohf_geom* g, int mode) {
  if (!g) return fail(ctx, ASOHF_ERR_INVALID, "geometry is NULL");
  if (g->nx < 3 || g->ny < 3 || g->nz < 3)
    return fail(ctx, ASOHF_ERR_INVALID, "base grid needs at least 3 cells per axis");
  if (g->nx > 16000 || g->ny > 16000 || g->nz > 16000)
    return fail(ctx, ASOHF_ERR_CAPACITY, "base grid side above 16000 cells");
  if (g->nl < 0 || g->nl > kNLevels) return fail(ctx, ASOHF_ERR_CAPACITY, "NL outside 0..9 (NLEVELS)");
  if (!g->npatch) return fail(ctx, ASOHF_ERR_INVALID, "npatch is NULL");
  long long P = 0;
  for (int l = 0; l <= g->nl; ++l) {
    if (g->npatch[l] < 0) return fail(ctx, ASOHF_ERR_INVALID, "negative NPATCH entry");
    P += g->npatch[l];
  }
  if (P > (1 << 26)) return fail(ctx, ASOHF_ERR_CAPACITY, "more than 2^26 patches");
  if (P > 0 && (!g->patchnx || !g->patchny || !g->patchnz || !g->patchrx || !g->patchry || !g->patchrz))
    return fail(ctx, ASOHF_ERR_INVALID, "patch arrays are NULL");
  for (long long p = 0; p < P; ++p) {
    const int a = g->patchnx[p], b = g->patchny[p], c = g->patchnz[p];
    if (a < 1 || b < 1 || c < 1) return fail(ctx, ASOHF_ERR_INVALID, "patch with a non-positive side");
    if (a > 16000 || b > 16000 || c > 16000) return fail(ctx, ASOHF_ERR_CAPACITY, "patch side above 16000 cells");
    if (!std::isfinite(g->patchrx[p]) || !std::isfinite(g->patchry[p]) || !std::isfinite(g->patchrz[p]))
      return fail(ctx, ASOHF_ERR_INVALID, "non-finite patch origin");
  }
  if (!(g->dx > 0 && g->dy > 0 && g->dz > 0 && g->lado0 > 0))
    return fail(ctx, ASOHF_ERR_INVALID, "cell sizes and LADO0 must be positive");
  if (!(g->rote > 0 && g->rete > 0 && g->fdm > 0) || !std::isfinite(g->rote) || !std::isfinite(g->rete))
    return fail(ctx, ASOHF_ERR_INVALID, "ROTE, RETE and FDM must be positive");
  if (mode != ASOHF_MODE_TABLE && mode != ASOHF_MODE_TSC) return fail(ctx, ASOHF_ERR_INVALID, "unknown mode");
  if (mode == ASOHF_MODE_TABLE && g->interp_degree != 1 && g->interp_degree != 2)
    return fail(ctx, ASOHF_ERR_INVALID, "interp_degree must be 1 or 2");
  return ASOHF_OK;
}

static uint64_t plan_key(const asohf_geom* g, int mode, const asohf_tune& tn, const asohf_opts* o) {
  uint64_t h = 1469598103934665603ull;
  long long P = 0;
  for (int l = 0; l <= g->nl; ++l) P += g->npatch[l];
  int hdr[16] = {g->nx, g->ny, g->nz, g->nl, mode, g->interp_degree, tn.tile[0], tn.tile[1], tn.tile[2],
                 o->u1_ld[0], o->u1_ld[1], o->u11_ld[0], o->u11_ld[1], o->u11_ld[2], 0, 0};
  h = fnv(h, hdr, sizeof(hdr));
  float sc[7] = {g->dx, g->dy, g->dz, g->lado0, g->rote, g->rete, g->fdm};
  h = fnv(h, sc, sizeof(sc));
  h = fnv(h, g->npatch, sizeof(int32_t) * (g->nl + 1));
  if (P) {
    h = fnv(h, g->patchnx, 4 * P); h = fnv(h, g->patchny, 4 * P); h = fnv(h, g->patchnz, 4 * P);
    h = fnv(h, g->patchrx, 4 * P); h = fnv(h, g->patchry, 4 * P); h = fnv(h, g->patchrz, 4 * P);
  }
  const float* r[3] = {g->radx, g->rady, g->radz};
  const int n[3] = {g->nx, g->ny, g->nz};
  for (int a = 0; a < 3; ++a) {
    if (r[a]) h = fnv(h, r[a], sizeof(float) * (n[a] + 2));
    else h = fnv(h, "null", 4);
  }
  return h;
}

static asohf_status build_plan(asohf_ctx* ctx, const asohf_geom* g, int mode, const asohf_opts* o) {
  const asohf_tune& tn = ctx->tune;
  const uint64_t key = plan_key(g, mode, tn, o);
  HostPlan& H = ctx->plan;
  if (H.valid && H.key == key) return ASOHF_OK;
  H = HostPlan();
  const int nl = g->nl;
  long long P = 0;
  std::vector<int> level_start(nl + 2, 0);
  for (int l = 0; l <= nl; ++l) { level_start[l] = (int)P; P += g->npatch[l]; }
  level_start[nl + 1] = (int)P;
  const int T[3] = {tn.tile[0], tn.tile[1], tn.tile[2]};

  // cell centres: COMMON /GRID/ or MALLA with LADO = LADO0 - LADO0/NX (asohf.f:323)
  const float lado = g->lado0 - g->lado0 / (float)g->nx;
  if (g->radx) H.radx.assign(g->radx, g->radx + g->nx + 2); else malla(g->nx, lado, H.radx);
  if (g->rady) H.rady.assign(g->rady, g->rady + g->ny + 2); else malla(g->ny, lado, H.rady);
  if (g->radz) H.radz.assign(g->radz, g->radz + g->nz + 2); else malla(g->nz, lado, H.radz);
  // kernel table (table mode), cached per (nl, degree)
  {
    auto k = std::make_pair(nl, g->interp_degree);
    auto it = ctx->ktab_cache.find(k);
    if (it == ctx->ktab_cache.end()) {
      std::vector<float> K;
      compute_kernels(nl, g->interp_degree == 1 ? 1 : 2, K);
      it = ctx->ktab_cache.emplace(k, std::move(K)).first;
    }
    H.ktab = it->second;
  }
  const float r3 = g->rete * (g->rete * g->rete);   // RETE**3
  // ---- grids ----
  H.grids.resize(1 + P);
  int ntiles = 0;
  long long u11_off = 0;
  const bool padded = o->u11_ld[0] > 0;
  {
    GridDesc& B = H.grids[0];
    memset(&B, 0, sizeof(B));
    B.n[0] = g->nx; B.n[1] = g->ny; B.n[2] = g->nz;
    B.level = 0;
    for (int a = 0; a < 3; ++a) {
      B.nt[a] = (B.n[a] + T[a] - 1) / T[a];
      B.org[a] = -g->lado0 / 2.0f;
    }
    B.d[0] = g->dx; B.d[1] = g->dy; B.d[2] = g->dz;
    for (int a = 0; a < 3; ++a) {
      B.inv_d[a] = 1.0f / B.d[a];
      B.inv_dd[a] = 1.0 / (double)B.d[a];
    }
    B.den = g->dx * g->dy * g->dz * g->rote * r3;         // DENBAS (grids.f:824)
    B.tile0 = 0;
    ntiles = B.nt[0] * B.nt[1] * B.nt[2];
    const int l0 = o->u1_ld[0] > 0 ? o->u1_ld[0] : g->nx;
    const int l1 = o->u1_ld[1] > 0 ? o->u1_ld[1] : g->ny;
    B.ld1 = l0;
    B.ld2 = l0 * l1;
    B.off = 0;
    H.u1_cells = (long long)l0 * l1 * g->nz;
  }
  for (int l = 0; l <= nl; ++l) {
    const float p2 = powi_f(2.0f, l);
    const float dxpa = g->dx / p2, dypa = g->dy / p2, dzpa = g->dz / p2;
    const float den = dxpa * dypa * dzpa * g->rote * r3;   // DENBAS of the level (grids.f:844)
    for (int p = level_start[l]; p < level_start[l + 1]; ++p) {
      GridDesc& G = H.grids[1 + p];
      memset(&G, 0, sizeof(G));
      G.n[0] = g->patchnx[p]; G.n[1] = g->patchny[p]; G.n[2] = g->patchnz[p];
      G.level = l;
      G.d[0] = dxpa; G.d[1] = dypa; G.d[2] = dzpa;
      for (int a = 0; a < 3; ++a) {
        G.inv_d[a] = 1.0f / G.d[a];
        G.inv_dd[a] = 1.0 / (double)G.d[a];
      }
      G.den = den;
      const float pr[3] = {g->patchrx[p], g->patchry[p], g->patchrz[p]};
      const float d[3] = {dxpa, dypa, dzpa};
      const float dbase = g->dx;   // MAXKERNELEXT = DX on every axis (grids.f:837, 880-885)
      for (int a = 0; a < 3; ++a) {
        G.pr[a] = pr[a];
        if (mode == ASOHF_MODE_TABLE) {
          const float xlfix = pr[a] - d[a];                        // grids.f:871-873
          const float xr = xlfix + (float)G.n[a] * d[a];           // grids.f:878-880
          G.org[a] = xlfix;
          G.lo[a] = xlfix - dbase;                                 // grids.f:881-886
          G.hi[a] = xr + dbase;
        } else {
          const float xl = pr[a] - 2.0f * d[a];                    // (1+BUF)*DXPA, grids.f:1294
          G.org[a] = xl;
          G.lo[a] = xl;
          G.hi[a] = xl + (float)(G.n[a] + 2) * d[a];               // (N1+2*BUF)*DXPA, grids.f:1300
        }
      }
      if (padded) {
        G.ld1 = o->u11_ld[0];
        G.ld2 = o->u11_ld[0] * o->u11_ld[1];
        G.off = (long long)p * o->u11_ld[0] * o->u11_ld[1] * o->u11_ld[2];
      } else {
        G.ld1 = G.n[0];
        G.ld2 = G.n[0] * G.n[1];
        G.off = u11_off;
      }
      u11_off += (long long)G.n[0] * G.n[1] * G.n[2];
      H.ncells += (long long)G.n[0] * G.n[1] * G.n[2];
      if (l == 0) {   // patches numbered before level 1 are never deposited (grids.f:846)
        G.tile0 = ntiles;
        continue;
      }
      for (int a = 0; a < 3; ++a) G.nt[a] = (G.n[a] + T[a] - 1) / T[a];
      G.tile0 = ntiles;
      const long long nt = (long long)G.nt[0] * G.nt[1] * G.nt[2];
      if ((long long)ntiles + nt > (1ll << 30)) return fail(ctx, ASOHF_ERR_CAPACITY, "more than 2^30 tiles");
      ntiles += (int)nt;
    }
  }
  // TSC weights are taken from the reference's three FP32 centres of the
  // stencil, each rounded on its own; their spacing differs from the cell
  // size by up to two ulps of the coordinate.  Up to a quarter of a cell every
  // weight stays non-negative and every non-centre product below 1/2 (the
  // exact deposit's term bounds); beyond it the grid is too fine for FP32
  // coordinates at its position and the build refuses it.
  if (mode == ASOHF_MODE_TSC) {
    for (size_t gi = 0; gi < H.grids.size(); ++gi) {
      const GridDesc& G = H.grids[gi];
      if (gi > 0 && G.level == 0) continue;
      for (int a = 0; a < 3; ++a) {
        const double lo = gi == 0 ? -0.5 * g->lado0 - G.d[a] : G.lo[a] - G.d[a];
        const double hi = gi == 0 ? 0.5 * g->lado0 + G.d[a] : G.hi[a] + G.d[a];
        const double mx = std::max(std::fabs(lo), std::fabs(hi));
        const double ulp = mx > 0 ? std::ldexp(1.0, std::ilogb(mx) - 23) : 0.0;
        if (2.0 * ulp > 0.25 * (double)G.d[a])
          return fail(ctx, ASOHF_ERR_CAPACITY,
                      "a grid's cells are too small for FP32 coordinates at its position (TSC centres)");
      }
    }
  }
  H.u11_cells = padded ? P * (long long)o->u11_ld[0] * o->u11_ld[1] * o->u11_ld[2] : u11_off;
  if (padded) {
    for (long long p = 0; p < P; ++p)
      if (g->patchnx[p] > o->u11_ld[0] || g->patchny[p] > o->u11_ld[1] || g->patchnz[p] > o->u11_ld[2])
        return fail(ctx, ASOHF_ERR_CAPACITY, "patch larger than NAMRX/NAMRY/NAMRZ of the padded layout");
  }
  if (o->u1_ld[0] > 0 && (o->u1_ld[0] < g->nx || o->u1_ld[1] < g->ny))
    return fail(ctx, ASOHF_ERR_CAPACITY, "NX/NY above the leading dimensions of U1");
  H.ncells += (long long)g->nx * g->ny * g->nz;
  H.tile_grid.resize(ntiles);
  for (size_t gi = 0; gi < H.grids.size(); ++gi) {
    const GridDesc& G = H.grids[gi];
    if (gi > 0 && G.level == 0) continue;
    const int nt = G.nt[0] * G.nt[1] * G.nt[2];
    for (int k = 0; k < nt; ++k) H.tile_grid[G.tile0 + k] = (int)gi;
  }
  // ---- per-level spatial hash of the membership boxes ----
  Plan& D = H.dev;
  memset(&D, 0, sizeof(D));
  const float x0 = -g->lado0 / 2.0f;
  std::vector<std::pair<uint64_t, int>> ent;
  for (int l = 1; l <= nl; ++l) {
    LevelHash& lh = D.lh[l];
    const int lo = level_start[l], hi = level_start[l + 1];
    if (hi == lo) { lh.x0 = x0; lh.inv_bs = 1.0f / g->lado0; lh.nb = 1; continue; }
    std::vector<float> w;
    w.reserve(hi - lo);
    for (int p = lo; p < hi; ++p) {
      const GridDesc& G = H.grids[1 + p];
      w.push_back(std::max({G.hi[0] - G.lo[0], G.hi[1] - G.lo[1], G.hi[2] - G.lo[2]}));
    }
    std::nth_element(w.begin(), w.begin() + w.size() / 2, w.end());
    const double bs = std::max<double>(w[w.size() / 2], g->lado0 / 16384.0);
    int nb = (int)std::ceil(g->lado0 / bs);
    nb = std::min(std::max(nb, 1), 16384);
    lh.x0 = x0;
    lh.inv_bs = (float)((double)nb / (double)g->lado0);
    lh.nb = nb;
    for (int p = lo; p < hi; ++p) {
      const GridDesc& G = H.grids[1 + p];
      int b0[3], b1[3];
      for (int a = 0; a < 3; ++a) {
        double t0 = ((double)G.lo[a] - (double)lh.x0) * (double)lh.inv_bs - 0.01;
        double t1 = ((double)G.hi[a] - (double)lh.x0) * (double)lh.inv_bs + 0.01;
        t0 = std::min(std::max(t0, -1.0), 1.0e6);
        t1 = std::min(std::max(t1, -1.0), 1.0e6);
        b0[a] = std::min(std::max((int)std::floor(t0), 0), nb - 1);
        b1[a] = std::min(std::max((int)std::floor(t1), 0), nb - 1);
      }
      for (int c = b0[2]; c <= b1[2]; ++c)
        for (int b = b0[1]; b <= b1[1]; ++b)
          for (int a = b0[0]; a <= b1[0]; ++a) ent.emplace_back(bucket_key(l, a, b, c), 1 + p);
    }
  }
  std::sort(ent.begin(), ent.end());
  std::vector<uint64_t> ukeys;
  std::vector<int2> uvals;
  H.hlist.resize(ent.size());
  for (size_t i = 0; i < ent.size(); ++i) {
    H.hlist[i] = ent[i].second;
    if (i == 0 || ent[i].first != ent[i - 1].first) {
      ukeys.push_back(ent[i].first);
      uvals.push_back(make_int2((int)i, 0));
    }
    uvals.back().y++;
  }
  size_t cap = 16;
  while (cap < 2 * ukeys.size() + 2) cap <<= 1;
  H.hslots.assign(cap, HashSlot{kHashEmpty, 0, 0});
  for (size_t i = 0; i < ukeys.size(); ++i) {
    const uint64_t k = ukeys[i];
    unsigned s = bucket_hash((int)(k >> 60), (int)((k >> 40) & 0xfffff), (int)((k >> 20) & 0xfffff),
                             (int)(k & 0xfffff)) & (unsigned)(cap - 1);
    while (H.hslots[s].key != kHashEmpty) s = (s + 1) & (unsigned)(cap - 1);
    H.hslots[s] = HashSlot{ukeys[i], uvals[i].x, uvals[i].y};
  }
  // nesting: does every level-l membership box lie inside some level-(l-1)
  // box (FP32 values, as the device compares them)?  Then a particle outside
  // every level-l box cannot be in a deeper one.
  D.nested = 1;
  for (int l = 2; l <= nl && D.nested; ++l) {
    const LevelHash& lh = D.lh[l - 1];
    for (int p = level_start[l]; p < level_start[l + 1] && D.nested; ++p) {
      const GridDesc& Cg = H.grids[1 + p];
      int b[3];
      for (int a = 0; a < 3; ++a) {
        const double mid = 0.5 * ((double)Cg.lo[a] + (double)Cg.hi[a]);
        b[a] = std::min(std::max((int)std::floor((mid - lh.x0) * (double)lh.inv_bs), 0), lh.nb - 1);
      }
      const uint64_t key = bucket_key(l - 1, b[0], b[1], b[2]);
      bool inside = false;
      auto itk = std::lower_bound(ukeys.begin(), ukeys.end(), key);
      if (itk != ukeys.end() && *itk == key) {
        const int2 se = uvals[itk - ukeys.begin()];
        for (int q = 0; q < se.y && !inside; ++q) {
          const GridDesc& Q = H.grids[H.hlist[se.x + q]];
          inside = Q.lo[0] <= Cg.lo[0] && Cg.hi[0] <= Q.hi[0] && Q.lo[1] <= Cg.lo[1] &&
                   Cg.hi[1] <= Q.hi[1] && Q.lo[2] <= Cg.lo[2] && Cg.hi[2] <= Q.hi[2];
        }
      }
      if (!inside) D.nested = 0;
    }
  }
  if (H.hlist.empty()) H.hlist.push_back(0);
  auto make_rec = [&](int gi) {
    const GridDesc& G = H.grids[gi];
    BoxRec r;
    memset(&r, 0, sizeof(r));
    for (int a = 0; a < 3; ++a) { r.lo[a] = G.lo[a]; r.hi[a] = G.hi[a]; }
    r.g = gi;
    r.level = G.level;
    for (int a = 0; a < 3; ++a) {
      r.org[a] = G.org[a];
      r.d[a] = G.d[a];
      r.inv_d[a] = G.inv_d[a];
      r.n[a] = G.n[a];
    }
    r.tile0 = G.tile0;
    r.nt0 = G.nt[0];
    r.nt1 = G.nt[1];
    r.den = G.den;
    for (int a = 0; a < 3; ++a) r.pr[a] = G.pr[a];
    r.ld1 = G.ld1;
    r.ld2 = G.ld2;
    r.pad = 0;
    r.off = G.off;
    return r;
  };
  H.hrec.resize(H.hlist.size());
  for (size_t i = 0; i < H.hlist.size(); ++i) H.hrec[i] = make_rec(H.hlist[i]);
  // ---- first tile of every level (grids, and so tiles, are numbered level
  // by level; patches listed at level 0 have no tiles) ----
  H.level_tile0.assign(nl + 2, ntiles);
  for (int l = nl; l >= 1; --l)
    H.level_tile0[l] = g->npatch[l] > 0 ? H.grids[1 + level_start[l]].tile0 : H.level_tile0[l + 1];
  H.level_tile0[0] = 0;
  // ---- TSC binning along the hierarchy: every level-l patch (l >= 1) gets an
  // emitter, the deepest grid found whose membership box holds the patch's box
  // in FP32 (so every member of the patch is a member of the emitter): a
  // patch of the level above whose box holds it, else one two levels up, ...,
  // else the base grid, which holds every particle.  The patch is then a
  // target of every emitter tile whose home region (the positions whose
  // clamped cell lies in the tile) can meet its box. ----
  if (mode == ASOHF_MODE_TSC) {
    H.grec.resize(H.grids.size());
    for (size_t gi = 0; gi < H.grids.size(); ++gi) H.grec[gi] = make_rec((int)gi);
    auto contains = [&](const GridDesc& Q, const GridDesc& Cg) {
      return Q.lo[0] <= Cg.lo[0] && Cg.hi[0] <= Q.hi[0] && Q.lo[1] <= Cg.lo[1] && Cg.hi[1] <= Q.hi[1] &&
             Q.lo[2] <= Cg.lo[2] && Cg.hi[2] <= Q.hi[2];
    };
    auto find_container = [&](int lq, const GridDesc& Cg) -> int {
      const LevelHash& lh = D.lh[lq];
      int b[3];
      for (int a = 0; a < 3; ++a) {
        const double mid = 0.5 * ((double)Cg.lo[a] + (double)Cg.hi[a]);
        double t = (mid - lh.x0) * (double)lh.inv_bs;
        t = std::min(std::max(t, -1.0), 1.0e6);
        b[a] = std::min(std::max((int)std::floor(t), 0), lh.nb - 1);
      }
      const uint64_t key = bucket_key(lq, b[0], b[1], b[2]);
      auto itk = std::lower_bound(ukeys.begin(), ukeys.end(), key);
      if (itk == ukeys.end() || *itk != key) return -1;
      const int2 se = uvals[itk - ukeys.begin()];
      for (int q = 0; q < se.y; ++q) {
        const int gq = H.hlist[se.x + q];
        if (contains(H.grids[gq], Cg)) return gq;
      }
      return -1;
    };
    std::vector<std::pair<int, int>> tp;   // (emitter tile, target grid)
    for (int l = 1; l <= nl; ++l) {
      for (int p = level_start[l]; p < level_start[l + 1]; ++p) {
        const int gc = 1 + p;
        const GridDesc& Cg = H.grids[gc];
        int ge = 0;
        for (int lq = l - 1; lq >= 1 && ge == 0; --lq) {
          const int f = find_container(lq, Cg);
          if (f > 0) ge = f;
        }
        const GridDesc& E = H.grids[ge];
        H.emit_levels |= 1u << E.level;
        int t0[3], t1[3];
        for (int a = 0; a < 3; ++a) {
          // cells of E whose home region can meet (lo, hi), one cell of margin
          const double u0 = ((double)Cg.lo[a] - (double)E.org[a]) / (double)E.d[a];
          const double u1 = ((double)Cg.hi[a] - (double)E.org[a]) / (double)E.d[a];
          const double c0 = std::floor(std::min(std::max(u0, -4.0), (double)E.n[a] + 4.0)) + (ge == 0 ? 0.0 : -1.0);
          const double c1 = std::floor(std::min(std::max(u1, -4.0), (double)E.n[a] + 4.0)) + (ge == 0 ? 2.0 : 1.0);
          const int i0 = std::min(std::max((int)c0, 1), E.n[a]);
          const int i1 = std::min(std::max((int)c1, 1), E.n[a]);
          t0[a] = (i0 - 1) / T[a];
          t1[a] = (i1 - 1) / T[a];
        }
        for (int c = t0[2]; c <= t1[2]; ++c)
          for (int b = t0[1]; b <= t1[1]; ++b)
            for (int a = t0[0]; a <= t1[0]; ++a)
              tp.emplace_back(E.tile0 + (c * E.nt[1] + b) * E.nt[0] + a, gc);
      }
    }
    H.tgt_off.assign(ntiles + 1, 0u);
    for (const auto& e : tp) H.tgt_off[e.first + 1]++;
    for (int t = 0; t < ntiles; ++t) H.tgt_off[t + 1] += H.tgt_off[t];
    H.tgt.assign(std::max<size_t>(tp.size(), 1), 0);
    {
      std::vector<unsigned> fill(H.tgt_off.begin(), H.tgt_off.end() - 1);
      for (const auto& e : tp) H.tgt[fill[e.first]++] = e.second;
    }
  } else {
    H.tgt_off.assign(ntiles + 1, 0u);
    H.tgt.assign(1, 0);
    H.grec.assign(1, BoxRec{});
  }
  // ---- upload ----
  size_t off = 0;
  auto place = [&](size_t bytes) { size_t o2 = off; off = align_up(off + bytes, 256); return o2; };
  const size_t o_grids = place(sizeof(GridDesc) * H.grids.size());
  const size_t o_tg = place(sizeof(int) * std::max<size_t>(H.tile_grid.size(), 1));
  const size_t o_hs = place(sizeof(HashSlot) * H.hslots.size());
  const size_t o_hr = place(sizeof(BoxRec) * H.hrec.size());
  const size_t o_rx = place(sizeof(float) * H.radx.size());
  const size_t o_ry = place(sizeof(float) * H.rady.size());
  const size_t o_rz = place(sizeof(float) * H.radz.size());
  const size_t o_kt = place(sizeof(float) * H.ktab.size());
  const size_t o_to = place(sizeof(unsigned) * H.tgt_off.size());
  const size_t o_tt = place(sizeof(int) * H.tgt.size());
  const size_t o_gr = place(sizeof(BoxRec) * H.grec.size());
  HIPCHK(ctx->plan_blob.ensure(off), "allocating the geometry plan");
  char* base = ctx->plan_blob.as<char>();
  auto up = [&](size_t o2, const void* src, size_t bytes) {
    return hipMemcpyAsync(base + o2, src, bytes, hipMemcpyHostToDevice, ctx->stream);
  };
  HIPCHK(up(o_grids, H.grids.data(), sizeof(GridDesc) * H.grids.size()), "uploading grids");
  if (!H.tile_grid.empty()) HIPCHK(up(o_tg, H.tile_grid.data(), sizeof(int) * H.tile_grid.size()), "uploading tiles");
  HIPCHK(up(o_hs, H.hslots.data(), sizeof(HashSlot) * H.hslots.size()), "uploading hash");
  HIPCHK(up(o_hr, H.hrec.data(), sizeof(BoxRec) * H.hrec.size()), "uploading hash");
  HIPCHK(up(o_rx, H.radx.data(), sizeof(float) * H.radx.size()), "uploading RADX");
  HIPCHK(up(o_ry, H.rady.data(), sizeof(float) * H.rady.size()), "uploading RADY");
  HIPCHK(up(o_rz, H.radz.data(), sizeof(float) * H.radz.size()), "uploading RADZ");
  HIPCHK(up(o_kt, H.ktab.data(), sizeof(float) * H.ktab.size()), "uploading kernel table");
  HIPCHK(up(o_to, H.tgt_off.data(), sizeof(unsigned) * H.tgt_off.size()), "uploading targets");
  HIPCHK(up(o_tt, H.tgt.data(), sizeof(int) * H.tgt.size()), "uploading targets");
  HIPCHK(up(o_gr, H.grec.data(), sizeof(BoxRec) * H.grec.size()), "uploading grid records");
  HIPCHK(hipStreamSynchronize(ctx->stream), "uploading the plan");
  D.grids = (const GridDesc*)(base + o_grids);
  D.tile_grid = (const int*)(base + o_tg);
  D.hslots = (const HashSlot*)(base + o_hs);
  D.hrec = (const BoxRec*)(base + o_hr);
  D.radx = (const float*)(base + o_rx);
  D.rady = (const float*)(base + o_ry);
  D.radz = (const float*)(base + o_rz);
  D.ktab = (const float*)(base + o_kt);
  D.tgt_off = (const unsigned*)(base + o_to);
  D.tgt = (const int*)(base + o_tt);
  D.grec = (const BoxRec*)(base + o_gr);
  D.nl = nl;
  D.ngrids = (int)H.grids.size();
  D.ntiles = ntiles;
  D.kmax = 1 << nl;
  D.kw = 2 * D.kmax + 1;
  D.hmask = (unsigned)(cap - 1);
  for (int a = 0; a < 3; ++a) { D.tile[a] = T[a]; D.tshift[a] = __builtin_ctz(T[a]); }
  D.mode = mode;
  // lane merging of narrow deposits sums up to four non-centre terms in 32
  // bits: TSC non-centre weight products stay below 1/2 (the centre check
  // above); the table's row 0 (centred at kmax) is checked here
  {
    const float* k0 = H.ktab.data() + D.kmax - 1;
    float mx = 0.f;
    bool nonneg = k0[0] >= 0.f && k0[1] >= 0.f && k0[2] >= 0.f;
    for (int c = 0; c < 3; ++c)
      for (int b = 0; b < 3; ++b)
        for (int a = 0; a < 3; ++a)
          if (!(a == 1 && b == 1 && c == 1)) mx = std::max(mx, k0[a] * k0[b] * k0[c]);
    D.merge_narrow = mode == 2 || (nonneg && mx <= 0.45f) ? 1 : 0;
  }
  // are the centre tables MALLA's recurrence?  (always, unless the caller
  // passed its own COMMON /GRID/ values that differ)
  D.rad_malla = 1;
  {
    const std::vector<float>* rv[3] = {&H.radx, &H.rady, &H.radz};
    const int nn[3] = {g->nx, g->ny, g->nz};
    for (int a = 0; a < 3; ++a) {
      const std::vector<float>& r = *rv[a];
      const int n = nn[a];
      const float r1 = r[1];
      const float c = r[2] - r[1] == 0.f ? 0.f : (r[n] - r[1]) / (float)(n - 1);
      // recover MALLA's C exactly: C = (B-A)/(N-1) with A = -LADO/2, B = LADO/2
      const float cm = (lado / 2.0f - (-lado / 2.0f)) / (float)(n - 1);
      (void)c;
      D.rad1[a] = r1;
      D.radc[a] = cm;
      bool same = true;
      for (int i = 0; i <= n && same; ++i) same = r[i] == r1 + (float)(i - 1) * cm;
      same = same && r[n + 1] == (r1 + (float)(n - 1) * cm) + cm;
      if (!same) D.rad_malla = 0;
    }
  }
  D.xl = -g->lado0 / 2.0f;
  D.fdm = g->fdm;
  H.key = key;
  H.valid = true;
  return ASOHF_OK;
}

static asohf_opts default_opts() {
  asohf_opts o;
  memset(&o, 0, sizeof(o));
  o.mode = ASOHF_MODE_TSC;
  o.reduction = ASOHF_REDUCE_DETERMINISTIC;
  o.path = ASOHF_PATH_BATCHED;
  return o;
}

static asohf_status make_parts(asohf_ctx* ctx, const asohf_opts& o, int64_t n, const float* rx,
                               const float* ry, const float* rz, const float* m, const int32_t* esp,
                               Parts& A) {
  if (n < 0) return fail(ctx, ASOHF_ERR_INVALID, "negative particle count");
  if (n > 0 && (!rx || !ry || !rz || !m)) return fail(ctx, ASOHF_ERR_INVALID, "particle arrays are NULL");
  memset(&A, 0, sizeof(A));
  A.x = rx; A.y = ry; A.z = rz; A.m = m;
  A.n = n;
  if (o.mode == ASOHF_MODE_TABLE) {
    if (!esp) return fail(ctx, ASOHF_ERR_INVALID, "npart_esp is NULL (table mode)");
    long long s = 0;
    A.sp[0] = 0;
    for (int k = 0; k < kNEsp; ++k) {
      if (esp[k] < 0) return fail(ctx, ASOHF_ERR_INVALID, "negative NPART_ESP entry");
      s += esp[k];
      A.sp[k + 1] = s;
    }
    if (s > n) return fail(ctx, ASOHF_ERR_INVALID, "SUM(NPART_ESP) exceeds the particle count");
    A.n = s;   // DENSITY deposits particles 1..SUM(NPART_ESP) (grids.f:784-785)
  }
  return ASOHF_OK;
}

static asohf_status decode_err(asohf_ctx* ctx, unsigned e) {
  if (!e) return ASOHF_OK;
  std::string m = "invalid particle data:";
  if (e & kErrNonFinite) m += " non-finite position or mass;";
  if (e & kErrNegMass) m += " negative mass;";
  if (e & kErrWrap)
    m += " base-grid position outside the range the mode accepts (kernel-table mode: cell index outside "
         "the single periodic wrap of DENSITY; TSC mode: more than half a cell past the first or last "
         "cell centre);";
  return fail(ctx, ASOHF_ERR_INVALID, m);
}

// ---------------------------------------------------------------------------
// batched pipeline
// ---------------------------------------------------------------------------
// Workspaces whose size depends only on the geometry and the particle count,
// and the arguments of every kernel of the call.  The entry, item and scratch
// buffers keep the size of the largest earlier call; the gate kernel checks
// them on the device.
static asohf_status prepare_args(asohf_ctx* ctx, const asohf_opts& o, const Parts& A, float* u1, float* u11,
                                 double* grid_mass, PipeArgs& a, bool want_elist = true) {
  HostPlan& H = ctx->plan;
  memset(&a, 0, sizeof(a));
  a.P = H.dev;
  a.A = A;
  a.ntiles = a.P.ntiles;
  a.nkeys = 2ll * a.ntiles;
  a.red = o.reduction;
  a.elem = a.red == 0 ? 8 : 4;
  // the per-block tables serve both binning forms: sized for the larger
  a.nblk = std::max(bin_blocks(A.n), hier_bin_blocks(A.n));
  const long long nkeys = a.nkeys, nblk = a.nblk;
  const int ntiles = a.ntiles;
  HIPCHK(ctx->cnt.ensure(sizeof(unsigned) * std::max<long long>(nkeys, 1)), "allocating counts");
  HIPCHK(ctx->start.ensure(sizeof(unsigned long long) * (nkeys + 1)), "allocating offsets");
  HIPCHK(ctx->cursor.ensure(sizeof(unsigned long long) * std::max<long long>(nkeys, 1)), "allocating cursors");
  HIPCHK(ctx->bsum.ensure(sizeof(unsigned long long) * scan_scratch_elems(nkeys)), "allocating scan");
  HIPCHK(ctx->counters.ensure(sizeof(Counters)), "allocating counters");
  HIPCHK(ctx->stiles.ensure(sizeof(int) * std::max(ntiles, 1)), "allocating tile state");
  HIPCHK(ctx->isplit.ensure(sizeof(unsigned) * std::max(ntiles, 1)), "allocating tile state");
  HIPCHK(ctx->srank.ensure(sizeof(unsigned long long) * (ntiles + 1)), "allocating tile state");
  HIPCHK(ctx->nitems.ensure(sizeof(unsigned) * std::max(ntiles, 1)), "allocating tile state");
  HIPCHK(ctx->nsplit.ensure(sizeof(unsigned) * std::max(ntiles, 1)), "allocating tile state");
  HIPCHK(ctx->item_off.ensure(sizeof(unsigned long long) * (ntiles + 1)), "allocating tile state");
  HIPCHK(ctx->split_off.ensure(sizeof(unsigned long long) * (ntiles + 1)), "allocating tile state");
  HIPCHK(ctx->split.ensure(sizeof(int) * std::max(ntiles, 1)), "allocating tile state");
  HIPCHK(ctx->tmass.ensure(sizeof(double) * std::max(ntiles, 1)), "allocating tile sums");
  // saved per-block tables: sized from the previous call (a block whose table
  // does not fit its sub-pool recounts in the scatter pass)
  const long long pool_cap = std::max<long long>(ctx->pool_want + ctx->pool_want / 8, 256 * nblk);
  HIPCHK(ctx->pool.ensure(sizeof(int2) * pool_cap), "allocating block tables");
  HIPCHK(ctx->pool_off.ensure(sizeof(long long) * std::max<long long>(nblk, 1)), "allocating block tables");
  HIPCHK(ctx->pool_nk.ensure(sizeof(int) * std::max<long long>(nblk, 1)), "allocating block tables");
  HIPCHK(ctx->blk_pairs.ensure(sizeof(unsigned long long) * std::max<long long>(nblk, 1)), "allocating block sums");
  HIPCHK(ctx->blk_mmax.ensure(sizeof(unsigned) * std::max<long long>(nblk, 1)), "allocating block sums");
  // emission lists: about four per particle until a call has measured them
  a.BT.use_elist = want_elist && 2ll * ntiles < (1ll << 23) ? 1 : 0;
  if (a.BT.use_elist) {
    const long long ecap = std::max<long long>(ctx->epool_want + ctx->epool_want / 8, 4 * A.n + 4096);
    HIPCHK(ctx->elist.ensure(sizeof(unsigned) * ecap), "allocating emission lists");
    HIPCHK(ctx->eoff.ensure(sizeof(long long) * std::max<long long>(nblk, 1)), "allocating emission lists");
    HIPCHK(ctx->ecount.ensure(sizeof(int) * std::max<long long>(nblk, 1)), "allocating emission lists");
    a.BT.elist = ctx->elist.as<unsigned>();
    a.BT.ecap = (long long)(ctx->elist.cap / sizeof(unsigned));
    a.BT.eoff = ctx->eoff.as<long long>();
    a.BT.ecount = ctx->ecount.as<int>();
  }
  a.BT.pool = ctx->pool.as<int2>();
  a.BT.cap = (long long)(ctx->pool.cap / sizeof(int2));
  a.BT.off = ctx->pool_off.as<long long>();
  a.BT.nkeys = ctx->pool_nk.as<int>();
  a.BT.pairs = ctx->blk_pairs.as<unsigned long long>();
  a.BT.mmax = ctx->blk_mmax.as<unsigned>();
  a.C = ctx->counters.as<Counters>();
  a.cnt = ctx->cnt.as<unsigned>();
  a.start = ctx->start.as<unsigned long long>();
  a.cursor = ctx->cursor.as<unsigned long long>();
  a.bsum = ctx->bsum.as<unsigned long long>();
  a.nitems = ctx->nitems.as<unsigned>();
  a.nsplit = ctx->nsplit.as<unsigned>();
  a.item_off = ctx->item_off.as<unsigned long long>();
  a.split_off = ctx->split_off.as<unsigned long long>();
  a.split = ctx->split.as<int>();
  a.is_split = ctx->isplit.as<unsigned>();
  a.split_rank = ctx->srank.as<unsigned long long>();
  a.split_tiles = ctx->stiles.as<int>();
  a.grid_mass = grid_mass;
  a.chunk = std::max(ctx->tune.chunk, 64);
  // wide table-mode entries: a tile's classes are deposited in the
  // histogram form (a fixed cost per tile and class) when one work item
  // holds all of the tile's wide entries, otherwise (and for the widest
  // classes) entry by entry; items of up to 8 chunks of entries
  a.chunk_wide = a.chunk * 8;
  a.cus = ctx->cus;
  const int waves = std::max(ctx->tune.waves_per_cu, 1);
  a.block = ctx->tune.block;
  a.grid = std::max(1, ctx->cus * waves / std::max(a.block / 64, 1));
  a.D.nitems_tile = ctx->nitems.as<int>();
  a.D.split_slot = ctx->split.as<int>();
  a.D.split_tiles = ctx->stiles.as<int>();
  a.D.n_split_tiles = ctx->srank.as<unsigned long long>() + ntiles;
  a.D.C = a.C;
  a.D.u1 = u1;
  a.D.u11 = u11;
  a.D.tile_mass_i = ctx->tmass.as<long long>();
  a.D.tile_mass_d = ctx->tmass.as<double>();
  a.D.n_queues = std::min(kQueues, a.grid);
  return ASOHF_OK;
}

// the buffers sized by the planning totals, as they are now
static void bind_sized(asohf_ctx* ctx, PipeArgs& a) {
  const int T = a.P.tile[0] * a.P.tile[1] * a.P.tile[2];
  a.D.entries = ctx->entries.as<int4>();
  a.D.items = ctx->items.as<Item>();
  a.D.scratch = ctx->scratch.p;
  a.cap_entries = ctx->entries.cap / sizeof(int4);
  a.cap_items = ctx->items.cap / sizeof(Item);
  a.cap_split = ctx->scratch.cap / ((size_t)T * a.elem);
}

// count pass: per-(tile, kind) counts, block sums, offsets and the item plan
static asohf_status enqueue_count(asohf_ctx* ctx, const PipeArgs& a, hipStream_t s) {
  HIPCHK(launch_fill_u32((unsigned*)a.C, 0u, sizeof(Counters) / 4, s), "clearing counters");
  HIPCHK(launch_fill_u32(a.cnt, 0u, a.nkeys, s), "clearing counts");
  HIPCHK(launch_count(a.P, a.A, a.cnt, a.C, a.BT, a.cus, s), "count kernel");
  if (a.A.n > 0) HIPCHK(launch_reduce_blocks(a.BT, bin_blocks(a.A.n), a.C, a.cus, s), "block sums");
  HIPCHK(launch_scan(a.cnt, a.nkeys, a.bsum, a.start, a.cursor, s), "scan");
  HIPCHK(launch_plan_count(a.start, a.ntiles, a.chunk, a.chunk_wide, a.nitems, a.nsplit, a.is_split, s),
         "planning items");
  HIPCHK(launch_scan(a.nitems, a.ntiles, a.bsum, a.item_off, nullptr, s), "item offsets");
  HIPCHK(launch_scan(a.nsplit, a.ntiles, a.bsum, a.split_off, nullptr, s), "split offsets");
  HIPCHK(launch_scan(a.is_split, a.ntiles, a.bsum, a.split_rank, nullptr, s), "split tiles");
  return ASOHF_OK;
}

// gate, scatter, item plan, persistent deposit, grid sums; events between the
// phases unless the sequence is being captured
static asohf_status enqueue_deposit(asohf_ctx* ctx, const PipeArgs& a, hipStream_t s, bool events) {
  HIPCHK(launch_gate(a.start + a.nkeys, a.item_off + a.ntiles, a.split_off + a.ntiles, a.cap_entries,
                     a.cap_items, a.cap_split, a.C, s), "gate");
  if (events) HIPCHK(hipEventRecord(ctx->ev[2], s), "event");
  HIPCHK(launch_scatter(a.P, a.A, a.cursor, const_cast<int4*>(a.D.entries), a.cap_entries, a.BT, a.C, a.cus, s), "scatter kernel");
  if (events) HIPCHK(hipEventRecord(ctx->ev[3], s), "event");
  HIPCHK(launch_plan_write(a.start, a.ntiles, a.chunk, a.chunk_wide, a.item_off, a.split_off, a.split_rank,
                           const_cast<Item*>(a.D.items), a.split, a.split_tiles, a.D.tile_mass_i, a.C, s),
         "writing items");
  HIPCHK(launch_zero_split(a.P, a.D, a.red, a.cus, s), "clearing split tiles");
  if (events) HIPCHK(hipEventRecord(ctx->ev[4], s), "event");
  HIPCHK(launch_deposit(a.P, a.D, a.red, a.grid, a.block, s), "deposit kernel");
  if (events) HIPCHK(hipEventRecord(ctx->ev[8], s), "event");
  HIPCHK(launch_split_reduce(a.P, a.D, a.red, a.cus, s), "split tiles");
  if (events) HIPCHK(hipEventRecord(ctx->ev[5], s), "event");
  if (a.grid_mass) HIPCHK(launch_grid_sums(a.P, a.D, a.red, a.grid_mass, s), "grid sums");
  return ASOHF_OK;
}

static void set_stats(asohf_ctx* ctx, const Parts& A, const Counters& hc, bool phases) {
  ctx->stats = asohf_stats{};
  ctx->stats.n_particles = A.n;
  ctx->stats.n_pairs = (int64_t)hc.n_pairs;
  ctx->stats.n_entries = (int64_t)hc.n_entries;
  ctx->stats.n_tiles = ctx->plan.dev.ntiles;
  ctx->stats.n_items = (int64_t)hc.n_items;
  ctx->stats.n_cells = ctx->plan.ncells;
  ctx->stats_pending = true;
  ctx->stats_phases = phases;
}

static void note_pool(asohf_ctx* ctx, const Counters& hc) {
  unsigned long long sub_max = 0;
  for (int k = 0; k < kPoolSubs; ++k) sub_max = std::max(sub_max, hc.pool_fill[k * (kLine / 2)]);
  ctx->pool_want = (long long)(sub_max * kPoolSubs);
  unsigned long long esub = 0;
  for (int k = 0; k < kPoolSubs; ++k) esub = std::max(esub, hc.epool_fill[k * (kLine / 2)]);
  ctx->epool_want = std::max(ctx->epool_want, (long long)(esub * kPoolSubs));
}

// Exact form: the count pass, one host read of its totals, buffers grown to
// fit, then the rest.  Used for the first call on new sizes and whenever the
// buffers of an earlier call turn out too small.
static asohf_status run_exact(asohf_ctx* ctx, const asohf_opts& o, const Parts& A, float* u1, float* u11,
                              double* grid_mass, hipStream_t s) {
  PipeArgs a;
  asohf_status st = prepare_args(ctx, o, A, u1, u11, grid_mass, a);
  if (st != ASOHF_OK) return st;
  HIPCHK(hipEventRecord(ctx->ev[0], s), "event");
  st = enqueue_count(ctx, a, s);
  if (st != ASOHF_OK) return st;
  HIPCHK(hipEventRecord(ctx->ev[1], s), "event");
  HIPCHK(hipMemcpyAsync(ctx->hcount, a.C, sizeof(Counters), hipMemcpyDeviceToHost, s), "reading counters");
  HIPCHK(hipMemcpyAsync(ctx->hscalar, a.start + a.nkeys, sizeof(unsigned long long), hipMemcpyDeviceToHost, s),
         "reading totals");
  HIPCHK(hipMemcpyAsync(ctx->hscalar + 1, a.item_off + a.ntiles, sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, s), "reading totals");
  HIPCHK(hipMemcpyAsync(ctx->hscalar + 2, a.split_off + a.ntiles, sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, s), "reading totals");
  HIPCHK(hipStreamSynchronize(s), "count pass");
  Counters hc = *ctx->hcount;
  hc.n_entries = ctx->hscalar[0];
  hc.n_items = ctx->hscalar[1];
  hc.n_split = ctx->hscalar[2];
  st = decode_err(ctx, hc.err);
  if (st != ASOHF_OK) return st;
  if (hc.n_items >= (1ull << 31)) return fail(ctx, ASOHF_ERR_CAPACITY, "more than 2^31 work items");
  note_pool(ctx, hc);
  const int T = a.P.tile[0] * a.P.tile[1] * a.P.tile[2];
  HIPCHK(ctx->entries.ensure(sizeof(int4) * std::max<unsigned long long>(hc.n_entries, 1)), "allocating entries");
  HIPCHK(ctx->items.ensure(sizeof(Item) * std::max<unsigned long long>(hc.n_items, 1)), "allocating work items");
  HIPCHK(ctx->scratch.ensure((size_t)std::max<unsigned long long>(hc.n_split, 1) * T * a.elem),
         "allocating split-tile scratch");
  bind_sized(ctx, a);
  a.grid = (int)std::min<long long>(a.grid, std::max<unsigned long long>(hc.n_items, 1));
  a.D.n_queues = std::min(kQueues, a.grid);
  st = enqueue_deposit(ctx, a, s, true);
  if (st != ASOHF_OK) return st;
  HIPCHK(hipEventRecord(ctx->ev[6], s), "event");
  set_stats(ctx, A, hc, true);
  ctx->sized = true;
  return ASOHF_OK;
}

// ---------------------------------------------------------------------------
// TSC: binning along the hierarchy (hier.hip)
// ---------------------------------------------------------------------------
static uint64_t hier_key(const asohf_ctx* ctx, const Parts& A, int red) {
  uint64_t h = ctx->plan.key;
  const long long v[6] = {A.n, ctx->tune.chunk, ctx->tune.tile[0], ctx->tune.tile[1], ctx->tune.tile[2], red};
  return fnv(h, v, sizeof(v));
}

// Exact form of the capacities: the count pass (every particle enumerated
// over the patch hash), a capacity per tile from its count, offsets, the work
// items planned on the capacities, and the buffers sized to them.  The items
// stay valid for later calls on the same plan, particle count and tuning.
static asohf_status hier_plan(asohf_ctx* ctx, PipeArgs& a, hipStream_t s) {
  HostPlan& H = ctx->plan;
  HIPCHK(launch_fill_u32((unsigned*)a.C, 0u, sizeof(Counters) / 4, s), "clearing counters");
  HIPCHK(launch_fill_u32(a.cnt, 0u, a.nkeys, s), "clearing counts");
  // a member's home-tile entry counts under key 2t, its others under 2t + 1
  Plan pc = a.P;
  pc.home_split = 1;
  HIPCHK(launch_count(pc, a.A, a.cnt, a.C, a.BT, a.cus, s), "count kernel");
  HIPCHK(launch_count_to_capacity(a.cnt, a.nkeys, a.cus, s), "capacities");
  HIPCHK(launch_scan(a.cnt, a.nkeys, a.bsum, a.start, a.cursor, s), "scan");
  HIPCHK(launch_plan_count(a.start, a.ntiles, a.chunk, 0, a.nitems, a.nsplit, a.is_split, s),
         "planning items");
  HIPCHK(launch_scan(a.nitems, a.ntiles, a.bsum, a.item_off, nullptr, s), "item offsets");
  HIPCHK(launch_scan(a.nsplit, a.ntiles, a.bsum, a.split_off, nullptr, s), "split offsets");
  HIPCHK(launch_scan(a.is_split, a.ntiles, a.bsum, a.split_rank, nullptr, s), "split tiles");
  const int nl = a.P.nl;
  HIPCHK(hipMemcpyAsync(ctx->hcount, a.C, sizeof(Counters), hipMemcpyDeviceToHost, s), "reading counters");
  HIPCHK(hipMemcpyAsync(ctx->hscalar, a.start + a.nkeys, sizeof(unsigned long long), hipMemcpyDeviceToHost, s),
         "reading totals");
  HIPCHK(hipMemcpyAsync(ctx->hscalar + 1, a.split_off + a.ntiles, sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, s), "reading totals");
  for (int l = 0; l <= nl + 1; ++l)
    HIPCHK(hipMemcpyAsync(ctx->hscalar + 2 + l, a.item_off + H.level_tile0[l], sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, s), "reading item offsets");
  HIPCHK(hipStreamSynchronize(s), "count pass");
  asohf_status st = decode_err(ctx, ctx->hcount->err);
  if (st != ASOHF_OK) return st;
  const unsigned long long cap_entries = ctx->hscalar[0];
  const unsigned long long nsplit = ctx->hscalar[1];
  const long long n_items = (long long)ctx->hscalar[2 + nl + 1];
  if (n_items >= (1ll << 31)) return fail(ctx, ASOHF_ERR_CAPACITY, "more than 2^31 work items");
  if (cap_entries >= (1ull << 40)) return fail(ctx, ASOHF_ERR_CAPACITY, "more than 2^40 tile entries");
  for (int l = 0; l <= nl + 1; ++l) ctx->hier_items[l] = (long long)ctx->hscalar[2 + l];
  const int T = a.P.tile[0] * a.P.tile[1] * a.P.tile[2];
  HIPCHK(ctx->entries.ensure(sizeof(int4) * std::max<unsigned long long>(cap_entries, 1)), "allocating entries");
  HIPCHK(ctx->items.ensure(sizeof(Item) * std::max<long long>(n_items, 1)), "allocating work items");
  HIPCHK(ctx->scratch.ensure((size_t)std::max<unsigned long long>(nsplit, 1) * T * a.elem),
         "allocating split-tile scratch");
  HIPCHK(ctx->ipairs.ensure(sizeof(unsigned) * std::max<long long>(n_items, 1)), "allocating item sums");
  a.D.entries = ctx->entries.as<int4>();
  a.D.items = ctx->items.as<Item>();
  a.D.scratch = ctx->scratch.p;
  HIPCHK(launch_plan_write(a.start, a.ntiles, a.chunk, 0, a.item_off, a.split_off, a.split_rank,
                           const_cast<Item*>(a.D.items), a.split, a.split_tiles, a.D.tile_mass_i, a.C, s),
         "writing items");
  ctx->hier_nitems = n_items;
  ctx->hier_nsplit = nsplit;
  return ASOHF_OK;
}

// the arguments of a call on the planned capacities
static void hier_bind(asohf_ctx* ctx, PipeArgs& a) {
  a.D.entries = ctx->entries.as<int4>();
  a.D.items = ctx->items.as<Item>();
  a.D.scratch = ctx->scratch.p;
  a.D.cursor = a.cursor;
  a.D.start = a.start;
  a.item_pairs = ctx->ipairs.as<unsigned>();
  a.n_items = ctx->hier_nitems;
  for (int l = 0; l <= kNLevels + 1; ++l) a.lvl_items[l] = ctx->hier_items[std::min(l, a.P.nl + 1)];
  a.emit_levels = ctx->plan.emit_levels;
  a.nl = a.P.nl;
  a.grid = (int)std::min<long long>(a.grid, std::max<long long>(a.n_items, 1));
  a.D.n_queues = std::min(kQueues, a.grid);
}

// One call on the planned capacities: base-grid binning, one emission pass
// per level, the persistent deposit over every item, the split tiles, the
// grid sums and the call's statistics; no host round trip (graph-capturable).
static asohf_status enqueue_hier(asohf_ctx* ctx, const PipeArgs& a, hipStream_t s, bool events) {
  HIPCHK(launch_fill_u32((unsigned*)a.C, 0u, sizeof(Counters) / 4, s), "clearing counters");
  HIPCHK(launch_fill_u32(&a.C->n_items_eff, (unsigned)a.n_items, 1, s), "queue length");
  HIPCHK(launch_fill_u32(a.item_pairs, 0u, a.n_items, s), "clearing item sums");
  HIPCHK(launch_cursor_reset(a.start, a.cursor, a.nkeys, a.cus, s), "cursors");
  HIPCHK(launch_bin_base_tsc(a.P, a.A, a.cursor, a.start, const_cast<int4*>(a.D.entries), a.C, a.BT, a.cus, s),
         "base-grid binning");
  if (a.A.n > 0) HIPCHK(launch_reduce_blocks(a.BT, hier_bin_blocks(a.A.n), a.C, a.cus, s), "block sums");
  if (events) {
    HIPCHK(hipEventRecord(ctx->ev[1], s), "event");
    HIPCHK(hipEventRecord(ctx->ev[2], s), "event");
  }
  for (int l = 0; l < a.nl; ++l) {
    if (!(a.emit_levels & (1u << l))) continue;
    HIPCHK(launch_emit_tsc(a.P, a.D.items, a.lvl_items[l], a.lvl_items[l + 1], l, a.cursor, a.start,
                           const_cast<int4*>(a.D.entries), a.item_pairs, a.C, a.cus, s),
           "emission pass");
  }
  if (events) HIPCHK(hipEventRecord(ctx->ev[3], s), "event");
  HIPCHK(launch_zero_split(a.P, a.D, a.red, a.cus, s), "clearing split tiles");
  if (events) HIPCHK(hipEventRecord(ctx->ev[4], s), "event");
  HIPCHK(launch_deposit(a.P, a.D, a.red, a.grid, a.block, s), "deposit kernel");
  if (events) HIPCHK(hipEventRecord(ctx->ev[8], s), "event");
  HIPCHK(launch_split_reduce(a.P, a.D, a.red, a.cus, s), "split tiles");
  if (events) HIPCHK(hipEventRecord(ctx->ev[5], s), "event");
  if (a.grid_mass) HIPCHK(launch_grid_sums(a.P, a.D, a.red, a.grid_mass, s), "grid sums");
  HIPCHK(launch_hier_stats(a.item_pairs, a.n_items, a.start, a.cursor, a.nkeys, a.C, a.cus, s), "statistics");
  return ASOHF_OK;
}

static asohf_status run_hier(asohf_ctx* ctx, const asohf_opts& o, const Parts& A, float* u1, float* u11,
                             double* grid_mass, hipStream_t s) {
  for (int attempt = 0; attempt < 2; ++attempt) {
    PipeArgs a;
    // the count pass keeps no emission lists here
    asohf_status st = prepare_args(ctx, o, A, u1, u11, grid_mass, a, false);
    if (st != ASOHF_OK) return st;
    const uint64_t hk = hier_key(ctx, A, o.reduction);
    bool planned = false;
    if (!ctx->hier_sized || ctx->hier_key != hk) {
      ctx->hier_sized = false;
      if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
      ctx->gexec = nullptr;
      HIPCHK(hipEventRecord(ctx->ev[0], s), "event");
      st = hier_plan(ctx, a, s);
      if (st != ASOHF_OK) return st;
      ctx->hier_key = hk;
      ctx->hier_sized = true;
      planned = true;
    }
    hier_bind(ctx, a);
    hipStream_t rs = s;
    bool phases = true;
    if (o.use_graph && !planned) {
      phases = false;
      rs = ctx->stream;
      if (!ctx->gexec || memcmp(&ctx->gkey, &a, sizeof(a)) != 0) {
        if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
        ctx->gexec = nullptr;
        hipGraph_t graph = nullptr;
        HIPCHK(hipStreamBeginCapture(rs, hipStreamCaptureModeThreadLocal), "graph capture");
        st = enqueue_hier(ctx, a, rs, false);
        const hipError_t e = hipStreamEndCapture(rs, &graph);
        if (st != ASOHF_OK) {
          if (e == hipSuccess && graph) (void)hipGraphDestroy(graph);
          return st;
        }
        HIPCHK(e, "graph capture");
        const hipError_t ei = hipGraphInstantiate(&ctx->gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) ctx->gexec = nullptr;
        HIPCHK(ei, "instantiating the graph");
        ctx->gkey = a;
      }
      if (s != rs) {
        HIPCHK(hipEventRecord(ctx->ev[7], s), "event");
        HIPCHK(hipStreamWaitEvent(rs, ctx->ev[7], 0), "stream wait");
      }
      HIPCHK(hipEventRecord(ctx->ev[0], rs), "event");
      HIPCHK(hipGraphLaunch(ctx->gexec, rs), "graph launch");
    } else {
      if (!planned) HIPCHK(hipEventRecord(ctx->ev[0], rs), "event");
      st = enqueue_hier(ctx, a, rs, true);
      if (st != ASOHF_OK) return st;
    }
    HIPCHK(hipEventRecord(ctx->ev[6], rs), "event");
    HIPCHK(hipMemcpyAsync(ctx->hcount, a.C, sizeof(Counters), hipMemcpyDeviceToHost, rs), "reading counters");
    HIPCHK(hipStreamSynchronize(rs), "batched build");
    Counters hc = *ctx->hcount;
    st = decode_err(ctx, hc.err);
    if (st != ASOHF_OK) return st;
    if (hc.overflow) {
      // a tile received more entries than its capacity: plan them again
      ctx->hier_sized = false;
      continue;
    }
    hc.n_items = (unsigned long long)a.n_items;
    set_stats(ctx, A, hc, phases);
    return ASOHF_OK;
  }
  return fail(ctx, ASOHF_ERR_CAPACITY, "tile capacities overflowed right after an exact count");
}

// Batched build.  Once a call has sized the buffers, later calls run without
// any host round trip before the deposit (the gate kernel checks capacities
// on the device) and read the counters once at the end; with opts.use_graph
// the whole sequence is one hipGraph, captured on first use and replayed
// while every kernel argument stays the same.
static asohf_status run_batched(asohf_ctx* ctx, const asohf_opts& o, const Parts& A, float* u1, float* u11,
                                double* grid_mass, hipStream_t s) {
  if (A.n > (1ll << 32)) return fail(ctx, ASOHF_ERR_CAPACITY, "more than 2^32 particles");
  if (o.mode == ASOHF_MODE_TSC && !ctx->hier_off) return run_hier(ctx, o, A, u1, u11, grid_mass, s);
  ctx->hier_sized = false;   // this path re-plans the offsets and items the hierarchy binning keeps
  if (!ctx->sized) return run_exact(ctx, o, A, u1, u11, grid_mass, s);
  PipeArgs a;
  asohf_status st = prepare_args(ctx, o, A, u1, u11, grid_mass, a);
  if (st != ASOHF_OK) return st;
  bind_sized(ctx, a);
  hipStream_t rs = s;   // the stream the work runs on
  bool phases = true;
  if (o.use_graph) {
    phases = false;
    rs = ctx->stream;   // captured and replayed on the library's own stream
    if (!ctx->gexec || memcmp(&ctx->gkey, &a, sizeof(a)) != 0) {
      if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
      ctx->gexec = nullptr;
      hipGraph_t graph = nullptr;
      HIPCHK(hipStreamBeginCapture(rs, hipStreamCaptureModeThreadLocal), "graph capture");
      st = enqueue_count(ctx, a, rs);
      if (st == ASOHF_OK) st = enqueue_deposit(ctx, a, rs, false);
      const hipError_t e = hipStreamEndCapture(rs, &graph);
      if (st != ASOHF_OK) {
        if (e == hipSuccess && graph) (void)hipGraphDestroy(graph);
        return st;
      }
      HIPCHK(e, "graph capture");
      const hipError_t ei = hipGraphInstantiate(&ctx->gexec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ei != hipSuccess) ctx->gexec = nullptr;
      HIPCHK(ei, "instantiating the graph");
      ctx->gkey = a;
    }
    if (s != rs) {   // order after the caller's earlier work on its stream
      HIPCHK(hipEventRecord(ctx->ev[7], s), "event");
      HIPCHK(hipStreamWaitEvent(rs, ctx->ev[7], 0), "stream wait");
    }
    HIPCHK(hipEventRecord(ctx->ev[0], rs), "event");
    HIPCHK(hipGraphLaunch(ctx->gexec, rs), "graph launch");
  } else {
    HIPCHK(hipEventRecord(ctx->ev[0], rs), "event");
    st = enqueue_count(ctx, a, rs);
    if (st != ASOHF_OK) return st;
    HIPCHK(hipEventRecord(ctx->ev[1], rs), "event");
    st = enqueue_deposit(ctx, a, rs, true);
    if (st != ASOHF_OK) return st;
  }
  HIPCHK(hipEventRecord(ctx->ev[6], rs), "event");
  HIPCHK(hipMemcpyAsync(ctx->hcount, a.C, sizeof(Counters), hipMemcpyDeviceToHost, rs), "reading counters");
  HIPCHK(hipStreamSynchronize(rs), "batched build");
  const Counters hc = *ctx->hcount;
  st = decode_err(ctx, hc.err);
  if (st != ASOHF_OK) return st;
  note_pool(ctx, hc);
  if (hc.overflow) {
    // the buffers of earlier calls were too small: nothing was deposited
    ctx->sized = false;
    return run_exact(ctx, o, A, u1, u11, grid_mass, s);
  }
  set_stats(ctx, A, hc, phases);
  return ASOHF_OK;
}

// ---------------------------------------------------------------------------
// plain form: one launch per grid
// ---------------------------------------------------------------------------
static asohf_status run_plain(asohf_ctx* ctx, const asohf_opts& o, const Parts& A, const asohf_geom* g,
                              float* u1, float* u11, double* grid_mass, hipStream_t s) {
  (void)o;
  ctx->hier_sized = false;   // the offsets buffer is reused below
  const Plan& P = ctx->plan.dev;
  HIPCHK(ctx->counters.ensure(sizeof(Counters)), "allocating counters");
  Counters* C = ctx->counters.as<Counters>();
  HIPCHK(hipEventRecord(ctx->ev[0], s), "event");
  HIPCHK(launch_fill_u32((unsigned*)C, 0u, sizeof(Counters) / 4, s), "clearing counters");
  HIPCHK(launch_plain_validate(P, A, C, ctx->cus, s), "validation");
  HIPCHK(hipMemcpyAsync(ctx->hcount, C, sizeof(Counters), hipMemcpyDeviceToHost, s), "reading counters");
  HIPCHK(hipStreamSynchronize(s), "validation");
  asohf_status st = decode_err(ctx, ctx->hcount->err);
  if (st != ASOHF_OK) return st;
  HIPCHK(hipEventRecord(ctx->ev[1], s), "event");
  HIPCHK(launch_plain_zero(P, u1, u11, s), "zeroing");
  HIPCHK(launch_plain_base(P, A, u1, ctx->cus, s), "base grid");
  if (g->nl > 0) {
    const long long n = A.n;
    HIPCHK(ctx->flags.ensure(sizeof(unsigned) * std::max<long long>(n, 1)), "allocating flags");
    HIPCHK(ctx->start.ensure(sizeof(unsigned long long) * (n + 1)), "allocating offsets");
    HIPCHK(ctx->bsum.ensure(sizeof(unsigned long long) * scan_scratch_elems(n)), "allocating scan");
    HIPCHK(ctx->idx.ensure(sizeof(long long) * std::max<long long>(n, 1)), "allocating candidates");
    int p0 = 0;
    for (int l = 0; l < 1; ++l) p0 += g->npatch[l];
    for (int l = 1; l <= g->nl; ++l) {
      if (g->npatch[l] == 0) continue;
      HIPCHK(launch_plain_flags(P, A, l, ctx->flags.as<unsigned>(), ctx->cus, s), "level flags");
      HIPCHK(launch_scan(ctx->flags.as<unsigned>(), n, ctx->bsum.as<unsigned long long>(),
                         ctx->start.as<unsigned long long>(), nullptr, s),
             "level scan");
      HIPCHK(launch_plain_compact(ctx->flags.as<unsigned>(), ctx->start.as<unsigned long long>(), n,
                                  ctx->idx.as<long long>(), ctx->cus, s),
             "level compaction");
      for (int p = p0; p < p0 + g->npatch[l]; ++p)
        HIPCHK(launch_plain_patch(P, A, 1 + p, ctx->idx.as<long long>(), ctx->start.as<unsigned long long>() + n,
                                  n, u11, ctx->cus, s),
               "patch kernel");
      p0 += g->npatch[l];
    }
  }
  HIPCHK(launch_plain_finish(P, u1, u11, grid_mass, s), "normalisation");
  HIPCHK(hipEventRecord(ctx->ev[6], s), "event");
  ctx->stats = asohf_stats{};
  ctx->stats.n_particles = A.n;
  ctx->stats.n_tiles = 0;
  ctx->stats.n_cells = ctx->plan.ncells;
  ctx->stats_pending = true;
  ctx->stats_phases = false;
  return ASOHF_OK;
}

// ---------------------------------------------------------------------------
// extern "C"
// ---------------------------------------------------------------------------
extern "C" {

const char* asohf_hip_version(void) { return "asohf_hip 0.1 (gfx950)"; }

asohf_status asohf_hip_create(asohf_ctx** out, int32_t device) {
  if (!out) return ASOHF_ERR_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return ASOHF_ERR_NODEVICE;
  asohf_ctx* ctx = new asohf_ctx();
  if (device < 0) {
    if (hipGetDevice(&ctx->device) != hipSuccess) { delete ctx; return ASOHF_ERR_HIP; }
  } else {
    if (device >= ndev) { delete ctx; return ASOHF_ERR_NODEVICE; }
    ctx->device = device;
  }
  if (hipSetDevice(ctx->device) != hipSuccess) { delete ctx; return ASOHF_ERR_HIP; }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, ctx->device) != hipSuccess) { delete ctx; return ASOHF_ERR_HIP; }
  ctx->cus = prop.multiProcessorCount;
  {
    // ASOHF_TSC_BINNING=hash: the TSC build bins with the patch-hash count
    // and scatter passes (the kernel-table mode always does)
    const char* b = getenv("ASOHF_TSC_BINNING");
    ctx->hier_off = b && strcmp(b, "hash") == 0;
  }
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) { delete ctx; return ASOHF_ERR_HIP; }
  for (auto& e : ctx->ev)
    if (hipEventCreate(&e) != hipSuccess) { asohf_hip_destroy(ctx); return ASOHF_ERR_HIP; }
  if (hipHostMalloc((void**)&ctx->hcount, sizeof(Counters)) != hipSuccess ||
      hipHostMalloc((void**)&ctx->hscalar, sizeof(unsigned long long) * 16) != hipSuccess) {
    asohf_hip_destroy(ctx);
    return ASOHF_ERR_OOM;
  }
  // the device address of this GPU's no-allocation scratch array (taken once
  // per device, with the device current)
  if (noalloc_scratch_init(ctx->device, &ctx->noalloc) != hipSuccess) {
    asohf_hip_destroy(ctx);
    return ASOHF_ERR_HIP;
  }
  *out = ctx;
  return ASOHF_OK;
}

void asohf_hip_destroy(asohf_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
  DBuf* bufs[] = {&ctx->pool, &ctx->pool_off, &ctx->pool_nk, &ctx->blk_pairs, &ctx->blk_mmax, &ctx->nsplit,
                  &ctx->item_off, &ctx->split_off, &ctx->plan_blob, &ctx->cnt, &ctx->start, &ctx->cursor, &ctx->bsum, &ctx->counters,
                  &ctx->stiles, &ctx->isplit, &ctx->srank, &ctx->nitems, &ctx->split, &ctx->tmass, &ctx->entries, &ctx->items,
                  &ctx->scratch, &ctx->flags, &ctx->idx, &ctx->h_in[0], &ctx->h_in[1], &ctx->h_in[2],
                  &ctx->h_in[3], &ctx->h_out[0], &ctx->h_out[1], &ctx->h_mass, &ctx->ord[0], &ctx->ord[1],
                  &ctx->ord[2], &ctx->ord[3], &ctx->ord_tmp, &ctx->elist, &ctx->eoff, &ctx->ecount, &ctx->sp_k[0],
                  &ctx->sp_k[1], &ctx->sp_k[2], &ctx->sp_k[3], &ctx->sp_dens, &ctx->sp_cnt, &ctx->sp_tmp,
                  &ctx->pk_r[0], &ctx->pk_r[1], &ctx->pk_k[0], &ctx->pk_k[1], &ctx->pk_tmp, &ctx->pk_cnt,
                  &ctx->pk_geo, &ctx->mk_geo, &ctx->mk_err, &ctx->vg_cnt, &ctx->vg_start, &ctx->vg_pairs};
  for (DBuf* b : bufs) b->release();
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->hcount) (void)hipHostFree(ctx->hcount);
  if (ctx->hscalar) (void)hipHostFree(ctx->hscalar);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* asohf_hip_last_error(const asohf_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

asohf_status asohf_hip_density(asohf_ctx* ctx, const asohf_geom* geom, int64_t n, const float* rx,
                               const float* ry, const float* rz, const float* m, const int32_t* npart_esp,
                               float* u1, float* u11, double* grid_mass, const asohf_opts* opts) {
  if (!ctx) return ASOHF_ERR_INVALID;
  ctx->err.clear();
  const asohf_opts o = opts ? *opts : default_opts();
  if (o.reduction != 0 && o.reduction != 1) return fail(ctx, ASOHF_ERR_INVALID, "unknown reduction");
  if (o.path != 0 && o.path != 1) return fail(ctx, ASOHF_ERR_INVALID, "unknown path");
  asohf_status st = check_geom(ctx, geom, o.mode);
  if (st != ASOHF_OK) return st;
  if (!u1) return fail(ctx, ASOHF_ERR_INVALID, "u1 is NULL");
  long long P = 0;
  for (int l = 0; l <= geom->nl; ++l) P += geom->npatch[l];
  if (P > 0 && !u11) return fail(ctx, ASOHF_ERR_INVALID, "u11 is NULL but there are patches");
  if (!tile_shape_supported(ctx->tune.tile)) return fail(ctx, ASOHF_ERR_INVALID, "unsupported tile shape");
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  Parts A;
  st = make_parts(ctx, o, n, rx, ry, rz, m, npart_esp, A);
  if (st != ASOHF_OK) return st;
  st = build_plan(ctx, geom, o.mode, &o);
  if (st != ASOHF_OK) return st;
  hipStream_t s = (hipStream_t)o.stream;   // NULL: the default stream
  if (o.path == ASOHF_PATH_PLAIN) return run_plain(ctx, o, A, geom, u1, u11, grid_mass, s);
  return run_batched(ctx, o, A, u1, u11, grid_mass, s);
}

asohf_status asohf_hip_density_host(asohf_ctx* ctx, const asohf_geom* geom, int64_t n, const float* rx,
                                    const float* ry, const float* rz, const float* m,
                                    const int32_t* npart_esp, float* u1, float* u11, double* grid_mass,
                                    const asohf_opts* opts) {
  if (!ctx) return ASOHF_ERR_INVALID;
  ctx->err.clear();
  asohf_opts o = opts ? *opts : default_opts();
  asohf_status st = check_geom(ctx, geom, o.mode);
  if (st != ASOHF_OK) return st;
  if (n < 0 || (n > 0 && (!rx || !ry || !rz || !m)) || !u1)
    return fail(ctx, ASOHF_ERR_INVALID, "NULL particle or output array");
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  st = build_plan(ctx, geom, o.mode, &o);
  if (st != ASOHF_OK) return st;
  const HostPlan& H = ctx->plan;
  hipStream_t s = o.stream ? (hipStream_t)o.stream : ctx->stream;   // host arrays: the context's stream
  const size_t pb = sizeof(float) * std::max<int64_t>(n, 1);
  const float* src[4] = {rx, ry, rz, m};
  for (int k = 0; k < 4; ++k) {
    HIPCHK(ctx->h_in[k].ensure(pb), "allocating particles");
    if (n > 0) HIPCHK(hipMemcpyAsync(ctx->h_in[k].p, src[k], sizeof(float) * n, hipMemcpyHostToDevice, s), "H2D");
  }
  const size_t b1 = sizeof(float) * H.u1_cells, b11 = sizeof(float) * std::max<long long>(H.u11_cells, 1);
  HIPCHK(ctx->h_out[0].ensure(b1), "allocating U1");
  HIPCHK(ctx->h_out[1].ensure(b11), "allocating U11");
  // untouched padding keeps the caller's values (INIT_GRIDVARS' -1, alloc.f:162-186)
  HIPCHK(hipMemcpyAsync(ctx->h_out[0].p, u1, b1, hipMemcpyHostToDevice, s), "H2D");
  if (H.u11_cells > 0 && u11) HIPCHK(hipMemcpyAsync(ctx->h_out[1].p, u11, b11, hipMemcpyHostToDevice, s), "H2D");
  double* dmass = nullptr;
  if (grid_mass) {
    HIPCHK(ctx->h_mass.ensure(sizeof(double) * H.grids.size()), "allocating masses");
    dmass = ctx->h_mass.as<double>();
  }
  o.stream = (void*)s;
  st = asohf_hip_density(ctx, geom, n, ctx->h_in[0].as<float>(), ctx->h_in[1].as<float>(),
                         ctx->h_in[2].as<float>(), ctx->h_in[3].as<float>(), npart_esp,
                         ctx->h_out[0].as<float>(), ctx->h_out[1].as<float>(), dmass, &o);
  if (st != ASOHF_OK) return st;
  HIPCHK(hipMemcpyAsync(u1, ctx->h_out[0].p, b1, hipMemcpyDeviceToHost, s), "D2H");
  if (H.u11_cells > 0 && u11) HIPCHK(hipMemcpyAsync(u11, ctx->h_out[1].p, b11, hipMemcpyDeviceToHost, s), "D2H");
  if (grid_mass) HIPCHK(hipMemcpyAsync(grid_mass, dmass, sizeof(double) * H.grids.size(), hipMemcpyDeviceToHost, s), "D2H");
  HIPCHK(hipStreamSynchronize(s), "host density call");
  return ASOHF_OK;
}

asohf_status asohf_hip_density_adjoint(asohf_ctx* ctx, const asohf_geom* geom, int64_t n, const float* rx,
                                       const float* ry, const float* rz, const float* m, const float* g1,
                                       const float* g11, float* dm, float* dxp, float* dyp, float* dzp,
                                       const asohf_opts* opts) {
  if (!ctx) return ASOHF_ERR_INVALID;
  ctx->err.clear();
  asohf_opts o = opts ? *opts : default_opts();
  if (o.mode != ASOHF_MODE_TSC)
    return fail(ctx, ASOHF_ERR_INVALID, "the adjoint is defined for the TSC mode (table weights are piecewise constant)");
  asohf_status st = check_geom(ctx, geom, o.mode);
  if (st != ASOHF_OK) return st;
  if (n < 0 || (n > 0 && (!rx || !ry || !rz || !m || !dm || !dxp || !dyp || !dzp)) || !g1)
    return fail(ctx, ASOHF_ERR_INVALID, "NULL array");
  long long npatch = 0;
  for (int l = 0; l <= geom->nl; ++l) npatch += geom->npatch[l];
  if (npatch > 0 && !g11) return fail(ctx, ASOHF_ERR_INVALID, "g11 is NULL but there are patches");
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  st = build_plan(ctx, geom, o.mode, &o);
  if (st != ASOHF_OK) return st;
  Parts A;
  st = make_parts(ctx, o, n, rx, ry, rz, m, nullptr, A);
  if (st != ASOHF_OK) return st;
  hipStream_t s = (hipStream_t)o.stream;   // NULL: the default stream
  HIPCHK(ctx->counters.ensure(sizeof(Counters)), "allocating counters");
  Counters* C = ctx->counters.as<Counters>();
  HIPCHK(hipEventRecord(ctx->ev[0], s), "event");
  // the forward's input checks (non-finite values, negative masses,
  // positions outside the TSC base range), read back after the gather
  HIPCHK(launch_fill_u32((unsigned*)C, 0u, sizeof(Counters) / 4, s), "clearing counters");
  // Large inputs are gathered in base-cell order (a radix sort of (cell,
  // index) pairs): neighbouring lanes then read the same gradient rows.  The
  // key pass makes the input checks; smaller inputs get the validation pass.
  const bool sorted = n >= kAdjointSortMin && n < (1ll << 32);
  if (n > 0 && !sorted) HIPCHK(launch_plain_validate(ctx->plan.dev, A, C, ctx->cus, s), "validation");
  const unsigned* order = nullptr;
  if (sorted) {
    const long long nc = (long long)geom->nx * geom->ny * geom->nz;
    int bits = 1;
    while (bits < 32 && (1ll << bits) < nc) ++bits;
    const size_t b4 = sizeof(unsigned) * (size_t)n;
    HIPCHK(ctx->ord[0].ensure(b4), "allocating the gather order");
    HIPCHK(ctx->ord[1].ensure(b4), "allocating the gather order");
    HIPCHK(ctx->ord[2].ensure(b4), "allocating the gather order");
    HIPCHK(ctx->ord[3].ensure(b4), "allocating the gather order");
    size_t tb = 0;
    HIPCHK(sort_cell_pairs(nullptr, tb, nullptr, nullptr, nullptr, nullptr, (unsigned long long)n, bits, s),
           "sizing the sort");
    HIPCHK(ctx->ord_tmp.ensure(std::max<size_t>(tb, 1)), "allocating the sort");
    HIPCHK(launch_cell_keys(ctx->plan.dev, A, ctx->ord[0].as<unsigned>(), ctx->ord[1].as<unsigned>(), C, ctx->cus, s),
           "cell keys");
    HIPCHK(sort_cell_pairs(ctx->ord_tmp.p, tb, ctx->ord[0].as<unsigned>(), ctx->ord[2].as<unsigned>(),
                           ctx->ord[1].as<unsigned>(), ctx->ord[3].as<unsigned>(), (unsigned long long)n, bits, s),
           "sorting by cell");
    order = ctx->ord[3].as<unsigned>();
  }
  if (n > 0) HIPCHK(launch_adjoint(ctx->plan.dev, A, g1, g11, dm, dxp, dyp, dzp, order, ctx->cus, s), "adjoint kernel");
  HIPCHK(hipEventRecord(ctx->ev[6], s), "event");
  HIPCHK(hipMemcpyAsync(ctx->hcount, C, sizeof(Counters), hipMemcpyDeviceToHost, s), "reading counters");
  HIPCHK(hipStreamSynchronize(s), "adjoint");
  st = decode_err(ctx, ctx->hcount->err);
  if (st != ASOHF_OK) return st;
  ctx->stats = asohf_stats{};
  ctx->stats.n_particles = n;
  ctx->stats_pending = true;
  ctx->stats_phases = false;
  return ASOHF_OK;
}

asohf_status asohf_hip_membership(asohf_ctx* ctx, const asohf_geom* geom, int64_t n, const float* rx,
                                  const float* ry, const float* rz, const int32_t* npart_esp, int32_t mode,
                                  int32_t* out, int64_t cap, int64_t* count) {
  if (!ctx) return ASOHF_ERR_INVALID;
  ctx->err.clear();
  asohf_opts o = default_opts();
  o.mode = mode;
  asohf_status st = check_geom(ctx, geom, mode);
  if (st != ASOHF_OK) return st;
  if (!count || (cap > 0 && !out)) return fail(ctx, ASOHF_ERR_INVALID, "NULL output");
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  st = build_plan(ctx, geom, mode, &o);
  if (st != ASOHF_OK) return st;
  Parts A;
  // positions stand in for the masses: the membership pass never reads them
  st = make_parts(ctx, o, n, rx, ry, rz, rx, npart_esp, A);
  if (st != ASOHF_OK) return st;
  HIPCHK(ctx->counters.ensure(sizeof(Counters)), "allocating counters");
  unsigned long long* c = (unsigned long long*)ctx->counters.p;
  hipStream_t s = ctx->stream;
  HIPCHK(launch_fill_u32((unsigned*)c, 0u, 4, s), "clearing");   // the count and the validation flags
  HIPCHK(launch_membership(ctx->plan.dev, A, out, cap, c, ctx->cus, s), "membership kernel");
  HIPCHK(hipMemcpyAsync(ctx->hscalar, c, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s), "D2H");
  HIPCHK(hipStreamSynchronize(s), "membership");
  st = decode_err(ctx, (unsigned)ctx->hscalar[1]);
  if (st != ASOHF_OK) return st;
  *count = (int64_t)ctx->hscalar[0];
  return ASOHF_OK;
}

asohf_status asohf_hip_get_tune(const asohf_ctx* ctx, asohf_tune* out) {
  if (!ctx || !out) return ASOHF_ERR_INVALID;
  *out = ctx->tune;
  return ASOHF_OK;
}

asohf_status asohf_hip_set_tune(asohf_ctx* ctx, const asohf_tune* t) {
  if (!ctx || !t) return ASOHF_ERR_INVALID;
  if (!tile_shape_supported(t->tile)) return fail(ctx, ASOHF_ERR_INVALID, "unsupported tile shape");
  if (t->block < 64 || t->block > 1024 || t->block % 64)
    return fail(ctx, ASOHF_ERR_INVALID, "block must be a multiple of 64 threads, at most 1024");
  if (t->chunk < 64 || t->waves_per_cu < 1 || t->waves_per_cu > 64)
    return fail(ctx, ASOHF_ERR_INVALID, "chunk >= 64 and 1 <= waves_per_cu <= 64");
  const bool reshape = memcmp(t->tile, ctx->tune.tile, sizeof(t->tile)) != 0;
  ctx->tune = *t;
  if (reshape) ctx->plan.valid = false;
  return ASOHF_OK;
}

asohf_status asohf_hip_autotune(asohf_ctx* ctx, const asohf_geom* geom, int64_t n, const float* rx,
                                const float* ry, const float* rz, const float* m, const int32_t* npart_esp,
                                float* u1, float* u11, const asohf_opts* opts, asohf_tune* out) {
  if (!ctx) return ASOHF_ERR_INVALID;
  asohf_opts o = opts ? *opts : default_opts();
  o.path = ASOHF_PATH_BATCHED;
  // tile shape and workgroup size together (the deposit's LDS tile and its
  // threads go together), then item chunk, then persistent grid; each
  // candidate is one warm-up call and two timed ones
  static const int shapes[][3] = {{8, 8, 8}, {16, 8, 8}, {16, 16, 8}, {16, 16, 16}};
  static const int blocks[] = {256, 512, 1024};
  static const int chunks[] = {2048, 8192, 32768};
  static const int waves[] = {8, 16, 32};
  asohf_tune best = ctx->tune;
  double best_ms = 1e30;
  auto trial = [&](const asohf_tune& t) -> asohf_status {
    asohf_status st = asohf_hip_set_tune(ctx, &t);
    if (st != ASOHF_OK) return st;
    double ms = 0;
    for (int rep = 0; rep < 3; ++rep) {
      st = asohf_hip_density(ctx, geom, n, rx, ry, rz, m, npart_esp, u1, u11, nullptr, &o);
      if (st != ASOHF_OK) return st;
      asohf_stats stt;
      st = asohf_hip_get_stats(ctx, &stt);
      if (st != ASOHF_OK) return st;
      if (rep > 0) ms += stt.ms_total;
    }
    if (ms < best_ms) { best_ms = ms; best = t; }
    return ASOHF_OK;
  };
  asohf_status st0;
  const asohf_tune start = best;
  for (auto& sh : shapes)
    for (int bl : blocks) {
      asohf_tune t = start;
      t.tile[0] = sh[0]; t.tile[1] = sh[1]; t.tile[2] = sh[2];
      t.block = bl;
      if ((st0 = trial(t)) != ASOHF_OK) return st0;
    }
  for (int ch : chunks) {
    asohf_tune t = best;
    t.chunk = ch;
    if ((st0 = trial(t)) != ASOHF_OK) return st0;
  }
  for (int wv : waves) {
    asohf_tune t = best;
    t.waves_per_cu = wv;
    if ((st0 = trial(t)) != ASOHF_OK) return st0;
  }
  asohf_status st = asohf_hip_set_tune(ctx, &best);
  if (st != ASOHF_OK) return st;
  if (out) *out = best;
  return ASOHF_OK;
}

asohf_status asohf_hip_species(asohf_ctx* ctx, const int32_t nmax[3], const float d[3], float rodo, float re0,
                               int64_t n, const float* rx, const float* ry, const float* rz, const float* m,
                               int32_t* species, int64_t* order, int32_t* npart_esp, void* stream) {
  if (!ctx) return ASOHF_ERR_INVALID;
  ctx->err.clear();
  if (!nmax || !d || !npart_esp || n < 0 || (n > 0 && (!rx || !ry || !rz || !m || !order)))
    return fail(ctx, ASOHF_ERR_INVALID, "NULL argument");
  for (int a = 0; a < 3; ++a) {
    if (nmax[a] < 1 || nmax[a] > 4096) return fail(ctx, ASOHF_ERR_INVALID, "NMAX outside 1..4096");
    if (!(d[a] > 0.f) || !std::isfinite(d[a])) return fail(ctx, ASOHF_ERR_INVALID, "cell size must be > 0");
  }
  if (n >= (1ll << 32)) return fail(ctx, ASOHF_ERR_CAPACITY, "2^32 or more particles");
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  hipStream_t s = (hipStream_t)stream;
  const long long nc = (long long)nmax[0] * nmax[1] * nmax[2];
  // BAS = DX*DY*DZ*RODO*RE0**3 in FP32, left to right (reader.f:203)
  const float bas = (((d[0] * d[1]) * d[2]) * rodo) * powi_f(re0, 3);
  const size_t b4 = sizeof(unsigned) * (size_t)std::max<int64_t>(n, 1);
  for (int k = 0; k < 4; ++k) HIPCHK(ctx->sp_k[k].ensure(b4), "allocating the species sort");
  HIPCHK(ctx->sp_dens.ensure(sizeof(float) * (size_t)nc), "allocating the species grid");
  HIPCHK(ctx->sp_cnt.ensure(sizeof(unsigned) * (kNEsp + 2)), "allocating the species counts");
  size_t tb = 0;
  HIPCHK(species_sort_bytes((unsigned long long)std::max<int64_t>(n, 1), tb), "sizing the species sort");
  HIPCHK(ctx->sp_tmp.ensure(std::max<size_t>(tb, 1)), "allocating the species sort");
  HIPCHK(launch_species(nmax, d, bas, n, rx, ry, rz, m, species, (long long*)order, ctx->sp_cnt.as<unsigned>(),
                        ctx->sp_dens.as<float>(), ctx->sp_k[0].as<unsigned>(), ctx->sp_k[1].as<unsigned>(),
                        ctx->sp_k[2].as<unsigned>(), ctx->sp_k[3].as<unsigned>(), ctx->sp_tmp.p,
                        ctx->sp_tmp.cap, ctx->cus, s),
         "species kernels");
  unsigned cnt[kNEsp];
  HIPCHK(hipMemcpyAsync(cnt, ctx->sp_cnt.p, sizeof(cnt), hipMemcpyDeviceToHost, s), "reading the counts");
  HIPCHK(hipStreamSynchronize(s), "species partition");
  for (int k = 0; k < kNEsp; ++k) npart_esp[k] = (int32_t)cnt[k];
  return ASOHF_OK;
}

asohf_status asohf_hip_species_noalloc(asohf_ctx* ctx, const int32_t nmax[3], const float d[3], float rodo,
                                       float re0, int64_t n, const float* rx, const float* ry, const float* rz,
                                       const float* m, int32_t* species, int64_t* order, int32_t* npart_esp,
                                       void* stream) {
  if (!ctx) return ASOHF_ERR_INVALID;
  ctx->err.clear();
  if (!nmax || !d || !npart_esp || n < 0 || (n > 0 && (!rx || !ry || !rz || !m || !order)))
    return fail(ctx, ASOHF_ERR_INVALID, "NULL argument");
  for (int a = 0; a < 3; ++a) {
    if (nmax[a] < 1 || nmax[a] > 4096) return fail(ctx, ASOHF_ERR_INVALID, "NMAX outside 1..4096");
    if (!(d[a] > 0.f) || !std::isfinite(d[a])) return fail(ctx, ASOHF_ERR_INVALID, "cell size must be > 0");
  }
  const long long nc = (long long)nmax[0] * nmax[1] * nmax[2];
  if (n > kNoAllocMaxParticles)
    return fail(ctx, ASOHF_ERR_CAPACITY, "asohf_hip_species_noalloc takes at most 2^20 particles");
  if (nc > kNoAllocMaxCells)
    return fail(ctx, ASOHF_ERR_CAPACITY, "asohf_hip_species_noalloc takes at most 2^21 grid cells");
  if (!ctx->noalloc) return fail(ctx, ASOHF_ERR_HIP, "no scratch array address for this device");
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  hipStream_t s = (hipStream_t)stream;
  const float bas = (((d[0] * d[1]) * d[2]) * rodo) * powi_f(re0, 3);   // reader.f:203
  unsigned* counts = nullptr;
  size_t need = 0;
  HIPCHK(launch_species_noalloc(ctx->noalloc, nmax, d, bas, n, rx, ry, rz, m, species, (long long*)order,
                                ctx->cus, s, &counts, &need),
         "species kernels");
  if (need > noalloc_temp_bytes())
    return fail(ctx, ASOHF_ERR_CAPACITY, "the radix sort's temporary storage exceeds the scratch array");
  // the counts come back through the context's pinned buffer
  HIPCHK(hipMemcpyAsync(ctx->hscalar, counts, sizeof(unsigned) * kNEsp, hipMemcpyDeviceToHost, s),
         "reading the counts");
  HIPCHK(hipStreamSynchronize(s), "species partition");
  const unsigned* cnt = reinterpret_cast<const unsigned*>(ctx->hscalar);
  for (int k = 0; k < kNEsp; ++k) npart_esp[k] = (int32_t)cnt[k];
  return ASOHF_OK;
}

asohf_status asohf_hip_peak_candidates(asohf_ctx* ctx, const asohf_geom* geom, int32_t level, const float* u1,
                                       const float* u11, const int32_t* conta, float contrastec, float* dens,
                                       int32_t* ix, int32_t* iy, int32_t* iz, int32_t* patch, int64_t cap,
                                       int64_t* count, const asohf_opts* opts) {
  if (!ctx) return ASOHF_ERR_INVALID;
  ctx->err.clear();
  if (!geom || !count || cap < 0 || (cap > 0 && (!dens || !ix || !iy || !iz)))
    return fail(ctx, ASOHF_ERR_INVALID, "NULL argument");
  *count = 0;
  if (geom->nx < 1 || geom->ny < 1 || geom->nz < 1 || geom->nl < 0 || geom->nl > kNLevels)
    return fail(ctx, ASOHF_ERR_INVALID, "grid sizes or NL out of range");
  if (level < 0 || level > geom->nl) return fail(ctx, ASOHF_ERR_INVALID, "level outside 0..NL");
  if (geom->nl > 0 && (!geom->npatch || !geom->patchnx || !geom->patchny || !geom->patchnz))
    return fail(ctx, ASOHF_ERR_INVALID, "NULL patch arrays");
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  hipStream_t s = opts ? (hipStream_t)opts->stream : (hipStream_t)0;
  // CONTRASTECPEAK = CONTRASTEC/6.0 in FP32 (haloes_grids.f:288)
  const float peak = contrastec / 6.0f;
  long long ld1 = 0, ld2 = 0;
  int np = 0, patch0 = 0;
  std::vector<PeakPatchRec> recs;
  if (level == 0) {
    if (!u1) return fail(ctx, ASOHF_ERR_INVALID, "NULL u1");
    const int l1 = (opts && opts->u1_ld[0] > 0) ? opts->u1_ld[0] : geom->nx;
    const int l2 = (opts && opts->u1_ld[1] > 0) ? opts->u1_ld[1] : geom->ny;
    if (l1 < geom->nx || l2 < geom->ny) return fail(ctx, ASOHF_ERR_INVALID, "u1 leading dimensions below NX, NY");
    ld1 = l1;
    ld2 = (long long)l1 * l2;
  } else {
    long long before = 0, total = 0;
    for (int l = 1; l <= geom->nl; ++l) {
      if (geom->npatch[l] < 0) return fail(ctx, ASOHF_ERR_INVALID, "negative NPATCH");
      if (l < level) before += geom->npatch[l];
      total += geom->npatch[l];
    }
    patch0 = (int)before;
    np = geom->npatch[level];
    const bool padded = opts && opts->u11_ld[0] > 0;
    if (padded) {
      ld1 = opts->u11_ld[0];
      ld2 = (long long)opts->u11_ld[0] * opts->u11_ld[1];
    }
    long long off = 0, rank = 0;
    for (long long p = 0; p < before + np; ++p) {
      const int a = geom->patchnx[p], b = geom->patchny[p], c = geom->patchnz[p];
      if (a < 1 || b < 1 || c < 1) return fail(ctx, ASOHF_ERR_INVALID, "patch sizes must be >= 1");
      if ((long long)a * b * c > (1ll << 30)) return fail(ctx, ASOHF_ERR_CAPACITY, "patch too large");
      if (padded && (a > opts->u11_ld[0] || b > opts->u11_ld[1] || c > opts->u11_ld[2]))
        return fail(ctx, ASOHF_ERR_INVALID, "a patch exceeds the padded U11 dimensions");
      const long long cells = (long long)a * b * c;
      const long long first = padded ? p * ld2 * opts->u11_ld[2] : off;
      if (p >= before) {
        PeakPatchRec r;
        r.n[0] = a;
        r.n[1] = b;
        r.n[2] = c;
        r.pad = 0;
        r.off = first;
        r.rank = rank;
        recs.push_back(r);
        rank += cells;
      }
      off += cells;
    }
    (void)total;
    if (np > 0 && !u11) return fail(ctx, ASOHF_ERR_INVALID, "NULL u11");
  }
  const bool empty = level == 0 ? (geom->nx < 3 || geom->ny < 3 || geom->nz < 3) : np == 0;
  if (empty) return ASOHF_OK;
  if (np > 0) {
    HIPCHK(ctx->pk_geo.ensure(sizeof(PeakPatchRec) * recs.size()), "allocating the patch records");
    HIPCHK(hipMemcpyAsync(ctx->pk_geo.p, recs.data(), sizeof(PeakPatchRec) * recs.size(), hipMemcpyHostToDevice, s),
           "uploading the patch records");
  }
  HIPCHK(ctx->pk_cnt.ensure(sizeof(unsigned long long)), "allocating the candidate counter");
  const PeakPatchRec* dpp = np > 0 ? ctx->pk_geo.as<PeakPatchRec>() : nullptr;
  unsigned long long want = (unsigned long long)std::max<int64_t>(cap, 1 << 16), k = 0;
  for (int pass = 0; pass < 2; ++pass) {
    for (int b = 0; b < 2; ++b) {
      HIPCHK(ctx->pk_r[b].ensure(sizeof(unsigned long long) * want), "allocating the candidates");
      HIPCHK(ctx->pk_k[b].ensure(sizeof(unsigned) * want), "allocating the candidates");
    }
    const unsigned long long have = std::min(ctx->pk_r[0].cap / sizeof(unsigned long long),
                                             ctx->pk_k[0].cap / sizeof(unsigned));
    HIPCHK(launch_peak_scan(level, level == 0 ? u1 : u11, conta, geom->nx, geom->ny, geom->nz, ld1, ld2, dpp, np,
                            peak, ctx->pk_r[0].as<unsigned long long>(), ctx->pk_k[0].as<unsigned>(), have,
                            ctx->pk_cnt.as<unsigned long long>(), ctx->cus, s),
           "scanning for peak candidates");
    HIPCHK(hipMemcpyAsync(&k, ctx->pk_cnt.p, sizeof(k), hipMemcpyDeviceToHost, s), "reading the count");
    HIPCHK(hipStreamSynchronize(s), "peak scan");
    if (k <= have) break;
    want = k;   // every candidate did not fit: grow and scan again
  }
  *count = (int64_t)k;
  if ((int64_t)k > cap) return fail(ctx, ASOHF_ERR_CAPACITY, "more peak candidates than cap (count set)");
  size_t tb = 0;
  HIPCHK(peak_sort_bytes(std::max<unsigned long long>(k, 1), tb), "sizing the candidate sort");
  HIPCHK(ctx->pk_tmp.ensure(std::max<size_t>(tb, 1)), "allocating the candidate sort");
  HIPCHK(launch_peak_sort_write(ctx->pk_r[0].as<unsigned long long>(), ctx->pk_k[0].as<unsigned>(),
                                ctx->pk_r[1].as<unsigned long long>(), ctx->pk_k[1].as<unsigned>(), (long long)k,
                                ctx->pk_tmp.p, ctx->pk_tmp.cap, level, geom->nx, geom->ny, dpp, np, patch0, dens,
                                ix, iy, iz, patch, ctx->cus, s),
         "sorting the peak candidates");
  HIPCHK(hipStreamSynchronize(s), "peak candidates");
  return ASOHF_OK;
}

// ---------------------------------------------------------------------------
// Grid masks and the density dump (asohf.f:515-545)
// ---------------------------------------------------------------------------
// The patch records of every level, offsets in the U11 layout of opts.
static asohf_status mask_patches(asohf_ctx* ctx, const asohf_geom* g, const int32_t* px, const int32_t* py,
                                 const int32_t* pz, const int32_t* pare, const asohf_opts* opts,
                                 std::vector<MaskPatch>& mp, std::vector<int>& lev_lo, std::vector<int>& lev_n,
                                 long long& s1p, long long& s2p, long long& cells11) {
  if (g->nx < 1 || g->ny < 1 || g->nz < 1 || g->nl < 0 || g->nl > kNLevels)
    return fail(ctx, ASOHF_ERR_INVALID, "grid sizes or NL out of range");
  if (g->nl > 0 && (!g->npatch || !g->patchnx || !g->patchny || !g->patchnz || !g->patchrx || !g->patchry ||
                    !g->patchrz))
    return fail(ctx, ASOHF_ERR_INVALID, "NULL patch arrays");
  lev_lo.assign(g->nl + 2, 0);
  lev_n.assign(g->nl + 2, 0);
  long long P = 0;
  for (int l = 1; l <= g->nl; ++l) {
    if (g->npatch[l] < 0) return fail(ctx, ASOHF_ERR_INVALID, "negative NPATCH");
    lev_lo[l] = (int)P;
    lev_n[l] = g->npatch[l];
    P += g->npatch[l];
  }
  if (P > (1ll << 30)) return fail(ctx, ASOHF_ERR_CAPACITY, "too many patches");
  const bool padded = opts && opts->u11_ld[0] > 0;
  s1p = padded ? opts->u11_ld[0] : 0;
  s2p = padded ? (long long)opts->u11_ld[0] * opts->u11_ld[1] : 0;
  const long long slab = padded ? s2p * opts->u11_ld[2] : 0;
  mp.resize((size_t)P);
  long long off = 0;
  for (long long p = 0; p < P; ++p) {
    MaskPatch& m = mp[(size_t)p];
    m.n[0] = g->patchnx[p];
    m.n[1] = g->patchny[p];
    m.n[2] = g->patchnz[p];
    if (m.n[0] < 1 || m.n[1] < 1 || m.n[2] < 1) return fail(ctx, ASOHF_ERR_INVALID, "patch sizes must be >= 1");
    if ((long long)m.n[0] * m.n[1] * m.n[2] > (1ll << 30)) return fail(ctx, ASOHF_ERR_CAPACITY, "patch too large");
    if (padded && (m.n[0] > opts->u11_ld[0] || m.n[1] > opts->u11_ld[1] || m.n[2] > opts->u11_ld[2]))
      return fail(ctx, ASOHF_ERR_INVALID, "a patch exceeds the padded U11 dimensions");
    m.pc[0] = px ? px[p] : 0;
    m.pc[1] = py ? py[p] : 0;
    m.pc[2] = pz ? pz[p] : 0;
    m.pare = pare ? pare[p] : 0;
    m.self = (int)(p + 1);
    m.pr[0] = g->patchrx[p];
    m.pr[1] = g->patchry[p];
    m.pr[2] = g->patchrz[p];
    m.pad = 0;
    m.off = padded ? p * slab : off;
    off += (long long)m.n[0] * m.n[1] * m.n[2];
  }
  cells11 = padded ? P * slab : off;
  return ASOHF_OK;
}

static asohf_status base_ld(asohf_ctx* ctx, const asohf_geom* g, const asohf_opts* opts, long long& b1,
                            long long& b2) {
  const int l1 = (opts && opts->u1_ld[0] > 0) ? opts->u1_ld[0] : g->nx;
  const int l2 = (opts && opts->u1_ld[1] > 0) ? opts->u1_ld[1] : g->ny;
  if (l1 < g->nx || l2 < g->ny) return fail(ctx, ASOHF_ERR_INVALID, "u1 leading dimensions below NX, NY");
  b1 = l1;
  b2 = (long long)l1 * l2;
  return ASOHF_OK;
}

static asohf_status mask_error(asohf_ctx* ctx, hipStream_t s) {
  int e = 0;
  HIPCHK(hipMemcpyAsync(&e, ctx->mk_err.p, sizeof(int), hipMemcpyDeviceToHost, s), "reading the mask status");
  HIPCHK(hipStreamSynchronize(s), "grid masks");
  if (e & 1) return fail(ctx, ASOHF_ERR_INVALID, "PARE of a level >= 2 patch is not a patch");
  if (e & 2) return fail(ctx, ASOHF_ERR_INVALID, "a mother cell (PATCHX..Z) falls outside its parent grid");
  if (e & 4) return fail(ctx, ASOHF_ERR_CAPACITY, "more than 512 coarser patches near one finer patch");
  if (e & 8) return fail(ctx, ASOHF_ERR_CAPACITY, "more than 16 patches of one level cover one cell");
  return ASOHF_OK;
}

asohf_status asohf_hip_refinement_masks(asohf_ctx* ctx, const asohf_geom* geom, const int32_t* patchx,
                                        const int32_t* patchy, const int32_t* patchz, const int32_t* pare,
                                        int32_t* cr0amr, int32_t* cr0amr11, const asohf_opts* opts) {
  if (!ctx) return ASOHF_ERR_INVALID;
  ctx->err.clear();
  if (!geom || !cr0amr) return fail(ctx, ASOHF_ERR_INVALID, "NULL argument");
  std::vector<MaskPatch> mp;
  std::vector<int> lo, ln;
  long long s1p, s2p, cells11, b1, b2;
  asohf_status st = mask_patches(ctx, geom, patchx, patchy, patchz, pare, opts, mp, lo, ln, s1p, s2p, cells11);
  if (st != ASOHF_OK) return st;
  if ((st = base_ld(ctx, geom, opts, b1, b2)) != ASOHF_OK) return st;
  if (!mp.empty() && (!patchx || !patchy || !patchz || !pare || !cr0amr11))
    return fail(ctx, ASOHF_ERR_INVALID, "NULL PATCHX..Z, PARE or cr0amr11");
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  hipStream_t s = opts ? (hipStream_t)opts->stream : (hipStream_t)0;
  HIPCHK(ctx->mk_err.ensure(sizeof(int)), "allocating the mask status");
  HIPCHK(launch_fill_i32(ctx->mk_err.as<int>(), 1, 0, ctx->cus, s), "clearing the mask status");
  if (!mp.empty()) {
    HIPCHK(ctx->mk_geo.ensure(sizeof(MaskPatch) * mp.size()), "allocating the patch records");
    HIPCHK(hipMemcpyAsync(ctx->mk_geo.p, mp.data(), sizeof(MaskPatch) * mp.size(), hipMemcpyHostToDevice, s),
           "uploading the patch records");
  }
  HIPCHK(launch_cr0amr(ctx->mk_geo.as<MaskPatch>(), (int)mp.size(), geom->nl >= 1 ? ln[1] : 0, lo.data(), ln.data(),
                       geom->nl, geom->lado0, geom->nx, geom->ny, geom->nz, cr0amr, b1, b2, cr0amr11, cells11, s1p,
                       s2p, ctx->mk_err.as<int>(), ctx->cus, s),
         "refinement masks");
  return mask_error(ctx, s);
}

asohf_status asohf_hip_overlap_mask(asohf_ctx* ctx, const asohf_geom* geom, int32_t* solap, const asohf_opts* opts) {
  if (!ctx) return ASOHF_ERR_INVALID;
  ctx->err.clear();
  if (!geom) return fail(ctx, ASOHF_ERR_INVALID, "NULL argument");
  std::vector<MaskPatch> mp;
  std::vector<int> lo, ln;
  long long s1p, s2p, cells11;
  asohf_status st = mask_patches(ctx, geom, nullptr, nullptr, nullptr, nullptr, opts, mp, lo, ln, s1p, s2p, cells11);
  if (st != ASOHF_OK) return st;
  if (mp.empty()) return ASOHF_OK;
  if (!solap) return fail(ctx, ASOHF_ERR_INVALID, "NULL solap");
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  hipStream_t s = opts ? (hipStream_t)opts->stream : (hipStream_t)0;
  HIPCHK(ctx->mk_err.ensure(sizeof(int)), "allocating the mask status");
  HIPCHK(launch_fill_i32(ctx->mk_err.as<int>(), 1, 0, ctx->cus, s), "clearing the mask status");
  HIPCHK(ctx->mk_geo.ensure(sizeof(MaskPatch) * mp.size()), "allocating the patch records");
  HIPCHK(hipMemcpyAsync(ctx->mk_geo.p, mp.data(), sizeof(MaskPatch) * mp.size(), hipMemcpyHostToDevice, s),
         "uploading the patch records");
  // SOLAP(:,:,:,I) = 1 over the whole (padded) array of every patch
  if (s1p > 0) HIPCHK(launch_fill_i32(solap, cells11, 1, ctx->cus, s), "filling SOLAP");
  // RADX(1) and DX from COMMON /GRID/ and /ESPACIADO/ (MALLA's when RADX is not given)
  const float lado = geom->lado0 - geom->lado0 / (float)geom->nx;
  std::vector<float> rad[3];
  const float* given[3] = {geom->radx, geom->rady, geom->radz};
  const int nn[3] = {geom->nx, geom->ny, geom->nz};
  for (int a = 0; a < 3; ++a) {
    if (given[a]) rad[a].assign(given[a], given[a] + nn[a] + 2);
    else malla(nn[a], lado, rad[a]);
  }
  const float dd[3] = {geom->dx, geom->dy, geom->dz};
  std::vector<unsigned> cnt;
  std::vector<unsigned long long> start;
  for (int ir = 1; ir <= geom->nl; ++ir) {
    const int np = ln[ir];
    if (np == 0) continue;
    VgGeo G;
    float p2 = 1.0f;
    for (int k = 0; k < ir; ++k) p2 *= 2.0f;
    for (int a = 0; a < 3; ++a) {
      G.d[a] = dd[a] / p2;                                     // DXPA = DX/(2.0**IR)
      G.fix[a] = (rad[a][1] - dd[a] * 0.5f) + 0.5f * G.d[a];   // RXFIX (grids.f:2285-2287)
    }
    HIPCHK(ctx->vg_cnt.ensure(sizeof(unsigned) * np), "allocating the neighbour counts");
    HIPCHK(ctx->vg_start.ensure(sizeof(unsigned long long) * (np + 1)), "allocating the neighbour offsets");
    HIPCHK(launch_vg_neighbours(ctx->mk_geo.as<MaskPatch>(), lo[ir], np, G, 0, ctx->vg_cnt.as<unsigned>(), nullptr,
                                nullptr, s),
           "counting neighbours");
    cnt.resize(np);
    HIPCHK(hipMemcpyAsync(cnt.data(), ctx->vg_cnt.p, sizeof(unsigned) * np, hipMemcpyDeviceToHost, s),
           "reading the neighbour counts");
    HIPCHK(hipStreamSynchronize(s), "neighbour counts");
    start.assign(np + 1, 0);
    for (int k = 0; k < np; ++k) start[k + 1] = start[k] + cnt[k];
    HIPCHK(hipMemcpyAsync(ctx->vg_start.p, start.data(), sizeof(unsigned long long) * (np + 1),
                          hipMemcpyHostToDevice, s),
           "uploading the neighbour offsets");
    HIPCHK(ctx->vg_pairs.ensure(sizeof(VgPair) * std::max<unsigned long long>(start[np], 1)),
           "allocating the neighbour pairs");
    HIPCHK(launch_vg_neighbours(ctx->mk_geo.as<MaskPatch>(), lo[ir], np, G, 1, nullptr,
                                ctx->vg_start.as<unsigned long long>(), ctx->vg_pairs.as<VgPair>(), s),
           "listing neighbours");
    HIPCHK(launch_vg_solap(ctx->mk_geo.as<MaskPatch>(), lo[ir], np, ctx->vg_start.as<unsigned long long>(),
                           ctx->vg_pairs.as<VgPair>(), solap, s1p, s2p, ctx->mk_err.as<int>(), ctx->cus, s),
           "overlap mask");
    // the next level reuses the offsets and pairs
    HIPCHK(hipStreamSynchronize(s), "overlap mask");
  }
  return mask_error(ctx, s);
}

// Unformatted sequential records as the reference's compiler writes them:
// a 4-byte length before and after the data.  Records longer than
// 2,147,483,639 bytes are split into subrecords whose leading marker is
// negative when another subrecord follows and whose trailing marker is
// negative when one precedes (the convention of gfortran's runtime).
namespace {
struct DumpWriter {
  FILE* f = nullptr;
  hipStream_t s = nullptr;
  char* stage = nullptr;   // pinned staging buffer
  size_t cap = 0;
  bool ok = true;
  static constexpr long long kSub = 2147483639ll;
  long long rec_left = 0, sub_left = 0, sub_len = 0;
  bool first_sub = true;
  void marker(int32_t v) { ok = ok && fwrite(&v, 4, 1, f) == 1; }
  void begin(long long bytes) {
    rec_left = bytes;
    first_sub = true;
    open_sub();
    if (bytes == 0) close_sub();
  }
  void open_sub() {
    sub_len = std::min(rec_left, kSub);
    sub_left = sub_len;
    marker((int32_t)(rec_left > kSub ? -sub_len : sub_len));
  }
  void close_sub() {
    marker((int32_t)(first_sub ? sub_len : -sub_len));
    first_sub = false;
  }
  void put(const char* p, long long n) {
    while (n > 0 && ok) {
      const long long k = std::min(n, sub_left);
      ok = ok && fwrite(p, 1, (size_t)k, f) == (size_t)k;
      p += k;
      n -= k;
      sub_left -= k;
      rec_left -= k;
      if (sub_left == 0) {
        close_sub();
        if (rec_left > 0) open_sub();
      }
    }
  }
};
}  // namespace

asohf_status asohf_hip_write_density_dump(asohf_ctx* ctx, const char* path, const asohf_geom* geom, const float* u1,
                                          const int32_t* cr0amr, const float* u11, const int32_t* cr0amr11,
                                          const int32_t* solap, const asohf_opts* opts) {
  if (!ctx) return ASOHF_ERR_INVALID;
  ctx->err.clear();
  if (!path || !geom || !u1 || !cr0amr) return fail(ctx, ASOHF_ERR_INVALID, "NULL argument");
  std::vector<MaskPatch> mp;
  std::vector<int> lo, ln;
  long long s1p, s2p, cells11, b1, b2;
  asohf_status st = mask_patches(ctx, geom, nullptr, nullptr, nullptr, nullptr, opts, mp, lo, ln, s1p, s2p, cells11);
  if (st != ASOHF_OK) return st;
  if ((st = base_ld(ctx, geom, opts, b1, b2)) != ASOHF_OK) return st;
  if (!mp.empty() && (!u11 || !cr0amr11 || !solap)) return fail(ctx, ASOHF_ERR_INVALID, "NULL patch arrays");
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  DumpWriter W;
  W.s = opts ? (hipStream_t)opts->stream : (hipStream_t)0;
  W.cap = (size_t)64 << 20;
  HIPCHK(hipHostMalloc((void**)&W.stage, W.cap, hipHostMallocDefault), "allocating the dump staging buffer");
  W.f = fopen(path, "wb");
  if (!W.f) {
    (void)hipHostFree(W.stage);
    return fail(ctx, ASOHF_ERR_INVALID, std::string("cannot open ") + path);
  }
  std::vector<char> fbuf((size_t)8 << 20);
  setvbuf(W.f, fbuf.data(), _IOFBF, fbuf.size());
  hipError_t he = hipSuccess;
  auto fetch = [&](const void* src, size_t bytes) {
    he = hipMemcpyAsync(W.stage, src, bytes, hipMemcpyDeviceToHost, W.s);
    if (he == hipSuccess) he = hipStreamSynchronize(W.s);
    return he == hipSuccess;
  };
  const long long nx = geom->nx, ny = geom->ny, nz = geom->nz;
  // U1 and CR0AMR: interiors, x fastest, one z-plane (or several) at a time
  for (int which = 0; which < 2 && W.ok && he == hipSuccess; ++which) {
    const char* base = which == 0 ? (const char*)u1 : (const char*)cr0amr;
    W.begin(4 * nx * ny * nz);
    const bool packed = b1 == nx && b2 == nx * ny;
    if (packed) {
      const long long total = 4 * nx * ny * nz;
      for (long long o = 0; o < total && W.ok; o += (long long)W.cap) {
        const long long k = std::min<long long>(W.cap, total - o);
        if (!fetch(base + o, (size_t)k)) break;
        W.put(W.stage, k);
      }
    } else {
      const long long plane = 4 * b2;
      if ((size_t)plane > W.cap) {
        he = hipErrorInvalidValue;
        break;
      }
      for (long long kz = 0; kz < nz && W.ok; ++kz) {
        if (!fetch(base + kz * plane, (size_t)plane)) break;
        for (long long jy = 0; jy < ny; ++jy) W.put(W.stage + 4 * b1 * jy, 4 * nx);
      }
    }
  }
  // per patch: U11, CR0AMR11, SOLAP, in batches of patches that fit a third
  // of the staging buffer each
  const size_t third = W.cap / 3;
  size_t p = 0;
  while (p < mp.size() && W.ok && he == hipSuccess) {
    const long long first = mp[p].off;
    size_t q = p;
    long long span = 0;
    while (q < mp.size()) {
      const long long cells = s1p > 0 ? s2p * (opts ? opts->u11_ld[2] : 1)
                                      : (long long)mp[q].n[0] * mp[q].n[1] * mp[q].n[2];
      const long long end = mp[q].off + cells - first;
      if ((size_t)(4 * end) > third) break;
      span = end;
      ++q;
    }
    if (q == p) {
      he = hipErrorInvalidValue;   // one patch larger than the staging buffer
      break;
    }
    const char* src[3] = {(const char*)(u11 + first), (const char*)(cr0amr11 + first), (const char*)(solap + first)};
    for (int a = 0; a < 3 && he == hipSuccess; ++a) {
      he = hipMemcpyAsync(W.stage + a * third, src[a], (size_t)(4 * span), hipMemcpyDeviceToHost, W.s);
    }
    if (he == hipSuccess) he = hipStreamSynchronize(W.s);
    if (he != hipSuccess) break;
    for (size_t r = p; r < q && W.ok; ++r) {
      const MaskPatch& m = mp[r];
      const long long rel = m.off - first;
      for (int a = 0; a < 3; ++a) {
        const char* b = W.stage + a * third + 4 * rel;
        W.begin(4ll * m.n[0] * m.n[1] * m.n[2]);
        if (s1p > 0) {
          for (int kz = 0; kz < m.n[2]; ++kz)
            for (int jy = 0; jy < m.n[1]; ++jy) W.put(b + 4 * (s1p * jy + s2p * kz), 4ll * m.n[0]);
        } else {
          W.put(b, 4ll * m.n[0] * m.n[1] * m.n[2]);
        }
      }
    }
    p = q;
  }
  const bool closed = fclose(W.f) == 0;
  (void)hipHostFree(W.stage);
  if (he != hipSuccess) return hip_fail(ctx, he, "copying the fields for the dump");
  if (!W.ok || !closed) return fail(ctx, ASOHF_ERR_INVALID, std::string("write error on ") + path);
  return ASOHF_OK;
}

// ---------------------------------------------------------------------------
// The mesh builder (CREATE_MESH)
// ---------------------------------------------------------------------------
asohf_status asohf_hip_create_mesh(asohf_ctx* ctx, const asohf_geom* base, int64_t n, const float* rx,
                                   const float* ry, const float* rz, const int32_t* npart_esp,
                                   const asohf_mesh_params* prm, int32_t* nl_out, int32_t* npatch_out,
                                   int32_t* patchnx, int32_t* patchny, int32_t* patchnz, int32_t* patchx,
                                   int32_t* patchy, int32_t* patchz, int32_t* pare, float* patchrx, float* patchry,
                                   float* patchrz, int64_t cap, const asohf_opts* opts) {
  if (!ctx) return ASOHF_ERR_INVALID;
  ctx->err.clear();
  if (!base || !prm || !npart_esp || !nl_out || !npatch_out || n < 0 || (n > 0 && (!rx || !ry || !rz)))
    return fail(ctx, ASOHF_ERR_INVALID, "NULL argument");
  if (cap > 0 && (!patchnx || !patchny || !patchnz || !patchx || !patchy || !patchz || !pare || !patchrx ||
                  !patchry || !patchrz))
    return fail(ctx, ASOHF_ERR_INVALID, "NULL patch output");
  if (base->nx < 3 || base->ny < 3 || base->nz < 3) return fail(ctx, ASOHF_ERR_INVALID, "grid below 3 cells");
  if (prm->nl_mesh < 1 || prm->nl_mesh > kNLevels) return fail(ctx, ASOHF_ERR_INVALID, "NL_MESH outside 1..9");
  for (int a = 0; a < 3; ++a)
    if (prm->namr[a] < 4) return fail(ctx, ASOHF_ERR_INVALID, "NAMRX..Z below 4");
  if (prm->npalev < 1) return fail(ctx, ASOHF_ERR_INVALID, "NPALEV below 1");
  if (n >= (1ll << 31)) return fail(ctx, ASOHF_ERR_CAPACITY, "2^31 or more particles");
  MeshParams pr;
  pr.nl_mesh = prm->nl_mesh;
  pr.refine_thr = prm->refine_thr;
  pr.min_patchsize = prm->min_patchsize;
  pr.frac_refinable = prm->frac_refinable;
  pr.bor = prm->bor;
  pr.boramr = prm->boramr;
  pr.bor_ovlp = prm->bor_ovlp;
  for (int a = 0; a < 3; ++a) pr.namr[a] = prm->namr[a];
  pr.npalev = prm->npalev;
  for (int l = 0; l < kNLevels; ++l) pr.mpapolev[l] = prm->mpapolev[l];
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  hipStream_t s = opts ? (hipStream_t)opts->stream : (hipStream_t)0;
  const int nx = base->nx, ny = base->ny, nz = base->nz;
  const float lado = base->lado0 - base->lado0 / (float)nx;
  std::vector<float> rad[3];
  const float* given[3] = {base->radx, base->rady, base->radz};
  const int nn[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a) {
    if (given[a]) rad[a].assign(given[a], given[a] + nn[a] + 2);
    else malla(nn[a], lado, rad[a]);
  }
  Parts A;
  memset(&A, 0, sizeof(A));
  A.x = rx;
  A.y = ry;
  A.z = rz;
  A.m = rx;   // never read
  A.n = n;
  long long acc = 0;
  for (int k = 0; k < kNEsp; ++k) {
    if (npart_esp[k] < 0) return fail(ctx, ASOHF_ERR_INVALID, "negative NPART_ESP entry");
    acc += npart_esp[k];
    A.sp[k + 1] = acc;
  }
  // ---- level 1: base-grid counts on the GPU, growth on the host ----
  const long long nc = (long long)nx * ny * nz;
  HIPCHK(ctx->mk_geo.ensure(sizeof(int) * (size_t)nc), "allocating the base counts");
  int* dcount = ctx->mk_geo.as<int>();
  HIPCHK(launch_fill_i32(dcount, nc, 0, ctx->cus, s), "clearing the base counts");
  if (n > 0)
    HIPCHK(launch_mesh_count_base(A, nx, ny, nz, base->dx, base->dy, base->dz, pr.refine_thr, dcount, ctx->cus, s),
           "base-grid counts");
  std::vector<int> conta((size_t)nc);
  HIPCHK(hipMemcpyAsync(conta.data(), dcount, sizeof(int) * (size_t)nc, hipMemcpyDeviceToHost, s),
         "reading the base counts");
  HIPCHK(hipStreamSynchronize(s), "base-grid counts");
  std::vector<MeshPatchHost> all;
  std::string e = mesh_first_level(nx, ny, nz, rad[0].data(), rad[1].data(), rad[2].data(), conta, pr, all);
  if (!e.empty()) return fail(ctx, ASOHF_ERR_INVALID, e);
  std::vector<int> npatch(kNLevels + 1, 0);
  npatch[1] = (int)all.size();
  const float rad1[3] = {rad[0][1], rad[1][1], rad[2][1]};
  // ---- levels 2..NL_MESH: counts in every patch of the level above ----
  for (int irpa = 1; irpa <= pr.nl_mesh - 1; ++irpa) {
    if (npatch[irpa] == 0) break;
    const int total = (int)all.size();
    const int first = total - npatch[irpa];
    // the geometry so far, for the level hash of the density plan
    std::vector<int32_t> np_(irpa + 1, 0), gx(total), gy(total), gz(total);
    std::vector<float> fx(total), fy(total), fz(total);
    for (int l = 1; l <= irpa; ++l) np_[l] = npatch[l];
    for (int p = 0; p < total; ++p) {
      gx[p] = all[p].n[0];
      gy[p] = all[p].n[1];
      gz[p] = all[p].n[2];
      fx[p] = all[p].pr[0];
      fy[p] = all[p].pr[1];
      fz[p] = all[p].pr[2];
    }
    asohf_geom g = *base;
    g.nl = irpa;
    g.npatch = np_.data();
    g.patchnx = gx.data();
    g.patchny = gy.data();
    g.patchnz = gz.data();
    g.patchrx = fx.data();
    g.patchry = fy.data();
    g.patchrz = fz.data();
    asohf_opts o = default_opts();
    o.mode = ASOHF_MODE_TSC;
    asohf_status st = check_geom(ctx, &g, ASOHF_MODE_TSC);
    if (st != ASOHF_OK) return st;
    st = build_plan(ctx, &g, ASOHF_MODE_TSC, &o);
    if (st != ASOHF_OK) return st;
    std::vector<MeshPatchDev> md((size_t)total);
    std::vector<long long> off((size_t)npatch[irpa] + 1, 0);
    for (int p = 0; p < total; ++p) {
      for (int a = 0; a < 3; ++a) {
        md[p].n[a] = p >= first ? all[p].n[a] : 0;   // only this level's patches are counted into
        md[p].pr[a] = all[p].pr[a];
      }
      md[p].off = 0;
      if (p >= first) {
        md[p].off = off[p - first];
        off[p - first + 1] = off[p - first] + (long long)all[p].n[0] * all[p].n[1] * all[p].n[2];
      }
    }
    const long long cells = off[npatch[irpa]];
    HIPCHK(ctx->vg_pairs.ensure(sizeof(MeshPatchDev) * md.size()), "allocating the patch records");
    HIPCHK(hipMemcpyAsync(ctx->vg_pairs.p, md.data(), sizeof(MeshPatchDev) * md.size(), hipMemcpyHostToDevice, s),
           "uploading the patch records");
    HIPCHK(ctx->mk_geo.ensure(sizeof(int) * (size_t)std::max(cells, 1ll)), "allocating the patch counts");
    dcount = ctx->mk_geo.as<int>();
    HIPCHK(launch_fill_i32(dcount, cells, 0, ctx->cus, s), "clearing the patch counts");
    float p2 = 1.0f;
    for (int k = 0; k < irpa; ++k) p2 *= 2.0f;
    if (n > 0)
      HIPCHK(launch_mesh_count_level(ctx->plan.dev, A, irpa, ctx->vg_pairs.as<MeshPatchDev>(), base->dx / p2,
                                     base->dy / p2, base->dz / p2, pr.refine_thr, dcount, ctx->cus, s),
             "patch counts");
    std::vector<int> c11((size_t)cells);
    HIPCHK(hipMemcpyAsync(c11.data(), dcount, sizeof(int) * (size_t)cells, hipMemcpyDeviceToHost, s),
           "reading the patch counts");
    HIPCHK(hipStreamSynchronize(s), "patch counts");
    // BORAMR border (grids.f:391-401)
    for (int k = 0; k < npatch[irpa]; ++k) {
      const MeshPatchHost& P = all[(size_t)(first + k)];
      for (int kz = 1; kz <= P.n[2]; ++kz)
        for (int jy = 1; jy <= P.n[1]; ++jy)
          for (int ix = 1; ix <= P.n[0]; ++ix)
            if (ix <= pr.boramr || ix >= P.n[0] - pr.boramr + 1 || jy <= pr.boramr || jy >= P.n[1] - pr.boramr + 1 ||
                kz <= pr.boramr || kz >= P.n[2] - pr.boramr + 1)
              c11[(size_t)off[k] + (ix - 1) + (size_t)P.n[0] * ((jy - 1) + (size_t)P.n[1] * (kz - 1))] = 0;
    }
    std::vector<MeshPatchHost> next;
    e = mesh_next_level(irpa, all, first, npatch[irpa], c11, off, base->dx, base->dy, base->dz, rad1, pr, total,
                        next);
    if (!e.empty()) return fail(ctx, ASOHF_ERR_CAPACITY, e);
    npatch[irpa + 1] = (int)next.size();
    all.insert(all.end(), next.begin(), next.end());
  }
  int nl = pr.nl_mesh;
  for (int l = 1; l <= pr.nl_mesh; ++l)
    if (npatch[l] == 0) {
      nl = l - 1;
      break;
    }
  *nl_out = nl;
  for (int l = 0; l <= kNLevels; ++l) npatch_out[l] = npatch[l];
  if ((int64_t)all.size() > cap) return fail(ctx, ASOHF_ERR_CAPACITY, "more patches than cap (npatch set)");
  for (size_t p = 0; p < all.size(); ++p) {
    patchnx[p] = all[p].n[0];
    patchny[p] = all[p].n[1];
    patchnz[p] = all[p].n[2];
    patchx[p] = all[p].pc[0];
    patchy[p] = all[p].pc[1];
    patchz[p] = all[p].pc[2];
    pare[p] = all[p].pare;
    patchrx[p] = all[p].pr[0];
    patchry[p] = all[p].pr[1];
    patchrz[p] = all[p].pr[2];
  }
  return ASOHF_OK;
}

asohf_status asohf_hip_release_workspaces(asohf_ctx* ctx) {
  if (!ctx) return ASOHF_ERR_INVALID;
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  HIPCHK(hipStreamSynchronize(ctx->stream), "waiting for the context's stream");
  if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
  ctx->gexec = nullptr;
  DBuf* bufs[] = {&ctx->pool, &ctx->pool_off, &ctx->pool_nk, &ctx->blk_pairs, &ctx->blk_mmax, &ctx->nsplit,
                  &ctx->item_off, &ctx->split_off, &ctx->cnt, &ctx->start, &ctx->cursor, &ctx->bsum,
                  &ctx->stiles, &ctx->isplit, &ctx->srank, &ctx->nitems, &ctx->split, &ctx->tmass, &ctx->entries,
                  &ctx->items, &ctx->scratch, &ctx->flags, &ctx->idx, &ctx->h_in[0], &ctx->h_in[1],
                  &ctx->h_in[2], &ctx->h_in[3], &ctx->h_out[0], &ctx->h_out[1], &ctx->h_mass, &ctx->ord[0],
                  &ctx->ord[1], &ctx->ord[2], &ctx->ord[3], &ctx->ord_tmp, &ctx->elist, &ctx->eoff, &ctx->ecount,
                  &ctx->sp_k[0], &ctx->sp_k[1], &ctx->sp_k[2], &ctx->sp_k[3], &ctx->sp_dens, &ctx->sp_cnt,
                  &ctx->sp_tmp, &ctx->pk_r[0], &ctx->pk_r[1], &ctx->pk_k[0], &ctx->pk_k[1], &ctx->pk_tmp,
                  &ctx->pk_cnt, &ctx->pk_geo, &ctx->mk_geo, &ctx->mk_err, &ctx->vg_cnt, &ctx->vg_start,
                  &ctx->vg_pairs};
  for (DBuf* b : bufs) b->release();
  ctx->ipairs.release();
  ctx->sized = false;   // the next batched call sizes its buffers again
  ctx->hier_sized = false;
  return ASOHF_OK;
}

asohf_status asohf_hip_get_stats(const asohf_ctx* cctx, asohf_stats* out) {
  if (!cctx || !out) return ASOHF_ERR_INVALID;
  asohf_ctx* ctx = const_cast<asohf_ctx*>(cctx);
  if (ctx->stats_pending) {
    HIPCHK(hipEventSynchronize(ctx->ev[6]), "waiting for the call");
    auto ms = [&](int a, int b, double* out) -> hipError_t {
      float v = 0;
      hipError_t e = hipEventElapsedTime(&v, ctx->ev[a], ctx->ev[b]);
      *out = v;
      return e;
    };
    HIPCHK(ms(0, 6, &ctx->stats.ms_total), "timing");
    if (ctx->stats_phases) {
      double t4 = 0;
      HIPCHK(ms(0, 1, &ctx->stats.ms_bin), "timing");
      HIPCHK(ms(2, 3, &ctx->stats.ms_scatter), "timing");
      HIPCHK(ms(4, 8, &ctx->stats.ms_deposit), "timing");
      HIPCHK(ms(8, 5, &ctx->stats.ms_split), "timing");
      HIPCHK(ms(3, 4, &t4), "timing");
      double t6 = 0;
      HIPCHK(ms(5, 6, &t6), "timing");
      ctx->stats.ms_sums = t4 + t6;
    }
    ctx->stats_pending = false;
  }
  *out = ctx->stats;
  return ASOHF_OK;
}

}  // extern "C"
