// mesh.hip -- the mesh builder CREATE_MESH (grids.f:2-707).
//
// The expensive part of the routine is counting particles into cells: on the
// base grid (grids.f:107-129) and, for every level IRPA, into the cells of
// each of its patches (grids.f:363-405: every patch scans every particle in
// the reference, O(P*N)).  Those counts run on the GPU: one thread per
// particle, NGP index as the reference writes it, integer atomics (exact in
// any order); on AMR levels the candidate patches come from the density
// plan's per-level spatial hash, whose TSC member boxes contain each patch's
// counted cells with a cell of margin, and the reference's FLOOR index test
// decides.
//
// What follows the counts is sequential by construction (each new patch
// changes the counts the next one sees), and is restated here on the host:
// the level-1 growth over the cells sorted as SORT_CELLS does, VEINSGRID_
// REDUCED's overlap cleaning, the growth inside each parent patch from its
// MAXLOC cell, and the assembly of a level's patches by decreasing seed
// count (MAXLOC over LVAL).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <string>
#include <thread>
#include <vector>

#include "dev.h"
#include "launch.h"

namespace asohf {

// ---------------------------------------------------------------------------
// counts on the GPU
// ---------------------------------------------------------------------------
// PLEV (grids.f:73-81): the species block a particle sits in, 0 past
// SUM(NPART_ESP)
__device__ __forceinline__ int mesh_plev(const Parts& A, long long i) {
  return i < A.sp[kNEsp] ? species_of(A, i) : 0;
}

// Base grid (grids.f:107-129): IX = INT((X-XL)/DX)+1 clamped to [1, NX],
// XL = -NX*DX/2; species 0 counts 1, the others REFINE_THR.
__global__ void __launch_bounds__(256) mesh_count_base(Parts A, int nx, int ny, int nz, float dx, float dy, float dz,
                                                       int thr, int* conta) {
  const float xl = -(((float)nx * dx) / 2.0f), yl = -(((float)ny * dy) / 2.0f), zl = -(((float)nz * dz) / 2.0f);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < A.n;
       i += (long long)gridDim.x * blockDim.x) {
    const float p[3] = {A.x[i], A.y[i], A.z[i]};
    const float lo[3] = {xl, yl, zl}, d[3] = {dx, dy, dz};
    const int nn[3] = {nx, ny, nz};
    int c[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float t = (p[a] - lo[a]) / d[a];
      int k = (t < 2.0e9f && t > -2.0e9f) ? (int)t + 1 : (t > 0.f ? nn[a] : 1);
      c[a] = k < 1 ? 1 : (k > nn[a] ? nn[a] : k);
    }
    const int w = mesh_plev(A, i) == 0 ? 1 : thr;
    atomicAdd(&conta[(c[0] - 1) + (long long)nx * ((c[1] - 1) + (long long)ny * (c[2] - 1))], w);
  }
}

// Level l patches (grids.f:363-389): IX = FLOOR((X-XL)/DXPA)+1 in [1, N1],
// XL = PATCHRX - DXPA; a particle of species <= l counts 1, else REFINE_THR.
__global__ void __launch_bounds__(256) mesh_count_level(Plan P, Parts A, int l, const MeshPatchDev* mp, float dxpa,
                                                        float dypa, float dzpa, int thr, int* conta) {
  const float d[3] = {dxpa, dypa, dzpa};
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < A.n;
       i += (long long)gridDim.x * blockDim.x) {
    const float p[3] = {A.x[i], A.y[i], A.z[i]};
    const int2 se = bucket_lookup(P, l, p[0], p[1], p[2]);
    if (se.y == 0) continue;
    const int w = mesh_plev(A, i) <= l ? 1 : thr;
    for (int r = 0; r < se.y; ++r) {
      const int q = P.hrec[se.x + r].g - 1;   // grid g = 1 + patch
      if (q < 0) continue;
      const MeshPatchDev M = mp[q];            // patches of other levels have n = 0
      int c[3];
      bool in = true;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        c[a] = (int)floorf(fminf(fmaxf((p[a] - (M.pr[a] - d[a])) / d[a], -4.0f), 1.0e6f)) + 1;
        in = in && c[a] >= 1 && c[a] <= M.n[a];
      }
      if (in)
        atomicAdd(&conta[M.off + (c[0] - 1) + (long long)M.n[0] * ((c[1] - 1) + (long long)M.n[1] * (c[2] - 1))], w);
    }
  }
}

hipError_t launch_mesh_count_base(const Parts& A, int nx, int ny, int nz, float dx, float dy, float dz, int thr,
                                  int* conta, int cus, hipStream_t s) {
  long long b = (A.n + 255) / 256;
  if (b > (long long)cus * 32) b = (long long)cus * 32;
  mesh_count_base<<<(int)(b > 0 ? b : 1), 256, 0, s>>>(A, nx, ny, nz, dx, dy, dz, thr, conta);
  return hipGetLastError();
}

hipError_t launch_mesh_count_level(const Plan& P, const Parts& A, int l, const MeshPatchDev* mp, float dxpa,
                                   float dypa, float dzpa, int thr, int* conta, int cus, hipStream_t s) {
  long long b = (A.n + 255) / 256;
  if (b > (long long)cus * 32) b = (long long)cus * 32;
  mesh_count_level<<<(int)(b > 0 ? b : 1), 256, 0, s>>>(P, A, l, mp, dxpa, dypa, dzpa, thr, conta);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// the sequential part, on the host
// ---------------------------------------------------------------------------
namespace {

// A grid's cell array, x fastest, 1-based accessor.
struct Grid3 {
  int* v;
  int n[3];
  int& at(int i, int j, int k) { return v[(size_t)(i - 1) + (size_t)n[0] * ((size_t)(j - 1) + (size_t)n[1] * (k - 1))]; }
};

// COUNT(g(i1:i2, j1:j2, k1:k2) >= thr)
long long count_ge(Grid3& g, int i1, int i2, int j1, int j2, int k1, int k2, int thr) {
  long long c = 0;
  for (int k = k1; k <= k2; ++k)
    for (int j = j1; j <= j2; ++j)
      for (int i = i1; i <= i2; ++i) c += g.at(i, j, k) >= thr;
  return c;
}

// The growth of grids.f:181-248 (and 449-558 inside a parent): faces are
// added while the fraction of refinable cells on the new face, computed as
// FLOAT(4*COUNT)/(N2*N3) in FP32, exceeds FRAC_REFINABLE and the patch stays
// below NAMR-2 cells; lo/hi are the allowed cell range per axis.
void grow(Grid3& conta, int b[6], int n[3], const int lo[3], const int hi[3], const MeshParams& pr) {
  int marca = 1;
  while (marca == 1) {
    marca = 0;
    for (int a = 0; a < 3; ++a) {
      const int o1 = (a + 1) % 3, o2 = (a + 2) % 3;
      const int na = pr.namr[a] - 2;
      for (int side = 0; side < 2; ++side) {
        const bool can = side == 0 ? b[2 * a] > lo[a] : b[2 * a + 1] < hi[a];
        if (!(n[a] <= na && can)) continue;
        int q[6] = {b[0], b[1], b[2], b[3], b[4], b[5]};
        const int plane = side == 0 ? b[2 * a] - 1 : b[2 * a + 1] + 1;
        q[2 * a] = q[2 * a + 1] = plane;
        const long long c = count_ge(conta, q[0], q[1], q[2], q[3], q[4], q[5], pr.refine_thr);
        const float bas = (float)(4 * c) / (float)(n[o1] * n[o2]);
        if (bas > pr.frac_refinable) {
          if (side == 0) b[2 * a] -= 1;
          else b[2 * a + 1] += 1;
          n[a] = 2 * (b[2 * a + 1] - b[2 * a] + 1);
          marca = 1;
        }
      }
    }
  }
}

// the BOR_OVLP-shrunk box, kept at the seed cell's side where the box starts
// or ends at it (grids.f:270-281)
void bis_box(const int b[6], const int seed[3], int ovlp, int o[6]) {
  for (int a = 0; a < 3; ++a) {
    o[2 * a] = b[2 * a] + ovlp;
    o[2 * a + 1] = b[2 * a + 1] - ovlp;
    if (b[2 * a] == seed[a]) o[2 * a] = b[2 * a];
    if (b[2 * a + 1] == seed[a]) o[2 * a + 1] = b[2 * a + 1];
  }
}

void fill_box(Grid3& g, const int b[6], int v) {
  for (int k = b[4]; k <= b[5]; ++k)
    for (int j = b[2]; j <= b[3]; ++j)
      for (int i = b[0]; i <= b[1]; ++i) g.at(i, j, k) = v;
}

}  // namespace

// Level 1 from the base counts (grids.f:131-297); conta is zeroed on the
// border here.  Returns "" or an error.
std::string mesh_first_level(int nx, int ny, int nz, const float* radx, const float* rady, const float* radz,
                             std::vector<int>& conta1, const MeshParams& pr, std::vector<MeshPatchHost>& out) {
  Grid3 ct{conta1.data(), {nx, ny, nz}};
  std::vector<int> cr0v(conta1.size());
  Grid3 cr0{cr0v.data(), {nx, ny, nz}};
  const int bor = pr.bor;
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        if (i <= bor || i >= nx - bor + 1 || j <= bor || j >= ny - bor + 1 || k <= bor || k >= nz - bor + 1)
          ct.at(i, j, k) = 0;
        cr0.at(i, j, k) = ct.at(i, j, k);
      }
  // SORT_CELLS: keys 1/FLOAT(CR0) ascending; of equal keys the later cell
  // in scan order first (ARGSORT's merge)
  std::vector<long long> cand;
  for (size_t c = 0; c < cr0v.size(); ++c)
    if (cr0v[c] >= pr.refine_thr) cand.push_back((long long)c);
  std::vector<float> key(cand.size());
  for (size_t q = 0; q < cand.size(); ++q) key[q] = 1.0f / (float)cr0v[(size_t)cand[q]];
  std::vector<size_t> ord(cand.size());
  for (size_t q = 0; q < ord.size(); ++q) ord[q] = ord.size() - 1 - q;   // scan order reversed
  std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return key[a] < key[b]; });
  const int lo[3] = {bor + 1, bor + 1, bor + 1}, hi[3] = {nx - bor, ny - bor, nz - bor};
  int ipatch = 0;
  for (size_t q : ord) {
    const long long c = cand[q];
    const int seed[3] = {(int)(c % nx) + 1, (int)((c / nx) % ny) + 1, (int)(c / ((long long)nx * ny)) + 1};
    if (cr0.at(seed[0], seed[1], seed[2]) < pr.refine_thr) continue;   // CYCLE skips the MPAPOLEV test
    {
      int b[6], n[3];
      for (int a = 0; a < 3; ++a) {
        b[2 * a] = std::max(seed[a] - 1, lo[a]);
        b[2 * a + 1] = std::min(seed[a] + 1, hi[a]);
        n[a] = 2 * (b[2 * a + 1] - b[2 * a] + 1);
      }
      grow(ct, b, n, lo, hi, pr);
      if (std::min(n[0], std::min(n[1], n[2])) < pr.min_patchsize) {
        for (int k = b[4]; k <= b[5]; ++k)
          for (int j = b[2]; j <= b[3]; ++j)
            for (int i = b[0]; i <= b[1]; ++i)
              if (cr0.at(i, j, k) > 0) cr0.at(i, j, k) = 0;
        fill_box(ct, b, 0);
      } else {
        ++ipatch;
        int o[6];
        bis_box(b, seed, pr.bor_ovlp, o);
        fill_box(ct, o, 0);
        fill_box(cr0, o, -1);
        MeshPatchHost m;
        for (int a = 0; a < 3; ++a) {
          m.n[a] = n[a];
          m.pc[a] = b[2 * a];
        }
        m.pr[0] = radx[b[0]];
        m.pr[1] = rady[b[2]];
        m.pr[2] = radz[b[4]];
        m.pare = 0;
        out.push_back(m);
      }
    }
    if (ipatch >= pr.mpapolev[0]) break;
  }
  return "";
}

// VEINSGRID_REDUCED (grids.f:1771-2073) on the level's counts (cr and
// conta per patch, packed) and then the growth inside every parent and the
// assembly of the next level (grids.f:413-646).  `lev` holds the parent
// level's patches (global numbers first+1 ..), `dx..dz` = COMMON /ESPACIADO/,
// `rad1` = RADX(1), RADY(1), RADZ(1).
std::string mesh_next_level(int irpa, const std::vector<MeshPatchHost>& all, int first, int np,
                            std::vector<int>& conta, const std::vector<long long>& off, float dx, float dy, float dz,
                            const float rad1[3], const MeshParams& pr, int total_before, std::vector<MeshPatchHost>& out) {
  float p2 = 1.0f;
  for (int k = 0; k < irpa; ++k) p2 *= 2.0f;
  const float d[3] = {dx / p2, dy / p2, dz / p2};
  std::vector<int> cr(conta);   // CR01 = CONTA11 after the borders were cleared
  // ---- VEINSGRID_REDUCED ----
  {
    const float fix[3] = {(rad1[0] - dx * 0.5f) + 0.5f * d[0], (rad1[1] - dy * 0.5f) + 0.5f * d[1],
                          (rad1[2] - dz * 0.5f) + 0.5f * d[2]};
    std::vector<int> g((size_t)6 * np);
    for (int k = 0; k < np; ++k) {
      const MeshPatchHost& P = all[(size_t)(first + k)];
      for (int a = 0; a < 3; ++a) {
        const float r1 = P.pr[a] - 0.5f * d[a];
        g[6 * k + a] = (int)(((r1 - fix[a]) / d[a]) + 0.5f) + 1;
        g[6 * k + 3 + a] = g[6 * k + a] + P.n[a] - 1;
      }
    }
    const int npalev2 = std::max(100, pr.npalev / 5);
    std::vector<char> solap(conta.size(), 0);
    for (int k = 0; k < np; ++k) {
      const MeshPatchHost& P = all[(size_t)(first + k)];
      float r1[3], r2[3];
      for (int a = 0; a < 3; ++a) {
        r1[a] = P.pr[a] - 0.5f * d[a];
        r2[a] = (P.pr[a] - 0.5f * d[a]) + (float)(P.n[a] - 1) * d[a];
      }
      int nv = 0;
      for (int j = 0; j < np; ++j) {
        if (j == k) continue;
        if (!(g[6 * k] <= g[6 * j + 3] && g[6 * j] <= g[6 * k + 3] && g[6 * k + 1] <= g[6 * j + 4] &&
              g[6 * j + 1] <= g[6 * k + 4] && g[6 * k + 2] <= g[6 * j + 5] && g[6 * j + 2] <= g[6 * k + 5]))
          continue;
        if (++nv > npalev2) return "VEINSGRID_REDUCED: too many neighbours";
        const MeshPatchHost& Q = all[(size_t)(first + j)];
        int c1[3], cc1[3], cc2[3];
        for (int a = 0; a < 3; ++a) {
          const float q1 = Q.pr[a] - 0.5f * d[a];
          const float q2 = (Q.pr[a] - 0.5f * d[a]) + (float)(Q.n[a] - 1) * d[a];
          if (q1 >= r1[a] && q2 <= r2[a]) {
            c1[a] = (int)(((q1 - r1[a]) / d[a]) + 0.5f) + 1;
            cc1[a] = 1;
            cc2[a] = Q.n[a];
          }
          if (q1 >= r1[a] && q2 > r2[a]) {
            c1[a] = (int)(((q1 - r1[a]) / d[a]) + 0.5f) + 1;
            cc1[a] = 1;
            cc2[a] = (int)(((r2[a] - q1) / d[a]) + 0.5f) + 1;
          }
          if (q2 <= r2[a] && q1 < r1[a]) {
            c1[a] = 1;
            cc1[a] = (int)(((r1[a] - q1) / d[a]) + 0.5f) + 1;
            cc2[a] = Q.n[a];
          }
          if (q1 < r1[a] && q2 > r2[a]) {
            c1[a] = 1;
            cc1[a] = (int)(((r1[a] - q1) / d[a]) + 0.5f) + 1;
            cc2[a] = (int)(((r2[a] - q1) / d[a]) + 0.5f) + 1;
          }
        }
        // fraction of J's volume in the overlap, FP32 as the reference: an
        // INTEGER product stored in REAL, then divided by NN1, NN2, NN3
        float ov = (float)((cc2[2] - cc1[2] + 1) * (cc2[1] - cc1[1] + 1) * (cc2[0] - cc1[0] + 1));
        ov = std::fabs(ov) / (float)Q.n[0] / (float)Q.n[1] / (float)Q.n[2];
        if (!(ov > 0.1f)) continue;
        for (int kk = cc1[2]; kk <= cc2[2]; ++kk)
          for (int jj = cc1[1]; jj <= cc2[1]; ++jj)
            for (int ii = cc1[0]; ii <= cc2[0]; ++ii) {
              const int ix = ii - cc1[0] + c1[0], jy = jj - cc1[1] + c1[1], kz = kk - cc1[2] + c1[2];
              if (ix < 1 || ix > P.n[0] || jy < 1 || jy > P.n[1] || kz < 1 || kz > P.n[2] || ii < 1 ||
                  ii > Q.n[0] || jj < 1 || jj > Q.n[1] || kk < 1 || kk > Q.n[2])
                return "VEINSGRID_REDUCED: an overlap maps a cell outside its patch";
              const size_t ci = (size_t)off[k] + (ix - 1) + (size_t)P.n[0] * ((jy - 1) + (size_t)P.n[1] * (kz - 1));
              const size_t cj = (size_t)off[j] + (ii - 1) + (size_t)Q.n[0] * ((jj - 1) + (size_t)Q.n[1] * (kk - 1));
              if (solap[ci] == 0) {
                solap[cj] = 1;
                cr[cj] = 0;
                conta[cj] = 0;
              }
            }
      }
    }
  }
  // ---- growth inside every parent (grids.f:449-588).  Parents are
  // independent (each touches only its own cells, as the reference's
  // OpenMP loop over IPARE assumes), so they run on host threads; their
  // patches are concatenated in parent order, as the sequential loop
  // produces them.  The refinable count is kept up to date from the boxes
  // each step overwrites, instead of recounting the parent's cells. ----
  const int npalev3 = (pr.namr[0] / 5) * (pr.namr[0] / 5) * (pr.namr[0] / 5) + 1;
  struct Cand {
    MeshPatchHost m;
    int val, parent, rank;
  };
  std::vector<std::vector<Cand>> per(np);
  auto grow_parent = [&](int k) {
    const MeshPatchHost& P = all[(size_t)(first + k)];
    Grid3 c11{cr.data() + off[k], {P.n[0], P.n[1], P.n[2]}};
    Grid3 k11{conta.data() + off[k], {P.n[0], P.n[1], P.n[2]}};
    const long long cells = (long long)P.n[0] * P.n[1] * P.n[2];
    long long refine_count = 0;
    for (long long q = 0; q < cells; ++q) refine_count += c11.v[q] >= pr.refine_thr;
    // overwrite a box of c11 with v, keeping refine_count = COUNT(c11 >= thr)
    auto set_c11 = [&](const int bx[6], int v) {
      for (int kk = bx[4]; kk <= bx[5]; ++kk)
        for (int jj = bx[2]; jj <= bx[3]; ++jj)
          for (int ii = bx[0]; ii <= bx[1]; ++ii) {
            int& cv = c11.at(ii, jj, kk);
            refine_count += (long long)(v >= pr.refine_thr) - (long long)(cv >= pr.refine_thr);
            cv = v;
          }
    };
    int ipatch = 0;
    const int lo[3] = {pr.boramr + 1, pr.boramr + 1, pr.boramr + 1};
    const int hi[3] = {P.n[0] - pr.boramr, P.n[1] - pr.boramr, P.n[2] - pr.boramr};
    while (refine_count > 0 && ipatch < npalev3) {
      // MAXLOC: the first maximum in x-fastest order (padding cells, 0 in
      // the reference's NAMR-sized array, never win: the maximum is >= thr)
      long long best = 0;
      for (long long q = 1; q < cells; ++q)
        if (c11.v[q] > c11.v[best]) best = q;
      const int seed[3] = {(int)(best % P.n[0]) + 1, (int)((best / P.n[0]) % P.n[1]) + 1,
                           (int)(best / ((long long)P.n[0] * P.n[1])) + 1};
      const int basint = c11.v[best];
      int bb[6], n[3];
      for (int ax = 0; ax < 3; ++ax) {
        bb[2 * ax] = std::max(seed[ax] - 1, lo[ax]);
        bb[2 * ax + 1] = std::min(seed[ax] + 1, hi[ax]);
        n[ax] = 2 * (bb[2 * ax + 1] - bb[2 * ax] + 1);
      }
      grow(k11, bb, n, lo, hi, pr);
      if (std::min(n[0], std::min(n[1], n[2])) < pr.min_patchsize) {
        set_c11(bb, 0);
        fill_box(k11, bb, 0);
      } else {
        ++ipatch;
        int o[6];
        bis_box(bb, seed, pr.bor_ovlp, o);
        fill_box(k11, o, 0);
        set_c11(o, -1);
        Cand cd;
        for (int ax = 0; ax < 3; ++ax) {
          cd.m.n[ax] = n[ax];
          cd.m.pc[ax] = bb[2 * ax];
          // the parent's cell size: PATCHRX(IPARE)+(I1-1.5)*DXPA
          cd.m.pr[ax] = P.pr[ax] + ((float)bb[2 * ax] - 1.5f) * d[ax];
        }
        cd.m.pare = first + k + 1;
        cd.val = basint;
        cd.parent = k;
        cd.rank = ipatch;
        per[(size_t)k].push_back(cd);
      }
    }
  };
  {
    const int nthreads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::atomic<int> next{0};
    auto worker = [&]() {
      for (int k = next.fetch_add(1); k < np; k = next.fetch_add(1)) grow_parent(k);
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < std::min(nthreads, np); ++t) pool.emplace_back(worker);
    worker();
    for (std::thread& th : pool) th.join();
  }
  std::vector<Cand> cands;
  for (int k = 0; k < np; ++k) cands.insert(cands.end(), per[(size_t)k].begin(), per[(size_t)k].end());
  // ---- assembly: MAXLOC over LVAL(I, IPARE), i.e. decreasing seed count,
  // then parent, then order within the parent (grids.f:593-621) ----
  std::stable_sort(cands.begin(), cands.end(), [](const Cand& a, const Cand& b) { return a.val > b.val; });
  const int cap = pr.mpapolev[std::min(irpa, kNLevels - 1)];   // MPAPOLEV(IR), IR = IRPA+1
  int ipatch = total_before;
  for (const Cand& c : cands) {
    if (c.val <= 0 || ipatch >= pr.npalev) break;   // COUNT(LVAL > 0) and the NPALEV bound
    ++ipatch;
    out.push_back(c.m);
    if (ipatch - total_before >= cap) break;
  }
  if (total_before + (int)out.size() == pr.npalev) return "NPALEV too small";
  return "";
}

}  // namespace asohf
