// masks.hip -- the grid masks the density dump carries (asohf.f:515-545):
//   CR0AMR / CR0AMR11  COMPUTE_CR0AMR grids.f:2076-2226: 1 for cells that are
//                      not refined, 0 for mother cells of finer patches
//   SOLAP              VEINSGRID grids.f:2229-2506: 1, or 0 where a patch
//                      cell is overlapped by another patch of its level and
//                      the other one keeps it
//
// COMPUTE_CR0AMR's marks are idempotent zeros (a cell is 0 when it is the
// mother of a child patch's cell, or its centre lies between the first and
// last cell centres of a finer patch that is not its child), so the kernels
// write them in any order: one workgroup per child patch for the mothers,
// one workgroup per finer patch for the cross marks, each cell tested with
// the reference's FP32 expressions.
//
// VEINSGRID resolves overlaps sequentially: patches I ascending, for each
// its neighbours J ascending (boxes on the level's lattice that intersect),
// for each shared cell: if I and J both still keep it, I drops it when the
// cell is on I's border, else J drops it.  A shared cell's outcome depends
// only on the patches that cover that cell and on its border status in each,
// so every (patch, cell) can replay that short sequence on its own: the
// thread of cell X in patch I lists the patches covering X (I and those
// neighbours whose overlap range, from the reference's CORN expressions for
// the pair (I, J), holds X), replays the pairs in the reference's order, and
// writes I's outcome.
#include <hip/hip_runtime.h>

#include "launch.h"

namespace asohf {

// ---------------------------------------------------------------------------
// COMPUTE_CR0AMR
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) cr_fill(int* p, long long n, int v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// base interior 1..nx x 1..ny x 1..nz with leading dims ld1, ld2
__global__ void __launch_bounds__(256) cr_fill_base(int* p, int nx, int ny, int nz, long long ld1, long long ld2,
                                                    int v) {
  const long long total = (long long)nx * ny * nz;
  for (long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x; r < total;
       r += (long long)gridDim.x * blockDim.x) {
    const long long a = r % nx, q = r / nx;
    p[a + ld1 * (q % ny) + ld2 * (q / ny)] = v;
  }
}

// Every cell of every patch of level >= 1 marks its mother cell: level 1 on
// the base grid (the loop of grids.f:2123-2135 steps by two, which marks the
// same cells), deeper levels in PARE (grids.f:2138-2160).  A mother outside
// its grid raises the error flag instead of writing.
__global__ void __launch_bounds__(256) cr_mothers(const MaskPatch* mp, int np, int nlev1, int* cr0, int nx, int ny,
                                                  int nz, long long b1, long long b2, int* cr11, long long s1p,
                                                  long long s2p, int* err) {
  for (int p = blockIdx.x; p < np; p += gridDim.x) {
    const MaskPatch P = mp[p];
    const int q = P.pare - 1;
    int qn[3] = {nx, ny, nz};
    long long qoff = 0, q1 = b1, q2 = b2;
    if (p >= nlev1) {
      if (q < 0 || q >= np) {
        if (threadIdx.x == 0) atomicOr(err, 1);
        continue;
      }
      const MaskPatch Q = mp[q];
      qn[0] = Q.n[0];
      qn[1] = Q.n[1];
      qn[2] = Q.n[2];
      qoff = Q.off;
      q1 = s1p > 0 ? s1p : Q.n[0];
      q2 = s2p > 0 ? s2p : (long long)Q.n[0] * Q.n[1];
    }
    int* dst = p >= nlev1 ? cr11 : cr0;
    const int cells = P.n[0] * P.n[1] * P.n[2];
    for (int l = threadIdx.x; l < cells; l += blockDim.x) {
      const int ix = l % P.n[0] + 1, jy = (l / P.n[0]) % P.n[1] + 1, kz = l / (P.n[0] * P.n[1]) + 1;
      const int c1 = P.pc[0] - 1 + (ix + 1) / 2, c2 = P.pc[1] - 1 + (jy + 1) / 2, c3 = P.pc[2] - 1 + (kz + 1) / 2;
      if (c1 < 1 || c1 > qn[0] || c2 < 1 || c2 > qn[1] || c3 < 1 || c3 > qn[2]) {
        atomicOr(err, 2);
        continue;
      }
      dst[qoff + (c1 - 1) + q1 * (c2 - 1) + q2 * (c3 - 1)] = 0;
    }
  }
}

// Cells of level ir inside a non-child patch of level ir+1 (grids.f:
// 2162-2224).  One workgroup per finer patch II; its threads first list the
// level-ir patches whose cell centres can reach II's box (a test widened by
// two cells), then mark, per listed patch, the cells in a window around the
// box whose centres pass the reference's exact test.
constexpr int kCrList = 512;
constexpr int kCrGroup = 4;   // fine patches tested together against each coarse patch
__global__ void __launch_bounds__(256) cr_cross(const MaskPatch* mp, int lo, int n_ir, int lo2, int n_ir1,
                                                float dxpa, float dypa, float dzpa, int* cr11, long long s1p,
                                                long long s2p, int* err) {
  __shared__ int s_list[kCrGroup][kCrList];
  __shared__ int s_n[kCrGroup];
  const float d[3] = {dxpa, dypa, dzpa};
  const int ngroups = (n_ir1 + kCrGroup - 1) / kCrGroup;
  // a workgroup takes kCrGroup consecutive fine patches at a time, so each
  // coarse patch is loaded and its box formed once for all of them
  for (int grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    const int wg = grp * kCrGroup;
    const int gn = min(kCrGroup, n_ir1 - wg);
    float b1[kCrGroup][3], b2[kCrGroup][3];
    int par[kCrGroup];
#pragma unroll
    for (int g = 0; g < kCrGroup; ++g) {
      const MaskPatch F = mp[lo2 + wg + min(g, gn - 1)];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        // XX1 = PATCHRX(II) - 0.5*(DXPA*0.5); XX2 = XX1 + (PATCHNX(II)-1)*DXPA*0.5
        b1[g][a] = F.pr[a] - 0.5f * (d[a] * 0.5f);
        b2[g][a] = b1[g][a] + ((float)(F.n[a] - 1) * d[a]) * 0.5f;
      }
      par[g] = F.pare;
    }
    if (threadIdx.x < kCrGroup) s_n[threadIdx.x] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < n_ir; i += blockDim.x) {
      const MaskPatch P = mp[lo + i];
      float c0[3], c1[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        c0[a] = P.pr[a] - 0.5f * d[a];
        c1[a] = c0[a] + (float)(P.n[a] - 1) * d[a];
      }
#pragma unroll
      for (int g = 0; g < kCrGroup; ++g) {
        if (g >= gn || P.self == par[g]) continue;   // children were marked as mothers
        bool hit = true;
#pragma unroll
        for (int a = 0; a < 3; ++a)
          hit = hit && (c1[a] >= b1[g][a] - 2.0f * d[a]) && (c0[a] <= b2[g][a] + 2.0f * d[a]);
        if (hit) {
          const int k = atomicAdd(&s_n[g], 1);
          if (k < kCrList) s_list[g][k] = lo + i;
          else atomicOr(err, 4);
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int g = 0; g < kCrGroup; ++g) {   // unrolled: the boxes stay in registers
      if (g >= gn) break;
      const int nl = min(s_n[g], kCrList);
      for (int k = 0; k < nl; ++k) {
        const MaskPatch P = mp[s_list[g][k]];
        int w0[3], wn[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const float c0 = P.pr[a] - 0.5f * d[a];
          int i0 = (int)floorf((b1[g][a] - c0) / d[a]) - 1, i1 = (int)ceilf((b2[g][a] - c0) / d[a]) + 3;
          i0 = max(i0 + 1, 1);
          i1 = min(i1, P.n[a]);
          w0[a] = i0;
          wn[a] = max(i1 - i0 + 1, 0);
        }
        const long long t1 = s1p > 0 ? s1p : P.n[0], t2 = s2p > 0 ? s2p : (long long)P.n[0] * P.n[1];
        const int cells = wn[0] * wn[1] * wn[2];
        for (int l = threadIdx.x; l < cells; l += blockDim.x) {
          const int ix = w0[0] + l % wn[0], jy = w0[1] + (l / wn[0]) % wn[1], kz = w0[2] + l / (wn[0] * wn[1]);
          const int idx[3] = {ix, jy, kz};
          bool in = true;
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            // XX = PATCHRX(I) - 0.5*DXPA + (IX-1)*DXPA; (XX-XX1)*(XX2-XX) >= 0
            const float xx = (P.pr[a] - 0.5f * d[a]) + (float)(idx[a] - 1) * d[a];
            in = in && ((xx - b1[g][a]) * (b2[g][a] - xx) >= 0.0f);
          }
          if (in) cr11[P.off + (ix - 1) + t1 * (jy - 1) + t2 * (kz - 1)] = 0;
        }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// VEINSGRID
// ---------------------------------------------------------------------------
// IG1..KG2 on the level's lattice and RX1, RX2 (per axis) of one patch
__device__ __forceinline__ void vg_box(const MaskPatch& P, const float d[3], const float fix[3], int g[6],
                                       float r1[3], float r2[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    r1[a] = P.pr[a] - 0.5f * d[a];
    r2[a] = (P.pr[a] - 0.5f * d[a]) + (float)(P.n[a] - 1) * d[a];
    g[a] = (int)(((r1[a] - fix[a]) / d[a]) + 0.5f) + 1;
    g[3 + a] = g[a] + P.n[a] - 1;
  }
}

constexpr int kVgTile = 256;   // threads of vg_neighbours, and candidates per LDS tile

// Neighbours of each level patch (J != I, lattice boxes intersecting), in
// ascending J.  Pass 0 counts them; pass 1 writes them with the pair's
// overlap from the reference's CORN expressions (grids.f:2372-2453): per
// axis I's first shared cell c1, J's first cc1, and the count - 1 (len).
__global__ void __launch_bounds__(256) vg_neighbours(const MaskPatch* mp, int lo, int np, VgGeo G, int pass,
                                                     unsigned* count, const unsigned long long* start,
                                                     VgPair* out) {
  // candidates J come in tiles of kVgTile: each one's lattice box is computed
  // once per workgroup into LDS and tested by every thread; a pair that
  // overlaps recomputes J's box in full for the overlap expressions.  The
  // tiles and the candidates within them run in ascending J.
  __shared__ int sbox[6][kVgTile];
  const int k = blockIdx.x * kVgTile + threadIdx.x;
  const bool active = k < np;
  int gi[6] = {0, 0, 0, 0, 0, 0}, gj[6];
  float r1[3] = {0.f, 0.f, 0.f}, r2[3] = {0.f, 0.f, 0.f}, q1[3], q2[3];
  if (active) vg_box(mp[lo + k], G.d, G.fix, gi, r1, r2);
  unsigned c = 0;
  const unsigned long long w = pass && active ? start[k] : 0;
  for (int j0 = 0; j0 < np; j0 += kVgTile) {
    __syncthreads();   // every thread is done with the previous tile
    if (j0 + (int)threadIdx.x < np) {
      int gt[6];
      float t1[3], t2[3];
      vg_box(mp[lo + j0 + threadIdx.x], G.d, G.fix, gt, t1, t2);
#pragma unroll
      for (int a = 0; a < 6; ++a) sbox[a][threadIdx.x] = gt[a];
    }
    __syncthreads();
    if (!active) continue;
    const int jn = min(kVgTile, np - j0);
    for (int t = 0; t < jn; ++t) {
      const int j = j0 + t;
      if (j == k) continue;
      if (!(gi[0] <= sbox[3][t] && sbox[0][t] <= gi[3] && gi[1] <= sbox[4][t] && sbox[1][t] <= gi[4] &&
            gi[2] <= sbox[5][t] && sbox[2][t] <= gi[5]))
        continue;
      if (!pass) {
        ++c;
        continue;
      }
      const MaskPatch Q = mp[lo + j];
      vg_box(Q, G.d, G.fix, gj, q1, q2);
      VgPair v;
      v.j = j;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const float d = G.d[a];
        int c1 = 1, cc1 = 1, cc2 = 1;
        if (q1[a] >= r1[a] && q2[a] <= r2[a]) {
          c1 = (int)(((q1[a] - r1[a]) / d) + 0.5f) + 1;
          cc1 = 1;
          cc2 = Q.n[a];
        }
        if (q1[a] >= r1[a] && q2[a] > r2[a]) {
          c1 = (int)(((q1[a] - r1[a]) / d) + 0.5f) + 1;
          cc1 = 1;
          cc2 = (int)(((r2[a] - q1[a]) / d) + 0.5f) + 1;
        }
        if (q2[a] <= r2[a] && q1[a] < r1[a]) {
          c1 = 1;
          cc1 = (int)(((r1[a] - q1[a]) / d) + 0.5f) + 1;
          cc2 = Q.n[a];
        }
        if (q1[a] < r1[a] && q2[a] > r2[a]) {
          c1 = 1;
          cc1 = (int)(((r1[a] - q1[a]) / d) + 0.5f) + 1;
          cc2 = (int)(((r2[a] - q1[a]) / d) + 0.5f) + 1;
        }
        // the reference walks J's cells cc1..cc2 and maps each to I's
        // c1 + (jj - cc1)
        v.c1[a] = c1;
        v.cc1[a] = cc1;
        v.len[a] = cc2 - cc1;
      }
      out[w + c] = v;
      ++c;
    }
  }
  if (active && !pass) count[k] = c;
}

constexpr int kVgCover = 16;   // patches of one level covering one cell
__global__ void __launch_bounds__(256) vg_solap(const MaskPatch* mp, int lo, int np, const unsigned long long* start,
                                                const VgPair* pairs, int* solap, long long s1p, long long s2p,
                                                int* err) {
  for (int k = blockIdx.x; k < np; k += gridDim.x) {
    const MaskPatch P = mp[lo + k];
    const unsigned long long b = start[k], e = start[k + 1];
    const long long t1 = s1p > 0 ? s1p : P.n[0], t2 = s2p > 0 ? s2p : (long long)P.n[0] * P.n[1];
    const int cells = P.n[0] * P.n[1] * P.n[2];
    for (int l = threadIdx.x; l < cells; l += blockDim.x) {
      const int x[3] = {l % P.n[0] + 1, (l / P.n[0]) % P.n[1] + 1, l / (P.n[0] * P.n[1]) + 1};
      // the covering patches in ascending order, with the cell's border
      // status in each; this patch is inserted at its place
      int who[kVgCover];
      bool border[kVgCover], keep[kVgCover];
      int nc = 0;
      bool placed = false, overflow = false;
      auto push = [&](int id, bool bd) {
        if (nc < kVgCover) {
          who[nc] = id;
          border[nc] = bd;
          keep[nc] = true;
          ++nc;
        } else {
          overflow = true;
        }
      };
      const bool my_border = x[0] == 1 || x[0] == P.n[0] || x[1] == 1 || x[1] == P.n[1] || x[2] == 1 || x[2] == P.n[2];
      for (unsigned long long q = b; q < e; ++q) {
        const VgPair v = pairs[q];
        if (x[0] < v.c1[0] || x[0] > v.c1[0] + v.len[0] || x[1] < v.c1[1] || x[1] > v.c1[1] + v.len[1] ||
            x[2] < v.c1[2] || x[2] > v.c1[2] + v.len[2])
          continue;
        if (!placed && v.j > k) {
          push(k, my_border);
          placed = true;
        }
        const MaskPatch Q = mp[lo + v.j];
        bool bd = false;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const int y = x[a] - v.c1[a] + v.cc1[a];
          bd = bd || y == 1 || y == Q.n[a];
        }
        push(v.j, bd);
      }
      if (!placed) push(k, my_border);
      if (overflow) atomicOr(err, 8);
      int mine = 0;
      for (int a = 0; a < nc; ++a)
        if (who[a] == k) mine = a;
      // the reference's order: I ascending, then its neighbours J ascending
      for (int a = 0; a < nc; ++a)
        for (int c = 0; c < nc; ++c) {
          if (c == a || !keep[a] || !keep[c]) continue;
          if (border[a]) keep[a] = false;
          else keep[c] = false;
        }
      solap[P.off + (x[0] - 1) + t1 * (x[1] - 1) + t2 * (x[2] - 1)] = keep[mine] ? 1 : 0;
    }
  }
}

// ---------------------------------------------------------------------------
static int grid_1d(long long n, int cus) {
  long long b = (n + 255) / 256;
  if (b > (long long)cus * 32) b = (long long)cus * 32;
  return (int)(b > 0 ? b : 1);
}

hipError_t launch_cr0amr(const MaskPatch* mp, int np, int nlev1, const int* lev_lo, const int* lev_n, int nl,
                         float lado0, int nx, int ny, int nz, int* cr0, long long b1, long long b2, int* cr11,
                         long long cells11, long long s1p, long long s2p, int* err, int cus, hipStream_t s) {
  cr_fill_base<<<grid_1d((long long)nx * ny * nz, cus), 256, 0, s>>>(cr0, nx, ny, nz, b1, b2, 1);
  if (cells11 > 0) cr_fill<<<grid_1d(cells11, cus), 256, 0, s>>>(cr11, cells11, 1);
  if (np > 0)
    cr_mothers<<<np < cus * 16 ? np : cus * 16, 256, 0, s>>>(mp, np, nlev1, cr0, nx, ny, nz, b1, b2, cr11, s1p,
                                                             s2p, err);
  // DX = LADO0/NX (grids.f:2162-2164), DXPA = DX/2**IR
  const float dx = lado0 / (float)nx, dy = lado0 / (float)ny, dz = lado0 / (float)nz;
  for (int ir = 1; ir <= nl - 1; ++ir) {
    float p2 = 1.0f;
    for (int k = 0; k < ir; ++k) p2 *= 2.0f;
    const int n1 = lev_n[ir + 1];
    if (lev_n[ir] == 0 || n1 == 0) continue;
    const int ng = (n1 + kCrGroup - 1) / kCrGroup;
    cr_cross<<<ng < cus * 16 ? ng : cus * 16, 256, 0, s>>>(mp, lev_lo[ir], lev_n[ir], lev_lo[ir + 1], n1, dx / p2,
                                                           dy / p2, dz / p2, cr11, s1p, s2p, err);
  }
  return hipGetLastError();
}

hipError_t launch_vg_neighbours(const MaskPatch* mp, int lo, int np, const VgGeo& G, int pass, unsigned* count,
                                const unsigned long long* start, VgPair* out, hipStream_t s) {
  if (np <= 0) return hipSuccess;
  vg_neighbours<<<(np + kVgTile - 1) / kVgTile, kVgTile, 0, s>>>(mp, lo, np, G, pass, count, start, out);
  return hipGetLastError();
}

hipError_t launch_vg_solap(const MaskPatch* mp, int lo, int np, const unsigned long long* start, const VgPair* pairs,
                           int* solap, long long s1p, long long s2p, int* err, int cus, hipStream_t s) {
  if (np <= 0) return hipSuccess;
  vg_solap<<<np < cus * 16 ? np : cus * 16, 256, 0, s>>>(mp, lo, np, start, pairs, solap, s1p, s2p, err);
  return hipGetLastError();
}

hipError_t launch_fill_i32(int* p, long long n, int v, int cus, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  cr_fill<<<grid_1d(n, cus), 256, 0, s>>>(p, n, v);
  return hipGetLastError();
}

}  // namespace asohf
