// hier.hip -- TSC binning along the patch hierarchy (gfx950).
//
// The batched TSC build bins each (particle, grid) member into the tiles its
// 27-cell stencil touches.  Instead of testing every particle against a hash
// of every level's patch boxes, the binning follows the hierarchy:
//
//   bin_base_tsc  particles -> base-grid tile entries (16 B: x, y, z, m);
//                 every particle is a member of the base grid
//                 (grids.f:1159-1246), one entry per base tile its stencil
//                 touches (1 to 8)
//   emit_tsc      for one level at a time: the entries of a level-l tile whose
//                 particle has its home cell in that tile (its clamped cell
//                 index lies there, so each member of a grid is visited once)
//                 are tested against the patch boxes of the tile's target
//                 list: the patches whose membership box lies inside this
//                 grid's box (found on the host in FP32, the values the
//                 device compares), so every member of such a patch is a
//                 member of this grid and has exactly one home entry here.
//                 A member gets entries in the tiles of its stencil in that
//                 patch, exactly as the reference's strict box test and index
//                 expressions give them (grids.f:1294-1323).
//
// Entries go to per-tile ranges whose capacities come from an exact count
// pass (the first call on a geometry, or after an overflow): every workgroup
// counts its emissions per tile in LDS, reserves each tile's range with one
// atomic add, and writes.  An emission that finds the LDS table or list full
// reserves its own slot directly.  A reservation past a tile's capacity sets
// the overflow flag (nothing is written past it) and the host rebuilds the
// capacities exactly.
#include "dev.h"
#include "launch.h"

namespace asohf {

namespace {

constexpr int kHbThreads = 256;
constexpr int kHbPer = 2;                        // particles per thread
constexpr int kHbChunk = kHbThreads * kHbPer;    // particles per bin_base workgroup
constexpr int kHbCap = 1024;                     // LDS key slots (power of two)
constexpr int kHbList = 8 * kHbChunk;            // at most 8 base tiles per particle: never full
constexpr int kHbProbe = 32;
constexpr int kEmThreads = 256;
constexpr int kEmRound = 1024;                   // entries staged per round of emit_tsc
constexpr int kEmCap = 1024;
constexpr int kEmList = 4096;
constexpr int kEmTargets = 32;                   // target records staged in LDS at once

__device__ __forceinline__ unsigned slot_hash(unsigned k, unsigned mask) {
  k ^= k >> 15;
  k *= 0x2c1b3c6du;
  k ^= k >> 12;
  k *= 0x297a2d39u;
  k ^= k >> 15;
  return k & mask;
}
// slot of key k in an LDS open-addressing table of CAP slots, inserting it
// when absent; -1 when the probe path is full
template <int CAP>
__device__ __forceinline__ int slot_insert(int* keys, int k) {
  unsigned h = slot_hash((unsigned)k, CAP - 1);
#pragma unroll 1
  for (int p = 0; p < kHbProbe; ++p) {
    const int cur = keys[h];
    if (cur == k) return (int)h;
    if (cur == -1) {
      const int old = atomicCAS(&keys[h], -1, k);
      if (old == -1 || old == k) return (int)h;
    }
    h = (h + 1) & (CAP - 1);
  }
  return -1;
}

__device__ __forceinline__ void set_overflow(Counters* C) {
  C->overflow = 1u;
  C->n_items_eff = 0u;   // the deposit of this call then does nothing
}

// one entry written at its own reserved slot (the path of an emission that
// found the LDS table or list full)
__device__ __forceinline__ void emit_direct(unsigned long long* cursor, const unsigned long long* start,
                                            int4* entries, Counters* C, int key, int4 e) {
  const unsigned long long p = atomicAdd(&cursor[key], 1ull);
  if (p < start[key + 1]) entries[p] = e;
  else set_overflow(C);
}

}  // namespace

// cursor[k] = start[k]: every tile's range starts empty
__global__ void __launch_bounds__(256) cursor_reset(const unsigned long long* start, unsigned long long* cursor,
                                                    long long nkeys) {
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < nkeys;
       k += (long long)gridDim.x * blockDim.x)
    cursor[k] = start[k];
}

// per-key capacity from an exact count: the count plus an eighth plus 32
// entries, so a later snapshot with a similar distribution fits
__global__ void __launch_bounds__(256) count_to_capacity(unsigned* cnt, long long nkeys) {
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < nkeys;
       k += (long long)gridDim.x * blockDim.x) {
    const unsigned c = cnt[k];
    cnt[k] = c == 0u ? 0u : c + (c >> 3) + 32u;
  }
}

__global__ void __launch_bounds__(kHbThreads) bin_base_tsc(Plan P, Parts A, unsigned long long* cursor,
                                                           const unsigned long long* start, int4* entries,
                                                           Counters* C, BlockTables BT) {
  __shared__ int keys[kHbCap];
  __shared__ unsigned vals[kHbCap];
  __shared__ unsigned long long base[kHbCap];
  __shared__ unsigned list[kHbList];
  __shared__ int4 parts[kHbChunk];
  __shared__ int s_n;
  __shared__ unsigned long long s_valid[kHbThreads / 64];
  __shared__ float s_mmax[kHbThreads / 64];
  const int tid = threadIdx.x;
  for (int k = tid; k < kHbCap; k += kHbThreads) { keys[k] = -1; vals[k] = 0u; }
  if (tid == 0) s_n = 0;
  __syncthreads();
  const GridDesc& B = P.grids[0];
  const int bn[3] = {B.n[0], B.n[1], B.n[2]};
  const float bd[3] = {B.d[0], B.d[1], B.d[2]};
  const float bid[3] = {B.inv_d[0], B.inv_d[1], B.inv_d[2]};
  const int bnt0 = B.nt[0], bnt1 = B.nt[1];
  const long long c0 = (long long)blockIdx.x * kHbChunk;
  unsigned err = 0;
  unsigned long long valid = 0;
  float mmax = 0.f;
  for (int r = 0; r < kHbPer; ++r) {
    const int li = r * kHbThreads + tid;
    const long long i = c0 + li;
    if (i >= A.n) break;
    const float x = A.x[i], y = A.y[i], z = A.z[i], m = A.m[i];
    const int4 e = make_int4(__float_as_int(x), __float_as_int(y), __float_as_int(z), __float_as_int(m));
    parts[li] = e;
    if (!(isfinite(x) && isfinite(y) && isfinite(z) && isfinite(m))) { err |= kErrNonFinite; continue; }
    if (m < 0.f) { err |= kErrNegMass; continue; }
    if (!(tsc_base_in_range(x, P.xl, bid[0], bn[0]) && tsc_base_in_range(y, P.xl, bid[1], bn[1]) &&
          tsc_base_in_range(z, P.xl, bid[2], bn[2]))) {
      err |= kErrWrap;
      continue;
    }
    ++valid;
    mmax = fmaxf(mmax, m);
    const int ix = tsc_base_index(x, P.xl, bd[0], bid[0], bn[0]);
    const int jy = tsc_base_index(y, P.xl, bd[1], bid[1], bn[1]);
    const int kz = tsc_base_index(z, P.xl, bd[2], bid[2], bn[2]);
    int tx[3], ty[3], tz[3];
    const int nx_ = base_axis_tiles(ix, bn[0], P.tshift[0], tx);
    const int ny_ = base_axis_tiles(jy, bn[1], P.tshift[1], ty);
    const int nz_ = base_axis_tiles(kz, bn[2], P.tshift[2], tz);
    for (int c = 0; c < nz_; ++c)
      for (int b = 0; b < ny_; ++b)
        for (int a = 0; a < nx_; ++a) {
          const int key = 2 * ((tz[c] * bnt1 + ty[b]) * bnt0 + tx[a]);
          const int h = slot_insert<kHbCap>(keys, key);
          if (h >= 0) {
            atomicAdd(&vals[h], 1u);
            list[atomicAdd(&s_n, 1)] = ((unsigned)h << 9) | (unsigned)li;   // at most 8 per particle
          } else {
            emit_direct(cursor, start, entries, C, key, e);
          }
        }
  }
  __syncthreads();
  // one reservation per tile of the workgroup
  for (int k = tid; k < kHbCap; k += kHbThreads) {
    const int key = keys[k];
    if (key >= 0) {
      const unsigned long long b0 = atomicAdd(&cursor[key], (unsigned long long)vals[k]);
      base[k] = b0;
      if (b0 + vals[k] > start[key + 1]) set_overflow(C);
    }
    vals[k] = 0u;
  }
  __syncthreads();
  const int ne = s_n;
  for (int j = tid; j < ne; j += kHbThreads) {
    const unsigned v = list[j];
    const int h = (int)(v >> 9);
    const int key = keys[h];
    const unsigned long long pos = base[h] + atomicAdd(&vals[h], 1u);
    if (pos < start[key + 1]) entries[pos] = parts[v & 511u];
  }
  // per-workgroup partials (members of the base grid, largest mass)
  valid = wave_sum(valid);
  mmax = wave_max(mmax);
  err = wave_or(err);
  if ((tid & 63) == 0) {
    s_valid[tid >> 6] = valid;
    s_mmax[tid >> 6] = mmax;
    if (err) atomicOr(&C->err, err);
  }
  __syncthreads();
  if (tid == 0) {
    unsigned long long p = 0;
    float mm = 0.f;
    for (int w = 0; w < kHbThreads / 64; ++w) { p += s_valid[w]; mm = fmaxf(mm, s_mmax[w]); }
    BT.pairs[blockIdx.x] = p;
    BT.mmax[blockIdx.x] = __float_as_uint(mm);
  }
}

// The emission pass of one level: items [item_lo, item_hi) are the work items
// (chunks of one tile's entry range) of the level's tiles.
__global__ void __launch_bounds__(kEmThreads) emit_tsc(Plan P, const Item* items, long long item_lo,
                                                       long long item_hi, unsigned long long* cursor,
                                                       const unsigned long long* start, int4* entries,
                                                       unsigned* item_pairs, Counters* C) {
  __shared__ int keys[kEmCap];
  __shared__ unsigned vals[kEmCap];
  __shared__ unsigned long long base[kEmCap];
  __shared__ unsigned list[kEmList];
  __shared__ int4 stage[kEmRound];
  __shared__ BoxRec recs[kEmTargets];
  __shared__ int s_n;
  __shared__ unsigned s_red[kEmThreads / 64];
  const int tid = threadIdx.x;
  for (long long ii = item_lo + blockIdx.x; ii < item_hi; ii += gridDim.x) {
    const Item it = items[ii];
    const int t = it.tile;
    const unsigned q0 = P.tgt_off[t], q1 = P.tgt_off[t + 1];
    if (q0 == q1) {
      if (tid == 0) item_pairs[ii] = 0u;
      continue;
    }
    const int g = P.tile_grid[t];
    const GridDesc& G = P.grids[g];
    const int gn[3] = {G.n[0], G.n[1], G.n[2]};
    const float gorg[3] = {G.org[0], G.org[1], G.org[2]};
    const float gd[3] = {G.d[0], G.d[1], G.d[2]};
    const float gid[3] = {G.inv_d[0], G.inv_d[1], G.inv_d[2]};
    const int local = t - G.tile0;
    const int tc[3] = {local % G.nt[0], (local / G.nt[0]) % G.nt[1], local / (G.nt[0] * G.nt[1])};
    const long long e0 = it.begin;
    const long long e1 = min(it.end, (long long)cursor[2 * t]);
    for (int k = tid; k < kEmCap; k += kEmThreads) { keys[k] = -1; vals[k] = 0u; }
    unsigned pairs = 0;
    for (long long r0 = e0; r0 < e1; r0 += kEmRound) {
      const int nr = (int)min((long long)kEmRound, e1 - r0);
      if (tid == 0) s_n = 0;
      for (int k = tid; k < nr; k += kEmThreads) stage[k] = entries[r0 + k];
      __syncthreads();
      for (unsigned q = q0; q < q1; q += kEmTargets) {
        const int nq = (int)min((unsigned)kEmTargets, q1 - q);
        for (int k = tid; k < nq * 8; k += kEmThreads)
          reinterpret_cast<int4*>(recs)[k] = reinterpret_cast<const int4*>(P.grec + P.tgt[q + (k >> 3)])[k & 7];
        __syncthreads();
        for (int k = tid; k < nr; k += kEmThreads) {
          const int4 e = stage[k];
          const float p[3] = {__int_as_float(e.x), __int_as_float(e.y), __int_as_float(e.z)};
          // home: the clamped cell of this grid lies in this tile
          bool home = true;
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            int ci;
            if (g == 0) ci = tsc_base_index(p[a], P.xl, gd[a], gid[a], gn[a]);
            else ci = clampi((int)index_quotient(p[a] - gorg[a], gd[a], gid[a]), 1, gn[a]);
            home = home && ((ci - 1) >> P.tshift[a]) == tc[a];
          }
          if (!home) continue;
          for (int j = 0; j < nq; ++j) {
            const BoxRec& R = recs[j];
            if (!in_box(R, p[0], p[1], p[2])) continue;
            ++pairs;
            int lo[3], hi[3];
            bool empty = false;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              const int ci = (int)index_quotient(p[a] - R.org[a], R.d[a], R.inv_d[a]);
              const int c0 = max(ci - 1, 1), c1 = min(ci + 1, R.n[a]);
              // a member whose quotient rounds onto the far face of the box
              // (ci = N + 2) has no stencil cell in the patch
              empty = empty || c1 < c0;
              lo[a] = (c0 - 1) >> P.tshift[a];
              hi[a] = (c1 - 1) >> P.tshift[a];
            }
            if (empty) continue;
            for (int cz = lo[2]; cz <= hi[2]; ++cz)
              for (int cy = lo[1]; cy <= hi[1]; ++cy)
                for (int cx = lo[0]; cx <= hi[0]; ++cx) {
                  const int key = 2 * (R.tile0 + (cz * R.nt1 + cy) * R.nt0 + cx);
                  const int h = slot_insert<kEmCap>(keys, key);
                  int pos = -1;
                  if (h >= 0) {
                    pos = atomicAdd(&s_n, 1);
                    if (pos < kEmList) {
                      atomicAdd(&vals[h], 1u);
                      list[pos] = ((unsigned)h << 10) | (unsigned)k;   // h < 1024, k < 1024
                    }
                  }
                  if (h < 0 || pos >= kEmList) emit_direct(cursor, start, entries, C, key, e);
                }
          }
        }
        __syncthreads();
      }
      // one reservation per tile key of the round
      for (int k = tid; k < kEmCap; k += kEmThreads) {
        const int key = keys[k];
        const unsigned v = vals[k];
        if (key >= 0 && v) {
          const unsigned long long b0 = atomicAdd(&cursor[key], (unsigned long long)v);
          base[k] = b0;
          if (b0 + v > start[key + 1]) set_overflow(C);
        }
        vals[k] = 0u;
      }
      __syncthreads();
      const int ne = min(s_n, kEmList);
      for (int j = tid; j < ne; j += kEmThreads) {
        const unsigned v = list[j];
        const int h = (int)(v >> 10);
        const int key = keys[h];
        const unsigned long long pos = base[h] + atomicAdd(&vals[h], 1u);
        if (pos < start[key + 1]) entries[pos] = stage[v & 1023u];
      }
      __syncthreads();
      for (int k = tid; k < kEmCap; k += kEmThreads) vals[k] = 0u;
      __syncthreads();
    }
    // members found by this item
    unsigned s = wave_sum(pairs);
    if ((tid & 63) == 0) s_red[tid >> 6] = s;
    __syncthreads();
    if (tid == 0) {
      unsigned tot = 0;
      for (int w = 0; w < kEmThreads / 64; ++w) tot += s_red[w];
      item_pairs[ii] = tot;
    }
    __syncthreads();
  }
}

// statistics of the call: members = base members + the members every emission
// item found; entries = the filled part of every tile's range
__global__ void __launch_bounds__(1024) hier_stats(const unsigned* item_pairs, long long n_items,
                                                  const unsigned long long* start,
                                                  const unsigned long long* cursor, long long nkeys, Counters* C) {
  __shared__ unsigned long long sp[16], se[16];
  unsigned long long p = 0, e = 0;
  for (long long i = threadIdx.x; i < n_items; i += blockDim.x) p += item_pairs[i];
  for (long long k = threadIdx.x; k < nkeys; k += blockDim.x) e += cursor[k] - start[k];
  p = wave_sum(p);
  e = wave_sum(e);
  if ((threadIdx.x & 63) == 0) { sp[threadIdx.x >> 6] = p; se[threadIdx.x >> 6] = e; }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tp = 0, te = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) { tp += sp[w]; te += se[w]; }
    C->n_pairs += tp;
    C->n_entries = te;
  }
}

long long hier_bin_blocks(long long n) { return (n + kHbChunk - 1) / kHbChunk; }

hipError_t launch_cursor_reset(const unsigned long long* start, unsigned long long* cursor, long long nkeys,
                               int cus, hipStream_t s) {
  if (nkeys <= 0) return hipSuccess;
  const long long want = (nkeys + 255) / 256;
  cursor_reset<<<(int)std::min<long long>(want, (long long)cus * 16), 256, 0, s>>>(start, cursor, nkeys);
  return hipGetLastError();
}

hipError_t launch_count_to_capacity(unsigned* cnt, long long nkeys, int cus, hipStream_t s) {
  if (nkeys <= 0) return hipSuccess;
  const long long want = (nkeys + 255) / 256;
  count_to_capacity<<<(int)std::min<long long>(want, (long long)cus * 16), 256, 0, s>>>(cnt, nkeys);
  return hipGetLastError();
}

hipError_t launch_bin_base_tsc(const Plan& P, const Parts& A, unsigned long long* cursor,
                               const unsigned long long* start, int4* entries, Counters* C,
                               const BlockTables& BT, hipStream_t s) {
  const long long nb = hier_bin_blocks(A.n);
  if (nb == 0) return hipSuccess;
  bin_base_tsc<<<(unsigned)nb, kHbThreads, 0, s>>>(P, A, cursor, start, entries, C, BT);
  return hipGetLastError();
}

hipError_t launch_emit_tsc(const Plan& P, const Item* items, long long item_lo, long long item_hi,
                           unsigned long long* cursor, const unsigned long long* start, int4* entries,
                           unsigned* item_pairs, Counters* C, int cus, hipStream_t s) {
  if (item_hi <= item_lo) return hipSuccess;
  const int grid = (int)std::min<long long>(item_hi - item_lo, (long long)cus * 8);
  emit_tsc<<<grid, kEmThreads, 0, s>>>(P, items, item_lo, item_hi, cursor, start, entries, item_pairs, C);
  return hipGetLastError();
}

hipError_t launch_hier_stats(const unsigned* item_pairs, long long n_items, const unsigned long long* start,
                             const unsigned long long* cursor, long long nkeys, Counters* C, hipStream_t s) {
  hier_stats<<<1, 1024, 0, s>>>(item_pairs, n_items, start, cursor, nkeys, C);
  return hipGetLastError();
}

}  // namespace asohf
