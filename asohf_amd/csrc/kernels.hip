// kernels.hip -- gfx950 kernels of the density build and their launchers.
//
// Batched pipeline (one call):
//   count_kernel    particles -> (grid, tile) membership counts     [reads x,y,z,m]
//   scan_*          exclusive scan of the counts -> entry offsets
//   scatter_kernel  particles -> 16-byte entries grouped by tile    [reads x,y,z,m]
//   plan_items      tiles -> work items (chunks of <= chunk entries)
//   deposit_kernel  ONE persistent launch over the item queue: LDS tile
//                   accumulation, normalised coalesced stores, FP64-exact
//                   tile mass partials
//   grid_sums       fixed-order per-grid mass sums
// Plain form: one launch per grid with global FP32 atomics (plain_*).
#include <hip/hip_runtime.h>

#include "dev.h"
#include "launch.h"

namespace asohf {

// LDS copy of the kernel table rows 0..kNLevels: row irk holds 2^(irk+1) + 1
// values and starts after the rows below it
__host__ __device__ constexpr int ktab_off(int irk) { return (2 << irk) - 2 + irk; }
constexpr int kKtabLds = ktab_off(kNLevels + 1);

// ---------------------------------------------------------------------------
// fills (used instead of hipMemsetAsync so every step can live in a graph)
// ---------------------------------------------------------------------------
__global__ void fill_u32_kernel(unsigned* p, unsigned v, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---------------------------------------------------------------------------
// count / scatter
// ---------------------------------------------------------------------------
// Block-local aggregation of (tile) keys.  A workgroup takes a contiguous
// run of particles (consecutive particles of one halo share tiles) and counts
// its emissions per key in an LDS open-addressing table; only one global
// atomic per distinct key leaves the workgroup.  A key whose probe path is
// full is never inserted later (the table only fills), so every emission of
// that key takes the direct global path consistently in all phases.
constexpr int kBinThreads = 256;
constexpr int kBinPer = 2;                       // particles per thread per block
constexpr int kBinChunk = kBinThreads * kBinPer;
constexpr int kBinCap = 1024;                    // LDS table slots (power of two)
constexpr int kBinProbe = 32;

// tile keys are < 2^24 (2 * tiles); 24-bit multiplies are full rate
__device__ __forceinline__ unsigned key_hash(unsigned k) {
  k ^= k >> 13;
  k = __umul24(k, 0x9E3779u) ^ (k >> 11);
  k ^= k >> 7;
  return k & (kBinCap - 1);
}
// slot of key k, inserting it when absent; -1 when its probe path is full.
// A plain read first: most emissions find their key already present, and a
// read costs less than a compare-and-swap.
__device__ __forceinline__ int lds_insert(int* keys, int k) {
  unsigned h = key_hash((unsigned)k);
#pragma unroll 1
  for (int p = 0; p < kBinProbe; ++p) {
    const int cur = keys[h];
    if (cur == k) return (int)h;
    if (cur == -1) {
      const int old = atomicCAS(&keys[h], -1, k);
      if (old == -1 || old == k) return (int)h;
    }
    h = (h + 1) & (kBinCap - 1);
  }
  return -1;
}
__device__ __forceinline__ int lds_find(const int* keys, int k) {
  unsigned h = key_hash((unsigned)k);
  for (int p = 0; p < kBinProbe; ++p) {
    const int v = keys[h];
    if (v == k) return (int)h;
    if (v == -1) return -1;
    h = (h + 1) & (kBinCap - 1);
  }
  return -1;
}

// one particle of lookahead: the next particle's loads are in flight while
// the current one is enumerated
struct Prefetch {
  float x, y, z, m;
  __device__ Prefetch(const Parts& A, long long i) : x(0.f), y(0.f), z(0.f), m(0.f) {
    if (i < A.n) { x = A.x[i]; y = A.y[i]; z = A.z[i]; m = A.m[i]; }
  }
  __device__ __forceinline__ void next(const Parts& A, long long inext, bool more, float& cx, float& cy,
                                       float& cz, float& cm) {
    cx = x; cy = y; cz = z; cm = m;
    if (more && inext < A.n) { x = A.x[inext]; y = A.y[inext]; z = A.z[inext]; m = A.m[inext]; }
  }
};

// emission lists: (key << 9) | particle index within the block
constexpr int kEmitCap = 4096;
static_assert(kBinChunk <= 512, "the emission list keeps 9 bits of particle index");

template <int MODE>
__global__ void __launch_bounds__(kBinThreads) count_kernel(Plan P, Parts A, unsigned* cnt, Counters* C,
                                                            BlockTables BT) {
  __shared__ int keys[kBinCap];
  __shared__ unsigned vals[kBinCap];
  __shared__ unsigned elist_s[kEmitCap];
  __shared__ int s_ne;
  for (int k = threadIdx.x; k < kBinCap; k += kBinThreads) { keys[k] = -1; vals[k] = 0u; }
  if (threadIdx.x == 0) s_ne = 0;
  __syncthreads();
  const bool list = BT.use_elist;
  unsigned err = 0;
  unsigned long long pairs = 0;
  float mmax = 0.f;
  const long long c0 = (long long)blockIdx.x * kBinChunk;
  Prefetch pf(A, c0 + threadIdx.x);
  for (int r = 0; r < kBinPer; ++r) {
    const long long i = c0 + (long long)r * kBinThreads + threadIdx.x;
    if (i >= A.n) break;
    float x, y, z, m;
    pf.next(A, i + kBinThreads, r + 1 < kBinPer, x, y, z, m);
    if (!(isfinite(x) && isfinite(y) && isfinite(z) && isfinite(m))) { err |= kErrNonFinite; continue; }
    if (m < 0.f) { err |= kErrNegMass; continue; }
    mmax = fmaxf(mmax, m);
    const int s = MODE == 1 ? species_of(A, i) : 0;
    enumerate_particle<MODE>(
        P, i, x, y, z, m, s,
        [&](int, int t, int wide, int4) {
          const int key = 2 * t + wide;
          const int h = lds_insert(keys, key);
          if (h >= 0) atomicAdd(&vals[h], 1u);
          else atomicAdd(&cnt[key], 1u);
          if (list) {
            const int pos = atomicAdd(&s_ne, 1);
            if (pos < kEmitCap) elist_s[pos] = ((unsigned)key << 9) | (unsigned)(i - c0);
          }
        },
        [&](int, int, int, int) { ++pairs; }, err);
  }
  __syncthreads();
  // the emission list, when it fits, for the scatter pass
  if (list) {
    __shared__ long long s_eoff;
    const int ne = s_ne;
    if (threadIdx.x == 0) {
      long long eo = -1;
      if (ne == 0) {
        eo = 0;
      } else if (ne <= kEmitCap) {
        const int sub = blockIdx.x & (kPoolSubs - 1);
        const long long subcap = BT.ecap / kPoolSubs;
        const long long off = (long long)atomicAdd(&C->epool_fill[sub * (kLine / 2)], (unsigned long long)ne);
        eo = off + ne <= subcap ? sub * subcap + off : -1;
      }
      s_eoff = eo;
      BT.eoff[blockIdx.x] = eo;
      BT.ecount[blockIdx.x] = ne;
    }
    __syncthreads();
    const long long eo = s_eoff;
    if (eo >= 0)
      for (int j = threadIdx.x; j < ne; j += kBinThreads) BT.elist[eo + j] = elist_s[j];
  }
  // flush the counts, and save the table (compacted) for the scatter pass
  __shared__ int s_n;
  __shared__ long long s_off;
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  for (int k = threadIdx.x; k < kBinCap; k += kBinThreads)
    if (keys[k] >= 0) {
      atomicAdd(&cnt[keys[k]], vals[k]);
      atomicAdd(&s_n, 1);
    }
  __syncthreads();
  if (threadIdx.x == 0) {
    // sub-pool of this block: 1/kPoolSubs of the pool, its own counter line
    const int sub = blockIdx.x & (kPoolSubs - 1);
    const long long subcap = BT.cap / kPoolSubs;
    const long long off = (long long)atomicAdd(&C->pool_fill[sub * (kLine / 2)], (unsigned long long)s_n);
    s_off = off + s_n <= subcap ? sub * subcap + off : -1;
    BT.off[blockIdx.x] = s_off;
    BT.nkeys[blockIdx.x] = s_n;
    s_n = 0;
  }
  __syncthreads();
  if (s_off >= 0)
    for (int k = threadIdx.x; k < kBinCap; k += kBinThreads)
      if (keys[k] >= 0) BT.pool[s_off + atomicAdd(&s_n, 1)] = make_int2(keys[k], (int)vals[k]);
  // per-block partials (no contended global atomics); reduce_blocks folds them
  __shared__ unsigned long long s_pairs[kBinThreads / 64];
  __shared__ float s_mmax[kBinThreads / 64];
  pairs = wave_sum(pairs);
  mmax = wave_max(mmax);
  err = wave_or(err);
  if ((threadIdx.x & 63) == 0) {
    s_pairs[threadIdx.x >> 6] = pairs;
    s_mmax[threadIdx.x >> 6] = mmax;
    if (err) atomicOr(&C->err, err);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long p = 0;
    float mm = 0.f;
    for (int w = 0; w < kBinThreads / 64; ++w) { p += s_pairs[w]; mm = fmaxf(mm, s_mmax[w]); }
    BT.pairs[blockIdx.x] = p;
    BT.mmax[blockIdx.x] = __float_as_uint(mm);
  }
}

// The call's base members and largest mass from the per-block partials:
// every workgroup folds a grid-stride slice and adds its part to the
// counters (cleared at the start of the call); integer sums and the maximum
// do not depend on the order.  The mass maximum compares FP32 bit patterns,
// which order like the values for masses >= 0; the sign bit is cleared so a
// -0 mass counts as 0.
__global__ void __launch_bounds__(256) reduce_blocks(BlockTables BT, long long nblk, Counters* C) {
  __shared__ unsigned long long sp[4];
  __shared__ float sm[4];
  unsigned long long p = 0;
  float mm = 0.f;
  for (long long b = blockIdx.x * (long long)blockDim.x + threadIdx.x; b < nblk;
       b += (long long)gridDim.x * blockDim.x) {
    p += BT.pairs[b];
    mm = fmaxf(mm, __uint_as_float(BT.mmax[b]));
  }
  p = wave_sum(p);
  mm = wave_max(mm);
  if ((threadIdx.x & 63) == 0) { sp[threadIdx.x >> 6] = p; sm[threadIdx.x >> 6] = mm; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) { p = sp[0] + sp[w]; sp[0] = p; sm[0] = fmaxf(sm[0], sm[w]); }
    if (sp[0]) atomicAdd(&C->n_pairs, sp[0]);
    atomicMax(&C->mmax_bits, __float_as_uint(sm[0]) & 0x7fffffffu);
  }
}

template <int MODE>
__global__ void __launch_bounds__(kBinThreads) scatter_kernel(Plan P, Parts A, unsigned long long* cursor,
                                                              int4* entries, unsigned long long cap,
                                                              BlockTables BT, const Counters* C) {
  __shared__ int keys[kBinCap];
  __shared__ unsigned vals[kBinCap];
  __shared__ unsigned long long base[kBinCap];
  __shared__ int4 parts_s[kBinChunk];
  if (C->overflow) return;   // the entry buffer is too small: the host rebuilds exactly
  for (int k = threadIdx.x; k < kBinCap; k += kBinThreads) { keys[k] = -1; vals[k] = 0u; }
  __syncthreads();
  unsigned err = 0;
  const long long c0 = (long long)blockIdx.x * kBinChunk;
  const long long saved = BT.off[blockIdx.x];
  if (saved >= 0) {
    // the count pass's table of this block: reinsert its keys
    const int nk = BT.nkeys[blockIdx.x];
    for (int j = threadIdx.x; j < nk; j += kBinThreads) {
      const int2 kv = BT.pool[saved + j];
      const int h = lds_insert(keys, kv.x);
      if (h >= 0) vals[h] = (unsigned)kv.y;
    }
  } else {
    // phase A: count this block's emissions per key
    Prefetch pf(A, c0 + threadIdx.x);
    for (int r = 0; r < kBinPer; ++r) {
      const long long i = c0 + (long long)r * kBinThreads + threadIdx.x;
      if (i >= A.n) break;
      float x, y, z, m;
      pf.next(A, i + kBinThreads, r + 1 < kBinPer, x, y, z, m);
      if (!(isfinite(x) && isfinite(y) && isfinite(z) && isfinite(m)) || m < 0.f) continue;
      const int s = MODE == 1 ? species_of(A, i) : 0;
      enumerate_particle<MODE>(
          P, i, x, y, z, m, s,
          [&](int, int t, int wide, int4) {
            const int h = lds_insert(keys, 2 * t + wide);
            if (h >= 0) atomicAdd(&vals[h], 1u);
          },
          [&](int, int, int, int) {}, err);
    }
  }
  __syncthreads();
  // phase B: one global reservation per key
  for (int k = threadIdx.x; k < kBinCap; k += kBinThreads) {
    if (keys[k] >= 0) base[k] = atomicAdd(&cursor[keys[k]], (unsigned long long)vals[k]);
    vals[k] = 0u;
  }
  __syncthreads();
  // phase C, from the count pass's emission list: the block's particles staged
  // in LDS, each listed emission written to its reserved slot
  const long long eo = BT.use_elist ? BT.eoff[blockIdx.x] : -1;
  if (eo >= 0) {
    for (int j = threadIdx.x; j < kBinChunk; j += kBinThreads) {
      const long long i = c0 + j;
      if (i < A.n)
        parts_s[j] = make_int4(__float_as_int(A.x[i]), __float_as_int(A.y[i]), __float_as_int(A.z[i]),
                               __float_as_int(A.m[i]));
    }
    __syncthreads();
    const int ne = BT.ecount[blockIdx.x];
    for (int j = threadIdx.x; j < ne; j += kBinThreads) {
      const unsigned v = BT.elist[eo + j];
      const int key = (int)(v >> 9);
      const int h = lds_find(keys, key);
      const unsigned long long slot =
          h >= 0 ? base[h] + atomicAdd(&vals[h], 1u) : atomicAdd(&cursor[key], 1ull);
      const int li = (int)(v & 511u);
      int4 e = parts_s[li];
      if (MODE == 1) {
        // the table-mode entry: cell indices in the tile's grid
        const float x = __int_as_float(e.x), y = __int_as_float(e.y), z = __int_as_float(e.z),
                    m = __int_as_float(e.w);
        const int g = P.tile_grid[key >> 1];
        e = g == 0 ? table_base_entry(P, x, y, z, m)
                   : table_patch_entry(P.grids[g], x, y, z, m, species_of(A, c0 + li));
      }
      if (slot < cap) entries[slot] = e;
    }
    return;
  }
  // phase C: the same emissions again, written to their reserved slots
  Prefetch pc(A, c0 + threadIdx.x);
  for (int r = 0; r < kBinPer; ++r) {
    const long long i = c0 + (long long)r * kBinThreads + threadIdx.x;
    if (i >= A.n) break;
    float x, y, z, m;
    pc.next(A, i + kBinThreads, r + 1 < kBinPer, x, y, z, m);
    if (!(isfinite(x) && isfinite(y) && isfinite(z) && isfinite(m)) || m < 0.f) continue;
    const int s = MODE == 1 ? species_of(A, i) : 0;
    enumerate_particle<MODE>(
        P, i, x, y, z, m, s,
        [&](int, int t, int wide, int4 e) {
          const int key = 2 * t + wide;
          const int h = lds_find(keys, key);
          const unsigned long long slot =
              h >= 0 ? base[h] + atomicAdd(&vals[h], 1u) : atomicAdd(&cursor[key], 1ull);
          if (slot < cap) entries[slot] = e;   // the gate checked the buffer against the count pass
        },
        [&](int, int, int, int) {}, err);
  }
}

// membership records {grid, particle, ix, jy, kz}
template <int MODE>
__global__ void __launch_bounds__(256) membership_kernel(Plan P, Parts A, int* out, long long cap,
                                                         unsigned long long* count) {
  unsigned err = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < A.n;
       i += (long long)gridDim.x * blockDim.x) {
    const float x = A.x[i], y = A.y[i], z = A.z[i], m = A.m[i];
    if (!(isfinite(x) && isfinite(y) && isfinite(z))) continue;
    const int s = MODE == 1 ? species_of(A, i) : 0;
    enumerate_particle<MODE>(
        P, i, x, y, z, m, s, [&](int, int, int, int4) {},
        [&](int g, int ix, int jy, int kz) {
          const unsigned long long slot = atomicAdd(count, 1ull);
          if ((long long)slot < cap) {
            int* r = out + 5 * slot;
            r[0] = g; r[1] = (int)i; r[2] = ix; r[3] = jy; r[4] = kz;
          }
        },
        err);
  }
  // validation flags of the enumeration in count[1]
  err = wave_or(err);
  if ((threadIdx.x & 63) == 0 && err) atomicOr(count + 1, (unsigned long long)err);
}

// ---------------------------------------------------------------------------
// exclusive scan of u32 counts into u64 offsets (3 kernels)
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 1024;
constexpr int kScanPer = 4;
constexpr int kScanTile = kScanThreads * kScanPer;

__device__ unsigned long long block_exclusive_scan(unsigned long long v, unsigned long long* sh,
                                                   unsigned long long* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  unsigned long long inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    unsigned long long t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  if (lane == 63) sh[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long run = 0;
    for (int k = 0; k < nw; ++k) { unsigned long long t = sh[k]; sh[k] = run; run += t; }
    sh[nw] = run;
  }
  __syncthreads();
  const unsigned long long r = sh[w] + inc - v;
  *total = sh[nw];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(kScanThreads) scan_reduce(const unsigned* in, long long n,
                                                           unsigned long long* bsum) {
  __shared__ unsigned long long sh[kScanThreads / 64 + 1];
  const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanPer;
  unsigned long long s = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k)
    if (base + k < n) s += in[base + k];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) bsum[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kScanThreads) scan_bsums(unsigned long long* bsum, int nb,
                                                          unsigned long long* total_out) {
  __shared__ unsigned long long sh[kScanThreads / 64 + 1];
  unsigned long long carry = 0;
  for (int b0 = 0; b0 < nb; b0 += kScanThreads) {
    const int b = b0 + threadIdx.x;
    const unsigned long long v = b < nb ? bsum[b] : 0ull;
    unsigned long long tot;
    const unsigned long long ex = block_exclusive_scan(v, sh, &tot);
    if (b < nb) bsum[b] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) *total_out = carry;
}

__global__ void __launch_bounds__(kScanThreads) scan_apply(const unsigned* in, long long n,
                                                          const unsigned long long* bsum,
                                                          unsigned long long* start,
                                                          unsigned long long* cursor) {
  __shared__ unsigned long long sh[kScanThreads / 64 + 1];
  const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanPer;
  unsigned v[kScanPer];
  unsigned long long s = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    v[k] = base + k < n ? in[base + k] : 0u;
    s += v[k];
  }
  unsigned long long tot;
  unsigned long long ex = block_exclusive_scan(s, sh, &tot) + bsum[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    if (base + k < n) {
      start[base + k] = ex;
      if (cursor) cursor[base + k] = ex;
    }
    ex += v[k];
  }
}

// ---------------------------------------------------------------------------
// work items
// ---------------------------------------------------------------------------
// Work items.  plan_count: items per tile and the partial-tile slots a split
// tile needs; after exclusive scans of both, plan_write places every tile's
// items at its offset (no atomics: the item order is the tile order).
// With chunk_wide = 0 the tile's whole range [start[2t], start[2t+2]) is
// chunked as narrow entries (the TSC hierarchy binning, whose keys 2t and
// 2t + 1 are a tile's home and other entries).
__device__ __forceinline__ void tile_chunks(const unsigned long long* start, int t, int chunk, int chunk_wide,
                                            long long& s0, long long& s1, long long& s2, int& cn, int& cw) {
  s0 = start[2 * t]; s1 = start[2 * t + 1]; s2 = start[2 * t + 2];
  if (chunk_wide == 0) {
    s1 = s2;
    cn = (int)((s2 - s0 + chunk - 1) / chunk);
    cw = 0;
    return;
  }
  cn = (int)((s1 - s0 + chunk - 1) / chunk);
  cw = (int)((s2 - s1 + chunk_wide - 1) / chunk_wide);
}

__global__ void plan_count(const unsigned long long* start, int ntiles, int chunk, int chunk_wide,
                           unsigned* nitems_tile, unsigned* nsplit_tile, unsigned* is_split) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  long long s0, s1, s2;
  int cn, cw;
  tile_chunks(start, t, chunk, chunk_wide, s0, s1, s2, cn, cw);
  const int tot = max(cn + cw, 1);   // an empty tile still writes its zeros
  nitems_tile[t] = (unsigned)tot;
  nsplit_tile[t] = tot > 1 ? 1u : 0u;   // one accumulator tile per split tile
  is_split[t] = tot > 1 ? 1u : 0u;
}

__global__ void plan_write(const unsigned long long* start, int ntiles, int chunk, int chunk_wide,
                           const unsigned long long* item_off, const unsigned long long* split_off,
                           const unsigned long long* split_rank, Item* items, int* split_slot,
                           int* split_tiles, long long* tile_mass, const Counters* C) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles || C->overflow) return;
  long long s0, s1, s2;
  int cn, cw;
  tile_chunks(start, t, chunk, chunk_wide, s0, s1, s2, cn, cw);
  const long long base = (long long)item_off[t];
  const int tot = max(cn + cw, 1);
  split_slot[t] = tot > 1 ? (int)split_off[t] : -1;
  if (tot > 1) {
    split_tiles[split_rank[t]] = t;
    tile_mass[t] = 0;   // summed by split_reduce (zero bits: 0 as int64 and as FP64)
  }
  if (cn + cw == 0) { items[base] = Item{t, 0, s0, s0}; return; }
  int k = 0;
  for (int c = 0; c < cn; ++c, ++k) {
    const long long b = s0 + (long long)c * chunk;
    items[base + k] = Item{t, k << 8, b, min(b + (long long)chunk, s1)};
  }
  // bit 1: the item holds every wide entry of its tile (the histogram form
  // is used only then, so its one rounding per cell does not depend on how
  // the entries fall into items)
  for (int c = 0; c < cw; ++c, ++k) {
    const long long b = s1 + (long long)c * chunk_wide;
    items[base + k] = Item{t, 1 | (cw == 1 ? 2 : 0) | (k << 8), b, min(b + (long long)chunk_wide, s2)};
  }
}

// The totals of the planning scans against the workspaces of this call.  A
// call whose buffers were sized by an earlier call finds out here, on the
// device, whether they still suffice; if not, every later kernel of the call
// does nothing and the host rebuilds with exact sizes.
__global__ void gate_kernel(const unsigned long long* tot_entries, const unsigned long long* tot_items,
                            const unsigned long long* tot_split, unsigned long long cap_entries,
                            unsigned long long cap_items, unsigned long long cap_split, Counters* C) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const unsigned long long ne = *tot_entries, ni = *tot_items, ns = *tot_split;
  C->n_entries = ne;
  C->n_items = ni;
  C->n_split = ns;
  const bool over = ne > cap_entries || ni > cap_items || ns > cap_split || ni >= (1ull << 31);
  C->overflow = over ? 1u : 0u;
  C->n_items_eff = over ? 0u : (unsigned)ni;
}

// Fixed-point scale of the deterministic reduction: the power of two with
// max(m) * scale in [2^29, 2^30), so every quantised term is below 2^30,
// eight non-centre terms (each below half of it) sum below 2^32, and a sum
// over <= 2^32 particles fits an int64.  The unit is at most max(m) * 2^-29.
__device__ __forceinline__ double fixed_scale(unsigned mmax_bits) {
  const float mm = __uint_as_float(mmax_bits);
  int e = 0;
  if (mm > 0.f) (void)frexpf(mm, &e);
  return ldexp(1.0, 30 - e);
}

// ---------------------------------------------------------------------------
// the persistent deposit
// ---------------------------------------------------------------------------
template <int RED> struct AccT;
template <> struct AccT<0> { using T = long long; };
template <> struct AccT<1> { using T = float; };

__device__ __forceinline__ void lds_add(long long* a, long long v) {
  atomicAdd(reinterpret_cast<unsigned long long*>(a), (unsigned long long)v);
}
__device__ __forceinline__ void lds_add(float* a, float v) { atomicAdd(a, v); }
__device__ __forceinline__ void glb_add(long long* a, long long v) {
  atomicAdd(reinterpret_cast<unsigned long long*>(a), (unsigned long long)v);
}
__device__ __forceinline__ void glb_add(float* a, float v) { atomicAdd(a, v); }
__device__ __forceinline__ long long glb_load(long long* a) {
  return __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float glb_load(float* a) {
  return __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A deterministic term: ((m*scale*wx)*wy)*wz rounded to the nearest integer
// (halves up) as one fused multiply-add of 1/2 and a truncating conversion.
// The terms are >= 0 and below 2^30; a term of 2^23 or more comes out within
// one unit of the exact product's rounding.  Every term is a function of its
// particle alone, so the sums stay independent of the order of the adds.
__device__ __forceinline__ unsigned term_u32(float mxy, float wz) {
  return (unsigned)__fmaf_rn(mxy, wz, 0.5f);
}

// 27-cell deposit of one particle into the LDS tile, which carries a 2-cell
// halo on every side: the particle's centre cell lies within one cell of the
// tile (it was binned to the tile because one of its 27 cells is inside), so
// all 27 cells are in the array and the deposit needs no per-cell test, one
// base address and 27 constant offsets.  The terms are the reference's FP32
// products ((m*wx)*wy)*wz.  Deterministic mode scales the mass by the
// power-of-two fixed-point scale first (exact), rounds every term to an
// integer and gives the centre cell the remainder Q(m) - (sum of the other
// 26), so each particle deposits exactly Q(m) over its stencil.
template <int RED, int HX, int HY>
__device__ __forceinline__ void deposit27(typename AccT<RED>::T* base, float m, const float wx[3],
                                          const float wy[3], const float wz[3], float scale) {
  const float ms = RED == 0 ? m * scale : m;
  float mx[3], mxy[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a) mx[a] = ms * wx[a];
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int a = 0; a < 3; ++a) mxy[b][a] = mx[a] * wy[b];
  if (RED == 1) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int a = 0; a < 3; ++a)
          lds_add(reinterpret_cast<float*>(base) + (a + HX * (b + HY * c)), mxy[b][a] * wz[c]);
  } else {
    // every term is >= 0 (masses >= 0, TSC and kernel weights >= 0) and
    // below 2^30, so it converts to an unsigned word whose 64-bit form needs
    // no sign extension
    unsigned rest = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          if (a == 1 && b == 1 && c == 1) continue;
          const unsigned q = term_u32(mxy[b][a], wz[c]);
          rest += q;
          lds_add(reinterpret_cast<long long*>(base) + (a + HX * (b + HY * c)), (long long)(unsigned long long)q);
        }
    lds_add(reinterpret_cast<long long*>(base) + (1 + HX * (1 + HY * 1)),
            (long long)__float2uint_rn(ms) - (long long)rest);
  }
}

// a wave-uniform 64-bit value, moved to scalar registers
__device__ __forceinline__ long long uniform64(long long v) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)((unsigned long long)v >> 32));
  return (long long)(((unsigned long long)hi << 32) | lo);
}

// the value of lane l^1 (l^2) of the wave, through a quad permutation
__device__ __forceinline__ unsigned from_xor1(unsigned v) {
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
}
__device__ __forceinline__ unsigned from_xor2(unsigned v) {
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
}
// the value of lane l+4 (l-4) of the same row of 16 lanes; a lane whose
// source lies outside its row gets 0 (the DPP default)
__device__ __forceinline__ unsigned from_plus4(unsigned v) {
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x104, 0xF, 0xF, false);   // row_shl:4
}
__device__ __forceinline__ unsigned from_minus4(unsigned v) {
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);   // row_shr:4
}
__device__ __forceinline__ long long from_plus4(long long v) {
  const unsigned lo = from_plus4((unsigned)v), hi = from_plus4((unsigned)((unsigned long long)v >> 32));
  return (long long)(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ float from_xor1(float v) { return __uint_as_float(from_xor1(__float_as_uint(v))); }
__device__ __forceinline__ float from_xor2(float v) { return __uint_as_float(from_xor2(__float_as_uint(v))); }
__device__ __forceinline__ float from_plus4(float v) { return __uint_as_float(from_plus4(__float_as_uint(v))); }
__device__ __forceinline__ long long from_xor1(long long v) {
  const unsigned lo = from_xor1((unsigned)v), hi = from_xor1((unsigned)((unsigned long long)v >> 32));
  return (long long)(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ long long from_xor2(long long v) {
  const unsigned lo = from_xor2((unsigned)v), hi = from_xor2((unsigned)((unsigned long long)v >> 32));
  return (long long)(((unsigned long long)hi << 32) | lo);
}

// Lanes of a merge step needed before it pays (each merged pair saves one
// lane's 27 LDS adds and costs every lane two instructions per term).
constexpr int kMergeMin = 4;

// The 27-cell deposit (TSC, or a table-mode narrow entry) with lanes merged
// first.  Clustered particles put many lanes of a wave on the same stencil,
// and the LDS serialises adds to one address; the terms can be summed in
// registers instead.  Deterministic mode: the terms are exact integers, so
// the result is the same.  Atomic mode: an FP32 sum in another order, as any
// order of its LDS atomics is.
// Three butterfly steps: lane l takes lane l^1's terms when both stencils
// start at the same cell (l even), then lane l^2's (bit 1 of l clear), then
// lane l^4's (bit 2 clear).  Every non-centre TSC term is below half of
// Q(m) < 2^30 (one factor is an outer weight: at most 0.5 * 0.75 * 0.75 with
// exact centres, below 0.48 with the centre rounding the plan admits;
// table mode merges only when the plan found every non-centre product of its
// row-0 weights at most 0.45), so eight of them sum below 2^32 and all three
// steps add in 32 bits; the centre remainder, negative for a light enough
// particle, merges as int64.
// Every lane of the wave must be active (the steps read other lanes); ok =
// the lane holds an entry, corner = its stencil's first cell in the tile.
template <int RED, int HX, int HY>
__device__ __forceinline__ void deposit27_merged(typename AccT<RED>::T* acc, int corner, bool ok, float m,
                                                 const float wx[3], const float wy[3], const float wz[3],
                                                 float scale) {
  const unsigned lane = __lane_id();
  // lanes without an entry get keys that match no other lane
  const unsigned key0 = ok ? (unsigned)corner : (0x80000000u | lane);
  const bool m1 = from_xor1(key0) == key0;
  if (__popcll(__ballot(m1)) < kMergeMin) {
    if (ok) deposit27<RED, HX, HY>(acc + corner, m, wx, wy, wz, scale);
    return;
  }
  // a lane absorbed by a step sends its value and every lane adds what it
  // receives: keepers receive their partner's, everyone else receives 0
  const bool send1 = m1 && (lane & 1u);
  const bool alive1 = ok && !send1;
  const unsigned key1 = alive1 ? (unsigned)corner : (0x80000000u | lane);
  const bool m2 = from_xor2(key1) == key1;
  const bool step2 = __popcll(__ballot(m2)) >= kMergeMin;
  const bool send2 = step2 && m2 && (lane & 2u);
  const bool alive2 = alive1 && !send2;
  // lane l^4 stays in the row of 16 lanes, so one row shift reaches it.  Both
  // shifts run on every lane and the lane picks its value afterwards: under a
  // branch, half the lanes would be inactive and their values unreadable.
  const unsigned key2 = alive2 ? (unsigned)corner : (0x80000000u | lane);
  const unsigned k_minus = from_minus4(key2), k_plus = from_plus4(key2);
  const unsigned pk = (lane & 4u) ? k_minus : k_plus;
  const bool m3 = pk == key2;
  const bool step3 = __popcll(__ballot(m3)) >= kMergeMin;
  const bool send3 = step3 && m3 && (lane & 4u);
  const bool alive = alive2 && !send3;

  const float ms = RED == 0 ? m * scale : m;
  float mx[3], mxy[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a) mx[a] = ms * wx[a];
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int a = 0; a < 3; ++a) mxy[b][a] = mx[a] * wy[b];
  if (RED == 1) {
    float* fbase = reinterpret_cast<float*>(acc) + corner;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          float t = mxy[b][a] * wz[c];
          t += from_xor1(send1 ? t : 0.f);
          if (step2) t += from_xor2(send2 ? t : 0.f);
          if (step3) t += from_plus4(send3 ? t : 0.f);
          if (alive) lds_add(fbase + (a + HX * (b + HY * c)), t);
        }
    return;
  }
  long long* base = reinterpret_cast<long long*>(acc) + corner;
  unsigned rest = 0;
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        if (a == 1 && b == 1 && c == 1) continue;
        unsigned q = term_u32(mxy[b][a], wz[c]);
        rest += q;
        q += from_xor1(send1 ? q : 0u);
        if (step2) q += from_xor2(send2 ? q : 0u);
        if (step3) q += from_plus4(send3 ? q : 0u);
        if (alive) lds_add(base + (a + HX * (b + HY * c)), (long long)(unsigned long long)q);
      }
  long long cen = (long long)__float2uint_rn(ms) - (long long)rest;
  cen += from_xor1(send1 ? cen : 0ll);
  if (step2) cen += from_xor2(send2 ? cen : 0ll);
  if (step3) cen += from_plus4(send3 ? cen : 0ll);
  if (alive) lds_add(base + (1 + HX * (1 + HY * 1)), cen);
}

// base-grid cell centre RAD(i), i = 0..n+1: MALLA's own FP32 recurrence
// (RAD(i) = RAD(1)+(i-1)*C, ghost RAD(n+1) = RAD(n)+C) when the plan found the
// table to be exactly that, a table read otherwise
__device__ __forceinline__ float base_centre(const Plan& P, int a, int i, int n) {
  if (P.rad_malla) {
    if (i <= n) return P.rad1[a] + (float)(i - 1) * P.radc[a];
    return (P.rad1[a] + (float)(n - 1) * P.radc[a]) + P.radc[a];
  }
  const float* r = a == 0 ? P.radx : (a == 1 ? P.rady : P.radz);
  return r[i];
}

// tile-local coordinate of the stencil centre: on the periodic base grid the
// image of the (wrapped) centre that lies next to the tile
__device__ __forceinline__ int centre_local(int c, int org, int ext, int n, bool periodic) {
  if (!periodic) return c - org;
  int l = wrap1(c, n) - org;
  if (l > ext) l -= n;
  if (l < -1) l += n;
  return l;
}

// The wide entries of one class (kernel level irk, half-width N) of a work
// item, in the histogram form: the fixed-point masses of the class's members
// summed at their cells (the reference's IX, JY, KZ) on the tile's box
// extended by N cells (every member of the item has its cell there: its box
// meets the tile), exact int64 sums; then the kernel row applied along x, y
// and z in FP64, each pass in place on the buffer; the result, rounded once
// per cell, is added to the tile's accumulator.  Cells of the tile outside
// the grid receive values that are never written out.
template <int RED, int TX, int TY, int TZ, int HX, int HY>
__device__ void wide_histogram_form(const DepArgs& D, const Item& it, int irk, int i0, int j0, int k0c,
                                    const float* K, float scale, double dscale,
                                    typename AccT<RED>::T* acc, long long* hist) {
  using Acc = typename AccT<RED>::T;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int nk = 1 << irk;
  const int bx = TX + 2 * nk, by = TY + 2 * nk, bz = TZ + 2 * nk, nb = bx * by * bz;
  double* hd = reinterpret_cast<double*>(hist);
  auto at = [](int a, int b, int c) { return (a + 2) + HX * ((b + 2) + HY * (c + 2)); };
  for (int c = tid; c < nb; c += nthr) hist[c] = 0;
  __syncthreads();
  for (long long e = it.begin + tid; e < it.end; e += nthr) {
    int ci[3], irk_e;
    float m;
    unpack_table(D.entries[e], ci[0], ci[1], ci[2], irk_e, m);
    if (irk_e != irk) continue;
    const int hx = ci[0] - i0 + nk, hy = ci[1] - j0 + nk, hz = ci[2] - k0c + nk;
    if (hx < 0 || hx >= bx || hy < 0 || hy >= by || hz < 0 || hz >= bz) continue;
    atomicAdd(reinterpret_cast<unsigned long long*>(hist + hx + bx * (hy + by * hz)),
              (unsigned long long)__float2ll_rn(m * scale));
  }
  __syncthreads();
  for (int c = tid; c < nb; c += nthr) hd[c] = (double)hist[c];
  __syncthreads();
  // x: every (y, z) row of the box; out(xo) = sum_d row(xo + N - d) K(d),
  // written over row(xo), which no later output of the row reads
  for (int r = tid; r < by * bz; r += nthr) {
    double* row = hd + r * bx;
    for (int xo = 0; xo < TX; ++xo) {
      double s = 0.0;
      for (int d = -nk; d <= nk; ++d) s += row[xo + nk - d] * (double)K[d];
      row[xo] = s;
    }
  }
  __syncthreads();
  // y: every column (x of the tile, z of the box)
  for (int r = tid; r < TX * bz; r += nthr) {
    double* col = hd + (r % TX) + bx * by * (r / TX);
    for (int yo = 0; yo < TY; ++yo) {
      double s = 0.0;
      for (int d = -nk; d <= nk; ++d) s += col[bx * (yo + nk - d)] * (double)K[d];
      col[bx * yo] = s;
    }
  }
  __syncthreads();
  // z: every (x, y) of the tile, into the accumulator
  const double inv_scale = 1.0 / dscale;   // a power of two: exact
  for (int r = tid; r < TX * TY; r += nthr) {
    const int xo = r % TX, yo = r / TX;
    const double* col = hd + xo + bx * yo;
    for (int zo = 0; zo < TZ; ++zo) {
      double s = 0.0;
      for (int d = -nk; d <= nk; ++d) s += col[bx * by * (zo + nk - d)] * (double)K[d];
      if (s != 0.0) {
        Acc* dst = acc + at(xo, yo, zo);
        if (RED == 0) *reinterpret_cast<long long*>(dst) += __double2ll_rn(s);
        else *reinterpret_cast<float*>(dst) += (float)(s * inv_scale);
      }
    }
  }
  __syncthreads();
}

template <int MODE, int RED, int TX, int TY, int TZ>
__global__ void __launch_bounds__(1024) deposit_kernel(Plan P, DepArgs D) {
  using Acc = typename AccT<RED>::T;
  constexpr int T = TX * TY * TZ;
  // the tile with a 2-cell halo; x and y rows padded by one cell (a row of
  // 21 int64 cells instead of 20 spreads the stencil's adds over more LDS
  // banks: 6% less deposit time on clustered particles)
  constexpr int HX = TX + 5, HY = TY + 5, HZ = TZ + 4, H = HX * HY * HZ;
  __shared__ Acc acc[H];
  __shared__ unsigned s_item;
  __shared__ long long s_redi[16];
  __shared__ double s_redd[16];
  // table mode: the used part of the kernel table, row irk (2^(irk+1) + 1
  // values) at ktab_off(irk), centred at ktab_off(irk) + 2^irk
  __shared__ float s_ktab[MODE == 1 ? kKtabLds : 1];
  __shared__ unsigned s_ccnt[kNLevels + 1];
  // table mode: the histogram buffer of the wide classes (dynamic LDS,
  // D.hist_cells int64 cells)
  extern __shared__ long long s_hist[];
  const int tid = threadIdx.x, nthr = blockDim.x;
  const double dscale = fixed_scale(D.C->mmax_bits);
  const float scale = (float)dscale;
  const unsigned n_items = D.C->n_items_eff;
  const float k0[3] = {P.ktab[P.kmax - 1], P.ktab[P.kmax], P.ktab[P.kmax + 1]};
  if (MODE == 1) {
    for (int irk = 0; irk <= P.nl; ++irk) {
      const int w = 2 * (1 << irk) + 1;
      for (int j = tid; j < w; j += nthr)
        s_ktab[ktab_off(irk) + j] = P.ktab[irk * P.kw + P.kmax - (1 << irk) + j];
    }
    // the first item's __syncthreads orders these writes before any read
  }
  // interior cell (a, b, c) of the tile
  auto at = [](int a, int b, int c) { return (a + 2) + HX * ((b + 2) + HY * (c + 2)); };

  // workgroup w serves queue w mod n_queues, which hands out items
  // q, q + n_queues, ...; consecutive workgroups land on different XCDs, so
  // each queue counter lives in one XCD's L2
  const int q = blockIdx.x % D.n_queues;
  for (;;) {
    if (tid == 0) s_item = atomicAdd(&D.C->queue[q * kLine], 1u) * D.n_queues + q;
    __syncthreads();
    const unsigned idx = __builtin_amdgcn_readfirstlane(s_item);
    if (idx >= n_items) break;
    const Item it = D.items[idx];
    const int t = __builtin_amdgcn_readfirstlane(it.tile);
    const int g = __builtin_amdgcn_readfirstlane(P.tile_grid[t]);
    const GridDesc G = P.grids[g];
    const int local = t - G.tile0;
    const int tx = local % G.nt[0], ty = (local / G.nt[0]) % G.nt[1], tz = local / (G.nt[0] * G.nt[1]);
    const int i0 = tx * TX + 1, j0 = ty * TY + 1, k0c = tz * TZ + 1;
    const int ex = min(TX, G.n[0] - i0 + 1), ey = min(TY, G.n[1] - j0 + 1), ez = min(TZ, G.n[2] - k0c + 1);
    const bool periodic = g == 0;

    for (int c = tid; c < H; c += nthr) acc[c] = 0;
    __syncthreads();

    if (!(it.flags & 1)) {
      // ---- narrow entries: one lane per entry, 27 cells.  Lane j of the
      // workgroup takes entries j, j + nthr, ... of the item, so a wave reads
      // 64 consecutive entries (1 KB, every entry fetched once).  Two entries
      // per pass, in two registers refilled right after use (register copies
      // would wait on the loads); the prefetch indices are clamped into the
      // item instead of tested, so every pass issues the same loads, and the
      // loop runs the same number of passes on every lane (ok = the lane holds
      // an entry), which the lane merging of the TSC deposit needs.  An empty
      // item reads nothing: no pass runs.
      // the item's filled entries: [bA, bA + la) and [bB, bB + (len - la));
      // with the hierarchy binning the tile's home entries end at cursor[2t]
      // and its other entries run from start[2t+1] to cursor[2t+1]
      long long la = it.end - it.begin, gap = 0, len = la;
      if (D.cursor) {
        const long long s1 = (long long)D.start[2 * t + 1];
        la = max(min(min(it.end, s1), (long long)D.cursor[2 * t]) - it.begin, 0ll);
        const long long bB = max(it.begin, s1);
        gap = bB - (it.begin + la);
        len = la + max(min(it.end, (long long)D.cursor[2 * t + 1]) - bB, 0ll);
      }
      // wave-uniform: kept in scalar registers
      la = uniform64(la);
      gap = uniform64(gap);
      len = uniform64(len);
      const int4* const eb = D.entries + uniform64(it.begin);
      auto narrow = [&](const int4 v, const bool ok) {
        int ci[3];
        float wx[3], wy[3], wz[3], m;
        if (MODE == 2) {
          const float p[3] = {__int_as_float(v.x), __int_as_float(v.y), __int_as_float(v.z)};
          m = __int_as_float(v.w);
          float* w[3] = {wx, wy, wz};
          // weights from the three centres RAD(I-1..I+1) on the base grid,
          // MINIRADX(I-1..I+1) on a patch; the indices are the reference's
          if (g == 0) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              ci[a] = tsc_base_index(p[a], P.xl, G.d[a], G.inv_d[a], G.n[a]);
              tsc_weights_f64(p[a], base_centre(P, a, ci[a] - 1, G.n[a]), base_centre(P, a, ci[a], G.n[a]),
                              base_centre(P, a, ci[a] + 1, G.n[a]), G.inv_dd[a], w[a]);
            }
          } else {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              ci[a] = (int)index_quotient(p[a] - G.org[a], G.d[a], G.inv_d[a]);
              tsc_weights_f64(p[a], minirad(G.pr[a], ci[a] - 1, G.d[a]), minirad(G.pr[a], ci[a], G.d[a]),
                              minirad(G.pr[a], ci[a] + 1, G.d[a]), G.inv_dd[a], w[a]);
            }
          }
        } else {
          int irk;
          unpack_table(v, ci[0], ci[1], ci[2], irk, m);
#pragma unroll
          for (int k = 0; k < 3; ++k) { wx[k] = k0[k]; wy[k] = k0[k]; wz[k] = k0[k]; }
        }
        // in [-1, ext] by construction of the binning; clamped so the array
        // bounds never depend on that
        const int lx = clampi(centre_local(ci[0], i0, ex, G.n[0], periodic), -1, TX);
        const int ly = clampi(centre_local(ci[1], j0, ey, G.n[1], periodic), -1, TY);
        const int lz = clampi(centre_local(ci[2], k0c, ez, G.n[2], periodic), -1, TZ);
        // corner (-1,-1,-1) of the stencil; halo coordinates are local + 2
        const int corner = (lx + 1) + HX * ((ly + 1) + HY * (lz + 1));
        if (MODE == 2 || P.merge_narrow)
          deposit27_merged<RED, HX, HY>(acc, corner, ok, m, wx, wy, wz, scale);
        else if (ok)
          deposit27<RED, HX, HY>(acc + corner, m, wx, wy, wz, scale);
      };
      // entry j of the item (j clamped to the last one)
      auto at_j = [&](long long j) {
        j = min(j, len - 1);
        return eb[j + (j < la ? 0ll : gap)];
      };
      int4 va = make_int4(0, 0, 0, 0), vb = va;
      if (len > 0) {
        va = at_j(tid);
        vb = at_j(tid + nthr);
      }
      for (long long o = 0; o < len; o += 2 * nthr) {
        narrow(va, o + tid < len);
        va = at_j(o + 2 * nthr + tid);
        narrow(vb, o + nthr + tid < len);
        vb = at_j(o + 3 * nthr + tid);
      }
    } else if (MODE == 1) {
      // ---- wide table-mode entries (kernel level irk >= 1, half-width
      // N = 2^irk), one class at a time.  Classes whose halo-extended box
      // fits the histogram buffer: an NGP histogram of the members' masses
      // at their cells, then the separable kernel, three 1-D passes (the
      // histogram form of grids.f:888-951).  Larger classes: the direct walk
      // of every entry's clipped box. ----
      // entries per class (wave ballots, one LDS add per wave and class)
      if (tid <= kNLevels) s_ccnt[tid] = 0u;
      __syncthreads();
      {
        unsigned cnt[kNLevels + 1] = {};
        for (long long b0 = it.begin; b0 < it.end; b0 += nthr) {
          const long long e = b0 + tid;
          const int irk_e = e < it.end ? (D.entries[e].y >> 16) : 0;
#pragma unroll
          for (int k = 1; k <= kNLevels; ++k) cnt[k] += (unsigned)__popcll(__ballot(irk_e == k));
        }
        if ((tid & 63) == 0)
          for (int k = 1; k <= kNLevels; ++k)
            if (cnt[k]) atomicAdd(&s_ccnt[k], cnt[k]);
      }
      __syncthreads();
      for (int irk = 1; irk <= kNLevels; ++irk) {
        const unsigned ncls = s_ccnt[irk];
        if (ncls == 0u) continue;
        const int nk = 1 << irk;
        const float* K = s_ktab + ktab_off(irk) + nk;   // K[d], d in [-N, N]
        const int bx = TX + 2 * nk, by = TY + 2 * nk, bz = TZ + 2 * nk;
        // the histogram form when its buffer fits and its three passes cost
        // less than the entries' direct adds (an LDS add counted as four
        // FP64 multiply-adds)
        const long long w = 2 * nk + 1;
        const long long passes = w * ((long long)bx * by * TX + (long long)TX * bz * TY + (long long)TX * TY * TZ);
        if ((it.flags & 2) && bx * by * bz <= D.hist_cells && 4ll * ncls * w * w * w > passes) {
          wide_histogram_form<RED, TX, TY, TZ, HX, HY>(D, it, irk, i0, j0, k0c, K, scale, dscale, acc, s_hist);
          continue;
        }
        // direct: one wave per entry.  The entry's box clipped to the tile
        // (bx x by x bz cells) is walked as one flat range, lane l taking
        // cells l, l + 64, ...; its (x, y, z) counters advance by constant
        // steps, so the loop has no divisions.  Kernel values come from the
        // LDS copy of the table; the next entry is loaded while this one
        // deposits.  Terms are ((m*Kx)*Ky)*Kz as in the reference's loop.
        const int wave = tid >> 6, lane = tid & 63, nwaves = nthr >> 6;
        long long e = it.begin + wave;
        int4 nxt = e < it.end ? D.entries[e] : make_int4(0, 0, 0, 0);
        for (; e < it.end; e += nwaves) {
          int ci[3], irk_e;
          float m;
          unpack_table(nxt, ci[0], ci[1], ci[2], irk_e, m);
          if (e + nwaves < it.end) nxt = D.entries[e + nwaves];
          if (irk_e != irk) continue;
          const int xa = max(max(ci[0] - nk, 1), i0), xb = min(min(ci[0] + nk, G.n[0]), i0 + ex - 1);
          const int ya = max(max(ci[1] - nk, 1), j0), yb = min(min(ci[1] + nk, G.n[1]), j0 + ey - 1);
          const int za = max(max(ci[2] - nk, 1), k0c), zb = min(min(ci[2] + nk, G.n[2]), k0c + ez - 1);
          if (xb < xa || yb < ya || zb < za) continue;
          const int cx = xb - xa + 1, cy = yb - ya + 1, total = cx * cy * (zb - za + 1);
          // q / n for q <= 64 * 16 and n <= 16: (q + 1/2) / n is at least 1/32
          // away from an integer, far beyond the FP32 error of the product
          const float rbx = 1.0f / (float)cx, rby = 1.0f / (float)cy;
          auto divn = [](int q, float rn) { return (int)(((float)q + 0.5f) * rn); };
          const int r0 = divn(lane, rbx);
          int a = lane - r0 * cx;
          int c = divn(r0, rby);
          int b = r0 - c * cy;
          const int dr = divn(64, rbx), da = 64 - dr * cx;
          const int dc0 = divn(dr, rby), db0 = dr - dc0 * cy;
          const int dc1 = divn(dr + 1, rby), db1 = dr + 1 - dc1 * cy;
          const float ms = RED == 0 ? m * scale : m;
          const float* Kx = K + (xa - ci[0]);
          const float* Ky = K + (ya - ci[1]);
          const float* Kz = K + (za - ci[2]);
          Acc* base = acc + at(xa - i0, ya - j0, za - k0c);
          for (int q2 = lane; q2 < total; q2 += 64) {
            const float term = ((ms * Kx[a]) * Ky[b]) * Kz[c];
            Acc* dst = base + (a + HX * (b + HY * c));
            if (RED == 0) lds_add(reinterpret_cast<long long*>(dst), (long long)__float2int_rn(term));
            else lds_add(reinterpret_cast<float*>(dst), term);
            a += da;
            const bool carry = a >= cx;
            if (carry) a -= cx;
            b += carry ? db1 : db0;
            c += carry ? dc1 : dc0;
            if (b >= cy) {
              b -= cy;
              ++c;
            }
          }
        }
        __syncthreads();
      }
    }
    __syncthreads();

    // ---- a tile spanning a whole periodic axis: its halo cells are its own
    // interior cells under the reference's wrap; fold them in ----
    if (periodic && (G.nt[0] == 1 || G.nt[1] == 1 || G.nt[2] == 1)) {
      for (int c = tid; c < H; c += nthr) {
        const int a = c % HX - 2, b = (c / HX) % HY - 2, cz = c / (HX * HY) - 2;
        if (a >= 0 && a < ex && b >= 0 && b < ey && cz >= 0 && cz < ez) continue;
        const int ga = wrap1(i0 + a, G.n[0]) - i0, gb = wrap1(j0 + b, G.n[1]) - j0,
                  gc = wrap1(k0c + cz, G.n[2]) - k0c;
        if (ga < 0 || ga >= ex || gb < 0 || gb >= ey || gc < 0 || gc >= ez) continue;
        const Acc v = acc[c];
        if (v != 0) lds_add(acc + at(ga, gb, gc), v);
      }
      __syncthreads();
    }

    // ---- split tile: every item adds its interior into the tile's
    // accumulator in global memory (int64 adds are exact, so the sum does not
    // depend on the order of the items); split_reduce writes the cells ----
    const int nit = D.nitems_tile[t];
    const bool writer = nit == 1;
    if (!writer) {
      Acc* dst = reinterpret_cast<Acc*>(D.scratch) + (long long)D.split_slot[t] * T;
      for (int c = tid; c < T; c += nthr) {
        const Acc v = acc[at(c % TX, (c / TX) % TY, c / (TX * TY))];
        if (v != 0) glb_add(dst + c, v);
      }
    }
    if (writer) {
      float* out = (g == 0 ? D.u1 : D.u11) + G.off;
      long long msi = 0;
      double msd = 0.0;
      const double inv_scale = 1.0 / dscale;   // a power of two: exact
      for (int c = tid; c < T; c += nthr) {
        const int a = c % TX, b = (c / TX) % TY, cz = c / (TX * TY);
        if (a >= ex || b >= ey || cz >= ez) continue;
        const Acc v = acc[at(a, b, cz)];
        if (RED == 0) msi += (long long)v;
        else msd += (double)v;
        float f = RED == 0 ? (float)((double)v * inv_scale) : (float)v;
        f = f / G.den;
        f = f / P.fdm;
        out[(long long)(i0 - 1 + a) + (long long)G.ld1 * (j0 - 1 + b) + (long long)G.ld2 * (k0c - 1 + cz)] = f;
      }
      if (RED == 0) {
        msi = block_sum(msi, s_redi);
        if (tid == 0) D.tile_mass_i[t] = msi;
      } else {
        msd = block_sum(msd, s_redd);
        if (tid == 0) D.tile_mass_d[t] = msd;
      }
    }
    __syncthreads();
  }
}

// Split tiles: the accumulated interiors normalised and written as the
// deposit's writer does.  Work unit = (split tile, kPer * 256 consecutive
// cells, kPer per thread with every load issued before the first store), so a
// heavy tile is spread over many workgroups; the exact int64 (FP64 in atomic
// mode) tile mass is accumulated per unit.
template <int RED, int TX, int TY, int TZ>
__global__ void __launch_bounds__(256) split_reduce(Plan P, DepArgs D) {
  using Acc = typename AccT<RED>::T;
  constexpr int T = TX * TY * TZ;
  constexpr int kPer = 4;
  constexpr int kCells = 256 * kPer;
  constexpr int cpt = (T + kCells - 1) / kCells;
  __shared__ long long s_redi[4];
  __shared__ double s_redd[4];
  if (D.C->overflow) return;
  const double inv_scale = 1.0 / fixed_scale(D.C->mmax_bits);
  const unsigned long long units = *D.n_split_tiles * cpt;
  for (unsigned long long u = blockIdx.x; u < units; u += gridDim.x) {
    const int t = D.split_tiles[u / cpt];
    const int c0 = (int)(u % cpt) * kCells + threadIdx.x;
    const int g = P.tile_grid[t];
    const GridDesc& G = P.grids[g];
    const int local = t - G.tile0;
    const int tx = local % G.nt[0], ty = (local / G.nt[0]) % G.nt[1], tz = local / (G.nt[0] * G.nt[1]);
    const int i0 = tx * TX + 1, j0 = ty * TY + 1, k0c = tz * TZ + 1;
    const int ex = min(TX, G.n[0] - i0 + 1), ey = min(TY, G.n[1] - j0 + 1), ez = min(TZ, G.n[2] - k0c + 1);
    const Acc* part = reinterpret_cast<const Acc*>(D.scratch) + (long long)D.split_slot[t] * T;
    float* out = (g == 0 ? D.u1 : D.u11) + G.off;
    Acc v[kPer];
    bool in[kPer];
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
      const int c = c0 + r * 256;
      const int a = c % TX, b = (c / TX) % TY, cz = c / (TX * TY);
      in[r] = c < T && a < ex && b < ey && cz < ez;
      v[r] = in[r] ? part[c] : (Acc)0;
    }
    long long msi = 0;
    double msd = 0.0;
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
      if (!in[r]) continue;
      const int c = c0 + r * 256;
      const int a = c % TX, b = (c / TX) % TY, cz = c / (TX * TY);
      if (RED == 0) msi += (long long)v[r];
      else msd += (double)v[r];
      float f = RED == 0 ? (float)((double)v[r] * inv_scale) : (float)v[r];
      f = f / G.den;
      f = f / P.fdm;
      out[(long long)(i0 - 1 + a) + (long long)G.ld1 * (j0 - 1 + b) + (long long)G.ld2 * (k0c - 1 + cz)] = f;
    }
    if (RED == 0) {
      msi = block_sum(msi, s_redi);
      if (threadIdx.x == 0 && msi != 0)
        atomicAdd(reinterpret_cast<unsigned long long*>(D.tile_mass_i + t), (unsigned long long)msi);
    } else {
      msd = block_sum(msd, s_redd);
      if (threadIdx.x == 0 && msd != 0.0) atomicAdd(D.tile_mass_d + t, msd);
    }
  }
}

// clears the accumulator tiles of this call's split tiles (their count is on
// the device); a kernel, not a memset, so the build can be captured
__global__ void __launch_bounds__(256) zero_split(DepArgs D, int T, int elem_words) {
  if (D.C->overflow) return;
  const unsigned long long ns = *D.n_split_tiles;
  const unsigned long long n = ns * (unsigned long long)T * (unsigned long long)elem_words;
  unsigned* p = reinterpret_cast<unsigned*>(D.scratch);
  for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < n;
       i += (unsigned long long)gridDim.x * blockDim.x)
    p[i] = 0u;
  // the split tiles' masses, which split_reduce accumulates (zero bits are 0
  // as int64 and as FP64)
  for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < ns;
       i += (unsigned long long)gridDim.x * blockDim.x)
    D.tile_mass_i[D.split_tiles[i]] = 0;
}

// per-grid mass in a fixed order; FP64 (int64-exact in deterministic mode)
template <int RED>
__global__ void __launch_bounds__(256) grid_sums(Plan P, DepArgs D, double* grid_mass) {
  __shared__ long long shi[16];
  __shared__ double shd[16];
  const int g = blockIdx.x;
  const GridDesc& G = P.grids[g];
  const int nt = G.level == 0 && g != 0 ? 0 : G.nt[0] * G.nt[1] * G.nt[2];
  if (RED == 0) {
    long long s = 0;
    for (int k = threadIdx.x; k < nt; k += blockDim.x) s += D.tile_mass_i[G.tile0 + k];
    s = block_sum(s, shi);
    if (threadIdx.x == 0) grid_mass[g] = (double)s / fixed_scale(D.C->mmax_bits);
  } else {
    double s = 0;
    for (int k = threadIdx.x; k < nt; k += blockDim.x) s += D.tile_mass_d[G.tile0 + k];
    s = block_sum(s, shd);
    if (threadIdx.x == 0) grid_mass[g] = s;
  }
}

// ---------------------------------------------------------------------------
// adjoint of the TSC mode: one thread per particle, gathers the upstream
// gradient at every stencil cell of every grid the particle belongs to
// ---------------------------------------------------------------------------
// three consecutive cells of a row in one load (dword alignment suffices)
struct __attribute__((aligned(4))) Row3 {
  float v[3];
};
__device__ __forceinline__ void load_row3(const float* p, float out[3]) {
  const Row3 r = *reinterpret_cast<const Row3*>(p);
  out[0] = r.v[0];
  out[1] = r.v[1];
  out[2] = r.v[2];
}

// Base-cell key (x fastest) of every particle, for the gather order of the
// adjoint: lanes of a wave that take particles of one cell read the same
// rows of the upstream gradient.  The pass also makes the TSC input checks
// (non-finite values, negative masses, positions outside the base range:
// the flags of the forward's validation), as it reads the positions anyway.
__global__ void __launch_bounds__(256) cell_keys(Plan P, Parts A, unsigned* key, unsigned* idx, Counters* C) {
  const GridDesc& G = P.grids[0];
  unsigned err = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < A.n;
       i += (long long)gridDim.x * blockDim.x) {
    const float p[3] = {A.x[i], A.y[i], A.z[i]};
    const float m = A.m[i];
    if (!(isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2]) && isfinite(m))) {
      err |= kErrNonFinite;
    } else {
      if (m < 0.f) err |= kErrNegMass;
#pragma unroll
      for (int a = 0; a < 3; ++a)
        if (!tsc_base_in_range(p[a], P.xl, G.inv_d[a], G.n[a])) err |= kErrWrap;
    }
    int c[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) c[a] = clampi(tsc_base_index(p[a], P.xl, G.d[a], G.inv_d[a], G.n[a]), 1, G.n[a]) - 1;
    key[i] = (unsigned)c[0] + (unsigned)G.n[0] * ((unsigned)c[1] + (unsigned)G.n[1] * (unsigned)c[2]);
    idx[i] = (unsigned)i;
  }
  err = wave_or(err);
  if ((threadIdx.x & 63) == 0 && err) atomicOr(&C->err, err);
}

// One (particle, grid) term of the adjoint: the nine rows of three gradient
// values around the centre cell, weighted.  Per row (b, c):
// sx = sum_a wx_a g, sdx = sum_a wx'_a g; then am += wy wz sx,
// ax += wy wz sdx, ay += wy' wz sx, az += wy dwz sx.
struct AdjAcc {
  float am = 0.f, ax = 0.f, ay = 0.f, az = 0.f;
  __device__ __forceinline__ void row(const float gv[3], const float w[3][3], const float dw[3][3], int b, int c,
                                      float sc) {
    const float sx = (w[0][0] * gv[0] + w[0][1] * gv[1] + w[0][2] * gv[2]) * sc;
    const float sdx = (dw[0][0] * gv[0] + dw[0][1] * gv[1] + dw[0][2] * gv[2]) * sc;
    const float yz = w[1][b] * w[2][c];
    am += yz * sx;
    ax += yz * sdx;
    ay += dw[1][b] * w[2][c] * sx;
    az += w[1][b] * dw[2][c] * sx;
  }
};

// One thread per particle, in the given order (NULL: index order).  The
// result of a particle does not depend on the order.  Eight workgroups per
// CU (at most 64 VGPRs, a few of them spilled): the gather waits on memory,
// and the extra waves in flight gain more than the spills cost.
__global__ void __launch_bounds__(256, 8) adjoint_gather(Plan P, Parts A, const float* g1, const float* g11,
                                                      float* dm, float* dxo, float* dyo, float* dzo,
                                                      const unsigned* order) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < A.n;
       t += (long long)gridDim.x * blockDim.x) {
    const long long i = order ? (long long)order[t] : t;
    if (i >= A.n) continue;   // the order is a permutation of [0, n); checked so a bad one cannot read out of range
    const float p[3] = {A.x[i], A.y[i], A.z[i]};
    const float m = A.m[i];
    AdjAcc acc;
    // the first level's hash slot is loaded while the base grid is handled
    Probe nxt;
    if (P.nl >= 1) nxt = probe_start(P, 1, p[0], p[1], p[2]);
    // base grid
    {
      const GridDesc& G = P.grids[0];
      int ci[3];
      float w[3][3], dw[3][3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        ci[a] = tsc_base_index(p[a], P.xl, G.d[a], G.inv_d[a], G.n[a]);
        tsc_weights_dc(p[a], base_centre(P, a, ci[a] - 1, G.n[a]), base_centre(P, a, ci[a], G.n[a]),
                       base_centre(P, a, ci[a] + 1, G.n[a]), G.inv_d[a], w[a], dw[a]);
      }
      const float sc = 1.0f / (G.den * P.fdm);
      // the three x cells of a row are contiguous unless the row wraps
      const bool xin = ci[0] >= 2 && ci[0] <= G.n[0] - 1;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const long long kz = wrap1(ci[2] - 1 + c, G.n[2]) - 1;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          const long long jy = wrap1(ci[1] - 1 + b, G.n[1]) - 1;
          const float* row = g1 + G.off + G.ld1 * jy + G.ld2 * kz;
          float gv[3];
          if (xin) {
            load_row3(row + (ci[0] - 2), gv);
          } else {
#pragma unroll
            for (int a = 0; a < 3; ++a) gv[a] = row[wrap1(ci[0] - 1 + a, G.n[0]) - 1];
          }
          acc.row(gv, w, dw, b, c, sc);
        }
      }
    }
    // the descent stops where the forward enumeration stops (nested boxes)
    for (int l = 1; l <= P.nl; ++l) {
      const int2 se = probe_finish(P, nxt);
      if (l < P.nl) nxt = probe_start(P, l + 1, p[0], p[1], p[2]);
      bool any = false;
      for (int q = 0; q < se.y; ++q) {
        const BoxRec& G = P.hrec[se.x + q];   // the grid's fields travel in its candidate record
        if (!in_box(G, p[0], p[1], p[2])) continue;
        any = true;
        int ci[3];
        float w[3][3], dw[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          ci[a] = (int)index_quotient(p[a] - G.org[a], G.d[a], G.inv_d[a]);
          tsc_weights_dc(p[a], minirad(G.pr[a], ci[a] - 1, G.d[a]), minirad(G.pr[a], ci[a], G.d[a]),
                         minirad(G.pr[a], ci[a] + 1, G.d[a]), G.inv_d[a], w[a], dw[a]);
        }
        const float sc = 1.0f / (G.den * P.fdm);
        const float* gp = g11 + G.off;
        const bool xin = ci[0] >= 2 && ci[0] <= G.n[0] - 1;
        const long long ld1 = G.ld1, ld2 = G.ld2;
        const int gn0 = G.n[0], gn1 = G.n[1], gn2 = G.n[2];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const int kz = ci[2] - 1 + c;
          if (kz < 1 || kz > gn2) continue;
#pragma unroll
          for (int b = 0; b < 3; ++b) {
            const int jy = ci[1] - 1 + b;
            if (jy < 1 || jy > gn1) continue;
            const float* row = gp + ld1 * (jy - 1) + ld2 * (kz - 1);
            float gv[3];
            if (xin) {
              load_row3(row + (ci[0] - 2), gv);
            } else {
#pragma unroll
              for (int a = 0; a < 3; ++a) {
                const int ix = ci[0] - 1 + a;
                gv[a] = (ix >= 1 && ix <= gn0) ? row[ix - 1] : 0.f;
              }
            }
            acc.row(gv, w, dw, b, c, sc);
          }
        }
      }
      if (!any && P.nested) break;
    }
    dm[i] = acc.am;
    dxo[i] = m * acc.ax;
    dyo[i] = m * acc.ay;
    dzo[i] = m * acc.az;
  }
}

// ---------------------------------------------------------------------------
// plain form: one launch per grid, global FP32 atomics into the outputs
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) plain_zero(Plan P, float* u1, float* u11) {
  const GridDesc& G = P.grids[blockIdx.x];
  if (G.level == 0 && blockIdx.x != 0) return;
  float* out = (blockIdx.x == 0 ? u1 : u11) + G.off;
  const long long nc = (long long)G.n[0] * G.n[1] * G.n[2];
  for (long long c = threadIdx.x; c < nc; c += blockDim.x) {
    const long long a = c % G.n[0], r = c / G.n[0], b = r % G.n[1], cz = r / G.n[1];
    out[a + G.ld1 * b + G.ld2 * cz] = 0.f;
  }
}

template <int MODE>
__global__ void __launch_bounds__(256) plain_base(Plan P, Parts A, float* u1) {
  const GridDesc& G = P.grids[0];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < A.n;
       i += (long long)gridDim.x * blockDim.x) {
    const float p[3] = {A.x[i], A.y[i], A.z[i]};
    const float m = A.m[i];
    int ci[3];
    float w[3][3];
    if (MODE == 2) {
      const float* rad[3] = {P.radx, P.rady, P.radz};
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        ci[a] = tsc_base_index(p[a], P.xl, G.d[a], G.inv_d[a], G.n[a]);
        tsc_weights(p[a], rad[a][ci[a] - 1], rad[a][ci[a]], rad[a][ci[a] + 1], G.d[a], w[a]);
      }
    } else {
      bool ok = true;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const float t = (p[a] - P.xl) / G.d[a];
        ok = ok && t > -1.0e9f && t < 1.0e9f;
        ci[a] = ok ? (int)t + 1 : 1;
        ok = ok && ci[a] >= 2 - G.n[a] && ci[a] <= 2 * G.n[a] - 1;
        w[a][0] = P.ktab[P.kmax - 1]; w[a][1] = P.ktab[P.kmax]; w[a][2] = P.ktab[P.kmax + 1];
      }
      if (!ok) continue;   // rejected by the count pass of the same call
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long kz = wrap1(ci[2] - 1 + c, G.n[2]) - 1;
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const long long jy = wrap1(ci[1] - 1 + b, G.n[1]) - 1;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const long long ix = wrap1(ci[0] - 1 + a, G.n[0]) - 1;
          atomicAdd(u1 + G.off + ix + G.ld1 * jy + G.ld2 * kz, m * w[0][a] * w[1][b] * w[2][c]);
        }
      }
    }
  }
}

// input validation of the plain form (same flags as the count pass)
template <int MODE>
__global__ void __launch_bounds__(256) plain_validate(Plan P, Parts A, Counters* C) {
  unsigned err = 0;
  const GridDesc& B = P.grids[0];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < A.n;
       i += (long long)gridDim.x * blockDim.x) {
    const float p[3] = {A.x[i], A.y[i], A.z[i]};
    const float m = A.m[i];
    if (!(isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2]) && isfinite(m))) { err |= kErrNonFinite; continue; }
    if (m < 0.f) err |= kErrNegMass;
    if (MODE == 2) {
#pragma unroll
      for (int a = 0; a < 3; ++a)
        if (!tsc_base_in_range(p[a], P.xl, B.inv_d[a], B.n[a])) err |= kErrWrap;
    }
    if (MODE == 1) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const float t = (p[a] - P.xl) / B.d[a];
        if (!(t > -1.0e9f && t < 1.0e9f)) { err |= kErrWrap; break; }
        const int c = (int)t + 1;
        if (c < 2 - B.n[a] || c > 2 * B.n[a] - 1) err |= kErrWrap;
      }
    }
  }
  err = wave_or(err);
  if ((threadIdx.x & 63) == 0 && err) atomicOr(&C->err, err);
}

// level membership flags for the plain form's per-level compaction: flag = 1
// when the particle lies in some level-l box (TSC: the PLEV rule of
// grids.f:1114-1157; table mode: the DX-extended boxes of grids.f:871-909)
__global__ void __launch_bounds__(256) plain_level_flags(Plan P, Parts A, int l, unsigned* flag) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < A.n;
       i += (long long)gridDim.x * blockDim.x) {
    const float x = A.x[i], y = A.y[i], z = A.z[i];
    unsigned f = 0;
    const int2 se = bucket_lookup(P, l, x, y, z);
    for (int q = 0; q < se.y && !f; ++q)
      if (in_box(P.hrec[se.x + q], x, y, z)) f = 1;
    flag[i] = f;
  }
}

__global__ void plain_compact(const unsigned* flag, const unsigned long long* start, long long n,
                              long long* idx) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    if (flag[i]) idx[start[i]] = i;
}

template <int MODE>
__global__ void __launch_bounds__(256) plain_patch(Plan P, Parts A, int g, const long long* idx,
                                                   const unsigned long long* nidx_p, float* u11) {
  const GridDesc& G = P.grids[g];
  const long long nidx = (long long)*nidx_p;
  float* out = u11 + G.off;
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < nidx;
       k += (long long)gridDim.x * blockDim.x) {
    const long long i = idx[k];
    const float p[3] = {A.x[i], A.y[i], A.z[i]};
    if (!(G.lo[0] < p[0] && p[0] < G.hi[0] && G.lo[1] < p[1] && p[1] < G.hi[1] && G.lo[2] < p[2] && p[2] < G.hi[2]))
      continue;
    const float m = A.m[i];
    if (MODE == 2) {
      int ci[3];
      float w[3][3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        ci[a] = (int)((p[a] - G.org[a]) / G.d[a]);
        tsc_weights(p[a], minirad(G.pr[a], ci[a] - 1, G.d[a]), minirad(G.pr[a], ci[a], G.d[a]),
                    minirad(G.pr[a], ci[a] + 1, G.d[a]), G.d[a], w[a]);
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int kz = ci[2] - 1 + c;
        if (kz < 1 || kz > G.n[2]) continue;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          const int jy = ci[1] - 1 + b;
          if (jy < 1 || jy > G.n[1]) continue;
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            const int ix = ci[0] - 1 + a;
            if (ix < 1 || ix > G.n[0]) continue;
            atomicAdd(out + (ix - 1) + (long long)G.ld1 * (jy - 1) + (long long)G.ld2 * (kz - 1),
                      m * w[0][a] * w[1][b] * w[2][c]);
          }
        }
      }
    } else {
      const int s = species_of(A, i);
      const int irk = max(G.level - s, 0), nk = 1 << irk;
      const float* K = P.ktab + irk * P.kw + P.kmax;
      int ci[3], lo[3], hi[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        ci[a] = (int)floorf((p[a] - G.org[a]) / G.d[a]) + 1;
        lo[a] = max(ci[a] - nk, 1);
        hi[a] = min(ci[a] + nk, G.n[a]);
      }
      for (int kk = lo[2]; kk <= hi[2]; ++kk)
        for (int jj = lo[1]; jj <= hi[1]; ++jj)
          for (int ii = lo[0]; ii <= hi[0]; ++ii)
            atomicAdd(out + (ii - 1) + (long long)G.ld1 * (jj - 1) + (long long)G.ld2 * (kk - 1),
                      m * K[ii - ci[0]] * K[jj - ci[1]] * K[kk - ci[2]]);
    }
  }
}

// normalise every grid (/DENBAS, then /FDM) and sum its mass in FP64
__global__ void __launch_bounds__(256) plain_finish(Plan P, float* u1, float* u11, double* grid_mass) {
  __shared__ double sh[16];
  const int g = blockIdx.x;
  const GridDesc& G = P.grids[g];
  if (G.level == 0 && g != 0) {
    if (threadIdx.x == 0 && grid_mass) grid_mass[g] = 0.0;
    return;
  }
  float* out = (g == 0 ? u1 : u11) + G.off;
  const long long nc = (long long)G.n[0] * G.n[1] * G.n[2];
  double s = 0.0;
  for (long long c = threadIdx.x; c < nc; c += blockDim.x) {
    const long long a = c % G.n[0], r = c / G.n[0], b = r % G.n[1], cz = r / G.n[1];
    float* q = out + a + G.ld1 * b + G.ld2 * cz;
    const float v = *q;
    s += v;
    *q = (v / G.den) / P.fdm;
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0 && grid_mass) grid_mass[g] = s;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static int grid_for(long long n, int block, int cap) {
  long long b = (n + block - 1) / block;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (int)b;
}

hipError_t launch_fill_u32(unsigned* p, unsigned v, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  fill_u32_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(p, v, n);
  return hipGetLastError();
}

long long bin_blocks(long long n) { return (n + kBinChunk - 1) / kBinChunk; }

hipError_t launch_count(const Plan& P, const Parts& A, unsigned* cnt, Counters* C, const BlockTables& BT,
                        int cus, hipStream_t s) {
  (void)cus;
  if (A.n <= 0) return hipSuccess;
  const int grid = (int)bin_blocks(A.n);
  if (P.mode == 2) count_kernel<2><<<grid, 256, 0, s>>>(P, A, cnt, C, BT);
  else count_kernel<1><<<grid, 256, 0, s>>>(P, A, cnt, C, BT);
  return hipGetLastError();
}

hipError_t launch_scatter(const Plan& P, const Parts& A, unsigned long long* cursor, int4* entries,
                          unsigned long long cap, const BlockTables& BT, const Counters* C, int cus,
                          hipStream_t s) {
  (void)cus;
  if (A.n <= 0) return hipSuccess;
  const int grid = (int)bin_blocks(A.n);
  if (P.mode == 2) scatter_kernel<2><<<grid, 256, 0, s>>>(P, A, cursor, entries, cap, BT, C);
  else scatter_kernel<1><<<grid, 256, 0, s>>>(P, A, cursor, entries, cap, BT, C);
  return hipGetLastError();
}

hipError_t launch_membership(const Plan& P, const Parts& A, int* out, long long cap,
                             unsigned long long* count, int cus, hipStream_t s) {
  const int grid = grid_for(A.n, 256, cus * 16);
  if (P.mode == 2) membership_kernel<2><<<grid, 256, 0, s>>>(P, A, out, cap, count);
  else membership_kernel<1><<<grid, 256, 0, s>>>(P, A, out, cap, count);
  return hipGetLastError();
}

long long scan_scratch_elems(long long n) { return (n + kScanTile - 1) / kScanTile + 1; }

hipError_t launch_scan(const unsigned* in, long long n, unsigned long long* bsum, unsigned long long* start,
                       unsigned long long* cursor, hipStream_t s) {
  const int nb = (int)((n + kScanTile - 1) / kScanTile);
  if (nb > 0) scan_reduce<<<nb, kScanThreads, 0, s>>>(in, n, bsum);
  scan_bsums<<<1, kScanThreads, 0, s>>>(bsum, nb, start + n);
  if (nb > 0) scan_apply<<<nb, kScanThreads, 0, s>>>(in, n, bsum, start, cursor);
  return hipGetLastError();
}

hipError_t launch_reduce_blocks(const BlockTables& BT, long long nblk, Counters* C, int cus, hipStream_t s) {
  if (nblk <= 0) return hipSuccess;
  const long long want = (nblk + 255) / 256;
  reduce_blocks<<<(int)std::min<long long>(want, (long long)std::max(cus, 1) * 4), 256, 0, s>>>(BT, nblk, C);
  return hipGetLastError();
}

hipError_t launch_plan_count(const unsigned long long* start, int ntiles, int chunk, int chunk_wide,
                             unsigned* nitems_tile, unsigned* nsplit_tile, unsigned* is_split, hipStream_t s) {
  plan_count<<<grid_for(ntiles, 256, 1 << 30), 256, 0, s>>>(start, ntiles, chunk, chunk_wide, nitems_tile,
                                                           nsplit_tile, is_split);
  return hipGetLastError();
}

hipError_t launch_plan_write(const unsigned long long* start, int ntiles, int chunk, int chunk_wide,
                             const unsigned long long* item_off, const unsigned long long* split_off,
                             const unsigned long long* split_rank, Item* items, int* split_slot,
                             int* split_tiles, long long* tile_mass, const Counters* C, hipStream_t s) {
  plan_write<<<grid_for(ntiles, 256, 1 << 30), 256, 0, s>>>(start, ntiles, chunk, chunk_wide, item_off,
                                                           split_off, split_rank, items, split_slot,
                                                           split_tiles, tile_mass, C);
  return hipGetLastError();
}

hipError_t launch_gate(const unsigned long long* tot_entries, const unsigned long long* tot_items,
                       const unsigned long long* tot_split, unsigned long long cap_entries,
                       unsigned long long cap_items, unsigned long long cap_split, Counters* C, hipStream_t s) {
  gate_kernel<<<1, 64, 0, s>>>(tot_entries, tot_items, tot_split, cap_entries, cap_items, cap_split, C);
  return hipGetLastError();
}

// Table mode: the histogram buffer of the wide classes takes the LDS left
// beside the kernel's static arrays, sized for the largest class of the
// geometry whose extended box fits there (smaller classes use a part of it).
template <int MODE, int RED, int TX, int TY, int TZ>
static hipError_t dep(const Plan& P, const DepArgs& D, int grid, int block, hipStream_t s) {
  DepArgs d2 = D;
  d2.hist_cells = 0;
  size_t dyn = 0;
  if (MODE == 1) {
    static long long avail = -1;   // int64 cells of LDS beside the static arrays (per instantiation)
    if (avail < 0) {
      int dev = 0, maxl = 0;
      hipFuncAttributes fa;
      hipError_t e = hipGetDevice(&dev);
      if (e == hipSuccess) e = hipDeviceGetAttribute(&maxl, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
      if (e == hipSuccess)
        e = hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&deposit_kernel<MODE, RED, TX, TY, TZ>));
      if (e != hipSuccess) return e;
      avail = std::max<long long>(((long long)maxl - (long long)fa.sharedSizeBytes - 256) / 8, 0);
      if (avail > 0) {
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&deposit_kernel<MODE, RED, TX, TY, TZ>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)(avail * 8));
        if (e != hipSuccess) return e;
      }
    }
    for (int irk = 1; irk <= P.nl; ++irk) {
      const long long nk = 1ll << irk;
      const long long nb = (TX + 2 * nk) * (TY + 2 * nk) * (TZ + 2 * nk);
      if (nb <= avail) d2.hist_cells = (int)nb;
    }
    dyn = sizeof(long long) * (size_t)d2.hist_cells;
  }
  deposit_kernel<MODE, RED, TX, TY, TZ><<<grid, block, dyn, s>>>(P, d2);
  return hipGetLastError();
}

template <int MODE, int RED>
static hipError_t dep_shape(const Plan& P, const DepArgs& D, int grid, int block, hipStream_t s) {
  const int* t = P.tile;
  if (t[0] == 8 && t[1] == 8 && t[2] == 8) return dep<MODE, RED, 8, 8, 8>(P, D, grid, block, s);
  if (t[0] == 16 && t[1] == 8 && t[2] == 8) return dep<MODE, RED, 16, 8, 8>(P, D, grid, block, s);
  if (t[0] == 16 && t[1] == 16 && t[2] == 8) return dep<MODE, RED, 16, 16, 8>(P, D, grid, block, s);
  if (t[0] == 16 && t[1] == 16 && t[2] == 16) return dep<MODE, RED, 16, 16, 16>(P, D, grid, block, s);
  return hipErrorInvalidValue;
}

bool tile_shape_supported(const int t[3]) {
  return (t[0] == 8 && t[1] == 8 && t[2] == 8) || (t[0] == 16 && t[1] == 8 && t[2] == 8) ||
         (t[0] == 16 && t[1] == 16 && t[2] == 8) || (t[0] == 16 && t[1] == 16 && t[2] == 16);
}

template <int RED>
static hipError_t split_shape(const Plan& P, const DepArgs& D, int grid, hipStream_t s) {
  const int* t = P.tile;
  if (t[0] == 8 && t[1] == 8 && t[2] == 8) split_reduce<RED, 8, 8, 8><<<grid, 256, 0, s>>>(P, D);
  else if (t[0] == 16 && t[1] == 8 && t[2] == 8) split_reduce<RED, 16, 8, 8><<<grid, 256, 0, s>>>(P, D);
  else if (t[0] == 16 && t[1] == 16 && t[2] == 8) split_reduce<RED, 16, 16, 8><<<grid, 256, 0, s>>>(P, D);
  else if (t[0] == 16 && t[1] == 16 && t[2] == 16) split_reduce<RED, 16, 16, 16><<<grid, 256, 0, s>>>(P, D);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_zero_split(const Plan& P, const DepArgs& D, int reduction, int cus, hipStream_t s) {
  const int T = P.tile[0] * P.tile[1] * P.tile[2];
  zero_split<<<(cus > 1 ? cus : 1) * 8, 256, 0, s>>>(D, T, reduction == 0 ? 2 : 1);
  return hipGetLastError();
}

hipError_t launch_split_reduce(const Plan& P, const DepArgs& D, int reduction, int cus, hipStream_t s) {
  const int grid = (cus > 1 ? cus : 1) * 8;
  return reduction == 0 ? split_shape<0>(P, D, grid, s) : split_shape<1>(P, D, grid, s);
}

hipError_t launch_deposit(const Plan& P, const DepArgs& D, int reduction, int grid, int block, hipStream_t s) {
  if (P.mode == 2) return reduction == 0 ? dep_shape<2, 0>(P, D, grid, block, s) : dep_shape<2, 1>(P, D, grid, block, s);
  return reduction == 0 ? dep_shape<1, 0>(P, D, grid, block, s) : dep_shape<1, 1>(P, D, grid, block, s);
}

hipError_t launch_grid_sums(const Plan& P, const DepArgs& D, int reduction, double* grid_mass, hipStream_t s) {
  if (P.ngrids <= 0) return hipSuccess;
  if (reduction == 0) grid_sums<0><<<P.ngrids, 256, 0, s>>>(P, D, grid_mass);
  else grid_sums<1><<<P.ngrids, 256, 0, s>>>(P, D, grid_mass);
  return hipGetLastError();
}

hipError_t launch_cell_keys(const Plan& P, const Parts& A, unsigned* key, unsigned* idx, Counters* C, int cus,
                            hipStream_t s) {
  cell_keys<<<grid_for(A.n, 256, cus * 16), 256, 0, s>>>(P, A, key, idx, C);
  return hipGetLastError();
}

hipError_t launch_adjoint(const Plan& P, const Parts& A, const float* g1, const float* g11, float* dm,
                          float* dx, float* dy, float* dz, const unsigned* order, int cus, hipStream_t s) {
  adjoint_gather<<<grid_for(A.n, 256, cus * 16), 256, 0, s>>>(P, A, g1, g11, dm, dx, dy, dz, order);
  return hipGetLastError();
}

hipError_t launch_plain_zero(const Plan& P, float* u1, float* u11, hipStream_t s) {
  plain_zero<<<P.ngrids, 256, 0, s>>>(P, u1, u11);
  return hipGetLastError();
}
hipError_t launch_plain_base(const Plan& P, const Parts& A, float* u1, int cus, hipStream_t s) {
  const int grid = grid_for(A.n, 256, cus * 16);
  if (P.mode == 2) plain_base<2><<<grid, 256, 0, s>>>(P, A, u1);
  else plain_base<1><<<grid, 256, 0, s>>>(P, A, u1);
  return hipGetLastError();
}
hipError_t launch_plain_validate(const Plan& P, const Parts& A, Counters* C, int cus, hipStream_t s) {
  const int grid = grid_for(A.n, 256, cus * 16);
  if (P.mode == 2) plain_validate<2><<<grid, 256, 0, s>>>(P, A, C);
  else plain_validate<1><<<grid, 256, 0, s>>>(P, A, C);
  return hipGetLastError();
}
hipError_t launch_plain_flags(const Plan& P, const Parts& A, int l, unsigned* flag, int cus, hipStream_t s) {
  plain_level_flags<<<grid_for(A.n, 256, cus * 16), 256, 0, s>>>(P, A, l, flag);
  return hipGetLastError();
}
hipError_t launch_plain_compact(const unsigned* flag, const unsigned long long* start, long long n,
                                long long* idx, int cus, hipStream_t s) {
  plain_compact<<<grid_for(n, 256, cus * 16), 256, 0, s>>>(flag, start, n, idx);
  return hipGetLastError();
}
hipError_t launch_plain_patch(const Plan& P, const Parts& A, int g, const long long* idx,
                              const unsigned long long* nidx, long long nmax, float* u11, int cus,
                              hipStream_t s) {
  if (nmax <= 0) return hipSuccess;
  const int grid = grid_for(nmax, 256, cus * 8);
  if (P.mode == 2) plain_patch<2><<<grid, 256, 0, s>>>(P, A, g, idx, nidx, u11);
  else plain_patch<1><<<grid, 256, 0, s>>>(P, A, g, idx, nidx, u11);
  return hipGetLastError();
}
hipError_t launch_plain_finish(const Plan& P, float* u1, float* u11, double* grid_mass, hipStream_t s) {
  plain_finish<<<P.ngrids, 256, 0, s>>>(P, u1, u11, grid_mass);
  return hipGetLastError();
}

}  // namespace asohf
