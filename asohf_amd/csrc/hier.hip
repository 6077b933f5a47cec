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
constexpr int kHbChunk = kHbThreads * kHbPer;    // particles per bin_base workgroup (< 2^16)
constexpr int kHbCap = 1024;                     // LDS key slots (power of two, < 2^16)
constexpr int kHbWList = 384;                    // list words per wave
constexpr int kHbProbe = 32;
constexpr int kEmThreads = 256;
constexpr int kEmPer = 2;                        // entries per thread and round
constexpr int kEmRound = kEmPer * kEmThreads;    // entries per round (< 2^16: a list word's position)
constexpr int kEmCap = 512;                      // LDS key slots of an item (< 2^16)
constexpr int kEmWList = 384;                    // list words per wave and round
constexpr unsigned long long kEmSkip = ~0ull;    // list word of an emission written directly
static_assert(kEmRound <= 65536 && kEmCap <= 65536, "a list word holds (entry, slot, position)");
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

// c / e for 0 <= c <= 26 and e in {1, 2, 3} (the tiles of a 3-cell stencil
// per axis), by a multiply and a shift: c * 86 / 256 lies within 0.07 below
// c / 3 there, and c / 3 is never within 0.07 above an integer
__device__ __forceinline__ int div_small(int c, int e) {
  return (c * (e == 1 ? 256 : (e == 2 ? 128 : 86))) >> 8;
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

// The position of one emission per lane that wants one (key = destination
// tile key) in its key's range of the workgroup's round: the lane finds the
// key's table slot and adds 1 to its count, whose old value is the position.
// Lanes of a wave with the same key meet at one LDS address, which the LDS
// serialises at less cost than grouping the lanes first (ballots and lane
// reads per group, measured slower with 1 to 8 groups).  h = -1 when the
// table is full (the emission is then not counted).
template <int CAP>
__device__ __forceinline__ void count_round(int* keys, unsigned* vals, bool want, int key, int& h, unsigned& at) {
  h = -1;
  at = 0;
  if (want) {
    h = slot_insert<CAP>(keys, key);
    if (h >= 0) at = atomicAdd(&vals[h], 1u);
  }
}

// exclusive prefix over the wave's lanes of v (0 <= v < 2^BITS) and the total
template <int BITS>
__device__ __forceinline__ int wave_prefix(int v, int& total) {
  const unsigned long long below = (1ull << __lane_id()) - 1ull;
  int off = 0, tot = 0;
#pragma unroll
  for (int b = 0; b < BITS; ++b) {
    const unsigned long long m = __ballot((v >> b) & 1);
    off += __popcll(m & below) << b;
    tot += __popcll(m) << b;
  }
  total = tot;
  return off;
}
__device__ __forceinline__ int wave_prefix4(int v, int& total) { return wave_prefix<4>(v, total); }
__device__ __forceinline__ int wave_prefix5(int v, int& total) { return wave_prefix<5>(v, total); }

// a reserved range in one word: its first slot (below 2^40; the host rejects
// larger entry buffers) and, above it, how many of its slots lie within the
// key's capacity (all of them unless the reservation overflowed)
constexpr int kSlotBits = 40;
constexpr unsigned long long kSlotMask = (1ull << kSlotBits) - 1ull;

// one reservation per tile key of the round: base[k] = the key's first slot
// and its room
template <int CAP, int NTHR>
__device__ __forceinline__ void reserve_round(const int* keys, unsigned* vals, unsigned long long* base,
                                              unsigned long long* cursor, const unsigned long long* start,
                                              Counters* C) {
  for (int k = threadIdx.x; k < CAP; k += NTHR) {
    const int key = keys[k];
    const unsigned v = vals[k];
    if (key >= 0 && v) {
      const unsigned long long b0 = atomicAdd(&cursor[key], (unsigned long long)v);
      const unsigned long long lim = start[key + 1];
      const unsigned long long room = b0 >= lim ? 0ull : (lim - b0 < v ? lim - b0 : (unsigned long long)v);
      base[k] = (b0 & kSlotMask) | (room << kSlotBits);   // room 0 when b0 is past the capacity
      if (b0 + v > lim) set_overflow(C);
    }
    vals[k] = 0u;
  }
}

// the slot of position `at` of a reserved range, or -1 past the key's capacity
__device__ __forceinline__ long long reserved_slot(unsigned long long b, unsigned at) {
  return at < (unsigned)(b >> kSlotBits) ? (long long)((b & kSlotMask) + at) : -1ll;
}

// Base-grid binning, persistent: each workgroup takes chunks of kHbChunk
// particles (kHbPer per thread) in turn and loads its next chunk while it
// bins this one.  Each valid particle
// (finite, mass >= 0, within the TSC range of the base grid) emits one entry
// per base tile its wrapped stencil touches.  Every wave lists its emissions
// (tile key, particle) in its own LDS list, then counts the list 64 at a time
// (a position in the key's range of the workgroup, written back over the
// list word); one global add per key reserves the workgroup's ranges and
// every wave writes its listed entries, taking each particle from the lane
// that holds it.
// Per-chunk partials: base members and the largest mass.
__global__ void __launch_bounds__(kHbThreads) bin_base_tsc(Plan P, Parts A, unsigned long long* cursor,
                                                           const unsigned long long* start, int4* entries,
                                                           Counters* C, BlockTables BT) {
  __shared__ int keys[kHbCap];
  __shared__ unsigned vals[kHbCap];
  __shared__ unsigned long long base[kHbCap];
  __shared__ unsigned long long wlist_all[(kHbThreads / 64) * kHbWList];
  __shared__ unsigned long long s_valid[kHbThreads / 64];
  __shared__ float s_mmax[kHbThreads / 64];
  const int tid = threadIdx.x, lane = (int)__lane_id();
  unsigned long long* wl = wlist_all + (tid >> 6) * kHbWList;
  const GridDesc& B = P.grids[0];
  const int bn[3] = {B.n[0], B.n[1], B.n[2]};
  const float bd[3] = {B.d[0], B.d[1], B.d[2]};
  const float bid[3] = {B.inv_d[0], B.inv_d[1], B.inv_d[2]};
  const int bnt0 = B.nt[0], bnt1 = B.nt[1];
  auto particle = [&](long long i) {
    return make_int4(__float_as_int(A.x[i]), __float_as_int(A.y[i]), __float_as_int(A.z[i]), __float_as_int(A.m[i]));
  };
  const long long nchunks = (A.n + kHbChunk - 1) / kHbChunk;
  // the workgroup's next chunk of particles, loaded while this one is binned
  float fx[kHbPer], fy[kHbPer], fz[kHbPer], fm[kHbPer];
  auto fetch = [&](long long cc) {
#pragma unroll
    for (int r = 0; r < kHbPer; ++r) {
      const long long i = cc * kHbChunk + r * kHbThreads + tid;
      fx[r] = fy[r] = fz[r] = fm[r] = 0.f;
      if (cc < nchunks && i < A.n) { fx[r] = A.x[i]; fy[r] = A.y[i]; fz[r] = A.z[i]; fm[r] = A.m[i]; }
    }
  };
  fetch(blockIdx.x);
  for (long long chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const long long c0 = chunk * kHbChunk;
    float px[kHbPer], py[kHbPer], pz[kHbPer], pm[kHbPer];
#pragma unroll
    for (int r = 0; r < kHbPer; ++r) { px[r] = fx[r]; py[r] = fy[r]; pz[r] = fz[r]; pm[r] = fm[r]; }
    fetch(chunk + gridDim.x);
    for (int k = tid; k < kHbCap; k += kHbThreads) { keys[k] = -1; vals[k] = 0u; }
    __syncthreads();
    unsigned err = 0;
    unsigned long long valid = 0;
    float mmax = 0.f;
    int nw = 0;   // the wave's list length (uniform)
#pragma unroll
    for (int r = 0; r < kHbPer; ++r) {
      const int k = r * kHbThreads + tid;
      bool ok = c0 + k < A.n;
      const float x = px[r], y = py[r], z = pz[r], m = pm[r];
      if (ok && !(isfinite(x) && isfinite(y) && isfinite(z) && isfinite(m))) { err |= kErrNonFinite; ok = false; }
      if (ok && m < 0.f) { err |= kErrNegMass; ok = false; }
      if (ok && !(tsc_base_in_range(x, P.xl, bid[0], bn[0]) && tsc_base_in_range(y, P.xl, bid[1], bn[1]) &&
                  tsc_base_in_range(z, P.xl, bid[2], bn[2]))) {
        err |= kErrWrap;
        ok = false;
      }
      int tx[3] = {0, 0, 0}, ty[3] = {0, 0, 0}, tz[3] = {0, 0, 0};
      int nt = 0, nx_ = 1, ny_ = 1;
      int hb[3] = {0, 0, 0};   // the home tile
      if (ok) {
        ++valid;
        mmax = fmaxf(mmax, m);
        nx_ = base_axis_tiles(tsc_base_index(x, P.xl, bd[0], bid[0], bn[0]), bn[0], P.tshift[0], tx);
        ny_ = base_axis_tiles(tsc_base_index(y, P.xl, bd[1], bid[1], bn[1]), bn[1], P.tshift[1], ty);
        const int nz_ = base_axis_tiles(tsc_base_index(z, P.xl, bd[2], bid[2], bn[2]), bn[2], P.tshift[2], tz);
        nt = nx_ * ny_ * nz_;
        hb[0] = (home_cell(x, P.xl, bid[0], true, bn[0]) - 1) >> P.tshift[0];
        hb[1] = (home_cell(y, P.xl, bid[1], true, bn[1]) - 1) >> P.tshift[1];
        hb[2] = (home_cell(z, P.xl, bid[2], true, bn[2]) - 1) >> P.tshift[2];
      }
      int tot;
      const int off = nw + wave_prefix5(nt, tot);
      for (int c = 0; c < nt; ++c) {
        const int q = div_small(c, nx_), qz = div_small(q, ny_);
        const int a = c - q * nx_, b = q - qz * ny_;
        const int ta = a == 0 ? tx[0] : (a == 1 ? tx[1] : tx[2]);
        const int tb = b == 0 ? ty[0] : (b == 1 ? ty[1] : ty[2]);
        const int tcz = qz == 0 ? tz[0] : (qz == 1 ? tz[1] : tz[2]);
        const int key = 2 * ((tcz * bnt1 + tb) * bnt0 + ta) + ((ta == hb[0] && tb == hb[1] && tcz == hb[2]) ? 0 : 1);
        if (off + c < kHbWList) wl[off + c] = (unsigned long long)(unsigned)key | ((unsigned long long)k << 32);
        else emit_direct(cursor, start, entries, C, key, particle(c0 + k));
      }
      nw = min(nw + tot, kHbWList);
    }
    // count the list: a position in the key's range of the workgroup
    for (int b0 = 0; b0 < nw; b0 += 64) {
      const int idx = b0 + lane;
      const bool want = idx < nw;
      const unsigned long long w = want ? wl[idx] : 0ull;
      int h;
      unsigned at;
      count_round<kHbCap>(keys, vals, want, (int)(unsigned)w, h, at);
      if (want) {
        if (h >= 0) {
          wl[idx] = (w & 0xffffffff00000000ull) | ((unsigned long long)h << 16) | at;
        } else {
          wl[idx] = kEmSkip;
          emit_direct(cursor, start, entries, C, (int)(unsigned)w, particle(c0 + (long long)(w >> 32)));
        }
      }
    }
    __syncthreads();
    reserve_round<kHbCap, kHbThreads>(keys, vals, base, cursor, start, C);
    __syncthreads();
    // the listed particle comes from the lane of this wave that loaded it
    // (k = r * kHbThreads + tid: lane k % 64, register slot k / kHbThreads);
    // every lane of the wave takes part in the shuffles
    for (int b0 = 0; b0 < nw; b0 += 64) {
      const int idx = b0 + lane;
      const unsigned long long w = idx < nw ? wl[idx] : kEmSkip;
      const int k = (int)(unsigned)(w >> 32);
      const int src = k & 63, slot = (int)((unsigned)k / (unsigned)kHbThreads);
      int4 e = make_int4(0, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < kHbPer; ++r) {
        const int4 u = make_int4(__shfl(__float_as_int(px[r]), src), __shfl(__float_as_int(py[r]), src),
                                 __shfl(__float_as_int(pz[r]), src), __shfl(__float_as_int(pm[r]), src));
        if (slot == r) e = u;
      }
      if (w == kEmSkip) continue;
      const long long pos = reserved_slot(base[(w >> 16) & 0xffffu], (unsigned)(w & 0xffffu));
      if (pos >= 0) entries[pos] = e;
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
      BT.pairs[chunk] = p;
      BT.mmax[chunk] = __float_as_uint(mm);
    }
    __syncthreads();   // the chunk's writes read base; its partials s_valid
  }
}

// The emission pass of one level: items [item_lo, item_hi) are the work items
// (chunks of one tile's entry range) of the level's tiles.  A workgroup takes
// `batch` consecutive items at a time from the level's queue counter and
// works through those whose tile has targets (the others' member counts stay
// 0).  Per item the target records are staged in LDS (kEmTargets at a time);
// the item's entries are taken kEmPer per thread per round (one per round
// where overlapping patches make the lists long).  In a
// round every wave tests its entries against the targets and lists its
// member emissions (destination tile key, entry) in its own LDS list, then
// counts the list 64 emissions at a time (a position in the key's range of
// the round for each, written back over the list word); one global add per
// key reserves the round's ranges and every wave writes its listed entries,
// reading each again from its source tile.
__global__ void __launch_bounds__(kEmThreads, 5) emit_tsc(Plan P, const Item* items, long long item_lo,
                                                       long long item_hi, int level, int batch,
                                                       unsigned long long* cursor,
                                                       const unsigned long long* start, int4* entries,
                                                       unsigned* item_pairs, Counters* C) {
  __shared__ int keys[kEmCap];
  __shared__ unsigned vals[kEmCap];
  __shared__ unsigned long long base[kEmCap];
  __shared__ unsigned long long wlist_all[(kEmThreads / 64) * kEmWList];
  __shared__ unsigned mlist_all[(kEmThreads / 64) * 128];
  __shared__ BoxRec recs[kEmTargets];
  __shared__ long long s_i0;
  __shared__ unsigned s_red[kEmThreads / 64];
  __shared__ int s_nwmax[2];   // the waves' longest list of a round, by round parity
  const int tid = threadIdx.x, lane = (int)__lane_id();
  unsigned long long* wl = wlist_all + (tid >> 6) * kEmWList;
  // wave-steps per round: kEmPer, or 1 after a round whose lists would have
  // overflowed with kEmPer (members of overlapping patches emit several
  // entries each); kept from one item to the next
  int steps = kEmPer;
  unsigned* ml = mlist_all + (tid >> 6) * 128;
  unsigned* queue = &C->equeue[level * kLine];
  for (;;) {
    if (tid == 0) s_i0 = item_lo + (long long)atomicAdd(queue, (unsigned)batch);
    __syncthreads();
    const long long i0 = s_i0;
    __syncthreads();   // every wave has read s_i0 before thread 0 takes the next batch
    if (i0 >= item_hi) break;
    const int nb = (int)min((long long)batch, item_hi - i0);
    for (int bi = 0; bi < nb; ++bi) {
      const long long ii = i0 + bi;
      const Item it = items[ii];
      const int t = it.tile;
      const unsigned q0 = P.tgt_off[t], q1 = P.tgt_off[t + 1];
      if (q0 == q1) continue;
      // the item's part of the tile's home entries: every member of this
      // grid has exactly one, so each member is visited once
      const long long e0 = it.begin;
      const long long e1 = min(min(it.end, (long long)start[2 * t + 1]), (long long)cursor[2 * t]);
      if (e1 <= e0) continue;
      for (int k = tid; k < kEmCap; k += kEmThreads) { keys[k] = -1; vals[k] = 0u; }
      if (tid < 2) s_nwmax[tid] = 0;
      int par = 0;
      unsigned pairs = 0;
      for (unsigned q = q0; q < q1; q += kEmTargets) {
        const int nq = (int)min((unsigned)kEmTargets, q1 - q);
        for (int k = tid; k < nq * 8; k += kEmThreads)
          reinterpret_cast<int4*>(recs)[k] = reinterpret_cast<const int4*>(P.grec + P.tgt[q + (k >> 3)])[k & 7];
        __syncthreads();
        for (long long r0 = e0, rn = 0; r0 < e1; r0 += rn) {
          const int nsteps = steps;
          rn = (long long)nsteps * kEmThreads;
          // ---- list the wave's member emissions
          int nw = 0;   // the wave's list length (uniform)
#pragma unroll 1
          for (int i = 0; i < nsteps; ++i) {
            const int k = i * kEmThreads + tid;
            const bool ok = r0 + k < e1;
            const int4 v = ok ? entries[r0 + k] : make_int4(0, 0, 0, 0);
            const float p[3] = {__int_as_float(v.x), __int_as_float(v.y), __int_as_float(v.z)};
            const bool home = ok;   // the range holds the tile's home entries only
            // the item's targets, one per pass: the box tests, with the
            // members listed (lane, target) in the wave's member list; every
            // 64 listed members, and at the end, the wave works through them
            // one per lane
            int nm = 0;   // the member list's length (uniform)
            auto members = [&](int n) {
              const int src = (int)(ml[lane < n ? lane : 0] & 63u);
              const int j = (int)(ml[lane < n ? lane : 0] >> 8);
              const int4 u = make_int4(__shfl(v.x, src), __shfl(v.y, src), __shfl(v.z, src), __shfl(v.w, src));
              const float q[3] = {__int_as_float(u.x), __int_as_float(u.y), __int_as_float(u.z)};
              const BoxRec& R = recs[j];
              bool mem = lane < n;
              int lo[3] = {0, 0, 0}, ext[3] = {1, 1, 1};
              if (mem) {
                ++pairs;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                  const int ci = (int)index_quotient(q[a] - R.org[a], R.d[a], R.inv_d[a]);
                  const int c0 = max(ci - 1, 1), c1 = min(ci + 1, R.n[a]);
                  // a member whose quotient rounds onto the far face of the
                  // box (ci = N + 2) has no stencil cell in the patch
                  mem = mem && c0 <= c1;
                  lo[a] = (c0 - 1) >> P.tshift[a];
                  ext[a] = ((c1 - 1) >> P.tshift[a]) - lo[a] + 1;
                }
              }
              const int nt = mem ? ext[0] * ext[1] * ext[2] : 0;
              int hp[3] = {0, 0, 0};   // the member's home tile in the patch
#pragma unroll
              for (int a = 0; a < 3; ++a) hp[a] = (home_cell(q[a], R.org[a], R.inv_d[a], false, R.n[a]) - 1) >> P.tshift[a];
              const int ks = i * kEmThreads + (tid & ~63) + src;
              int tot;
              const int off = nw + wave_prefix4(nt, tot);
              for (int c = 0; c < nt; ++c) {
                const int r = div_small(c, ext[0]), rz = div_small(r, ext[1]);
                const int cx = lo[0] + c - r * ext[0], cy = lo[1] + r - rz * ext[1], cz = lo[2] + rz;
                const int key = 2 * (R.tile0 + (cz * R.nt1 + cy) * R.nt0 + cx) +
                                ((cx == hp[0] && cy == hp[1] && cz == hp[2]) ? 0 : 1);
                if (off + c < kEmWList) wl[off + c] = (unsigned long long)(unsigned)key | ((unsigned long long)ks << 32);
                else emit_direct(cursor, start, entries, C, key, u);
              }
              nw = min(nw + tot, kEmWList);
            };
            for (int j = 0; j < nq; ++j) {
              const bool mem = home && in_box(recs[j], p[0], p[1], p[2]);
              const unsigned long long mb = __ballot(mem);
              if (mem) ml[nm + __popcll(mb & ((1ull << lane) - 1ull))] = (unsigned)lane | ((unsigned)j << 8);
              nm += __popcll(mb);
              if (nm >= 64) {
                members(64);
                nm -= 64;
                if (lane < nm) ml[lane] = ml[64 + lane];
              }
            }
            if (nm > 0) members(nm);
          }
          // ---- count the list: a position in the key's range of the round
          for (int b0 = 0; b0 < nw; b0 += 64) {
            const int idx = b0 + lane;
            const bool want = idx < nw;
            const unsigned long long w = want ? wl[idx] : 0ull;
            int h;
            unsigned at;
            count_round<kEmCap>(keys, vals, want, (int)(unsigned)w, h, at);
            if (want) {
              if (h >= 0) {
                wl[idx] = (w & 0xffffffff00000000ull) | ((unsigned long long)h << 16) | at;
              } else {
                wl[idx] = kEmSkip;
                emit_direct(cursor, start, entries, C, (int)(unsigned)w, entries[r0 + (long long)(w >> 32)]);
              }
            }
          }
          if (lane == 0) atomicMax(&s_nwmax[par], nw);
          if (tid == 0) s_nwmax[par ^ 1] = 0;   // the next round's, before this barrier
          __syncthreads();
          {
            const int mx = s_nwmax[par];
            steps = (mx * kEmPer > nsteps * (kEmWList * 3 / 4)) ? 1 : kEmPer;
            par ^= 1;
          }
          reserve_round<kEmCap, kEmThreads>(keys, vals, base, cursor, start, C);
          __syncthreads();
          // ---- write the listed entries
          for (int idx = lane; idx < nw; idx += 64) {
            const unsigned long long w = wl[idx];
            if (w == kEmSkip) continue;
            const long long pos = reserved_slot(base[(w >> 16) & 0xffffu], (unsigned)(w & 0xffffu));
            if (pos >= 0) entries[pos] = entries[r0 + (long long)(w >> 32)];
          }
        }
      }
      // members found by this item
      const unsigned sw = wave_sum(pairs);
      if ((tid & 63) == 0) s_red[tid >> 6] = sw;
      __syncthreads();
      if (tid == 0) {
        unsigned tot = 0;
        for (int w = 0; w < kEmThreads / 64; ++w) tot += s_red[w];
        item_pairs[ii] = tot;
      }
      __syncthreads();
    }
  }
}

// statistics of the call: members = base members + the members every emission
// item found; entries = the filled part of every tile's range.  Every
// workgroup adds its grid-stride slice to the counters (cleared at the start
// of the call; integer sums, so the order does not matter).
__global__ void __launch_bounds__(256) hier_stats(const unsigned* item_pairs, long long n_items,
                                                 const unsigned long long* start,
                                                 const unsigned long long* cursor, long long nkeys, Counters* C) {
  __shared__ unsigned long long sp[4], se[4];
  unsigned long long p = 0, e = 0;
  const long long i0 = blockIdx.x * (long long)blockDim.x + threadIdx.x, di = (long long)gridDim.x * blockDim.x;
  for (long long i = i0; i < n_items; i += di) p += item_pairs[i];
  for (long long k = i0; k < nkeys; k += di) e += cursor[k] - start[k];
  p = wave_sum(p);
  e = wave_sum(e);
  if ((threadIdx.x & 63) == 0) { sp[threadIdx.x >> 6] = p; se[threadIdx.x >> 6] = e; }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tp = 0, te = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) { tp += sp[w]; te += se[w]; }
    if (tp) atomicAdd(&C->n_pairs, tp);
    if (te) atomicAdd(&C->n_entries, te);
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
                               const BlockTables& BT, int cus, hipStream_t s) {
  const long long nb = hier_bin_blocks(A.n);
  if (nb == 0) return hipSuccess;
  static int per_cu = 0;
  if (per_cu <= 0) {
    int k = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&k, bin_base_tsc, kHbThreads, 0) != hipSuccess || k <= 0) k = 4;
    per_cu = k;
  }
  // the workgroups that fit the chip at once, each taking chunks in turn
  const long long grid = std::min<long long>(nb, (long long)cus * per_cu);
  bin_base_tsc<<<(unsigned)grid, kHbThreads, 0, s>>>(P, A, cursor, start, entries, C, BT);
  return hipGetLastError();
}

hipError_t launch_emit_tsc(const Plan& P, const Item* items, long long item_lo, long long item_hi, int level,
                           unsigned long long* cursor, const unsigned long long* start, int4* entries,
                           unsigned* item_pairs, Counters* C, int cus, hipStream_t s) {
  if (item_hi <= item_lo) return hipSuccess;
  if (level < 0 || level >= kNLevels) return hipErrorInvalidValue;
  // the workgroups that fit the chip at once; the queue balances the items,
  // taken up to 8 at a time while there are many per workgroup
  static int per_cu = 0;
  if (per_cu <= 0) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, emit_tsc, kEmThreads, 0) != hipSuccess || nb <= 0) nb = 4;
    per_cu = nb;
  }
  const long long n = item_hi - item_lo;
  const int grid = (int)std::min<long long>(n, (long long)cus * per_cu);
  const int batch = (int)std::max<long long>(1, std::min<long long>(8, n / ((long long)grid * 16)));
  emit_tsc<<<grid, kEmThreads, 0, s>>>(P, items, item_lo, item_hi, level, batch, cursor, start, entries, item_pairs,
                                       C);
  return hipGetLastError();
}

hipError_t launch_hier_stats(const unsigned* item_pairs, long long n_items, const unsigned long long* start,
                             const unsigned long long* cursor, long long nkeys, Counters* C, int cus,
                             hipStream_t s) {
  const long long want = (std::max(n_items, nkeys) + 255) / 256;
  if (want <= 0) return hipSuccess;
  hier_stats<<<(int)std::min<long long>(want, (long long)std::max(cus, 1) * 4), 256, 0, s>>>(item_pairs, n_items, start,
                                                                                          cursor, nkeys, C);
  return hipGetLastError();
}

}  // namespace asohf
