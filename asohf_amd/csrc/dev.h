// dev.h -- device helpers of the density build (gfx950).
//
// Index expressions are the reference's FP32 expressions evaluated in the
// same operation order (the library is compiled with -ffp-contract=off and
// IEEE-correct FP32 division), so every cell and patch index matches the
// reference bit for bit:
//   table mode base:  IX = INT((XP-XL)/DX)+1, single periodic wrap  (grids.f:798-813)
//   table mode patch: strict box with the DX extension, IX = FLOOR((XP-XLFIX)/DXPA)+1,
//                     stencil half-width N = 2**max(IR-species,0)  (grids.f:871-935)
//   TSC base:         IX = INT((XP-XL)/DX)+1 clamped to [1,NX]      (grids.f:1180-1192)
//   TSC patch:        strict BUF=1 box, IX = INT((XP-XL)/DXPA)+1-1  (grids.f:1294-1323)
#pragma once
#include "internal.h"

namespace asohf {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// ---------------------------------------------------------------------------
// wave / block reductions
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ unsigned wave_or(unsigned v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
  return v;
}

// fixed-shape block sum (deterministic for a given blockDim)
template <class T>
__device__ T block_sum(T v, T* sh) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  v = wave_sum(v);
  if (lane == 0) sh[w] = v;
  __syncthreads();
  T r = 0;
  if (threadIdx.x == 0) {
    for (int k = 0; k < nw; ++k) r += sh[k];
    sh[0] = r;
  }
  __syncthreads();
  r = sh[0];
  __syncthreads();
  return r;
}

// single periodic wrap of grids.f:808-813 (the TSC wrap of grids.f:1221-1226
// is the same map on the clamped index range)
__device__ __forceinline__ int wrap1(int i, int n) {
  if (i > n) i -= n;
  if (i < 1) i += n;
  return i;
}

// TSC weights of grids.f:1194-1213 with centres c0 = RAD(I-1), c1 = RAD(I), c2 = RAD(I+1)
__device__ __forceinline__ void tsc_weights(float xp, float c0, float c1, float c2, float d, float w[3]) {
  float b, t;
  b = fabsf(xp - c0) / d; t = 1.5f - b; w[0] = 0.5f * (t * t);
  b = fabsf(xp - c1) / d; w[1] = 0.75f - b * b;
  b = fabsf(xp - c2) / d; t = 1.5f - b; w[2] = 0.5f * (t * t);
}
// The same weights with the reciprocal of the cell size: each b within 1.5
// ulp of the reference's quotient.  The three centres are the reference's
// own FP32 values (RADX(I-1..I+1), MINIRADX(I-1..I+1)): each is rounded on
// its own, and at large coordinates their spacing differs from the cell size
// by a noticeable fraction of a cell (ulp(250) is 0.4% of a 1/128 cell of a
// 500 Mpc/h, 1024-cell grid), which the weights must see as the reference
// does.
__device__ __forceinline__ void tsc_weights_c(float xp, float c0, float c1, float c2, float inv_d, float w[3]) {
  float b, t;
  b = fabsf(xp - c0) * inv_d; t = 1.5f - b; w[0] = 0.5f * (t * t);
  b = fabsf(xp - c1) * inv_d; w[1] = 0.75f - b * b;
  b = fabsf(xp - c2) * inv_d; t = 1.5f - b; w[2] = 0.5f * (t * t);
}
// The deposit's form: distances to the three centres in FP64 (exact
// differences of FP32 values), the weights rounded once to FP32.  The
// reference's FP32 evaluation cancels in 1.5 - b as b approaches 3/2, where
// a corner weight is tiny and its FP32 value keeps few correct digits; in
// FP64 every weight is within one FP32 rounding of the exact value on the
// reference's operands, so even a cell reached only by such a corner holds
// its FP64 value to FP32 accuracy.
__device__ __forceinline__ void tsc_weights_f64(float xp, float c0, float c1, float c2, double inv_d, float w[3]) {
  const double x = (double)xp;
  double b, t;
  b = fabs(x - (double)c0) * inv_d; t = 1.5 - b; w[0] = (float)(0.5 * (t * t));
  b = fabs(x - (double)c1) * inv_d; w[1] = (float)(0.75 - b * b);
  b = fabs(x - (double)c2) * inv_d; t = 1.5 - b; w[2] = (float)(0.5 * (t * t));
}
// and their derivatives with respect to xp
__device__ __forceinline__ void tsc_weights_dc(float xp, float c0, float c1, float c2, float inv_d, float w[3],
                                               float dw[3]) {
  tsc_weights_c(xp, c0, c1, c2, inv_d, w);
  float t;
  t = xp - c0; dw[0] = -(1.5f - fabsf(t) * inv_d) * (t > 0.f ? 1.f : (t < 0.f ? -1.f : 0.f)) * inv_d;
  t = xp - c1; dw[1] = -2.0f * t * (inv_d * inv_d);
  t = xp - c2; dw[2] = -(1.5f - fabsf(t) * inv_d) * (t > 0.f ? 1.f : (t < 0.f ? -1.f : 0.f)) * inv_d;
}
// The quotient a/d whose integer part is a cell index, exact where it
// matters: t = a*(1/d) is within |q|*1.2e-7 of q = a/d and RN(q) within
// |q|*6e-8, so when t is farther than |t|*2^-20 from every integer, RN(q) lies
// strictly between the same two integers and truncating or flooring either
// gives the same index; otherwise the IEEE division is taken.
// The division runs only in waves where some lane needs it (a uniform
// branch: computed for every lane and selected, it costs about a dozen
// instructions per call).
__device__ __forceinline__ float index_quotient(float a, float d, float inv_d) {
  const float t = a * inv_d;
  const bool near = !(fabsf(t - rintf(t)) > fabsf(t) * 9.5367431640625e-7f);
  if (__ballot(near)) return near ? a / d : t;
  return t;
}

// MINIRADX(II) = PATCHRX + (FLOAT(II)-1.5)*DXPA  (grids.f:1304-1312)
__device__ __forceinline__ float minirad(float pr, int ii, float d) { return pr + ((float)ii - 1.5f) * d; }

// The cell that decides a member's home tile: INT((x - origin) / d) + 1
// clamped (the base grid) or INT((x - origin) / d) clamped (a patch) with
// the quotient taken as a product with the reciprocal.  It is within one cell
// of the reference's index, so it lies in the member's stencil and its tile
// holds one of the member's entries; every entry of the member computes the
// same value, so exactly one of them is home.
__device__ __forceinline__ int home_cell(float x, float org, float inv_d, bool base, int n) {
  float t = (x - org) * inv_d;
  t = fminf(fmaxf(t, -1.0e9f), 1.0e9f);
  return clampi((int)t + (base ? 1 : 0), 1, n);
}

// INT((XP-XL)/DX)+1 clamped, on the base grid; t is guarded so the
// conversion is always defined.
__device__ __forceinline__ int tsc_base_index(float xp, float xl, float d, float inv_d, int n) {
  float t = index_quotient(xp - xl, d, inv_d);
  if (!(t < 1.0e9f)) t = 1.0e9f;
  if (!(t > -1.0e9f)) t = -1.0e9f;
  return clampi((int)t + 1, 1, n);
}

// Positions the TSC base deposit accepts.  The reference clamps the base
// index to [1,N] (grids.f:1184-1192) and evaluates the weights at that cell's
// centre (grids.f:1194-1213); more than half a cell past the first or last
// centre the middle weight 0.75 - b^2 shrinks and then turns negative, and
// the outer ones grow past 1/2, which the exact fixed-point deposit (terms
// rounded to unsigned 32-bit integers, four of them summed in 32 bits) does
// not represent.  Such positions lie outside the box [-LADO0/2, LADO0/2] by
// more than rounding; they are rejected (kErrWrap), as the table mode
// rejects positions outside its wrap range.  The slack of 2^-10 cells keeps
// every position of the box (whose quotient may round a few ulps past 0 or
// N) and bounds |offset| by 1/2 + 2^-10, where every non-centre weight
// product stays below 0.2822 (below 1/2 with the centres' own rounding).
constexpr float kTscBaseSlack = 9.765625e-4f;
__device__ __forceinline__ bool tsc_base_in_range(float xp, float xl, float inv_d, int n) {
  const float t = (xp - xl) * inv_d;
  return t >= -kTscBaseSlack && t <= (float)n + kTscBaseSlack;
}

__host__ __device__ __forceinline__ uint64_t bucket_key(int level, int bx, int by, int bz) {
  return ((uint64_t)level << 60) | ((uint64_t)bx << 40) | ((uint64_t)by << 20) | (uint64_t)bz;
}

// low 32 bits of the product of the low 24 bits of a and b (one full-rate
// v_mul_u32_u24 on the device)
__host__ __device__ __forceinline__ unsigned mul24(unsigned a, unsigned b) {
#ifdef __HIP_DEVICE_COMPILE__
  return __umul24(a, b);
#else
  return (a & 0xffffffu) * (b & 0xffffffu);
#endif
}

// slot hash of a bucket (coordinates < 2^14, level < 16), 32-bit arithmetic
__host__ __device__ __forceinline__ unsigned bucket_hash(int level, int bx, int by, int bz) {
  unsigned h = mul24((unsigned)bx, 0x9E3779u) ^ mul24((unsigned)by, 0x85EBCBu) ^
               mul24((unsigned)bz, 0xC2B2AFu) ^ mul24((unsigned)level + 1u, 0x27D4EBu);
  h ^= h >> 15;
  h = mul24(h, 0x2C1B3Du) ^ (h >> 17);
  h ^= h >> 13;
  return h;
}

// A hash lookup in two steps, so the first slot load of the next level can be
// in flight while the current level's candidates are tested.
struct Probe {
  uint64_t key;
  HashSlot h;
  unsigned s;
};
__device__ __forceinline__ Probe probe_start(const Plan& P, int l, float x, float y, float z) {
  const LevelHash lh = P.lh[l];
  const int bx = clampi((int)floorf(fminf(fmaxf((x - lh.x0) * lh.inv_bs, -1.f), 1.0e6f)), 0, lh.nb - 1);
  const int by = clampi((int)floorf(fminf(fmaxf((y - lh.x0) * lh.inv_bs, -1.f), 1.0e6f)), 0, lh.nb - 1);
  const int bz = clampi((int)floorf(fminf(fmaxf((z - lh.x0) * lh.inv_bs, -1.f), 1.0e6f)), 0, lh.nb - 1);
  Probe p;
  p.key = bucket_key(l, bx, by, bz);
  p.s = bucket_hash(l, bx, by, bz) & P.hmask;
  p.h = P.hslots[p.s];
  return p;
}
// candidates (start, count) in hrec; the table is at most half full, so an
// empty slot ends every probe
__device__ __forceinline__ int2 probe_finish(const Plan& P, Probe p) {
  for (;;) {
    if (p.h.key == p.key) return make_int2(p.h.start, p.h.count);
    if (p.h.key == kHashEmpty) return make_int2(0, 0);
    p.s = (p.s + 1) & P.hmask;
    p.h = P.hslots[p.s];
  }
}
// candidate patches of level l whose (inflated) membership box may contain (x,y,z)
__device__ __forceinline__ int2 bucket_lookup(const Plan& P, int l, float x, float y, float z) {
  return probe_finish(P, probe_start(P, l, x, y, z));
}

// strict membership test of the reference's box rule on a candidate record
__device__ __forceinline__ bool in_box(const BoxRec& r, float x, float y, float z) {
  return r.lo[0] < x && x < r.hi[0] && r.lo[1] < y && y < r.hi[1] && r.lo[2] < z && z < r.hi[2];
}

__device__ __forceinline__ int species_of(const Parts& A, long long i) {
  int s = 0;
#pragma unroll
  for (int k = 1; k < kNEsp; ++k) s += (i >= A.sp[k]) ? 1 : 0;
  return s;
}

// Table-mode entry: cell index triple (int16 each), kernel level, mass.
__device__ __forceinline__ int4 pack_table(int ix, int jy, int kz, int irk, float m) {
  int4 e;
  e.x = (ix & 0xffff) | (jy << 16);
  e.y = (kz & 0xffff) | (irk << 16);
  e.z = __float_as_int(m);
  e.w = 0;
  return e;
}
// The table-mode entries of a particle that enumerate_particle produces,
// rebuilt from the particle and the grid alone (same expressions): the base
// cell INT((XP-XL)/DX)+1 (range checked by the enumeration), and on a patch
// FLOOR((XP-XLFIX)/DXPA)+1 with the stencil level max(IR - species, 0).
__device__ __forceinline__ int4 table_base_entry(const Plan& P, float x, float y, float z, float m) {
  const GridDesc& B = P.grids[0];
  const int ix = (int)index_quotient(x - P.xl, B.d[0], B.inv_d[0]) + 1;
  const int jy = (int)index_quotient(y - P.xl, B.d[1], B.inv_d[1]) + 1;
  const int kz = (int)index_quotient(z - P.xl, B.d[2], B.inv_d[2]) + 1;
  return pack_table(ix, jy, kz, 0, m);
}
__device__ __forceinline__ int4 table_patch_entry(const GridDesc& G, float x, float y, float z, float m,
                                                  int species) {
  const int ix = (int)floorf(index_quotient(x - G.org[0], G.d[0], G.inv_d[0])) + 1;
  const int jy = (int)floorf(index_quotient(y - G.org[1], G.d[1], G.inv_d[1])) + 1;
  const int kz = (int)floorf(index_quotient(z - G.org[2], G.d[2], G.inv_d[2])) + 1;
  return pack_table(ix, jy, kz, max(G.level - species, 0), m);
}
__device__ __forceinline__ void unpack_table(int4 e, int& ix, int& jy, int& kz, int& irk, float& m) {
  ix = (int)(short)(e.x & 0xffff);
  jy = e.x >> 16;
  kz = (int)(short)(e.y & 0xffff);
  irk = e.y >> 16;
  m = __int_as_float(e.z);
}

// Tiles touched by the wrapped cells {c-1, c, c+1} along one base axis.
// Returns the number of distinct tile indices (1..3) in t[].
__device__ __forceinline__ int base_axis_tiles(int c, int n, int sh, int t[3]) {
  int a = (wrap1(c - 1, n) - 1) >> sh, b = (wrap1(c, n) - 1) >> sh, d = (wrap1(c + 1, n) - 1) >> sh;
  int k = 0;
  t[k++] = a;
  if (b != a) t[k++] = b;
  if (d != a && d != b) t[k++] = d;
  return k;
}

// Enumerates every (grid, tile) a particle deposits into, calling
//   emit(g, tile, wide, entry)
// once per touched tile, and member(g, ix, jy, kz) once per (particle, grid)
// member of the mode's box rule with the reference's cell index of that grid
// (before any wrap or clipping).  err receives validation flags.
template <int MODE, class Emit, class Member>
__device__ __forceinline__ void enumerate_particle(const Plan& P, long long i, float x, float y,
                                                   float z, float m, int species, Emit&& emit,
                                                   Member&& member, unsigned& err) {
  // the first level's hash slot is loaded while the base grid is handled
  Probe nxt;
  if (P.nl >= 1) nxt = probe_start(P, 1, x, y, z);
  // Grid fields are copied into registers before the emission loops: the
  // callbacks write memory (atomics, entries), after which the compiler would
  // otherwise reload every field from the descriptor.
  const GridDesc& Bref = P.grids[0];
  const int bn[3] = {Bref.n[0], Bref.n[1], Bref.n[2]};
  const float bd[3] = {Bref.d[0], Bref.d[1], Bref.d[2]};
  const float bid[3] = {Bref.inv_d[0], Bref.inv_d[1], Bref.inv_d[2]};
  const int btile0 = Bref.tile0, bnt0 = Bref.nt[0], bnt1 = Bref.nt[1];
  // ---------------- base grid ----------------
  int ix, jy, kz;
  if (MODE == 2) {
    if (!(tsc_base_in_range(x, P.xl, bid[0], bn[0]) && tsc_base_in_range(y, P.xl, bid[1], bn[1]) &&
          tsc_base_in_range(z, P.xl, bid[2], bn[2]))) {
      err |= kErrWrap;
      return;
    }
    ix = tsc_base_index(x, P.xl, bd[0], bid[0], bn[0]);
    jy = tsc_base_index(y, P.xl, bd[1], bid[1], bn[1]);
    kz = tsc_base_index(z, P.xl, bd[2], bid[2], bn[2]);
  } else {
    const float tx = index_quotient(x - P.xl, bd[0], bid[0]);
    const float ty = index_quotient(y - P.xl, bd[1], bid[1]);
    const float tz = index_quotient(z - P.xl, bd[2], bid[2]);
    // INT() is defined only on a bounded range; beyond it the reference would
    // index outside U1 anyway
    if (!(tx > -1.0e9f && tx < 1.0e9f && ty > -1.0e9f && ty < 1.0e9f && tz > -1.0e9f && tz < 1.0e9f)) {
      err |= kErrWrap;
      return;
    }
    ix = (int)tx + 1; jy = (int)ty + 1; kz = (int)tz + 1;
    // the single wrap must land in 1..N for all three offsets
    if (ix < 2 - bn[0] || ix > 2 * bn[0] - 1 || jy < 2 - bn[1] || jy > 2 * bn[1] - 1 ||
        kz < 2 - bn[2] || kz > 2 * bn[2] - 1) { err |= kErrWrap; return; }
  }
  member(0, ix, jy, kz);
  {
    int tx[3], ty[3], tz[3];
    const int nx_ = base_axis_tiles(ix, bn[0], P.tshift[0], tx);
    const int ny_ = base_axis_tiles(jy, bn[1], P.tshift[1], ty);
    const int nz_ = base_axis_tiles(kz, bn[2], P.tshift[2], tz);
    int4 e;
    if (MODE == 2) e = make_int4(__float_as_int(x), __float_as_int(y), __float_as_int(z), __float_as_int(m));
    else e = pack_table(ix, jy, kz, 0, m);
    // the hierarchy binning's split: the home tile's entry under 2t, the
    // others under 2t + 1
    int hb[3] = {-1, -1, -1};
    if (MODE == 2 && P.home_split) {
      hb[0] = (home_cell(x, P.xl, bid[0], true, bn[0]) - 1) >> P.tshift[0];
      hb[1] = (home_cell(y, P.xl, bid[1], true, bn[1]) - 1) >> P.tshift[1];
      hb[2] = (home_cell(z, P.xl, bid[2], true, bn[2]) - 1) >> P.tshift[2];
    }
    for (int c = 0; c < nz_; ++c)
      for (int b = 0; b < ny_; ++b)
        for (int a = 0; a < nx_; ++a) {
          const int away = (MODE == 2 && P.home_split && !(tx[a] == hb[0] && ty[b] == hb[1] && tz[c] == hb[2]));
          emit(0, btile0 + (tz[c] * bnt1 + ty[b]) * bnt0 + tx[a], away, e);
        }
  }
  // ---------------- patches, level by level ----------------
  // When every patch's membership box lies inside a box of the level above
  // (checked on the host in FP32, P.nested), a particle in no level-l box is
  // in no deeper box either and the descent stops.
  for (int l = 1; l <= P.nl; ++l) {
    const int2 se = probe_finish(P, nxt);
    if (l < P.nl) nxt = probe_start(P, l + 1, x, y, z);
    bool any = false;
    for (int q = 0; q < se.y; ++q) {
      const BoxRec r = P.hrec[se.x + q];
      if (!in_box(r, x, y, z)) continue;
      const int g = r.g;
      const float gorg[3] = {r.org[0], r.org[1], r.org[2]};
      const float gd[3] = {r.d[0], r.d[1], r.d[2]};
      const float gid[3] = {r.inv_d[0], r.inv_d[1], r.inv_d[2]};
      const int gn[3] = {r.n[0], r.n[1], r.n[2]};
      const int gtile0 = r.tile0, gnt0 = r.nt0, gnt1 = r.nt1;
      const int glevel = r.level;
      any = true;
      int lo[3], hi[3], ci[3];
      int irk = 0;
      if (MODE == 2) {
        ci[0] = (int)index_quotient(x - gorg[0], gd[0], gid[0]);
        ci[1] = (int)index_quotient(y - gorg[1], gd[1], gid[1]);
        ci[2] = (int)index_quotient(z - gorg[2], gd[2], gid[2]);
        member(g, ci[0], ci[1], ci[2]);
#pragma unroll
        for (int a = 0; a < 3; ++a) { lo[a] = max(ci[a] - 1, 1); hi[a] = min(ci[a] + 1, gn[a]); }
      } else {
        ci[0] = (int)floorf(index_quotient(x - gorg[0], gd[0], gid[0])) + 1;
        ci[1] = (int)floorf(index_quotient(y - gorg[1], gd[1], gid[1])) + 1;
        ci[2] = (int)floorf(index_quotient(z - gorg[2], gd[2], gid[2])) + 1;
        member(g, ci[0], ci[1], ci[2]);
        irk = max(glevel - species, 0);
        const int nk = 1 << irk;
#pragma unroll
        for (int a = 0; a < 3; ++a) { lo[a] = max(ci[a] - nk, 1); hi[a] = min(ci[a] + nk, gn[a]); }
      }
      if (hi[0] < lo[0] || hi[1] < lo[1] || hi[2] < lo[2]) continue;   // cloud misses the patch
      int4 e;
      if (MODE == 2) e = make_int4(__float_as_int(x), __float_as_int(y), __float_as_int(z), __float_as_int(m));
      else e = pack_table(ci[0], ci[1], ci[2], irk, m);
      const int wide = (MODE == 1 && irk > 0) ? 1 : 0;
      const int t0x = (lo[0] - 1) >> P.tshift[0], t1x = (hi[0] - 1) >> P.tshift[0];
      const int t0y = (lo[1] - 1) >> P.tshift[1], t1y = (hi[1] - 1) >> P.tshift[1];
      const int t0z = (lo[2] - 1) >> P.tshift[2], t1z = (hi[2] - 1) >> P.tshift[2];
      int hp[3] = {-1, -1, -1};
      if (MODE == 2 && P.home_split) {
        hp[0] = (home_cell(x, gorg[0], gid[0], false, gn[0]) - 1) >> P.tshift[0];
        hp[1] = (home_cell(y, gorg[1], gid[1], false, gn[1]) - 1) >> P.tshift[1];
        hp[2] = (home_cell(z, gorg[2], gid[2], false, gn[2]) - 1) >> P.tshift[2];
      }
      for (int c = t0z; c <= t1z; ++c)
        for (int b = t0y; b <= t1y; ++b)
          for (int a = t0x; a <= t1x; ++a) {
            const int away = (MODE == 2 && P.home_split && !(a == hp[0] && b == hp[1] && c == hp[2]));
            emit(g, gtile0 + (c * gnt1 + b) * gnt0 + a, wide | away, e);
          }
    }
    if (!any && P.nested) break;
  }
}

}  // namespace asohf
