// internal.h -- structures shared by the host planner and the gfx950 kernels
// of libasohf_hip.  Everything here is plain data: kernels receive these by
// value or through device pointers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace asohf {

constexpr int kNEsp = 9;       // asohf_parameters.dat: N_ESP
constexpr int kNLevels = 9;    // asohf_parameters.dat: NLEVELS
constexpr uint64_t kHashEmpty = ~0ull;

// One grid of the build: the base grid (g = 0) or patch p (g = 1 + p).
// Box constants are FP32 values computed on the host with the reference's own
// expressions (grids.f:871-886 for the table mode, grids.f:1294-1302 for TSC),
// so the device only compares and indexes.
struct GridDesc {
  int n[3];          // cells per axis
  int level;         // IR (0 = base)
  int nt[3];         // tiles per axis
  int tile0;         // first tile id
  float lo[3];       // strict membership box: lo < x < hi
  float hi[3];
  float org[3];      // table mode: XLFIX (patch) / XL (base); TSC mode: XL of the BUF box / XL (base)
  float pr[3];       // PATCHRX/RY/RZ (TSC cell centres MINIRADX)
  float d[3];        // cell size DXPA/DYPA/DZPA (DX/DY/DZ on the base grid)
  float inv_d[3];    // 1/d in FP32 (weights only; indices use the IEEE division)
  float den;         // DENBAS of the level, FP32 as the reference forms it
  int ld1;           // output row stride (elements)
  int ld2;           // output plane stride (elements)
  long long off;     // output offset of cell (1,1,1) in u1 (g = 0) or u11
  double inv_dd[3];  // 1/d in FP64 (TSC weights of the batched deposit)
};

// Slot of the per-level spatial hash: bucket key and its candidate range, in
// one 16-byte load.
struct HashSlot {
  uint64_t key;              // kHashEmpty for a free slot
  int start;                 // candidates hrec[start, start + count)
  int count;
};

// A candidate patch of a bucket: its membership box and the grid fields a
// member particle needs, inline, so neither the box test nor the member's
// indices, tiles and output cells wait on a load of the grid descriptor
// (128 bytes, eight 16-byte loads issued together).
struct BoxRec {
  float lo[3];
  float hi[3];
  int g;                     // grid id (1 + patch)
  int level;                 // IR
  float org[3];              // GridDesc::org, d, inv_d, n, tile0, nt[0], nt[1], pr, den, ld1, ld2,
  float d[3];                // off of grid g
  float inv_d[3];
  int n[3];
  int tile0;
  int nt0, nt1;
  float den;
  float pr[3];
  int ld1;
  int ld2;
  int pad;
  long long off;
};
static_assert(sizeof(BoxRec) == 128, "eight 16-byte loads");

// Per-level spatial hash of patch membership boxes.
struct LevelHash {
  float x0, inv_bs;  // bucket b = clamp(floor((x - x0) * inv_bs), 0, nb-1)
  int nb;
  int pad;
};

// Device view of a planned geometry.
struct Plan {
  const GridDesc* grids;     // [1 + P]
  const int* tile_grid;      // [ntiles]
  const HashSlot* hslots;    // [hcap] open addressing, at most half full
  const BoxRec* hrec;        // candidates, bucket by bucket
  const float* radx;         // RADX(0:NX+1) etc.
  const float* rady;
  const float* radz;
  const float* ktab;         // kernel table [(nl+1) * kw], row ir at ir*kw + kmax
  int nl;
  int ngrids;
  int ntiles;
  int kmax;                  // 2^nl
  int kw;                    // 2*kmax + 1
  unsigned hmask;
  int tile[3];               // tile shape (powers of two)
  int tshift[3];             // log2 of the tile shape
  int mode;                  // 1 table, 2 TSC
  LevelHash lh[kNLevels + 1];
  float xl;                  // -LADO0/2
  float fdm;
  int nested;                // 1: every level-l box lies inside a level-(l-1) box (FP32 check)
  int rad_malla;             // 1: RAD tables are exactly MALLA's recurrence (rad1 + (i-1)*radc)
  int merge_narrow;          // 1: narrow deposits may merge lanes (every non-centre weight product
                             // <= 0.45, so eight terms of a mass below 2^30 sum below 2^32)
  int home_split;            // TSC count pass of the hierarchy binning: a member's entry in its
                             // home tile t counts under key 2t, its other entries under 2t + 1
  float rad1[3];
  float radc[3];
  // TSC binning along the hierarchy: the patches whose members are found
  // from tile t's home entries are tgt[tgt_off[t] .. tgt_off[t+1]) (grid
  // ids); grec[g] is grid g's box and fields in candidate-record form
  const unsigned* tgt_off;   // [ntiles + 1]
  const int* tgt;
  const BoxRec* grec;        // [1 + P]
};

// Particle inputs of one call.
struct Parts {
  const float* x;
  const float* y;
  const float* z;
  const float* m;
  long long n;               // particles visited (table mode: SUM(NPART_ESP))
  long long sp[kNEsp + 1];   // species prefix sums (table mode)
};

// Device counters of one call (one small buffer, cleared by a fill kernel).
// Counters that many workgroups update sit on their own 128-byte lines.
constexpr int kPoolSubs = 64;       // sub-pools of the saved block tables
constexpr int kQueues = 8;          // deposit work queues (one per XCD)
constexpr int kLine = 32;           // 128 bytes in 4-byte units
struct Counters {
  unsigned long long n_pairs;     // reduced from per-block partials
  unsigned long long n_entries;   // totals of the planning scans, copied by the gate kernel
  unsigned long long n_items;
  unsigned long long n_split;
  unsigned int err;               // bit 0: non-finite position/mass, bit 1: base index outside
                                  // the single-wrap range, bit 2: negative mass
  unsigned int mmax_bits;         // max particle mass, as FP32 bits (masses >= 0)
  unsigned int overflow;          // 1: a workspace was smaller than this call needs; the
                                  // scatter, planning and deposit kernels then do nothing
  unsigned int n_items_eff;       // items the deposit serves (0 after an overflow)
  unsigned int pad[kLine - 12];
  unsigned long long pool_fill[kPoolSubs * kLine / 2];   // sub-pool k at [k * kLine / 2]
  unsigned long long epool_fill[kPoolSubs * kLine / 2];  // emission lists, same layout
  unsigned int queue[kQueues * kLine];                   // queue q at [q * kLine]
  unsigned int equeue[kNLevels * kLine];                 // emission pass of level l at [l * kLine]
};

static_assert(sizeof(unsigned long long) * 4 + sizeof(unsigned) * (kLine - 8) == kLine * 4,
              "the scalar counters fill one 128-byte line");

// Work item of the persistent deposit: a chunk of one tile's entry list.
struct Item {
  int tile;
  int flags;                 // bit 0: wide (table-mode N >= 2) entries; bit 1: every wide entry
                             // of the tile; bits 8+: the item's rank within its tile
  long long begin;           // entry range [begin, end)
  long long end;
};

// Per-workgroup (key, count) tables saved by the count pass for the scatter
// pass, in a compact pool; off = -1 when the table did not fit.
struct BlockTables {
  int2* pool;                // (key, count) pairs, kPoolSubs equal sub-pools
  long long cap;             // pairs the pool holds
  long long* off;            // [blocks] start of each block's pairs, or -1
  int* nkeys;                // [blocks]
  unsigned long long* pairs; // [blocks] (particle, grid) members counted by the block
  unsigned* mmax;            // [blocks] max mass of the block, FP32 bits
  // every emission of the count pass, (tile key << 9) | particle within the
  // block, so the scatter pass need not enumerate again
  unsigned* elist;           // kPoolSubs equal sub-pools
  long long ecap;
  long long* eoff;           // [blocks] start of the block's list, or -1 (enumerate again)
  int* ecount;               // [blocks]
  int use_elist;             // 0 for table mode and for keys of 2^23 or more
  int pad;
};

// Arguments of the persistent deposit.
struct DepArgs {
  const int4* entries;       // binned entries, grouped by tile
  const Item* items;         // work queue
  const int* nitems_tile;    // items per tile
  const int* split_slot;     // first scratch slot of a split tile, -1 otherwise
  const int* split_tiles;    // the split tiles, in tile order
  const unsigned long long* n_split_tiles;   // their count (device)
  void* scratch;             // int64 (deterministic) or FP32 (atomic) tile accumulators
  Counters* C;               // n_items_eff (the queue length) and mmax_bits (the fixed-point
                             // scale) are read here, so no host round trip precedes the deposit
  float* u1;
  float* u11;
  long long* tile_mass_i;    // deterministic: exact fixed-point mass per tile
  double* tile_mass_d;       // atomic: FP64 mass per tile
  int n_queues;              // queues in use (<= kQueues, <= grid)
  // TSC binning along the hierarchy: tile t's range [start[2t], start[2t+2])
  // holds its members' home entries from start[2t] to cursor[2t] and their
  // other entries from start[2t+1] to cursor[2t+1]; items cover capacity
  // ranges (both NULL otherwise: every item's range is filled)
  const unsigned long long* cursor;
  const unsigned long long* start;
  int hist_cells;            // table mode: int64 cells of the wide classes' histogram buffer
                             // (dynamic LDS, set by the launcher; 0: every class walks directly)
  int pad2;
};

enum : int { kErrNonFinite = 1, kErrWrap = 2, kErrNegMass = 4, kErrIndexRange = 8 };

}  // namespace asohf
