// launch.h -- host launchers of the kernels in kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "internal.h"

namespace asohf {

hipError_t launch_fill_u32(unsigned* p, unsigned v, long long n, hipStream_t s);
long long bin_blocks(long long n);
// workgroups of the hierarchy path's base-grid binning (its per-block tables)
long long hier_bin_blocks(long long n);
hipError_t launch_count(const Plan& P, const Parts& A, unsigned* cnt, Counters* C, const BlockTables& BT,
                        int cus, hipStream_t s);
hipError_t launch_scatter(const Plan& P, const Parts& A, unsigned long long* cursor, int4* entries,
                          unsigned long long cap, const BlockTables& BT, const Counters* C, int cus,
                          hipStream_t s);
hipError_t launch_membership(const Plan& P, const Parts& A, int* out, long long cap,
                             unsigned long long* count, int cus, hipStream_t s);
long long scan_scratch_elems(long long n);
hipError_t launch_scan(const unsigned* in, long long n, unsigned long long* bsum, unsigned long long* start,
                       unsigned long long* cursor, hipStream_t s);
hipError_t launch_reduce_blocks(const BlockTables& BT, long long nblk, Counters* C, int cus, hipStream_t s);
hipError_t launch_plan_count(const unsigned long long* start, int ntiles, int chunk, int chunk_wide,
                             unsigned* nitems_tile, unsigned* nsplit_tile, unsigned* is_split, hipStream_t s);
hipError_t launch_plan_write(const unsigned long long* start, int ntiles, int chunk, int chunk_wide,
                             const unsigned long long* item_off, const unsigned long long* split_off,
                             const unsigned long long* split_rank, Item* items, int* split_slot,
                             int* split_tiles, long long* tile_mass, const Counters* C, hipStream_t s);
hipError_t launch_gate(const unsigned long long* tot_entries, const unsigned long long* tot_items,
                       const unsigned long long* tot_split, unsigned long long cap_entries,
                       unsigned long long cap_items, unsigned long long cap_split, Counters* C, hipStream_t s);
bool tile_shape_supported(const int t[3]);
hipError_t launch_deposit(const Plan& P, const DepArgs& D, int reduction, int grid, int block, hipStream_t s);
hipError_t launch_zero_split(const Plan& P, const DepArgs& D, int reduction, int cus, hipStream_t s);
hipError_t launch_split_reduce(const Plan& P, const DepArgs& D, int reduction, int cus, hipStream_t s);
hipError_t launch_grid_sums(const Plan& P, const DepArgs& D, int reduction, double* grid_mass, hipStream_t s);
// TSC binning along the hierarchy (hier.hip)
hipError_t launch_cursor_reset(const unsigned long long* start, unsigned long long* cursor, long long nkeys,
                               int cus, hipStream_t s);
hipError_t launch_count_to_capacity(unsigned* cnt, long long nkeys, int cus, hipStream_t s);
hipError_t launch_bin_base_tsc(const Plan& P, const Parts& A, unsigned long long* cursor,
                               const unsigned long long* start, int4* entries, Counters* C,
                               const BlockTables& BT, int cus, hipStream_t s);
hipError_t launch_emit_tsc(const Plan& P, const Item* items, long long item_lo, long long item_hi, int level,
                           unsigned long long* cursor, const unsigned long long* start, int4* entries,
                           unsigned* item_pairs, Counters* C, int cus, hipStream_t s);
hipError_t launch_hier_stats(const unsigned* item_pairs, long long n_items, const unsigned long long* start,
                             const unsigned long long* cursor, long long nkeys, Counters* C, int cus,
                             hipStream_t s);
hipError_t launch_adjoint(const Plan& P, const Parts& A, const float* g1, const float* g11, float* dm,
                          float* dx, float* dy, float* dz, const unsigned* order, int cus, hipStream_t s);
hipError_t launch_cell_keys(const Plan& P, const Parts& A, unsigned* key, unsigned* idx, Counters* C, int cus,
                            hipStream_t s);
hipError_t species_sort_bytes(unsigned long long n, size_t& bytes);
hipError_t launch_species(const int nmax[3], const float d[3], float bas, long long n, const float* x,
                          const float* y, const float* z, const float* m, int* species, long long* order,
                          unsigned* counts, float* dens, unsigned* k0, unsigned* k1, unsigned* k2, unsigned* k3,
                          void* temp, size_t temp_bytes, int cus, hipStream_t s);
// the species partition on the file-scope scratch array of order.hip
constexpr long long kNoAllocMaxParticles = 1ll << 20;
constexpr long long kNoAllocMaxCells = 1ll << 21;
constexpr size_t kNoAllocBytes = 64ull << 20;
hipError_t noalloc_scratch_init(int device, void** out);
hipError_t launch_species_noalloc(void* base, const int nmax[3], const float d[3], float bas, long long n,
                                  const float* x, const float* y, const float* z, const float* m, int* species,
                                  long long* order, int cus, hipStream_t s, unsigned** counts_out,
                                  size_t* temp_need);
size_t noalloc_temp_bytes();
hipError_t sort_cell_pairs(void* temp, size_t& temp_bytes, const unsigned* key_in, unsigned* key_out,
                           const unsigned* idx_in, unsigned* idx_out, unsigned long long n, int end_bit,
                           hipStream_t s);
// Peak candidates of the grid halo finder (peaks.hip)
struct PeakPatchRec {
  int n[3];
  int pad;
  long long off;
  long long rank;
};
hipError_t peak_sort_bytes(unsigned long long n, size_t& bytes);
hipError_t launch_peak_scan(int level, const float* u, const int* conta, int nx, int ny, int nz, long long ld1,
                            long long ld2, const PeakPatchRec* pp, int np, float peak, unsigned long long* rec_r,
                            unsigned* rec_k, unsigned long long cap, unsigned long long* count, int cus,
                            hipStream_t s);
hipError_t launch_peak_sort_write(unsigned long long* rec_r, unsigned* rec_k, unsigned long long* rec_r2,
                                  unsigned* rec_k2, long long k, void* temp, size_t temp_bytes, int level, int nx,
                                  int ny, const PeakPatchRec* pp, int np, int patch0, float* dens, int* ix,
                                  int* iy, int* iz, int* patch, int cus, hipStream_t s);
// Grid masks of the density dump (masks.hip)
struct MaskPatch {
  int n[3];        // PATCHNX, PATCHNY, PATCHNZ
  int pc[3];       // PATCHX, PATCHY, PATCHZ
  int pare;        // PARE (1-based, 0 = base grid)
  int self;        // this patch's 1-based number
  float pr[3];     // PATCHRX, PATCHRY, PATCHRZ
  int pad;
  long long off;   // first cell of the patch in the U11-layout arrays
};
struct VgGeo {
  float d[3];      // DXPA, DYPA, DZPA of the level
  float fix[3];    // RXFIX, RYFIX, RZFIX (grids.f:2285-2287)
};
struct VgPair {
  int j;           // neighbour, index within the level
  int c1[3], cc1[3], len[3];
};
hipError_t launch_cr0amr(const MaskPatch* mp, int np, int nlev1, const int* lev_lo, const int* lev_n, int nl,
                         float lado0, int nx, int ny, int nz, int* cr0, long long b1, long long b2, int* cr11,
                         long long cells11, long long s1p, long long s2p, int* err, int cus, hipStream_t s);
hipError_t launch_vg_neighbours(const MaskPatch* mp, int lo, int np, const VgGeo& G, int pass, unsigned* count,
                                const unsigned long long* start, VgPair* out, hipStream_t s);
hipError_t launch_vg_solap(const MaskPatch* mp, int lo, int np, const unsigned long long* start, const VgPair* pairs,
                           int* solap, long long s1p, long long s2p, int* err, int cus, hipStream_t s);
hipError_t launch_fill_i32(int* p, long long n, int v, int cus, hipStream_t s);
// Mesh builder (mesh.hip)
struct MeshParams {
  int nl_mesh, refine_thr, min_patchsize;
  float frac_refinable;
  int bor, boramr, bor_ovlp;
  int namr[3];
  int npalev;
  int mpapolev[kNLevels];
};
struct MeshPatchHost {
  int n[3];        // PATCHNX..Z
  int pc[3];       // PATCHX..Z
  int pare;        // PARE
  float pr[3];     // PATCHRX..Z
};
struct MeshPatchDev {
  int n[3];
  float pr[3];
  long long off;   // first count of the patch (level being counted)
};
hipError_t launch_mesh_count_base(const Parts& A, int nx, int ny, int nz, float dx, float dy, float dz, int thr,
                                  int* conta, int cus, hipStream_t s);
hipError_t launch_mesh_count_level(const Plan& P, const Parts& A, int l, const MeshPatchDev* mp, float dxpa,
                                   float dypa, float dzpa, int thr, int* conta, int cus, hipStream_t s);
std::string mesh_first_level(int nx, int ny, int nz, const float* radx, const float* rady, const float* radz,
                             std::vector<int>& conta1, const MeshParams& pr, std::vector<MeshPatchHost>& out);
std::string mesh_next_level(int irpa, const std::vector<MeshPatchHost>& all, int first, int np,
                            std::vector<int>& conta, const std::vector<long long>& off, float dx, float dy, float dz,
                            const float rad1[3], const MeshParams& pr, int total_before,
                            std::vector<MeshPatchHost>& out);
hipError_t launch_plain_zero(const Plan& P, float* u1, float* u11, hipStream_t s);
hipError_t launch_plain_base(const Plan& P, const Parts& A, float* u1, int cus, hipStream_t s);
hipError_t launch_plain_flags(const Plan& P, const Parts& A, int l, unsigned* flag, int cus, hipStream_t s);
hipError_t launch_plain_compact(const unsigned* flag, const unsigned long long* start, long long n,
                                long long* idx, int cus, hipStream_t s);
hipError_t launch_plain_validate(const Plan& P, const Parts& A, Counters* C, int cus, hipStream_t s);
hipError_t launch_plain_patch(const Plan& P, const Parts& A, int g, const long long* idx,
                              const unsigned long long* nidx, long long nmax, float* u11, int cus,
                              hipStream_t s);
hipError_t launch_plain_finish(const Plan& P, float* u1, float* u11, double* grid_mass, hipStream_t s);

}  // namespace asohf
