// order.hip -- particle orderings for gather kernels: a radix sort of
// (base-cell key, particle index) pairs with rocPRIM, kept in its own
// translation unit (the library headers are large).
#include <hip/hip_runtime.h>

#include <mutex>
#include <rocprim/device/device_radix_sort.hpp>

#include "launch.h"

namespace asohf {

hipError_t sort_cell_pairs(void* temp, size_t& temp_bytes, const unsigned* key_in, unsigned* key_out,
                           const unsigned* idx_in, unsigned* idx_out, unsigned long long n, int end_bit,
                           hipStream_t s) {
  // 10 bits per onesweep pass: a 30-bit cell key of a 1024^3 grid in three
  // passes instead of four
  using Cfg = rocprim::radix_sort_config<
      rocprim::default_config, rocprim::default_config,
      rocprim::radix_sort_onesweep_config<rocprim::kernel_config<1024, 16>, rocprim::kernel_config<1024, 16>, 10,
                                          rocprim::block_radix_rank_algorithm::match>>;
  return rocprim::radix_sort_pairs<Cfg>(temp, temp_bytes, key_in, key_out, idx_in, idx_out, n, 0u,
                                        (unsigned)end_bit, s);
}

// ---------------------------------------------------------------------------
// Species by local density (SORT_DM_PARTICLES_LOCALDENSITY, reader.f:140-336)
// ---------------------------------------------------------------------------
struct SpeciesGrid {
  int n[3];       // NMAX, NMAY, NMAZ: the reference's compile-time grid
  float d[3];     // DX, DY, DZ
  float vmin[3];  // -NMAX*DX/2 (reader.f:183-185)
  float bas;      // DX*DY*DZ*RODO*RE0**3 (reader.f:203)
};

// NGP cell INT((X-XMIN)/DX)+1 clamped to [1, NMAX] (reader.f:190-200), 0-based
__device__ __forceinline__ unsigned species_cell(const SpeciesGrid& S, float x, float y, float z) {
  const float v[3] = {x, y, z};
  int c[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float t = (v[a] - S.vmin[a]) / S.d[a];
    int i = (t < 2.0e9f && t > -2.0e9f) ? (int)t + 1 : (t > 0.f ? S.n[a] : 1);
    if (i < 1) i = 1;
    if (i > S.n[a]) i = S.n[a];
    c[a] = i - 1;
  }
  return (unsigned)c[0] + (unsigned)S.n[0] * ((unsigned)c[1] + (unsigned)S.n[1] * (unsigned)c[2]);
}

__global__ void __launch_bounds__(256) species_keys(SpeciesGrid S, const float* x, const float* y,
                                                    const float* z, long long n, unsigned* key, unsigned* idx) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    key[i] = species_cell(S, x[i], y[i], z[i]);
    idx[i] = (unsigned)i;
  }
}

// Runs of equal keys in the sorted cell keys: the last position of each run
// writes the run's end (exclusive) into runend[cell].  runend is the density
// array itself, read back as unsigned by the run's first position before the
// sums overwrite it (a cell's slot is touched by its own run only).
__global__ void __launch_bounds__(256) species_run_ends(const unsigned* key, long long n, unsigned* runend) {
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x) {
    const unsigned c = key[k];
    if (k + 1 == n || key[k + 1] != c) runend[c] = (unsigned)(k + 1);
  }
}

constexpr unsigned kLongRun = 32;  // runs longer than this are summed by a whole wave

// Cell densities: the masses of a cell summed in FP32 in particle order (the
// pairs are sorted stably by cell), as the reference's loop does, then
// divided by BAS.  Runs of up to kLongRun particles: one thread each, loads
// issued four at a time ahead of the (ordered) adds.  Longer runs are listed
// as {start, end, cell} for species_sums_long.
__global__ void __launch_bounds__(256) species_sums(SpeciesGrid S, const unsigned* key, const unsigned* idx,
                                                    const float* m, long long n, float* dens, uint4* longs,
                                                    unsigned* nlong) {
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x) {
    const unsigned c = key[k];
    if (k > 0 && key[k - 1] == c) continue;
    const unsigned b = (unsigned)k;
    const unsigned e = reinterpret_cast<const unsigned*>(dens)[c];
    if (e - b > kLongRun) {
      longs[atomicAdd(nlong, 1u)] = make_uint4(b, e, c, 0u);
      continue;
    }
    float s = 0.f;
    unsigned j = b;
    for (; j + 4 <= e; j += 4) {
      const unsigned i0 = idx[j], i1 = idx[j + 1], i2 = idx[j + 2], i3 = idx[j + 3];
      const float a0 = m[i0], a1 = m[i1], a2 = m[i2], a3 = m[i3];
      s += a0;
      s += a1;
      s += a2;
      s += a3;
    }
    for (; j < e; ++j) s += m[idx[j]];
    dens[c] = s / S.bas;
  }
}

// Long runs: one wave per run, taken from the list through a work counter.
// The wave gathers the run's masses 1024 at a time (lane l holds positions
// b + 64t + l, t = 0..15) with the next block's masses and the indices of the
// block after it in flight, stages the block in its own LDS slice, and every
// lane then adds the slice into one running sum in position order (broadcast
// reads, four masses each), so the result is the sequential FP32 sum.  Lanes
// past the run's end stage +0.0f, which leaves a sum of non-negative masses
// unchanged.
constexpr int kRunBlk = 16;  // 64-position groups per block
constexpr int kSumWaves = 4;
__global__ void __launch_bounds__(64 * kSumWaves) species_sums_long(SpeciesGrid S, const unsigned* idx,
                                                                    const float* m, float* dens,
                                                                    const uint4* longs, const unsigned* nlong,
                                                                    unsigned* next) {
  __shared__ float4 stage_s[kSumWaves][kRunBlk * 16];
  const unsigned lane = threadIdx.x & 63u;
  float4* stage = stage_s[threadIdx.x >> 6];
  float* stage1 = reinterpret_cast<float*>(stage);
  const unsigned total = *nlong;
  for (;;) {
    unsigned w = 0;
    if (lane == 0) w = atomicAdd(next, 1u);
    w = __builtin_amdgcn_readfirstlane(w);
    if (w >= total) break;
    const uint4 r = longs[w];
    // positions in 64 bits: blocks ahead of the run's end stay past it
    const unsigned long long b0 = r.x, e = r.y;
    auto ld_idx = [&](unsigned long long b, unsigned (&o)[kRunBlk]) {
#pragma unroll
      for (int t = 0; t < kRunBlk; ++t) {
        const unsigned long long p = b + 64u * t + lane;
        o[t] = p < e ? idx[p] : 0xffffffffu;
      }
    };
    auto ld_m = [&](const unsigned (&ix)[kRunBlk], float (&o)[kRunBlk]) {
#pragma unroll
      for (int t = 0; t < kRunBlk; ++t) o[t] = ix[t] != 0xffffffffu ? m[ix[t]] : 0.f;
    };
    unsigned ia[kRunBlk], ib[kRunBlk];
    float ma[kRunBlk], mb[kRunBlk];
    ld_idx(b0, ia);
    ld_m(ia, ma);
    ld_idx(b0 + 64u * kRunBlk, ib);
    float s = 0.f;
    for (unsigned long long b = b0; b < e; b += 64u * kRunBlk) {
      ld_m(ib, mb);                          // masses of the next block
      ld_idx(b + 2u * 64u * kRunBlk, ia);    // indices of the block after it
#pragma unroll
      for (int t = 0; t < kRunBlk; ++t) stage1[64 * t + lane] = ma[t];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
      for (int t = 0; t < kRunBlk; ++t) {
        if (b + 64u * t >= e) break;
        float4 v[16];  // the group's 64 masses read at once: one LDS latency per group
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = stage[16 * t + q];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          s += v[q].x;
          s += v[q].y;
          s += v[q].z;
          s += v[q].w;
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
      for (int t = 0; t < kRunBlk; ++t) {
        ma[t] = mb[t];
        ib[t] = ia[t];
      }
    }
    if (lane == 0) dens[r.z] = s / S.bas;
  }
}

// MAX(MIN(INT(LOG(D)/LOG(8.0)), N_ESP-1), 0) for D > 0, else 0 (reader.f:219-236)
__global__ void __launch_bounds__(256) species_assign(SpeciesGrid S, const float* x, const float* y,
                                                      const float* z, long long n, const float* dens,
                                                      int* species, unsigned* skey, unsigned* sidx,
                                                      unsigned* counts) {
  __shared__ unsigned cnt[kNEsp];
  if (threadIdx.x < kNEsp) cnt[threadIdx.x] = 0u;
  __syncthreads();
  const float l8 = logf(8.0f);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float b = dens[species_cell(S, x[i], y[i], z[i])];
    int s = 0;
    if (b > 0.f) {
      s = (int)(logf(b) / l8);
      s = s > kNEsp - 1 ? kNEsp - 1 : (s < 0 ? 0 : s);
    }
    if (species) species[i] = s;
    skey[i] = (unsigned)s;
    sidx[i] = (unsigned)i;
    atomicAdd(&cnt[s], 1u);
  }
  __syncthreads();
  if (threadIdx.x < kNEsp && cnt[threadIdx.x]) atomicAdd(&counts[threadIdx.x], cnt[threadIdx.x]);
}

__global__ void __launch_bounds__(256) widen_order(const unsigned* in, long long n, long long* out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = (long long)in[i];
}

__global__ void clear_counts(unsigned* counts) {
  if (threadIdx.x < kNEsp + 2) counts[threadIdx.x] = 0u;
}

static int species_grid(long long n, int cus) {
  long long b = (n + 255) / 256;
  const long long cap = (long long)(cus > 1 ? cus : 1) * 16;
  if (b > cap) b = cap;
  return (int)(b > 0 ? b : 1);
}

hipError_t species_sort_bytes(unsigned long long n, size_t& bytes) {
  size_t b1 = 0, b2 = 0;
  hipError_t e = rocprim::radix_sort_pairs(nullptr, b1, (const unsigned*)nullptr, (unsigned*)nullptr,
                                           (const unsigned*)nullptr, (unsigned*)nullptr, n, 0u, 32u, (hipStream_t)0);
  if (e != hipSuccess) return e;
  e = rocprim::radix_sort_pairs(nullptr, b2, (const unsigned*)nullptr, (unsigned*)nullptr, (const unsigned*)nullptr,
                                (unsigned*)nullptr, n, 0u, 4u, (hipStream_t)0);
  bytes = b1 > b2 ? b1 : b2;
  return e;
}

// The whole routine on the device: k[0..3] are four n-element u32 buffers,
// counts holds kNEsp + 2 u32, dens nmax*nmay*nmaz floats, temp the sort scratch.
hipError_t launch_species(const int nmax[3], const float d[3], float rodo_re0_bas, long long n, const float* x,
                          const float* y, const float* z, const float* m, int* species, long long* order,
                          unsigned* counts, float* dens, unsigned* k0, unsigned* k1, unsigned* k2, unsigned* k3,
                          void* temp, size_t temp_bytes, int cus, hipStream_t s) {
  SpeciesGrid S;
  for (int a = 0; a < 3; ++a) {
    S.n[a] = nmax[a];
    S.d[a] = d[a];
    S.vmin[a] = -(((float)nmax[a] * d[a]) / 2.0f);
  }
  S.bas = rodo_re0_bas;
  const int grid = species_grid(n, cus);
  clear_counts<<<1, 64, 0, s>>>(counts);
  if (n <= 0) return hipGetLastError();
  species_keys<<<grid, 256, 0, s>>>(S, x, y, z, n, k0, k1);
  int bits = 1;
  const long long nc = (long long)nmax[0] * nmax[1] * nmax[2];
  while (bits < 32 && (1ll << bits) < nc) ++bits;
  size_t tb = temp_bytes;
  hipError_t e = rocprim::radix_sort_pairs(temp, tb, k0, k2, k1, k3, (unsigned long long)n, 0u, (unsigned)bits, s);
  if (e != hipSuccess) return e;
  // k0 is free between the sort and species_assign: it holds the long-run list
  // (at most n / kLongRun records of 4 words); counts[kNEsp] counts the list,
  // counts[kNEsp + 1] hands its records out.
  species_run_ends<<<grid, 256, 0, s>>>(k2, n, reinterpret_cast<unsigned*>(dens));
  species_sums<<<grid, 256, 0, s>>>(S, k2, k3, m, n, dens, reinterpret_cast<uint4*>(k0), counts + kNEsp);
  species_sums_long<<<cus * 4, 64 * kSumWaves, 0, s>>>(S, k3, m, dens, reinterpret_cast<const uint4*>(k0), counts + kNEsp,
                                            counts + kNEsp + 1);
  species_assign<<<grid, 256, 0, s>>>(S, x, y, z, n, dens, species, k0, k1, counts);
  tb = temp_bytes;
  e = rocprim::radix_sort_pairs(temp, tb, k0, k2, k1, k3, (unsigned long long)n, 0u, 4u, s);
  if (e != hipSuccess) return e;
  widen_order<<<grid, 256, 0, s>>>(k3, n, order);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The species partition without allocation (asohf_hip_species_noalloc): its
// scratch (sort keys and indices, the NGP grid, the class counts, the radix
// sort's temporary storage) is carved from `scratch`, a file-scope device
// array that every GPU the library runs on holds until the process ends.
// In host code `scratch` names a host-side stand-in: its device address is
// taken once per device with hipGetSymbolAddress and only that address is
// handed to kernels.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kNoAllocDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kNoAllocDevices] = {};
constexpr size_t kNaKeys = sizeof(unsigned) * (size_t)kNoAllocMaxParticles;   // one key/index array
constexpr size_t kNaDens = sizeof(float) * (size_t)kNoAllocMaxCells;
constexpr size_t kNaCounts = 256;
constexpr size_t kNaTemp = kNoAllocBytes - 4 * kNaKeys - kNaDens - kNaCounts;
static_assert(4 * kNaKeys + kNaDens + kNaCounts < kNoAllocBytes, "the carve-out fits the array");
}  // namespace

hipError_t noalloc_scratch_init(int device, void** out) {
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (device < 0 || device >= kNoAllocDevices) return hipErrorInvalidDevice;
  if (!g_noalloc_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_noalloc_ptr[device] = p;
  }
  *out = g_noalloc_ptr[device];
  return hipSuccess;
}

hipError_t launch_species_noalloc(void* base, const int nmax[3], const float d[3], float bas, long long n,
                                  const float* x, const float* y, const float* z, const float* m, int* species,
                                  long long* order, int cus, hipStream_t s, unsigned** counts_out,
                                  size_t* temp_need) {
  unsigned char* b = static_cast<unsigned char*>(base);
  unsigned* k[4];
  for (int i = 0; i < 4; ++i) k[i] = reinterpret_cast<unsigned*>(b + i * kNaKeys);
  float* dens = reinterpret_cast<float*>(b + 4 * kNaKeys);
  unsigned* counts = reinterpret_cast<unsigned*>(b + 4 * kNaKeys + kNaDens);
  void* temp = b + 4 * kNaKeys + kNaDens + kNaCounts;
  *counts_out = counts;
  size_t tb = 0;
  hipError_t e = species_sort_bytes((unsigned long long)(n > 0 ? n : 1), tb);
  if (e != hipSuccess) return e;
  *temp_need = tb;
  if (tb > kNaTemp) return hipSuccess;   // the caller reports the capacity error
  return launch_species(nmax, d, bas, n, x, y, z, m, species, order, counts, dens, k[0], k[1], k[2], k[3], temp,
                        kNaTemp, cus, s);
}

size_t noalloc_temp_bytes() { return kNaTemp; }

}  // namespace asohf
