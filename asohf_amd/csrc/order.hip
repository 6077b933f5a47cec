// order.hip -- particle orderings for gather kernels: a radix sort of
// (base-cell key, particle index) pairs with rocPRIM, kept in its own
// translation unit (the library headers are large).
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "launch.h"

namespace asohf {

hipError_t sort_cell_pairs(void* temp, size_t& temp_bytes, const unsigned* key_in, unsigned* key_out,
                           const unsigned* idx_in, unsigned* idx_out, unsigned long long n, int end_bit,
                           hipStream_t s) {
  return rocprim::radix_sort_pairs(temp, temp_bytes, key_in, key_out, idx_in, idx_out, n, 0u,
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

// Cell densities: the masses of a cell summed in FP32 in particle order (the
// pairs are sorted stably by cell), as the reference's loop does, then
// divided by BAS.  One thread per cell, starting at its first sorted pair.
__global__ void __launch_bounds__(256) species_sums(SpeciesGrid S, const unsigned* key, const unsigned* idx,
                                                    const float* m, long long n, float* dens) {
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x) {
    const unsigned c = key[k];
    if (k > 0 && key[k - 1] == c) continue;
    float s = 0.f;
    for (long long j = k; j < n && key[j] == c; ++j) s += m[idx[j]];
    dens[c] = s / S.bas;
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
  if (threadIdx.x < kNEsp) counts[threadIdx.x] = 0u;
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
// counts holds kNEsp u32, dens nmax*nmay*nmaz floats, temp the sort scratch.
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
  species_sums<<<grid, 256, 0, s>>>(S, k2, k3, m, n, dens);
  species_assign<<<grid, 256, 0, s>>>(S, x, y, z, n, dens, species, k0, k1, counts);
  tb = temp_bytes;
  e = rocprim::radix_sort_pairs(temp, tb, k0, k2, k1, k3, (unsigned long long)n, 0u, 4u, s);
  if (e != hipSuccess) return e;
  widen_order<<<grid, 256, 0, s>>>(k3, n, order);
  return hipGetLastError();
}

}  // namespace asohf
