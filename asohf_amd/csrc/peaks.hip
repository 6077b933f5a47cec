// peaks.hip -- peak candidates of the grid halo finder (HALOFIND_GRID,
// haloes_grids.f:316-404 on the base grid, 700-790 on an AMR level) and their
// order (SORT_CELLS / SORT_CELLS_AMR, haloes_grids.f:1341-1425).
//
// A cell outside the one-cell border of its grid is a candidate when its
// density is at least CONTRASTEC/6, its CONTA1 mask (AMR levels) is 1, and
// it exceeds its six face neighbours.  The reference tests the neighbours by
// the sign of FP32 differences (U(x+1)-U(x) < 0, U(x)-U(x-1) > 0, ...); IEEE
// subtraction of finite values is negative exactly when the first operand is
// smaller, so the kernels compare directly, which stays exact where a flush
// of denormal differences to zero would not.
//
// The reference collects candidates in grid, z, y, x order (the "raster
// index" below) and sorts them with ARGSORT (num.f:150-200) on 1/density: a
// bottom-up merge that, of two equal keys, emits the one from the right-hand
// run first, i.e. the one with the larger raster index.  The result is the
// stable order of (1/density ascending, raster index descending).  Here the
// candidates are appended in any order with atomics, then sorted twice with
// rocPRIM's stable radix sort: by raster index descending, then by the FP32
// bits of 1/density (positive floats order as their bit patterns).  The
// density returned is 1/(1/density), as SORT_CELLS writes it back.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "launch.h"

namespace asohf {

// One AMR patch of the scanned level.
struct PeakPatch {
  int n[3];             // PATCHNX, PATCHNY, PATCHNZ
  int pad;
  long long off;        // first cell of the patch in the field (and the mask)
  long long rank;       // raster index of its first cell within the level
};

__device__ __forceinline__ bool peak_test(const float* u, long long c, long long s1, long long s2, float peak) {
  const float v = u[c];
  if (!(v >= peak)) return false;
  return u[c + 1] < v && u[c + s1] < v && u[c + s2] < v && v > u[c - 1] && v > u[c - s1] && v > u[c - s2];
}

__device__ __forceinline__ void peak_append(unsigned long long r, float v, unsigned long long* rec_r,
                                            unsigned* rec_k, unsigned long long cap, unsigned long long* count) {
  const unsigned long long slot = atomicAdd(count, 1ull);
  if (slot < cap) {
    rec_r[slot] = r;
    rec_k[slot] = __float_as_uint(1.0f / v);
  }
}

// Base grid: U1 with leading dimensions ld1 (>= nx) and ld2 = ld1 * ldy.
__global__ void __launch_bounds__(256) peak_scan_base(const float* u, int nx, int ny, int nz, long long ld1,
                                                      long long ld2, float peak, unsigned long long* rec_r,
                                                      unsigned* rec_k, unsigned long long cap,
                                                      unsigned long long* count) {
  const long long total = (long long)nx * ny * nz;
  for (long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x; r < total;
       r += (long long)gridDim.x * blockDim.x) {
    const int ix = (int)(r % nx);
    const long long q = r / nx;
    const int jy = (int)(q % ny), kz = (int)(q / ny);
    if (ix < 1 || ix > nx - 2 || jy < 1 || jy > ny - 2 || kz < 1 || kz > nz - 2) continue;
    const long long c = ix + ld1 * jy + ld2 * kz;
    if (peak_test(u, c, ld1, ld2, peak)) peak_append((unsigned long long)r, u[c], rec_r, rec_k, cap, count);
  }
}

// Patches of one level: one workgroup per patch at a time.
// AMR levels: a workgroup per patch at a time.  Deep levels hold millions of
// candidates, and one global counter taking an atomic per wave serialises in
// L2, so a workgroup collects its candidates in LDS over whole passes of 256
// cells and reserves space in the output once per kPeakBuf of them (the
// output order does not matter: the two sorts fix it).
constexpr int kPeakBuf = 1024;
__global__ void __launch_bounds__(256) peak_scan_patches(const float* u, const int* conta, const PeakPatch* pp,
                                                         int np, long long ld1, long long ld2, float peak,
                                                         unsigned long long* rec_r, unsigned* rec_k,
                                                         unsigned long long cap, unsigned long long* count) {
  __shared__ unsigned long long buf_r[kPeakBuf];
  __shared__ unsigned buf_k[kPeakBuf];
  __shared__ unsigned s_n;
  __shared__ unsigned long long s_base;
  const int tid = threadIdx.x, nthr = blockDim.x;
  if (tid == 0) s_n = 0;
  __syncthreads();
  // moves the buffer to the output; every thread of the workgroup calls it
  auto flush = [&]() {
    const unsigned n = s_n;
    if (n == 0) return;
    if (tid == 0) s_base = atomicAdd(count, (unsigned long long)n);
    __syncthreads();
    const unsigned long long b = s_base;
    for (unsigned j = tid; j < n; j += nthr)
      if (b + j < cap) {
        rec_r[b + j] = buf_r[j];
        rec_k[b + j] = buf_k[j];
      }
    __syncthreads();
    if (tid == 0) s_n = 0;
    __syncthreads();
  };
  for (int p = blockIdx.x; p < np; p += gridDim.x) {
    const PeakPatch P = pp[p];
    const long long s1 = ld1 > 0 ? ld1 : P.n[0];
    const long long s2 = ld2 > 0 ? ld2 : (long long)P.n[0] * P.n[1];
    const int cells = P.n[0] * P.n[1] * P.n[2];
    for (int l0 = 0; l0 < cells; l0 += nthr) {   // the same passes on every thread
      const int l = l0 + tid;
      bool cand = false;
      float v = 0.f;
      if (l < cells) {
        const int ix = l % P.n[0], jy = (l / P.n[0]) % P.n[1], kz = l / (P.n[0] * P.n[1]);
        if (ix >= 1 && ix <= P.n[0] - 2 && jy >= 1 && jy <= P.n[1] - 2 && kz >= 1 && kz <= P.n[2] - 2) {
          const long long c = P.off + ix + s1 * jy + s2 * kz;
          if (!conta || conta[c] == 1) {
            cand = peak_test(u, c, s1, s2, peak);
            v = u[c];
          }
        }
      }
      if (cand) {
        const unsigned j = atomicAdd(&s_n, 1u);   // below kPeakBuf: flushed before a pass could overflow
        buf_r[j] = (unsigned long long)(P.rank + l);
        buf_k[j] = __float_as_uint(1.0f / v);
      }
      __syncthreads();
      // every thread reads the count before any adds to it again, so all take
      // the same decision
      const unsigned n_now = s_n;
      __syncthreads();
      if (n_now > (unsigned)(kPeakBuf - nthr)) flush();
    }
  }
  flush();
}

// Sorted records to the caller's arrays: 1-based cell indices and the
// 1-based global patch number (0 on the base grid).
__global__ void __launch_bounds__(256) peak_write(const unsigned long long* rec_r, const unsigned* rec_k,
                                                  long long k, int level, int nx, int ny, const PeakPatch* pp,
                                                  int np, int patch0, float* dens, int* ix, int* iy, int* iz,
                                                  int* patch) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < k;
       i += (long long)gridDim.x * blockDim.x) {
    const unsigned long long r = rec_r[i];
    int a, b, c, g = 0;
    if (level == 0) {
      a = (int)(r % (unsigned long long)nx);
      const unsigned long long q = r / (unsigned long long)nx;
      b = (int)(q % (unsigned long long)ny);
      c = (int)(q / (unsigned long long)ny);
    } else {
      int lo = 0, hi = np - 1;   // last patch whose first raster index is <= r
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((unsigned long long)pp[mid].rank <= r) lo = mid;
        else hi = mid - 1;
      }
      const int l = (int)(r - (unsigned long long)pp[lo].rank);
      a = l % pp[lo].n[0];
      b = (l / pp[lo].n[0]) % pp[lo].n[1];
      c = l / (pp[lo].n[0] * pp[lo].n[1]);
      g = patch0 + lo + 1;
    }
    dens[i] = 1.0f / __uint_as_float(rec_k[i]);
    ix[i] = a + 1;
    iy[i] = b + 1;
    iz[i] = c + 1;
    if (patch) patch[i] = g;
  }
}

hipError_t peak_sort_bytes(unsigned long long n, size_t& bytes) {
  size_t b1 = 0, b2 = 0;
  hipError_t e = rocprim::radix_sort_pairs_desc(nullptr, b1, (const unsigned long long*)nullptr,
                                                (unsigned long long*)nullptr, (const unsigned*)nullptr,
                                                (unsigned*)nullptr, n, 0u, 64u, (hipStream_t)0);
  if (e != hipSuccess) return e;
  e = rocprim::radix_sort_pairs(nullptr, b2, (const unsigned*)nullptr, (unsigned*)nullptr,
                                (const unsigned long long*)nullptr, (unsigned long long*)nullptr, n, 0u, 32u,
                                (hipStream_t)0);
  bytes = b1 > b2 ? b1 : b2;
  return e;
}

hipError_t launch_peak_scan(int level, const float* u, const int* conta, int nx, int ny, int nz, long long ld1,
                            long long ld2, const PeakPatchRec* pp, int np, float peak, unsigned long long* rec_r,
                            unsigned* rec_k, unsigned long long cap, unsigned long long* count, int cus,
                            hipStream_t s) {
  hipError_t e = launch_fill_u32(reinterpret_cast<unsigned*>(count), 0u, 2, s);
  if (e != hipSuccess) return e;
  if (level == 0) {
    const long long total = (long long)nx * ny * nz;
    long long b = (total + 255) / 256;
    if (b > (long long)cus * 64) b = (long long)cus * 64;
    peak_scan_base<<<(int)(b > 0 ? b : 1), 256, 0, s>>>(u, nx, ny, nz, ld1, ld2, peak, rec_r, rec_k, cap, count);
  } else if (np > 0) {
    const int g = np < cus * 16 ? np : cus * 16;
    peak_scan_patches<<<g, 256, 0, s>>>(u, conta, reinterpret_cast<const PeakPatch*>(pp), np, ld1, ld2, peak,
                                        rec_r, rec_k, cap, count);
  }
  return hipGetLastError();
}

// Sorts k records (rec_r/rec_k, with rec_r2/rec_k2 as the second buffers)
// and writes them out.
hipError_t launch_peak_sort_write(unsigned long long* rec_r, unsigned* rec_k, unsigned long long* rec_r2,
                                  unsigned* rec_k2, long long k, void* temp, size_t temp_bytes, int level, int nx,
                                  int ny, const PeakPatchRec* pp, int np, int patch0, float* dens, int* ix,
                                  int* iy, int* iz, int* patch, int cus, hipStream_t s) {
  if (k <= 0) return hipSuccess;
  size_t tb = temp_bytes;
  // raster index descending (keys into rec_r2, densities' keys into rec_k2)
  hipError_t e = rocprim::radix_sort_pairs_desc(temp, tb, rec_r, rec_r2, rec_k, rec_k2, (unsigned long long)k, 0u,
                                                64u, s);
  if (e != hipSuccess) return e;
  tb = temp_bytes;
  // then stably by 1/density ascending
  e = rocprim::radix_sort_pairs(temp, tb, rec_k2, rec_k, rec_r2, rec_r, (unsigned long long)k, 0u, 32u, s);
  if (e != hipSuccess) return e;
  long long b = (k + 255) / 256;
  if (b > (long long)cus * 16) b = (long long)cus * 16;
  peak_write<<<(int)b, 256, 0, s>>>(rec_r, rec_k, k, level, nx, ny, reinterpret_cast<const PeakPatch*>(pp), np,
                                    patch0, dens, ix, iy, iz, patch);
  return hipGetLastError();
}

static_assert(sizeof(PeakPatch) == sizeof(PeakPatchRec), "the host-built patch records match the kernels' view");

}  // namespace asohf
