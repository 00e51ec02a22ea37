// Particle reordering for the native block loop on gfx950.
//
// The pair kernel's cost is set by how many distinct cache lines a
// wavefront's neighbour gather touches.  Inside a block of steps the loop
// keeps particles in cell order (the order of the cell list: cells in hash
// order, ids ascending inside a cell), so that the particles of a wavefront
// share a cell and their neighbours lie in a few contiguous runs.  The
// reference's user order is restored at the end of the block, and the
// neighbour list is translated back to user ids.
#include "jmd_common.hpp"

namespace jmd {

// out[r] = in[perm[r]] for rows of `row_bytes` bytes (multiple of 4).
__global__ __launch_bounds__(256) void k_gather_rows(const uint32_t* __restrict__ in,
                                                     uint32_t* __restrict__ out,
                                                     const int* __restrict__ perm, int n,
                                                     int row_words, int in_stride_words,
                                                     int out_stride_words) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long total = (long long)n * row_words;
  long long step = (long long)gridDim.x * blockDim.x;
  for (; t < total; t += step) {
    int r = (int)(t / row_words);
    int w = (int)(t - (long long)r * row_words);
    out[(long long)r * out_stride_words + w] = in[(long long)perm[r] * in_stride_words + w];
  }
}

// out[perm[r]] = in[r] (the inverse of k_gather_rows).
__global__ __launch_bounds__(256) void k_scatter_rows(const uint32_t* __restrict__ in,
                                                      uint32_t* __restrict__ out,
                                                      const int* __restrict__ perm, int n,
                                                      int row_words, int in_stride_words,
                                                      int out_stride_words) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long total = (long long)n * row_words;
  long long step = (long long)gridDim.x * blockDim.x;
  for (; t < total; t += step) {
    int r = (int)(t / row_words);
    int w = (int)(t - (long long)r * row_words);
    out[(long long)perm[r] * out_stride_words + w] = in[(long long)r * in_stride_words + w];
  }
}

// Dense list in permuted ids -> user ids: out[perm[r], k] = perm[in[r, k]]
// (padding n stays n).
__global__ __launch_bounds__(256) void k_remap_dense(const int* __restrict__ in,
                                                     int* __restrict__ out,
                                                     const int* __restrict__ perm, int n,
                                                     int width) {
  // one wavefront per row: the row's destination is read once, lanes walk the
  // entries (no per-entry division to find the row)
  const int lane = lane_id();
  const long long nw = (long long)gridDim.x * (blockDim.x / kWave);
  for (long long r = (long long)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; r < n;
       r += nw) {
    const int* src = in + r * width;
    int* dst = out + (long long)perm[r] * width;
    for (int k = lane; k < width; k += kWave) {
      const int j = src[k];
      dst[k] = j < n ? perm[j] : n;
    }
  }
}

// The same over the flat [n * width] array in 16-byte chunks: each lane
// loads four consecutive entries with one vector load, finds their row once
// (32-bit division; n * width < 2^31), gathers the four ids' new numbers
// independently and writes the entries into the destination row(s).  The
// rows' entries stay contiguous, so consecutive lanes write consecutive
// addresses except where a chunk crosses a row end.  in must be 16-byte
// aligned; total % 4 trailing entries are done by the first lanes.
__global__ __launch_bounds__(256) void k_remap_dense_flat(const int* __restrict__ in,
                                                          int* __restrict__ out,
                                                          const int* __restrict__ perm, int n,
                                                          int width) {
  const unsigned total = (unsigned)n * (unsigned)width;
  const unsigned nq = total / 4u;
  const unsigned step = gridDim.x * blockDim.x;
  for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += step) {
    const int4 v = reinterpret_cast<const int4*>(in)[q];
    const unsigned e = q * 4u;
    unsigned row = e / (unsigned)width;
    unsigned col = e - row * (unsigned)width;
    const int jv[4] = {v.x, v.y, v.z, v.w};
    int pj[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) pj[u] = jv[u] < n ? perm[jv[u]] : n;
    long long dst = (long long)perm[row] * width;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      out[dst + col] = pj[u];
      if (++col == (unsigned)width && u < 3) {
        col = 0;
        ++row;
        dst = (long long)perm[row] * width;
      }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < total % 4u) {
    const unsigned e = nq * 4u + threadIdx.x;
    const unsigned row = e / (unsigned)width, col = e - row * (unsigned)width;
    const int j = in[e];
    out[(long long)perm[row] * width + col] = j < n ? perm[j] : n;
  }
}

// Sparse list [2, width] through perm (entries keep their slots).
// OrderedSparse keeps receiver < sender (partition.py:998): a pair whose
// order the renumbering reversed is stored swapped.
__global__ __launch_bounds__(256) void k_remap_ids(const int* __restrict__ in,
                                                   int* __restrict__ out,
                                                   const int* __restrict__ perm, int n,
                                                   long long width, int ordered) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long step = (long long)gridDim.x * blockDim.x;
  for (; t < width; t += step) {
    const int a = in[t], b = in[width + t];
    int ra = a < n ? perm[a] : n;
    int sb = b < n ? perm[b] : n;
    if (ordered && ra < n && sb < n && ra > sb) {
      const int x = ra;
      ra = sb;
      sb = x;
    }
    out[t] = ra;
    out[width + t] = sb;
  }
}

static int grid_for(long long total) {
  long long b = (total + 255) / 256;
  if (b < 1) b = 1;
  if (b > 8192) b = 8192;
  return (int)b;
}

}  // namespace jmd

using namespace jmd;

extern "C" {

int jmd_permute_rows(const void* in, void* out, const int32_t* perm, int64_t n,
                     int32_t row_bytes, int32_t in_stride_bytes, int32_t out_stride_bytes,
                     int32_t inverse, void* stream) {
  if (n < 0 || n > 0x7fffffff || row_bytes <= 0 || (row_bytes & 3) || (in_stride_bytes & 3) ||
      (out_stride_bytes & 3) || in_stride_bytes < row_bytes || out_stride_bytes < row_bytes ||
      (n > 0 && (!in || !out || !perm)) || in == out)
    return JMD_ERR_INVALID;
  if (n == 0) return JMD_OK;
  hipStream_t st = (hipStream_t)stream;
  int rw = row_bytes / 4;
  long long total = n * (long long)rw;
  if (inverse)
    hipLaunchKernelGGL(k_scatter_rows, dim3(grid_for(total)), dim3(256), 0, st,
                       (const uint32_t*)in, (uint32_t*)out, perm, (int)n, rw,
                       in_stride_bytes / 4, out_stride_bytes / 4);
  else
    hipLaunchKernelGGL(k_gather_rows, dim3(grid_for(total)), dim3(256), 0, st,
                       (const uint32_t*)in, (uint32_t*)out, perm, (int)n, rw,
                       in_stride_bytes / 4, out_stride_bytes / 4);
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

int jmd_remap_neighbor_ids(const int32_t* idx_in, int32_t* idx_out, const int32_t* perm,
                           int64_t n, int32_t format, int64_t width, void* stream) {
  if (n < 0 || n > 0x7fffffff || width < 0 || (n > 0 && !perm) || idx_in == idx_out ||
      (width > 0 && (!idx_in || !idx_out)))
    return JMD_ERR_INVALID;
  if (format != JMD_DENSE && format != JMD_SPARSE && format != JMD_ORDERED_SPARSE)
    return JMD_ERR_INVALID;
  if (width == 0 || n == 0) return JMD_OK;
  hipStream_t st = (hipStream_t)stream;
  if (format == JMD_DENSE) {
    if (width > 0x7fffffff) return JMD_ERR_INVALID;
    if (n * width < 0x7fffffffLL && ((uintptr_t)idx_in & 15) == 0) {
      long long qb = (n * width / 4 + 255) / 256;
      if (qb < 1) qb = 1;
      if (qb > 16384) qb = 16384;
      hipLaunchKernelGGL(k_remap_dense_flat, dim3((unsigned)qb), dim3(256), 0, st, idx_in,
                         idx_out, perm, (int)n, (int)width);
    } else {
      long long rb = (n + 3) / 4;  // four rows (wavefronts) per workgroup
      hipLaunchKernelGGL(k_remap_dense,
                         dim3((unsigned)(rb < 65536 ? (rb < 1 ? 1 : rb) : 65536)), dim3(256), 0,
                         st, idx_in, idx_out, perm, (int)n, (int)width);
    }
  } else {
    hipLaunchKernelGGL(k_remap_ids, dim3(grid_for(width)), dim3(256), 0, st, idx_in, idx_out,
                       perm, (int)n, width, format == JMD_ORDERED_SPARSE ? 1 : 0);
  }
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

}  // extern "C"
