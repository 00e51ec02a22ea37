// Multi-image neighbour lists for small periodic cells
// (jax_md/custom_partition.py:147-626,800-1122).
//
// Every pair (i, j) is tested against every lattice shift s of a stencil
// [-n_0, n_0] x ... (the shift counts n_d from the reciprocal heights, on the
// host): d = (x_j + s T^T) - x_i in real space, float32, kept when
// d^2 < c^2 and not (i = j, s = 0).  Sparse lists hold the kept (s, i, j)
// triples in that flattened order (shift slowest) with receiver i, sender j
// and the integer shift; OrderedSparse keeps i < j for the zero shift and
// only canonical shifts (first nonzero component positive) otherwise; Dense
// rows hold, per atom, the max_neighbors nearest kept (s, j) in order of
// increasing d^2, ties by flat index s N + j (a stable argsort).
//
// Launches: real positions (x = T u for fractional input, products summed
// left to right), a count pass (one wavefront per (s, i) row, or per atom for
// Dense), an exclusive scan of the Sparse row counts, and the ordered write.

#include "jmd_common.hpp"

namespace jmd {

struct MIGeom {
  int dim;
  int n_shift;  // stencil size
  int zero;     // index of the zero shift
  int nmax[3];  // shifts -n..n per dimension (0 for non-periodic)
  float T[3][3];
  float cutoff_sq;
};

// integer shift vector of stencil index k (meshgrid 'ij': first dimension slowest)
__device__ __forceinline__ void mi_shift(const MIGeom& G, int k, int* s) {
  for (int d = G.dim - 1; d >= 0; --d) {
    const int w = 2 * G.nmax[d] + 1;
    s[d] = k % w - G.nmax[d];
    k /= w;
  }
}

// real-space shift: s T^T (products rounded, summed left to right)
__device__ __forceinline__ void mi_shift_real(const MIGeom& G, const int* s, float* r) {
#pragma clang fp contract(off)
  for (int a = 0; a < G.dim; ++a) {
    float acc = (float)s[0] * G.T[a][0];
    for (int b = 1; b < G.dim; ++b) acc = acc + (float)s[b] * G.T[a][b];
    r[a] = acc;
  }
}

__device__ __forceinline__ bool mi_canonical(const MIGeom& G, const int* s) {
  for (int d = 0; d < G.dim; ++d)
    if (s[d] != 0) return s[d] > 0;
  return true;
}

// d^2 of (x_j + sr) - x_i, left-to-right sums, no contraction
__device__ __forceinline__ float mi_d2(const float4 xi, const float4 xj, const float* sr, int dim) {
#pragma clang fp contract(off)
  const float d0 = (xj.x + sr[0]) - xi.x;
  const float d1 = (xj.y + sr[1]) - xi.y;
  float acc = d0 * d0 + d1 * d1;
  if (dim == 3) {
    const float d2 = (xj.z + sr[2]) - xi.z;
    acc = acc + d2 * d2;
  }
  return acc;
}

__global__ __launch_bounds__(256) void k_mi_real(const float* __restrict__ pos, int n, int dim,
                                                 int fractional, MIGeom G,
                                                 float4* __restrict__ xr) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float u[3] = {0.0f, 0.0f, 0.0f}, x[3] = {0.0f, 0.0f, 0.0f};
  for (int d = 0; d < dim; ++d) u[d] = pos[(long long)i * dim + d];
  if (fractional) {
    for (int a = 0; a < dim; ++a) {
      float acc = u[0] * G.T[a][0];
      for (int b = 1; b < dim; ++b) acc = acc + u[b] * G.T[a][b];
      x[a] = acc;
    }
  } else {
    for (int d = 0; d < dim; ++d) x[d] = u[d];
  }
  xr[i] = make_float4(x[0], x[1], x[2], 0.0f);
}

// Sparse formats: one wavefront per row r = s N + i; MODE 0 counts, 1 writes.
template <int MODE>
__global__ __launch_bounds__(256) void k_mi_sparse(const float4* __restrict__ xr, int n,
                                                   MIGeom G, int ordered,
                                                   int* __restrict__ row_count,
                                                   const int* __restrict__ row_offset,
                                                   long long capacity, int* __restrict__ idx,
                                                   int* __restrict__ eshift) {
  const int lane = lane_id();
  const long long nrow = (long long)G.n_shift * n;
  for (long long r = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kWave; r < nrow;
       r += (long long)gridDim.x * blockDim.x / kWave) {
    const int k = (int)(r / n), i = (int)(r % n);
    int s[3] = {0, 0, 0};
    mi_shift(G, k, s);
    float sr[3] = {0.0f, 0.0f, 0.0f};
    mi_shift_real(G, s, sr);
    const bool zero = k == G.zero;
    const bool canon = mi_canonical(G, s);
    const float4 xi = xr[i];
    long long at = MODE == 1 ? row_offset[r] : 0;
    int cnt = 0;
    for (int j0 = 0; j0 < n; j0 += kWave) {
      const int j = j0 + lane;
      bool keep = false;
      if (j < n) {
        keep = mi_d2(xi, xr[j], sr, G.dim) < G.cutoff_sq && !(zero && j == i);
        if (ordered) keep = keep && (zero ? i < j : canon);
      }
      const unsigned long long bal = __ballot(keep);
      if (MODE == 1 && keep) {
        const long long e = at + wave_prefix_count(bal);
        if (e < capacity) {
          idx[e] = i;             // receiver
          idx[capacity + e] = j;  // sender
          for (int d = 0; d < G.dim; ++d) eshift[e * G.dim + d] = s[d];
        }
      }
      at += __popcll(bal);
      cnt += __popcll(bal);
    }
    if (MODE == 0 && lane == 0) row_count[r] = cnt;
  }
}

// Exclusive scan of m counts (one workgroup of 1024), 64-bit running sums
// clamped to int range; total into out[0].
__global__ __launch_bounds__(1024) void k_mi_scan(const int* __restrict__ cnt,
                                                  int* __restrict__ off, long long m,
                                                  long long* __restrict__ total) {
  __shared__ long long s[1024];
  const int t = threadIdx.x;
  const long long per = (m + 1023) / 1024;
  const long long beg = t * per, end = beg + per < m ? beg + per : m;
  long long local = 0;
  for (long long c = beg; c < end; ++c) local += cnt[c];
  s[t] = local;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const long long v = t >= o ? s[t - o] : 0;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  long long run = t == 0 ? 0 : s[t - 1];
  for (long long c = beg; c < end; ++c) {
    off[c] = (int)(run < 0x7fffffffLL ? run : 0x7fffffffLL);
    run += cnt[c];
  }
  if (t == 1023) total[0] = s[1023];
}

// Sparse padding beyond the valid entries, overflow flag.
__global__ __launch_bounds__(256) void k_mi_pad(int* __restrict__ idx, int* __restrict__ eshift,
                                                long long capacity, int n, int dim,
                                                const long long* __restrict__ total,
                                                uint8_t* __restrict__ overflow) {
  const long long t0 = total[0];
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < capacity;
       e += (long long)gridDim.x * blockDim.x) {
    if (e >= t0) {
      idx[e] = n;
      idx[capacity + e] = n;
      for (int d = 0; d < dim; ++d) eshift[e * dim + d] = 0;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && overflow) overflow[0] = t0 > capacity ? 1 : 0;
}

__global__ void k_fill_u64(unsigned long long* p, int m, unsigned long long v) {
  if ((int)threadIdx.x < m) p[threadIdx.x] = v;
}

__global__ void k_fill_u8(uint8_t* p, int m, uint8_t v) {
  if ((int)threadIdx.x < m) p[threadIdx.x] = v;
}

__global__ void k_copy_i64_u64(const long long* a, unsigned long long* b) {
  if (threadIdx.x == 0) b[0] = (unsigned long long)a[0];
}

constexpr int kMIDenseLDS = 2048;  // kept candidates of one atom held in LDS

// Dense: one workgroup per atom.  MODE 0 counts the kept candidates (max into
// total); MODE 1 ranks them by (d^2, flat index) and writes the first width.
template <int MODE>
__global__ __launch_bounds__(256) void k_mi_dense(const float4* __restrict__ xr, int n,
                                                  MIGeom G, long long width,
                                                  int* __restrict__ idx,
                                                  int* __restrict__ eshift,
                                                  unsigned long long* __restrict__ total,
                                                  uint8_t* __restrict__ overflow) {
  __shared__ float s_d2[kMIDenseLDS];
  __shared__ int s_flat[kMIDenseLDS];
  __shared__ int s_wc[4];
  const int tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
  const long long ncand = (long long)G.n_shift * n;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const float4 xi = xr[i];
    int kept = 0;  // workgroup-uniform
    for (long long c0 = 0; c0 < ncand; c0 += blockDim.x) {
      const long long c = c0 + tid;
      bool keep = false;
      float d2 = 0.0f;
      if (c < ncand) {
        const int k = (int)(c / n), j = (int)(c % n);
        int s[3] = {0, 0, 0};
        mi_shift(G, k, s);
        float sr[3] = {0.0f, 0.0f, 0.0f};
        mi_shift_real(G, s, sr);
        d2 = mi_d2(xi, xr[j], sr, G.dim);
        keep = d2 < G.cutoff_sq && !(k == G.zero && j == i);
      }
      const unsigned long long bal = __ballot(keep);
      if (lane == 0) s_wc[wave] = (int)__popcll(bal);
      __syncthreads();
      int before = 0, all = 0;
      for (int w = 0; w < (int)(blockDim.x / kWave); ++w) {
        before += w < wave ? s_wc[w] : 0;
        all += s_wc[w];
      }
      if (MODE == 1 && keep) {
        const int q = kept + before + wave_prefix_count(bal);
        if (q < kMIDenseLDS) {
          s_d2[q] = d2;
          s_flat[q] = (int)c;
        }
      }
      kept += all;
      __syncthreads();
    }
    if (MODE == 0) {
      if (tid == 0) atomicMax(total, (unsigned long long)kept);
      continue;
    }
    const int held = kept < kMIDenseLDS ? kept : kMIDenseLDS;
    if (tid == 0 && (kept > width || kept > kMIDenseLDS)) overflow[0] = 1;
    // rank of each held candidate: number of (d^2, flat) keys before it
    for (int q = tid; q < held; q += blockDim.x) {
      const float dq = s_d2[q];
      const int fq = s_flat[q];
      int rank = 0;
      for (int p = 0; p < held; ++p) {
        const float dp = s_d2[p];
        rank += (dp < dq || (dp == dq && s_flat[p] < fq)) ? 1 : 0;
      }
      if (rank < width) {
        const int k = fq / n, j = fq % n;
        int s[3] = {0, 0, 0};
        mi_shift(G, k, s);
        idx[(long long)i * width + rank] = j;
        for (int d = 0; d < G.dim; ++d) eshift[((long long)i * width + rank) * G.dim + d] = s[d];
      }
    }
    for (long long q = held + tid; q < width; q += blockDim.x) {
      idx[(long long)i * width + q] = n;
      for (int d = 0; d < G.dim; ++d) eshift[((long long)i * width + q) * G.dim + d] = 0;
    }
    __syncthreads();  // LDS reuse by the next atom
  }
}

// max over atoms of |x - x_ref|^2 in real space, unwrapped (the rebuild test
// of custom_partition.py:961-985); float bits through an unsigned max.
__global__ __launch_bounds__(256) void k_mi_maxdisp(const float* __restrict__ pos,
                                                    const float* __restrict__ ref, int n,
                                                    int dim, int fractional, MIGeom G,
                                                    MIGeom Gr, unsigned* __restrict__ out) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float d2 = 0.0f;
  if (i < n) {
    float u[3] = {0, 0, 0}, v[3] = {0, 0, 0}, x[3] = {0, 0, 0}, y[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) {
      u[d] = pos[(long long)i * dim + d];
      v[d] = ref[(long long)i * dim + d];
    }
    for (int a = 0; a < dim; ++a) {
      if (fractional) {
        float p = u[0] * G.T[a][0], q = v[0] * Gr.T[a][0];
        for (int b = 1; b < dim; ++b) {
          p = p + u[b] * G.T[a][b];
          q = q + v[b] * Gr.T[a][b];
        }
        x[a] = p;
        y[a] = q;
      } else {
        x[a] = u[a];
        y[a] = v[a];
      }
    }
    for (int d = 0; d < dim; ++d) {
      const float e = x[d] - y[d];
      d2 = d == 0 ? e * e : d2 + e * e;
    }
  }
  for (int off = 32; off > 0; off >>= 1) d2 = fmaxf(d2, __shfl_xor(d2, off, kWave));
  if (lane_id() == 0) atomicMax(out, __float_as_uint(d2));
}

}  // namespace jmd

using namespace jmd;

extern "C" {

size_t jmd_multi_image_workspace_bytes(int64_t n, int32_t n_shift) {
  if (n < 0 || n_shift < 1) return 0;
  const size_t rows = (size_t)n * (size_t)n_shift;
  return 256 + (size_t)n * 16 + 2 * rows * sizeof(int) + 64;
}

int jmd_multi_image_build(const float* position, int64_t n, int32_t dim, int32_t fractional,
                          const float* box, const int32_t* shift_max, float cutoff_sq,
                          int32_t format, int32_t* idx, int32_t* edge_shift, int64_t capacity,
                          unsigned long long* occupancy, uint8_t* overflow, void* workspace,
                          void* stream) {
  if ((dim != 2 && dim != 3) || n < 1 || n > 0x7fffffff || !position || !box || !shift_max ||
      !occupancy || !workspace || capacity < 0 || !(cutoff_sq >= 0.0f))
    return JMD_ERR_INVALID;
  if (format != JMD_DENSE && format != JMD_SPARSE && format != JMD_ORDERED_SPARSE)
    return JMD_ERR_INVALID;
  const bool measure = capacity == 0 || !idx;
  if (!measure && (!edge_shift || !overflow)) return JMD_ERR_INVALID;
  MIGeom G;
  G.dim = dim;
  G.cutoff_sq = cutoff_sq;
  long long ns = 1;
  for (int d = 0; d < 3; ++d) {
    G.nmax[d] = d < dim ? shift_max[d] : 0;
    if (G.nmax[d] < 0 || G.nmax[d] > 64) return JMD_ERR_INVALID;
    ns *= 2 * G.nmax[d] + 1;
  }
  if (ns * n > 0x7fffffffLL) return JMD_ERR_INVALID;
  G.n_shift = (int)ns;
  // zero shift: the centre of the stencil
  int zi = 0;
  for (int d = 0; d < dim; ++d) zi = zi * (2 * G.nmax[d] + 1) + G.nmax[d];
  G.zero = zi;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) G.T[a][b] = (a < dim && b < dim) ? box[a * dim + b] : 0.0f;
  hipStream_t st = (hipStream_t)stream;
  char* ws = (char*)workspace;
  long long* total = (long long*)ws;                // 8 B
  float4* xr = (float4*)(ws + 256);
  int* row_count = (int*)(ws + 256 + (size_t)n * 16);
  int* row_offset = row_count + (size_t)n * ns;
  const int nn = (int)n;
  hipLaunchKernelGGL(k_mi_real, dim3(ceil_div(n, 256)), dim3(256), 0, st, position, nn, dim,
                     fractional ? 1 : 0, G, xr);
  JMD_LAUNCH_CHECK();
  if (format == JMD_DENSE) {
    int grid = nn < 4096 ? nn : 4096;
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, occupancy, 1, 0ull);
    JMD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_mi_dense<0>, dim3(grid), dim3(256), 0, st, xr, nn, G, capacity, idx,
                       edge_shift, occupancy, overflow);
    JMD_LAUNCH_CHECK();
    if (measure) return JMD_OK;
    hipLaunchKernelGGL(k_fill_u8, dim3(1), dim3(64), 0, st, overflow, 1, (uint8_t)0);
    JMD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_mi_dense<1>, dim3(grid), dim3(256), 0, st, xr, nn, G, capacity, idx,
                       edge_shift, occupancy, overflow);
    JMD_LAUNCH_CHECK();
    return JMD_OK;
  }
  const int ordered = format == JMD_ORDERED_SPARSE ? 1 : 0;
  const long long nrow = ns * n;
  long long gb = (nrow * kWave + 255) / 256;
  if (gb > 8192) gb = 8192;
  hipLaunchKernelGGL(k_mi_sparse<0>, dim3((unsigned)gb), dim3(256), 0, st, xr, nn, G, ordered,
                     row_count, row_offset, capacity, idx, edge_shift);
  JMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mi_scan, dim3(1), dim3(1024), 0, st, row_count, row_offset, nrow, total);
  JMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_copy_i64_u64, dim3(1), dim3(64), 0, st, total, occupancy);
  JMD_LAUNCH_CHECK();
  if (measure) return JMD_OK;
  hipLaunchKernelGGL(k_mi_sparse<1>, dim3((unsigned)gb), dim3(256), 0, st, xr, nn, G, ordered,
                     row_count, row_offset, capacity, idx, edge_shift);
  JMD_LAUNCH_CHECK();
  long long pb = (capacity + 255) / 256;
  if (pb < 1) pb = 1;
  if (pb > 4096) pb = 4096;
  hipLaunchKernelGGL(k_mi_pad, dim3((unsigned)pb), dim3(256), 0, st, idx, edge_shift, capacity,
                     nn, dim, total, overflow);
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

int jmd_multi_image_max_displacement_sq(const float* position, const float* reference,
                                        int64_t n, int32_t dim, int32_t fractional,
                                        const float* box, const float* reference_box,
                                        float* out, void* stream) {
  if ((dim != 2 && dim != 3) || n < 1 || n > 0x7fffffff || !position || !reference || !out ||
      (fractional && (!box || !reference_box)))
    return JMD_ERR_INVALID;
  MIGeom G{}, Gr{};
  G.dim = Gr.dim = dim;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      G.T[a][b] = (fractional && a < dim && b < dim) ? box[a * dim + b] : 0.0f;
      Gr.T[a][b] = (fractional && a < dim && b < dim) ? reference_box[a * dim + b] : 0.0f;
    }
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_fill_u8, dim3(1), dim3(64), 0, st, (uint8_t*)out, 4, (uint8_t)0);
  JMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mi_maxdisp, dim3(ceil_div(n, 256)), dim3(256), 0, st, position, reference,
                     (int)n, dim, fractional ? 1 : 0, G, Gr, (unsigned*)out);
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

}  // extern "C"
