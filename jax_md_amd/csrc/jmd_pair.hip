// Pair reduction over a neighbour list for gfx950: energy.lennard_jones and
// energy.soft_sphere mapped by smap.pair_neighbor_list
// (jax_md/smap.py:856-979, jax_md/energy.py:125-173,246-343,534-580), and its
// gradient, which the reference obtains by autodiff (quantity.py:58-60).
//
// Dense lists hold both orders of every pair, so the force on particle i is
// the full sum over its own row: a group of TPA lanes owns row i, gathers
// the neighbour positions (the [N, dim] float32 table stays L2-resident),
// accumulates in registers and reduces across the group with cross-lane
// shuffles -- no atomics, deterministic.  Pair arithmetic is float32, the
// per-particle accumulation float64; the system energy is reduced in float64
// per block and then over blocks in a fixed order.
//
// The LJ force and energy are evaluated from r^2 without a square root:
//   V = 4 eps (s^12 - s^6),  V'/r = -24 eps (2 s^12 - s^6) / r^2,
//   S = (rc^2 - r^2)^2 (rc^2 + 2 r^2 - 3 ro^2) / (rc^2 - ro^2)^3,
//   S'/r = 12 (rc^2 - r^2)(ro^2 - r^2) / (rc^2 - ro^2)^3   on [r_on, r_c),
//   F_i += (S V'/r + S'/r V) (R_j - R_i).
#include <mutex>

#include "jmd_pair_dense.hpp"

namespace jmd {

JMD_DENSE_INSTANCES(JMD_DENSE_EXTERN)

// Sparse / OrderedSparse: one lane per stored pair (smap.py:926-931,975-977).
// Lists are sender-major (partition.py:990-1008: row i of the candidate
// array becomes the pairs with idx[1] = i), so a wavefront holds one or two
// runs of equal sender ids: the sender's share is summed over each run with a
// segmented inclusive scan and added by the run's last lane (lanes adding to
// one address would be serialised).  The receivers of a chunk of pairs are
// neighbours of a few consecutive senders, so the chunk touches far fewer
// distinct receivers than it holds pairs: each workgroup sums the receiver
// shares (and per-particle energies, summed over idx[0]) of its chunk in an
// LDS hash table and adds each entry to memory once.  A receiver that finds
// no free slot within kProbe probes goes to memory directly.  Any pair order
// is correct; ordered lists only make the sums shorter.
constexpr int kSparseChunk = JMD_SPARSE_CHUNK;  // pairs per workgroup pass
constexpr int kSparseTable = JMD_SPARSE_TABLE;  // LDS hash entries (power of two)
constexpr int kProbe = 16;

template <int DIM, int STRIDE, int KIND, bool GENERAL, bool ENERGY>
__global__ __launch_bounds__(256) void k_pair_sparse(
    const float* __restrict__ pos, int n, Space s, PairArgs a, const int* __restrict__ idx,
    long long width, float inv_norm, float* __restrict__ force, float* __restrict__ e_atom,
    double* __restrict__ e_part) {
  constexpr int NT = DIM + (ENERGY ? 1 : 0);  // table values per entry
  constexpr int kEmpty = -1;
  __shared__ int s_key[kSparseTable];
  __shared__ float s_val[NT][kSparseTable];
  const int lane = lane_id();
  const int tid = threadIdx.x;
  const bool table_on = force || (ENERGY && e_atom);
  for (int t = tid; t < kSparseTable; t += blockDim.x) {
    s_key[t] = kEmpty;
#pragma unroll
    for (int c = 0; c < NT; ++c) s_val[c][t] = 0.0f;
  }
  __syncthreads();
  double e = 0.0;
  for (long long base = (long long)blockIdx.x * kSparseChunk; base < width;
       base += (long long)gridDim.x * kSparseChunk) {
    // every lane of a wavefront runs the same iterations (for the shuffles)
    for (int r = 0; r < kSparseChunk; r += blockDim.x) {
      const long long k = base + r + tid;
      int ra = -1, sb = -1;
      if (k < width) {
        ra = idx[k];          // receiver (idx[0])
        sb = idx[width + k];  // sender   (idx[1])
      }
      const bool ok = (unsigned)ra < (unsigned)n && (unsigned)sb < (unsigned)n;
      float v[DIM], w[NT];
#pragma unroll
      for (int c = 0; c < DIM; ++c) v[c] = 0.0f;
#pragma unroll
      for (int c = 0; c < NT; ++c) w[c] = 0.0f;
      if (ok) {
        float xa[3] = {0, 0, 0}, xb[3] = {0, 0, 0}, d[3];  // disp(R[idx0], R[idx1]) = R_a - R_b
        load_pos<DIM, STRIDE>(pos, ra, xa);
        load_pos<DIM, STRIDE>(pos, sb, xb);
#pragma unroll
        for (int c = 0; c < DIM; ++c)
          d[c] = s.periodic ? min_image_any(xa[c], xb[c], xb[c] - s.L[c], s.L[c], s.half[c],
                                            s.inbox)
                            : xa[c] - xb[c];
        if (s.general) box_apply<DIM>(s, d);
        float r2 = DIM == 3 ? fmaf(d[2], d[2], fmaf(d[1], d[1], d[0] * d[0]))
                            : fmaf(d[1], d[1], d[0] * d[0]);
        float sig = a.sig, eps = a.eps, alp = a.alp;
        if (GENERAL) {
          int s0 = a.species ? a.species[ra] : 0, s1 = a.species ? a.species[sb] : 0;
          sig = pick(a.sig_mode, a.sig, a.sig_v, ra, sb, s0, s1, a.ns);
          eps = pick(a.eps_mode, a.eps, a.eps_v, ra, sb, s0, s1, a.ns);
          if (KIND == JMD_SOFT_SPHERE) alp = pick(a.alp_mode, a.alp, a.alp_v, ra, sb, s0, s1, a.ns);
        }
        float dEr;
        float u = pair_eval<KIND>(r2, sig, eps, alp, a, &dEr) * inv_norm;
        dEr *= inv_norm;
        if (force) {
#pragma unroll
          for (int c = 0; c < DIM; ++c) {
            const float fc = dEr * d[c];  // dE/dR_a = dEr * (R_a - R_b)
            v[c] = fc;
            w[c] = -fc;
          }
        }
        if (ENERGY) {
          e += (double)u;
          w[NT - 1] = u;
        }
      }
      if (!table_on) continue;
      // receiver share into the table
      if (ok) {
        unsigned slot = ((unsigned)ra * 2654435761u) >> (32 - __builtin_ctz(kSparseTable));
        bool placed = false;
        for (int q = 0; q < kProbe; ++q, slot = (slot + 1) & (kSparseTable - 1)) {
          int cur = __hip_atomic_load(&s_key[slot], __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_WORKGROUP);
          if (cur == kEmpty) cur = atomicCAS(&s_key[slot], kEmpty, ra);
          if (cur == kEmpty || cur == ra) {
            placed = true;
            break;
          }
        }
        if (placed) {
#pragma unroll
          for (int c = 0; c < NT; ++c) atomicAdd(&s_val[c][slot], w[c]);
        } else {
          if (force) {
#pragma unroll
            for (int c = 0; c < DIM; ++c) atomicAdd(&force[(long long)ra * DIM + c], w[c]);
          }
          if (ENERGY && e_atom) atomicAdd(&e_atom[ra], w[NT - 1]);
        }
      }
      if (!force) continue;
      // sender share: segmented scan over runs of equal sender ids
      const int key = ok ? sb : -1;
      const int prev = __shfl_up(key, 1, kWave);
      const unsigned long long heads = __ballot(lane == 0 || prev != key);
      const int h = 63 - __clzll(heads & ((2ull << lane) - 1ull));
#pragma unroll
      for (int off = 1; off < kWave; off <<= 1) {
        float t[DIM];
#pragma unroll
        for (int c = 0; c < DIM; ++c) t[c] = __shfl_up(v[c], off, kWave);
        if (lane - off >= h) {
#pragma unroll
          for (int c = 0; c < DIM; ++c) v[c] += t[c];
        }
      }
      const bool tail = lane == kWave - 1 || ((heads >> (lane + 1)) & 1ull);
      if (tail && ok) {
#pragma unroll
        for (int c = 0; c < DIM; ++c) atomicAdd(&force[(long long)sb * DIM + c], v[c]);
      }
    }
    if (!table_on) continue;
    // flush the table and clear it for the next chunk
    __syncthreads();
    for (int t = tid; t < kSparseTable; t += blockDim.x) {
      const int key = s_key[t];
      if (key != kEmpty) {
        if (force) {
#pragma unroll
          for (int c = 0; c < DIM; ++c) atomicAdd(&force[(long long)key * DIM + c], s_val[c][t]);
        }
        if (ENERGY && e_atom) atomicAdd(&e_atom[key], s_val[NT - 1][t]);
        s_key[t] = kEmpty;
#pragma unroll
        for (int c = 0; c < NT; ++c) s_val[c][t] = 0.0f;
      }
    }
    __syncthreads();
  }
  if (ENERGY) {
    __shared__ double s_e[256 / kWave];
    double wsum = wave_reduce_sum(e);
    if (lane_id() == 0) s_e[threadIdx.x / kWave] = wsum;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int q = 0; q < (int)(blockDim.x / kWave); ++q) t += s_e[q];
      e_part[blockIdx.x] = t;
    }
  }
}

// Closing half kick after a Sparse-format force: p += half_dt F (the same
// expression the Dense kernel fuses into its epilogue).
__global__ __launch_bounds__(256) void k_kick_after(float* __restrict__ mom,
                                                    const float* __restrict__ force, long long m,
                                                    float half_dt,
                                                    const float* __restrict__ dt_dev) {
  if (dt_dev) half_dt = dt_dev[0] / 2.0f;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < m;
       q += (long long)gridDim.x * blockDim.x)
    mom[q] += half_dt * force[q];
}

// Fixed-order sum of block partials (one workgroup).
__global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ part, int m,
                                                      double* __restrict__ out) {
  __shared__ double s[256];
  double v = 0.0;
  for (int k = threadIdx.x; k < m; k += 256) v += part[k];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0];
}

static int g_tpa_override = 0;
constexpr long long kSparseBlocks = 2048;

// Virial W_ab = weight sum over the list's ordered pairs of (U'(r)/r) d_a d_b
// (the box derivative quantity.stress / pressure take by autodiff,
// quantity.py:202-282): float32 pair terms, float64 accumulation, block
// partials [6 per block] (xx, yy, zz, xy, xz, yz; 2-D uses xx, yy, xy).
// Dense: 8 lanes per row; Sparse / OrderedSparse: one thread per entry.
constexpr int kVirBlocks = 2048;

template <int DIM, int STRIDE, int KIND, bool GENERAL, bool DENSE>
__global__ __launch_bounds__(256) void k_pair_virial(const float* __restrict__ pos, int n,
                                                     Space s, PairArgs a,
                                                     const int* __restrict__ idx,
                                                     long long width, double weight,
                                                     double* __restrict__ part) {
#pragma clang fp contract(off)
  double w[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  auto add = [&](unsigned ii, unsigned jj) {
    float xi[3] = {0, 0, 0}, xj[3] = {0, 0, 0}, d[3] = {0, 0, 0};
    load_pos<DIM, STRIDE>(pos, ii, xi);
    load_pos<DIM, STRIDE>(pos, jj, xj);
#pragma unroll
    for (int c = 0; c < DIM; ++c)
      d[c] = s.periodic ? min_image_any(xj[c], xi[c], xi[c] - s.L[c], s.L[c], s.half[c], s.inbox)
                        : xj[c] - xi[c];
    if (s.general) box_apply<DIM>(s, d);
    const float r2 = DIM == 3 ? fmaf(d[2], d[2], fmaf(d[1], d[1], d[0] * d[0]))
                              : fmaf(d[1], d[1], d[0] * d[0]);
    float sig = a.sig, eps = a.eps, alp = a.alp;
    if (GENERAL) {
      const int si = a.species ? a.species[ii] : 0, sj = a.species ? a.species[jj] : 0;
      sig = pick(a.sig_mode, a.sig, a.sig_v, ii, jj, si, sj, a.ns);
      eps = pick(a.eps_mode, a.eps, a.eps_v, ii, jj, si, sj, a.ns);
      if (KIND == JMD_SOFT_SPHERE) alp = pick(a.alp_mode, a.alp, a.alp_v, ii, jj, si, sj, a.ns);
    }
    float dEr;
    pair_eval<KIND>(r2, sig, eps, alp, a, &dEr);
    const double g = (double)dEr;
    if (DIM == 3) {
      w[0] += g * ((double)d[0] * d[0]);
      w[1] += g * ((double)d[1] * d[1]);
      w[2] += g * ((double)d[2] * d[2]);
      w[3] += g * ((double)d[0] * d[1]);
      w[4] += g * ((double)d[0] * d[2]);
      w[5] += g * ((double)d[1] * d[2]);
    } else {
      w[0] += g * ((double)d[0] * d[0]);
      w[1] += g * ((double)d[1] * d[1]);
      w[3] += g * ((double)d[0] * d[1]);
    }
  };
  const long long nth = (long long)gridDim.x * blockDim.x;
  const long long t0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (DENSE) {
    for (long long t = t0; t < (long long)n * 8; t += nth) {
      const unsigned i = (unsigned)(t / 8), sub = (unsigned)(t % 8);
      const int* row = idx + (long long)i * width;
      for (long long k = sub; k < width; k += 8) {
        const unsigned j = (unsigned)row[k];
        if (j >= (unsigned)n) break;  // left-compacted rows, padding n
        add(i, j);
      }
    }
  } else {
    for (long long k = t0; k < width; k += nth) {
      const unsigned ra = (unsigned)idx[k], sb = (unsigned)idx[width + k];
      if (ra < (unsigned)n && sb < (unsigned)n) add(ra, sb);
    }
  }
  __shared__ double s_w[6][256 / kWave];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const double v = wave_reduce_sum(w[c]);
    if (lane_id() == 0) s_w[c][threadIdx.x / kWave] = v;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    double t = 0.0;
    for (int q = 0; q < (int)(blockDim.x / kWave); ++q) t += s_w[threadIdx.x][q];
    part[6LL * blockIdx.x + threadIdx.x] = weight * t;
  }
}

// out[dim x dim] (row-major, symmetric) from the block partials, fixed order.
__global__ __launch_bounds__(256) void k_virial_sum(const double* __restrict__ part, int nb,
                                                    int dim, double* __restrict__ out) {
  __shared__ double s[6][256];
  for (int c = 0; c < 6; ++c) {
    double v = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) v += part[6LL * b + c];
    s[c][threadIdx.x] = v;
  }
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off)
      for (int c = 0; c < 6; ++c) s[c][threadIdx.x] += s[c][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (dim == 3) {
      const double m[9] = {s[0][0], s[3][0], s[4][0], s[3][0], s[1][0],
                           s[5][0], s[4][0], s[5][0], s[2][0]};
      for (int k = 0; k < 9; ++k) out[k] = m[k];
    } else {
      out[0] = s[0][0];
      out[1] = s[3][0];
      out[2] = s[3][0];
      out[3] = s[1][0];
    }
  }
}

// Per-entry pair energies (reduce_axis=() of smap.pair_neighbor_list,
// smap.py:940-953): one thread per list entry, U(r_ij) * scale for a valid
// entry, 0 for padding.  Dense: entry t = i * width + k; Sparse: entry k of
// the [2, width] list (receiver idx[k], sender idx[width + k]).
template <int DIM, int STRIDE, int KIND, bool GENERAL, bool DENSE>
__global__ __launch_bounds__(256) void k_pair_edges(const float* __restrict__ pos, int n,
                                                    Space s, PairArgs a,
                                                    const int* __restrict__ idx,
                                                    long long width, float scale,
                                                    float* __restrict__ out) {
#pragma clang fp contract(off)
  const long long total = DENSE ? (long long)n * width : width;
  const long long nth = (long long)gridDim.x * blockDim.x;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += nth) {
    unsigned ii, jj;
    if (DENSE) {
      ii = (unsigned)(t / width);
      jj = (unsigned)idx[t];
    } else {
      ii = (unsigned)idx[t];
      jj = (unsigned)idx[width + t];
    }
    float u = 0.0f;
    if (ii < (unsigned)n && jj < (unsigned)n) {
      float xi[3] = {0, 0, 0}, xj[3] = {0, 0, 0}, d[3] = {0, 0, 0};
      load_pos<DIM, STRIDE>(pos, ii, xi);
      load_pos<DIM, STRIDE>(pos, jj, xj);
#pragma unroll
      for (int c = 0; c < DIM; ++c)
        d[c] = s.periodic ? min_image_any(xj[c], xi[c], xi[c] - s.L[c], s.L[c], s.half[c],
                                          s.inbox)
                          : xj[c] - xi[c];
      if (s.general) box_apply<DIM>(s, d);
      const float r2 = DIM == 3 ? fmaf(d[2], d[2], fmaf(d[1], d[1], d[0] * d[0]))
                                : fmaf(d[1], d[1], d[0] * d[0]);
      float sig = a.sig, eps = a.eps, alp = a.alp;
      if (GENERAL) {
        const int si = a.species ? a.species[ii] : 0, sj = a.species ? a.species[jj] : 0;
        sig = pick(a.sig_mode, a.sig, a.sig_v, ii, jj, si, sj, a.ns);
        eps = pick(a.eps_mode, a.eps, a.eps_v, ii, jj, si, sj, a.ns);
        if (KIND == JMD_SOFT_SPHERE) alp = pick(a.alp_mode, a.alp, a.alp_v, ii, jj, si, sj, a.ns);
      }
      float dEr;
      u = pair_eval<KIND>(r2, sig, eps, alp, a, &dEr) * scale;
    }
    out[t] = u;
  }
}

}  // namespace jmd

using namespace jmd;

extern "C" {

size_t jmd_pair_workspace_bytes(int64_t n) {
  // block partials: Dense uses at most ceil(n*16/256) padded to 8, Sparse at
  // most kSparseBlocks.
  long long blocks = (n * 16 + 255) / 256 + 64;
  if (blocks < kSparseBlocks) blocks = kSparseBlocks;
  return (size_t)(blocks * 8 + 256);
}

// Threads per Dense row (2, 4, 8 or 16); 0 restores the built-in choice.
int jmd_set_threads_per_atom(int32_t tpa) {
  if (tpa != 0 && tpa != 2 && tpa != 4 && tpa != 8 && tpa != 16) return JMD_ERR_INVALID;
  g_tpa_override = tpa;
  return JMD_OK;
}

}  // extern "C"

namespace {

int pair_energy_force(const float* position, int64_t n, int32_t dim, int32_t position_stride,
                      int32_t periodic, const float* side, const jmd_pair_potential* potential,
                      const int32_t* idx, int64_t idx_width, int32_t format, float* force,
                      float* energy_atom, double* energy_total, float* momentum, float half_dt,
                      const float* dt_device, void* workspace, void* stream) {
  if ((dim != 2 && dim != 3) || n < 0 || n > 0x7fffffff || idx_width < 0 ||
      (n > 0 && !position) || (idx_width > 0 && !idx) || (periodic && !side))
    return JMD_ERR_INVALID;
  if (n > 0 && !stride_ok(dim, position_stride, position)) return JMD_ERR_INVALID;
  if (format != JMD_DENSE && format != JMD_SPARSE && format != JMD_ORDERED_SPARSE)
    return JMD_ERR_INVALID;
  const bool energy = energy_total != nullptr || energy_atom != nullptr;
  if (energy && !workspace) return JMD_ERR_INVALID;
  if (format == JMD_ORDERED_SPARSE && energy_atom) return JMD_ERR_UNSUPPORTED;
  if (momentum && !force) return JMD_ERR_INVALID;
  PairArgs a;
  if (!make_pair_args(potential, &a)) return JMD_ERR_INVALID;
  {
    const int grc = space_check(dim, periodic, side, true);
    if (grc) return grc;
  }
  Space s = make_space(dim, periodic, side);
  hipStream_t st = (hipStream_t)stream;
  int nn = (int)n;
  double* part = (double*)workspace;
  const bool s4 = dim == 3 && position_stride == 4;
  if (format == JMD_DENSE) {
    if (idx_width > 0x7fffffff) return JMD_ERR_INVALID;
    if (nn == 0) {
      if (energy_total) {
        hipLaunchKernelGGL(fill_kernel<double>, dim3(1), dim3(64), 0, st, energy_total, 1LL, 0.0);
        JMD_LAUNCH_CHECK();
      }
      return JMD_OK;
    }
    int tpa = g_tpa_override ? g_tpa_override : (dim == 3 ? 8 : 2);
    long long threads = (long long)nn * tpa;
    int nb = ceil_div(threads, 256);
    nb = (nb + 7) & ~7;
    DenseLaunch L{dim3(nb), st, position, nn, s, a, idx, (int)idx_width, force, energy_atom,
                  energy ? part : nullptr, momentum, half_dt, dt_device, KickDrift{}};
    const bool kick = momentum != nullptr;
    if (dim == 3) {
      if (a.kind == JMD_LENNARD_JONES) {
        if (s4) dispatch_dense<3, 4, JMD_LENNARD_JONES>(tpa, a.general, L, energy, kick, false);
        else dispatch_dense<3, 3, JMD_LENNARD_JONES>(tpa, a.general, L, energy, kick, false);
      } else {
        if (s4) dispatch_dense<3, 4, JMD_SOFT_SPHERE>(tpa, a.general, L, energy, kick, false);
        else dispatch_dense<3, 3, JMD_SOFT_SPHERE>(tpa, a.general, L, energy, kick, false);
      }
    } else {
      if (a.kind == JMD_LENNARD_JONES)
        dispatch_dense<2, 2, JMD_LENNARD_JONES>(tpa, a.general, L, energy, kick, false);
      else
        dispatch_dense<2, 2, JMD_SOFT_SPHERE>(tpa, a.general, L, energy, kick, false);
    }
    JMD_LAUNCH_CHECK();
    if (energy_total) {
      hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, part, nb, energy_total);
      JMD_LAUNCH_CHECK();
    }
    return JMD_OK;
  }
  // Sparse / OrderedSparse
  float inv_norm = format == JMD_ORDERED_SPARSE ? 1.0f : 0.5f;
  if (force) {
    int rc = launch_fill<float>(force, (long long)nn * dim, 0.0f, st);
    if (rc) return rc;
  }
  if (energy_atom) {
    int rc = launch_fill<float>(energy_atom, (long long)nn, 0.0f, st);
    if (rc) return rc;
  }
  long long nb = (idx_width + kSparseChunk - 1) / kSparseChunk;
  if (nb == 0) nb = 1;
  if (nb > kSparseBlocks) nb = kSparseBlocks;
#define JMD_PS(D, S, K)                                                                        \
  do {                                                                                         \
    if (energy)                                                                                \
      hipLaunchKernelGGL((k_pair_sparse<D, S, K, true, true>), dim3(nb), dim3(256), 0, st,     \
                         position, nn, s, a, idx, idx_width, inv_norm, force, energy_atom,     \
                         part);                                                                \
    else                                                                                       \
      hipLaunchKernelGGL((k_pair_sparse<D, S, K, true, false>), dim3(nb), dim3(256), 0, st,    \
                         position, nn, s, a, idx, idx_width, inv_norm, force, energy_atom,     \
                         part);                                                                \
  } while (0)
  if (dim == 3) {
    if (a.kind == JMD_LENNARD_JONES) {
      if (s4) JMD_PS(3, 4, JMD_LENNARD_JONES); else JMD_PS(3, 3, JMD_LENNARD_JONES);
    } else {
      if (s4) JMD_PS(3, 4, JMD_SOFT_SPHERE); else JMD_PS(3, 3, JMD_SOFT_SPHERE);
    }
  } else {
    if (a.kind == JMD_LENNARD_JONES) JMD_PS(2, 2, JMD_LENNARD_JONES);
    else JMD_PS(2, 2, JMD_SOFT_SPHERE);
  }
#undef JMD_PS
  JMD_LAUNCH_CHECK();
  if (momentum && nn > 0) {
    const long long m = (long long)nn * dim;
    long long kb = (m + 255) / 256;
    if (kb > 4096) kb = 4096;
    hipLaunchKernelGGL(k_kick_after, dim3((unsigned)kb), dim3(256), 0, st, momentum, force, m,
                       half_dt, dt_device);
    JMD_LAUNCH_CHECK();
  }
  if (energy_total) {
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, part, (int)nb, energy_total);
    JMD_LAUNCH_CHECK();
  }
  return JMD_OK;
}

}  // namespace

// Scratch of jmd_pair_energy_force_noalloc: the float64 block partials of
// the energy sum, jmd_pair_workspace_bytes(n) bytes for n up to
// JMD_PAIR_NOALLOC_MAX_N.  Each device holds its own copy for the life of
// the process; host code reaches it only through the address that
// hipGetSymbolAddress returns on that device.
constexpr size_t kNoAllocScratchBytes = (size_t)1 << 20;
static_assert((size_t)((JMD_PAIR_NOALLOC_MAX_N * 16LL + 255) / 256 + 64) * 8 + 256 <=
                  kNoAllocScratchBytes,
              "scratch too small for JMD_PAIR_NOALLOC_MAX_N");
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kScratchDevices];

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
int scratch_address(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kScratchDevices)
    return JMD_ERR_LAUNCH;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return JMD_ERR_LAUNCH;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return JMD_OK;
}

}  // namespace

extern "C" {

int jmd_pair_energy_force(const float* position, int64_t n, int32_t dim,
                          int32_t position_stride, int32_t periodic, const float* side,
                          const jmd_pair_potential* potential, const int32_t* idx,
                          int64_t idx_width, int32_t format, float* force, float* energy_atom,
                          double* energy_total, float* momentum, float half_dt,
                          const float* dt_device, void* workspace, void* stream) {
  return pair_energy_force(position, n, dim, position_stride, periodic, side, potential, idx,
                           idx_width, format, force, energy_atom, energy_total, momentum,
                           half_dt, dt_device, workspace, stream);
}

int jmd_pair_energy_force_noalloc(const float* position, int64_t n, int32_t dim,
                                  int32_t position_stride, int32_t periodic, const float* side,
                                  const jmd_pair_potential* potential, const int32_t* idx,
                                  int64_t idx_width, int32_t format, float* force,
                                  float* energy_atom, double* energy_total, float* momentum,
                                  float half_dt, const float* dt_device, void* stream) {
  if (n < 0 || n > JMD_PAIR_NOALLOC_MAX_N) return JMD_ERR_INVALID;
  void* ws = nullptr;
  const int rc = scratch_address(&ws);
  if (rc) return rc;
  // every launch writes each block partial it later sums, so nothing in the
  // scratch needs clearing between calls
  return pair_energy_force(position, n, dim, position_stride, periodic, side, potential, idx,
                           idx_width, format, force, energy_atom, energy_total, momentum,
                           half_dt, dt_device, ws, stream);
}

int jmd_nve_force_kick_drift(const float* position, int64_t n, int32_t dim,
                             int32_t position_stride, int32_t periodic, const float* side,
                             const jmd_pair_potential* potential, const int32_t* idx,
                             int64_t idx_width, float* force, float* momentum, float half_dt,
                             const float* dt_device, double* position_hi, const float* mass,
                             float mass_scalar, int32_t shift_periodic, const float* shift_side,
                             float dt, const float* reference_position, float threshold_sq,
                             int32_t* gate_next, int32_t drift, float* position_out,
                             void* stream) {
  if ((dim != 2 && dim != 3) || n < 0 || n > 0x7fffffff || idx_width < 0 ||
      idx_width > 0x7fffffff || (n > 0 && (!position || !force || !momentum)) ||
      (idx_width > 0 && !idx) || (periodic && !side) || (shift_periodic && !shift_side) ||
      (gate_next && !reference_position))
    return JMD_ERR_INVALID;
  if (drift && (!position_hi || !position_out || position_out == position)) return JMD_ERR_INVALID;
  if (!mass && !(mass_scalar > 0.0f)) return JMD_ERR_INVALID;
  if (n == 0) return JMD_OK;
  if (!stride_ok(dim, position_stride, position)) return JMD_ERR_INVALID;
  if (drift && !stride_ok(dim, position_stride, position_out)) return JMD_ERR_INVALID;
  PairArgs a;
  if (!make_pair_args(potential, &a)) return JMD_ERR_INVALID;
  {
    int grc = space_check(dim, periodic, side, true);
    if (!grc) grc = space_check(dim, shift_periodic, shift_side, true);
    if (grc) return grc;
  }
  Space s = make_space(dim, periodic, side);
  KickDrift kd{};
  kd.active = drift ? 1 : 0;
  kd.xhi_in = position_hi;
  kd.xhi_out = position_hi;
  kd.mass = mass;
  kd.mass_scalar = mass_scalar;
  kick_drift_mass(&kd);
  kd.dt = dt;
  kd.half_dt = half_dt;
  kd.dt_dev = dt_device;
  kd.x_out = position_out;
  kd.ref = reference_position;
  kd.thr_sq = threshold_sq;
  kd.gate_next = gate_next;
  kd.s = make_space(dim, shift_periodic, shift_side);
  box_inverse(kd.s, kd.Ti);
  hipStream_t st = (hipStream_t)stream;
  const int nn = (int)n;
  const int tpa = g_tpa_override ? g_tpa_override : (dim == 3 ? 8 : 2);
  int nb = ceil_div((long long)nn * tpa, 256);
  nb = (nb + 7) & ~7;
  DenseLaunch L{dim3(nb), st, position, nn, s, a, idx, (int)idx_width, force, nullptr,
                nullptr, momentum, half_dt, dt_device, kd};
  const bool s4 = dim == 3 && position_stride == 4;
  if (dim == 3) {
    if (a.kind == JMD_LENNARD_JONES) {
      if (s4) dispatch_dense<3, 4, JMD_LENNARD_JONES>(tpa, a.general, L, false, true, true);
      else dispatch_dense<3, 3, JMD_LENNARD_JONES>(tpa, a.general, L, false, true, true);
    } else {
      if (s4) dispatch_dense<3, 4, JMD_SOFT_SPHERE>(tpa, a.general, L, false, true, true);
      else dispatch_dense<3, 3, JMD_SOFT_SPHERE>(tpa, a.general, L, false, true, true);
    }
  } else {
    if (a.kind == JMD_LENNARD_JONES)
      dispatch_dense<2, 2, JMD_LENNARD_JONES>(tpa, a.general, L, false, true, true);
    else
      dispatch_dense<2, 2, JMD_SOFT_SPHERE>(tpa, a.general, L, false, true, true);
  }
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

int jmd_pair_edge_energies(const float* position, int64_t n, int32_t dim,
                           int32_t position_stride, int32_t periodic, const float* side,
                           const jmd_pair_potential* potential, const int32_t* idx,
                           int64_t idx_width, int32_t format, float* out, void* stream) {
  if ((dim != 2 && dim != 3) || n < 0 || n > 0x7fffffff || idx_width < 0 ||
      (n > 0 && !position) || (idx_width > 0 && (!idx || !out)) || (periodic && !side))
    return JMD_ERR_INVALID;
  if (format != JMD_DENSE && format != JMD_SPARSE && format != JMD_ORDERED_SPARSE)
    return JMD_ERR_INVALID;
  if (n > 0 && !stride_ok(dim, position_stride, position)) return JMD_ERR_INVALID;
  PairArgs a;
  if (!make_pair_args(potential, &a)) return JMD_ERR_INVALID;
  {
    const int grc = space_check(dim, periodic, side, true);
    if (grc) return grc;
  }
  const bool dense = format == JMD_DENSE;
  const long long total = dense ? n * idx_width : idx_width;
  if (total == 0) return JMD_OK;
  Space s = make_space(dim, periodic, side);
  hipStream_t st = (hipStream_t)stream;
  // the reference divides by its normalization: 2 for lists holding both
  // orders, 1 for OrderedSparse
  const float scale = format == JMD_ORDERED_SPARSE ? 1.0f : 0.5f;
  long long nb = (total + 255) / 256;
  if (nb > 65536) nb = 65536;
  const int nn = (int)n;
#define JMD_EDGE(D, S, K)                                                                     \
  do {                                                                                       \
    if (a.general) {                                                                         \
      if (dense) hipLaunchKernelGGL((k_pair_edges<D, S, K, true, true>), dim3(nb), dim3(256),  \
                                    0, st, position, nn, s, a, idx, idx_width, scale, out);  \
      else hipLaunchKernelGGL((k_pair_edges<D, S, K, true, false>), dim3(nb), dim3(256), 0,   \
                              st, position, nn, s, a, idx, idx_width, scale, out);           \
    } else {                                                                                 \
      if (dense) hipLaunchKernelGGL((k_pair_edges<D, S, K, false, true>), dim3(nb), dim3(256), \
                                    0, st, position, nn, s, a, idx, idx_width, scale, out);  \
      else hipLaunchKernelGGL((k_pair_edges<D, S, K, false, false>), dim3(nb), dim3(256), 0,  \
                              st, position, nn, s, a, idx, idx_width, scale, out);           \
    }                                                                                        \
  } while (0)
  if (dim == 3 && position_stride == 4) {
    if (a.kind == JMD_LENNARD_JONES) JMD_EDGE(3, 4, JMD_LENNARD_JONES);
    else JMD_EDGE(3, 4, JMD_SOFT_SPHERE);
  } else if (dim == 3) {
    if (a.kind == JMD_LENNARD_JONES) JMD_EDGE(3, 3, JMD_LENNARD_JONES);
    else JMD_EDGE(3, 3, JMD_SOFT_SPHERE);
  } else {
    if (a.kind == JMD_LENNARD_JONES) JMD_EDGE(2, 2, JMD_LENNARD_JONES);
    else JMD_EDGE(2, 2, JMD_SOFT_SPHERE);
  }
#undef JMD_EDGE
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

size_t jmd_pair_virial_workspace_bytes(void) { return (size_t)kVirBlocks * 6 * sizeof(double) + 256; }

int jmd_pair_virial(const float* position, int64_t n, int32_t dim, int32_t position_stride,
                    int32_t periodic, const float* side, const jmd_pair_potential* potential,
                    const int32_t* idx, int64_t idx_width, int32_t format, double* virial,
                    void* workspace, void* stream) {
  if ((dim != 2 && dim != 3) || n < 0 || n > 0x7fffffff || idx_width < 0 || !virial ||
      !workspace || (n > 0 && !position) || (idx_width > 0 && !idx) || (periodic && !side))
    return JMD_ERR_INVALID;
  if (format != JMD_DENSE && format != JMD_SPARSE && format != JMD_ORDERED_SPARSE)
    return JMD_ERR_INVALID;
  if (n > 0 && !stride_ok(dim, position_stride, position)) return JMD_ERR_INVALID;
  PairArgs a;
  if (!make_pair_args(potential, &a)) return JMD_ERR_INVALID;
  {
    const int grc = space_check(dim, periodic, side, true);
    if (grc) return grc;
  }
  Space s = make_space(dim, periodic, side);
  hipStream_t st = (hipStream_t)stream;
  const bool dense = format == JMD_DENSE;
  const long long work = dense ? (long long)n * 8 : idx_width;
  long long nb = (work + 255) / 256;
  if (nb > kVirBlocks) nb = kVirBlocks;
  if (nb < 1) nb = 1;
  const double weight = format == JMD_ORDERED_SPARSE ? 1.0 : 0.5;
  double* part = (double*)workspace;
  const int nn = (int)n;
#define JMD_VIR(D, S, K)                                                                      \
  do {                                                                                       \
    if (a.general) {                                                                         \
      if (dense) hipLaunchKernelGGL((k_pair_virial<D, S, K, true, true>), dim3(nb), dim3(256), \
                                    0, st, position, nn, s, a, idx, idx_width, weight, part); \
      else hipLaunchKernelGGL((k_pair_virial<D, S, K, true, false>), dim3(nb), dim3(256), 0,   \
                              st, position, nn, s, a, idx, idx_width, weight, part);          \
    } else {                                                                                 \
      if (dense) hipLaunchKernelGGL((k_pair_virial<D, S, K, false, true>), dim3(nb),          \
                                    dim3(256), 0, st, position, nn, s, a, idx, idx_width,     \
                                    weight, part);                                            \
      else hipLaunchKernelGGL((k_pair_virial<D, S, K, false, false>), dim3(nb), dim3(256), 0,  \
                              st, position, nn, s, a, idx, idx_width, weight, part);          \
    }                                                                                        \
  } while (0)
  if (dim == 3 && position_stride == 4) {
    if (a.kind == JMD_LENNARD_JONES) JMD_VIR(3, 4, JMD_LENNARD_JONES);
    else JMD_VIR(3, 4, JMD_SOFT_SPHERE);
  } else if (dim == 3) {
    if (a.kind == JMD_LENNARD_JONES) JMD_VIR(3, 3, JMD_LENNARD_JONES);
    else JMD_VIR(3, 3, JMD_SOFT_SPHERE);
  } else {
    if (a.kind == JMD_LENNARD_JONES) JMD_VIR(2, 2, JMD_LENNARD_JONES);
    else JMD_VIR(2, 2, JMD_SOFT_SPHERE);
  }
#undef JMD_VIR
  JMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_virial_sum, dim3(1), dim3(256), 0, st, part, (int)nb, dim, virial);
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

}  // extern "C"
