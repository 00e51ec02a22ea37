// Ragged ensembles of independent 2-D systems in one persistent launch.
//
// Each workgroup owns one system of a CSR (particles [offset, offset + n)) and
// runs the reference's minimisation loop for it,
//
//   for step:  nbrs = nbrs.update(R)            partition.py:1141-1164
//              state = fire_apply(state, nbrs)  minimize.py:170-226
//
// entirely inside the workgroup, separated by workgroup barriers: the skin
// test, the Dense list rebuild on the system's own cell grid (LDS counts, the
// reference's candidate order and both-direction test, partition.py:911-980),
// the velocity-Verlet sub-step with float64 positions (simulate.py:227-243),
// the soft-sphere forces (energy.py:125-173 through smap.py:856-979) with the
// lane grouping of the plain Dense kernel (two accumulators per row, summed in
// float64), and the FIRE reductions and float16 bookkeeping
// (minimize.py:185-221).  No host round trip and no grid-wide
// synchronisation: systems never interact.  Workgroups take systems in the
// order given (largest first balances the tail).
#include "jmd_pair_common.hpp"

namespace jmd {

// Threads per system: 1024 when the systems average 4096 particles or more
// (16 wavefronts hide the gathers' latency of a large system), else 512
// (small systems, two workgroups per CU).
constexpr int kEnsThreads = 1024;  // the most a workgroup has
static int g_ens_threads = 0;      // jmd_fire_ensemble_set_threads (0: by size)
inline int ens_threads(int n_systems, long long n_total) {
  if (g_ens_threads) return g_ens_threads;
  return n_systems > 0 && n_total / n_systems >= 4096 ? 1024 : 512;
}
constexpr int kEnsMaxCells = 8192;
#ifndef JMD_ENS_ROW_CHUNK
#define JMD_ENS_ROW_CHUNK 4
#endif
constexpr int kEnsRowChunk = JMD_ENS_ROW_CHUNK;  // row entries gathered together (even)
static_assert(kEnsRowChunk % 2 == 0, "entry parity is the chunk position's");
constexpr int kEnsParMax = 64;  // species matrices staged in LDS (n_species^2 at most)
constexpr int kEnsPerThread = 8;  // particles per thread of a system staged in LDS
static_assert(kEnsThreads * kEnsPerThread <= (1 << 16), "staged ids are 16-bit");

struct EnsSys {
  int offset, n;
  int cps[2];
  float cell_size[2];
  int fill_cps[2];
  float fill_cell_size[2];
  float side[2];
  int cap, max_occ;
  float cutoff_sq, threshold_sq;  // list cutoff^2, (skin / 2)^2
  int param_offset;               // this system's block of species parameters
  int reserved;
};

// Per-step record of a forward segment for the adjoint (all system-major by
// particle, node k = state after k steps of the segment, k = 0..K):
//   x[k], F[k]: float32 positions and forces at node k ([K+1][ntot][2])
//   q[k]: momentum of step k after the second half kick, before the FIRE
//         mixing ([K][ntot][2])
//   idx[k]: a Dense list slot ([K+1][ntot][width]); slot 0 holds node 0's
//         list, slot k + 1 the list a system rebuilt at step k (a system's
//         other slots are not written)
//   sc[k][s]: step k's scalars of system s: dt, alpha, keep, sqrt(pp),
//         sqrt(ff + 1e-6), and the slot of the list the force at node k + 1
//         was computed with  ([K][nsys][kTapeScalars])
struct EnsTape {
  float* x;
  float* F;
  float* q;
  int* idx;
  float* sc;
  int K;
};
constexpr int kTapeScalars = 6;

struct FireP {
  float dt_start, dt_max;
  int n_min;
  float f_inc, f_dec, alpha_start, f_alpha;
};

enum { ENS_FILL = 0, ENS_OCCUPANCY = 1, ENS_INIT = 2, ENS_RUN = 3 };

__device__ __forceinline__ float f16r(float x) { return (float)(_Float16)x; }

// Workgroup-wide sum of three doubles (all threads call; result in all).
__device__ __forceinline__ void wg_sum3(double& a, double& b, double& c, double* s) {
  a = wave_reduce_sum(a);
  b = wave_reduce_sum(b);
  c = wave_reduce_sum(c);
  const int w = threadIdx.x / kWave, nw = blockDim.x / kWave;
  __syncthreads();
  if (lane_id() == 0) {
    s[3 * w] = a;
    s[3 * w + 1] = b;
    s[3 * w + 2] = c;
  }
  __syncthreads();
  a = b = c = 0.0;
  for (int k = 0; k < nw; ++k) {  // fixed order
    a += s[3 * k];
    b += s[3 * k + 1];
    c += s[3 * k + 2];
  }
  __syncthreads();
}

__device__ __forceinline__ int wg_max(int v, int* s) {
  v = wave_reduce_max(v);
  const int w = threadIdx.x / kWave, nw = blockDim.x / kWave;
  __syncthreads();
  if (lane_id() == 0) s[w] = v;
  __syncthreads();
  int r = 0;
  for (int k = 0; k < nw; ++k) r = max(r, s[k]);
  __syncthreads();
  return r;
}

// Ascending order of 8 values: Batcher's odd-even merge sort network
// (19 compare-exchanges, all indices static so the values stay in registers).
__device__ __forceinline__ void cmp_swap(int& a, int& b) {
  const int lo = min(a, b), hi = max(a, b);
  a = lo;
  b = hi;
}
__device__ __forceinline__ void sort8(int* v) {
  cmp_swap(v[0], v[1]); cmp_swap(v[2], v[3]); cmp_swap(v[4], v[5]); cmp_swap(v[6], v[7]);
  cmp_swap(v[0], v[2]); cmp_swap(v[1], v[3]); cmp_swap(v[4], v[6]); cmp_swap(v[5], v[7]);
  cmp_swap(v[1], v[2]); cmp_swap(v[5], v[6]);
  cmp_swap(v[0], v[4]); cmp_swap(v[1], v[5]); cmp_swap(v[2], v[6]); cmp_swap(v[3], v[7]);
  cmp_swap(v[2], v[4]); cmp_swap(v[3], v[5]);
  cmp_swap(v[1], v[2]); cmp_swap(v[3], v[4]); cmp_swap(v[5], v[6]);
}

// Ascending order of one cell's ids m[0..cnt): up to 8 through the sorting
// network in registers, more by insertion sort in place.
template <typename T>
__device__ __forceinline__ void ens_sort_cell(T* m, int cnt) {
  if (cnt <= 8) {
    int v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = u < cnt ? (int)m[u] : 0x7fffffff;
    sort8(v);
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (u < cnt) m[u] = (T)v[u];
    return;
  }
  for (int u = 1; u < cnt; ++u) {
    const T v = m[u];
    int w = u - 1;
    while (w >= 0 && m[w] > v) {
      m[w + 1] = m[w];
      --w;
    }
    m[w + 1] = v;
  }
}

// A system's float32 positions (and species ids) as the pair and list loops
// read them: from LDS when the workgroup staged the system there (l, ls), else
// from global memory.  base is the system's first global id.
template <bool STAGED>
struct EnsView {
  const float* __restrict__ g;
  const float* l;
  const int* __restrict__ gs;
  const uint8_t* ls;  // staged species ids (nullptr: gs)
  int base;
  __device__ __forceinline__ void pos(int i, float* xi) const {
    if constexpr (STAGED) {
      xi[0] = l[2 * (i - base)];
      xi[1] = l[2 * (i - base) + 1];
    } else {
      load_pos<2>(g, i, xi);
    }
  }
  __device__ __forceinline__ int species(int i) const {
    if constexpr (STAGED) {
      if (ls) return (int)ls[i - base];
    }
    return gs ? gs[i] : 0;
  }
};

// Bins the system on grid g: LDS counts, exclusive starts, arrival ranks,
// scatter into sorted[] (system-local), then each cell's ids in ascending
// order (the stable argsort of partition.py:432).  s_cs holds the counts,
// then, scanned in place, the starts (cell c spans s_cs[c] .. s_cs[c + 1]).
// Returns the max filling.
template <bool STAGED>
__device__ __forceinline__ int ens_bin(const EnsSys& S, const Grid& g, const EnsView<STAGED>& V,
                                       int* __restrict__ cellp, int* __restrict__ rankp,
                                       int* __restrict__ sorted, int* s_cs, int* s_red,
                                       uint16_t* ls) {
  const int nt = blockDim.x, t = threadIdx.x;
  for (int c = t; c < g.ncell; c += nt) s_cs[c] = 0;
  __syncthreads();
  // staged system (ls: its cell-sorted ids in LDS): counts here, then after
  // the scan a second pass of LDS atomics on the starts places each id (the
  // cells are sorted by id afterwards, so arrival order does not matter);
  // cells are recomputed from the LDS positions, nothing goes to memory
  if constexpr (STAGED) {
    for (int a = t; a < S.n; a += nt) {
      float xi[2];
      V.pos(S.offset + a, xi);
      int ci[2];
      atomicAdd(&s_cs[cell_of<2>(xi, g, ci)], 1);
    }
  } else {
    for (int a = t; a < S.n; a += nt) {
      float xi[2];
      V.pos(S.offset + a, xi);
      int ci[2];
      const int h = cell_of<2>(xi, g, ci);
      cellp[a] = h;
      rankp[a] = atomicAdd(&s_cs[h], 1);
    }
  }
  __syncthreads();
  // exclusive scan of the counts: per-thread chunks, a wavefront scan of the
  // chunk totals, then the wave totals in order
  const int per = (g.ncell + nt - 1) / nt;
  const int beg = min(t * per, g.ncell), end = min(beg + per, g.ncell);
  int local = 0, fill = 0;
  for (int c = beg; c < end; ++c) {
    local += s_cs[c];
    fill = max(fill, s_cs[c]);
  }
  const int lane = lane_id(), wv = t / kWave, nw = nt / kWave;
  int incl = local;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const int y = __shfl_up(incl, off, kWave);
    if (lane >= off) incl += y;
  }
  __shared__ int s_wave[kEnsThreads / kWave + 1];
  if (lane == kWave - 1) s_wave[wv] = incl;
  __syncthreads();
  if (t == 0) {
    int acc = 0;
    for (int k = 0; k < nw; ++k) {
      const int v = s_wave[k];
      s_wave[k] = acc;
      acc += v;
    }
    s_wave[nw] = acc;
  }
  __syncthreads();
  int run = s_wave[wv] + incl - local;
  for (int c = beg; c < end; ++c) {
    const int cnt = s_cs[c];  // this thread's own chunk: read, then overwrite
    s_cs[c] = run;
    run += cnt;
  }
  if (t == 0) s_cs[g.ncell] = s_wave[nw];
  __syncthreads();
  if constexpr (STAGED) {
    for (int a = t; a < S.n; a += nt) {
      float xi[2];
      V.pos(S.offset + a, xi);
      int ci[2];
      ls[atomicAdd(&s_cs[cell_of<2>(xi, g, ci)], 1)] = (uint16_t)a;
    }
    __syncthreads();
    // s_cs[c] now holds the end of cell c, the start of cell c + 1: shift
    // back by one, chunks from the top (a chunk reads its own and the next
    // lower chunk's values before any of them is written)
    for (int c0 = ((g.ncell - 1) / nt) * nt; c0 >= 0; c0 -= nt) {
      const int c = c0 + t;
      const int v = c < g.ncell && c > 0 ? s_cs[c - 1] : 0;
      __syncthreads();
      if (c < g.ncell) s_cs[c] = v;
      __syncthreads();
    }
  } else {
    for (int a = t; a < S.n; a += nt) sorted[s_cs[cellp[a]] + rankp[a]] = a;
  }
  __syncthreads();
  // order each cell by id (thread per cell): up to 8 ids through a sorting
  // network in registers, more by insertion sort in place
  if constexpr (STAGED) {
    for (int c = t; c < g.ncell; c += nt) ens_sort_cell(ls + s_cs[c], s_cs[c + 1] - s_cs[c]);
  } else {
    for (int c = t; c < g.ncell; c += nt) ens_sort_cell(sorted + s_cs[c], s_cs[c + 1] - s_cs[c]);
  }
  __syncthreads();
  return wg_max(fill, s_red);
}

// Dense rows of the system (global ids, padding pad) in the reference's
// candidate order; returns the largest row count.
template <bool STAGED>
__device__ __forceinline__ int ens_rows(const EnsSys& S, const Grid& g, const Space& sp,
                                        float cutoff_sq, const EnsView<STAGED>& V,
                                        const int* __restrict__ cellp,
                        const int* __restrict__ sorted, const int* s_cs,
                        int* __restrict__ idx, int width, int pad, int* s_red,
                        const uint16_t* ls) {
  const int nt = blockDim.x, t = threadIdx.x;
  // The candidate test: the minimum-image |x_j - x_i|^2 of positions in the
  // box decides every candidate farther than `margin` from the cutoff (the
  // bound of the main cell build on the few float32 roundings by which it
  // can differ from the reference's wrapped float32 d^2); the rest, and
  // positions outside [0, L], take the reference-exact two-way test.
  const float rc = sqrtf(cutoff_sq), Lmax = fmaxf(sp.L[0], sp.L[1]);
  const float margin =
      fmaxf(128.0f * (rc + 1.0f) * (Lmax + rc + 1.0f) * 1.1920929e-7f, 1e-4f * cutoff_sq);
  const float lo_thr = cutoff_sq - margin, hi_thr = cutoff_sq + margin;
  // rows whose stride is a multiple of 4 entries are written 16 bytes at a
  // time: a row's entries collect in a register quad, flushed when full
  const bool vec = width > 0 && (width & 3) == 0 && ((uintptr_t)idx & 15) == 0;
  int occ = 0;
  for (int a = t; a < S.n; a += nt) {
    const int i = S.offset + a;
    float xi[2];
    V.pos(i, xi);
    const float xim[2] = {xi[0] - sp.L[0], xi[1] - sp.L[1]};
    const bool in_i = xi[0] >= 0.0f && xi[0] <= sp.L[0] && xi[1] >= 0.0f && xi[1] <= sp.L[1];
    int ci[2];
    const int h = STAGED ? cell_of<2>(xi, g, ci) : cellp[a];  // the binning's cell
    const int hx = h % g.cps[0], hy = h / g.cps[0];
    // the stencil's columns and rows, wrapped by comparison (offsets of one)
    const int cx[3] = {hx == 0 ? g.cps[0] - 1 : hx - 1, hx, hx + 1 == g.cps[0] ? 0 : hx + 1};
    const int cy[3] = {hy == 0 ? g.cps[1] - 1 : hy - 1, hy, hy + 1 == g.cps[1] ? 0 : hy + 1};
    int cnt = 0;
    int4 quad = make_int4(pad, pad, pad, pad);
    int* row = idx + (long long)i * width;
    for (int st = 0; st < 9; ++st) {  // iota order, x offset slowest
      const int c = cx[st / 3] + g.cps[0] * cy[st % 3];
      const int start = s_cs[c];
      const int m = min(s_cs[c + 1] - start, g.cap);
      const int w = g.cap - (start % g.cap);
      // the cell's members in chunks: ids, then their positions, then the
      // tests in slot order
      for (int q0 = 0; q0 < m; q0 += kEnsRowChunk) {
        int jl[kEnsRowChunk];
        float xj[kEnsRowChunk][2];
#pragma unroll
        for (int c2 = 0; c2 < kEnsRowChunk; ++c2) {
          const int q = q0 + c2;
          int k = q;  // slot order: rank (start + k) mod cap
          if (m > w) k = (q < m - w) ? (w + q) : (q - (m - w));
          jl[c2] = q < m ? (STAGED ? (int)ls[start + k] : sorted[start + k]) : -1;
        }
#pragma unroll
        for (int c2 = 0; c2 < kEnsRowChunk; ++c2) {
          xj[c2][0] = xj[c2][1] = 0.0f;
          if (jl[c2] >= 0) V.pos(S.offset + jl[c2], xj[c2]);
        }
#pragma unroll
        for (int c2 = 0; c2 < kEnsRowChunk; ++c2) {
          if (jl[c2] < 0 || jl[c2] == a) continue;  // past the cell; mask_self
          const int j = S.offset + jl[c2];
          const float e0 = min_image_disp(xj[c2][0], xi[0], xim[0], sp.L[0], sp.half[0]);
          const float e1 = min_image_disp(xj[c2][1], xi[1], xim[1], sp.L[1], sp.half[1]);
          const float d2 = fmaf(e1, e1, e0 * e0);
          const bool in_j = xj[c2][0] >= 0.0f && xj[c2][0] <= sp.L[0] && xj[c2][1] >= 0.0f &&
                            xj[c2][1] <= sp.L[1];
          bool keep = in_i && in_j && d2 < lo_thr;
          // Dense with cells keeps a candidate when both wrap(Ri - Rj) and
          // wrap(Rj - Ri) are inside (partition.py:951,971)
          if (!keep && (!(in_i && in_j) || d2 <= hi_thr))
            keep = ref_metric_sq<2>(xi, xj[c2], sp) < cutoff_sq &&
                   ref_metric_sq<2>(xj[c2], xi, sp) < cutoff_sq;
          if (keep) {
            if (cnt < width) {
              if (vec) {
                const int slot = cnt & 3;
                quad.x = slot == 0 ? j : quad.x;
                quad.y = slot == 1 ? j : quad.y;
                quad.z = slot == 2 ? j : quad.z;
                quad.w = slot == 3 ? j : quad.w;
                if (slot == 3) {
                  *reinterpret_cast<int4*>(row + (cnt & ~3)) = quad;
                  quad = make_int4(pad, pad, pad, pad);
                }
              } else {
                row[cnt] = j;
              }
            }
            ++cnt;
          }
        }
      }
    }
    if (vec) {
      // the partial quad (pads after the last entry), then whole pad quads
      int k = cnt < width ? (cnt & ~3) : width;
      if (k < width && (cnt & 3) != 0) {
        *reinterpret_cast<int4*>(row + k) = quad;
        k += 4;
      }
      for (; k < width; k += 4) *reinterpret_cast<int4*>(row + k) = make_int4(pad, pad, pad, pad);
    } else {
      for (int k = cnt; k < width; ++k) row[k] = pad;
    }
    occ = max(occ, cnt);
  }
  return wg_max(occ, s_red);
}

// A system's species matrices (sigma, epsilon, alpha in species mode, and
// the derivative coefficients dsig) copied to LDS, so that the pair loops
// read them there instead of gathering them per pair.  Returns s_par, or
// nullptr when the matrices are too large (the loops then read memory).
// Called by the whole workgroup.
__device__ __forceinline__ const float* ens_stage_params(const PairArgs& a,
                                                         const float* __restrict__ dsig,
                                                         float* s_par) {
  const int nn = a.ns * a.ns;
  if (a.ns < 1 || nn > kEnsParMax) return nullptr;
  for (int q = threadIdx.x; q < nn; q += blockDim.x) {
    s_par[q] = a.sig_mode == JMD_PARAM_SPECIES ? a.sig_v[q] : 0.0f;
    s_par[kEnsParMax + q] = a.eps_mode == JMD_PARAM_SPECIES ? a.eps_v[q] : 0.0f;
    s_par[2 * kEnsParMax + q] = a.alp_mode == JMD_PARAM_SPECIES ? a.alp_v[q] : 0.0f;
    s_par[3 * kEnsParMax + q] = dsig ? dsig[q] : 0.0f;
  }
  __syncthreads();
  return s_par;
}

// sigma, epsilon, alpha (which = 0, 1, 2) of pair (i, j), species matrices
// from LDS when staged.
__device__ __forceinline__ float ens_pick(const PairArgs& a, int which, const float* par, int i,
                                          int j, int si, int sj) {
  const int mode = which == 0 ? a.sig_mode : (which == 1 ? a.eps_mode : a.alp_mode);
  if (par && mode == JMD_PARAM_SPECIES) return par[which * kEnsParMax + si * a.ns + sj];
  return which == 0 ? pick(a.sig_mode, a.sig, a.sig_v, i, j, si, sj, a.ns)
                    : (which == 1 ? pick(a.eps_mode, a.eps, a.eps_v, i, j, si, sj, a.ns)
                                  : pick(a.alp_mode, a.alp, a.alp_v, i, j, si, sj, a.ns));
}

// dsig[si, sj], from LDS when staged.
__device__ __forceinline__ float ens_dsig(const PairArgs& a, const float* __restrict__ dsig,
                                          const float* par, int si, int sj) {
  return par ? par[3 * kEnsParMax + si * a.ns + sj] : dsig[si * a.ns + sj];
}

// Soft-sphere force of row i: two accumulators (entries k even / odd) summed
// in float64, as the Dense kernel's two lanes per row do; optional energy.
template <bool STAGED>
__device__ __forceinline__ void ens_force_row(int i, const EnsView<STAGED>& V,
                                              const int* __restrict__ idx, int width, int pad,
                                              const Space& sp, const PairArgs& a,
                                              const float* par, float* F, double* e) {
#pragma clang fp contract(off)
  float xi[2], xim[2];
  V.pos(i, xi);
  xim[0] = xi[0] - sp.L[0];
  xim[1] = xi[1] - sp.L[1];
  float f[2][2] = {{0.0f, 0.0f}, {0.0f, 0.0f}};
  float en[2] = {0.0f, 0.0f};
  const int si = V.species(i);
  const int* row = idx + (long long)i * width;
  // The row in chunks: the chunk's ids, then all its gathers (position,
  // species), then the terms in entry order; a row ends at its first pad.
  // rows whose stride is a multiple of 4 entries are read 16 bytes at a time
  const bool vec = kEnsRowChunk == 4 && (width & 3) == 0 && ((uintptr_t)idx & 15) == 0;
  for (int k0 = 0; k0 < width; k0 += kEnsRowChunk) {
    int js[kEnsRowChunk], sjs[kEnsRowChunk];
    float xj[kEnsRowChunk][2];
    if (vec) {
      const int4 q = *reinterpret_cast<const int4*>(row + k0);
      js[0] = q.x;
      js[1 % kEnsRowChunk] = q.y;
      js[2 % kEnsRowChunk] = q.z;
      js[3 % kEnsRowChunk] = q.w;
    } else {
#pragma unroll
      for (int c = 0; c < kEnsRowChunk; ++c) js[c] = k0 + c < width ? row[k0 + c] : pad;
    }
#pragma unroll
    for (int c = 0; c < kEnsRowChunk; ++c) {
      xj[c][0] = xj[c][1] = 0.0f;
      sjs[c] = 0;
      if (js[c] != pad) {
        V.pos(js[c], xj[c]);
        sjs[c] = V.species(js[c]);
      }
    }
    bool end = false;
#pragma unroll
    for (int c = 0; c < kEnsRowChunk; ++c) {
      const int j = js[c];
      if (j == pad) {
        end = true;
        break;
      }
      float d[2];
      d[0] = min_image_disp(xj[c][0], xi[0], xim[0], sp.L[0], sp.half[0]);
      d[1] = min_image_disp(xj[c][1], xi[1], xim[1], sp.L[1], sp.half[1]);
      const float r2 = fmaf(d[1], d[1], d[0] * d[0]);
      const int sj = sjs[c];
      const float sig = ens_pick(a, 0, par, i, j, si, sj);
      const float eps = ens_pick(a, 1, par, i, j, si, sj);
      const float alp = ens_pick(a, 2, par, i, j, si, sj);
      float dEr;
      const float u = pair_eval<JMD_SOFT_SPHERE>(r2, sig, eps, alp, a, &dEr);
      const int l = c & 1;  // k0 is even: the parity of k = k0 + c
      f[l][0] = fmaf(dEr, d[0], f[l][0]);
      f[l][1] = fmaf(dEr, d[1], f[l][1]);
      en[l] += u;
    }
    if (end) break;
  }
  F[0] = (float)((double)f[0][0] + (double)f[1][0]);
  F[1] = (float)((double)f[0][1] + (double)f[1][1]);
  if (e) *e += 0.5 * ((double)en[0] + (double)en[1]);
}

// STAGED: the systems of at most lds_n particles, kept in LDS; else the
// larger ones, read from global memory.  Each instance skips the other's.
template <bool STAGED>
__global__ __launch_bounds__(kEnsThreads) void k_fire_ensemble(
    const EnsSys* __restrict__ systems, const int* __restrict__ order, int nsys, int ntot,
    int mode, int n_steps, double* __restrict__ x_hi, float* __restrict__ x,
    float* __restrict__ p, float* __restrict__ F, float* __restrict__ x_ref,
    int* __restrict__ idx, int width, int* __restrict__ scratch, PairArgs a0, FireP prm,
    float* __restrict__ fire_state, int* __restrict__ fire_npos, uint8_t* __restrict__ error,
    int* __restrict__ out_stat, double* __restrict__ energy, int* __restrict__ builds,
    EnsTape tape, int lds_n) {
  __shared__ int s_cs[kEnsMaxCells + 1];
  // dynamic LDS: [2 lds_n] float32 positions, [lds_n] uint16 cell-sorted
  // ids, [lds_n] uint8 species ids
  extern __shared__ float4 s_dyn[];
  __shared__ int s_red[kEnsThreads / kWave];
  __shared__ double s_d[3 * (kEnsThreads / kWave)];
  __shared__ float s_scal[4];  // dt, alpha, ratio, keep
  __shared__ int s_flag;
  const int sid = order[blockIdx.x];
  if (sid < 0 || sid >= nsys) return;
  const EnsSys S = systems[sid];
  if (STAGED != (S.n <= lds_n)) return;  // the other instance's system
  const int t = threadIdx.x, nt = blockDim.x;
  const float cutoff_sq = S.cutoff_sq, thr_sq = S.threshold_sq;
  PairArgs a = a0;  // this system's parameter block
  if (a.sig_mode == JMD_PARAM_SPECIES) a.sig_v += S.param_offset;
  if (a.eps_mode == JMD_PARAM_SPECIES) a.eps_v += S.param_offset;
  if (a.alp_mode == JMD_PARAM_SPECIES) a.alp_v += S.param_offset;
  __shared__ float s_par[4 * kEnsParMax];
  const float* par = ens_stage_params(a, nullptr, s_par);
  // A staged system keeps a copy of its positions (and species ids, when
  // they fit a byte) and its cell-sorted ids in LDS for the list and pair
  // loops; the integrator writes both position copies.
  float* lx = STAGED ? reinterpret_cast<float*>(s_dyn) : nullptr;
  uint16_t* lsorted = STAGED ? reinterpret_cast<uint16_t*>(lx + 2 * lds_n) : nullptr;
  uint8_t* lsp = STAGED && a.species && a.ns <= 256
                     ? reinterpret_cast<uint8_t*>(lsorted + lds_n) : nullptr;
  if constexpr (STAGED) {
    for (int k = threadIdx.x; k < 2 * S.n; k += blockDim.x) lx[k] = x[2LL * S.offset + k];
    if (lsp)
      for (int q = threadIdx.x; q < S.n; q += blockDim.x)
        lsp[q] = (uint8_t)a.species[S.offset + q];
    __syncthreads();
  }
  const EnsView<STAGED> V{x, lx, a.species, lsp, S.offset};
  int* cellp = scratch + S.offset;
  int* rankp = scratch + ntot + S.offset;
  int* sorted = scratch + 2LL * ntot + S.offset;
  Space sp;
  sp.dim = 2;
  sp.periodic = 1;
  sp.general = 0;
  for (int d = 0; d < 3; ++d) {
    sp.L[d] = d < 2 ? S.side[d] : 1.0f;
    sp.half[d] = sp.L[d] * 0.5f;
  }
  Grid g;
  g.cps[0] = S.cps[0];
  g.cps[1] = S.cps[1];
  g.cps[2] = 1;
  g.cell_size[0] = S.cell_size[0];
  g.cell_size[1] = S.cell_size[1];
  g.cell_size[2] = 1.0f;
  g.ncell = S.cps[0] * S.cps[1];
  g.cap = S.cap;
  const int pad = ntot;

  if (mode == ENS_FILL) {
    Grid g2 = g;
    g2.cps[0] = S.fill_cps[0];
    g2.cps[1] = S.fill_cps[1];
    g2.cell_size[0] = S.fill_cell_size[0];
    g2.cell_size[1] = S.fill_cell_size[1];
    g2.ncell = g2.cps[0] * g2.cps[1];
    const int fill = ens_bin(S, g2, V, cellp, rankp, sorted, s_cs, s_red, lsorted);
    if (t == 0) out_stat[sid] = fill;
    return;
  }

  uint8_t err = 0;
  int nbuilds = 0;
  // list build at the current positions (allocate / update)
  auto rebuild = [&](bool store) {
    const int fill = ens_bin(S, g, V, cellp, rankp, sorted, s_cs, s_red, lsorted);
    const int occ = ens_rows(S, g, sp, cutoff_sq, V, cellp, sorted, s_cs, idx,
                             store ? width : 0, pad, s_red, lsorted);
    if (fill > g.cap) err |= JMD_CELL_LIST_OVERFLOW;
    if (occ > S.max_occ) err |= JMD_NEIGHBOR_LIST_OVERFLOW;
    for (int q = t; q < S.n; q += nt) {
      float xi[2];
      V.pos(S.offset + q, xi);
      store_pos<2>(x_ref, S.offset + q, xi);
    }
    ++nbuilds;
    return occ;
  };

  if (mode == ENS_OCCUPANCY) {
    const int occ = rebuild(false);
    if (t == 0) out_stat[sid] = occ;
    return;
  }

  const double L0 = S.side[0], L1 = S.side[1];
  if (mode == ENS_INIT) {
    rebuild(true);
    __syncthreads();
    double e = 0.0, z1 = 0.0, z2 = 0.0;
    for (int q = t; q < S.n; q += nt) {
      const int i = S.offset + q;
      float f[2];
      ens_force_row(i, V, idx, width, pad, sp, a, par, f, &e);
      F[2LL * i] = f[0];
      F[2LL * i + 1] = f[1];
      p[2LL * i] = 0.0f;
      p[2LL * i + 1] = 0.0f;
    }
    wg_sum3(e, z1, z2, s_d);
    if (t == 0) {
      fire_state[2 * sid] = prm.dt_start;
      fire_state[2 * sid + 1] = prm.alpha_start;
      fire_npos[sid] = 0;
      energy[sid] = e;
      error[sid] = (uint8_t)(error[sid] | err);
      builds[sid] += nbuilds;
    }
    return;
  }

  // ENS_RUN
  if (t == 0) {
    s_scal[0] = fire_state[2 * sid];
    s_scal[1] = fire_state[2 * sid + 1];
    s_flag = fire_npos[sid];
  }
  __syncthreads();
  int n_pos = s_flag;
  const long long n2 = 2LL * ntot;
  auto record_node = [&](int k) {  // x, F, list at node k of the segment
    for (int q = t; q < 2 * S.n; q += nt) {
      const long long g = 2LL * S.offset + q;
      tape.x[k * n2 + g] = x[g];
      tape.F[k * n2 + g] = F[g];
    }
    const long long rows = (long long)S.n * width;
    int* dst = tape.idx + (long long)k * ntot * width + (long long)S.offset * width;
    const int* src = idx + (long long)S.offset * width;
    for (long long q = t; q < rows; q += nt) dst[q] = src[q];
  };
  if (tape.x) record_node(0);
  // The half-step pass also makes the next step's skin test (on the
  // positions it writes, against the reference positions, which only a
  // rebuild changes, before the half-step) and applies the previous step's
  // FIRE mixing to each momentum it reads (the mixing needs the step's
  // reductions, so it waits for the next pass over the momenta); the
  // per-element arithmetic is that of the separate passes.
  // the pending mixing's ratio, keep and alpha (workgroup-uniform: in LDS,
  // out of the registers the force loop needs)
  __shared__ double s_mix[3];
  int list_slot = 0;  // tape slot of the current list (record_node(0) fills 0)
  bool moved_next = false;
  bool mix_pending = false;
  for (int step = 0; step < n_steps; ++step) {
    // neighbour update: rebuild when any particle moved beyond half the skin
    bool moved = moved_next;
    if (step == 0) {
#pragma unroll 4
      for (int q = t; q < S.n; q += nt) {
        float xi[2], ri[2];
        V.pos(S.offset + q, xi);
        load_pos<2>(x_ref, S.offset + q, ri);
        moved = moved || ref_metric_sq<2>(xi, ri, sp) > thr_sq;
      }
    }
    const bool rebuilt = __syncthreads_or(moved);
    if (rebuilt) rebuild(true);
    __syncthreads();
    if (tape.x && rebuilt) {  // the list the force at node step + 1 will use
      list_slot = step + 1;
      const long long rows = (long long)S.n * width;
      int* dst = tape.idx + (long long)(step + 1) * ntot * width + (long long)S.offset * width;
      const int* src = idx + (long long)S.offset * width;
      for (long long q = t; q < rows; q += nt) dst[q] = src[q];
    }
    // velocity-Verlet sub-step at the FIRE time step (float16 value), one
    // particle per iteration
    const float dt = s_scal[0];
    const float half = dt / 2.0f;
    const double mix_ratio = s_mix[0], mix_keep = s_mix[1];
    const float mix_alpha = (float)s_mix[2];
    moved_next = false;
    for (int q = t; q < S.n; q += nt) {
      const long long g0 = 2LL * (S.offset + q);
      float xn[2];
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const long long gq = g0 + c;
        float p0 = p[gq];
        const float fq = F[gq];
        if (mix_pending) {  // the previous step's FIRE mixing
          double pv = p0;
          pv = pv + (double)mix_alpha * ((double)fq * mix_ratio - pv);
          p0 = (float)(pv * mix_keep);
        }
        const float pk = p0 + half * fq;
        double xv = x_hi[gq] + (double)dt * (double)pk;
        xv = ref_mod64(xv, c ? L1 : L0);
        p[gq] = pk;
        x_hi[gq] = xv;
        xn[c] = (float)xv;
        x[gq] = xn[c];
        if constexpr (STAGED) lx[2 * q + c] = xn[c];
      }
      float ri[2];
      load_pos<2>(x_ref, S.offset + q, ri);
      moved_next = moved_next || ref_metric_sq<2>(xn, ri, sp) > thr_sq;
    }
    __syncthreads();
    for (int q = t; q < S.n; q += nt) {
      const int i = S.offset + q;
      float f[2];
      ens_force_row(i, V, idx, width, pad, sp, a, par, f, nullptr);
      F[2LL * i] = f[0];
      F[2LL * i + 1] = f[1];
    }
    __syncthreads();
    double ff = 0.0, pp = 0.0, fp = 0.0;
    if (tape.x) {
      for (int k = t; k < 2 * S.n; k += nt) {
        const long long q = 2LL * S.offset + k;
        const float pk = p[q] + half * F[q];
        p[q] = pk;
        const double fv = F[q], pv = pk;
        ff += fv * fv;
        pp += pv * pv;
        fp += fv * pv;
        tape.q[step * n2 + q] = pk;
        tape.x[(step + 1) * n2 + q] = x[q];
        tape.F[(step + 1) * n2 + q] = F[q];
      }
    } else {
#pragma unroll 4
      for (int k = t; k < 2 * S.n; k += nt) {
        const long long q = 2LL * S.offset + k;
        const float pk = p[q] + half * F[q];
        p[q] = pk;
        const double fv = F[q], pv = pk;
        ff += fv * fv;
        pp += pv * pv;
        fp += fv * pv;
      }
    }
    wg_sum3(ff, pp, fp, s_d);
    // FIRE bookkeeping (minimize.py:190-219) with float16 dt / alpha
    const double ratio = sqrt(pp) / sqrt(ff + 1e-6);
    const float alpha_mix = s_scal[1];
    const double keep = fp >= 0.0 ? 1.0 : 0.0;
    if (tape.x && t == 0) {
      float* sc = tape.sc + ((long long)step * nsys + sid) * kTapeScalars;
      sc[0] = dt;
      sc[1] = alpha_mix;
      sc[2] = (float)keep;
      sc[3] = (float)sqrt(pp);
      sc[4] = (float)sqrt(ff + 1e-6);
      sc[5] = (float)list_slot;
    }
    n_pos = fp >= 0.0 ? n_pos + 1 : 0;
    __syncthreads();
    if (t == 0) {
      float dtn = s_scal[0], al = s_scal[1];
      if (fp > 0.0 && n_pos > prm.n_min) dtn = fminf(f16r(dtn * prm.f_inc), prm.dt_max);
      if (fp < 0.0) dtn = f16r(dtn * prm.f_dec);
      if (fp > 0.0 && n_pos > prm.n_min) al = f16r(al * prm.f_alpha);
      if (fp < 0.0) al = prm.alpha_start;
      s_scal[0] = dtn;
      s_scal[1] = al;
      s_mix[0] = ratio;
      s_mix[1] = keep;
      s_mix[2] = (double)alpha_mix;
    }
    mix_pending = true;  // applied by the next half-step, or below
    __syncthreads();
  }
  if (mix_pending) {  // the last step's FIRE mixing
    const double mix_ratio = s_mix[0], mix_keep = s_mix[1];
    const float mix_alpha = (float)s_mix[2];
#pragma unroll 4
    for (int k = t; k < 2 * S.n; k += nt) {
      const long long q = 2LL * S.offset + k;
      double pv = p[q];
      pv = pv + (double)mix_alpha * ((double)F[q] * mix_ratio - pv);
      p[q] = (float)(pv * mix_keep);
    }
    __syncthreads();
  }
  double e = 0.0, z1 = 0.0, z2 = 0.0;
  for (int q = t; q < S.n; q += nt) {
    float f[2];
    ens_force_row(S.offset + q, V, idx, width, pad, sp, a, par, f, &e);
  }
  wg_sum3(e, z1, z2, s_d);
  if (t == 0) {
    fire_state[2 * sid] = s_scal[0];
    fire_state[2 * sid + 1] = s_scal[1];
    fire_npos[sid] = n_pos;
    energy[sid] = e;
    error[sid] = (uint8_t)(error[sid] | err);
    builds[sid] += nbuilds;
  }
}

// ---------------------------------------------------------------------------
// Adjoint of the FIRE trajectory (per system, one workgroup each)
//
// Forward step k (dt, alpha, keep from the tape; float16 bookkeeping is
// piecewise constant and carries no derivative):
//   ph = p_k + h F_k,  x_{k+1} = x_k + dt ph,  F' = F(x_{k+1}),  q = ph + h F',
//   p_{k+1} = keep ((1 - alpha) q + alpha |q| / sqrt(|F'|^2 + 1e-6) F').
// Backward, with X, P the adjoints of x, p and W the pending vector of node
// k + 1 (W = h PH of the step after):
//   G = keep P,  s1 = F'.G,
//   Q  = (1 - alpha) G + alpha s1 q / (|q| sqrt(ff)),
//   Fb = alpha |q|/sqrt(ff) G - alpha s1 |q| F' / sqrt(ff)^3 + h Q,
//   X += J(x_{k+1}) (Fb + W),  theta += (Fb + W)^T dF(x_{k+1})/dtheta,
//   P = Q + dt X,  W = h P.
// One product with the force Jacobian per step.  The loss is
// scale * sum_s E_s(x_T), theta the derivative with respect to each system's
// parameter (the species matrix's derivative coefficients dsigma/dtheta).
//
// Box adjoint (optional): the derivative with respect to the side L of each
// system's square, positions held in real coordinates.  L enters through the
// minimum image, d = x_j - x_i - L n (n the pair's integer image vector, so
// dd/dL = -n), and through the fold of the shift, x_{k+1} = x_k + dt ph -
// L m_k (m_k = rint((x_k - x_{k+1}) / L), a step being far shorter than L/2):
//   Lb  = -scale/2 sum_ij g (d.n)                           at x_T,
//   Lb += -sum_ij V_i.(g n + 2 g' (d.n) d)    at every node (V = Fb + W),
//   Lb += -X_{k+1}.m_k                        for every step.
// ---------------------------------------------------------------------------

enum { BACK_INIT = 0, BACK_STEPS = 1, BACK_FINISH = 2 };

// Integer image of a pair along one axis: the fold min_image_disp applied to
// raw = x_j - x_i (d = raw - n L).
__device__ __forceinline__ float image_of(float raw, float half) {
  return raw >= half ? 1.0f : (raw < -half ? -1.0f : 0.0f);
}

// h_i = (J v)_i over the system's rows at positions x with list idx; returns
// this thread's share of v^T dF/dtheta (rows it owns) and, with box, adds its
// share of v^T dF/dL to lacc.
template <bool STAGED>
__device__ __forceinline__ double ens_hvp_rows(const EnsSys& S, const Space& sp,
                                               const PairArgs& a, const float* par,
                                               const float* __restrict__ dsig,
                                               const EnsView<STAGED>& XV,
                                               const int* __restrict__ idx, int width, int pad,
                                               const EnsView<STAGED>& VV,
                                               double* __restrict__ X, bool box,
                                               double& lacc) {
#pragma clang fp contract(off)
  double tacc = 0.0;
  // rows whose stride is a multiple of 4 entries are read 16 bytes at a time
  const bool vec = kEnsRowChunk == 4 && (width & 3) == 0 && ((uintptr_t)idx & 15) == 0;
  for (int q = threadIdx.x; q < S.n; q += blockDim.x) {
    const int i = S.offset + q;
    float xi[2], xim[2], vi[2];
    XV.pos(i, xi);
    VV.pos(i, vi);
    xim[0] = xi[0] - sp.L[0];
    xim[1] = xi[1] - sp.L[1];
    const int si = XV.species(i);
    float h0 = 0.0f, h1 = 0.0f, tp = 0.0f, lp = 0.0f;
    const int* row = idx + (long long)i * width;
    // the row in chunks: ids, then their gathers, then the terms in order
    for (int k0 = 0; k0 < width; k0 += kEnsRowChunk) {
      int js[kEnsRowChunk], sjs[kEnsRowChunk];
      float xj[kEnsRowChunk][2], vj[kEnsRowChunk][2];
      if (vec) {
        const int4 q4 = *reinterpret_cast<const int4*>(row + k0);
        js[0] = q4.x;
        js[1 % kEnsRowChunk] = q4.y;
        js[2 % kEnsRowChunk] = q4.z;
        js[3 % kEnsRowChunk] = q4.w;
      } else {
#pragma unroll
        for (int c = 0; c < kEnsRowChunk; ++c) js[c] = k0 + c < width ? row[k0 + c] : pad;
      }
#pragma unroll
      for (int c = 0; c < kEnsRowChunk; ++c) {
        xj[c][0] = xj[c][1] = vj[c][0] = vj[c][1] = 0.0f;
        sjs[c] = 0;
        if (js[c] != pad) {
          XV.pos(js[c], xj[c]);
          VV.pos(js[c], vj[c]);
          sjs[c] = XV.species(js[c]);
        }
      }
      bool end = false;
#pragma unroll
      for (int c = 0; c < kEnsRowChunk; ++c) {
        const int j = js[c];
        if (j == pad) {
          end = true;
          break;
        }
        float d[2], w[2];
        d[0] = min_image_disp(xj[c][0], xi[0], xim[0], sp.L[0], sp.half[0]);
        d[1] = min_image_disp(xj[c][1], xi[1], xim[1], sp.L[1], sp.half[1]);
        w[0] = vj[c][0] - vi[0];
        w[1] = vj[c][1] - vi[1];
        const float r2 = fmaf(d[1], d[1], d[0] * d[0]);
        const int sj = sjs[c];
        const float sig = ens_pick(a, 0, par, i, j, si, sj);
        const float eps = ens_pick(a, 1, par, i, j, si, sj);
        const float alp = ens_pick(a, 2, par, i, j, si, sj);
        const PairD2 D = pair_eval2<JMD_SOFT_SPHERE>(r2, sig, eps, alp, a);
        const float dw = fmaf(d[1], w[1], d[0] * w[0]);
        const float tt = 2.0f * D.gp * dw;
        h0 = fmaf(D.g, w[0], fmaf(tt, d[0], h0));
        h1 = fmaf(D.g, w[1], fmaf(tt, d[1], h1));
        if (dsig) {
          const float cc = ens_dsig(a, dsig, par, si, sj);
          const float lam = fmaf(vi[1], d[1], vi[0] * d[0]);
          tp = fmaf(cc * D.gs, lam, tp);
        }
        if (box) {
          const float n0 = image_of(xj[c][0] - xi[0], sp.half[0]);
          const float n1 = image_of(xj[c][1] - xi[1], sp.half[1]);
          const float dn = fmaf(d[1], n1, d[0] * n0);
          const float vn = fmaf(vi[1], n1, vi[0] * n0);
          const float vd = fmaf(vi[1], d[1], vi[0] * d[0]);
          lp = fmaf(D.g, vn, fmaf(2.0f * D.gp * dn, vd, lp));
        }
      }
      if (end) break;
    }
    X[2LL * i] += (double)h0;
    X[2LL * i + 1] += (double)h1;
    tacc += (double)tp;
    lacc -= (double)lp;
  }
  return tacc;
}

// STAGED: systems of at most lds_n particles, with the positions of the
// node and the vector of each Jacobian product copied to LDS; else the
// larger ones, read from memory.  Each instance skips the other's systems.
template <bool STAGED>
__global__ __launch_bounds__(kEnsThreads) void k_fire_ensemble_back(
    const EnsSys* __restrict__ systems, int nsys, int ntot, int mode, int K,
    const float* __restrict__ x, const float* __restrict__ F, const int* __restrict__ idx,
    int width, PairArgs a0, const float* __restrict__ dsig0, EnsTape tape,
    double* __restrict__ X, double* __restrict__ P, float* __restrict__ W,
    float* __restrict__ V, double* __restrict__ theta, double* __restrict__ box_adj,
    float scale, int lds_n) {
#pragma clang fp contract(off)
  // dynamic LDS: [2 lds_n] positions, [2 lds_n] product vector, [lds_n]
  // uint8 species ids
  extern __shared__ float4 s_dynb[];
  __shared__ double s_d[3 * (kEnsThreads / kWave)];
  const int sid = blockIdx.x;
  if (sid >= nsys) return;
  const EnsSys S = systems[sid];
  if (STAGED != (S.n <= lds_n)) return;  // the other instance's system
  const int t = threadIdx.x, nt = blockDim.x;
  PairArgs a = a0;
  if (a.sig_mode == JMD_PARAM_SPECIES) a.sig_v += S.param_offset;
  if (a.eps_mode == JMD_PARAM_SPECIES) a.eps_v += S.param_offset;
  if (a.alp_mode == JMD_PARAM_SPECIES) a.alp_v += S.param_offset;
  const float* dsig = dsig0 ? dsig0 + S.param_offset : nullptr;
  __shared__ float s_par[4 * kEnsParMax];
  const float* par = ens_stage_params(a, dsig, s_par);
  float* lx = STAGED ? reinterpret_cast<float*>(s_dynb) : nullptr;
  float* lv = STAGED ? lx + 2 * lds_n : nullptr;
  uint8_t* lsp = STAGED && a.species && a.ns <= 256
                     ? reinterpret_cast<uint8_t*>(lv + 2 * lds_n) : nullptr;
  if (lsp)
    for (int q = t; q < S.n; q += nt) lsp[q] = (uint8_t)a.species[S.offset + q];
  Space sp;
  sp.dim = 2;
  sp.periodic = 1;
  sp.general = 0;
  for (int d = 0; d < 3; ++d) {
    sp.L[d] = d < 2 ? S.side[d] : 1.0f;
    sp.half[d] = sp.L[d] * 0.5f;
  }
  const int pad = ntot;
  const long long n2 = 2LL * ntot;
  const long long b2 = 2LL * S.offset;
  const bool box = box_adj != nullptr;

  if (mode == BACK_INIT) {  // X = -scale F, P = W = 0, theta = scale dE/dtheta
    double e = 0.0, el = 0.0, z2 = 0.0;
    for (int k = t; k < 2 * S.n; k += nt) {
      X[b2 + k] = -(double)scale * (double)F[b2 + k];
      P[b2 + k] = 0.0;
      W[b2 + k] = 0.0f;
    }
    if (dsig || box) {
      for (int q = t; q < S.n; q += nt) {
        const int i = S.offset + q;
        float xi[2], xim[2];
        load_pos<2>(x, i, xi);
        xim[0] = xi[0] - sp.L[0];
        xim[1] = xi[1] - sp.L[1];
        const int si = a.species ? a.species[i] : 0;
        const int* row = idx + (long long)i * width;
        for (int k = 0; k < width; ++k) {
          const int j = row[k];
          if (j == pad) break;
          float xj[2], d[2];
          load_pos<2>(x, j, xj);
          d[0] = min_image_disp(xj[0], xi[0], xim[0], sp.L[0], sp.half[0]);
          d[1] = min_image_disp(xj[1], xi[1], xim[1], sp.L[1], sp.half[1]);
          const float r2 = fmaf(d[1], d[1], d[0] * d[0]);
          const int sj = a.species ? a.species[j] : 0;
          const float sig = ens_pick(a, 0, par, i, j, si, sj);
          const float eps = ens_pick(a, 1, par, i, j, si, sj);
          const float alp = ens_pick(a, 2, par, i, j, si, sj);
          const PairD2 D = pair_eval2<JMD_SOFT_SPHERE>(r2, sig, eps, alp, a);
          if (dsig) e += 0.5 * (double)(ens_dsig(a, dsig, par, si, sj) * D.us);
          if (box) {
            const float n0 = image_of(xj[0] - xi[0], sp.half[0]);
            const float n1 = image_of(xj[1] - xi[1], sp.half[1]);
            el -= 0.5 * (double)(D.g * fmaf(d[1], n1, d[0] * n0));
          }
        }
      }
    }
    wg_sum3(e, el, z2, s_d);
    if (t == 0) {
      theta[sid] = (double)scale * e;
      if (box) box_adj[sid] = (double)scale * el;
    }
    return;
  }
  if (mode == BACK_FINISH) {  // pending W applied at node 0 of the tape
    double el = 0.0, z2 = 0.0;
    if constexpr (STAGED) {
      for (int m = t; m < 2 * S.n; m += nt) {
        lx[m] = tape.x[b2 + m];
        lv[m] = W[b2 + m];
      }
    }
    __syncthreads();
    const EnsView<STAGED> XV{tape.x, lx, a.species, lsp, S.offset};
    const EnsView<STAGED> VV{W, lv, nullptr, nullptr, S.offset};
    double e = ens_hvp_rows(S, sp, a, par, dsig, XV, tape.idx, width, pad, VV, X, box, el);
    wg_sum3(e, el, z2, s_d);
    if (t == 0) {
      theta[sid] += e;
      if (box) box_adj[sid] += el;
    }
    return;
  }
  double tsum = 0.0, lsum = 0.0;
  for (int k = K - 1; k >= 0; --k) {
    const float* sc = tape.sc + ((long long)k * nsys + sid) * kTapeScalars;
    const float dt = sc[0], alpha = sc[1], keep = sc[2];
    const double spp = sc[3], sff = sc[4];
    const float h = dt / 2.0f;
    const float* Fn = tape.F + (long long)(k + 1) * n2;
    const float* qk = tape.q + (long long)k * n2;
    // (1) FIRE mixing, backward
    double s1 = 0.0, z1 = 0.0, z2 = 0.0;
    for (int m = t; m < 2 * S.n; m += nt) s1 += (double)Fn[b2 + m] * (double)keep * P[b2 + m];
    wg_sum3(s1, z1, z2, s_d);
    const double ratio = spp / sff;
    const double cq = spp > 0.0 ? (double)alpha * s1 / (spp * sff) : 0.0;
    const double cf = (double)alpha * s1 * spp / (sff * sff * sff);
    const float* xn = tape.x + (long long)(k + 1) * n2;
    for (int m = t; m < 2 * S.n; m += nt) {
      const long long g = b2 + m;
      const double G = (double)keep * P[g];
      const double Q = (1.0 - (double)alpha) * G + cq * (double)qk[g];
      const double Fb = (double)alpha * ratio * G - cf * (double)Fn[g] + (double)h * Q;
      P[g] = Q;  // PH until step (3)
      const float vg = (float)(Fb + (double)W[g]);
      if constexpr (STAGED) {
        lv[m] = vg;
        lx[m] = xn[g];
      } else {
        V[g] = vg;
      }
    }
    __syncthreads();
    // (2) product with the force Jacobian at node k + 1
    const EnsView<STAGED> XV{xn, lx, a.species, lsp, S.offset};
    const EnsView<STAGED> VV{V, lv, nullptr, nullptr, S.offset};
    tsum += ens_hvp_rows(S, sp, a, par, dsig, XV,
                         tape.idx + (long long)sc[5] * ntot * width, width, pad, VV, X, box,
                         lsum);
    __syncthreads();
    // (3) x_{k+1} = x_k + dt ph - L m_k, ph = p_k + h F_k
    const float* xk = tape.x + (long long)k * n2;
    const float* xk1 = tape.x + (long long)(k + 1) * n2;
    for (int m = t; m < 2 * S.n; m += nt) {
      const long long g = b2 + m;
      if (box) {
        const float fold = rintf((xk[g] - xk1[g]) / ((m & 1) ? sp.L[1] : sp.L[0]));
        lsum -= X[g] * (double)fold;
      }
      const double PH = P[g] + (double)dt * X[g];
      P[g] = PH;
      W[g] = (float)((double)h * PH);
    }
    __syncthreads();
  }
  double e = tsum, el = lsum, z2 = 0.0;
  wg_sum3(e, el, z2, s_d);
  if (t == 0) {
    theta[sid] += e;
    if (box) box_adj[sid] += el;
  }
}

}  // namespace jmd

using namespace jmd;

extern "C" {

// Allows a staged kernel more than the default dynamic LDS, once per device
// and size (the largest size granted so far is kept per device).
static bool ens_lds_allow(const void* kernel, size_t* granted, size_t bytes) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return false;
  if (bytes <= granted[dev]) return true;
  if (hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) !=
      hipSuccess)
    return false;
  granted[dev] = bytes;
  return true;
}

static bool ens_lds_ready(size_t bytes) {
  static size_t granted[64] = {};
  return ens_lds_allow(reinterpret_cast<const void*>(&k_fire_ensemble<true>), granted, bytes);
}

static bool ens_lds_ready_back(size_t bytes) {
  static size_t granted[64] = {};
  return ens_lds_allow(reinterpret_cast<const void*>(&k_fire_ensemble_back<true>), granted,
                       bytes);
}

int jmd_fire_ensemble_2d(const jmd_system2d* systems, const int32_t* order, int32_t n_systems,
                         int64_t n_total, int32_t mode, int32_t n_steps, double* position_hi,
                         float* position, float* momentum, float* force,
                         float* reference_position, int32_t* idx, int32_t width,
                         int32_t* scratch, const jmd_pair_potential* potential,
                         const float* fire_params, float* fire_state, int32_t* fire_n_pos,
                         uint8_t* error, int32_t* out_stat, double* energy, int32_t* builds,
                         float* tape_x, float* tape_force, float* tape_q, int32_t* tape_idx,
                         float* tape_scalars, void* stream) {
  if (n_systems < 0 || n_total < 0 || n_total >= 0x7fffffff || width < 0 || n_steps < 0 ||
      mode < ENS_FILL || mode > ENS_RUN || !fire_params)
    return JMD_ERR_INVALID;
  if (n_systems == 0) return JMD_OK;
  if (!systems || !order || !position || !scratch || !error || !builds) return JMD_ERR_INVALID;
  if ((mode == ENS_FILL || mode == ENS_OCCUPANCY) && !out_stat) return JMD_ERR_INVALID;
  if (mode >= ENS_INIT && (!position_hi || !momentum || !force || !reference_position ||
                           !idx || !fire_state || !fire_n_pos || !energy || width < 1))
    return JMD_ERR_INVALID;
  if (mode == ENS_OCCUPANCY && !reference_position) return JMD_ERR_INVALID;
  PairArgs a;
  if (!make_pair_args(potential, &a) || a.kind != JMD_SOFT_SPHERE) return JMD_ERR_INVALID;
  FireP prm{fire_params[0], fire_params[1], (int)fire_params[2], fire_params[3],
            fire_params[4], fire_params[5], fire_params[6]};
  static_assert(sizeof(jmd_system2d) == sizeof(EnsSys), "system record layout");
  const bool taped = tape_x != nullptr;
  if (taped && (mode != ENS_RUN || !tape_force || !tape_q || !tape_idx || !tape_scalars))
    return JMD_ERR_INVALID;
  EnsTape tape{tape_x, tape_force, tape_q, tape_idx, tape_scalars, n_steps};
  // LDS copy of each system's positions, cell-sorted ids and species ids:
  // systems of up to 8 particles per thread (8 Ki per 1024-thread workgroup,
  // one per CU; 4 Ki per 512, two per CU); larger systems use global memory
  const int threads = ens_threads(n_systems, n_total);
  const int lds_n = threads * kEnsPerThread;
  const size_t lds_bytes = (size_t)lds_n * (2 * sizeof(float) + sizeof(uint16_t) + 1);
  if (!ens_lds_ready(lds_bytes)) return JMD_ERR_LAUNCH;
  hipLaunchKernelGGL(k_fire_ensemble<true>, dim3(n_systems), dim3(threads), lds_bytes,
                     (hipStream_t)stream, reinterpret_cast<const EnsSys*>(systems), order,
                     n_systems, (int)n_total, mode, n_steps, position_hi, position, momentum,
                     force, reference_position, idx, width, scratch, a, prm, fire_state,
                     fire_n_pos, error, out_stat, energy, builds, tape, lds_n);
  JMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_fire_ensemble<false>, dim3(n_systems), dim3(threads), 0,
                     (hipStream_t)stream, reinterpret_cast<const EnsSys*>(systems), order,
                     n_systems, (int)n_total, mode, n_steps, position_hi, position, momentum,
                     force, reference_position, idx, width, scratch, a, prm, fire_state,
                     fire_n_pos, error, out_stat, energy, builds, tape, lds_n);
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

int jmd_fire_ensemble_max_cells(void) { return kEnsMaxCells; }

int jmd_fire_ensemble_set_threads(int32_t threads) {
  if (threads != 0 && threads != 256 && threads != 512 && threads != 1024) return JMD_ERR_INVALID;
  g_ens_threads = threads;
  return JMD_OK;
}

int jmd_fire_ensemble_adjoint_2d(const jmd_system2d* systems, int32_t n_systems,
                                 int64_t n_total, int32_t mode, int32_t n_steps,
                                 const float* position, const float* force, const int32_t* idx,
                                 int32_t width, const jmd_pair_potential* potential,
                                 const float* dsigma, const float* tape_x,
                                 const float* tape_force, const float* tape_q,
                                 const int32_t* tape_idx, const float* tape_scalars,
                                 double* x_adjoint, double* p_adjoint, float* pending,
                                 float* scratch, double* theta, double* box_adjoint,
                                 float scale, void* stream) {
  if (n_systems < 0 || n_total < 0 || n_total >= 0x7fffffff || width < 1 || n_steps < 0 ||
      mode < BACK_INIT || mode > BACK_FINISH)
    return JMD_ERR_INVALID;
  if (n_systems == 0) return JMD_OK;
  if (!systems || !x_adjoint || !p_adjoint || !pending || !scratch || !theta)
    return JMD_ERR_INVALID;
  if (mode == BACK_INIT && (!position || !force || !idx)) return JMD_ERR_INVALID;
  if (mode != BACK_INIT && (!tape_x || !tape_force || !tape_q || !tape_idx || !tape_scalars))
    return JMD_ERR_INVALID;
  PairArgs a;
  if (!make_pair_args(potential, &a) || a.kind != JMD_SOFT_SPHERE) return JMD_ERR_INVALID;
  if (dsigma && a.sig_mode != JMD_PARAM_SPECIES) return JMD_ERR_UNSUPPORTED;
  EnsTape tape{const_cast<float*>(tape_x), const_cast<float*>(tape_force),
               const_cast<float*>(tape_q), const_cast<int*>(tape_idx),
               const_cast<float*>(tape_scalars), n_steps};
  // LDS copy of each system's node positions, product vector and species
  // ids (systems of up to 8 particles per thread; larger ones use memory)
  const int threads = ens_threads(n_systems, n_total);
  const int lds_n = threads * kEnsPerThread;
  const size_t lds_bytes = (size_t)lds_n * (4 * sizeof(float) + 1);
  if (!ens_lds_ready_back(lds_bytes)) return JMD_ERR_LAUNCH;
  hipLaunchKernelGGL(k_fire_ensemble_back<true>, dim3(n_systems), dim3(threads), lds_bytes,
                     (hipStream_t)stream, reinterpret_cast<const EnsSys*>(systems), n_systems,
                     (int)n_total, mode, n_steps, position, force, idx, width, a, dsigma, tape,
                     x_adjoint, p_adjoint, pending, scratch, theta, box_adjoint, scale, lds_n);
  JMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_fire_ensemble_back<false>, dim3(n_systems), dim3(threads), 0,
                     (hipStream_t)stream, reinterpret_cast<const EnsSys*>(systems), n_systems,
                     (int)n_total, mode, n_steps, position, force, idx, width, a, dsigma, tape,
                     x_adjoint, p_adjoint, pending, scratch, theta, box_adjoint, scale, lds_n);
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

}  // extern "C"
