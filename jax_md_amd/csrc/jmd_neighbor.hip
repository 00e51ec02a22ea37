// Cell binning and neighbour-list build for gfx950.
//
// Reference: jax_md/partition.py (cell_list_fn :349-471, neighbor_list
// :801-1164).  The reference materialises [N, 3^d * cap] candidate arrays and
// compacts them with cumsum + scatter; here the candidates never leave LDS:
//
//   1. bin:     cell hash per particle, counts by atomics   (k_bin_count)
//   2. scan:    exclusive scan of the counts in one workgroup, max filling,
//               cell-overflow bit                           (k_bin_scan)
//   3. scatter: particle -> cell-ordered slot               (k_bin_scatter)
//   4. order:   one wavefront per cell sorts its members by id, so the
//               within-cell order equals the reference's stable argsort and
//               the build is deterministic                   (k_cell_order)
//   5. build:   one workgroup per home cell stages the 3^d stencil cells'
//               positions/ids in LDS in the reference's candidate order
//               (stencil offsets with the first component slowest, slots
//               rotated by start mod cap), each wavefront takes one home
//               particle at a time, lanes test 64 candidates at once, rows
//               are compacted with ballot + mbcnt            (k_nl_cells)
//   6. finish:  error bits, occupancy, gate reset            (k_nl_finish)
//
// The distance tests reproduce the reference's float32 arithmetic exactly
// (ref_metric_sq), so the neighbour sets and overflow bits are bit-exact.
#include "jmd_common.hpp"

namespace jmd {

// misc words inside the workspace
enum {
  M_OCC = 0,        // max row count (Dense) / total pairs (Sparse)
  M_FILL = 1,       // max cell filling
  M_FLAGS = 2,      // error bits raised by kernels
  M_TOTAL = 3,      // Sparse total (scan result)
  M_BUILDS = 4,     // completed gated/ungated builds (never reset)
  M_FAR = 5,        // some position lies >= 2 periods outside the box
  M_TICKET_BIN = 6, // finished workgroups of k_bin_slots (zero at rest)
  M_TICKET_NL = 7,  // finished workgroups of k_nl_cells (zero at rest)
  M_WORDS = 16
};

struct Work {
  int* cell_count;   // [ncell]   (zero at rest)
  int* cell_start;   // [ncell+1]
  int* atom_cell;    // [n]
  int* atom_rank;    // [n]
  int* sorted_tmp;   // [n]
  int* sorted_id;    // [n]
  float4* sorted_pos;// [n]
  int* row_count;    // [n]
  int* row_offset;   // [n+1]
  int* slots;        // [ncell * cap] ids per cell in arrival order (k_bin_slots)
  int* misc;         // [M_WORDS] (zero at rest except as documented)
};

#ifndef JMD_CELL_BLOCKS
#define JMD_CELL_BLOCKS 1024
#endif
#ifndef JMD_SORT_BLOCKS
#define JMD_SORT_BLOCKS 1024
#endif
constexpr int kCellBlocks = JMD_CELL_BLOCKS;  // persistent grid of the cell build
constexpr int kSortBlocks = JMD_SORT_BLOCKS;  // grid cap of the cell sort

static inline size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

static size_t carve(void* base, long long n, long long ncell, long long cap, Work* w) {
  size_t off = 0;
  char* b = (char*)base;
  auto take = [&](size_t bytes) -> char* {
    char* p = b ? b + off : nullptr;
    off += align_up(bytes);
    return p;
  };
  long long nc = ncell > 0 ? ncell : 1;
  Work t;
  t.misc = (int*)take(sizeof(int) * M_WORDS);
  t.cell_count = (int*)take(sizeof(int) * nc);
  t.cell_start = (int*)take(sizeof(int) * (nc + 1));
  t.atom_cell = (int*)take(sizeof(int) * n);
  t.atom_rank = (int*)take(sizeof(int) * n);
  t.sorted_tmp = (int*)take(sizeof(int) * n);
  t.sorted_id = (int*)take(sizeof(int) * n);
  t.sorted_pos = (float4*)take(sizeof(float4) * n);
  t.row_count = (int*)take(sizeof(int) * n);
  t.row_offset = (int*)take(sizeof(int) * (n + 1));
  t.slots = (int*)take(sizeof(int) * (ncell > 0 && cap > 0 ? nc * cap : 1));
  if (w) *w = t;
  return off;
}

// ---------------------------------------------------------------------------
// Phases, written as device functions over a (block, grid) decomposition.
// ---------------------------------------------------------------------------

template <int DIM, int STRIDE>
__device__ __forceinline__ void dev_bin_count(const float* __restrict__ pos, int n,
                                              const Grid& g, int* __restrict__ cell_count,
                                              int* __restrict__ atom_cell,
                                              int* __restrict__ atom_rank,
                                              float* __restrict__ ref_out,
                                              int* __restrict__ cell_id_out, int t0,
                                              int tstride) {
  for (int i = t0; i < n; i += tstride) {
    float x[3];
    load_pos<DIM, STRIDE>(pos, i, x);
    int ci[3];
    int h = cell_of<DIM>(x, g, ci);
    atom_cell[i] = h;
    atom_rank[i] = atomicAdd(&cell_count[h], 1);
    if (ref_out) store_pos<DIM>(ref_out, i, x);
    if (cell_id_out) {
#pragma unroll
      for (int d = 0; d < DIM; ++d) cell_id_out[(long long)i * DIM + d] = ci[d];
    }
  }
}

// One workgroup: exclusive scan of cell_count into cell_start[0..ncell], max
// filling, cell-overflow flag, and (reset_counts) cell_count reset to zero.
// The counts go through in chunks of blockDim * 16 cells; each thread holds
// 16 consecutive counts (four 16-byte loads in flight at once), so a chunk
// costs one memory round trip.
constexpr int kScanPer = 16;

__device__ __forceinline__ int coherent_load(const int* p);

__device__ __forceinline__ void dev_bin_scan(int* __restrict__ cell_count,
                                             int* __restrict__ cell_start, int ncell, int cap,
                                             int* __restrict__ misc, bool reset_counts = true,
                                             bool coherent = false) {
  __shared__ int s_wsum[1024 / kWave];
  __shared__ int s_max[1024 / kWave];
  const int t = threadIdx.x, nt = blockDim.x;
  const int lane = lane_id(), wv = t / kWave, nw = nt / kWave;
  int carry = 0, lmax = 0;
  for (int c0 = 0; c0 < ncell; c0 += nt * kScanPer) {
    const int beg = c0 + t * kScanPer;
    int v[kScanPer];
    if (coherent) {  // counts produced by atomics of the same kernel
#pragma unroll
      for (int k = 0; k < kScanPer; ++k)
        v[k] = beg + k < ncell ? coherent_load(cell_count + beg + k) : 0;
    } else if (beg + kScanPer <= ncell) {
      const int4* p = reinterpret_cast<const int4*>(cell_count + beg);
#pragma unroll
      for (int k = 0; k < kScanPer / 4; ++k) {
        int4 q = p[k];
        v[4 * k] = q.x; v[4 * k + 1] = q.y; v[4 * k + 2] = q.z; v[4 * k + 3] = q.w;
      }
    } else {
#pragma unroll
      for (int k = 0; k < kScanPer; ++k) v[k] = beg + k < ncell ? cell_count[beg + k] : 0;
    }
    int local = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
      local += v[k];
      lmax = max(lmax, v[k]);
    }
    int incl = local;  // inclusive scan over the wavefront
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      int y = __shfl_up(incl, off, kWave);
      if (lane >= off) incl += y;
    }
    if (lane == kWave - 1) s_wsum[wv] = incl;
    __syncthreads();
    int wbase = 0, ctot = 0;
    for (int k = 0; k < nw; ++k) {
      const int ws = s_wsum[k];
      wbase += k < wv ? ws : 0;
      ctot += ws;
    }
    __syncthreads();  // s_wsum is rewritten by the next chunk
    int run = carry + wbase + incl - local;
    if (beg + kScanPer <= ncell) {
      int4* ps = reinterpret_cast<int4*>(cell_start + beg);
#pragma unroll
      for (int k = 0; k < kScanPer / 4; ++k) {
        int4 q;
        q.x = run; run += v[4 * k];
        q.y = run; run += v[4 * k + 1];
        q.z = run; run += v[4 * k + 2];
        q.w = run; run += v[4 * k + 3];
        ps[k] = q;
      }
      if (reset_counts) {
        int4* pc = reinterpret_cast<int4*>(cell_count + beg);
#pragma unroll
        for (int k = 0; k < kScanPer / 4; ++k) pc[k] = make_int4(0, 0, 0, 0);
      }
    } else {
#pragma unroll
      for (int k = 0; k < kScanPer; ++k) {
        if (beg + k < ncell) {
          cell_start[beg + k] = run;
          if (reset_counts) cell_count[beg + k] = 0;
        }
        run += v[k];
      }
    }
    carry += ctot;
  }
  if (t == 0) cell_start[ncell] = carry;
  const int wmax = wave_reduce_max(lmax);
  if (lane == 0) s_max[wv] = wmax;
  __syncthreads();
  if (t == 0) {
    int m = 0;
    for (int k = 0; k < nw; ++k) m = max(m, s_max[k]);
    misc[M_FILL] = m;
    if (cap > 0 && m > cap) atomicOr(&misc[M_FLAGS], JMD_CELL_LIST_OVERFLOW);
  }
  __syncthreads();
}

__device__ __forceinline__ void dev_bin_scatter(int n, const int* __restrict__ atom_cell,
                                                const int* __restrict__ atom_rank,
                                                const int* __restrict__ cell_start,
                                                int* __restrict__ sorted_tmp, int t0,
                                                int tstride) {
  for (int i = t0; i < n; i += tstride) sorted_tmp[cell_start[atom_cell[i]] + atom_rank[i]] = i;
}

// One wavefront per cell: order members by id (the stable argsort of
// partition.py:432 keeps ids ascending inside a cell), gather positions.
// sorted_pos holds each position moved by the whole box periods implied by
// its cell assignment (x - L * floor(int(x / cs) / cps)), so that positions
// of a cell lie within that cell's range, and the id's bits in .w; the raw
// input stays untouched.
template <int DIM, int STRIDE>
__device__ __forceinline__ void dev_cell_order(const float* __restrict__ pos, const Grid& g,
                                               const Space& s,
                                               const int* __restrict__ cell_start,
                                               const int* __restrict__ sorted_tmp,
                                               int* __restrict__ sorted_id,
                                               float4* __restrict__ sorted_pos,
                                               int* __restrict__ misc, int w0, int wstride) {
  const int lane = lane_id();
  bool far = false;
  for (int cell = w0; cell < g.ncell; cell += wstride) {
    int beg = cell_start[cell];
    int cnt = cell_start[cell + 1] - beg;
    for (int a = lane; a < cnt; a += kWave) {
      int id = sorted_tmp[beg + a];
      int rank = 0;
      for (int b = 0; b < cnt; ++b) rank += (sorted_tmp[beg + b] < id) ? 1 : 0;
      sorted_id[beg + rank] = id;
      float x[3] = {0.0f, 0.0f, 0.0f};
      load_pos<DIM, STRIDE>(pos, id, x);
      if (s.periodic) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
          int q = (int)__fdiv_rn(x[d], g.cell_size[d]);
          int wraps = (q - floor_mod_int(q, g.cps[d])) / g.cps[d];
          if (wraps != 0) x[d] = x[d] - (float)wraps * s.L[d];
          // x - wraps * L rounds at the scale of |x| once |wraps| >= 2; the
          // build then skips its fast test (see dev_nl_cells).
          if (wraps > 1 || wraps < -1) far = true;
        }
      }
      sorted_pos[beg + rank] = make_float4(x[0], x[1], x[2], __int_as_float(id));
    }
  }
  if (__ballot(far) != 0ull && lane == 0) atomicOr(&misc[M_FAR], 1);
}

template <int DIM, int STRIDE>
__global__ __launch_bounds__(256) void k_bin_count(
    const float* __restrict__ pos, int n, Grid g, int* __restrict__ cell_count,
    int* __restrict__ atom_cell, int* __restrict__ atom_rank,
    float* __restrict__ ref_out, int* __restrict__ cell_id_out, const int* gate) {
  if (gate_closed(gate)) return;
  dev_bin_count<DIM, STRIDE>(pos, n, g, cell_count, atom_cell, atom_rank, ref_out, cell_id_out,
                             blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

__global__ __launch_bounds__(1024) void k_bin_scan(int* __restrict__ cell_count,
                                                   int* __restrict__ cell_start,
                                                   int ncell, int cap,
                                                   int* __restrict__ misc,
                                                   const int* gate) {
  if (gate_closed(gate)) return;
  dev_bin_scan(cell_count, cell_start, ncell, cap, misc);
}

__global__ __launch_bounds__(256) void k_bin_scatter(
    int n, const int* __restrict__ atom_cell, const int* __restrict__ atom_rank,
    const int* __restrict__ cell_start, int* __restrict__ sorted_tmp, const int* gate) {
  if (gate_closed(gate)) return;
  dev_bin_scatter(n, atom_cell, atom_rank, cell_start, sorted_tmp,
                  blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

template <int DIM, int STRIDE>
__global__ __launch_bounds__(256) void k_cell_order(
    const float* __restrict__ pos, int ncell, Grid g, Space s,
    const int* __restrict__ cell_start, const int* __restrict__ sorted_tmp,
    int* __restrict__ sorted_id, float4* __restrict__ sorted_pos, int* __restrict__ misc,
    const int* gate) {
  if (gate_closed(gate)) return;
  dev_cell_order<DIM, STRIDE>(pos, g, s, cell_start, sorted_tmp, sorted_id, sorted_pos, misc,
                              (blockIdx.x * blockDim.x + threadIdx.x) / kWave,
                              gridDim.x * blockDim.x / kWave);
}

// Public cell-list buffers (partition.py:441-456): slot of the k-th particle
// in (cell, id) order is hash*cap + (k mod cap).
template <int DIM, int STRIDE>
__global__ __launch_bounds__(256) void k_cell_emplace(
    const float* __restrict__ pos, int n, int cap, const int* __restrict__ cell_start,
    const int* __restrict__ sorted_id, int ncell, float* __restrict__ pos_buf,
    int* __restrict__ id_buf) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  // cell of sorted entry k: binary search over cell_start
  int lo = 0, hi = ncell;  // cell_start[lo] <= k < cell_start[hi]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (cell_start[mid] <= k) lo = mid; else hi = mid;
  }
  long long slot = (long long)lo * cap + (k % cap);
  int id = sorted_id[k];
  id_buf[slot] = id;
  float x[3];
  load_pos<DIM, STRIDE>(pos, id, x);
  store_pos<DIM>(pos_buf, slot, x);
}

// ---------------------------------------------------------------------------
// Neighbour-list build over the cell grid
// ---------------------------------------------------------------------------

enum { MODE_DENSE = 0, MODE_SPARSE_COUNT = 1, MODE_SPARSE_WRITE = 2, MODE_COUNT_MAX = 3 };
enum { TEST_IJ = 0, TEST_JI = 1, TEST_BOTH = 2 };

#ifndef JMD_NL_STAGE
#define JMD_NL_STAGE 1536
#endif
constexpr int kStage = JMD_NL_STAGE;  // candidates staged in LDS per pass
#ifndef JMD_NL_HOME_GROUP
#define JMD_NL_HOME_GROUP 512
#endif
constexpr int kHomeGroup = JMD_NL_HOME_GROUP;  // home particles per candidate staging
#ifndef JMD_NL_HOME_PASS
#define JMD_NL_HOME_PASS 4
#endif
constexpr int kHomePass = JMD_NL_HOME_PASS;  // home particles per candidate pass
constexpr int kNlBlock = 256;                 // k_nl_cells workgroup size
constexpr int kNlWaves = kNlBlock / kWave;
constexpr int kStageQ = (kStage + kNlBlock - 1) / kNlBlock;  // staging loads per lane

template <int DIM, class G>
__device__ __forceinline__ bool pair_test(const float* xi, const float* xj,
                                          const G& s, float cutoff_sq,
                                          float margin, int test) {
  if (test == TEST_IJ) return ref_metric_sq<DIM>(xi, xj, s) < cutoff_sq;
  if (test == TEST_JI) return ref_metric_sq<DIM>(xj, xi, s) < cutoff_sq;
  // Both directions (Dense with a cell list, partition.py:951,971).  The two
  // square distances differ by a few float32 roundings, so with margin > 0 the
  // second one is only evaluated when the first lies within `margin` of the
  // cutoff; margin = 0 always evaluates both.
  float a = ref_metric_sq<DIM>(xi, xj, s);
  bool pa = a < cutoff_sq;
  if (margin > 0.0f && fabsf(a - cutoff_sq) > margin) return pa;
  float b = ref_metric_sq<DIM>(xj, xi, s);
  return pa && (b < cutoff_sq);
}

// pair_test on the original positions of particles i and j, kept out of line
// for the cell build: it runs only for candidates within the margin of the
// cutoff, and inlined its jnp.mod emulation would hold registers across the
// whole candidate loop.
template <int DIM, int STRIDE>
__device__ __noinline__ bool exact_pair_test(const float* __restrict__ pos, int i, int j,
                                             SpaceO s, float cutoff_sq, int test) {
  float xo[3] = {0.0f, 0.0f, 0.0f}, xj[3] = {0.0f, 0.0f, 0.0f};
  load_pos<DIM, STRIDE>(pos, i, xo);
  load_pos<DIM, STRIDE>(pos, j, xj);
  return pair_test<DIM>(xo, xj, s, cutoff_sq, 0.0f, test);
}

// periodic_general (box matrix): inlined, in the folding build variant only.
template <int DIM, int STRIDE>
__device__ __forceinline__ bool exact_pair_test_general(const float* __restrict__ pos, int i,
                                                        int j, const Space& s, float cutoff_sq,
                                                        int test) {
  float xo[3] = {0.0f, 0.0f, 0.0f}, xj[3] = {0.0f, 0.0f, 0.0f};
  load_pos<DIM, STRIDE>(pos, i, xo);
  load_pos<DIM, STRIDE>(pos, j, xj);
  return pair_test<DIM>(xo, xj, s, cutoff_sq, 0.0f, test);
}

// Arguments of the cell-grid build.
struct BuildArgs {
  const float* pos;
  int n;
  Grid g;
  Space s;
  float cutoff_sq;
  float margin;
  int test;
  int mask_self;
  int ordered;
  int* idx;
  long long width;
  float* ref_out;
  int* occupancy;
  uint8_t* error;
  int* gate;
  Work w;
  KickDrift kd;  // fused opening half-step (k_bin_slots), inactive by default
};

// Candidates are staged as float4 (x, y, z, id bits) holding the periodic
// image nearest to the home cell: the shift of each stencil cell (0 or +-L
// per dimension) follows from the cell topology.  The common test is then a
// plain |x_i - x_j'|^2 against c^2 -/+ margin, where margin bounds the
// difference between this value and the reference's wrapped float32 d^2;
// only candidates within the margin run the reference-exact test on the
// original positions (pair_test), so sets and order equal the reference's.
// GEOM 0: staged images; 1 (MINIMG): some dimension has fewer than 5 cells,
// where the staged image is
// not always the nearest one and a minimum-image fold is applied.
template <int DIM, int STRIDE, int MODE, int GEOM>
__device__ int dev_nl_cells(const BuildArgs& A, int home0, int home_stride) {
  // GEOM 2: periodic_general with 5 or more cells a side: staged images in
  // fractional units, T applied to each difference, pruning by T's smallest
  // singular value
  constexpr bool MINIMG = GEOM == 1;
  constexpr bool BOXS = GEOM == 2;
  constexpr int NST = DIM == 3 ? 27 : 9;
  __shared__ int s_cs[NST], s_cc[NST], s_off[NST + 1], s_rot[NST];
  __shared__ float s_shift[NST][3];
  __shared__ float4 s_pos[kStage];
  __shared__ int s_cnt[kHomeGroup];
  __shared__ float4 s_home[kHomeGroup];  // the home group's sorted positions (id bits in w)
  __shared__ int s_occ;
  __shared__ int s_wc[kStageQ * kNlWaves];  // kept candidates per load round and wavefront

  const Grid& g = A.g;
  const Space& s = A.s;
  const SpaceO so = ortho_of(A.s);
  const int n = A.n;
  const long long width = A.width;
  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int wave = tid / kWave;
  const int nwave = blockDim.x / kWave;
  // Positions canonicalised with a rounding beyond the margin's bound: every
  // candidate takes the reference-exact test.
  const float margin = A.w.misc[M_FAR] ? __builtin_inff() : A.margin;
  // with an infinite margin every real candidate is "near"; the finite upper
  // bound keeps the staged padding (d^2 > 2e37) out of the exact test
  const float lo_thr = A.cutoff_sq - margin;
  const float hi_thr = fminf(A.cutoff_sq + margin, 1.0e36f);
  const float far = 3.0e18f;  // staged padding: never within the cutoff
  int occ_local = 0;          // largest row count seen by this wavefront

  for (int home = home0; home < g.ncell; home += home_stride) {
    __syncthreads();  // LDS reuse across home cells
    int ccell = 0;  // members of stencil cell tid (tid < NST)
    if (tid < NST) {
      int hc[3], h = home;
#pragma unroll
      for (int d = 0; d < DIM; ++d) { hc[d] = h % g.cps[d]; h /= g.cps[d]; }
      int off[3];
      int r = tid;
#pragma unroll
      for (int d = DIM - 1; d >= 0; --d) { off[d] = r % 3 - 1; r /= 3; }
      int c = 0, mult = 1;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        float sh = 0.0f;
        if (d < DIM) {
          int raw = hc[d] + off[d];
          c += floor_mod_int(raw, g.cps[d]) * mult;
          mult *= g.cps[d];
          if (s.periodic) sh = raw < 0 ? -s.L[d] : (raw >= g.cps[d] ? s.L[d] : 0.0f);
        }
        s_shift[tid][d] = sh;
      }
      const int cs0 = A.w.cell_start[c];
      ccell = A.w.cell_start[c + 1] - cs0;
      // k_cell_sort orders at most cap members of a cell (more: overflow)
      if (g.cap > 0) ccell = min(ccell, g.cap);
      // the slot rotation of a cell's members (start mod cap, the cell
      // slots of partition.py): zero when the members do not wrap
      int rot = 0;
      if (g.cap > 0) {
        const int w = g.cap - (cs0 % g.cap);
        rot = ccell > w ? w : 0;
      }
      s_cs[tid] = cs0;
      s_cc[tid] = ccell;
      s_rot[tid] = rot;
    }
    if constexpr (GEOM == 0) {
      if (tid < kWave) {
        // candidate offsets of the stencil cells: an exclusive scan of the
        // counts over the first 32 lanes (lane NST, count 0, gets the total)
        int incl = ccell;
#pragma unroll
        for (int o = 1; o < 32; o <<= 1) {
          const int u = __shfl_up(incl, o, kWave);
          if (lane >= o) incl += u;
        }
        if (tid <= NST) s_off[tid] = incl - ccell;
      }
    } else {
      // (the folding and general-box builds have fewer registers to spare:
      // one lane sums the counts)
      __syncthreads();
      if (tid == 0) {
        int acc = 0;
        for (int t = 0; t < NST; ++t) {
          s_off[t] = acc;
          acc += s_cc[t];
        }
        s_off[NST] = acc;
      }
    }
    __syncthreads();
    const int total = s_off[NST];
    // The home cell's box in the staged frame, widened for the truncating
    // cell index (x in (-cs, 0) belongs to cell 0) and its rounding: a
    // candidate farther than sqrt(c^2 + margin) from it reaches no home
    // particle by either test and is not staged.
    // (folding variant: never; general boxes: through T's smallest singular
    // value, |T v| >= sqrt(smin2) |v|)
    const bool prune = !MINIMG && s.periodic && margin < 3.0e38f;
    float blo[3] = {0.0f, 0.0f, 0.0f}, bhi[3] = {0.0f, 0.0f, 0.0f};
    {
      int h = home;
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const int c = h % g.cps[d];
        h /= g.cps[d];
        const float cs = g.cell_size[d], dl = 1e-4f * cs;
        blo[d] = (c == 0 ? -cs : (float)c * cs) - dl;
        bhi[d] = (float)(c + 1) * cs + dl;
      }
    }
    const float reach_sq = (A.cutoff_sq + margin) * 1.001f;
    const int hstart = s_cs[NST / 2];
    const int hcount = s_cc[NST / 2];

    for (int g0 = 0; g0 < hcount; g0 += kHomeGroup) {
      const int gcount = min(kHomeGroup, hcount - g0);
      for (int a = tid; a < gcount; a += blockDim.x) {
        s_cnt[a] = 0;
        s_home[a] = A.w.sorted_pos[hstart + g0 + a];  // read by the passes and the epilogue
      }
      for (int c0 = 0; c0 < total; c0 += kStage) {
        const int ccount = min(kStage, total - c0);
        __syncthreads();
        // Stage candidates c0..c0+ccount in reference order, keeping those
        // that can reach the home cell (compacted in order), padded to a
        // multiple of 64 with far-away entries.  Every lane's loads (kStageQ
        // per lane at most) are issued before any is used, then one
        // compaction over (load round, wavefront, lane) keeps the order.
        // Stage candidates c0..c0+ccount in reference order, keeping those
        // that can reach the home cell (compacted in order), padded to a
        // multiple of 64 with far-away entries.  A lane's loads of G rounds
        // (G = all of them for the staged-image build, one for the folding
        // and general-box builds, whose registers are tighter) are issued
        // before any is used, then one compaction over (round, wavefront,
        // lane) keeps the order.
        constexpr int G = GEOM == 0 ? kStageQ : 1;
        int kept = 0;  // workgroup-uniform
        for (int r0 = 0; r0 * kNlBlock < ccount; r0 += G) {
          float4 v[G];
          int vt[G];
#pragma unroll
          for (int u = 0; u < G; ++u) {
            const int q = (r0 + u) * kNlBlock + tid;
            v[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            vt[u] = 0;
            if (q < ccount) {
              const int p = c0 + q;
              // the stencil cell of candidate p: the last t with s_off[t] <= p
              // (empty cells have s_off[t] == s_off[t + 1] and are passed over)
              int t = 0;
#pragma unroll
              for (int step = 16; step > 0; step >>= 1)
                if (t + step < NST && s_off[t + step] <= p) t += step;
              const int m = p - s_off[t];
              int k = m;
              const int w = s_rot[t];
              if (w > 0) k = (m < s_cc[t] - w) ? (w + m) : (m - (s_cc[t] - w));
              v[u] = A.w.sorted_pos[s_cs[t] + k];  // w: the id bits
              vt[u] = t;
            }
          }
          unsigned keepm = 0;
#pragma unroll
          for (int u = 0; u < G; ++u) {
            const int q = (r0 + u) * kNlBlock + tid;
            if (q < ccount) {
              const int t = vt[u];
              v[u].x += s_shift[t][0];
              v[u].y += s_shift[t][1];
              v[u].z += s_shift[t][2];
              bool keep = true;
              if (prune) {
                const float cv[3] = {v[u].x, v[u].y, v[u].z};
                float D2 = 0.0f;
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                  const float e = fmaxf(0.0f, fmaxf(blo[d] - cv[d], cv[d] - bhi[d]));
                  D2 = fmaf(e, e, D2);
                }
                keep = (BOXS ? D2 * s.smin2 : D2) <= reach_sq;
              }
              if (keep) keepm |= 1u << u;
            }
          }
#pragma unroll
          for (int u = 0; u < G; ++u) {
            const unsigned long long bal = __ballot((keepm >> u) & 1u);
            if (lane == 0) s_wc[u * kNlWaves + wave] = (int)__popcll(bal);
          }
          __syncthreads();
#pragma unroll
          for (int u = 0; u < G; ++u) {
            const unsigned long long bal = __ballot((keepm >> u) & 1u);
            int before = 0, all = 0;
#pragma unroll
            for (int w = 0; w < kNlWaves; ++w) {
              const int c = s_wc[u * kNlWaves + w];
              before += w < wave ? c : 0;
              all += c;
            }
            if ((keepm >> u) & 1u) s_pos[kept + before + wave_prefix_count(bal)] = v[u];
            kept += all;
          }
          if (r0 + G < kStageQ) __syncthreads();  // s_wc is rewritten by the next group
        }
        const int cpad = (kept + kWave - 1) & ~(kWave - 1);
        for (int q = kept + tid; q < cpad; q += blockDim.x)
          s_pos[q] = make_float4(far, far, far, __int_as_float(n));
        __syncthreads();
        // Each wavefront takes kHomePass home particles at a time: one LDS
        // read of 64 candidates serves all of them, and the loop overhead is
        // shared.  Home particles are wavefront-uniform.
        for (int a0 = wave; a0 < gcount; a0 += kHomePass * nwave) {
          float xi[kHomePass][3];
          int ih[kHomePass], cnt[kHomePass];
          // allowed candidates: j < lim and j != ex (OrderedSparse keeps
          // j < i only; mask_self drops j = i)
          int lim[kHomePass], ex[kHomePass];
          long long base[kHomePass];
#pragma unroll
          for (int h = 0; h < kHomePass; ++h) {
            const int a = a0 + h * nwave;
            xi[h][0] = xi[h][1] = xi[h][2] = 0.0f;
            ih[h] = -1;
            lim[h] = 0;
            ex[h] = -1;
            cnt[h] = 0;
            base[h] = 0;
            if (a < gcount) {
              const float4 xp = s_home[a];
              xi[h][0] = xp.x;
              xi[h][1] = xp.y;
              xi[h][2] = xp.z;
              ih[h] = __float_as_int(xp.w);
              lim[h] = A.ordered ? ih[h] : 0x7fffffff;
              ex[h] = A.mask_self ? ih[h] : -1;
              cnt[h] = s_cnt[a];
              if (MODE == MODE_DENSE) base[h] = (long long)ih[h] * width;
              if (MODE == MODE_SPARSE_WRITE) base[h] = A.w.row_offset[ih[h]];
            }
          }
          // homes of this pass that exist (wavefront-uniform, a scalar bound)
          const int nh = min(kHomePass, (gcount - a0 + nwave - 1) / nwave);
          float4 cnext = s_pos[lane];  // cpad >= 64 when this loop runs
          for (int q0 = 0; q0 < cpad; q0 += kWave) {
            const float4 c = cnext;
            if (q0 + kWave < cpad) cnext = s_pos[q0 + kWave + lane];  // next chunk, ahead
            // staged padding sits 3e18 away: its d2 fails both tests below
            const int j = __float_as_int(c.w);
            const float cj[3] = {c.x, c.y, c.z};
#pragma unroll
            for (int h = 0; h < kHomePass; ++h) {
              if (h >= nh) break;
              const int i = ih[h];
              float d[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
              for (int e = 0; e < DIM; ++e)
                d[e] = MINIMG ? -min_image_disp(cj[e], xi[h][e], xi[h][e] - s.L[e], s.L[e],
                                                s.half[e])
                              : xi[h][e] - cj[e];
              // periodic_general: the fold (or staged image) is in fractional
              // units, then T
              if (BOXS || (MINIMG && s.general)) box_apply<DIM>(s, d);
              const float d2 = DIM == 3 ? fmaf(d[2], d[2], fmaf(d[1], d[1], d[0] * d[0]))
                                        : fmaf(d[1], d[1], d[0] * d[0]);
              // Dense rows never drop j > i: only the Sparse modes read lim
              constexpr bool kLim = MODE != MODE_DENSE && MODE != MODE_COUNT_MAX;
              const bool okj = (!kLim || j < lim[h]) && j != ex[h];
              const bool fast = d2 < lo_thr;
              bool pass = fast && okj;
              // ballots of comparisons stay lane masks (no VGPR round trip)
              const unsigned long long m_fast = __builtin_amdgcn_ballot_w64(d2 < lo_thr);
              unsigned long long bal = m_fast & __builtin_amdgcn_ballot_w64(j != ex[h]);
              if (kLim) bal &= __builtin_amdgcn_ballot_w64(j < lim[h]);
              // within the margin of the cutoff: too close to call from the
              // canonicalised positions (rare: reference-exact test)
              const unsigned long long near_any =
                  __builtin_amdgcn_ballot_w64(d2 <= hi_thr) & ~m_fast;
              if (near_any) {
                if (!fast && d2 <= hi_thr && okj)
                  pass = (BOXS || (MINIMG && s.general))
                             ? exact_pair_test_general<DIM, STRIDE>(A.pos, i, j, s, A.cutoff_sq,
                                                                    A.test)
                             : exact_pair_test<DIM, STRIDE>(A.pos, i, j, so, A.cutoff_sq,
                                                            A.test);
                bal = __builtin_amdgcn_ballot_w64(pass);
              }
              if ((MODE == MODE_DENSE || MODE == MODE_SPARSE_WRITE) && pass) {
                if (MODE == MODE_DENSE) {
                  // row slots stay below width < 2^31: 32-bit unsigned offsets
                  const unsigned slot = (unsigned)cnt[h] + (unsigned)wave_prefix_count(bal);
                  if (slot < (unsigned)width) A.idx[base[h] + slot] = j;
                } else {
                  const long long slot = cnt[h] + wave_prefix_count(bal);
                  const long long at = base[h] + slot;
                  if (at < width) {
                    A.idx[at] = j;          // receiver
                    A.idx[width + at] = i;  // sender
                  }
                }
              }
              cnt[h] += __popcll(bal);
            }
          }
#pragma unroll
          for (int h = 0; h < kHomePass; ++h) {
            const int a = a0 + h * nwave;
            if (a < gcount && lane == 0) s_cnt[a] = cnt[h];
          }
        }
      }
      __syncthreads();
      // Row epilogue: padding (Dense), counts, occupancy.
      for (int a = wave; a < gcount; a += nwave) {
        const int i = __float_as_int(s_home[a].w);
        const int cnt = s_cnt[a];
        if (MODE == MODE_DENSE) {
          for (long long k = (long long)cnt + lane; k < width; k += kWave)
            A.idx[(long long)i * width + k] = n;
        } else if (MODE == MODE_SPARSE_COUNT) {
          if (lane == 0) A.w.row_count[i] = cnt;
        }
        occ_local = max(occ_local, cnt);
      }
      __syncthreads();  // s_cnt is reused by the next home group
    }
  }
  // One device-scope atomic per workgroup (its return value orders the
  // workgroup's completion ticket after it, see last_workgroup).
  int ret = 0;
  if (MODE == MODE_DENSE || MODE == MODE_COUNT_MAX) {
    if (tid == 0) s_occ = 0;
    __syncthreads();
    if (lane == 0) atomicMax(&s_occ, occ_local);
    __syncthreads();
    if (tid == 0) ret = atomicMax(&A.w.misc[M_OCC], s_occ);
  }
  return ret;
}

// ---------------------------------------------------------------------------
// Three-launch cell build of the neighbour lists:
//   k_bin_slots:  each particle takes a slot of its cell (wave-aggregated
//                 atomics); the last workgroup scans the counts into cell
//                 starts (the sorted-array offsets of the reference).
//   k_cell_sort:  one wavefront per cell orders the cell's ids (the stable
//                 argsort order) into sorted_id / sorted_pos.
//   k_nl_cells:   the build; its last workgroup finishes (error bits,
//                 occupancy, gate reset).
// ---------------------------------------------------------------------------

// Completion ticket: returns true (in all threads) in the workgroup that
// finishes last.  The last workgroup reads only values written by
// device-scope atomics, which are performed at the device's coherence point;
// each workgroup issues its ticket after its atomics have returned (`dep` is
// the value returned by thread 0's last one, and every other wave consumed
// its own returns before the barrier), so no cache write-back fence is needed
// (an agent-scope release fence writes back the whole L2 of the XCD).
__device__ __forceinline__ bool last_workgroup(int* ticket, int dep) {
  __shared__ int s_last;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned inc = dep == (int)0x80000000 ? 2u : 1u;  // never 2: orders after dep
    const unsigned t = atomicAdd(reinterpret_cast<unsigned*>(ticket), inc);
    s_last = (t == gridDim.x - 1) ? 1 : 0;
  }
  __syncthreads();
  return s_last != 0;
}

__device__ __forceinline__ int coherent_load(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int DIM, int STRIDE>
__global__ __launch_bounds__(256) void k_bin_slots(const float* __restrict__ pos, int n,
                                                   Grid g, int* __restrict__ cell_count,
                                                   int* __restrict__ cell_start,
                                                   int* __restrict__ slots,
                                                   float* __restrict__ ref_out,
                                                   int* __restrict__ misc, const int* gate,
                                                   KickDrift kd) {
  // kd.active: every step also applies the opening half of the next
  // velocity-Verlet step to each particle it reads, R' into kd.x_out (a
  // second position buffer, so the binning and the later list stages still
  // read R), and raises kd.gate_next when R' moved beyond the skin from the
  // list's reference positions (R itself when this call rebuilds).
  const bool open = !gate_closed(gate);
  if (!open && !kd.active) return;
  float kdt = kd.dt, khalf = kd.half_dt;
  if (kd.active && kd.dt_dev) {
    kdt = kd.dt_dev[0];
    khalf = kdt / 2.0f;
  }
  bool moved = false;
  const int lane = lane_id();
  for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const int i = base + threadIdx.x;
    const bool valid = i < n;
    int h = 0;
    if (valid) {
      float x[3];
      load_pos<DIM, STRIDE>(pos, i, x);
      if (open) {
        int ci[3];
        h = cell_of<DIM>(x, g, ci);
        if (ref_out) store_pos<DIM>(ref_out, i, x);
      }
      if (kd.active) {
        float xn[3], r[3];
        kick_drift_atom<DIM>(kd, i, kdt, khalf, xn);
        store_pos<DIM, STRIDE>(kd.x_out, i, xn);
        if (kd.gate_next) {
          if (open) {
#pragma unroll
            for (int d = 0; d < DIM; ++d) r[d] = x[d];
          } else {
            load_pos<DIM>(kd.ref, i, r);
          }
          moved = moved || ref_metric_sq<DIM>(xn, r, kd.s) > kd.thr_sq;
        }
      }
    }
    if (!open) continue;  // wavefront-uniform
    // one atomic per distinct cell in the wavefront (few when particles are
    // stored in cell order), all issued by one instruction: first group the
    // lanes by cell (leader = lowest lane of the group), then the leaders add
    // their group sizes, then every lane offsets its leader's result
    unsigned long long todo = __ballot(valid);
    int my_leader = lane, my_off = 0, grp = 0;
    while (todo) {
      const int leader = __ffsll(todo) - 1;
      const int hl = __shfl(h, leader);
      const unsigned long long peers = __ballot(valid && h == hl) & todo;
      if ((peers >> lane) & 1ull) {
        my_leader = leader;
        my_off = (int)__popcll(peers & ((1ull << lane) - 1ull));
      }
      if (lane == leader) grp = (int)__popcll(peers);
      todo &= ~peers;
    }
    int b = 0;
    if (valid && lane == my_leader) b = atomicAdd(&cell_count[h], grp);
    b = __shfl(b, my_leader);
    if (valid) {
      const int rank = b + my_off;
      if (rank < g.cap) slots[(long long)h * g.cap + rank] = i;
    }
  }
  if (kd.active && kd.gate_next) {
    const unsigned long long bal = __ballot(moved);
    if (bal != 0ull && lane == 0) atomicOr(kd.gate_next, 1);
  }
  if (!open) return;
  // every count was added by the device-scope atomics above
  if (!last_workgroup(&misc[M_TICKET_BIN], 0)) return;
  if (threadIdx.x == 0) misc[M_TICKET_BIN] = 0;
  dev_bin_scan(cell_count, cell_start, g.ncell, g.cap, misc, true, /*coherent=*/true);
}

// One wavefront per cell: ranks by id (a lane's id against every member,
// broadcast with readlane), canonical float4 position at that rank.
template <int DIM, int STRIDE>
__global__ __launch_bounds__(256) void k_cell_sort(const float* __restrict__ pos, Grid g,
                                                   Space s, const int* __restrict__ cell_start,
                                                   const int* __restrict__ slots,
                                                   int* __restrict__ sorted_id,
                                                   float4* __restrict__ sorted_pos,
                                                   int* __restrict__ misc, const int* gate) {
  if (gate_closed(gate)) return;
  const int lane = lane_id();
  const int nw = gridDim.x * blockDim.x / kWave;
  bool far = false;
  for (int cell = (blockIdx.x * blockDim.x + threadIdx.x) / kWave; cell < g.ncell; cell += nw) {
    const int start = cell_start[cell];
    const int cnt = min(cell_start[cell + 1] - start, g.cap);
    const int* mem = slots + (long long)cell * g.cap;
    for (int a0 = 0; a0 < cnt; a0 += kWave) {
      const int a = a0 + lane;
      const int id = a < cnt ? mem[a] : 0x7fffffff;
      int rank = 0;
      for (int b0 = 0; b0 < cnt; b0 += kWave) {
        const int idb = b0 + lane < cnt ? mem[b0 + lane] : 0x7fffffff;
        const int m = min(kWave, cnt - b0);
        for (int k = 0; k < m; ++k) rank += __builtin_amdgcn_readlane(idb, k) < id ? 1 : 0;
      }
      if (a < cnt) {
        float x[3] = {0.0f, 0.0f, 0.0f};
        load_pos<DIM, STRIDE>(pos, id, x);
        if (s.periodic) {
#pragma unroll
          for (int d = 0; d < DIM; ++d) {
            // 0 <= x < L - cell_size: int(x / cell_size) < cps, no wrap
            if (x[d] >= 0.0f && x[d] < s.L[d] - g.cell_size[d]) continue;
            const int q = (int)__fdiv_rn(x[d], g.cell_size[d]);
            const int wraps = (q - floor_mod_int(q, g.cps[d])) / g.cps[d];
            if (wraps != 0) x[d] = x[d] - (float)wraps * s.L[d];
            // x - wraps * L rounds at the scale of |x| once |wraps| >= 2
            if (wraps > 1 || wraps < -1) far = true;
          }
        }
        sorted_id[start + rank] = id;
        sorted_pos[start + rank] = make_float4(x[0], x[1], x[2], __int_as_float(id));
      }
    }
  }
  if (__ballot(far) != 0ull && lane == 0) atomicOr(&misc[M_FAR], 1);
}

__device__ __forceinline__ void dev_finish(int* __restrict__ misc, int occ,
                                           int* __restrict__ occupancy,
                                           uint8_t* __restrict__ error,
                                           long long max_occupancy, int* gate);

// The list build; with `finish`, its last workgroup completes the build
// (error bits, occupancy, gate reset).
template <int DIM, int STRIDE, int MODE, int GEOM>
__global__ __launch_bounds__(256) void k_nl_cells(BuildArgs A, int finish) {
  if (gate_closed(A.gate)) return;
  if (MODE == MODE_SPARSE_WRITE) {
    // Sparse padding idx[:, total:width] = n beside the pair writes (which
    // stay below total, the row scan's sum); every workgroup reads total
    // before it takes its completion ticket, so before the finish resets it.
    const long long total = A.w.misc[M_OCC];
    for (long long k = total + (long long)blockIdx.x * blockDim.x + threadIdx.x; k < A.width;
         k += (long long)gridDim.x * blockDim.x) {
      A.idx[k] = A.n;
      A.idx[A.width + k] = A.n;
    }
  }
  const int dep = dev_nl_cells<DIM, STRIDE, MODE, GEOM>(A, blockIdx.x, gridDim.x);
  if (!finish) return;
  if (!last_workgroup(&A.w.misc[M_TICKET_NL], dep)) return;
  if (threadIdx.x == 0) {
    A.w.misc[M_TICKET_NL] = 0;
    dev_finish(A.w.misc, coherent_load(&A.w.misc[M_OCC]), A.occupancy, A.error,
               A.error ? A.width : 0x7fffffffffffLL, A.gate);
  }
}

// Error bits, occupancy output, gate and counter reset (partition.py:1066,1110).
__device__ __forceinline__ void dev_finish(int* __restrict__ misc, int occ,
                                           int* __restrict__ occupancy,
                                           uint8_t* __restrict__ error,
                                           long long max_occupancy, int* gate) {
  int flags = misc[M_FLAGS];
  if ((long long)occ > max_occupancy) flags |= JMD_NEIGHBOR_LIST_OVERFLOW;
  if (occupancy) occupancy[0] = occ;
  if (error) error[0] = (uint8_t)(error[0] | flags);
  misc[M_OCC] = 0;
  misc[M_FLAGS] = 0;
  misc[M_FAR] = 0;
  if (error) misc[M_BUILDS] += 1;
  if (gate) gate[0] = 0;
}

// Gated copy of the positions into reference_position (no-cell path).
template <int DIM, int STRIDE>
__global__ __launch_bounds__(256) void k_copy_gated(const float* __restrict__ src,
                                                    float* __restrict__ dst, int n,
                                                    const int* gate) {
  if (gate_closed(gate)) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float x[3];
    load_pos<DIM, STRIDE>(src, i, x);
    store_pos<DIM>(dst, i, x);
  }
}

// ---------------------------------------------------------------------------
// Neighbour-list build without a cell list (all-pairs candidates,
// partition.py:904-909): one wavefront per row, candidates j in order.
// ---------------------------------------------------------------------------

template <int DIM, int STRIDE, int MODE>
__global__ __launch_bounds__(256) void k_nl_allpairs(
    const float* __restrict__ pos, int n, Space s, float cutoff_sq, int test,
    int mask_self, int ordered, int* __restrict__ idx, long long width,
    int* __restrict__ row_count, const int* __restrict__ row_offset,
    int* __restrict__ misc, const int* gate) {
  if (gate_closed(gate)) return;
  __shared__ int s_occ;
  if (threadIdx.x == 0) s_occ = 0;
  __syncthreads();
  int i = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  int lane = lane_id();
  int cnt = 0;
  if (i < n) {
  float xi[3] = {0, 0, 0};
  load_pos<DIM, STRIDE>(pos, i, xi);
  long long base = 0;
  if (MODE == MODE_DENSE) base = (long long)i * width;
  if (MODE == MODE_SPARSE_WRITE) base = row_offset[i];
  for (int j0 = 0; j0 < n; j0 += kWave) {
    int j = j0 + lane;
    bool pass = false;
    if (j < n) {
      float xj[3] = {0, 0, 0};
      load_pos<DIM, STRIDE>(pos, j, xj);
      pass = (!mask_self || j != i) && (!ordered || j < i) &&
             pair_test<DIM>(xi, xj, s, cutoff_sq, 0.0f, test);
    }
    unsigned long long bal = __ballot(pass);
    if ((MODE == MODE_DENSE || MODE == MODE_SPARSE_WRITE) && pass) {
      long long slot = cnt + wave_prefix_count(bal);
      if (MODE == MODE_DENSE) {
        if (slot < width) idx[base + slot] = j;
      } else {
        long long at = base + slot;
        if (at < width) {
          idx[at] = j;
          idx[width + at] = i;
        }
      }
    }
    cnt += __popcll(bal);
  }
  if (MODE == MODE_DENSE) {
    for (long long k = (long long)cnt + lane; k < width; k += kWave) idx[base + k] = n;
  } else if (MODE == MODE_SPARSE_COUNT) {
    if (lane == 0) row_count[i] = cnt;
  }
  }
  if (MODE == MODE_DENSE || MODE == MODE_COUNT_MAX) {
    if (lane == 0 && cnt > 0) atomicMax(&s_occ, cnt);
    __syncthreads();
    if (threadIdx.x == 0 && s_occ > 0) atomicMax(&misc[M_OCC], s_occ);
  }
}

// Inclusive scan over the 64 lanes of a wavefront (64-bit values, moved as
// two 32-bit halves).
__device__ __forceinline__ long long wave_inclusive_scan(long long x) {
  const int lane = lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const int lo = __shfl_up((int)(unsigned)(unsigned long long)x, off, kWave);
    const int hi = __shfl_up((int)((unsigned long long)x >> 32), off, kWave);
    const long long y = (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
    if (lane >= off) x += y;
  }
  return x;
}

// Exclusive scan of row_count[0..n) into row_offset[0..n] in one workgroup
// (Sparse formats: flattened row-major order, partition.py:1000-1005).
// Chunks of kRowScanChunk counts are loaded into LDS with coalesced reads; each
// thread sums kRowScanPer consecutive counts, the 1024 sums are scanned with
// wavefront scans (two barriers), and the offsets go back to memory from LDS
// with coalesced writes.  Offsets saturate at INT32_MAX as before.
constexpr int kRowScanPer = 8;
constexpr int kRowScanChunk = kRowScanPer * 1024;
__global__ __launch_bounds__(1024) void k_row_scan(const int* __restrict__ row_count,
                                                   int* __restrict__ row_offset, int n,
                                                   int* __restrict__ misc,
                                                   const int* gate) {
  if (gate_closed(gate)) return;
  __shared__ int s_v[kRowScanChunk + kRowScanChunk / 32];  // one pad word per 32
  __shared__ long long s_w[1024 / kWave];
  const int t = threadIdx.x;
  const int lane = lane_id(), wave = t / kWave;
  const long long kMax = 0x7fffffffLL;
  auto at = [](int k) { return k + (k >> 5); };  // padded LDS index
  long long carry = 0;
  for (int base = 0; base < n; base += kRowScanChunk) {
    const int m = min(kRowScanChunk, n - base);
    for (int k = t; k < kRowScanChunk; k += 1024) s_v[at(k)] = k < m ? row_count[base + k] : 0;
    __syncthreads();
    long long local = 0;
#pragma unroll
    for (int q = 0; q < kRowScanPer; ++q) local += s_v[at(t * kRowScanPer + q)];
    const long long incl = wave_inclusive_scan(local);
    if (lane == kWave - 1) s_w[wave] = incl;
    __syncthreads();
    if (wave == 0) {
      long long w = lane < 1024 / kWave ? s_w[lane] : 0;
      w = wave_inclusive_scan(w);
      if (lane < 1024 / kWave) s_w[lane] = w;
    }
    __syncthreads();
    long long run = carry + (wave > 0 ? s_w[wave - 1] : 0) + (incl - local);
    const long long total = s_w[1024 / kWave - 1];
#pragma unroll
    for (int q = 0; q < kRowScanPer; ++q) {
      const int k = at(t * kRowScanPer + q);
      const int c = s_v[k];
      s_v[k] = (int)min(run, kMax);
      run += c;
    }
    __syncthreads();
    for (int k = t; k < m; k += 1024) row_offset[base + k] = s_v[at(k)];
    carry += total;
    __syncthreads();  // s_v and s_w are refilled by the next chunk
  }
  if (t == 0) {
    row_offset[n] = (int)min(carry, kMax);
    misc[M_OCC] = (int)min(carry, kMax);
  }
}

// Sparse padding: idx[:, total:width] = n.
__global__ __launch_bounds__(256) void k_sparse_pad(int* __restrict__ idx, long long width,
                                                    int n, const int* __restrict__ misc,
                                                    const int* gate) {
  if (gate_closed(gate)) return;
  long long total = misc[M_OCC];
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (k += total; k < width; k += stride) {
    idx[k] = n;
    idx[width + k] = n;
  }
}

// Error word, occupancy output and resets after a multi-kernel build.
__global__ void k_nl_finish(int* __restrict__ misc, int* __restrict__ occupancy,
                            uint8_t* __restrict__ error, long long max_occupancy, int* gate) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (gate_closed(gate)) return;
  dev_finish(misc, misc[M_OCC], occupancy, error, max_occupancy, gate);
}

// Skin criterion (partition.py:1141-1154).
template <int DIM, int STRIDE>
__global__ __launch_bounds__(256) void k_skin_check(const float* __restrict__ pos,
                                                    const float* __restrict__ ref, int n,
                                                    Space s, float threshold_sq,
                                                    int* __restrict__ gate) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool moved = false;
  if (i < n) {
    float a[3], b[3];
    load_pos<DIM, STRIDE>(pos, i, a);
    load_pos<DIM>(ref, i, b);
    moved = ref_metric_sq<DIM>(a, b, s) > threshold_sq;
  }
  unsigned long long bal = __ballot(moved);
  if (bal != 0ull && lane_id() == 0) atomicOr(gate, 1);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

static bool make_grid(int dim, const int32_t* cps, const float* cell_size, int cap,
                      Grid* g) {
  long long nc = 1;
  for (int d = 0; d < 3; ++d) {
    g->cps[d] = d < dim ? cps[d] : 1;
    g->cell_size[d] = d < dim ? cell_size[d] : 1.0f;
    if (d < dim && (cps[d] < 1 || !(cell_size[d] > 0.0f))) return false;
    nc *= g->cps[d];
  }
  if (nc > 0x7fffffffLL) return false;
  g->ncell = (int)nc;
  g->cap = cap;
  return true;
}

static Space grid_space_free(int dim) {
  float one[3] = {1.0f, 1.0f, 1.0f};
  return make_space(dim, 0, one);
}

template <int DIM, int STRIDE>
static int bin_and_order(const float* pos, int n, const Grid& g, const Space& s,
                         const Work& w, float* ref_out, int* cell_id_out, const int* gate,
                         hipStream_t st) {
  int nb = ceil_div(n, 256);
  if (n > 0) {
    hipLaunchKernelGGL((k_bin_count<DIM, STRIDE>), dim3(nb), dim3(256), 0, st, pos, n, g,
                       w.cell_count, w.atom_cell, w.atom_rank, ref_out, cell_id_out, gate);
    JMD_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, st, w.cell_count, w.cell_start,
                     g.ncell, g.cap, w.misc, gate);
  JMD_LAUNCH_CHECK();
  if (n > 0) {
    hipLaunchKernelGGL(k_bin_scatter, dim3(nb), dim3(256), 0, st, n, w.atom_cell,
                       w.atom_rank, w.cell_start, w.sorted_tmp, gate);
    JMD_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL((k_cell_order<DIM, STRIDE>),
                     dim3(ceil_div((long long)g.ncell * kWave, 256)), dim3(256), 0, st, pos,
                     g.ncell, g, s, w.cell_start, w.sorted_tmp, w.sorted_id, w.sorted_pos, w.misc,
                     gate);
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

template <int DIM, int STRIDE, int GEOM>
static void launch_cells(const BuildArgs& A, int mode, int finish, hipStream_t st) {
  dim3 grid(A.g.ncell < kCellBlocks ? A.g.ncell : kCellBlocks), block(256);
  switch (mode) {
    case MODE_DENSE:
      hipLaunchKernelGGL((k_nl_cells<DIM, STRIDE, MODE_DENSE, GEOM>), grid, block, 0, st, A,
                         finish);
      break;
    case MODE_SPARSE_COUNT:
      hipLaunchKernelGGL((k_nl_cells<DIM, STRIDE, MODE_SPARSE_COUNT, GEOM>), grid, block, 0,
                         st, A, finish);
      break;
    case MODE_SPARSE_WRITE:
      hipLaunchKernelGGL((k_nl_cells<DIM, STRIDE, MODE_SPARSE_WRITE, GEOM>), grid, block, 0,
                         st, A, finish);
      break;
    default:
      hipLaunchKernelGGL((k_nl_cells<DIM, STRIDE, MODE_COUNT_MAX, GEOM>), grid, block, 0, st,
                         A, finish);
      break;
  }
}

template <int DIM, int STRIDE>
static int cells_dispatch(const BuildArgs& A, int mode, int finish, hipStream_t st) {
  // The staged image of a stencil cell is the nearest one when every
  // periodic dimension has at least 5 cells (|d| < 3 cell sizes < L - c).
  // periodic_general (fractional units) applies T to each difference.
  bool minimg = false;
  if (A.s.periodic)
    for (int d = 0; d < DIM; ++d) minimg = minimg || A.g.cps[d] < 5;
  if (minimg) launch_cells<DIM, STRIDE, 1>(A, mode, finish, st);
  else if (A.s.general) launch_cells<DIM, STRIDE, 2>(A, mode, finish, st);
  else launch_cells<DIM, STRIDE, 0>(A, mode, finish, st);
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

#ifndef JMD_BIN_BLOCKS
#define JMD_BIN_BLOCKS 512
#endif
constexpr int kBinBlocks = JMD_BIN_BLOCKS;

// Cell-list build: k_bin_slots, k_cell_sort, k_nl_cells (Sparse adds the
// row scan, the padding and a second k_nl_cells pass).
template <int DIM, int STRIDE>
static int build_cells(const BuildArgs& A, bool sparse, bool count_only, hipStream_t st) {
  const Work& w = A.w;
  int nb = ceil_div(A.n > 0 ? A.n : 1, 256);
  if (nb > kBinBlocks) nb = kBinBlocks;
  hipLaunchKernelGGL((k_bin_slots<DIM, STRIDE>), dim3(nb), dim3(256), 0, st, A.pos, A.n, A.g,
                     w.cell_count, w.cell_start, w.slots, A.ref_out, w.misc, A.gate, A.kd);
  JMD_LAUNCH_CHECK();
  long long sb = ceil_div((long long)A.g.ncell * kWave, 256);
  if (sb > kSortBlocks) sb = kSortBlocks;
  hipLaunchKernelGGL((k_cell_sort<DIM, STRIDE>), dim3((unsigned)sb),
                     dim3(256), 0, st, A.pos, A.g, A.s, w.cell_start, w.slots, w.sorted_id,
                     w.sorted_pos, w.misc, A.gate);
  JMD_LAUNCH_CHECK();
  if (!sparse)
    return cells_dispatch<DIM, STRIDE>(A, count_only ? MODE_COUNT_MAX : MODE_DENSE, 1, st);
  int rc = cells_dispatch<DIM, STRIDE>(A, MODE_SPARSE_COUNT, 0, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_row_scan, dim3(1), dim3(1024), 0, st, w.row_count, w.row_offset, A.n,
                     w.misc, A.gate);
  JMD_LAUNCH_CHECK();
  if (count_only) {
    hipLaunchKernelGGL(k_nl_finish, dim3(1), dim3(64), 0, st, w.misc, A.occupancy, nullptr,
                       (long long)0x7fffffffffffLL, A.gate);
    JMD_LAUNCH_CHECK();
    return JMD_OK;
  }
  // the write launch also pads idx[:, total:width]
  return cells_dispatch<DIM, STRIDE>(A, MODE_SPARSE_WRITE, 1, st);
}

template <int DIM, int STRIDE>
static int build_impl(const float* pos, int n, const Space& s, float cutoff_sq,
                      int use_cells, const Grid& g, int format, int mask_self,
                      int* idx, long long width, int* occupancy, uint8_t* error,
                      float* ref_out, int* gate, const Work& w, hipStream_t st,
                      const KickDrift* kd = nullptr) {
  // JMD_DENSE_SPARSE_TEST: Dense rows kept by the Sparse formats' pair test
  const bool ij_rows = (format & JMD_DENSE_SPARSE_TEST) != 0;
  format &= ~JMD_DENSE_SPARSE_TEST;
  const bool sparse = format != JMD_DENSE;
  const int ordered = format == JMD_ORDERED_SPARSE;
  const bool count_only = error == nullptr;
  if (use_cells) {
    BuildArgs A{};
    if (kd) A.kd = *kd;
    A.pos = pos;
    A.n = n;
    A.g = g;
    A.s = s;
    A.cutoff_sq = cutoff_sq;
    // Bound on |fast d^2 - reference wrapped d^2| in float32: a few roundings
    // at the scale of 2L on each of the components of d.
    float Lmax = 0.0f;
    for (int d = 0; d < DIM; ++d) Lmax = fmaxf(Lmax, s.periodic ? s.L[d] : 0.0f);
    if (s.general) {  // real-space extent of the cell: largest row sum of |T|
      Lmax = 0.0f;
      for (int a = 0; a < DIM; ++a) {
        float r = 0.0f;
        for (int b = 0; b < DIM; ++b) r += fabsf(s.T[a][b]);
        Lmax = fmaxf(Lmax, r);
      }
    }
    float c = sqrtf(cutoff_sq);
    float margin = 128.0f * (c + 1.0f) * (Lmax + c + 1.0f) * 1.1920929e-7f;
    A.margin = fmaxf(margin, 1e-4f * cutoff_sq);
    A.test = (sparse || ij_rows) ? TEST_IJ : TEST_BOTH;
    A.mask_self = mask_self;
    A.ordered = ordered;
    A.idx = idx;
    A.width = width;
    A.ref_out = ref_out;
    A.occupancy = occupancy;
    A.error = error;
    A.gate = gate;
    A.w = w;
    return build_cells<DIM, STRIDE>(A, sparse, count_only, st);
  }
  if (ref_out && n > 0) {
    int nb = ceil_div(n, 256);
    hipLaunchKernelGGL((k_copy_gated<DIM, STRIDE>), dim3(nb > 2048 ? 2048 : nb), dim3(256),
                       0, st, pos, ref_out, n, gate);
    JMD_LAUNCH_CHECK();
  }
  // Dense without cells tests wrap(Rj - Ri) (partition.py:965-971);
  // Sparse tests wrap(Ri - Rj) (partition.py:990-995).
  const int test = (sparse || ij_rows) ? TEST_IJ : TEST_JI;
  dim3 grid(ceil_div((long long)n * kWave, 256)), block(256);
#define JMD_ALLP(MODE)                                                                      \
  hipLaunchKernelGGL((k_nl_allpairs<DIM, STRIDE, MODE>), grid, block, 0, st, pos, n, s,      \
                     cutoff_sq, test, mask_self, ordered, idx, width, w.row_count,           \
                     w.row_offset, w.misc, gate);                                            \
  JMD_LAUNCH_CHECK();
  if (n > 0) {
    if (!sparse && count_only) {
      JMD_ALLP(MODE_COUNT_MAX)
    } else if (!sparse) {
      JMD_ALLP(MODE_DENSE)
    } else {
      JMD_ALLP(MODE_SPARSE_COUNT)
      hipLaunchKernelGGL(k_row_scan, dim3(1), dim3(1024), 0, st, w.row_count, w.row_offset, n,
                         w.misc, gate);
      JMD_LAUNCH_CHECK();
      if (!count_only) { JMD_ALLP(MODE_SPARSE_WRITE) }
    }
  }
#undef JMD_ALLP
  if (sparse && !count_only) {
    int pb = ceil_div(width, 256);
    pb = pb > 4096 ? 4096 : pb + 1;
    hipLaunchKernelGGL(k_sparse_pad, dim3(pb), dim3(256), 0, st, idx, width, n, w.misc, gate);
    JMD_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_nl_finish, dim3(1), dim3(64), 0, st, w.misc, occupancy, error,
                     count_only ? (long long)0x7fffffffffffLL : width, gate);
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

__global__ __launch_bounds__(256) void k_order_out(const int* __restrict__ sorted_id, int n,
                                                   int* __restrict__ order,
                                                   int* __restrict__ inverse) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int id = sorted_id[s];
  if (order) order[s] = id;
  if (inverse) inverse[id] = s;
}

// Runs F<DIM, STRIDE>() for the supported (dim, stride) layouts.
#define JMD_DISPATCH_LAYOUT(dim, stride, CALL3, CALL34, CALL2) \
  ((dim) == 3 ? ((stride) == 4 ? (CALL34) : (CALL3)) : (CALL2))

}  // namespace jmd

using namespace jmd;

extern "C" {

const char* jmd_version(void) { return "jaxmd_native 0.3.0 (gfx950)"; }

size_t jmd_neighbor_workspace_bytes(int64_t n, int32_t dim, int64_t cell_count,
                                    int32_t cell_capacity, int32_t format) {
  (void)dim;
  (void)format;
  if (n < 0 || cell_count < 0 || cell_capacity < 0) return 0;
  return carve(nullptr, n, cell_count, cell_capacity, nullptr);
}

int jmd_cell_filling_max(const float* position, int64_t n, int32_t dim,
                         int32_t position_stride, const int32_t* cells_per_side,
                         const float* cell_size, int32_t* filling_max, void* workspace,
                         void* stream) {
  if ((dim != 2 && dim != 3) || n < 0 || n > 0x7fffffff || !cells_per_side || !cell_size ||
      !filling_max || !workspace || (n > 0 && !position))
    return JMD_ERR_INVALID;
  if (n > 0 && !stride_ok(dim, position_stride, position)) return JMD_ERR_INVALID;
  Grid g;
  if (!make_grid(dim, cells_per_side, cell_size, 0, &g)) return JMD_ERR_INVALID;
  Work w;
  carve(workspace, n, g.ncell, 0, &w);
  hipStream_t st = (hipStream_t)stream;
  int nn = (int)n;
  if (nn > 0) {
    dim3 grid(ceil_div(nn, 256)), block(256);
    if (dim == 3 && position_stride == 4)
      hipLaunchKernelGGL((k_bin_count<3, 4>), grid, block, 0, st, position, nn, g,
                         w.cell_count, w.atom_cell, w.atom_rank, nullptr, nullptr, nullptr);
    else if (dim == 3)
      hipLaunchKernelGGL((k_bin_count<3, 3>), grid, block, 0, st, position, nn, g,
                         w.cell_count, w.atom_cell, w.atom_rank, nullptr, nullptr, nullptr);
    else
      hipLaunchKernelGGL((k_bin_count<2, 2>), grid, block, 0, st, position, nn, g,
                         w.cell_count, w.atom_cell, w.atom_rank, nullptr, nullptr, nullptr);
    JMD_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, st, w.cell_count, w.cell_start,
                     g.ncell, 0, w.misc, nullptr);
  JMD_LAUNCH_CHECK();
  JMD_RETURN_IF_HIP(hipMemcpyAsync(filling_max, w.misc + M_FILL, sizeof(int),
                                   hipMemcpyDeviceToDevice, st));
  return JMD_OK;
}

int jmd_cell_list(const float* position, int64_t n, int32_t dim, int32_t position_stride,
                  const int32_t* cells_per_side, const float* cell_size,
                  int32_t cell_capacity, float* position_buffer, int32_t* id_buffer,
                  int32_t* particle_cell_id, int32_t* filling_max, uint8_t* overflow,
                  void* workspace, void* stream) {
  if ((dim != 2 && dim != 3) || n < 0 || n > 0x7fffffff || cell_capacity < 1 ||
      !cells_per_side || !cell_size || !position_buffer || !id_buffer || !workspace ||
      (n > 0 && (!position || !particle_cell_id)))
    return JMD_ERR_INVALID;
  if (n > 0 && !stride_ok(dim, position_stride, position)) return JMD_ERR_INVALID;
  Grid g;
  if (!make_grid(dim, cells_per_side, cell_size, cell_capacity, &g)) return JMD_ERR_INVALID;
  Work w;
  carve(workspace, n, g.ncell, 0, &w);
  hipStream_t st = (hipStream_t)stream;
  int nn = (int)n;
  long long slots = (long long)g.ncell * cell_capacity;
  int rc = launch_fill<int>(id_buffer, slots, nn, st);
  if (rc) return rc;
  rc = launch_fill<float>(position_buffer, slots * dim, 0.0f, st);
  if (rc) return rc;
  Space s = grid_space_free(dim);
  rc = JMD_DISPATCH_LAYOUT(
      dim, position_stride,
      (bin_and_order<3, 3>(position, nn, g, s, w, nullptr, particle_cell_id, nullptr, st)),
      (bin_and_order<3, 4>(position, nn, g, s, w, nullptr, particle_cell_id, nullptr, st)),
      (bin_and_order<2, 2>(position, nn, g, s, w, nullptr, particle_cell_id, nullptr, st)));
  if (rc) return rc;
  if (nn > 0) {
    dim3 grid(ceil_div(nn, 256)), block(256);
    if (dim == 3 && position_stride == 4)
      hipLaunchKernelGGL((k_cell_emplace<3, 4>), grid, block, 0, st, position, nn,
                         cell_capacity, w.cell_start, w.sorted_id, g.ncell, position_buffer,
                         id_buffer);
    else if (dim == 3)
      hipLaunchKernelGGL((k_cell_emplace<3, 3>), grid, block, 0, st, position, nn,
                         cell_capacity, w.cell_start, w.sorted_id, g.ncell, position_buffer,
                         id_buffer);
    else
      hipLaunchKernelGGL((k_cell_emplace<2, 2>), grid, block, 0, st, position, nn,
                         cell_capacity, w.cell_start, w.sorted_id, g.ncell, position_buffer,
                         id_buffer);
    JMD_LAUNCH_CHECK();
  }
  if (filling_max) {
    JMD_RETURN_IF_HIP(hipMemcpyAsync(filling_max, w.misc + M_FILL, sizeof(int),
                                     hipMemcpyDeviceToDevice, st));
  }
  // overflow byte from the scan's flag word, then clear the flag word
  hipLaunchKernelGGL(k_nl_finish, dim3(1), dim3(64), 0, st, w.misc, nullptr, overflow,
                     (long long)0x7fffffffffffLL, nullptr);
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

int jmd_neighbor_list_build(const float* position, int64_t n, int32_t dim,
                            int32_t position_stride, int32_t periodic, const float* side,
                            float cutoff_sq, int32_t use_cell_list,
                            const int32_t* cells_per_side, const float* cell_size,
                            int32_t cell_capacity, int32_t format, int32_t mask_self,
                            int32_t* idx, int64_t max_occupancy, int32_t* occupancy,
                            uint8_t* error, float* reference_position, int32_t* gate,
                            void* workspace, void* stream) {
  const int base_format = format == (JMD_DENSE | JMD_DENSE_SPARSE_TEST) ? JMD_DENSE : format;
  if ((dim != 2 && dim != 3) || n < 0 || n > 0x7fffffff || max_occupancy < 0 ||
      (base_format != JMD_DENSE && base_format != JMD_SPARSE &&
       base_format != JMD_ORDERED_SPARSE) ||
      !workspace || (n > 0 && !position) || (periodic && !side))
    return JMD_ERR_INVALID;
  if (n > 0 && !stride_ok(dim, position_stride, position)) return JMD_ERR_INVALID;
  if (error == nullptr && !occupancy) return JMD_ERR_INVALID;
  if (error != nullptr && max_occupancy > 0 && !idx) return JMD_ERR_INVALID;
  if (base_format == JMD_DENSE && (long long)n * max_occupancy > 0x7fffffffffffLL)
    return JMD_ERR_INVALID;
  Grid g{};
  if (use_cell_list) {
    if (!cells_per_side || !cell_size || cell_capacity < 1) return JMD_ERR_INVALID;
    if (!make_grid(dim, cells_per_side, cell_size, cell_capacity, &g)) return JMD_ERR_INVALID;
    for (int d = 0; d < dim; ++d)
      if (cells_per_side[d] < 3) return JMD_ERR_INVALID;
  }
  {
    const int grc = space_check(dim, periodic, side, true);
    if (grc) return grc;
  }
  Space s = make_space(dim, periodic, side);
  Work w;
  carve(workspace, n, use_cell_list ? g.ncell : 1, use_cell_list ? cell_capacity : 0, &w);
  hipStream_t st = (hipStream_t)stream;
  int nn = (int)n;
  return JMD_DISPATCH_LAYOUT(
      dim, position_stride,
      (build_impl<3, 3>(position, nn, s, cutoff_sq, use_cell_list, g, format, mask_self, idx,
                        max_occupancy, occupancy, error, reference_position, gate, w, st)),
      (build_impl<3, 4>(position, nn, s, cutoff_sq, use_cell_list, g, format, mask_self, idx,
                        max_occupancy, occupancy, error, reference_position, gate, w, st)),
      (build_impl<2, 2>(position, nn, s, cutoff_sq, use_cell_list, g, format, mask_self, idx,
                        max_occupancy, occupancy, error, reference_position, gate, w, st)));
}

int jmd_nve_list_update_kick_drift(
    const float* position, int64_t n, int32_t dim, int32_t position_stride, int32_t periodic,
    const float* side, float cutoff_sq, int32_t use_cell_list, const int32_t* cells_per_side,
    const float* cell_size, int32_t cell_capacity, int32_t format, int32_t mask_self,
    int32_t* idx, int64_t max_occupancy, int32_t* occupancy, uint8_t* error,
    float* reference_position, int32_t* gate, void* workspace, double* position_hi,
    float* momentum, const float* force, const float* mass, float mass_scalar,
    int32_t shift_periodic, const float* shift_side, float dt, float half_dt,
    const float* dt_device, float* position_out, float threshold_sq, int32_t* gate_next,
    void* stream) {
  if (n <= 0) return n < 0 ? JMD_ERR_INVALID : JMD_OK;
  if (!position_hi || !momentum || !force || !position_out || position_out == position ||
      !error || !gate || !gate_next || gate_next == gate || !reference_position ||
      (shift_periodic && !shift_side) || (!mass && !(mass_scalar > 0.0f)))
    return JMD_ERR_INVALID;
  if (!use_cell_list) {
    // no binning pass to share: the two stages one after the other
    int rc = jmd_neighbor_list_build(position, n, dim, position_stride, periodic, side,
                                     cutoff_sq, use_cell_list, cells_per_side, cell_size,
                                     cell_capacity, format, mask_self, idx, max_occupancy,
                                     occupancy, error, reference_position, gate, workspace,
                                     stream);
    if (rc) return rc;
    return jmd_nve_kick_drift(position_hi, momentum, force, mass, mass_scalar, n, dim,
                              shift_periodic, shift_side, dt, half_dt, dt_device, position_hi,
                              position_out, position_stride, momentum, reference_position,
                              threshold_sq, gate_next, stream);
  }
  const int base_format = format == (JMD_DENSE | JMD_DENSE_SPARSE_TEST) ? JMD_DENSE : format;
  if ((dim != 2 && dim != 3) || n > 0x7fffffff || max_occupancy < 0 ||
      (base_format != JMD_DENSE && base_format != JMD_SPARSE &&
       base_format != JMD_ORDERED_SPARSE) ||
      !workspace || (periodic && !side) || (max_occupancy > 0 && !idx))
    return JMD_ERR_INVALID;
  if (!stride_ok(dim, position_stride, position) ||
      !stride_ok(dim, position_stride, position_out))
    return JMD_ERR_INVALID;
  if (base_format == JMD_DENSE && (long long)n * max_occupancy > 0x7fffffffffffLL)
    return JMD_ERR_INVALID;
  if (!cells_per_side || !cell_size || cell_capacity < 1) return JMD_ERR_INVALID;
  Grid g{};
  if (!make_grid(dim, cells_per_side, cell_size, cell_capacity, &g)) return JMD_ERR_INVALID;
  for (int d = 0; d < dim; ++d)
    if (cells_per_side[d] < 3) return JMD_ERR_INVALID;
  {
    const int grc = space_check(dim, periodic, side, true);
    if (grc) return grc;
  }
  Space s = make_space(dim, periodic, side);
  KickDrift kd{};
  kd.active = 1;
  kd.xhi_in = position_hi;
  kd.xhi_out = position_hi;
  kd.p_in = momentum;
  kd.p_out = momentum;
  kd.F = force;
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
  {
    const int grc = space_check(dim, shift_periodic, shift_side, true);
    if (grc) return grc;
  }
  kd.s = make_space(dim, shift_periodic, shift_side);
  box_inverse(kd.s, kd.Ti);
  Work w;
  carve(workspace, n, g.ncell, cell_capacity, &w);
  hipStream_t st = (hipStream_t)stream;
  int nn = (int)n;
  return JMD_DISPATCH_LAYOUT(
      dim, position_stride,
      (build_impl<3, 3>(position, nn, s, cutoff_sq, 1, g, format, mask_self, idx, max_occupancy,
                        occupancy, error, reference_position, gate, w, st, &kd)),
      (build_impl<3, 4>(position, nn, s, cutoff_sq, 1, g, format, mask_self, idx, max_occupancy,
                        occupancy, error, reference_position, gate, w, st, &kd)),
      (build_impl<2, 2>(position, nn, s, cutoff_sq, 1, g, format, mask_self, idx, max_occupancy,
                        occupancy, error, reference_position, gate, w, st, &kd)));
}

int jmd_cell_order(const float* position, int64_t n, int32_t dim, int32_t position_stride,
                   const int32_t* cells_per_side, const float* cell_size, int32_t* order,
                   int32_t* inverse, void* workspace, void* stream) {
  if ((dim != 2 && dim != 3) || n < 0 || n > 0x7fffffff || !cells_per_side || !cell_size ||
      !workspace || (n > 0 && (!position || (!order && !inverse))))
    return JMD_ERR_INVALID;
  if (n > 0 && !stride_ok(dim, position_stride, position)) return JMD_ERR_INVALID;
  Grid g;
  if (!make_grid(dim, cells_per_side, cell_size, 0, &g)) return JMD_ERR_INVALID;
  if (n == 0) return JMD_OK;
  Work w;
  carve(workspace, n, g.ncell, 0, &w);
  hipStream_t st = (hipStream_t)stream;
  int nn = (int)n;
  Space s = grid_space_free(dim);
  int rc = JMD_DISPATCH_LAYOUT(
      dim, position_stride,
      (bin_and_order<3, 3>(position, nn, g, s, w, nullptr, nullptr, nullptr, st)),
      (bin_and_order<3, 4>(position, nn, g, s, w, nullptr, nullptr, nullptr, st)),
      (bin_and_order<2, 2>(position, nn, g, s, w, nullptr, nullptr, nullptr, st)));
  if (rc) return rc;
  hipLaunchKernelGGL(k_order_out, dim3(ceil_div(nn, 256)), dim3(256), 0, st, w.sorted_id, nn,
                     order, inverse);
  JMD_LAUNCH_CHECK();
  // clear the flag word the scan may have raised (capacity 0: none)
  hipLaunchKernelGGL(k_nl_finish, dim3(1), dim3(64), 0, st, w.misc, nullptr, nullptr,
                     (long long)0x7fffffffffffLL, nullptr);
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

int jmd_neighbor_list_check(const float* position, int32_t position_stride,
                            const float* reference_position, int64_t n, int32_t dim,
                            int32_t periodic, const float* side, float threshold_sq,
                            int32_t* gate, void* stream) {
  if ((dim != 2 && dim != 3) || n < 0 || n > 0x7fffffff || !gate ||
      (n > 0 && (!position || !reference_position)) || (periodic && !side))
    return JMD_ERR_INVALID;
  if (n == 0) return JMD_OK;
  if (!stride_ok(dim, position_stride, position)) return JMD_ERR_INVALID;
  {
    const int grc = space_check(dim, periodic, side, true);
    if (grc) return grc;
  }
  Space s = make_space(dim, periodic, side);
  hipStream_t st = (hipStream_t)stream;
  int nn = (int)n;
  dim3 grid(ceil_div(nn, 256)), block(256);
  if (dim == 3 && position_stride == 4)
    hipLaunchKernelGGL((k_skin_check<3, 4>), grid, block, 0, st, position, reference_position,
                       nn, s, threshold_sq, gate);
  else if (dim == 3)
    hipLaunchKernelGGL((k_skin_check<3, 3>), grid, block, 0, st, position, reference_position,
                       nn, s, threshold_sq, gate);
  else
    hipLaunchKernelGGL((k_skin_check<2, 2>), grid, block, 0, st, position, reference_position,
                       nn, s, threshold_sq, gate);
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

}  // extern "C"
