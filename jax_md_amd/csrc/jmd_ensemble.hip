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
//   idx[k]: the Dense list the force at node k was computed with
//         ([K+1][ntot][width])
//   sc[k][s]: step k's scalars of system s: dt, alpha, keep, sqrt(pp),
//         sqrt(ff + 1e-6)  ([K][nsys][kTapeScalars])
struct EnsTape {
  float* x;
  float* F;
  float* q;
  int* idx;
  float* sc;
  int K;
};
constexpr int kTapeScalars = 5;

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

// Bins the system on grid g: LDS counts, exclusive starts, arrival ranks,
// scatter into sorted[] (system-local), then each cell's ids in ascending
// order (the stable argsort of partition.py:432).  s_cs holds the counts,
// then, scanned in place, the starts (cell c spans s_cs[c] .. s_cs[c + 1]).
// Returns the max filling.
__device__ int ens_bin(const EnsSys& S, const Grid& g, const float* __restrict__ x,
                       int* __restrict__ cellp, int* __restrict__ rankp,
                       int* __restrict__ sorted, int* s_cs, int* s_red) {
  const int nt = blockDim.x, t = threadIdx.x;
  for (int c = t; c < g.ncell; c += nt) s_cs[c] = 0;
  __syncthreads();
  for (int a = t; a < S.n; a += nt) {
    float xi[2];
    load_pos<2>(x, S.offset + a, xi);
    int ci[2];
    const int h = cell_of<2>(xi, g, ci);
    cellp[a] = h;
    rankp[a] = atomicAdd(&s_cs[h], 1);
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
  for (int a = t; a < S.n; a += nt) sorted[s_cs[cellp[a]] + rankp[a]] = a;
  __syncthreads();
  // order each cell by id (thread per cell, insertion sort of a few ids)
  for (int c = t; c < g.ncell; c += nt) {
    int* m = sorted + s_cs[c];
    const int cnt = s_cs[c + 1] - s_cs[c];
    for (int u = 1; u < cnt; ++u) {
      const int v = m[u];
      int w = u - 1;
      while (w >= 0 && m[w] > v) {
        m[w + 1] = m[w];
        --w;
      }
      m[w + 1] = v;
    }
  }
  __syncthreads();
  return wg_max(fill, s_red);
}

// Dense rows of the system (global ids, padding pad) in the reference's
// candidate order; returns the largest row count.
__device__ int ens_rows(const EnsSys& S, const Grid& g, const Space& sp, float cutoff_sq,
                        const float* __restrict__ x, const int* __restrict__ cellp,
                        const int* __restrict__ sorted, const int* s_cs,
                        int* __restrict__ idx, int width, int pad, int* s_red) {
  const int nt = blockDim.x, t = threadIdx.x;
  int occ = 0;
  for (int a = t; a < S.n; a += nt) {
    const int i = S.offset + a;
    float xi[2];
    load_pos<2>(x, i, xi);
    const int h = cellp[a];
    const int hx = h % g.cps[0], hy = h / g.cps[0];
    int cnt = 0;
    for (int st = 0; st < 9; ++st) {  // iota order, x offset slowest
      const int ox = st / 3 - 1, oy = st % 3 - 1;
      const int c = floor_mod_int(hx + ox, g.cps[0]) + g.cps[0] * floor_mod_int(hy + oy, g.cps[1]);
      const int start = s_cs[c];
      const int m = min(s_cs[c + 1] - start, g.cap);
      const int w = g.cap - (start % g.cap);
      for (int q = 0; q < m; ++q) {
        int k = q;  // slot order: rank (start + k) mod cap
        if (m > w) k = (q < m - w) ? (w + q) : (q - (m - w));
        const int jl = sorted[start + k];
        if (jl == a) continue;  // mask_self
        const int j = S.offset + jl;
        float xj[2];
        load_pos<2>(x, j, xj);
        // Dense with cells keeps a candidate when both wrap(Ri - Rj) and
        // wrap(Rj - Ri) are inside (partition.py:951,971)
        if (ref_metric_sq<2>(xi, xj, sp) < cutoff_sq && ref_metric_sq<2>(xj, xi, sp) < cutoff_sq) {
          if (cnt < width) idx[(long long)i * width + cnt] = j;
          ++cnt;
        }
      }
    }
    for (int k = cnt; k < width; ++k) idx[(long long)i * width + k] = pad;
    occ = max(occ, cnt);
  }
  return wg_max(occ, s_red);
}

// Soft-sphere force of row i: two accumulators (entries k even / odd) summed
// in float64, as the Dense kernel's two lanes per row do; optional energy.
__device__ __forceinline__ void ens_force_row(int i, const float* __restrict__ x,
                                              const int* __restrict__ idx, int width, int pad,
                                              const Space& sp, const PairArgs& a, float* F,
                                              double* e) {
#pragma clang fp contract(off)
  float xi[2], xim[2];
  load_pos<2>(x, i, xi);
  xim[0] = xi[0] - sp.L[0];
  xim[1] = xi[1] - sp.L[1];
  float f[2][2] = {{0.0f, 0.0f}, {0.0f, 0.0f}};
  float en[2] = {0.0f, 0.0f};
  const int si = a.species ? a.species[i] : 0;
  const int* row = idx + (long long)i * width;
  for (int k = 0; k < width; ++k) {
    const int j = row[k];
    if (j == pad) break;
    float xj[2];
    load_pos<2>(x, j, xj);
    float d[2];
    d[0] = min_image_disp(xj[0], xi[0], xim[0], sp.L[0], sp.half[0]);
    d[1] = min_image_disp(xj[1], xi[1], xim[1], sp.L[1], sp.half[1]);
    const float r2 = fmaf(d[1], d[1], d[0] * d[0]);
    const int sj = a.species ? a.species[j] : 0;
    const float sig = pick(a.sig_mode, a.sig, a.sig_v, i, j, si, sj, a.ns);
    const float eps = pick(a.eps_mode, a.eps, a.eps_v, i, j, si, sj, a.ns);
    const float alp = pick(a.alp_mode, a.alp, a.alp_v, i, j, si, sj, a.ns);
    float dEr;
    const float u = pair_eval<JMD_SOFT_SPHERE>(r2, sig, eps, alp, a, &dEr);
    const int l = k & 1;
    f[l][0] = fmaf(dEr, d[0], f[l][0]);
    f[l][1] = fmaf(dEr, d[1], f[l][1]);
    en[l] += u;
  }
  F[0] = (float)((double)f[0][0] + (double)f[1][0]);
  F[1] = (float)((double)f[0][1] + (double)f[1][1]);
  if (e) *e += 0.5 * ((double)en[0] + (double)en[1]);
}

__global__ __launch_bounds__(kEnsThreads) void k_fire_ensemble(
    const EnsSys* __restrict__ systems, const int* __restrict__ order, int nsys, int ntot,
    int mode, int n_steps, double* __restrict__ x_hi, float* __restrict__ x,
    float* __restrict__ p, float* __restrict__ F, float* __restrict__ x_ref,
    int* __restrict__ idx, int width, int* __restrict__ scratch, PairArgs a0, FireP prm,
    float* __restrict__ fire_state, int* __restrict__ fire_npos, uint8_t* __restrict__ error,
    int* __restrict__ out_stat, double* __restrict__ energy, int* __restrict__ builds,
    EnsTape tape) {
  __shared__ int s_cs[kEnsMaxCells + 1];
  __shared__ int s_red[kEnsThreads / kWave];
  __shared__ double s_d[3 * (kEnsThreads / kWave)];
  __shared__ float s_scal[4];  // dt, alpha, ratio, keep
  __shared__ int s_flag;
  const int sid = order[blockIdx.x];
  if (sid < 0 || sid >= nsys) return;
  const EnsSys S = systems[sid];
  const int t = threadIdx.x, nt = blockDim.x;
  const float cutoff_sq = S.cutoff_sq, thr_sq = S.threshold_sq;
  PairArgs a = a0;  // this system's parameter block
  if (a.sig_mode == JMD_PARAM_SPECIES) a.sig_v += S.param_offset;
  if (a.eps_mode == JMD_PARAM_SPECIES) a.eps_v += S.param_offset;
  if (a.alp_mode == JMD_PARAM_SPECIES) a.alp_v += S.param_offset;
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
    const int fill = ens_bin(S, g2, x, cellp, rankp, sorted, s_cs, s_red);
    if (t == 0) out_stat[sid] = fill;
    return;
  }

  uint8_t err = 0;
  int nbuilds = 0;
  // list build at the current positions (allocate / update)
  auto rebuild = [&](bool store) {
    const int fill = ens_bin(S, g, x, cellp, rankp, sorted, s_cs, s_red);
    const int occ = ens_rows(S, g, sp, cutoff_sq, x, cellp, sorted, s_cs, idx,
                             store ? width : 0, pad, s_red);
    if (fill > g.cap) err |= JMD_CELL_LIST_OVERFLOW;
    if (occ > S.max_occ) err |= JMD_NEIGHBOR_LIST_OVERFLOW;
    for (int q = t; q < S.n; q += nt) {
      float xi[2];
      load_pos<2>(x, S.offset + q, xi);
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
      ens_force_row(i, x, idx, width, pad, sp, a, f, &e);
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
  for (int step = 0; step < n_steps; ++step) {
    // neighbour update: rebuild when any particle moved beyond half the skin
    bool moved = false;
    for (int q = t; q < S.n; q += nt) {
      float xi[2], ri[2];
      load_pos<2>(x, S.offset + q, xi);
      load_pos<2>(x_ref, S.offset + q, ri);
      moved = moved || ref_metric_sq<2>(xi, ri, sp) > thr_sq;
    }
    if (__syncthreads_or(moved)) rebuild(true);
    __syncthreads();
    if (tape.x) {  // the list the force at node step + 1 will use
      const long long rows = (long long)S.n * width;
      int* dst = tape.idx + (long long)(step + 1) * ntot * width + (long long)S.offset * width;
      const int* src = idx + (long long)S.offset * width;
      for (long long q = t; q < rows; q += nt) dst[q] = src[q];
    }
    // velocity-Verlet sub-step at the FIRE time step (float16 value)
    const float dt = s_scal[0];
    const float half = dt / 2.0f;
    for (int k = t; k < 2 * S.n; k += nt) {
      const long long q = 2LL * S.offset + k;
      const float pk = p[q] + half * F[q];
      double xv = x_hi[q] + (double)dt * (double)pk;
      xv = ref_mod64(xv, (k & 1) ? L1 : L0);
      p[q] = pk;
      x_hi[q] = xv;
      x[q] = (float)xv;
    }
    __syncthreads();
    for (int q = t; q < S.n; q += nt) {
      const int i = S.offset + q;
      float f[2];
      ens_force_row(i, x, idx, width, pad, sp, a, f, nullptr);
      F[2LL * i] = f[0];
      F[2LL * i + 1] = f[1];
    }
    __syncthreads();
    double ff = 0.0, pp = 0.0, fp = 0.0;
    for (int k = t; k < 2 * S.n; k += nt) {
      const long long q = 2LL * S.offset + k;
      const float pk = p[q] + half * F[q];
      p[q] = pk;
      const double fv = F[q], pv = pk;
      ff += fv * fv;
      pp += pv * pv;
      fp += fv * pv;
      if (tape.x) {
        tape.q[step * n2 + q] = pk;
        tape.x[(step + 1) * n2 + q] = x[q];
        tape.F[(step + 1) * n2 + q] = F[q];
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
    }
    for (int k = t; k < 2 * S.n; k += nt) {
      const long long q = 2LL * S.offset + k;
      double pv = p[q];
      pv = pv + (double)alpha_mix * ((double)F[q] * ratio - pv);
      p[q] = (float)(pv * keep);
    }
    __syncthreads();
  }
  double e = 0.0, z1 = 0.0, z2 = 0.0;
  for (int q = t; q < S.n; q += nt) {
    float f[2];
    ens_force_row(S.offset + q, x, idx, width, pad, sp, a, f, &e);
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
__device__ double ens_hvp_rows(const EnsSys& S, const Space& sp, const PairArgs& a,
                               const float* __restrict__ dsig, const float* __restrict__ x,
                               const int* __restrict__ idx, int width, int pad,
                               const float* __restrict__ v, double* __restrict__ X, bool box,
                               double& lacc) {
#pragma clang fp contract(off)
  double tacc = 0.0;
  for (int q = threadIdx.x; q < S.n; q += blockDim.x) {
    const int i = S.offset + q;
    float xi[2], xim[2], vi[2];
    load_pos<2>(x, i, xi);
    load_pos<2>(v, i, vi);
    xim[0] = xi[0] - sp.L[0];
    xim[1] = xi[1] - sp.L[1];
    const int si = a.species ? a.species[i] : 0;
    float h0 = 0.0f, h1 = 0.0f, tp = 0.0f, lp = 0.0f;
    const int* row = idx + (long long)i * width;
    for (int k = 0; k < width; ++k) {
      const int j = row[k];
      if (j == pad) break;
      float xj[2], vj[2];
      load_pos<2>(x, j, xj);
      load_pos<2>(v, j, vj);
      float d[2], w[2];
      d[0] = min_image_disp(xj[0], xi[0], xim[0], sp.L[0], sp.half[0]);
      d[1] = min_image_disp(xj[1], xi[1], xim[1], sp.L[1], sp.half[1]);
      w[0] = vj[0] - vi[0];
      w[1] = vj[1] - vi[1];
      const float r2 = fmaf(d[1], d[1], d[0] * d[0]);
      const int sj = a.species ? a.species[j] : 0;
      const float sig = pick(a.sig_mode, a.sig, a.sig_v, i, j, si, sj, a.ns);
      const float eps = pick(a.eps_mode, a.eps, a.eps_v, i, j, si, sj, a.ns);
      const float alp = pick(a.alp_mode, a.alp, a.alp_v, i, j, si, sj, a.ns);
      const PairD2 D = pair_eval2<JMD_SOFT_SPHERE>(r2, sig, eps, alp, a);
      const float dw = fmaf(d[1], w[1], d[0] * w[0]);
      const float tt = 2.0f * D.gp * dw;
      h0 = fmaf(D.g, w[0], fmaf(tt, d[0], h0));
      h1 = fmaf(D.g, w[1], fmaf(tt, d[1], h1));
      if (dsig) {
        const float c = dsig[si * a.ns + sj];
        const float lam = fmaf(vi[1], d[1], vi[0] * d[0]);
        tp = fmaf(c * D.gs, lam, tp);
      }
      if (box) {
        const float n0 = image_of(xj[0] - xi[0], sp.half[0]);
        const float n1 = image_of(xj[1] - xi[1], sp.half[1]);
        const float dn = fmaf(d[1], n1, d[0] * n0);
        const float vn = fmaf(vi[1], n1, vi[0] * n0);
        const float vd = fmaf(vi[1], d[1], vi[0] * d[0]);
        lp = fmaf(D.g, vn, fmaf(2.0f * D.gp * dn, vd, lp));
      }
    }
    X[2LL * i] += (double)h0;
    X[2LL * i + 1] += (double)h1;
    tacc += (double)tp;
    lacc -= (double)lp;
  }
  return tacc;
}

__global__ __launch_bounds__(kEnsThreads) void k_fire_ensemble_back(
    const EnsSys* __restrict__ systems, int nsys, int ntot, int mode, int K,
    const float* __restrict__ x, const float* __restrict__ F, const int* __restrict__ idx,
    int width, PairArgs a0, const float* __restrict__ dsig0, EnsTape tape,
    double* __restrict__ X, double* __restrict__ P, float* __restrict__ W,
    float* __restrict__ V, double* __restrict__ theta, double* __restrict__ box_adj,
    float scale) {
#pragma clang fp contract(off)
  __shared__ double s_d[3 * (kEnsThreads / kWave)];
  const int sid = blockIdx.x;
  if (sid >= nsys) return;
  const EnsSys S = systems[sid];
  const int t = threadIdx.x, nt = blockDim.x;
  PairArgs a = a0;
  if (a.sig_mode == JMD_PARAM_SPECIES) a.sig_v += S.param_offset;
  if (a.eps_mode == JMD_PARAM_SPECIES) a.eps_v += S.param_offset;
  if (a.alp_mode == JMD_PARAM_SPECIES) a.alp_v += S.param_offset;
  const float* dsig = dsig0 ? dsig0 + S.param_offset : nullptr;
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
          const float sig = pick(a.sig_mode, a.sig, a.sig_v, i, j, si, sj, a.ns);
          const float eps = pick(a.eps_mode, a.eps, a.eps_v, i, j, si, sj, a.ns);
          const float alp = pick(a.alp_mode, a.alp, a.alp_v, i, j, si, sj, a.ns);
          const PairD2 D = pair_eval2<JMD_SOFT_SPHERE>(r2, sig, eps, alp, a);
          if (dsig) e += 0.5 * (double)(dsig[si * a.ns + sj] * D.us);
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
    double e = ens_hvp_rows(S, sp, a, dsig, tape.x, tape.idx, width, pad, W, X, box, el);
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
    for (int m = t; m < 2 * S.n; m += nt) {
      const long long g = b2 + m;
      const double G = (double)keep * P[g];
      const double Q = (1.0 - (double)alpha) * G + cq * (double)qk[g];
      const double Fb = (double)alpha * ratio * G - cf * (double)Fn[g] + (double)h * Q;
      P[g] = Q;  // PH until step (3)
      V[g] = (float)(Fb + (double)W[g]);
    }
    __syncthreads();
    // (2) product with the force Jacobian at node k + 1
    tsum += ens_hvp_rows(S, sp, a, dsig, tape.x + (long long)(k + 1) * n2,
                         tape.idx + (long long)(k + 1) * ntot * width, width, pad, V, X, box,
                         lsum);
    __syncthreads();
    // (3) x_{k+1} = x_k + dt ph - L m_k, ph = p_k + h F_k
    const float* xk = tape.x + (long long)k * n2;
    const float* xk1 = tape.x + (long long)(k + 1) * n2;
    for (int m = t; m < 2 * S.n; m += nt) {
      const long long g = b2 + m;
      if (box) {
        const float fold = rintf((xk[g] - xk1[g]) / sp.L[m & 1]);
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
  hipLaunchKernelGGL(k_fire_ensemble, dim3(n_systems), dim3(ens_threads(n_systems, n_total)), 0,
                     (hipStream_t)stream, reinterpret_cast<const EnsSys*>(systems), order,
                     n_systems, (int)n_total, mode, n_steps, position_hi, position, momentum,
                     force, reference_position, idx, width, scratch, a, prm, fire_state,
                     fire_n_pos, error, out_stat, energy, builds, tape);
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
  hipLaunchKernelGGL(k_fire_ensemble_back, dim3(n_systems),
                     dim3(ens_threads(n_systems, n_total)), 0,
                     (hipStream_t)stream, reinterpret_cast<const EnsSys*>(systems), n_systems,
                     (int)n_total, mode, n_steps, position, force, idx, width, a, dsigma, tape,
                     x_adjoint, p_adjoint, pending, scratch, theta, box_adjoint, scale);
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

}  // extern "C"
