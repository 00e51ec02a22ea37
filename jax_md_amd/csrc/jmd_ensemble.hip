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

constexpr int kEnsThreads = 512;
constexpr int kEnsMaxCells = 4096;

struct EnsSys {
  int offset, n;
  int cps[2];
  float cell_size[2];
  int fill_cps[2];
  float fill_cell_size[2];
  float side[2];
  int cap, max_occ;
};

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
// order (the stable argsort of partition.py:432).  Returns the max filling.
__device__ int ens_bin(const EnsSys& S, const Grid& g, const float* __restrict__ x,
                       int* __restrict__ cellp, int* __restrict__ rankp,
                       int* __restrict__ sorted, int* s_cnt, int* s_start, int* s_red) {
  const int nt = blockDim.x, t = threadIdx.x;
  for (int c = t; c < g.ncell; c += nt) s_cnt[c] = 0;
  __syncthreads();
  for (int a = t; a < S.n; a += nt) {
    float xi[2];
    load_pos<2>(x, S.offset + a, xi);
    int ci[2];
    const int h = cell_of<2>(xi, g, ci);
    cellp[a] = h;
    rankp[a] = atomicAdd(&s_cnt[h], 1);
  }
  __syncthreads();
  // exclusive scan of the counts: per-thread chunks, a wavefront scan of the
  // chunk totals, then the wave totals in order
  const int per = (g.ncell + nt - 1) / nt;
  const int beg = min(t * per, g.ncell), end = min(beg + per, g.ncell);
  int local = 0, fill = 0;
  for (int c = beg; c < end; ++c) {
    local += s_cnt[c];
    fill = max(fill, s_cnt[c]);
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
    s_start[c] = run;
    run += s_cnt[c];
  }
  if (t == 0) s_start[g.ncell] = s_wave[nw];
  __syncthreads();
  for (int a = t; a < S.n; a += nt) sorted[s_start[cellp[a]] + rankp[a]] = a;
  __syncthreads();
  // order each cell by id (thread per cell, insertion sort of a few ids)
  for (int c = t; c < g.ncell; c += nt) {
    int* m = sorted + s_start[c];
    const int cnt = s_cnt[c];
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
                        const int* __restrict__ sorted, const int* s_cnt, const int* s_start,
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
      const int start = s_start[c];
      const int m = min(s_cnt[c], g.cap);
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
    int* __restrict__ idx, int width, int* __restrict__ scratch, PairArgs a, float cutoff_sq,
    float thr_sq, FireP prm, float* __restrict__ fire_state, int* __restrict__ fire_npos,
    uint8_t* __restrict__ error, int* __restrict__ out_stat, double* __restrict__ energy,
    int* __restrict__ builds) {
  __shared__ int s_cnt[kEnsMaxCells];
  __shared__ int s_start[kEnsMaxCells + 1];
  __shared__ int s_red[kEnsThreads / kWave];
  __shared__ double s_d[3 * (kEnsThreads / kWave)];
  __shared__ float s_scal[4];  // dt, alpha, ratio, keep
  __shared__ int s_flag;
  const int sid = order[blockIdx.x];
  if (sid < 0 || sid >= nsys) return;
  const EnsSys S = systems[sid];
  const int t = threadIdx.x, nt = blockDim.x;
  int* cellp = scratch + S.offset;
  int* rankp = scratch + ntot + S.offset;
  int* sorted = scratch + 2LL * ntot + S.offset;
  Space sp;
  sp.dim = 2;
  sp.periodic = 1;
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
    const int fill = ens_bin(S, g2, x, cellp, rankp, sorted, s_cnt, s_start, s_red);
    if (t == 0) out_stat[sid] = fill;
    return;
  }

  uint8_t err = 0;
  int nbuilds = 0;
  // list build at the current positions (allocate / update)
  auto rebuild = [&](bool store) {
    const int fill = ens_bin(S, g, x, cellp, rankp, sorted, s_cnt, s_start, s_red);
    const int occ = ens_rows(S, g, sp, cutoff_sq, x, cellp, sorted, s_cnt, s_start, idx,
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
    }
    wg_sum3(ff, pp, fp, s_d);
    // FIRE bookkeeping (minimize.py:190-219) with float16 dt / alpha
    const double ratio = sqrt(pp) / sqrt(ff + 1e-6);
    const float alpha_mix = s_scal[1];
    const double keep = fp >= 0.0 ? 1.0 : 0.0;
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

}  // namespace jmd

using namespace jmd;

extern "C" {

int jmd_fire_ensemble_2d(const jmd_system2d* systems, const int32_t* order, int32_t n_systems,
                         int64_t n_total, int32_t mode, int32_t n_steps, double* position_hi,
                         float* position, float* momentum, float* force,
                         float* reference_position, int32_t* idx, int32_t width,
                         int32_t* scratch, const jmd_pair_potential* potential,
                         float cutoff_sq, float threshold_sq, const float* fire_params,
                         float* fire_state, int32_t* fire_n_pos, uint8_t* error,
                         int32_t* out_stat, double* energy, int32_t* builds, void* stream) {
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
  hipLaunchKernelGGL(k_fire_ensemble, dim3(n_systems), dim3(kEnsThreads), 0,
                     (hipStream_t)stream, reinterpret_cast<const EnsSys*>(systems), order,
                     n_systems, (int)n_total, mode, n_steps, position_hi, position, momentum,
                     force, reference_position, idx, width, scratch, a, cutoff_sq,
                     threshold_sq, prm, fire_state, fire_n_pos, error, out_stat, energy,
                     builds);
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

int jmd_fire_ensemble_max_cells(void) { return kEnsMaxCells; }

}  // extern "C"
