// Dense-list pair kernels (templates only): row_sum, k_pair_dense and their
// launch helpers.  The host entry points live in jmd_pair.hip; the
// dispatch_dense instantiations are compiled in jmd_pair_dense_*.hip (one
// translation unit per position layout and potential, built in parallel).
#pragma once

#include "jmd_pair_common.hpp"

namespace jmd {

// Sum over one Dense row slice (entries sub, sub+TPA, ...) into f, e.
// INTERIOR: every particle of the wavefront lies at least the interaction
// range from each box face, so for any pair within range the plain
// difference x_j - x_i is already the minimum image and no wrap is needed
// (pairs whose plain difference is not the minimum image are then beyond
// range and contribute zero either way).
template <int DIM, int STRIDE, int KIND, int TPA, bool GENERAL, bool ENERGY, bool INTERIOR,
          int SMOOTH, bool BOX = false>
__device__ __forceinline__ void row_sum(const float* __restrict__ pos, unsigned n,
                                        const Space& s, const PairArgs& a,
                                        const int* __restrict__ row, int width, int sub,
                                        unsigned i, const float* xi, const float* xim,
                                        float* f, float* e) {
#pragma clang fp contract(off)
  float sig = a.sig, eps = a.eps, alp = a.alp;
  int si = 0;
  if (GENERAL && a.species) si = a.species[i];
#ifndef JMD_SPARSE_CHUNK
#define JMD_SPARSE_CHUNK 4096
#endif
#ifndef JMD_SPARSE_TABLE
#define JMD_SPARSE_TABLE 2048
#endif
#ifndef JMD_PAIR_UNROLL
#define JMD_PAIR_UNROLL 4
#endif
  constexpr int UNR = JMD_PAIR_UNROLL;
  // Interior waves of the Lennard-Jones kernel with scalar parameters gather
  // through a buffer descriptor of exactly n rows: a padding entry (j = n)
  // then reads a zero position, so d = -x_i and r^2 >= 3 r_c^2 (every
  // coordinate of an interior particle is at least r_c), where the switched
  // or truncated pair term is exactly zero -- no per-pair validity mask and
  // no index select.  The caller keeps n below 2^27 on this path.
  constexpr bool kNoMask = INTERIOR && KIND == JMD_LENNARD_JONES && !BOX && !GENERAL &&
                           DIM == 3 && STRIDE == 4;
  const __amdgpu_buffer_rsrc_t prs = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<float*>(pos), (short)0, (int)(n * 16u), 0x00020000);
  auto gather = [&](unsigned jraw, float* xj) {
    if (kNoMask) {
      const auto v = __builtin_amdgcn_raw_buffer_load_b128(prs, (int)(jraw * 16u), 0, 0);
      xj[0] = __uint_as_float(v[0]);
      xj[1] = __uint_as_float(v[1]);
      xj[2] = __uint_as_float(v[2]);
    } else {
      load_pos<DIM, STRIDE>(pos, jraw < n ? jraw : i, xj);
    }
  };
  // One pair term; padding entries (j = n) read particle i (or, without the
  // mask, a zero position) and add nothing.
  auto term = [&](unsigned jraw, const float* xj) {
    const bool ok = kNoMask || jraw < n;
    float d[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int c = 0; c < DIM; ++c)
      d[c] = (INTERIOR || !s.periodic)
                 ? xj[c] - xi[c]
                 : min_image_any(xj[c], xi[c], xim[c], s.L[c], s.half[c], s.inbox);
    if (BOX) box_apply<DIM>(s, d);  // periodic_general: fractional fold, then T
    float r2 = DIM == 3 ? fmaf(d[2], d[2], fmaf(d[1], d[1], d[0] * d[0]))
                        : fmaf(d[1], d[1], d[0] * d[0]);
    if (GENERAL) {
      unsigned j = ok ? jraw : i;
      int sj = a.species ? a.species[j] : 0;
      sig = pick(a.sig_mode, a.sig, a.sig_v, i, j, si, sj, a.ns);
      eps = pick(a.eps_mode, a.eps, a.eps_v, i, j, si, sj, a.ns);
      if (KIND == JMD_SOFT_SPHERE) alp = pick(a.alp_mode, a.alp, a.alp_v, i, j, si, sj, a.ns);
    }
    float dEr;
    float uu = pair_eval<KIND, SMOOTH>(r2, sig, eps, alp, a, &dEr);
    dEr = ok ? dEr : 0.0f;
#pragma unroll
    for (int c = 0; c < DIM; ++c) f[c] = fmaf(dEr, d[c], f[c]);
    if (ENERGY) *e += ok ? uu : 0.0f;
  };
  // This lane's entries are sub, sub + TPA, ... < width: nfull blocks of UNR
  // entries, then fewer than UNR in a tail.  In the blocks every entry lies
  // in the row, so they are read through one pointer with constant offsets;
  // the next block's entries are loaded before this block's terms are
  // computed (software pipeline: one dependent round trip per block, the
  // position gather).  Rows are left-compacted with padding n, so the row
  // ends at the first block whose entries are all padding.
  const int nent = width > sub ? (width - sub + TPA - 1) / TPA : 0;
  const int nfull = nent / UNR;
  const int* p = row + sub;
  if (nfull > 0) {
    unsigned nxt[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) nxt[u] = (unsigned)p[u * TPA];
    for (int b = 0; b < nfull; ++b) {
      unsigned jj[UNR];
      unsigned jmin = n;
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        jj[u] = nxt[u];
        jmin = min(jmin, jj[u]);
      }
      if (jmin >= n) return;
      float xj[UNR][3];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        xj[u][0] = xj[u][1] = xj[u][2] = 0.0f;
        gather(jj[u], xj[u]);
      }
      // next block, clamped to the last full one (re-read, unused, at the end)
      const int* q = p + min(b + 1, nfull - 1) * (UNR * TPA);
#pragma unroll
      for (int u = 0; u < UNR; ++u) nxt[u] = (unsigned)q[u * TPA];
      if (__builtin_amdgcn_ballot_w64(jj[UNR - 1] >= n) == 0ull) {
#pragma unroll
        for (int u = 0; u < UNR; ++u) term(jj[u], xj[u]);
      } else {
        // some row of the wavefront ends in this block: terms only while a
        // lane still holds an entry (validity is monotone in u per lane)
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
          if (u > 0 && __builtin_amdgcn_ballot_w64(jj[u] < n) == 0ull) break;
          term(jj[u], xj[u]);
        }
      }
    }
  }
  for (int t = nfull * UNR; t < nent; ++t) {
    const unsigned j = (unsigned)p[t * TPA];
    if (j >= n) return;
    float xj[3] = {0.0f, 0.0f, 0.0f};
    gather(j, xj);
    term(j, xj);
  }
}

// STEP (with KICK): after the closing half kick, the start of the next
// velocity-Verlet step for each row's particle, from the force and momentum
// in registers: the skin criterion of the positions just used against the
// list's reference positions raises kd.gate_next (the next step's update
// decision), and with kd.active the opening half kick and the drift write
// the next positions to kd.x_out (float64 master positions in place).
template <int DIM, int STRIDE, int KIND, int TPA, bool GENERAL, bool ENERGY, bool KICK,
          bool STEP = false>
__global__ __launch_bounds__(256) void k_pair_dense(
    const float* __restrict__ pos, int n, Space s, PairArgs a, const int* __restrict__ idx,
    int width, float* __restrict__ force, float* __restrict__ e_atom,
    double* __restrict__ e_part, float* __restrict__ mom, float half_dt,
    const float* __restrict__ dt_dev, KickDrift kd) {
  float kdt = kd.dt;
  if (KICK && dt_dev) {
    if (STEP) kdt = dt_dev[0];
    half_dt = dt_dev[0] / 2.0f;
  }
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int gt = tile * blockDim.x + threadIdx.x;
  const int i = gt / TPA;
  const int sub = gt % TPA;
  const bool active = i < n;
  float f[3] = {0.0f, 0.0f, 0.0f};
  float e = 0.0f;
  float xi[3] = {0, 0, 0}, xim[3] = {0, 0, 0};
  bool interior = true;
  // STEP with TPA >= DIM: lane `sub` of a row's group takes component `sub`
  // of the next step's start; its inputs are loaded here, ahead of the row
  // loop, so their latency hides behind it.
  constexpr bool SPREAD = STEP && TPA >= DIM;
  const bool mine = SPREAD && active && sub < DIM;
  double st_xh = 0.0;
  float st_ref = 0.0f, st_p = 0.0f, st_m = 0.0f;
  if (mine) {
    st_p = mom[(long long)i * DIM + sub];
    if (kd.active) {
      st_xh = kd.xhi_in[(long long)i * DIM + sub];
      if (kd.mass) st_m = kd.mass[i];
    }
    if (kd.gate_next) st_ref = kd.ref[(long long)i * DIM + sub];
  }
  if (active) {
    load_pos<DIM, STRIDE>(pos, i, xi);
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      xim[c] = xi[c] - s.L[c];
      interior = interior && xi[c] >= a.rc && xi[c] <= s.L[c] - a.rc;
    }
  }
  if (s.general && active) {
    // fractional coordinates: a neighbour within rc of particle i differs by
    // at most rc |row c of T^-1| in u_c, so no fold is needed when u_c stays
    // that far from both faces
    interior = true;
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      const float w = a.rc * s.fw[c];
      interior = interior && xi[c] >= w && xi[c] <= 1.0f - w;
    }
  }
  // (n < 2^27 keeps the interior path's buffer offsets j * 16 in range;
  // the fold-free path is exact only when every neighbour lies in the box)
  const bool wave_interior =
      s.inbox && __all(interior || !active) && n < (1 << 27);
  if (active && s.general) {
    const int* row = idx + (long long)i * width;
#define JMD_ROWG(INT, SM)                                                                      \
  row_sum<DIM, STRIDE, KIND, TPA, GENERAL, ENERGY, INT, SM, true>(pos, n, s, a, row, width, sub, \
                                                                  i, xi, xim, f, &e)
    if (KIND == JMD_LENNARD_JONES && a.smooth) {
      if (wave_interior) JMD_ROWG(true, 1); else JMD_ROWG(false, 1);
    } else {
      if (wave_interior) JMD_ROWG(true, 0); else JMD_ROWG(false, 0);
    }
#undef JMD_ROWG
  } else if (active) {
    const int* row = idx + (long long)i * width;
    // the smoothing switch is uniform: taken once here, not per pair
    if (KIND == JMD_LENNARD_JONES && a.smooth) {
      if (wave_interior)
        row_sum<DIM, STRIDE, KIND, TPA, GENERAL, ENERGY, true, 1>(pos, n, s, a, row, width, sub,
                                                                  i, xi, xim, f, &e);
      else
        row_sum<DIM, STRIDE, KIND, TPA, GENERAL, ENERGY, false, 1>(pos, n, s, a, row, width, sub,
                                                                   i, xi, xim, f, &e);
    } else {
      if (wave_interior)
        row_sum<DIM, STRIDE, KIND, TPA, GENERAL, ENERGY, true, 0>(pos, n, s, a, row, width, sub,
                                                                  i, xi, xim, f, &e);
      else
        row_sum<DIM, STRIDE, KIND, TPA, GENERAL, ENERGY, false, 0>(pos, n, s, a, row, width, sub,
                                                                   i, xi, xim, f, &e);
    }
  }
  double fd[3];
#pragma unroll
  for (int c = 0; c < DIM; ++c) fd[c] = group_sum<TPA>((double)f[c]);
  double ed = 0.0;
  if (ENERGY) ed = group_sum<TPA>((double)e);
  bool moved = false;
  if (SPREAD) {
    // the opening half kick, drift and skin test round every product and
    // sum as kick_drift_atom and ref_metric_sq do (no contraction)
#pragma clang fp contract(off)
    // lane `sub` < DIM: component sub of the row's particle
    const int c = sub;
    float fc = (float)fd[0], xc = xi[0];
#pragma unroll
    for (int q = 1; q < DIM; ++q) {
      fc = c == q ? (float)fd[q] : fc;
      xc = c == q ? xi[q] : xc;
    }
    const int base = lane_id() - sub;  // first lane of the row's group
    float pk = 0.0f;
    if (mine) {
      if (force) force[(long long)i * DIM + c] = fc;
      pk = fmaf(half_dt, fc, st_p);  // the closing half kick, fused as in the other forms
    }
    if (kd.gate_next) {
      // the skin criterion of ref_metric_sq: component terms, T for a
      // general box, squares summed in component order
      float v = 0.0f;
      if (mine) {
        v = xc - st_ref;
        if (kd.s.periodic) v = ref_wrap1(v, kd.s.L[c], kd.s.half[c]);
      }
      float vv[3];
#pragma unroll
      for (int q = 0; q < DIM; ++q) vv[q] = __shfl(v, base + q, kWave);
      if (kd.s.general) {
        box_apply_ref<DIM>(kd.s, vv);
      }
      if (sub == 0 && active) {
        float acc = 0.0f;
#pragma unroll
        for (int q = 0; q < DIM; ++q) {
          const float sq = vv[q] * vv[q];
          acc = q == 0 ? sq : acc + sq;
        }
        moved = acc > kd.thr_sq;
      }
    }
    float pn = pk;
    if (kd.active) {
      // the opening half kick and the float64 drift of kick_drift_regs
      pn = pk + half_dt * fc;
      const double w = (double)kdt * (double)pn;
      double dx = kd.mass_pow2 ? w * kd.inv_mass
                               : w / (kd.mass ? (double)st_m : (double)kd.mass_scalar);
      if (kd.s.general) {
        double dv[3];
#pragma unroll
        for (int q = 0; q < DIM; ++q) dv[q] = __shfl(dx, base + q, kWave);
        double acc = 0.0;
#pragma unroll
        for (int b = 0; b < DIM; ++b) {
          float t = kd.Ti[0][b];
#pragma unroll
          for (int q = 1; q < DIM; ++q) t = c == q ? kd.Ti[q][b] : t;
          acc += (double)t * dv[b];
        }
        dx = acc;
      }
      if (mine) {
        double x = st_xh + dx;
        if (kd.s.periodic) {
          float Lc = kd.s.L[0];
#pragma unroll
          for (int q = 1; q < DIM; ++q) Lc = c == q ? kd.s.L[q] : Lc;
          x = ref_mod64(x, (double)Lc);
        }
        kd.xhi_out[(long long)i * DIM + c] = x;
        kd.x_out[(long long)i * STRIDE + c] = (float)x;
      }
    }
    if (mine) mom[(long long)i * DIM + c] = pn;
  }
  if (active && sub == 0 && !SPREAD) {
    float fcv[3] = {0.0f, 0.0f, 0.0f}, pk[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      float fc = (float)fd[c];
      if (force) force[(long long)i * DIM + c] = fc;
      if (KICK && !STEP) mom[(long long)i * DIM + c] = fmaf(half_dt, fc, mom[(long long)i * DIM + c]);
      if (STEP) {
        fcv[c] = fc;
        pk[c] = fmaf(half_dt, fc, mom[(long long)i * DIM + c]);
      }
    }
    if (STEP) {
      if (kd.gate_next) {
        float r[3];
        load_pos<DIM>(kd.ref, i, r);
        moved = ref_metric_sq<DIM>(xi, r, kd.s) > kd.thr_sq;
      }
      if (kd.active) {
        float xn[3];
        kick_drift_regs<DIM>(kd, i, kdt, half_dt, pk, fcv, xn);
        store_pos<DIM, STRIDE>(kd.x_out, i, xn);
      }
#pragma unroll
      for (int c = 0; c < DIM; ++c) mom[(long long)i * DIM + c] = pk[c];
    }
  }
  if (ENERGY && active && sub == 0 && e_atom) e_atom[i] = (float)(0.5 * ed);
  if (STEP && kd.gate_next) {
    if (__ballot(moved) != 0ull && lane_id() == 0) atomicOr(kd.gate_next, 1);
  }
  if (ENERGY) {
    // block partial of 0.5 * sum_i e_i in float64
    __shared__ double s_e[256 / kWave];
    double v = (active && sub == 0) ? 0.5 * ed : 0.0;
    v = wave_reduce_sum(v);
    if (lane_id() == 0) s_e[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int w = 0; w < (int)(blockDim.x / kWave); ++w) t += s_e[w];
      e_part[tile] = t;
    }
  }
}

struct DenseLaunch {
  dim3 grid;
  hipStream_t st;
  const float* pos;
  int n;
  Space s;
  PairArgs a;
  const int* idx;
  int width;
  float* force;
  float* e_atom;
  double* e_part;
  float* mom;
  float half_dt;
  const float* dt_dev;
  KickDrift kd;  // the next step's start (STEP launches only)
};

template <int DIM, int STRIDE, int KIND, int TPA, bool GENERAL>
void launch_dense(const DenseLaunch& L, bool energy, bool kick, bool step) {
#define JMD_PD(E, K, S)                                                                      \
  hipLaunchKernelGGL((k_pair_dense<DIM, STRIDE, KIND, TPA, GENERAL, E, K, S>), L.grid,       \
                     dim3(256), 0, L.st, L.pos, L.n, L.s, L.a, L.idx, L.width, L.force,      \
                     L.e_atom, L.e_part, L.mom, L.half_dt, L.dt_dev, L.kd)
  if (energy) JMD_PD(true, false, false);
  else if (step) JMD_PD(false, true, true);
  else if (kick) JMD_PD(false, true, false);
  else JMD_PD(false, false, false);
#undef JMD_PD
}

template <int DIM, int STRIDE, int KIND>
void dispatch_dense(int tpa, bool general, const DenseLaunch& L, bool energy, bool kick,
                    bool step) {
#define JMD_TPA(T)                                                               \
  (general ? launch_dense<DIM, STRIDE, KIND, T, true>(L, energy, kick, step)     \
           : launch_dense<DIM, STRIDE, KIND, T, false>(L, energy, kick, step))
  switch (tpa) {
    case 2: JMD_TPA(2); break;
    case 8: JMD_TPA(8); break;
    case 16: JMD_TPA(16); break;
    default: JMD_TPA(4); break;
  }
#undef JMD_TPA
}

// Explicit instantiations live in jmd_pair_dense_*.hip.
#define JMD_DENSE_INSTANCES(X)                      \
  X(3, 4, JMD_LENNARD_JONES) X(3, 4, JMD_SOFT_SPHERE) \
  X(3, 3, JMD_LENNARD_JONES) X(3, 3, JMD_SOFT_SPHERE) \
  X(2, 2, JMD_LENNARD_JONES) X(2, 2, JMD_SOFT_SPHERE)
#define JMD_DENSE_EXTERN(D, S, K)                                                     \
  extern template void dispatch_dense<D, S, K>(int, bool, const DenseLaunch&, bool, bool, \
                                               bool);
#define JMD_DENSE_INSTANTIATE(D, S, K)                                         \
  template void dispatch_dense<D, S, K>(int, bool, const DenseLaunch&, bool, bool, bool);

}  // namespace jmd
