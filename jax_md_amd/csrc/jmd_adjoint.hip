// Adjoint of the pair force over a Dense neighbour list for gfx950.
//
// The reference differentiates through the force with jax.grad
// (quantity.py:58-60, simulate.py:227-243).  Here the vector-Jacobian product
// of F(R; sigma, epsilon) = -dE/dR is a kernel of its own: for a cotangent v,
//
//   (J v)_i = sum_j [ g w + 2 g' (d . w) d ],     d = R_j - R_i,  w = v_j - v_i,
//
// with g(s) = U'(r)/r and g' = dg/ds as functions of s = r^2 (J = dF/dR is
// symmetric, so the VJP is this product), and the parameter cotangents
//
//   v^T dF/dtheta = sum_i sum_j (dg_ij/dtheta) (v_i . d_ij)
//
// for scalar sigma and epsilon (r_onset and r_cutoff stay fixed, as the
// reference fixes them at construction, energy.py:319-320).  Rows are owned by
// TPA lanes exactly as in the force kernel: register accumulation, shuffle
// reduction, no atomics, deterministic.
//
// For a velocity-Verlet trajectory (x_{k+1} = x_k + dt ph_k / m,
// ph_k = p_k + dt/2 F(x_k), p_{k+1} = ph_k + dt/2 F(x_{k+1})) the backward
// step needs one product J(x_k) PH_k per step: the product at x_{k+1} with
// P_{k+1} = PH_{k+1} serves both adjacent steps.  k_pair_adjoint's step mode
// fuses that product with the elementwise recursion
//   X_k = X_{k+1} + c_x J(x_k) PH_k,   PH_{k-1} = PH_k + (dt/m) X_k,
// and accumulates the parameter cotangents per workgroup.
#include "jmd_pair_common.hpp"

namespace jmd {

enum { ADJ_VJP = 0, ADJ_STEP = 1 };

// Per-particle parameters p_i, combined as p_ij = (p_i + p_j) / 2
// (smap.py:836), differentiate row by row: with a symmetric list and
// symmetric pair terms,
//   v^T dF/dp_k = -1/2 sum_j (dg_kj/dp) (d_kj . w_kj),   w = v_j - v_k,
//   dE/dp_k     =  1/2 sum_j dU_kj/dp,
// so each row's owner writes particle k's entry (no scatter, deterministic).
//
// STRIDE 0: positions are float64 [n, DIM] (an x64 trajectory); each
// displacement is formed in float64 and rounded once, the rest is float32.
// One launch over all rows.  ADJ_VJP: hout = J v, part[4 b + k] = block sums
// of (v^T dF/dsigma, v^T dF/deps, dE/dsigma, dE/deps) for scalar parameters;
// per-particle ones go to pp_out[k n + i] (same four quantities).  ADJ_STEP:
// X += c_x (J v) (float64), vout = v + c_p X, part[2 b + k] += c_t (v^T dF/dtheta)
// for scalar parameters, pp_out[k n + i] += c_t (v^T dF/dp_i) for per-particle
// ones (k = 0 sigma, 1 epsilon).
template <int DIM, int STRIDE, int KIND, int TPA, bool GENERAL, int MODE>
__global__ __launch_bounds__(256) void k_pair_adjoint(
    const float* __restrict__ pos, int n, Space s, PairArgs a, const int* __restrict__ idx,
    int width, const float* __restrict__ vin, float* __restrict__ hout,
    double* __restrict__ part, int want_params, double* __restrict__ xadj,
    float* __restrict__ vout, float cx, float cp, float ct, double* __restrict__ pp_out) {
#pragma clang fp contract(off)
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int gt = tile * blockDim.x + threadIdx.x;
  const int i = gt / TPA;
  const int sub = gt % TPA;
  const bool active = i < n;
  float h[3] = {0.0f, 0.0f, 0.0f};
  float ts = 0.0f, te = 0.0f, us = 0.0f, ue = 0.0f;
  float ps = 0.0f, pe = 0.0f;  // sum_j (dg/dp) (d . w), per-particle parameters
  const bool pp_s = GENERAL && a.sig_mode == JMD_PARAM_PER_PARTICLE;
  const bool pp_e = GENERAL && a.eps_mode == JMD_PARAM_PER_PARTICLE;
  float xi[3] = {0, 0, 0}, xim[3] = {0, 0, 0}, vi[3] = {0, 0, 0};
  double xi64[3] = {0.0, 0.0, 0.0};
  const double* pos64 = reinterpret_cast<const double*>(pos);
  if (active) {
    if (STRIDE == 0) {
#pragma unroll
      for (int c = 0; c < DIM; ++c) xi64[c] = pos64[(long long)i * DIM + c];
    } else {
      load_pos<DIM, STRIDE>(pos, i, xi);
    }
    load_pos<DIM>(vin, i, vi);
#pragma unroll
    for (int c = 0; c < DIM; ++c) xim[c] = xi[c] - s.L[c];
  }
  if (active && width > 0) {
    const int* row = idx + (long long)i * width;
    float sig = a.sig, eps = a.eps, alp = a.alp;
    int si = 0;
    if (GENERAL && a.species) si = a.species[i];
    for (int k = sub; k < width; k += TPA) {
      const unsigned jr = (unsigned)row[k];
      if (jr >= (unsigned)n) break;  // left-compacted rows, padding n
      const unsigned j = jr;
      float xj[3] = {0, 0, 0}, vj[3] = {0, 0, 0};
      if (STRIDE != 0) load_pos<DIM, STRIDE>(pos, j, xj);
      load_pos<DIM>(vin, j, vj);
      float d[3] = {0.0f, 0.0f, 0.0f}, w[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        if (STRIDE == 0) {
          const double v = pos64[(long long)j * DIM + c] - xi64[c];
          d[c] = (float)(s.periodic ? wrap64(v, (double)s.L[c]) : v);
        } else {
          d[c] = s.periodic ? min_image_any(xj[c], xi[c], xim[c], s.L[c], s.half[c], s.inbox)
                            : xj[c] - xi[c];
        }
        w[c] = vj[c] - vi[c];
      }
      const float r2 = DIM == 3 ? fmaf(d[2], d[2], fmaf(d[1], d[1], d[0] * d[0]))
                                : fmaf(d[1], d[1], d[0] * d[0]);
      if (GENERAL) {
        const int sj = a.species ? a.species[j] : 0;
        sig = pick(a.sig_mode, a.sig, a.sig_v, i, j, si, sj, a.ns);
        eps = pick(a.eps_mode, a.eps, a.eps_v, i, j, si, sj, a.ns);
        if (KIND == JMD_SOFT_SPHERE) alp = pick(a.alp_mode, a.alp, a.alp_v, i, j, si, sj, a.ns);
      }
      const PairD2 D = pair_eval2<KIND>(r2, sig, eps, alp, a);
      const float dw = DIM == 3 ? fmaf(d[2], w[2], fmaf(d[1], w[1], d[0] * w[0]))
                                : fmaf(d[1], w[1], d[0] * w[0]);
      const float t = 2.0f * D.gp * dw;
#pragma unroll
      for (int c = 0; c < DIM; ++c) h[c] = fmaf(D.g, w[c], fmaf(t, d[c], h[c]));
      if (want_params) {
        const float lam = DIM == 3 ? fmaf(vi[2], d[2], fmaf(vi[1], d[1], vi[0] * d[0]))
                                   : fmaf(vi[1], d[1], vi[0] * d[0]);
        ts = fmaf(D.gs, lam, ts);
        te = fmaf(D.ge, lam, te);
        us += D.us;
        ue += D.ue;
        if (GENERAL) {
          ps = fmaf(D.gs, dw, ps);
          pe = fmaf(D.ge, dw, pe);
        }
      }
    }
  }
  double hd[3];
#pragma unroll
  for (int c = 0; c < DIM; ++c) hd[c] = group_sum<TPA>((double)h[c]);
  double pd[4] = {0.0, 0.0, 0.0, 0.0};
  if (want_params) {
    pd[0] = group_sum<TPA>((double)ts);
    pd[1] = group_sum<TPA>((double)te);
    pd[2] = group_sum<TPA>((double)us);
    pd[3] = group_sum<TPA>((double)ue);
  }
  double pq[2] = {0.0, 0.0};
  if (GENERAL && want_params) {
    pq[0] = group_sum<TPA>((double)ps);
    pq[1] = group_sum<TPA>((double)pe);
  }
  const bool owner = active && sub == 0;
  if (GENERAL && want_params && owner) {
    // per-particle parameters: this row's entries (scalar ones go to part)
    const bool pp[2] = {pp_s, pp_e};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (!pp[k]) continue;
      const double vjp = -0.5 * pq[k];
      if (MODE == ADJ_VJP) {
        pp_out[(long long)k * n + i] = vjp;
        pp_out[(long long)(2 + k) * n + i] = 0.5 * pd[2 + k];
      } else {
        pp_out[(long long)k * n + i] += (double)ct * vjp;
      }
    }
    if (pp_s) {
      pd[0] = 0.0;
      pd[2] = 0.0;
    }
    if (pp_e) {
      pd[1] = 0.0;
      pd[3] = 0.0;
    }
  }
  if (owner) {
    if (MODE == ADJ_VJP) {
#pragma unroll
      for (int c = 0; c < DIM; ++c) hout[(long long)i * DIM + c] = (float)hd[c];
    } else {
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        const long long q = (long long)i * DIM + c;
        const double X = xadj[q] + (double)cx * hd[c];
        xadj[q] = X;
        vout[q] = (float)((double)vin[q] + (double)cp * X);
      }
    }
  }
  if (want_params) {
    __shared__ double s_p[4][256 / kWave];
    const int nv = MODE == ADJ_VJP ? 4 : 2;
    for (int k = 0; k < nv; ++k) {
      double v = owner ? pd[k] : 0.0;
      v = wave_reduce_sum(v);
      if (lane_id() == 0) s_p[k][threadIdx.x / kWave] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int k = 0; k < nv; ++k) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x / kWave); ++w) t += s_p[k][w];
        if (MODE == ADJ_VJP) part[4 * tile + k] = k < 2 ? t : 0.5 * t;  // E = sum / 2
        else part[2 * tile + k] += (double)ct * t;
      }
    }
  }
}

// Fixed-order column sums of block partials (one workgroup):
// sums[k] = sum_b part[ncol b + k]; then out_a[0..1] = sums[0..1],
// out_b[0..1] = sums[2..3] (either may be null).
__global__ __launch_bounds__(256) void k_sum_columns(const double* __restrict__ part, int nb,
                                                     int ncol, double* __restrict__ out_a,
                                                     double* __restrict__ out_b) {
  __shared__ double s[256];
  for (int k = 0; k < ncol; ++k) {
    double v = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) v += part[(long long)ncol * b + k];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
      if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      if (k < 2 && out_a) out_a[k] = s[0];
      if (k >= 2 && out_b) out_b[k - 2] = s[0];
    }
    __syncthreads();
  }
}

// The parameter outputs in the layout [sigma block][epsilon block], each 1
// value (scalar parameter: the column sum `tot[k]`) or n values (per-particle:
// pp[k n + i]); `koff` selects the columns (0: force VJP / step
// accumulation, 2: energy gradient).
__global__ __launch_bounds__(256) void k_param_assemble(const double* __restrict__ tot,
                                                        const double* __restrict__ pp, int n,
                                                        int pp_s, int pp_e, int koff,
                                                        double* __restrict__ out) {
  const int ns = pp_s ? n : 1, ne = pp_e ? n : 1;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < ns + ne;
       t += (long long)gridDim.x * blockDim.x) {
    const int k = t < ns ? 0 : 1;
    const long long q = t < ns ? t : t - ns;
    const bool per = k == 0 ? pp_s : pp_e;
    out[t] = per ? pp[(long long)(koff + k) * n + q] : tot[koff + k];
  }
}

constexpr int kAdjTPA = 8;  // lanes per row (3-D); 2-D uses 4

static int adj_tpa(int dim) { return dim == 3 ? kAdjTPA : 4; }

static int adj_blocks(long long n, int dim) {
  int nb = ceil_div(n * adj_tpa(dim), 256);
  return (nb + 7) & ~7;
}

struct AdjLaunch {
  dim3 grid;
  hipStream_t st;
  const float* pos;
  int n;
  Space s;
  PairArgs a;
  const int* idx;
  int width;
  const float* vin;
  float* hout;
  double* part;
  int want_params;
  double* xadj;
  float* vout;
  float cx, cp, ct;
  double* pp_out;
};

template <int DIM, int STRIDE, int KIND, int TPA, int MODE>
static void launch_adj(const AdjLaunch& L) {
  if (L.a.general)
    hipLaunchKernelGGL((k_pair_adjoint<DIM, STRIDE, KIND, TPA, true, MODE>), L.grid, dim3(256), 0,
                       L.st, L.pos, L.n, L.s, L.a, L.idx, L.width, L.vin, L.hout, L.part,
                       L.want_params, L.xadj, L.vout, L.cx, L.cp, L.ct, L.pp_out);
  else
    hipLaunchKernelGGL((k_pair_adjoint<DIM, STRIDE, KIND, TPA, false, MODE>), L.grid, dim3(256),
                       0, L.st, L.pos, L.n, L.s, L.a, L.idx, L.width, L.vin, L.hout, L.part,
                       L.want_params, L.xadj, L.vout, L.cx, L.cp, L.ct, L.pp_out);
}

template <int MODE>
static void dispatch_adj(int dim, int stride, const AdjLaunch& L) {
  if (dim == 3) {
    if (L.a.kind == JMD_LENNARD_JONES) {
      if (stride == 4) launch_adj<3, 4, JMD_LENNARD_JONES, kAdjTPA, MODE>(L);
      else if (stride == 0) launch_adj<3, 0, JMD_LENNARD_JONES, kAdjTPA, MODE>(L);
      else launch_adj<3, 3, JMD_LENNARD_JONES, kAdjTPA, MODE>(L);
    } else {
      if (stride == 4) launch_adj<3, 4, JMD_SOFT_SPHERE, kAdjTPA, MODE>(L);
      else if (stride == 0) launch_adj<3, 0, JMD_SOFT_SPHERE, kAdjTPA, MODE>(L);
      else launch_adj<3, 3, JMD_SOFT_SPHERE, kAdjTPA, MODE>(L);
    }
  } else {
    if (L.a.kind == JMD_LENNARD_JONES) {
      if (stride == 0) launch_adj<2, 0, JMD_LENNARD_JONES, 4, MODE>(L);
      else launch_adj<2, 2, JMD_LENNARD_JONES, 4, MODE>(L);
    } else {
      if (stride == 0) launch_adj<2, 0, JMD_SOFT_SPHERE, 4, MODE>(L);
      else launch_adj<2, 2, JMD_SOFT_SPHERE, 4, MODE>(L);
    }
  }
}

static int adj_check(const float* position, int64_t n, int32_t dim, int32_t position_stride,
                     int32_t periodic, const float* side, const jmd_pair_potential* potential,
                     const int32_t* idx, int64_t idx_width, PairArgs* a) {
  if ((dim != 2 && dim != 3) || n < 0 || n > 0x7fffffff || idx_width < 0 ||
      idx_width > 0x7fffffff || (n > 0 && !position) || (idx_width > 0 && n > 0 && !idx) ||
      (periodic && !side))
    return JMD_ERR_INVALID;
  if (n > 0 && position_stride != 0 && !stride_ok(dim, position_stride, position))
    return JMD_ERR_INVALID;
  if ((long long)n * idx_width > 0x7fffffffffffLL) return JMD_ERR_INVALID;
  if (!make_pair_args(potential, a)) return JMD_ERR_INVALID;
  return JMD_OK;
}

}  // namespace jmd

using namespace jmd;

extern "C" {

size_t jmd_adjoint_workspace_bytes(int64_t n, int32_t dim) {
  if (n < 0 || (dim != 2 && dim != 3)) return 0;
  // block partials, per-particle outputs [4][n], column sums [4]
  return ((size_t)adj_blocks(n, dim) * 4 + 4 * (size_t)n + 4) * sizeof(double) + 256;
}

int64_t jmd_adjoint_param_count(const jmd_pair_potential* potential, int64_t n) {
  if (!potential || n < 0) return -1;
  int64_t c = 0;
  for (const jmd_param* p : {&potential->sigma, &potential->epsilon}) {
    if (p->mode == JMD_PARAM_SCALAR) c += 1;
    else if (p->mode == JMD_PARAM_PER_PARTICLE) c += n;
    else return -1;
  }
  return c;
}

int jmd_pair_force_vjp(const float* position, int64_t n, int32_t dim, int32_t position_stride,
                       int32_t periodic, const float* side, const jmd_pair_potential* potential,
                       const int32_t* idx, int64_t idx_width, int32_t format,
                       const float* cotangent, float* position_cotangent,
                       double* param_cotangent, double* energy_param_grad, void* workspace,
                       void* stream) {
  PairArgs a;
  int rc = adj_check(position, n, dim, position_stride, periodic, side, potential, idx,
                     idx_width, &a);
  if (rc) return rc;
  if (format != JMD_DENSE) return JMD_ERR_UNSUPPORTED;
  const bool params = param_cotangent != nullptr || energy_param_grad != nullptr;
  // scalar or per-particle sigma / epsilon (species matrices: unsupported)
  if (params && jmd_adjoint_param_count(potential, n) < 0) return JMD_ERR_UNSUPPORTED;
  if (n > 0 && !cotangent) return JMD_ERR_INVALID;
  if (params && !workspace) return JMD_ERR_INVALID;
  hipStream_t st = (hipStream_t)stream;
  const int nb = adj_blocks(n, dim);
  double* part = (double*)workspace;
  double* pp = part + (long long)nb * 4;
  double* tot = pp + 4 * (long long)n;
  const int pp_s = a.sig_mode == JMD_PARAM_PER_PARTICLE;
  const int pp_e = a.eps_mode == JMD_PARAM_PER_PARTICLE;
  if (n == 0) {
    double* outs[2] = {param_cotangent, energy_param_grad};
    for (double* o : outs)
      if (o) {
        rc = launch_fill<double>(o, 2, 0.0, st);
        if (rc) return rc;
      }
    return JMD_OK;
  }
  float* hout = position_cotangent;
  if (!hout) return JMD_ERR_INVALID;  // the product is the kernel's main output
  {
    const int grc = space_check(dim, periodic, side, false);
    if (grc) return grc;
  }
  Space s = make_space(dim, periodic, side);
  AdjLaunch L{dim3(nb), st, position, (int)n, s, a, idx, (int)idx_width, cotangent, hout,
              part, params ? 1 : 0, nullptr, nullptr, 0.0f, 0.0f, 0.0f, pp};
  dispatch_adj<ADJ_VJP>(dim, position_stride, L);
  JMD_LAUNCH_CHECK();
  if (params) {
    hipLaunchKernelGGL(k_sum_columns, dim3(1), dim3(256), 0, st, part, nb, 4, tot, tot + 2);
    JMD_LAUNCH_CHECK();
    const long long m = (pp_s ? n : 1) + (pp_e ? n : 1);
    const unsigned g = (unsigned)((m + 255) / 256 < 1024 ? (m + 255) / 256 : 1024);
    if (param_cotangent)
      hipLaunchKernelGGL(k_param_assemble, dim3(g), dim3(256), 0, st, tot, pp, (int)n, pp_s, pp_e,
                         0, param_cotangent);
    if (energy_param_grad)
      hipLaunchKernelGGL(k_param_assemble, dim3(g), dim3(256), 0, st, tot, pp, (int)n, pp_s, pp_e,
                         2, energy_param_grad);
    JMD_LAUNCH_CHECK();
  }
  return JMD_OK;
}

size_t jmd_adjoint_param_acc_doubles(int64_t n, int32_t dim) {
  if (n < 0 || (dim != 2 && dim != 3)) return 0;
  // block partials (scalar parameters), per-particle accumulators [2][n], sums [2]
  return (size_t)adj_blocks(n, dim) * 2 + 2 * (size_t)n + 2;
}

int jmd_nve_adjoint_step(const float* position, int64_t n, int32_t dim, int32_t position_stride,
                         int32_t periodic, const float* side,
                         const jmd_pair_potential* potential, const int32_t* idx,
                         int64_t idx_width, const float* ph_in, float* ph_out,
                         double* x_adjoint, float coef_x, float coef_p, double* param_acc,
                         float coef_param, void* stream) {
  PairArgs a;
  int rc = adj_check(position, n, dim, position_stride, periodic, side, potential, idx,
                     idx_width, &a);
  if (rc) return rc;
  if (n == 0) return JMD_OK;
  if (!ph_in || !ph_out || !x_adjoint || ph_in == ph_out) return JMD_ERR_INVALID;
  if (param_acc && jmd_adjoint_param_count(potential, n) < 0) return JMD_ERR_UNSUPPORTED;
  hipStream_t st = (hipStream_t)stream;
  {
    const int grc = space_check(dim, periodic, side, false);
    if (grc) return grc;
  }
  Space s = make_space(dim, periodic, side);
  AdjLaunch L{dim3(adj_blocks(n, dim)), st, position, (int)n, s, a, idx, (int)idx_width, ph_in,
              nullptr, param_acc, param_acc ? 1 : 0, x_adjoint, ph_out, coef_x, coef_p,
              coef_param, param_acc ? param_acc + (long long)adj_blocks(n, dim) * 2 : nullptr};
  dispatch_adj<ADJ_STEP>(dim, position_stride, L);
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

int jmd_nve_adjoint_step_x64(const double* position, int64_t n, int32_t dim, int32_t periodic,
                             const float* side, const jmd_pair_potential* potential,
                             const int32_t* idx, int64_t idx_width, const float* ph_in,
                             float* ph_out, double* x_adjoint, float coef_x, float coef_p,
                             double* param_acc, float coef_param, void* stream) {
  // stride code 0: float64 [n, dim] positions (k_pair_adjoint<.., 0, ..>)
  return jmd_nve_adjoint_step(reinterpret_cast<const float*>(position), n, dim, 0, periodic,
                              side, potential, idx, idx_width, ph_in, ph_out, x_adjoint, coef_x,
                              coef_p, param_acc, coef_param, stream);
}

int jmd_adjoint_param_sum(const double* param_acc, int64_t n, int32_t dim, double* out,
                          void* stream) {
  if (n < 0 || (dim != 2 && dim != 3) || !param_acc || !out) return JMD_ERR_INVALID;
  hipLaunchKernelGGL(k_sum_columns, dim3(1), dim3(256), 0, (hipStream_t)stream, param_acc,
                     adj_blocks(n, dim), 2, out, nullptr);
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

int jmd_adjoint_param_gather(double* param_acc, int64_t n, int32_t dim,
                             const jmd_pair_potential* potential, double* out, void* stream) {
  if (n < 0 || (dim != 2 && dim != 3) || !param_acc || !out) return JMD_ERR_INVALID;
  if (jmd_adjoint_param_count(potential, n) < 0) return JMD_ERR_UNSUPPORTED;
  hipStream_t st = (hipStream_t)stream;
  const int nb = adj_blocks(n, dim);
  double* pp = param_acc + (long long)nb * 2;
  double* tot = pp + 2 * (long long)n;
  hipLaunchKernelGGL(k_sum_columns, dim3(1), dim3(256), 0, st, param_acc, nb, 2, tot, nullptr);
  JMD_LAUNCH_CHECK();
  const int pp_s = potential->sigma.mode == JMD_PARAM_PER_PARTICLE;
  const int pp_e = potential->epsilon.mode == JMD_PARAM_PER_PARTICLE;
  const long long m = (pp_s ? n : 1) + (pp_e ? n : 1);
  const unsigned g = (unsigned)((m + 255) / 256 < 1024 ? (m + 255) / 256 : 1024);
  hipLaunchKernelGGL(k_param_assemble, dim3(g > 0 ? g : 1), dim3(256), 0, st, tot, pp, (int)n,
                     pp_s, pp_e, 0, out);
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

}  // extern "C"
