// The float64 ("x64") form of the NVE path for gfx950: the reference's
// JAX_ENABLE_X64 run, where positions, momenta, forces and the pair
// arithmetic are float64 (jax_md/energy.py:246-272,534-580,
// smap.py:856-979, simulate.py:227-243), while the parameters the reference
// casts with maybe_downcast stay float32-valued (util.py:109-112).  Used
// where float32 forces are not enough: a gradient through a long chaotic
// trajectory, whose rounding the trajectory amplifies exponentially.
//
// Dense lists only, built as usual from the float32 mirror of the positions
// (the list is a superset of the pairs within r_c; which pairs it holds does
// not change the forces).
#include "jmd_pair_common.hpp"

namespace jmd {

// U(r) and U'(r)/r from r^2 in float64, the reference's piecewise form.
template <int KIND>
__device__ __forceinline__ double pair_eval64(double r2, double sig, double eps, double alp,
                                              const PairArgs& a, double* dEr) {
  if (KIND == JMD_LENNARD_JONES) {
    const double rc2 = (double)a.rc * (double)a.rc, ro2 = (double)a.ro * (double)a.ro;
    if (!(r2 < rc2)) {
      *dEr = 0.0;
      return 0.0;
    }
    const double inv = 1.0 / r2;
    const double t = sig * sig * inv;
    const double t3 = t * t * t;
    const double t6 = t3 * t3;
    const double V = 4.0 * eps * (t6 - t3);
    const double dV = -24.0 * eps * (2.0 * t6 - t3) * inv;  // V'(r) / r
    if (a.smooth && r2 >= ro2) {
      const double D = (rc2 - ro2) * (rc2 - ro2) * (rc2 - ro2);
      const double u = rc2 - r2;
      const double S = u * u * (rc2 + 2.0 * r2 - 3.0 * ro2) / D;
      const double dS = 12.0 * u * (ro2 - r2) / D;  // S'(r) / r
      *dEr = S * dV + dS * V;
      return S * V;
    }
    *dEr = dV;
    return V;
  } else {
    const double r = sqrt(r2);
    if (!(r < sig) || !(r2 > 0.0)) {
      *dEr = 0.0;
      return 0.0;
    }
    const double b = 1.0 - r / sig;
    const double um = alp == 2.0 ? b : pow(b, alp - 1.0);
    *dEr = -eps / sig * um / r;
    return eps / alp * um * b;
  }
}

// One Dense row per TPA lanes; float64 accumulation, no atomics.  KICK: the
// closing half kick p += half_dt F on float64 momenta.
template <int DIM, int KIND, int TPA, bool GENERAL, bool KICK>
__global__ __launch_bounds__(256) void k_pair_dense64(
    const double* __restrict__ pos, int n, int periodic, double L0, double L1, double L2,
    PairArgs a, const int* __restrict__ idx, int width, double* __restrict__ force,
    double* __restrict__ e_part, double* __restrict__ mom, double half_dt) {
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = gt / TPA;
  const int sub = gt % TPA;
  const bool active = i < n;
  const double L[3] = {L0, L1, L2};
  double f[3] = {0.0, 0.0, 0.0}, e = 0.0;
  if (active) {
    double xi[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < DIM; ++c) xi[c] = pos[(long long)i * DIM + c];
    float sig = a.sig, eps = a.eps, alp = a.alp;
    const int si = (GENERAL && a.species) ? a.species[i] : 0;
    const int* row = idx + (long long)i * width;
    for (int k = sub; k < width; k += TPA) {
      const unsigned j = (unsigned)row[k];
      if (j >= (unsigned)n) break;  // left-compacted rows, padding n
      double d[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        const double v = pos[(long long)j * DIM + c] - xi[c];
        d[c] = periodic ? wrap64(v, L[c]) : v;
      }
      const double r2 = DIM == 3 ? d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
                                 : d[0] * d[0] + d[1] * d[1];
      if (GENERAL) {
        const int sj = a.species ? a.species[j] : 0;
        sig = pick(a.sig_mode, a.sig, a.sig_v, i, j, si, sj, a.ns);
        eps = pick(a.eps_mode, a.eps, a.eps_v, i, j, si, sj, a.ns);
        if (KIND == JMD_SOFT_SPHERE) alp = pick(a.alp_mode, a.alp, a.alp_v, i, j, si, sj, a.ns);
      }
      double dEr;
      e += pair_eval64<KIND>(r2, (double)sig, (double)eps, (double)alp, a, &dEr);
#pragma unroll
      for (int c = 0; c < DIM; ++c) f[c] += dEr * d[c];
    }
  }
#pragma unroll
  for (int c = 0; c < DIM; ++c) f[c] = group_sum<TPA>(f[c]);
  e = group_sum<TPA>(e);
  if (active && sub == 0) {
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      const long long q = (long long)i * DIM + c;
      if (force) force[q] = f[c];
      if (KICK) mom[q] += half_dt * f[c];
    }
  }
  if (e_part) {
    __shared__ double s_e[256 / kWave];
    double v = (active && sub == 0) ? 0.5 * e : 0.0;
    v = wave_reduce_sum(v);
    if (lane_id() == 0) s_e[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int w = 0; w < (int)(blockDim.x / kWave); ++w) t += s_e[w];
      e_part[blockIdx.x] = t;
    }
  }
}

__global__ __launch_bounds__(256) void k_sum_partials64(const double* __restrict__ part, int nb,
                                                        double* __restrict__ out) {
  __shared__ double s[256];
  double v = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) v += part[b];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0];
}

// p += half_dt F, R = mod(R + dt p / m, L) in float64; float32 mirror of R
// (for the list) and the skin test against the list's reference positions.
template <int DIM, int STRIDE>
__global__ __launch_bounds__(256) void k_kick_drift64(
    double* __restrict__ x, double* __restrict__ p, const double* __restrict__ F,
    const double* __restrict__ mass, double mass_scalar, int n, Space s, double dt,
    double half_dt, float* __restrict__ x32, const float* __restrict__ xref, float thr_sq,
    int* __restrict__ gate) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool moved = false;
  if (i < n) {
    const double m = mass ? mass[i] : mass_scalar;
    float xn[3];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const long long k = (long long)i * DIM + d;
      const double pk = p[k] + half_dt * F[k];
      double xv = x[k] + dt * pk / m;
      if (s.periodic) xv = ref_mod64(xv, (double)s.L[d]);
      p[k] = pk;
      x[k] = xv;
      xn[d] = (float)xv;
    }
    store_pos<DIM, STRIDE>(x32, i, xn);
    if (gate && xref) {
      float r[3];
      load_pos<DIM>(xref, i, r);
      moved = ref_metric_sq<DIM>(xn, r, s) > thr_sq;
    }
  }
  if (gate) {
    const unsigned long long bal = __ballot(moved);
    if (bal != 0ull && lane_id() == 0) atomicOr(gate, 1);
  }
}

constexpr int kTPA64 = 8;

template <int DIM, int KIND>
static void launch_dense64(dim3 grid, hipStream_t st, const double* pos, int n, int periodic,
                           const double* L, const PairArgs& a, const int* idx, int width,
                           double* force, double* part, double* mom, double half_dt) {
#define JMD_D64(G, K)                                                                        \
  hipLaunchKernelGGL((k_pair_dense64<DIM, KIND, kTPA64, G, K>), grid, dim3(256), 0, st, pos, n, \
                     periodic, L[0], L[1], L[2], a, idx, width, force, part, mom, half_dt)
  if (a.general) {
    if (mom) JMD_D64(true, true);
    else JMD_D64(true, false);
  } else {
    if (mom) JMD_D64(false, true);
    else JMD_D64(false, false);
  }
#undef JMD_D64
}

}  // namespace jmd

using namespace jmd;

extern "C" {

size_t jmd_pair_workspace_bytes_f64(int64_t n) {
  if (n < 0) return 0;
  return (size_t)ceil_div(n * kTPA64, 256) * sizeof(double) + 256;
}

int jmd_pair_energy_force_f64(const double* position, int64_t n, int32_t dim, int32_t periodic,
                              const double* side, const jmd_pair_potential* potential,
                              const int32_t* idx, int64_t idx_width, int32_t format,
                              double* force, double* energy_total, double* momentum,
                              double half_dt, void* workspace, void* stream) {
  periodic &= ~JMD_POSITIONS_IN_BOX;  // the float64 fold (rint) takes any position
  if (periodic < 0 || periodic > 2) return JMD_ERR_INVALID;
  if (periodic == 2) return JMD_ERR_UNSUPPORTED;  // periodic_general: float32 path only
  if ((dim != 2 && dim != 3) || n < 0 || n > 0x7fffffff || idx_width < 0 ||
      idx_width > 0x7fffffff || (n > 0 && !position) || (idx_width > 0 && !idx) ||
      (periodic && !side))
    return JMD_ERR_INVALID;
  if (format != JMD_DENSE) return JMD_ERR_UNSUPPORTED;
  if (energy_total && !workspace) return JMD_ERR_INVALID;
  PairArgs a;
  if (!make_pair_args(potential, &a)) return JMD_ERR_INVALID;
  hipStream_t st = (hipStream_t)stream;
  if (n == 0) {
    if (energy_total) return launch_fill<double>(energy_total, 1, 0.0, st);
    return JMD_OK;
  }
  double L[3] = {1.0, 1.0, 1.0};
  for (int d = 0; d < dim; ++d) L[d] = periodic ? side[d] : 1.0;
  const int nn = (int)n;
  const int nb = ceil_div((long long)nn * kTPA64, 256);
  double* part = energy_total ? (double*)workspace : nullptr;
  if (dim == 3) {
    if (a.kind == JMD_LENNARD_JONES)
      launch_dense64<3, JMD_LENNARD_JONES>(dim3(nb), st, position, nn, periodic, L, a, idx,
                                           (int)idx_width, force, part, momentum, half_dt);
    else
      launch_dense64<3, JMD_SOFT_SPHERE>(dim3(nb), st, position, nn, periodic, L, a, idx,
                                         (int)idx_width, force, part, momentum, half_dt);
  } else {
    if (a.kind == JMD_LENNARD_JONES)
      launch_dense64<2, JMD_LENNARD_JONES>(dim3(nb), st, position, nn, periodic, L, a, idx,
                                           (int)idx_width, force, part, momentum, half_dt);
    else
      launch_dense64<2, JMD_SOFT_SPHERE>(dim3(nb), st, position, nn, periodic, L, a, idx,
                                         (int)idx_width, force, part, momentum, half_dt);
  }
  JMD_LAUNCH_CHECK();
  if (energy_total) {
    hipLaunchKernelGGL(k_sum_partials64, dim3(1), dim3(256), 0, st, part, nb, energy_total);
    JMD_LAUNCH_CHECK();
  }
  return JMD_OK;
}

int jmd_nve_kick_drift_f64(double* position, double* momentum, const double* force,
                           const double* mass, double mass_scalar, int64_t n, int32_t dim,
                           int32_t periodic, const float* side, double dt, double half_dt,
                           float* position_f32, int32_t position_stride,
                           const float* reference_position, float threshold_sq, int32_t* gate,
                           void* stream) {
  if ((dim != 2 && dim != 3) || n < 0 || n > 0x7fffffff || (periodic && !side) ||
      (n > 0 && (!position || !momentum || !force || !position_f32)))
    return JMD_ERR_INVALID;
  if (n == 0) return JMD_OK;
  if (!stride_ok(dim, position_stride, position_f32)) return JMD_ERR_INVALID;
  if (!mass && !(mass_scalar > 0.0)) return JMD_ERR_INVALID;
  {
    const int grc = space_check(dim, periodic, side, false);
    if (grc) return grc;
  }
  Space s = make_space(dim, periodic, side);
  hipStream_t st = (hipStream_t)stream;
  const int nn = (int)n;
  const float* xref = gate ? reference_position : nullptr;
  int* g = xref ? gate : nullptr;
  dim3 grid(ceil_div(nn, 256)), block(256);
#define JMD_KD64(D, S)                                                                      \
  hipLaunchKernelGGL((k_kick_drift64<D, S>), grid, block, 0, st, position, momentum, force, \
                     mass, mass_scalar, nn, s, dt, half_dt, position_f32, xref, threshold_sq, g)
  if (dim == 3 && position_stride == 4) JMD_KD64(3, 4);
  else if (dim == 3) JMD_KD64(3, 3);
  else JMD_KD64(2, 2);
#undef JMD_KD64
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

}  // extern "C"
