// Integrator kernels for gfx950: the velocity-Verlet NVE step of
// jax_md/simulate.py:227-317 and the FIRE descent bookkeeping of
// jax_md/minimize.py:124-226.
//
// Positions are accumulated in float64 (position_hi) and mirrored to the
// float32 table the neighbour list and pair kernels read.  The half kick and
// drift of a step run in one pass that also evaluates the skin criterion on
// the new positions; the closing half kick is fused into the pair kernel.
#include "jmd_common.hpp"

namespace jmd {

struct FireParamsDev {
  float dt_start, dt_max;
  int n_min;
  float f_inc, f_dec, alpha_start, f_alpha;
};

template <int DIM, int STRIDE>
__global__ __launch_bounds__(256) void k_kick_drift(KickDrift kd, int n) {
  float dt = kd.dt, half_dt = kd.half_dt;
  if (kd.dt_dev) {
    dt = kd.dt_dev[0];
    half_dt = dt / 2.0f;  // f32(dt / 2) of velocity_verlet (simulate.py:236)
  }
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool moved = false;
  if (i < n) {
    float xn[3];
    kick_drift_atom<DIM>(kd, i, dt, half_dt, xn);
    store_pos<DIM, STRIDE>(kd.x_out, i, xn);
    if (kd.gate_next) {
      float r[3];
      load_pos<DIM>(kd.ref, i, r);
      moved = ref_metric_sq<DIM>(xn, r, kd.s) > kd.thr_sq;
    }
  }
  if (kd.gate_next) {
    const unsigned long long bal = __ballot(moved);
    if (bal != 0ull && lane_id() == 0) atomicOr(kd.gate_next, 1);
  }
}

// FIRE: per-block float64 partials of sum F^2, sum P^2, sum F.P.
template <int DIM>
__global__ __launch_bounds__(256) void k_fire_partials(const float* __restrict__ p,
                                                       const float* __restrict__ F, int n,
                                                       double* __restrict__ part) {
  double ff = 0.0, pp = 0.0, fp = 0.0;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < (long long)n * DIM;
       k += (long long)gridDim.x * blockDim.x) {
    double f = F[k], q = p[k];
    ff += f * f;
    pp += q * q;
    fp += f * q;
  }
  __shared__ double s[3][256 / kWave];
  ff = wave_reduce_sum(ff);
  pp = wave_reduce_sum(pp);
  fp = wave_reduce_sum(fp);
  if (lane_id() == 0) {
    s[0][threadIdx.x / kWave] = ff;
    s[1][threadIdx.x / kWave] = pp;
    s[2][threadIdx.x / kWave] = fp;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w) t += s[threadIdx.x][w];
    part[(long long)blockIdx.x * 3 + threadIdx.x] = t;
  }
}

__device__ __forceinline__ float to_f16(float x) { return (float)(_Float16)x; }

// One thread: reduce partials, then the FIRE scalar update in the order of
// minimize.py:190-219 with float16 dt/alpha (util.static_cast).
// scalars: [0] sum F^2, [1] sum P^2, [2] F.P, [3] alpha used for the mix,
// [4] |P| / |F|, [5] keep (F.P >= 0).
__global__ void k_fire_scalars(const double* __restrict__ part, int nb,
                               float* __restrict__ st, int* __restrict__ npos,
                               FireParamsDev prm, double* __restrict__ scalars) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double ff = 0.0, pp = 0.0, fp = 0.0;
  for (int b = 0; b < nb; ++b) {
    ff += part[3 * b];
    pp += part[3 * b + 1];
    fp += part[3 * b + 2];
  }
  double F_norm = sqrt(ff + 1e-6);
  double P_norm = sqrt(pp);
  float dt = st[0], alpha = st[1];
  scalars[0] = ff;
  scalars[1] = pp;
  scalars[2] = fp;
  scalars[3] = (double)alpha;
  scalars[4] = P_norm / F_norm;
  scalars[5] = fp >= 0.0 ? 1.0 : 0.0;
  int n_pos = fp >= 0.0 ? npos[0] + 1 : 0;
  if (fp > 0.0 && n_pos > prm.n_min) dt = fminf(to_f16(dt * prm.f_inc), prm.dt_max);
  if (fp < 0.0) dt = to_f16(dt * prm.f_dec);
  if (fp > 0.0 && n_pos > prm.n_min) alpha = to_f16(alpha * prm.f_alpha);
  if (fp < 0.0) alpha = prm.alpha_start;
  st[0] = dt;
  st[1] = alpha;
  npos[0] = n_pos;
}

template <int DIM>
__global__ __launch_bounds__(256) void k_fire_mix(float* __restrict__ p,
                                                  const float* __restrict__ F, int n,
                                                  const double* __restrict__ scalars) {
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long long)n * DIM) return;
  double alpha = scalars[3], ratio = scalars[4], keep = scalars[5];
  double q = p[k];
  q = q + alpha * ((double)F[k] * ratio - q);
  p[k] = (float)(q * keep);
}

}  // namespace jmd

using namespace jmd;

extern "C" {

int jmd_nve_kick_drift(const double* position_hi, const float* momentum, const float* force,
                       const float* mass, float mass_scalar, int64_t n, int32_t dim,
                       int32_t periodic, const float* side, float dt, float half_dt,
                       const float* dt_device, double* position_hi_out, float* position_out,
                       int32_t position_stride, float* momentum_out,
                       const float* reference_position, float threshold_sq, int32_t* gate,
                       void* stream) {
  if ((dim != 2 && dim != 3) || n < 0 || n > 0x7fffffff || (periodic && !side) ||
      (n > 0 && (!position_hi || !momentum || !force || !position_hi_out || !position_out ||
                 !momentum_out)))
    return JMD_ERR_INVALID;
  if (n == 0) return JMD_OK;
  if (!stride_ok(dim, position_stride, position_out)) return JMD_ERR_INVALID;
  if (!mass && !(mass_scalar > 0.0f)) return JMD_ERR_INVALID;
  Space s = make_space(dim, periodic, side);
  hipStream_t st = (hipStream_t)stream;
  int nn = (int)n;
  KickDrift kd{};
  kd.active = 1;
  kd.xhi_in = position_hi;
  kd.xhi_out = position_hi_out;
  kd.p_in = momentum;
  kd.p_out = momentum_out;
  kd.F = force;
  kd.mass = mass;
  kd.mass_scalar = mass_scalar;
  kick_drift_mass(&kd);
  kd.dt = dt;
  kd.half_dt = half_dt;
  kd.dt_dev = dt_device;
  kd.x_out = position_out;
  kd.ref = gate ? reference_position : nullptr;
  kd.thr_sq = threshold_sq;
  kd.gate_next = kd.ref ? gate : nullptr;
  kd.s = s;
  dim3 grid(ceil_div(nn, 256)), block(256);
#define JMD_KD(D, S) hipLaunchKernelGGL((k_kick_drift<D, S>), grid, block, 0, st, kd, nn)
  if (dim == 3 && position_stride == 4) JMD_KD(3, 4);
  else if (dim == 3) JMD_KD(3, 3);
  else JMD_KD(2, 2);
#undef JMD_KD
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

static const int kFireBlocks = 1024;

size_t jmd_fire_workspace_bytes(int64_t n) {
  (void)n;
  return (size_t)kFireBlocks * 3 * sizeof(double) + 256;
}

int jmd_fire_update(float* momentum, const float* force, int64_t n, int32_t dim,
                    float* fire_state, int32_t* fire_n_pos, const float* params,
                    double* scalars, void* workspace, void* stream) {
  if ((dim != 2 && dim != 3) || n < 0 || n > 0x7fffffff || !fire_state || !fire_n_pos ||
      !params || !scalars || !workspace || (n > 0 && (!momentum || !force)))
    return JMD_ERR_INVALID;
  hipStream_t st = (hipStream_t)stream;
  int nn = (int)n;
  long long m = (long long)nn * dim;
  int nb = ceil_div(m, 256);
  if (nb > kFireBlocks) nb = kFireBlocks;
  if (nb < 1) nb = 1;
  double* part = (double*)workspace;
  FireParamsDev prm;
  prm.dt_start = params[0];
  prm.dt_max = params[1];
  prm.n_min = (int)params[2];
  prm.f_inc = params[3];
  prm.f_dec = params[4];
  prm.alpha_start = params[5];
  prm.f_alpha = params[6];
  if (dim == 3)
    hipLaunchKernelGGL(k_fire_partials<3>, dim3(nb), dim3(256), 0, st, momentum, force, nn, part);
  else
    hipLaunchKernelGGL(k_fire_partials<2>, dim3(nb), dim3(256), 0, st, momentum, force, nn, part);
  JMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_fire_scalars, dim3(1), dim3(64), 0, st, part, nb, fire_state, fire_n_pos,
                     prm, scalars);
  JMD_LAUNCH_CHECK();
  if (m > 0) {
    if (dim == 3)
      hipLaunchKernelGGL(k_fire_mix<3>, dim3(ceil_div(m, 256)), dim3(256), 0, st, momentum, force,
                         nn, scalars);
    else
      hipLaunchKernelGGL(k_fire_mix<2>, dim3(ceil_div(m, 256)), dim3(256), 0, st, momentum, force,
                         nn, scalars);
    JMD_LAUNCH_CHECK();
  }
  return JMD_OK;
}

}  // extern "C"
