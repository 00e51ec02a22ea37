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

// ---------------------------------------------------------------------------
// Nose-Hoover chain thermostat (simulate.py:328-347,384-672): the chain is a
// few float64 scalars, updated by one thread; it hands the particles one
// momentum factor per half step.  Chain layout: chain[0..L) positions xi,
// chain[L..2L) momenta p_xi, chain[2L..3L) masses Q.
// ---------------------------------------------------------------------------

__device__ const double kSYWeights[4][7] = {
    {1.0, 0, 0, 0, 0, 0, 0},
    {0.828981543588751, -0.657963087177502, 0.828981543588751, 0, 0, 0, 0},
    {0.2967324292201065, 0.2967324292201065, -0.186929716880426, 0.2967324292201065,
     0.2967324292201065, 0, 0},
    {0.784513610477560, 0.235573213359357, -1.17767998417887, 1.31518632068391,
     -1.17767998417887, 0.235573213359357, 0.784513610477560}};

constexpr int kMaxChain = 16;

// exp(x) in float64: the chain's arguments are small (|x| << 0.05), where
// eight Taylor terms are exact to rounding (x^9/9! < 2e-17 relative) and
// several times cheaper than the library call on one thread.
__device__ __forceinline__ double chain_exp(double x) {
  if (fabs(x) > 0.03125) return exp(x);
  double t = fma(x, 1.0 / 9.0, 1.0);
  t = fma(x * (1.0 / 8.0), t, 1.0);
  t = fma(x * (1.0 / 7.0), t, 1.0);
  t = fma(x * (1.0 / 6.0), t, 1.0);
  t = fma(x * (1.0 / 5.0), t, 1.0);
  t = fma(x * (1.0 / 4.0), t, 1.0);
  t = fma(x * (1.0 / 3.0), t, 1.0);
  t = fma(x * 0.5, t, 1.0);
  return fma(x, t, 1.0);
}

// One substep of length delta (Martyna et al.: backward sweep of the chain
// momenta, rescaling of the particle momenta, chain positions, forward
// sweep); returns the particle momentum factor and updates ke.  The chain
// length is a template parameter so the chain stays in registers.
template <int L>
__device__ __forceinline__ double nhc_substep(double delta, double (&xi)[L], double (&p)[L],
                                              const double (&iQ)[L], double* ke, double dof,
                                              double kT) {
  const double d2 = delta / 2.0, d4 = d2 / 2.0, d8 = d4 / 2.0;
  constexpr int M = L - 1;
  double old[L];
#pragma unroll
  for (int m = 0; m < L; ++m) old[m] = p[m];
  p[M] += d4 * (old[M - 1] * old[M - 1] * iQ[M - 1] - kT);
  double carry = p[M];
#pragma unroll
  for (int m = M - 1; m >= 1; --m) {
    const double G = old[m - 1] * old[m - 1] * iQ[m - 1] - kT;
    const double sc = chain_exp(-d8 * carry * iQ[m + 1]);
    carry = sc * (sc * old[m] + d4 * G);
    p[m] = carry;
  }
  double sc = chain_exp(-d8 * p[1] * iQ[1]);
  p[0] = sc * (sc * p[0] + d4 * (2.0 * *ke - dof * kT));
  const double scale = chain_exp(-d2 * p[0] * iQ[0]);
  *ke = *ke * scale * scale;
#pragma unroll
  for (int m = 0; m < L; ++m) xi[m] += d2 * p[m] * iQ[m];
  double G = 2.0 * *ke - dof * kT;
  double q[L];
#pragma unroll
  for (int m = 0; m < L; ++m) q[m] = p[m];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    sc = chain_exp(-d8 * p[m + 1] * iQ[m + 1]);
    q[m] = sc * (sc * p[m] + d4 * G);
    G = q[m] * q[m] * iQ[m] - kT;
  }
  q[M] += d4 * G;
#pragma unroll
  for (int m = 0; m < L; ++m) p[m] = q[m];
  return scale;
}

template <int L>
__device__ double nhc_half_step(double (&xi)[L], double (&p)[L], const double (&iQ)[L],
                                int chain_steps, int sy_steps, float dt, double* ke, double dof,
                                double kT) {
  if (chain_steps == 1 && sy_steps == 1) return nhc_substep<L>((double)dt, xi, p, iQ, ke, dof, kT);
  double total = 1.0;
  const float delta = dt / (float)chain_steps;
  const int row = sy_steps == 1 ? 0 : sy_steps == 3 ? 1 : sy_steps == 5 ? 2 : 3;
  for (int i = 0; i < chain_steps * sy_steps; ++i) {
    const float d = (float)((double)delta * kSYWeights[row][i % sy_steps]);
    total *= nhc_substep<L>((double)d, xi, p, iQ, ke, dof, kT);
  }
  return total;
}

enum { NHC_FINISH = 1, NHC_START = 2, NHC_INIT = 4 };

// mode NHC_INIT: ke = sum of the partials, chain at rest.  NHC_FINISH: ke =
// sum of the partials (after the closing kick), half step.  NHC_START:
// chain masses for kT, half step.  scale = the product of the factors.
// chain = [xi (L), p_xi (L), Q (L)] in float64.  One thread; ksum is the
// partials' sum (used by NHC_FINISH and NHC_INIT).
template <int L>
__device__ void nhc_chain(int mode, double* __restrict__ chain, int chain_steps, int sy_steps,
                          float dt, float kT, float tau, double dof, double* __restrict__ ke,
                          double ksum, double* __restrict__ scale) {
  const double kt = (double)kT;
  double k = (mode & (NHC_FINISH | NHC_INIT)) ? ksum : ke[0];
  double xi[L], p[L], Q[L], iQ[L];
#pragma unroll
  for (int m = 0; m < L; ++m) {
    const bool init = (mode & NHC_INIT) != 0;
    xi[m] = init ? 0.0 : chain[m];
    p[m] = init ? 0.0 : chain[L + m];
    Q[m] = chain[2 * L + m];
  }
  if (mode & (NHC_INIT | NHC_START)) {
    const double t = (double)tau;
#pragma unroll
    for (int m = 0; m < L; ++m) Q[m] = (double)kT * t * t * (m == 0 ? dof : 1.0);
  }
#pragma unroll
  for (int m = 0; m < L; ++m) iQ[m] = 1.0 / Q[m];
  double total = 1.0;
  if (mode & NHC_FINISH) total *= nhc_half_step<L>(xi, p, iQ, chain_steps, sy_steps, dt, &k, dof, kt);
  if (mode & NHC_START) total *= nhc_half_step<L>(xi, p, iQ, chain_steps, sy_steps, dt, &k, dof, kt);
#pragma unroll
  for (int m = 0; m < L; ++m) {
    chain[m] = xi[m];
    chain[L + m] = p[m];
    chain[2 * L + m] = Q[m];
  }
  ke[0] = k;
  scale[0] = total;
}

// Sum of the partials in a fixed order: thread t adds entries t, t + 256,
// ..., then a tree over the 256 sums (block of 256 threads).  LOAD reads one
// partial.
template <typename LOAD>
__device__ __forceinline__ double partials_sum(int nparts, LOAD load) {
  __shared__ double s[256];
  double v = 0.0;
  for (int b = threadIdx.x; b < nparts; b += 256) v += load(b);
  s[threadIdx.x] = v;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  return s[0];
}

template <int L>
__global__ __launch_bounds__(256) void k_nhc(int mode, double* __restrict__ chain,
                                             int chain_steps, int sy_steps, float dt, float kT,
                                             float tau, double dof, double* __restrict__ ke,
                                             const double* __restrict__ part, int nparts,
                                             double* __restrict__ scale) {
  const bool sum = (mode & (NHC_FINISH | NHC_INIT)) != 0;
  const double ksum = partials_sum(sum ? nparts : 0, [&](int b) { return part[b]; });
  if (threadIdx.x != 0) return;
  nhc_chain<L>(mode, chain, chain_steps, sy_steps, dt, kT, tau, dof, ke, ksum, scale);
}

// p += half_dt F (optional), and per-block float64 partials of sum p^2 / m.
template <int DIM>
__global__ __launch_bounds__(256) void k_kick_ke(float* __restrict__ p, const float* __restrict__ F,
                                                 const float* __restrict__ mass, float mass_scalar,
                                                 int n, float half_dt, double* __restrict__ part) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double k = 0.0;
  if (i < n) {
    const double m = mass ? (double)mass[i] : (double)mass_scalar;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const long long q = (long long)i * DIM + d;
      float v = p[q];
      if (F) {
        v += half_dt * F[q];
        p[q] = v;
      }
      k += (double)v * (double)v;
    }
    k /= m;
  }
  __shared__ double s_k[256 / kWave];
  k = wave_reduce_sum(k);
  if (lane_id() == 0) s_k[threadIdx.x / kWave] = k;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w) t += s_k[w];
    part[blockIdx.x] = 0.5 * t;
  }
}

// k_kick_ke followed by k_nhc in one launch: each workgroup exchanges its
// partial into part[] with a device-scope atomic (performed at the device's
// coherence point) and takes a completion ticket only after that exchange has
// returned; the workgroup that takes the last ticket sums the partials with
// coherent loads in k_nhc's order, advances the chain and resets the ticket.
// The results equal those of the two launches bit for bit.
template <int DIM, int L>
__global__ __launch_bounds__(256) void k_kick_ke_nhc(
    float* __restrict__ p, const float* __restrict__ F, const float* __restrict__ mass,
    float mass_scalar, int n, float half_dt, double* __restrict__ part,
    unsigned* __restrict__ ticket, int mode, double* __restrict__ chain, int chain_steps,
    int sy_steps, float dt, float kT, float tau, double dof, double* __restrict__ ke,
    double* __restrict__ scale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double k = 0.0;
  if (i < n) {
    const double m = mass ? (double)mass[i] : (double)mass_scalar;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const long long q = (long long)i * DIM + d;
      float v = p[q];
      if (F) {
        v += half_dt * F[q];
        p[q] = v;
      }
      k += (double)v * (double)v;
    }
    k /= m;
  }
  __shared__ double s_k[256 / kWave];
  __shared__ int s_last;
  k = wave_reduce_sum(k);
  if (lane_id() == 0) s_k[threadIdx.x / kWave] = k;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w) t += s_k[w];
    unsigned long long* slot = reinterpret_cast<unsigned long long*>(part + blockIdx.x);
    const unsigned long long old =
        __hip_atomic_exchange(slot, (unsigned long long)__double_as_longlong(0.5 * t),
                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // the ticket is issued only once the exchange has returned
    asm volatile("" ::"v"(old) : "memory");
    const unsigned tk =
        __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_last = tk == gridDim.x - 1 ? 1 : 0;
  }
  __syncthreads();
  if (!s_last) return;
  const int nparts = (int)gridDim.x;
  const bool sum = (mode & (NHC_FINISH | NHC_INIT)) != 0;
  const double ksum = partials_sum(sum ? nparts : 0, [&](int b) {
    const unsigned long long bits = __hip_atomic_load(
        reinterpret_cast<const unsigned long long*>(part + b), __ATOMIC_RELAXED,
        __HIP_MEMORY_SCOPE_AGENT);
    return __longlong_as_double((long long)bits);
  });
  if (threadIdx.x != 0) return;
  __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  nhc_chain<L>(mode, chain, chain_steps, sy_steps, dt, kT, tau, dof, ke, ksum, scale);
}

__global__ __launch_bounds__(256) void k_scale_momentum(float* __restrict__ p, long long m,
                                                        const double* __restrict__ scale) {
  const float sc = (float)scale[0];
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < m;
       k += (long long)gridDim.x * blockDim.x)
    p[k] = p[k] * sc;
}

}  // namespace jmd

using namespace jmd;

extern "C" {

static int kick_drift_impl(const double* position_hi, const float* momentum,
                           const float* force, const float* mass, float mass_scalar, int64_t n,
                           int32_t dim, int32_t periodic, const float* side, float dt,
                           float half_dt, const float* dt_device, double* position_hi_out,
                           float* position_out, int32_t position_stride, float* momentum_out,
                           const float* reference_position, float threshold_sq, int32_t* gate,
                           const double* momentum_scale, void* stream);

int jmd_nve_kick_drift(const double* position_hi, const float* momentum, const float* force,
                       const float* mass, float mass_scalar, int64_t n, int32_t dim,
                       int32_t periodic, const float* side, float dt, float half_dt,
                       const float* dt_device, double* position_hi_out, float* position_out,
                       int32_t position_stride, float* momentum_out,
                       const float* reference_position, float threshold_sq, int32_t* gate,
                       void* stream) {
  return kick_drift_impl(position_hi, momentum, force, mass, mass_scalar, n, dim, periodic,
                         side, dt, half_dt, dt_device, position_hi_out, position_out,
                         position_stride, momentum_out, reference_position, threshold_sq, gate,
                         nullptr, stream);
}

int jmd_nvt_kick_drift(const double* position_hi, const float* momentum, const float* force,
                       const float* mass, float mass_scalar, int64_t n, int32_t dim,
                       int32_t periodic, const float* side, float dt, float half_dt,
                       double* position_hi_out, float* position_out, int32_t position_stride,
                       float* momentum_out, const float* reference_position, float threshold_sq,
                       int32_t* gate, const double* momentum_scale, void* stream) {
  if (!momentum_scale) return JMD_ERR_INVALID;
  return kick_drift_impl(position_hi, momentum, force, mass, mass_scalar, n, dim, periodic,
                         side, dt, half_dt, nullptr, position_hi_out, position_out,
                         position_stride, momentum_out, reference_position, threshold_sq, gate,
                         momentum_scale, stream);
}

static int kick_drift_impl(const double* position_hi, const float* momentum,
                           const float* force, const float* mass, float mass_scalar, int64_t n,
                           int32_t dim, int32_t periodic, const float* side, float dt,
                           float half_dt, const float* dt_device, double* position_hi_out,
                           float* position_out, int32_t position_stride, float* momentum_out,
                           const float* reference_position, float threshold_sq, int32_t* gate,
                           const double* momentum_scale, void* stream) {
  if ((dim != 2 && dim != 3) || n < 0 || n > 0x7fffffff || (periodic && !side) ||
      (n > 0 && (!position_hi || !momentum || !force || !position_hi_out || !position_out ||
                 !momentum_out)))
    return JMD_ERR_INVALID;
  if (n == 0) return JMD_OK;
  if (!stride_ok(dim, position_stride, position_out)) return JMD_ERR_INVALID;
  if (!mass && !(mass_scalar > 0.0f)) return JMD_ERR_INVALID;
  {
    const int grc = space_check(dim, periodic, side, true);
    if (grc) return grc;
  }
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
  box_inverse(s, kd.Ti);
  kd.p_scale = momentum_scale;
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

int jmd_kick_kinetic_partials(int64_t n) { return n <= 0 ? 1 : ceil_div(n, 256); }

int jmd_kick_kinetic(float* momentum, const float* force, const float* mass, float mass_scalar,
                     int64_t n, int32_t dim, float half_dt, double* ke_partials, void* stream) {
  if ((dim != 2 && dim != 3) || n < 0 || n > 0x7fffffff || !ke_partials ||
      (n > 0 && !momentum) || (!mass && !(mass_scalar > 0.0f)))
    return JMD_ERR_INVALID;
  hipStream_t st = (hipStream_t)stream;
  const int nn = (int)n;
  dim3 grid(jmd_kick_kinetic_partials(n)), block(256);
  if (dim == 3)
    hipLaunchKernelGGL(k_kick_ke<3>, grid, block, 0, st, momentum, force, mass, mass_scalar, nn,
                       half_dt, ke_partials);
  else
    hipLaunchKernelGGL(k_kick_ke<2>, grid, block, 0, st, momentum, force, mass, mass_scalar, nn,
                       half_dt, ke_partials);
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

int jmd_nhc_step(int32_t mode, double* chain, int32_t chain_length, int32_t chain_steps,
                 int32_t sy_steps, float dt, float kT, float tau, double dof, double* kinetic,
                 const double* ke_partials, int32_t n_partials, double* momentum_scale,
                 void* stream) {
  if (!chain || !kinetic || !momentum_scale || chain_length < 2 || chain_length > kMaxChain ||
      chain_steps < 1 || (sy_steps != 1 && sy_steps != 3 && sy_steps != 5 && sy_steps != 7) ||
      mode < 1 || mode > 7 || ((mode & (NHC_FINISH | NHC_INIT)) && (!ke_partials || n_partials < 1)))
    return JMD_ERR_INVALID;
  hipStream_t st = (hipStream_t)stream;
#define JMD_NHC(LEN)                                                                         \
  case LEN:                                                                                  \
    hipLaunchKernelGGL(k_nhc<LEN>, dim3(1), dim3(256), 0, st, (int)mode, chain,              \
                       (int)chain_steps, (int)sy_steps, dt, kT, tau, dof, kinetic, ke_partials, \
                       (int)n_partials, momentum_scale);                                     \
    break;
  switch (chain_length) {
    JMD_NHC(2) JMD_NHC(3) JMD_NHC(4) JMD_NHC(5) JMD_NHC(6) JMD_NHC(7) JMD_NHC(8) JMD_NHC(9)
    JMD_NHC(10) JMD_NHC(11) JMD_NHC(12) JMD_NHC(13) JMD_NHC(14) JMD_NHC(15) JMD_NHC(16)
    default:
      return JMD_ERR_INVALID;
  }
#undef JMD_NHC
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

int jmd_kick_kinetic_nhc(float* momentum, const float* force, const float* mass,
                         float mass_scalar, int64_t n, int32_t dim, float half_dt,
                         double* ke_partials, uint32_t* ticket, int32_t mode, double* chain,
                         int32_t chain_length, int32_t chain_steps, int32_t sy_steps, float dt,
                         float kT, float tau, double dof, double* kinetic,
                         double* momentum_scale, void* stream) {
  if ((dim != 2 && dim != 3) || n < 0 || n > 0x7fffffff || !ke_partials || !ticket ||
      (n > 0 && !momentum) || (!mass && !(mass_scalar > 0.0f)) || !chain || !kinetic ||
      !momentum_scale || chain_length < 2 || chain_length > kMaxChain || chain_steps < 1 ||
      (sy_steps != 1 && sy_steps != 3 && sy_steps != 5 && sy_steps != 7) || mode < 1 || mode > 7)
    return JMD_ERR_INVALID;
  hipStream_t st = (hipStream_t)stream;
  const int nn = (int)n;
  dim3 grid(jmd_kick_kinetic_partials(n)), block(256);
#define JMD_KNHC(LEN)                                                                        \
  case LEN:                                                                                  \
    if (dim == 3)                                                                            \
      hipLaunchKernelGGL((k_kick_ke_nhc<3, LEN>), grid, block, 0, st, momentum, force, mass, \
                         mass_scalar, nn, half_dt, ke_partials, ticket, (int)mode, chain,    \
                         (int)chain_steps, (int)sy_steps, dt, kT, tau, dof, kinetic,         \
                         momentum_scale);                                                    \
    else                                                                                     \
      hipLaunchKernelGGL((k_kick_ke_nhc<2, LEN>), grid, block, 0, st, momentum, force, mass, \
                         mass_scalar, nn, half_dt, ke_partials, ticket, (int)mode, chain,    \
                         (int)chain_steps, (int)sy_steps, dt, kT, tau, dof, kinetic,         \
                         momentum_scale);                                                    \
    break;
  switch (chain_length) {
    JMD_KNHC(2) JMD_KNHC(3) JMD_KNHC(4) JMD_KNHC(5) JMD_KNHC(6) JMD_KNHC(7) JMD_KNHC(8)
    JMD_KNHC(9) JMD_KNHC(10) JMD_KNHC(11) JMD_KNHC(12) JMD_KNHC(13) JMD_KNHC(14) JMD_KNHC(15)
    JMD_KNHC(16)
    default:
      return JMD_ERR_INVALID;
  }
#undef JMD_KNHC
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

int jmd_scale_momentum(float* momentum, int64_t n, int32_t dim, const double* scale,
                       void* stream) {
  if ((dim != 2 && dim != 3) || n < 0 || !scale || (n > 0 && !momentum)) return JMD_ERR_INVALID;
  const long long m = (long long)n * dim;
  long long nb = (m + 255) / 256;
  if (nb > 2048) nb = 2048;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(k_scale_momentum, dim3((unsigned)nb), dim3(256), 0, (hipStream_t)stream,
                     momentum, m, scale);
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

}  // extern "C"
