// Shared device helpers for the gfx950 neighbour-list MD kernels.
//
// Geometry follows jax_md/space.py: the periodic displacement is
// mod(d + L/2, L) - L/2 with jnp.mod's floor semantics (space.py:213-224),
// the square distance is the left-to-right sum of squares (space.py:227-235).
// The neighbour-list build must reproduce those float32 roundings exactly,
// so the "ref_" helpers below forbid multiply-add contraction.  The force
// kernels use a minimum-image difference that avoids the rounding of the
// wrap (see min_image_disp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#include "../../include/jaxmd_native.h"

#define JMD_RETURN_IF_HIP(expr)                          \
  do {                                                   \
    hipError_t _e = (expr);                              \
    if (_e != hipSuccess) return JMD_ERR_LAUNCH;         \
  } while (0)

#define JMD_LAUNCH_CHECK() JMD_RETURN_IF_HIP(hipGetLastError())

namespace jmd {

constexpr int kWave = 64;

// Geometry passed by value to kernels.  periodic_general with fractional
// coordinates (space.py:332-472) is the unit cube (L = 1) plus the box
// matrix T: displacements are T wrap(u_a - u_b), shifts mod(u + T^-1 dR, 1).
struct Space {
  int dim;
  int periodic;
  float L[3];
  float half[3];  // float32(L) * 0.5, exact
  int general = 0;  // 1: positions are fractional, d = T d_unit
  int upper = 1;    // T is upper triangular (the zero products are skipped)
  float T[3][3];  // box matrix, row-major (general only)
  float fw[3];    // |row d of T^-1|: fractional reach per unit of real distance
  float smin2;    // smallest squared singular value of T: |T v|^2 >= smin2 |v|^2
  int inbox = 0;  // every position lies in [0, L] (JMD_POSITIONS_IN_BOX)
  // (kept small: it is passed by value, in registers, to the exact pair
  // test; the inverse lives with the integrator's parameters)
};


// T^-1 of a general box (float64 cofactors, rounded once); identity otherwise.
inline void box_inverse(const Space& s, float Ti[3][3]) {
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) Ti[a][b] = a == b ? 1.0f : 0.0f;
  if (!s.general) return;
  double m[3][3];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) m[a][b] = (double)s.T[a][b];
  const double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                     m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                     m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  const double inv[3][3] = {
      {(m[1][1] * m[2][2] - m[1][2] * m[2][1]) / det, (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / det,
       (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / det},
      {(m[1][2] * m[2][0] - m[1][0] * m[2][2]) / det, (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / det,
       (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / det},
      {(m[1][0] * m[2][1] - m[1][1] * m[2][0]) / det, (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / det,
       (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / det}};
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) Ti[a][b] = (float)inv[a][b];
}

// periodic: 0 free, 1 periodic box of sides side[0..dim), 2 periodic_general
// with fractional coordinates, side = the [dim, dim] box matrix (row-major);
// JMD_POSITIONS_IN_BOX may be ORed in.
inline Space make_space(int dim, int periodic_arg, const float* side) {
  const int periodic = periodic_arg & ~JMD_POSITIONS_IN_BOX;
  Space s;
  s.dim = dim;
  s.periodic = periodic != 0;
  s.general = periodic == 2;
  s.inbox = periodic != 0 && (periodic_arg & JMD_POSITIONS_IN_BOX) != 0;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) s.T[a][b] = a == b ? 1.0f : 0.0f;
  for (int d = 0; d < 3; ++d) {
    float L = (!s.general && side && d < dim) ? side[d] : 1.0f;
    s.L[d] = L;
    s.half[d] = L * 0.5f;
  }
  if (s.general && side)
    for (int a = 0; a < dim; ++a)
      for (int b = 0; b < dim; ++b) s.T[a][b] = side[a * dim + b];
  s.upper = s.T[1][0] == 0.0f && s.T[2][0] == 0.0f && s.T[2][1] == 0.0f;
  float Ti[3][3];
  box_inverse(s, Ti);
  for (int a = 0; a < 3; ++a)
    s.fw[a] = sqrtf(Ti[a][0] * Ti[a][0] + Ti[a][1] * Ti[a][1] + Ti[a][2] * Ti[a][2]);
  // smallest eigenvalue of T^T T over the dim x dim block (cyclic Jacobi),
  // lowered by 1e-5 relative so that the pruning bound stays conservative
  {
    double M[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = 0; a < dim; ++a)
      for (int b = 0; b < dim; ++b)
        for (int k = 0; k < dim; ++k) M[a][b] += (double)s.T[k][a] * (double)s.T[k][b];
    for (int sweep = 0; sweep < 30; ++sweep)
      for (int p = 0; p < dim; ++p)
        for (int q = p + 1; q < dim; ++q) {
          if (M[p][q] == 0.0) continue;
          const double th = 0.5 * atan2(2.0 * M[p][q], M[q][q] - M[p][p]);
          const double c = cos(th), sn = sin(th);
          for (int k = 0; k < dim; ++k) {  // columns p, q
            const double mp = M[k][p], mq = M[k][q];
            M[k][p] = c * mp - sn * mq;
            M[k][q] = sn * mp + c * mq;
          }
          for (int k = 0; k < dim; ++k) {  // rows p, q
            const double mp = M[p][k], mq = M[q][k];
            M[p][k] = c * mp - sn * mq;
            M[q][k] = sn * mp + c * mq;
          }
        }
    double lo = M[0][0];
    for (int a = 1; a < dim; ++a) lo = lo < M[a][a] ? lo : M[a][a];
    s.smin2 = s.general ? (float)(lo * (1.0 - 1e-5)) : 1.0f;
  }
  return s;
}

// True when a periodic_general box matrix is usable: finite and invertible.
inline bool box_matrix_ok(int dim, const float* T) {
  if (!T) return false;
  double m[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int a = 0; a < dim; ++a)
    for (int b = 0; b < dim; ++b) {
      const float v = T[a * dim + b];
      if (!(v == v) || v > 3.0e38f || v < -3.0e38f) return false;
      m[a][b] = v;
    }
  const double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                     m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                     m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  return det != 0.0;
}

// Geometry arguments of an entry point: periodic 0 / 1 (side[dim] when 1), or
// 2 (periodic_general, side = an invertible [dim, dim] box matrix) where the
// entry point supports it.
inline int space_check(int dim, int periodic_arg, const float* side, bool allow_general) {
  const int periodic = periodic_arg & ~JMD_POSITIONS_IN_BOX;
  if (periodic < 0 || periodic > 2) return JMD_ERR_INVALID;
  if (periodic && !side) return JMD_ERR_INVALID;
  if (periodic == 2) {
    if (!allow_general) return JMD_ERR_UNSUPPORTED;
    if (!box_matrix_ok(dim, side)) return JMD_ERR_INVALID;
  }
  return JMD_OK;
}

// d <- T d for the force kernels (tolerance parity: contraction allowed).
template <int DIM>
__device__ __forceinline__ void box_apply(const Space& s, float* d) {
  float r[3];
  if (s.upper) {  // row a has nonzero entries from column a on
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      float acc = s.T[a][a] * d[a];
#pragma unroll
      for (int b = a + 1; b < DIM; ++b) acc = fmaf(s.T[a][b], d[b], acc);
      r[a] = acc;
    }
  } else {
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      float acc = s.T[a][0] * d[0];
#pragma unroll
      for (int b = 1; b < DIM; ++b) acc = fmaf(s.T[a][b], d[b], acc);
      r[a] = acc;
    }
  }
#pragma unroll
  for (int a = 0; a < DIM; ++a) d[a] = r[a];
}

// d <- T d as the reference's transform (einsum 'ij,j->i', space.py:128-153):
// products rounded, summed left to right.
template <int DIM>
__device__ __forceinline__ void box_apply_ref(const Space& s, float* d) {
#pragma clang fp contract(off)
  float r[3];
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    float acc = s.T[a][0] * d[0];
#pragma unroll
    for (int b = 1; b < DIM; ++b) acc = acc + s.T[a][b] * d[b];
    r[a] = acc;
  }
#pragma unroll
  for (int a = 0; a < DIM; ++a) d[a] = r[a];
}

// Cell grid (jax_md/partition.py:146-188,211-223).
struct Grid {
  int cps[3];          // cells per side, x fastest in the hash
  float cell_size[3];  // float32 box / cells_per_side
  int ncell;
  int cap;             // reference cell capacity (slot rotation, overflow)
};

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

// jnp.mod(t, L) - half for float32, bit-exact with XLA's rem + sign fix.
// fmod(t, L) is exact; the common ranges are handled without the libcall.
__device__ __forceinline__ float ref_wrap1(float d, float L, float half) {
#pragma clang fp contract(off)
  float t = d + half;
  float r;
  if (t >= 0.0f && t < L) {
    r = t;
  } else if (t >= L && t < 2.0f * L) {
    r = t - L;  // exact (Sterbenz)
  } else if (t < 0.0f && t > -L) {
    r = t;
  } else {
    r = fmodf(t, L);
  }
  if (r != 0.0f && r < 0.0f) r = r + L;
  return r - half;
}

// The orthorhombic part of a Space (28 bytes): small enough to travel in
// registers to an out-of-line function.
struct SpaceO {
  int periodic;
  float L[3];
  float half[3];
};

inline __host__ __device__ SpaceO ortho_of(const Space& s) {
  SpaceO o;
  o.periodic = s.periodic;
  for (int d = 0; d < 3; ++d) {
    o.L[d] = s.L[d];
    o.half[d] = s.half[d];
  }
  return o;
}

// Reference square distance of disp(a, b) = wrap(a - b) in float32 (a
// periodic box or free space).
template <int DIM>
__device__ __forceinline__ float ref_metric_sq(const float* a, const float* b,
                                               const SpaceO& s) {
#pragma clang fp contract(off)
  float acc = 0.0f;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    float v = a[d] - b[d];
    if (s.periodic) v = ref_wrap1(v, s.L[d], s.half[d]);
    float sq = v * v;
    acc = (d == 0) ? sq : acc + sq;
  }
  return acc;
}

// The same for any Space, periodic_general included (T wrap(a - b)).
template <int DIM>
__device__ __forceinline__ float ref_metric_sq(const float* a, const float* b,
                                               const Space& s) {
#pragma clang fp contract(off)
  float v[3];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    v[d] = a[d] - b[d];
    if (s.periodic) v[d] = ref_wrap1(v[d], s.L[d], s.half[d]);
  }
  if (s.general) box_apply_ref<DIM>(s, v);
  float acc = 0.0f;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    float sq = v[d] * v[d];
    acc = (d == 0) ? sq : acc + sq;
  }
  return acc;
}

// Minimum-image displacement x_j - x_i for positions in [0, L), used by the
// pair kernels (tolerance parity only).  A plain x_j - x_i rounds at the scale
// of L when the pair straddles the boundary (16.5 - 0.3 loses ~1e-6), which
// the r^-13 of LJ amplifies.  xim = x_i - L is exact (Sterbenz, x_i >= L/2)
// and precomputed per row; x_j - L is exact for x_j >= L/2; both crossing
// forms then round once at the scale of the small result.
__device__ __forceinline__ float min_image_disp(float xj, float xi, float xim, float L,
                                                float half) {
#pragma clang fp contract(off)
  float raw = xj - xi;
  float lo = xj - xim;         // x_j near 0, x_i near L
  float hi = (xj - L) - xi;    // x_j near L, x_i near 0
  float d = raw < -half ? lo : raw;
  return raw >= half ? hi : d;
}

// The same for positions that may lie anywhere (not known to be in [0, L]):
// when one fold leaves |d| > L/2 (a difference of 1.5 periods or more), the
// remainder is folded by rint.  With inbox (wave-uniform) the check is not
// made: one fold is exact for positions in [0, L].
__device__ __forceinline__ float min_image_any(float xj, float xi, float xim, float L,
                                               float half, int inbox) {
#pragma clang fp contract(off)
  float d = min_image_disp(xj, xi, xim, L, half);
  if (!inbox && !(fabsf(d) <= half)) d = d - L * rintf(d / L);
  return d;
}

// Minimum image of a float64 difference (the x64 form of the periodic
// displacement, space.py:213-224); exact for |d| < L/2, one rounding beyond.
__device__ __forceinline__ double wrap64(double d, double L) {
  const double h = 0.5 * L;
  if (d >= h || d < -h) d = fma(-L, rint(d / L), d);
  return d;
}

// jnp.mod(x, L) for the position update (space.py:250-252), float64.
__device__ __forceinline__ double ref_mod64(double x, double L) {
  if (x >= 0.0 && x < L) return x;  // fmod's own result there, without its cost
  double r = fmod(x, L);
  if (r != 0.0 && r < 0.0) r = r + L;
  return r;
}

// The opening half of simulate.velocity_verlet for one particle
// (simulate.py:168-188,227-239): p' = p + half_dt F, R' = mod(R + dt p' / m, L)
// with R accumulated in float64; writes p', R' (float64) and returns the
// float32 copy of R' in xn.
struct KickDrift {
  int active;
  const double* xhi_in;
  double* xhi_out;
  const float* p_in;
  float* p_out;
  const float* F;
  const float* mass;  // NULL: mass_scalar
  float mass_scalar;
  int mass_pow2;       // scalar mass a power of two: x / m == x * inv_mass exactly
  double inv_mass;
  float dt, half_dt;
  const float* dt_dev;  // device dt overriding dt (half_dt = f32(dt / 2)), or NULL
  float* x_out;         // float32 R', row layout of the kernel's positions
  const float* ref;     // [n, dim] reference positions of the list
  float thr_sq;         // skin criterion (squared)
  int* gate_next;       // raised when any particle moved beyond the skin
  Space s;
  const double* p_scale;  // thermostat factor applied to p before the kick, or NULL
  float Ti[3][3];         // T^-1 of a general box (box_inverse), for the shift
};

inline void kick_drift_mass(KickDrift* kd) {
  int e = 0;
  const double mant = kd->mass ? 0.0 : frexp((double)kd->mass_scalar, &e);
  kd->mass_pow2 = mant == 0.5;
  kd->inv_mass = kd->mass_pow2 ? 1.0 / (double)kd->mass_scalar : 0.0;
}

template <int DIM>
__device__ __forceinline__ void kick_drift_atom(const KickDrift& kd, long long i, float dt,
                                                float half_dt, float* xn) {
#pragma clang fp contract(off)
  const double m = kd.mass ? (double)kd.mass[i] : (double)kd.mass_scalar;
  double dx[3];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const long long k = i * DIM + d;
    const float pin = kd.p_scale ? kd.p_in[k] * (float)kd.p_scale[0] : kd.p_in[k];
    const float pk = pin + half_dt * kd.F[k];
    const double v = (double)dt * (double)pk;
    dx[d] = kd.mass_pow2 ? v * kd.inv_mass : v / m;
    kd.p_out[k] = pk;
  }
  if (kd.s.general) {  // fractional positions: u += T^-1 dR (space.py:450-470)
    double w[3];
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      double acc = 0.0;
#pragma unroll
      for (int b = 0; b < DIM; ++b) acc += (double)kd.Ti[a][b] * dx[b];
      w[a] = acc;
    }
#pragma unroll
    for (int d = 0; d < DIM; ++d) dx[d] = w[d];
  }
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const long long k = i * DIM + d;
    double x = kd.xhi_in[k] + dx[d];
    if (kd.s.periodic) x = ref_mod64(x, (double)kd.s.L[d]);
    kd.xhi_out[k] = x;
    xn[d] = (float)x;
  }
}

// kick_drift_atom with the momentum and force in registers (the force pass
// runs the next step's opening half): p <- p + half_dt F, then the same
// drift in float64, the same roundings; p is left to the caller to store.
template <int DIM>
__device__ __forceinline__ void kick_drift_regs(const KickDrift& kd, long long i, float dt,
                                                float half_dt, float* p, const float* F,
                                                float* xn) {
#pragma clang fp contract(off)
  const double m = kd.mass ? (double)kd.mass[i] : (double)kd.mass_scalar;
  double dx[3];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    p[d] = p[d] + half_dt * F[d];
    const double v = (double)dt * (double)p[d];
    dx[d] = kd.mass_pow2 ? v * kd.inv_mass : v / m;
  }
  if (kd.s.general) {  // fractional positions: u += T^-1 dR (space.py:450-470)
    double w[3];
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      double acc = 0.0;
#pragma unroll
      for (int b = 0; b < DIM; ++b) acc += (double)kd.Ti[a][b] * dx[b];
      w[a] = acc;
    }
#pragma unroll
    for (int d = 0; d < DIM; ++d) dx[d] = w[d];
  }
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const long long k = i * DIM + d;
    double x = kd.xhi_in[k] + dx[d];
    if (kd.s.periodic) x = ref_mod64(x, (double)kd.s.L[d]);
    kd.xhi_out[k] = x;
    xn[d] = (float)x;
  }
}

__device__ __forceinline__ int floor_mod_int(int a, int m) {
  int r = a % m;
  return r < 0 ? r + m : r;
}

// Cell coordinates of a particle: mod(int32(x / cell_size), cps)
// (partition.py:421-422), with IEEE division as XLA emits it.
template <int DIM>
__device__ __forceinline__ int cell_of(const float* x, const Grid& g, int* ci) {
  int h = 0, mult = 1;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    float q = __fdiv_rn(x[d], g.cell_size[d]);
    int iq = (int)q;  // truncation toward zero, as convert_element_type
    int c = floor_mod_int(iq, g.cps[d]);
    ci[d] = c;
    h += c * mult;
    mult *= g.cps[d];
  }
  return h;
}

// Position rows have a stride of STRIDE floats (DIM, or 4 for the padded
// [N, 4] layout whose rows are single 16-byte loads).
template <int DIM, int STRIDE = DIM>
__device__ __forceinline__ void load_pos(const float* __restrict__ p, long long i, float* x) {
  if (DIM == 3 && STRIDE == 4) {
    float4 v = *reinterpret_cast<const float4*>(p + i * 4);
    x[0] = v.x;
    x[1] = v.y;
    x[2] = v.z;
  } else {
#pragma unroll
    for (int d = 0; d < DIM; ++d) x[d] = p[i * STRIDE + d];
  }
}

template <int DIM, int STRIDE = DIM>
__device__ __forceinline__ void store_pos(float* __restrict__ p, long long i, const float* x) {
#pragma unroll
  for (int d = 0; d < DIM; ++d) p[i * STRIDE + d] = x[d];
}

// Valid position strides: dim itself, or 4 for 3-D (padded rows, 16-byte
// aligned base).
inline bool stride_ok(int dim, int stride, const void* base) {
  if (stride == dim) return true;
  return dim == 3 && stride == 4 && ((uintptr_t)base & 15) == 0;
}

// Wavefront helpers (64 lanes).
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ int wave_prefix_count(unsigned long long mask) {
  // Number of set bits in mask below this lane (mbcnt).
  return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

template <typename T>
__device__ __forceinline__ T wave_reduce_max(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    T o = __shfl_xor(v, off, kWave);
    v = v > o ? v : o;
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_reduce_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Fill kernel used instead of hipMemsetAsync (safe under graph capture).
template <typename T>
__global__ void fill_kernel(T* __restrict__ p, long long n, T v) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

template <typename T>
inline int launch_fill(T* p, long long n, T v, hipStream_t st) {
  if (n <= 0) return JMD_OK;
  int blocks = ceil_div(n, 256);
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(fill_kernel<T>, dim3(blocks), dim3(256), 0, st, p, n, v);
  JMD_LAUNCH_CHECK();
  return JMD_OK;
}

// Gate: rebuild work runs only when *gate != 0 (or gate == nullptr).
__device__ __forceinline__ bool gate_closed(const int* gate) {
  return gate != nullptr && *gate == 0;
}

}  // namespace jmd
