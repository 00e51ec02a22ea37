// Pair-potential arithmetic shared by the force and adjoint kernels:
// energy.lennard_jones with multiplicative_isotropic_cutoff and
// energy.soft_sphere (jax_md/energy.py:125-173,246-272,534-580) evaluated
// from r^2 in float32, parameter selection (smap.py:697-823), and helpers.
#pragma once

#include "jmd_common.hpp"

namespace jmd {

struct PairArgs {
  int kind;
  int general;          // any non-scalar parameter
  int sig_mode, eps_mode, alp_mode;
  float sig, eps, alp;
  const float* sig_v;
  const float* eps_v;
  const float* alp_v;
  const int* species;
  int ns;
  float ro2, rc2, inv_d;  // LJ smoothing: r_on^2, r_c^2, 1/(rc^2-ro^2)^3
  float c1;               // rc^2 - 3 ro^2
  float ro;               // LJ r_on
  int smooth;             // r_on < r_c
  float rc;               // interaction range (LJ r_c, soft sphere max sigma)
};

__device__ __forceinline__ float pick(int mode, float scalar, const float* v, int i, int j,
                                      int si, int sj, int ns) {
  if (mode == JMD_PARAM_SCALAR) return scalar;
  if (mode == JMD_PARAM_PER_PARTICLE) return 0.5f * (v[i] + v[j]);
  return v[si * ns + sj];
}

// 1/x from v_rcp_f32 plus one Newton step (~0.5 ulp, no IEEE division).
__device__ __forceinline__ float fast_rcp(float x) {
#pragma clang fp contract(off)
  float r = __builtin_amdgcn_rcpf(x);
  float e = fmaf(-x, r, 1.0f);
  return fmaf(r, e, r);
}

// 1/sqrt(x) from v_rsq_f32 plus one Newton step.
__device__ __forceinline__ float fast_rsqrt(float x) {
#pragma clang fp contract(off)
  float y = __builtin_amdgcn_rsqf(x);
  float h = 0.5f * x * y;
  float e = fmaf(-h, y, 0.5f);
  return fmaf(y, e, y);
}

// U(r) and dU/dr / r for one pair from r^2, without branches; callers mask
// invalid lanes with a select.  Multiply-adds are explicit (contraction off)
// so every kernel variant rounds alike.
template <int KIND, int SMOOTH = -1>  // SMOOTH: -1 read a.smooth, 0/1 fixed
__device__ __forceinline__ float pair_eval(float r2, float sig, float eps, float alp,
                                           const PairArgs& a, float* dEr) {
#pragma clang fp contract(off)
  if (KIND == JMD_LENNARD_JONES) {
    // With t = sigma^2 / r^2, Vp = t^6 - t^3 (V = 4 eps Vp) and
    // Wp = 2 t^6 - t^3 (V'/r = -24 eps Wp / r^2).  The smoothing factor is
    // evaluated at rr = r^2 clamped to [ro^2, rc^2] (one med3): S(rr) = 1
    // below r_on and 0 beyond r_c, and S'(rr) = 0 at both ends, so one
    // branch-free expression covers all three regions and the cutoff:
    //   u = rc^2 - rr, q = ro^2 - rr, p = rc^2 + 2 rr - 3 ro^2,
    //   S = u^2 p / D,  S'/r = 12 u q / D,   D = (rc^2 - ro^2)^3,
    //   (S V)'/r = 24 eps u (2 q Vp - u p Wp / r^2) / D.
    // Below r_on this is -24 eps Wp / r^2 up to the rounding of u^2 p / D.
    // v_rcp_f32 alone (1 ulp): the Newton step moved the force error
    // against FP64 by under 5% and cost two operations per pair
    const float inv_r2 = __builtin_amdgcn_rcpf(r2);
    const float s2 = sig * sig * inv_r2;
    const float s6 = s2 * s2 * s2;
    const float s12 = s6 * s6;
    const float Vp = s12 - s6;
    const float Wp = fmaf(2.0f, s12, -s6);
    if (SMOOTH == 1 || (SMOOTH < 0 && a.smooth)) {
      const float rr = __builtin_amdgcn_fmed3f(r2, a.ro2, a.rc2);
      const float u = a.rc2 - rr;
      const float q = a.ro2 - rr;
      const float pp = fmaf(2.0f, rr, a.c1);
      const float t2 = u * pp * Wp * inv_r2;
      const float inner = fmaf(2.0f, q * Vp, -t2);
      *dEr = 24.0f * eps * a.inv_d * u * inner;
      return 4.0f * eps * a.inv_d * (u * u * pp) * Vp;
    }
    const bool valid = r2 < a.rc2;
    *dEr = valid ? -24.0f * eps * Wp * inv_r2 : 0.0f;
    return valid ? 4.0f * eps * Vp : 0.0f;
  } else {
    // soft sphere: eps/alpha (1 - r/sigma)^alpha for r < sigma
    float inv_r = fast_rsqrt(r2);
    float r = r2 * inv_r;
    float inv_sig = fast_rcp(sig);
    float base = 1.0f - r * inv_sig;
    float up, um;  // base^alpha, base^(alpha-1)
    if (alp == 2.0f) { um = base; up = base * base; }
    else { um = powf(fmaxf(base, 0.0f), alp - 1.0f); up = um * base; }
    bool valid = (r2 < sig * sig) && (r2 > 0.0f);
    *dEr = valid ? -eps * inv_sig * um * inv_r : 0.0f;
    return valid ? eps / alp * up : 0.0f;
  }
}

template <int TPA, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int off = TPA / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// XCD-aware block order: consecutive logical tiles go to the same XCD so each
// XCD's L2 holds a compact slab of the position table.  The grid is padded to
// a multiple of 8 by the caller.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  int per = nb >> 3;
  return (b & 7) * per + (b >> 3);
}

inline bool make_pair_args(const jmd_pair_potential* p, PairArgs* a) {
  if (!p) return false;
  if (p->kind != JMD_LENNARD_JONES && p->kind != JMD_SOFT_SPHERE) return false;
  a->kind = p->kind;
  a->sig_mode = p->sigma.mode;
  a->eps_mode = p->epsilon.mode;
  a->alp_mode = p->kind == JMD_SOFT_SPHERE ? p->alpha.mode : JMD_PARAM_SCALAR;
  for (int m : {a->sig_mode, a->eps_mode, a->alp_mode})
    if (m < JMD_PARAM_SCALAR || m > JMD_PARAM_SPECIES) return false;
  a->sig = p->sigma.value;
  a->eps = p->epsilon.value;
  a->alp = p->alpha.value;
  a->sig_v = p->sigma.values;
  a->eps_v = p->epsilon.values;
  a->alp_v = p->alpha.values;
  a->species = p->species;
  a->ns = p->n_species;
  bool need_species = false;
  const jmd_param* ps[3] = {&p->sigma, &p->epsilon, &p->alpha};
  for (int q = 0; q < 3; ++q) {
    if (q == 2 && p->kind != JMD_SOFT_SPHERE) continue;
    if (ps[q]->mode != JMD_PARAM_SCALAR && !ps[q]->values) return false;
    if (ps[q]->mode == JMD_PARAM_SPECIES) need_species = true;
  }
  if (need_species && (!p->species || p->n_species < 1)) return false;
  a->general = (a->sig_mode | a->eps_mode | a->alp_mode) != 0;
  float ro = p->r_onset, rc = p->r_cutoff;
  a->rc2 = rc * rc;
  a->ro2 = ro * ro;
  a->rc = rc;
  a->ro = ro;
  if (p->kind == JMD_SOFT_SPHERE) {
    // Interaction range (used only for the wrap-free interior test): a scalar
    // sigma, else the caller's r_cutoff bound on sigma (0: unknown, never
    // interior).
    ro = rc = 0.0f;
    a->rc = p->sigma.mode == JMD_PARAM_SCALAR ? p->sigma.value
                                               : (p->r_cutoff > 0.0f ? p->r_cutoff : 3.0e38f);
  }
  a->smooth = p->kind == JMD_LENNARD_JONES && ro < rc;
  a->inv_d = 0.0f;
  a->c1 = fmaf(-3.0f, a->ro2, a->rc2);
  if (a->smooth) {
    double dd = (double)a->rc2 - (double)a->ro2;
    a->inv_d = (float)(1.0 / (dd * dd * dd));
  }
  return true;
}

// Second-order pair terms from s = r^2 (float32, contraction off):
//   g = U'(r)/r, gp = dg/ds, gs = dg/dsigma, ge = dg/depsilon,
//   us = dU/dsigma, ue = dU/depsilon.
struct PairD2 {
  float g, gp, gs, ge, us, ue;
};

template <int KIND>
__device__ __forceinline__ PairD2 pair_eval2(float s, float sig, float eps, float alp,
                                             const PairArgs& a) {
#pragma clang fp contract(off)
  PairD2 o;
  if (KIND == JMD_LENNARD_JONES) {
    // V = 4 eps (q^6 - q^3), q = sigma^2 / s:
    //   g_V = 2 dV/ds = -24 eps (2 q^6 - q^3) / s
    //   dg_V/ds = 48 eps (7 q^6 - 2 q^3) / s^2
    //   dV/dsigma = 24 eps (2 q^6 - q^3) / sigma,  dg_V/dsigma = -144 eps (4 q^6 - q^3) / (sigma s)
    const float inv = fast_rcp(s);
    const float inv_sig = fast_rcp(sig);
    const float q = sig * sig * inv;
    const float s6 = q * q * q;
    const float s12 = s6 * s6;
    const float c2 = fmaf(2.0f, s12, -s6);
    const float V = 4.0f * eps * (s12 - s6);
    const float gV = -24.0f * eps * c2 * inv;
    const float gVp = 48.0f * eps * fmaf(7.0f, s12, -2.0f * s6) * inv * inv;
    const float V_sig = 24.0f * eps * c2 * inv_sig;
    const float gV_sig = -144.0f * eps * fmaf(4.0f, s12, -s6) * inv_sig * inv;
    const float V_eps = 4.0f * (s12 - s6);
    const float gV_eps = -24.0f * c2 * inv;
    o.g = gV;
    o.gp = gVp;
    o.gs = gV_sig;
    o.ge = gV_eps;
    o.us = V_sig;
    o.ue = V_eps;
    if (a.smooth && s >= a.ro2) {
      // U = V S(s), S = (rc^2 - s)^2 (rc^2 + 2 s - 3 ro^2) / (rc^2 - ro^2)^3
      const float u = a.rc2 - s;
      const float S = u * u * fmaf(2.0f, s, fmaf(-3.0f, a.ro2, a.rc2)) * a.inv_d;
      const float Ss = 6.0f * u * (a.ro2 - s) * a.inv_d;
      const float Sss = 6.0f * (2.0f * s - a.ro2 - a.rc2) * a.inv_d;
      o.g = fmaf(S, gV, 2.0f * Ss * V);
      o.gp = fmaf(gVp, S, fmaf(2.0f * gV, Ss, 2.0f * V * Sss));
      o.gs = fmaf(S, gV_sig, 2.0f * Ss * V_sig);
      o.ge = fmaf(S, gV_eps, 2.0f * Ss * V_eps);
      o.us = S * V_sig;
      o.ue = S * V_eps;
    }
    if (!(s < a.rc2)) o = PairD2{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  } else {
    // U = eps / alpha b^alpha, b = 1 - r / sigma, for 0 < r < sigma:
    //   g = -eps b^(alpha-1) / (sigma r)
    //   dg/ds = eps / (2 sigma r) [(alpha-1) b^(alpha-2) / (sigma r) + b^(alpha-1) / r^2]
    //   dU/dsigma = eps b^(alpha-1) r / sigma^2
    //   dg/dsigma = -eps / (r sigma^2) [(alpha-1) b^(alpha-2) r / sigma - b^(alpha-1)]
    const float inv_r = fast_rsqrt(s);
    const float r = s * inv_r;
    const float inv_sig = fast_rcp(sig);
    const float b = 1.0f - r * inv_sig;
    float bm1, bm2;
    if (alp == 2.0f) {
      bm1 = b;
      bm2 = 1.0f;
    } else {
      bm2 = powf(fmaxf(b, 0.0f), alp - 2.0f);
      bm1 = bm2 * b;
    }
    const bool valid = (s < sig * sig) && (s > 0.0f);
    o.g = -eps * inv_sig * bm1 * inv_r;
    o.gp = 0.5f * eps * inv_sig * inv_r *
           fmaf((alp - 1.0f) * bm2, inv_sig * inv_r, bm1 * inv_r * inv_r);
    o.ge = -inv_sig * bm1 * inv_r;
    o.ue = bm1 * b / alp;
    o.us = eps * bm1 * r * inv_sig * inv_sig;
    o.gs = -eps * inv_r * inv_sig * inv_sig * fmaf((alp - 1.0f) * bm2, r * inv_sig, -bm1);
    if (!valid) o = PairD2{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  }
  return o;
}

}  // namespace jmd
