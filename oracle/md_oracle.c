/*
 * md_oracle.c -- FP64 C restatement of the neighbour-list LJ NVE path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ as a large-N checker and by the
 * cpu_baseline leg of bench.py as the timed CPU port (the reference itself
 * cannot run here: JAX is not installed).  Not linked into the product.
 *
 * Follows the reference algorithm:
 *   neighbour list: cell grid with cells of size >= r_cutoff + skin, 3^3
 *     stencil, pairs with |wrap(R_j - R_i)| < cutoff kept in a Dense row per
 *     particle (jax_md/partition.py:911-980);
 *   skin criterion: rebuild when any |wrap(R - R_ref)|^2 > (skin/2)^2
 *     (partition.py:1141-1154);
 *   pair energy: LJ times the C1 switch of multiplicative_isotropic_cutoff
 *     (jax_md/energy.py:246-272,534-580), forces analytic, summed over the
 *     full list with weight 1/2 per ordered pair (smap.py:922-956);
 *   integrator: velocity Verlet with dt cast to float32
 *     (jax_md/simulate.py:227-243), positions wrapped with floor mod
 *     (space.py:250-252).
 * Everything is float64; parallel over particles with OpenMP.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  int n;
  double L;
  double cutoff;   /* list cutoff r_c + skin */
  double skin;
  int width;       /* Dense row width */
  int *idx;        /* [n * width], -1 padded */
  int *count;      /* [n] */
  double *ref;     /* [n * 3] */
  int overflow;
} md_list;

static double wrap(double d, double L) {
  double r = fmod(d + 0.5 * L, L);
  if (r < 0) r += L;
  return r - 0.5 * L;
}

static double fmod_pos(double x, double L) {
  double r = fmod(x, L);
  if (r != 0.0 && r < 0.0) r += L;
  return r;
}

int md_list_build(md_list *nl, const double *R) {
  const int n = nl->n;
  const double L = nl->L, c = nl->cutoff, c2 = c * c;
  int cps = (int)floor(L / c);
  if (cps < 3) cps = 1;
  const double cs = L / cps;
  const int ncell = cps * cps * cps;
  int *head = (int *)malloc(sizeof(int) * ncell);
  int *next = (int *)malloc(sizeof(int) * n);
  int *cell = (int *)malloc(sizeof(int) * n);
  if (!head || !next || !cell) return -1;
  for (int k = 0; k < ncell; ++k) head[k] = -1;
  for (int i = 0; i < n; ++i) {
    int ci[3];
    for (int d = 0; d < 3; ++d) {
      int q = (int)(R[3 * i + d] / cs);
      q %= cps;
      if (q < 0) q += cps;
      ci[d] = q;
    }
    int h = ci[0] + cps * (ci[1] + cps * ci[2]);
    cell[i] = h;
  }
  /* insert in reverse so each chain lists ids ascending */
  for (int i = n - 1; i >= 0; --i) {
    next[i] = head[cell[i]];
    head[cell[i]] = i;
  }
  int overflow = 0;
#pragma omp parallel for schedule(static) reduction(| : overflow)
  for (int i = 0; i < n; ++i) {
    int h = cell[i];
    int cx = h % cps, cy = (h / cps) % cps, cz = h / (cps * cps);
    int cnt = 0;
    int *row = nl->idx + (size_t)i * nl->width;
    int span = cps >= 3 ? 1 : 0;
    for (int dx = -span; dx <= span; ++dx)
      for (int dy = -span; dy <= span; ++dy)
        for (int dz = -span; dz <= span; ++dz) {
          int x = (cx + dx + cps) % cps, y = (cy + dy + cps) % cps, z = (cz + dz + cps) % cps;
          for (int j = head[x + cps * (y + cps * z)]; j >= 0; j = next[j]) {
            if (j == i) continue;
            double r2 = 0.0;
            for (int d = 0; d < 3; ++d) {
              double v = wrap(R[3 * j + d] - R[3 * i + d], L);
              r2 += v * v;
            }
            if (r2 < c2) {
              if (cnt < nl->width) row[cnt] = j;
              else overflow = 1;
              ++cnt;
            }
          }
        }
    for (int k = cnt; k < nl->width; ++k) row[k] = -1;
    nl->count[i] = cnt;
  }
  memcpy(nl->ref, R, sizeof(double) * 3 * n);
  nl->overflow |= overflow;
  free(head);
  free(next);
  free(cell);
  return 0;
}

md_list *md_list_create(int n, double L, double cutoff, double skin, int width) {
  md_list *nl = (md_list *)calloc(1, sizeof(md_list));
  nl->n = n;
  nl->L = L;
  nl->cutoff = cutoff;
  nl->skin = skin;
  nl->width = width;
  nl->idx = (int *)malloc(sizeof(int) * (size_t)n * width);
  nl->count = (int *)malloc(sizeof(int) * n);
  nl->ref = (double *)malloc(sizeof(double) * 3 * n);
  return nl;
}

void md_list_free(md_list *nl) {
  if (!nl) return;
  free(nl->idx);
  free(nl->count);
  free(nl->ref);
  free(nl);
}

int md_list_overflow(const md_list *nl) { return nl->overflow; }

int md_list_max_count(const md_list *nl) {
  int m = 0;
  for (int i = 0; i < nl->n; ++i)
    if (nl->count[i] > m) m = nl->count[i];
  return m;
}

/* LJ energy and force over the list; returns the total energy. */
double md_lj_forces(const md_list *nl, const double *R, double sigma, double eps,
                    double r_on, double r_c, double *F) {
  const int n = nl->n;
  const double L = nl->L;
  const double rc2 = r_c * r_c, ro2 = r_on * r_on;
  const double D = (rc2 > ro2) ? (rc2 - ro2) * (rc2 - ro2) * (rc2 - ro2) : 1.0;
  double E = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : E)
  for (int i = 0; i < n; ++i) {
    double f[3] = {0, 0, 0}, e = 0.0;
    const int *row = nl->idx + (size_t)i * nl->width;
    for (int k = 0; k < nl->width; ++k) {
      int j = row[k];
      if (j < 0) break;
      double d[3], r2 = 0.0;
      for (int c = 0; c < 3; ++c) {
        d[c] = wrap(R[3 * j + c] - R[3 * i + c], L);
        r2 += d[c] * d[c];
      }
      if (r2 >= rc2) continue;
      double s2 = sigma * sigma / r2, s6 = s2 * s2 * s2;
      double V = 4.0 * eps * (s6 * s6 - s6);
      double dV = -24.0 * eps * (2.0 * s6 * s6 - s6) / r2;
      double S = 1.0, dS = 0.0;
      if (rc2 > ro2 && r2 >= ro2) {
        double u = rc2 - r2;
        S = u * u * (rc2 + 2.0 * r2 - 3.0 * ro2) / D;
        dS = 12.0 * u * (ro2 - r2) / D;
      }
      double dEr = S * dV + dS * V;
      for (int c = 0; c < 3; ++c) f[c] += dEr * d[c];
      e += S * V;
    }
    for (int c = 0; c < 3; ++c) F[3 * i + c] = f[c];
    E += 0.5 * e;
  }
  return E;
}

static int needs_rebuild(const md_list *nl, const double *R) {
  const double thr = 0.25 * nl->skin * nl->skin;
  int any = 0;
#pragma omp parallel for schedule(static) reduction(| : any)
  for (int i = 0; i < nl->n; ++i) {
    double r2 = 0.0;
    for (int d = 0; d < 3; ++d) {
      double v = wrap(R[3 * i + d] - nl->ref[3 * i + d], nl->L);
      r2 += v * v;
    }
    if (r2 > thr) any = 1;
  }
  return any;
}

/* n_steps velocity-Verlet steps; returns the number of list rebuilds. */
int md_nve_run(md_list *nl, double *R, double *P, double *F, double sigma, double eps,
               double r_on, double r_c, double dt, int n_steps) {
  const int n = nl->n;
  const double dt32 = (double)(float)dt, half = (double)((float)dt32 / 2.0f);
  int rebuilds = 0;
  for (int s = 0; s < n_steps; ++s) {
    if (needs_rebuild(nl, R)) {
      md_list_build(nl, R);
      ++rebuilds;
    }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < 3 * n; ++k) {
      P[k] += half * F[k];
      R[k] = fmod_pos(R[k] + dt32 * P[k], nl->L);
    }
    md_lj_forces(nl, R, sigma, eps, r_on, r_c, F);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < 3 * n; ++k) P[k] += half * F[k];
  }
  return rebuilds;
}

int md_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
