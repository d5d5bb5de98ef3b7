/*
 * CPU baseline: FP64 C/OpenMP restatement of the reference's Numba CPU plugin
 * (OpenONDA/solvers/VPM/ParticlesPhysics.py, @njit(parallel=True,
 * fastmath=True) with prange over targets).  TEST INFRASTRUCTURE: built by
 * oracle/build_cpu.py, used only by tests and by bench.py's cpu_baseline leg.
 *
 * Built with -O3 -ffast-math -fopenmp to mirror fastmath=True and prange.
 * A**2.5, A**3.5 and A**4.5 are written as A^k sqrt(A), the form LLVM's
 * fast-math pow simplification gives Numba's code; it also keeps GCC's
 * vectorised pow out (GCC 11 -O3 -ffast-math -march=x86-64-v3 computes wrong
 * values with it).
 * The vpm_cpu_cuda_* functions restate the GPU plugin's forms instead
 * (ParticlesPhysicsCuda.py), for full-resolution solver-step checks.
 * Arrays are the reference's AoS (N,3) float64 rows.  Each function
 * evaluates the targets listed in tidx[0..m) against all n sources (the
 * full evaluation is tidx = 0..n-1), so a strided target sample can be timed.
 */
#include <math.h>
#include <stdint.h>

#define ONE_OVER_FOUR_PI 0.079577471545947667884441881686257181
#define EPSILON 1e-8

static inline double pow25(double a) { return a * a * sqrt(a); }
static inline double pow35(double a) { return a * a * a * sqrt(a); }
static inline double pow45(double a) { return (a * a) * (a * a) * sqrt(a); }

/* ParticlesPhysics.py:81-143: sigma = sigma_i for every j != i.  The j loop
 * is split around j = i instead of skipping it with a branch. */
static inline void vel_range(int64_t j0, int64_t j1, const double* x, const double* a,
                             double tx, double ty, double tz, double sigma, double sigma_sq,
                             double* ux, double* uy, double* uz) {
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int64_t j = j0; j < j1; ++j) {
    const double rx = tx - x[3 * j], ry = ty - x[3 * j + 1], rz = tz - x[3 * j + 2];
    const double r2 = rx * rx + ry * ry + rz * rz;
    const double rho_sq = r2 / sigma_sq;
    const double q = (rho_sq + 2.5) / (pow25(rho_sq + 1.0) * sigma * sigma_sq);
    const double ax = a[3 * j], ay = a[3 * j + 1], az = a[3 * j + 2];
    sx -= q * (ry * az - rz * ay);
    sy -= q * (rz * ax - rx * az);
    sz -= q * (rx * ay - ry * ax);
  }
  *ux += sx;
  *uy += sy;
  *uz += sz;
}

void vpm_cpu_selfinduced_velocity(int64_t n, const double* x, const double* a, const double* s,
                                  int64_t m, const int64_t* tidx, double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < m; ++k) {
    const int64_t i = tidx[k];
    const double tx = x[3 * i], ty = x[3 * i + 1], tz = x[3 * i + 2];
    const double sigma = s[i];
    const double sigma_sq = sigma * sigma;
    double ux = 0.0, uy = 0.0, uz = 0.0;
    vel_range(0, i, x, a, tx, ty, tz, sigma, sigma_sq, &ux, &uy, &uz);
    vel_range(i + 1, n, x, a, tx, ty, tz, sigma, sigma_sq, &ux, &uy, &uz);
    out[3 * k] = ux * ONE_OVER_FOUR_PI;
    out[3 * k + 1] = uy * ONE_OVER_FOUR_PI;
    out[3 * k + 2] = uz * ONE_OVER_FOUR_PI;
  }
}

/* ParticlesPhysics.py:268-312 (sigma_ij, j != i). */
void vpm_cpu_strength_gradients(int64_t n, const double* x, const double* a, const double* s,
                                int64_t m, const int64_t* tidx, double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < m; ++k) {
    const int64_t i = tidx[k];
    const double tx = x[3 * i], ty = x[3 * i + 1], tz = x[3 * i + 2];
    const double ix = a[3 * i], iy = a[3 * i + 1], iz = a[3 * i + 2];
    const double si = s[i];
    double dx = 0.0, dy = 0.0, dz = 0.0;
    for (int64_t j = 0; j < n; ++j) {
      if (j == i) continue;
      const double rx = tx - x[3 * j], ry = ty - x[3 * j + 1], rz = tz - x[3 * j + 2];
      const double d2 = rx * rx + ry * ry + rz * rz;
      const double sigma = (si + s[j]) * 0.5;
      const double sigma_cb = sigma * sigma * sigma;
      const double A = d2 / (sigma * sigma) + 1.0;
      const double f1 = (A + 1.5) / (sigma_cb * pow25(A));
      const double f2 = 3.0 * (A + 2.5) / (sigma_cb * sigma * sigma * pow35(A));
      const double ax = a[3 * j], ay = a[3 * j + 1], az = a[3 * j + 2];
      const double p1x = iy * az - iz * ay, p1y = iz * ax - ix * az, p1z = ix * ay - iy * ax;
      const double p2x = ry * az - rz * ay, p2y = rz * ax - rx * az, p2z = rx * ay - ry * ax;
      const double p3 = ix * p2x + iy * p2y + iz * p2z;
      dx += f1 * p1x + f2 * p3 * rx;
      dy += f1 * p1y + f2 * p3 * ry;
      dz += f1 * p1z + f2 * p3 * rz;
    }
    out[3 * k] = dx * ONE_OVER_FOUR_PI;
    out[3 * k + 1] = dy * ONE_OVER_FOUR_PI;
    out[3 * k + 2] = dz * ONE_OVER_FOUR_PI;
  }
}

/* ParticlesPhysics.py:318-382: nu_ij, A + 1e-8, 1/(4pi) at the end. */
void vpm_cpu_strength_gradients_pse(int64_t n, const double* x, const double* a, const double* s,
                                    const double* nu, int64_t m, const int64_t* tidx,
                                    double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < m; ++k) {
    const int64_t i = tidx[k];
    const double tx = x[3 * i], ty = x[3 * i + 1], tz = x[3 * i + 2];
    const double ix = a[3 * i], iy = a[3 * i + 1], iz = a[3 * i + 2];
    double dx = 0.0, dy = 0.0, dz = 0.0;
    for (int64_t j = 0; j < n; ++j) {
      if (j == i) continue;
      const double vis = (nu[i] + nu[j]) * 0.5;
      const double sigma = (s[i] + s[j]) * 0.5;
      const double rx = tx - x[3 * j], ry = ty - x[3 * j + 1], rz = tz - x[3 * j + 2];
      const double d2 = rx * rx + ry * ry + rz * rz;
      const double sigma_sq = sigma * sigma, sigma_cb = sigma_sq * sigma;
      const double A = d2 / sigma_sq + 1.0 + EPSILON;
      const double f3 = 105.0 * vis / (sigma_sq * sigma_cb * pow45(A));
      dx += f3 * sigma_cb * (a[3 * j] - ix);
      dy += f3 * sigma_cb * (a[3 * j + 1] - iy);
      dz += f3 * sigma_cb * (a[3 * j + 2] - iz);
    }
    out[3 * k] = dx * ONE_OVER_FOUR_PI;
    out[3 * k + 1] = dy * ONE_OVER_FOUR_PI;
    out[3 * k + 2] = dz * ONE_OVER_FOUR_PI;
  }
}

/* ---- GPU-plugin arithmetic (ParticlesPhysicsCuda.py), FP64 -------------
 * The same sums in the GPU plugin's forms, for checking full-resolution
 * solver steps in seconds where the NumPy oracle would take minutes.  They
 * are pinned against the NumPy oracle (variant "cuda") at small N. */

/* ParticlesPhysicsCuda.py:133-191: sigma_ij = (sigma_i + sigma_j)/2, j != i. */
void vpm_cpu_cuda_selfinduced_velocity(int64_t n, const double* x, const double* a,
                                       const double* s, int64_t m, const int64_t* tidx,
                                       double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < m; ++k) {
    const int64_t i = tidx[k];
    const double tx = x[3 * i], ty = x[3 * i + 1], tz = x[3 * i + 2];
    const double si = s[i];
    double ux = 0.0, uy = 0.0, uz = 0.0;
    for (int64_t j = 0; j < n; ++j) {
      if (j == i) continue;
      const double rx = tx - x[3 * j], ry = ty - x[3 * j + 1], rz = tz - x[3 * j + 2];
      const double r2 = rx * rx + ry * ry + rz * rz;
      const double sg = (si + s[j]) / 2.0;
      const double sg2 = sg * sg;
      const double rho2 = r2 / sg2;
      const double q = (rho2 + 2.5) / (pow25(rho2 + 1.0) * sg * sg2);
      const double ax = a[3 * j], ay = a[3 * j + 1], az = a[3 * j + 2];
      ux -= q * (ry * az - rz * ay);
      uy -= q * (rz * ax - rx * az);
      uz -= q * (rx * ay - ry * ax);
    }
    out[3 * k] = ux * ONE_OVER_FOUR_PI;
    out[3 * k + 1] = uy * ONE_OVER_FOUR_PI;
    out[3 * k + 2] = uz * ONE_OVER_FOUR_PI;
  }
}

/* ParticlesPhysicsCuda.py:291-346 (simplified form): sigma_j, r^2 > 1e-8.
 * Targets are the sources' own positions (tidx), as vpmModule.py calls it. */
void vpm_cpu_cuda_vorticity(int64_t n, const double* x, const double* a, const double* s,
                            int64_t m, const int64_t* tidx, double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < m; ++k) {
    const int64_t i = tidx[k];
    const double tx = x[3 * i], ty = x[3 * i + 1], tz = x[3 * i + 2];
    double wx = 0.0, wy = 0.0, wz = 0.0;
    for (int64_t j = 0; j < n; ++j) {
      const double rx = tx - x[3 * j], ry = ty - x[3 * j + 1], rz = tz - x[3 * j + 2];
      const double r2 = rx * rx + ry * ry + rz * rz;
      if (!(r2 > EPSILON)) continue;
      const double sj = s[j];
      const double rho2 = r2 / (sj * sj);
      const double xi = 7.5 / (pow35(rho2 + 1.0) * (sj * sj * sj));
      wx += xi * a[3 * j];
      wy += xi * a[3 * j + 1];
      wz += xi * a[3 * j + 2];
    }
    out[3 * k] = wx * ONE_OVER_FOUR_PI;
    out[3 * k + 1] = wy * ONE_OVER_FOUR_PI;
    out[3 * k + 2] = wz * ONE_OVER_FOUR_PI;
  }
}

/* ParticlesPhysicsCuda.py:478-530: nu_i, A = rho^2 + 1, j != i. */
void vpm_cpu_cuda_strength_gradients_pse(int64_t n, const double* x, const double* a,
                                         const double* s, const double* nu, int64_t m,
                                         const int64_t* tidx, double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < m; ++k) {
    const int64_t i = tidx[k];
    const double tx = x[3 * i], ty = x[3 * i + 1], tz = x[3 * i + 2];
    const double ix = a[3 * i], iy = a[3 * i + 1], iz = a[3 * i + 2];
    const double vis = nu[i];
    double dx = 0.0, dy = 0.0, dz = 0.0;
    for (int64_t j = 0; j < n; ++j) {
      if (j == i) continue;
      const double sg = (s[i] + s[j]) * 0.5;
      const double rx = tx - x[3 * j], ry = ty - x[3 * j + 1], rz = tz - x[3 * j + 2];
      const double r2 = rx * rx + ry * ry + rz * rz;
      const double sg2 = sg * sg, sg3 = sg2 * sg;
      const double A = r2 / sg2 + 1.0;
      const double w = 105.0 * vis / (sg2 * sg3 * pow45(A)) * sg3;
      dx += w * (a[3 * j] - ix);
      dy += w * (a[3 * j + 1] - iy);
      dz += w * (a[3 * j + 2] - iz);
    }
    out[3 * k] = dx * ONE_OVER_FOUR_PI;
    out[3 * k + 1] = dy * ONE_OVER_FOUR_PI;
    out[3 * k + 2] = dz * ONE_OVER_FOUR_PI;
  }
}

/* ParticlesPhysicsCuda.py:867-920 + :777: i = j included, 1/(8pi) 1/(4pi).
 * Per-target partial sums in out[0..n), added in index order by the caller
 * side (deterministic for any thread count). */
void vpm_cpu_cuda_enstrophy_rows(int64_t n, const double* x, const double* a, const double* s,
                                 double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const double tx = x[3 * i], ty = x[3 * i + 1], tz = x[3 * i + 2];
    const double ix = a[3 * i], iy = a[3 * i + 1], iz = a[3 * i + 2];
    double acc = 0.0;
    for (int64_t j = 0; j < n; ++j) {
      const double rx = tx - x[3 * j], ry = ty - x[3 * j + 1], rz = tz - x[3 * j + 2];
      const double r2 = rx * rx + ry * ry + rz * rz;
      const double sg = (s[i] + s[j]) * 0.5;
      const double sg2 = sg * sg;
      const double d = r2 + sg2;
      const double t1 = (r2 + 3.2 * sg2) / (d * sqrt(d));
      acc += t1 * (ix * a[3 * j] + iy * a[3 * j + 1] + iz * a[3 * j + 2]);
    }
    out[i] = acc;
  }
}

int vpm_cpu_max_threads(void) {
#ifdef _OPENMP
  extern int omp_get_max_threads(void);
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* threads used by the following parallel loops of the calling thread */
void vpm_cpu_set_threads(int n) {
#ifdef _OPENMP
  extern void omp_set_num_threads(int);
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}
