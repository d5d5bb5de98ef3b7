/* The C/OpenMP FP64 port (oracle/vpm_cpu.c) under AddressSanitizer and
 * UndefinedBehaviorSanitizer: every function on a small random set with a
 * coincident pair, on all targets and on a strided target subset, with the
 * outputs checked finite.  Exits non-zero on a non-finite output. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

void vpm_cpu_selfinduced_velocity(int64_t, const double*, const double*, const double*, int64_t,
                                  const int64_t*, double*);
void vpm_cpu_strength_gradients(int64_t, const double*, const double*, const double*, int64_t,
                                const int64_t*, double*);
void vpm_cpu_strength_gradients_pse(int64_t, const double*, const double*, const double*,
                                    const double*, int64_t, const int64_t*, double*);
void vpm_cpu_cuda_selfinduced_velocity(int64_t, const double*, const double*, const double*,
                                       int64_t, const int64_t*, double*);
void vpm_cpu_cuda_vorticity(int64_t, const double*, const double*, const double*, int64_t,
                            const int64_t*, double*);
void vpm_cpu_cuda_strength_gradients_pse(int64_t, const double*, const double*, const double*,
                                         const double*, int64_t, const int64_t*, double*);
void vpm_cpu_cuda_enstrophy_rows(int64_t, const double*, const double*, const double*, double*);
int vpm_cpu_max_threads(void);
void vpm_cpu_set_threads(int);

static int finite_all(const double* v, int64_t m) {
  for (int64_t i = 0; i < m; ++i)
    if (!isfinite(v[i])) return 0;
  return 1;
}

int main(void) {
  enum { N = 257, M = 19 };
  double* x = malloc(3 * N * sizeof(double));
  double* a = malloc(3 * N * sizeof(double));
  double* s = malloc(N * sizeof(double));
  double* nu = malloc(N * sizeof(double));
  double* out = malloc(3 * N * sizeof(double));
  int64_t* all = malloc(N * sizeof(int64_t));
  int64_t sub[M];
  srand(5);
  for (int i = 0; i < N; ++i) {
    for (int c = 0; c < 3; ++c) {
      x[3 * i + c] = (double)rand() / RAND_MAX;
      a[3 * i + c] = (double)rand() / RAND_MAX - 0.5;
    }
    s[i] = 0.05 + 0.05 * (double)rand() / RAND_MAX;
    nu[i] = 1e-3;
    all[i] = i;
  }
  for (int c = 0; c < 3; ++c) x[3 * 7 + c] = x[3 * 3 + c]; /* a coincident pair */
  for (int k = 0; k < M; ++k) sub[k] = (k * 13) % N;
  int bad = 0;
  vpm_cpu_set_threads(vpm_cpu_max_threads());
  for (int pass = 0; pass < 2; ++pass) {
    const int64_t m = pass ? M : N;
    const int64_t* t = pass ? sub : all;
    vpm_cpu_selfinduced_velocity(N, x, a, s, m, t, out);
    bad += !finite_all(out, 3 * m);
    vpm_cpu_strength_gradients(N, x, a, s, m, t, out);
    bad += !finite_all(out, 3 * m);
    vpm_cpu_strength_gradients_pse(N, x, a, s, nu, m, t, out);
    bad += !finite_all(out, 3 * m);
    vpm_cpu_cuda_selfinduced_velocity(N, x, a, s, m, t, out);
    bad += !finite_all(out, 3 * m);
    vpm_cpu_cuda_vorticity(N, x, a, s, m, t, out);
    bad += !finite_all(out, 3 * m);
    vpm_cpu_cuda_strength_gradients_pse(N, x, a, s, nu, m, t, out);
    bad += !finite_all(out, 3 * m);
  }
  vpm_cpu_cuda_enstrophy_rows(N, x, a, s, out);
  bad += !finite_all(out, N);
  printf("%s: %d non-finite outputs\n", bad ? "FAILED" : "ok", bad);
  free(x);
  free(a);
  free(s);
  free(nu);
  free(out);
  free(all);
  return bad ? 1 : 0;
}
