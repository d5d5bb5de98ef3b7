// Library-level entry points, error reporting and argument validation.

#include "vpm_common.h"

#include <string>

namespace vpm {

namespace {
thread_local char g_last_error[512] = "";
}

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

void clear_error() { g_last_error[0] = '\0'; }

int check_particles(const vpm_particles_t* p, bool need_nu, const char* who) {
  if (!p) { set_error("%s: particle set is null", who); return VPM_ERR_ARG; }
  if (p->n < 0) { set_error("%s: n = %lld < 0", who, (long long)p->n); return VPM_ERR_ARG; }
  if (p->n > (int64_t(1) << 31) - 1) { set_error("%s: n = %lld exceeds 2^31-1", who, (long long)p->n); return VPM_ERR_ARG; }
  if (p->ld < p->n) { set_error("%s: ld = %lld < n = %lld", who, (long long)p->ld, (long long)p->n); return VPM_ERR_ARG; }
  if (p->n > 0) {
    if (!p->pos || !p->str || !p->sigma) { set_error("%s: null pos/str/sigma pointer", who); return VPM_ERR_ARG; }
    if (need_nu && !p->nu) { set_error("%s: viscosity array required", who); return VPM_ERR_ARG; }
  }
  return VPM_OK;
}

int check_points(const vpm_points_t* t, const char* who) {
  if (!t) { set_error("%s: target set is null", who); return VPM_ERR_ARG; }
  if (t->m < 0 || t->ld < t->m) { set_error("%s: bad target count/stride (m=%lld ld=%lld)", who, (long long)t->m, (long long)t->ld); return VPM_ERR_ARG; }
  if (t->m > (int64_t(1) << 31) - 1) { set_error("%s: m exceeds 2^31-1", who); return VPM_ERR_ARG; }
  if (t->m > 0 && !t->pos) { set_error("%s: null target pointer", who); return VPM_ERR_ARG; }
  return VPM_OK;
}

int check_out(const double* out, int64_t ld, int64_t n, const char* who) {
  if (n > 0 && !out) { set_error("%s: null output pointer", who); return VPM_ERR_ARG; }
  if (ld < n) { set_error("%s: output ld = %lld < %lld", who, (long long)ld, (long long)n); return VPM_ERR_ARG; }
  return VPM_OK;
}

__global__ __launch_bounds__(1024) void final_sum_kernel(const double* __restrict__ part,
                                                         int64_t n, double scale,
                                                         double* __restrict__ out) {
  __shared__ double red[1024 / 64];
  double s = 0.0;
  for (int64_t k = threadIdx.x; k < n; k += 1024) s += part[k];
  s = block_sum<1024>(s, red);
  if (threadIdx.x == 0) out[0] = s * scale;
}

}  // namespace vpm

extern "C" {

int vpm_abi_version(void) { return VPM_ABI_VERSION; }

const char* vpm_last_error(void) { return vpm::g_last_error; }

int vpm_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

}  // extern "C"
