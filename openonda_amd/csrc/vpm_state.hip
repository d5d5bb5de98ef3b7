// Device-resident time-integration state: the O(N) parts of a solver step
// (position RK stages, the step-indexed strength update, core spreading, the
// dynamic LES filter, relaxation) and the conserved-quantity reductions, all
// FP64, deterministic and free of host synchronisation so that a whole step
// can be captured in a HIP graph (see vpm_graph_*).
//
// Reference: OpenONDA/solvers/VPM/vpmModule.py -- RK position updates
// :563-692, strength update :749-845, core spreading :851-862, relaxation
// :867-883, LES filter :887-962, impulses :264-284.

#include <algorithm>

#include "vpm_common.h"

namespace vpm {

namespace {

constexpr int kB = 256;

inline unsigned grid_for(int64_t n) { return static_cast<unsigned>(ceil_div(n > 0 ? n : 1, kB)); }

// x, y and out may alias (in-place updates).
__global__ __launch_bounds__(kB) void axpy3_kernel(int n, int64_t ld, const double* x,
                                                   const double* y, double c, double* out) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int k = 0; k < 3; ++k) out[k * ld + i] = x[k * ld + i] + c * y[k * ld + i];
}

// out = x0 + dt/6 (k1 + 2 k2 + 2 k3 + k4)   (vpmModule.py:688); out may be x0.
__global__ __launch_bounds__(kB) void rk4_combine_kernel(int n, int64_t ld, const double* x0,
                                                         const double* k1, const double* k2,
                                                         const double* k3, const double* k4,
                                                         double dt, double* out) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int64_t o = k * ld + i;
    out[o] = x0[o] + (dt / 6) * (k1[o] + 2 * k2[o] + 2 * k3[o] + k4[o]);
  }
}

// Step-indexed strength update and history push, vpmModule.py:749-791
// (identical for PSE, :797-845).  `s` is updated in place.
__global__ __launch_bounds__(kB) void bdf_kernel(int n, int64_t ld, int step, double dt,
                                                 double* __restrict__ s, double* __restrict__ p1,
                                                 double* __restrict__ p2, double* __restrict__ p3,
                                                 const double* __restrict__ rhs) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int64_t o = k * ld + i;
    const double sv = s[o], a1 = p1[o], a2 = p2[o], a3 = p3[o], r = rhs[o];
    double v;
    if (step == 0)
      v = sv + dt * r;
    else if (step == 1)
      v = (4 * sv - a1) / 3 + (2.0 / 3.0) * dt * r;
    else if (step == 2)
      v = (18 * sv - 9 * a1 + 2 * a2) / 11 + (6.0 / 11.0) * dt * r;
    else
      v = (48 * sv - 36 * a1 + 16 * a2 - 3 * a3) / 25 + (12.0 / 25.0) * dt * r;
    if (step >= 3) p3[o] = a2;
    if (step >= 2) p2[o] = a1;
    p1[o] = v;
    s[o] = v;
  }
}

// sigma <- sqrt(sigma^2 + 5.34 (nu + nu_t) dt), vpmModule.py:855-857.
__global__ __launch_bounds__(kB) void core_spreading_kernel(int n, double* __restrict__ sigma,
                                                            const double* __restrict__ nu,
                                                            const double* __restrict__ nu_t,
                                                            double dt) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  const double nue = nu[i] + (nu_t ? nu_t[i] : 0.0);
  sigma[i] = sqrt(sigma[i] * sigma[i] + 5.34 * nue * dt);
}

// alpha <- 0.95 alpha + 0.05 (omega/|omega|) |alpha|, vpmModule.py:868-879.
__global__ __launch_bounds__(kB) void relax_kernel(int n, int64_t ld, double* __restrict__ a,
                                                   const double* __restrict__ w) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  const double wx = w[i], wy = w[ld + i], wz = w[2 * ld + i];
  const double ax = a[i], ay = a[ld + i], az = a[2 * ld + i];
  const double wn = sqrt(wx * wx + wy * wy + wz * wz);
  const double an = sqrt(ax * ax + ay * ay + az * az);
  a[i] = ax * 0.95 + (1 - 0.95) * (wx / wn) * an;
  a[ld + i] = ay * 0.95 + (1 - 0.95) * (wy / wn) * an;
  a[2 * ld + i] = az * 0.95 + (1 - 0.95) * (wz / wn) * an;
}

// ---- deterministic column sums -------------------------------------------
// Each block writes NC partial sums; final_cols sums every column in fixed
// order with one block.
template <int NC>
__device__ void write_block_partials(double (&v)[NC], double* __restrict__ part, double* lds) {
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const double s = block_sum<kB>(v[c], lds);
    if (threadIdx.x == 0) part[blockIdx.x * NC + c] = s;
  }
}

__global__ __launch_bounds__(256) void final_cols_kernel(const double* __restrict__ part, int nb,
                                                         int nc, double* __restrict__ out,
                                                         double scale) {
  __shared__ double red[256 / 64];
  for (int c = 0; c < nc; ++c) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) s += part[b * nc + c];
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) out[c] = s * scale;
  }
}

// Linear impulse (3), angular impulse (3), total strength (3), sum |alpha| (1).
__global__ __launch_bounds__(kB) void impulse_partial_kernel(int n, int64_t ld,
                                                             const double* __restrict__ x,
                                                             const double* __restrict__ a,
                                                             double* __restrict__ part) {
  __shared__ double red[kB / 64];
  double v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
    const double px = x[i], py = x[ld + i], pz = x[2 * ld + i];
    const double ax = a[i], ay = a[ld + i], az = a[2 * ld + i];
    const double cx = py * az - pz * ay, cy = pz * ax - px * az, cz = px * ay - py * ax;
    v[0] += 0.5 * cx;
    v[1] += 0.5 * cy;
    v[2] += 0.5 * cz;
    v[3] += (py * cz - pz * cy) / 3.0;
    v[4] += (pz * cx - px * cz) / 3.0;
    v[5] += (px * cy - py * cx) / 3.0;
    v[6] += ax;
    v[7] += ay;
    v[8] += az;
    v[9] += sqrt(ax * ax + ay * ay + az * az);
  }
  write_block_partials<10>(v, part, red);
}

// ---- dynamic LES filter ---------------------------------------------------
// state[0] = C_nu, state[1] = enstrophy_ref, state[2] = 1 once enstrophy_ref exists.
__global__ void les_constant_kernel(const double* __restrict__ enstrophy,
                                    double* __restrict__ state) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double e = sqrt(enstrophy[0]);                       // vpmModule.py:894
  double ref = state[2] != 0.0 ? state[1] : e;                // :897-898
  ref = 0.1 * e + (1 - 0.1) * ref;                            // :901
  double c = state[0] * e / (ref + 1e-9);                     // :904
  c = fmax(1e-9, fmin(c, 0.0761));                            // :907
  state[0] = c;
  state[1] = ref;
  state[2] = 1.0;
}

__global__ __launch_bounds__(kB) void vort_sum_partial_kernel(int n, int64_t ld,
                                                              const double* __restrict__ w,
                                                              double* __restrict__ part) {
  __shared__ double red[kB / 64];
  double v[3] = {0, 0, 0};
  for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
    v[0] += w[i];
    v[1] += w[ld + i];
    v[2] += w[2 * ld + i];
  }
  write_block_partials<3>(v, part, red);
}

// nu_t candidate per particle (vpmModule.py:932-951) and per-block maxima
// (grid-stride: one partial per block).
__global__ __launch_bounds__(kB) void les_nut_kernel(int n, int64_t ld,
                                                     const double* __restrict__ w,
                                                     const double* __restrict__ radii,
                                                     const double* __restrict__ wsum,
                                                     const double* __restrict__ state,
                                                     double* __restrict__ cand,
                                                     double* __restrict__ maxpart) {
  __shared__ double red[kB / 64];
  const double beta0 = kPi / 4.0;
  const double mx = wsum[0] / n, my = wsum[1] / n, mz = wsum[2] / n;     // :939
  const double mm = sqrt(mx * mx + my * my + mz * mz);                   // :940
  const double ux = mx / mm, uy = my / mm, uz = mz / mm;                 // :941
  double vmax = 0.0;
  for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
    const double wx = w[i], wy = w[ld + i], wz = w[2 * ld + i];
    const double wm = sqrt(wx * wx + wy * wy + wz * wz);                 // :935
    const double c = (wx / wm) * ux + (wy / wm) * uy + (wz / wm) * uz;   // :936,944
    const double beta = acos(fmin(fmax(c, -1.0), 1.0));                  // :945
    const double psi = (beta0 <= beta && beta <= kPi - beta0) ? 1.0 : 0.0;  // :948 (NaN -> 0)
    const double cr = state[0] * radii[i];
    const double v = psi * (cr * cr) * wm;                               // :951
    cand[i] = v;
    vmax = fmax(vmax, v);
  }
  vmax = wave_max(vmax);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = vmax;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int k = 1; k < kB / 64; ++k) m = fmax(m, red[k]);
    maxpart[blockIdx.x] = m;
  }
}

__global__ __launch_bounds__(256) void max_final_kernel(const double* __restrict__ part, int nb,
                                                        double* __restrict__ out) {
  __shared__ double red[4];
  double m = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) m = fmax(m, part[b]);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// nu_t is written only if max(nu_t) > 1e-9 (vpmModule.py:954-960).
__global__ __launch_bounds__(kB) void les_commit_kernel(int n, const double* __restrict__ cand,
                                                        const double* __restrict__ vmax,
                                                        double* __restrict__ nu_t) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i < n && vmax[0] > 1e-9) nu_t[i] = cand[i];
}

// Particle-set scan: ORs VPM_BAD_* bits into *flags for non-finite
// positions or strengths, radii that are not finite and > 0, viscosities
// that are not finite and >= 0.  One flag write per wave that saw a problem.
__global__ __launch_bounds__(kB) void scan_kernel(int n, int64_t ld, const double* __restrict__ pos,
                                                  const double* __restrict__ str,
                                                  const double* __restrict__ sigma,
                                                  const double* __restrict__ nu,
                                                  unsigned* __restrict__ flags) {
  unsigned bad = 0u;
  for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (!isfinite(pos[k * ld + i])) bad |= VPM_BAD_POSITION;
      if (!isfinite(str[k * ld + i])) bad |= VPM_BAD_STRENGTH;
    }
    const double s = sigma[i];
    if (!(isfinite(s) && s > 0.0)) bad |= VPM_BAD_RADIUS;
    if (nu) {
      const double v = nu[i];
      if (!(isfinite(v) && v >= 0.0)) bad |= VPM_BAD_VISCOSITY;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) bad |= __shfl_xor(bad, off, 64);
  if ((threadIdx.x & 63) == 0 && bad) atomicOr(flags, bad);
}

inline int blocks_reduce(int64_t n) {
  return static_cast<int>(std::min<int64_t>(1024, ceil_div(n > 0 ? n : 1, kB)));
}

}  // namespace

}  // namespace vpm

using namespace vpm;

extern "C" {

int vpm_axpy3(int64_t n, int64_t ld, const double* x, const double* y, double c, double* out,
              void* stream) {
  clear_error();
  if (n < 0 || ld < n || n > INT32_MAX) return fail(VPM_ERR_ARG, "vpm_axpy3: bad n/ld");
  if (n == 0) return VPM_OK;
  if (!x || !y || !out) return fail(VPM_ERR_ARG, "vpm_axpy3: null pointer");
  hipLaunchKernelGGL(axpy3_kernel, dim3(grid_for(n)), dim3(kB), 0, as_stream(stream),
                     static_cast<int>(n), ld, x, y, c, out);
  VPM_LAUNCH_CHECK("axpy3_kernel");
  return VPM_OK;
}

int vpm_rk4_combine(int64_t n, int64_t ld, const double* x0, const double* k1, const double* k2,
                    const double* k3, const double* k4, double dt, double* out, void* stream) {
  clear_error();
  if (n < 0 || ld < n || n > INT32_MAX) return fail(VPM_ERR_ARG, "vpm_rk4_combine: bad n/ld");
  if (n == 0) return VPM_OK;
  if (!x0 || !k1 || !k2 || !k3 || !k4 || !out) return fail(VPM_ERR_ARG, "vpm_rk4_combine: null pointer");
  hipLaunchKernelGGL(rk4_combine_kernel, dim3(grid_for(n)), dim3(kB), 0, as_stream(stream),
                     static_cast<int>(n), ld, x0, k1, k2, k3, k4, dt, out);
  VPM_LAUNCH_CHECK("rk4_combine_kernel");
  return VPM_OK;
}

int vpm_strength_update(int64_t n, int64_t ld, int step, double dt, double* strengths,
                        double* prev1, double* prev2, double* prev3, const double* rhs,
                        void* stream) {
  clear_error();
  if (n < 0 || ld < n || n > INT32_MAX || step < 0)
    return fail(VPM_ERR_ARG, "vpm_strength_update: bad n/ld/step");
  if (n == 0) return VPM_OK;
  if (!strengths || !prev1 || !prev2 || !prev3 || !rhs)
    return fail(VPM_ERR_ARG, "vpm_strength_update: null pointer");
  hipLaunchKernelGGL(bdf_kernel, dim3(grid_for(n)), dim3(kB), 0, as_stream(stream),
                     static_cast<int>(n), ld, step, dt, strengths, prev1, prev2, prev3, rhs);
  VPM_LAUNCH_CHECK("bdf_kernel");
  return VPM_OK;
}

int vpm_core_spreading(int64_t n, double* sigma, const double* nu, const double* nu_t, double dt,
                       void* stream) {
  clear_error();
  if (n < 0 || n > INT32_MAX) return fail(VPM_ERR_ARG, "vpm_core_spreading: bad n");
  if (n == 0) return VPM_OK;
  if (!sigma || !nu) return fail(VPM_ERR_ARG, "vpm_core_spreading: null pointer");
  hipLaunchKernelGGL(core_spreading_kernel, dim3(grid_for(n)), dim3(kB), 0, as_stream(stream),
                     static_cast<int>(n), sigma, nu, nu_t, dt);
  VPM_LAUNCH_CHECK("core_spreading_kernel");
  return VPM_OK;
}

int vpm_relax_strengths(int64_t n, int64_t ld, double* strengths, const double* vorticity,
                        void* stream) {
  clear_error();
  if (n < 0 || ld < n || n > INT32_MAX) return fail(VPM_ERR_ARG, "vpm_relax_strengths: bad n/ld");
  if (n == 0) return VPM_OK;
  if (!strengths || !vorticity) return fail(VPM_ERR_ARG, "vpm_relax_strengths: null pointer");
  hipLaunchKernelGGL(relax_kernel, dim3(grid_for(n)), dim3(kB), 0, as_stream(stream),
                     static_cast<int>(n), ld, strengths, vorticity);
  VPM_LAUNCH_CHECK("relax_kernel");
  return VPM_OK;
}

int vpm_scan_particles(const vpm_particles_t* p, uint32_t* flags, void* stream) {
  clear_error();
  int rc = check_particles(p, false, "vpm_scan_particles");
  if (rc) return rc;
  if (!flags) return fail(VPM_ERR_ARG, "vpm_scan_particles: null flags");
  if (p->n == 0) return VPM_OK;
  hipLaunchKernelGGL(scan_kernel, dim3(blocks_reduce(p->n)), dim3(kB), 0, as_stream(stream),
                     static_cast<int>(p->n), p->ld, p->pos, p->str, p->sigma, p->nu,
                     reinterpret_cast<unsigned*>(flags));
  VPM_LAUNCH_CHECK("scan_kernel");
  return VPM_OK;
}

size_t vpm_state_workspace_bytes(int64_t n) {
  const size_t nb = static_cast<size_t>(blocks_reduce(n));
  const size_t nn = static_cast<size_t>(n > 0 ? n : 1);
  return (nb * 10 + 8 + nb + nn) * sizeof(double);
}

int vpm_diagnostics(int64_t n, int64_t ld, const double* pos, const double* str, double* out10,
                    void* workspace, size_t ws_bytes, void* stream) {
  clear_error();
  if (n < 0 || ld < n || n > INT32_MAX) return fail(VPM_ERR_ARG, "vpm_diagnostics: bad n/ld");
  if (!out10) return fail(VPM_ERR_ARG, "vpm_diagnostics: null output");
  if (n > 0 && (!pos || !str)) return fail(VPM_ERR_ARG, "vpm_diagnostics: null pointer");
  if (!workspace || ws_bytes < vpm_state_workspace_bytes(n))
    return fail(VPM_ERR_WORKSPACE, "vpm_diagnostics: workspace too small");
  const int nb = blocks_reduce(n);
  double* part = static_cast<double*>(workspace);
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(impulse_partial_kernel, dim3(nb), dim3(kB), 0, st, static_cast<int>(n), ld,
                     pos, str, part);
  VPM_LAUNCH_CHECK("impulse_partial_kernel");
  hipLaunchKernelGGL(final_cols_kernel, dim3(1), dim3(256), 0, st, part, nb, 10, out10, 1.0);
  VPM_LAUNCH_CHECK("final_cols_kernel");
  return VPM_OK;
}

int vpm_les_update(int64_t n, int64_t ld, const double* vorticity, const double* radii,
                   const double* enstrophy, double* state3, double* nu_t, int update_constant,
                   void* workspace, size_t ws_bytes, void* stream) {
  clear_error();
  if (n < 0 || ld < n || n > INT32_MAX) return fail(VPM_ERR_ARG, "vpm_les_update: bad n/ld");
  if (!state3 || (update_constant && !enstrophy))
    return fail(VPM_ERR_ARG, "vpm_les_update: null scalar pointer");
  if (n > 0 && (!vorticity || !radii || !nu_t)) return fail(VPM_ERR_ARG, "vpm_les_update: null pointer");
  if (!workspace || ws_bytes < vpm_state_workspace_bytes(n))
    return fail(VPM_ERR_WORKSPACE, "vpm_les_update: workspace too small");
  hipStream_t st = as_stream(stream);
  if (update_constant) {
    hipLaunchKernelGGL(les_constant_kernel, dim3(1), dim3(64), 0, st, enstrophy, state3);
    VPM_LAUNCH_CHECK("les_constant_kernel");
  }
  if (n == 0) return VPM_OK;
  const int nb = blocks_reduce(n);
  double* part = static_cast<double*>(workspace);      // nb*10
  double* wsum = part + nb * 10;                        // 3 (+pad to 8)
  double* vmax = wsum + 4;                              // 1
  double* maxpart = wsum + 8;                           // nb
  double* cand = maxpart + nb;                          // n
  hipLaunchKernelGGL(vort_sum_partial_kernel, dim3(nb), dim3(kB), 0, st, static_cast<int>(n), ld,
                     vorticity, part);
  VPM_LAUNCH_CHECK("vort_sum_partial_kernel");
  hipLaunchKernelGGL(final_cols_kernel, dim3(1), dim3(256), 0, st, part, nb, 3, wsum, 1.0);
  VPM_LAUNCH_CHECK("final_cols_kernel");
  hipLaunchKernelGGL(les_nut_kernel, dim3(nb), dim3(kB), 0, st, static_cast<int>(n), ld, vorticity,
                     radii, wsum, state3, cand, maxpart);
  VPM_LAUNCH_CHECK("les_nut_kernel");
  hipLaunchKernelGGL(max_final_kernel, dim3(1), dim3(256), 0, st, maxpart, nb, vmax);
  VPM_LAUNCH_CHECK("max_final_kernel");
  hipLaunchKernelGGL(les_commit_kernel, dim3(grid_for(n)), dim3(kB), 0, st, static_cast<int>(n),
                     cand, vmax, nu_t);
  VPM_LAUNCH_CHECK("les_commit_kernel");
  (void)maxpart;
  return VPM_OK;
}

// ---- HIP graph capture of a sequence of calls on one stream -----------------
int vpm_graph_begin(void* stream) {
  clear_error();
  if (!stream) return fail(VPM_ERR_ARG, "vpm_graph_begin: capture needs a non-default stream");
  VPM_HIP_TRY(hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeThreadLocal));
  return VPM_OK;
}

int vpm_graph_end(void* stream, void** graph_exec) {
  clear_error();
  if (!stream || !graph_exec) return fail(VPM_ERR_ARG, "vpm_graph_end: null argument");
  hipGraph_t g = nullptr;
  VPM_HIP_TRY(hipStreamEndCapture(as_stream(stream), &g));
  hipGraphExec_t ex = nullptr;
  hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    set_error("hipGraphInstantiate failed: %s", hipGetErrorString(e));
    return VPM_ERR_LAUNCH;
  }
  *graph_exec = ex;
  return VPM_OK;
}

int vpm_graph_launch(void* graph_exec, void* stream) {
  clear_error();
  if (!graph_exec) return fail(VPM_ERR_ARG, "vpm_graph_launch: null graph");
  VPM_HIP_TRY(hipGraphLaunch(static_cast<hipGraphExec_t>(graph_exec), as_stream(stream)));
  return VPM_OK;
}

int vpm_graph_destroy(void* graph_exec) {
  clear_error();
  if (graph_exec) VPM_HIP_TRY(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec)));
  return VPM_OK;
}

}  // extern "C"
