// Plain FP64 drop-in kernels: one per function of the reference physics
// module, with the reference arithmetic restated literally (IEEE division,
// pow with fractional exponents, the same masks) and an MI355X memory design:
// SoA sources staged through LDS in tiles of BLOCK and broadcast to every lane
// of the wavefront, one target per lane, deterministic two-pass reductions.
//
// Reference: OpenONDA/solvers/VPM/ParticlesPhysicsCuda.py (variant CUDA) and
// OpenONDA/solvers/VPM/ParticlesPhysics.py (variant CPU).

#include "vpm_common.h"

namespace vpm {

namespace {

constexpr int kPlainBlock = 256;

struct PlainArgs {
  int64_t m;            // number of targets
  int64_t n;            // number of sources
  const double* tpos;   // [3][tld] target positions
  int64_t tld;
  const double* pos;    // [3][ld] sources
  const double* str;    // [3][ld]
  int64_t ld;
  const double* sig;    // [n]
  const double* nu;     // [n] or null
  double* out;          // [3][ld_out] or null
  int64_t ld_out;
  double* partial;      // per-block partial sums (reductions)
};

struct Target {
  double x, y, z;       // position
  double ax, ay, az;    // strength (self evaluations)
  double s, nu;         // radius, viscosity (self evaluations)
  int64_t i;            // particle index, -1 for external targets
};

struct Src {
  double x, y, z, ax, ay, az, s, nu;
};

// ---------------------------------------------------------------------------
// Pair operators.  Each has Acc (per-target state), pair() and a result.
// ---------------------------------------------------------------------------

// ParticlesPhysicsCuda.py:65-97 / ParticlesPhysics.py:46-74.
struct OpVelAt {
  struct Acc { double ux = 0, uy = 0, uz = 0; };
  static __device__ inline void pair(Acc& a, const Target& t, const Src& s, int64_t) {
    const double rx = t.x - s.x, ry = t.y - s.y, rz = t.z - s.z;
    const double r2 = rx * rx + ry * ry + rz * rz;
    if (r2 > kEpsPair) {
      const double sigma = s.s;
      const double sigma_sq = sigma * sigma;
      const double rho2 = r2 / sigma_sq;
      const double q = (rho2 + 2.5) / (pow(rho2 + 1.0, 2.5) * sigma_sq * sigma);
      const double cx = ry * s.az - rz * s.ay;
      const double cy = rz * s.ax - rx * s.az;
      const double cz = rx * s.ay - ry * s.ax;
      a.ux -= q * cx;
      a.uy -= q * cy;
      a.uz -= q * cz;
    }
  }
  static constexpr bool kVector = true;
  static __device__ inline void result(const Acc& a, const Target&, double* v) {
    v[0] = a.ux * kOneOverFourPi;
    v[1] = a.uy * kOneOverFourPi;
    v[2] = a.uz * kOneOverFourPi;
  }
};

// CUDA: ParticlesPhysicsCuda.py:159-191 (sigma_ij).  CPU: ParticlesPhysics.py:109-143
// (sigma_i).
template <bool CPUV>
struct OpVelSelf {
  struct Acc { double ux = 0, uy = 0, uz = 0; };
  static __device__ inline void pair(Acc& a, const Target& t, const Src& s, int64_t j) {
    if (j == t.i) return;
    const double rx = t.x - s.x, ry = t.y - s.y, rz = t.z - s.z;
    const double r2 = rx * rx + ry * ry + rz * rz;
    const double sigma = CPUV ? t.s : (t.s + s.s) / 2;
    const double sigma_sq = sigma * sigma;
    const double rho_sq = r2 / sigma_sq;
    const double q = (rho_sq + 2.5) / (pow(rho_sq + 1.0, 2.5) * sigma * sigma_sq);
    const double cx = ry * s.az - rz * s.ay;
    const double cy = rz * s.ax - rx * s.az;
    const double cz = rx * s.ay - ry * s.ax;
    a.ux -= q * cx;
    a.uy -= q * cy;
    a.uz -= q * cz;
  }
  static constexpr bool kVector = true;
  static __device__ inline void result(const Acc& a, const Target&, double* v) {
    v[0] = kOneOverFourPi * a.ux;
    v[1] = kOneOverFourPi * a.uy;
    v[2] = kOneOverFourPi * a.uz;
  }
};

// CUDA: simplified zeta form, ParticlesPhysicsCuda.py:319-346.
// CPU : full curl form, ParticlesPhysics.py:236-262.
template <bool CPUV>
struct OpVortAt {
  struct Acc { double wx = 0, wy = 0, wz = 0; };
  static __device__ inline void pair(Acc& a, const Target& t, const Src& s, int64_t) {
    const double rx = t.x - s.x, ry = t.y - s.y, rz = t.z - s.z;
    const double r2 = rx * rx + ry * ry + rz * rz;
    if (!(r2 > kEpsPair)) return;
    const double sigma = s.s;
    if (!CPUV) {
      const double rho_sq = r2 / (sigma * sigma);
      const double sigma_cb = sigma * sigma * sigma;
      const double xi = 7.5 / (pow(rho_sq + 1.0, 3.5) * sigma_cb);
      a.wx += xi * s.ax;
      a.wy += xi * s.ay;
      a.wz += xi * s.az;
    } else {
      const double rho2 = r2 / (sigma * sigma);
      const double s3 = sigma * sigma * sigma;
      const double xi = 7.5 / (pow(rho2 + 1.0, 3.5) * s3);
      const double q = (rho2 + 2.5) / (pow(rho2 + 1.0, 2.5) * s3);
      const double term1 = xi - q;
      const double term2 = (3.0 * q - xi) * (rx * s.ax + ry * s.ay + rz * s.az) / r2;
      a.wx += term1 * s.ax + term2 * rx;
      a.wy += term1 * s.ay + term2 * ry;
      a.wz += term1 * s.az + term2 * rz;
    }
  }
  static constexpr bool kVector = true;
  static __device__ inline void result(const Acc& a, const Target&, double* v) {
    v[0] = a.wx * kOneOverFourPi;
    v[1] = a.wy * kOneOverFourPi;
    v[2] = a.wz * kOneOverFourPi;
  }
};

// ParticlesPhysicsCuda.py:396-445 / ParticlesPhysics.py:291-312.
struct OpStretch {
  struct Acc { double sx = 0, sy = 0, sz = 0; };
  static __device__ inline void pair(Acc& a, const Target& t, const Src& s, int64_t j) {
    if (j == t.i) return;
    const double rx = t.x - s.x, ry = t.y - s.y, rz = t.z - s.z;
    const double r2 = rx * rx + ry * ry + rz * rz;
    const double sigma = (t.s + s.s) * 0.5;
    const double sigma_sq = sigma * sigma;
    const double sigma_cb = sigma_sq * sigma;
    const double A = r2 / sigma_sq + 1.0;
    const double f1 = (A + 1.5) / (sigma_cb * pow(A, 2.5));
    const double f2 = 3.0 * (A + 2.5) / (sigma_cb * sigma_sq * pow(A, 3.5));
    const double p1x = t.ay * s.az - t.az * s.ay;
    const double p1y = t.az * s.ax - t.ax * s.az;
    const double p1z = t.ax * s.ay - t.ay * s.ax;
    const double p2x = ry * s.az - rz * s.ay;
    const double p2y = rz * s.ax - rx * s.az;
    const double p2z = rx * s.ay - ry * s.ax;
    const double p3 = t.ax * p2x + t.ay * p2y + t.az * p2z;
    a.sx += (f1 * p1x + f2 * p3 * rx) * kOneOverFourPi;
    a.sy += (f1 * p1y + f2 * p3 * ry) * kOneOverFourPi;
    a.sz += (f1 * p1z + f2 * p3 * rz) * kOneOverFourPi;
  }
  static constexpr bool kVector = true;
  static __device__ inline void result(const Acc& a, const Target&, double* v) {
    v[0] = a.sx;
    v[1] = a.sy;
    v[2] = a.sz;
  }
};

// CUDA: ParticlesPhysicsCuda.py:500-530 (nu_i).  CPU: ParticlesPhysics.py:353-382
// (nu_ij, A + 1e-8, 1/(4pi) once at the end).
template <bool CPUV>
struct OpPse {
  struct Acc { double dx = 0, dy = 0, dz = 0; };
  static __device__ inline void pair(Acc& a, const Target& t, const Src& s, int64_t j) {
    if (j == t.i) return;
    const double rx = t.x - s.x, ry = t.y - s.y, rz = t.z - s.z;
    const double d2 = rx * rx + ry * ry + rz * rz;
    const double sigma = (t.s + s.s) * 0.5;
    const double sigma_sq = sigma * sigma;
    const double sigma_cb = sigma_sq * sigma;
    if (!CPUV) {
      const double sigma_qu = sigma_sq * sigma_cb;
      const double A = d2 / sigma_sq + 1.0;
      const double f3 = 105.0 * t.nu / (sigma_qu * pow(A, 4.5));
      a.dx += f3 * sigma_cb * (s.ax - t.ax) * kOneOverFourPi;
      a.dy += f3 * sigma_cb * (s.ay - t.ay) * kOneOverFourPi;
      a.dz += f3 * sigma_cb * (s.az - t.az) * kOneOverFourPi;
    } else {
      const double vis = (t.nu + s.nu) * 0.5;
      const double A = (d2 / sigma_sq) + 1.0 + kEpsPair;
      const double f3 = 105.0 * vis / (sigma_sq * sigma_cb * pow(A, 4.5));
      a.dx += f3 * sigma_cb * (s.ax - t.ax);
      a.dy += f3 * sigma_cb * (s.ay - t.ay);
      a.dz += f3 * sigma_cb * (s.az - t.az);
    }
  }
  static constexpr bool kVector = true;
  static __device__ inline void result(const Acc& a, const Target&, double* v) {
    const double k = CPUV ? kOneOverFourPi : 1.0;
    v[0] = a.dx * k;
    v[1] = a.dy * k;
    v[2] = a.dz * k;
  }
};

// Kinetic energy pair term: ParticlesPhysicsCuda.py:585-631 (CUDA) and
// ParticlesPhysics.py:404-427 (CPU, A + 1e-8).  Scale 1/(16 pi) on the total.
template <bool CPUV>
struct OpEnergy {
  struct Acc { double e = 0; };
  static __device__ inline void pair(Acc& a, const Target& t, const Src& s, int64_t j) {
    if (j == t.i) return;
    const double rx = t.x - s.x, ry = t.y - s.y, rz = t.z - s.z;
    const double rn = sqrt(rx * rx + ry * ry + rz * rz);
    const double ux = rx / rn, uy = ry / rn, uz = rz / rn;
    const double sigma = (t.s + s.s) * 0.5;
    const double rho = rn / sigma;
    const double A = rho * rho + 1.0 + (CPUV ? kEpsPair : 0.0);
    const double term_1 = 2. * rho / pow(A, 0.5);
    const double term_2 = rho * rho * rho / pow(A, 1.5);
    const double p1 = t.ax * s.ax + t.ay * s.ay + t.az * s.az;
    const double p2 = ux * t.ax + uy * t.ay + uz * t.az;
    const double p3 = ux * s.ax + uy * s.ay + uz * s.az;
    a.e += (term_1 * p1 + term_2 * (p2 * p3 - p1)) / rn;
  }
  static constexpr bool kVector = false;
  static __device__ inline double value(const Acc& a) { return a.e; }
};

// Helicity pair term: ParticlesPhysicsCuda.py:707-739 (CUDA, rho as a vector)
// and ParticlesPhysics.py:448-466 (CPU, |r|/sigma).
template <bool CPUV>
struct OpHelicity {
  struct Acc { double h = 0; };
  static __device__ inline void pair(Acc& a, const Target& t, const Src& s, int64_t j) {
    if (j == t.i) return;
    const double rx = t.x - s.x, ry = t.y - s.y, rz = t.z - s.z;
    const double sigma = 0.5 * (t.s + s.s);
    double rho_sq;
    if (!CPUV) {
      const double px = rx / sigma, py = ry / sigma, pz = rz / sigma;
      rho_sq = px * px + py * py + pz * pz;
    } else {
      const double rho = sqrt(rx * rx + ry * ry + rz * rz) / sigma;
      rho_sq = rho * rho;
    }
    const double term = (rho_sq + 2.5) / (sigma * sigma * sigma * pow(rho_sq + 1.0, 2.5));
    const double p1x = t.ay * s.az - t.az * s.ay;
    const double p1y = t.az * s.ax - t.ax * s.az;
    const double p1z = t.ax * s.ay - t.ay * s.ax;
    const double p2 = rx * p1x + ry * p1y + rz * p1z;
    a.h += p2 * term;
  }
  static constexpr bool kVector = false;
  static __device__ inline double value(const Acc& a) { return a.h; }
};

// Enstrophy pair term (i = j included): ParticlesPhysicsCuda.py:886-909 and
// ParticlesPhysics.py:533-548.
struct OpEnstrophy {
  struct Acc { double e = 0; };
  static __device__ inline void pair(Acc& a, const Target& t, const Src& s, int64_t) {
    const double rx = t.x - s.x, ry = t.y - s.y, rz = t.z - s.z;
    const double r2 = rx * rx + ry * ry + rz * rz;
    const double sigma = (t.s + s.s) * 0.5;
    const double sigma_sq = sigma * sigma;
    const double term_1 = (r2 + 3.2 * sigma_sq) / pow(r2 + sigma_sq, 1.5);
    const double p1 = t.ax * s.ax + t.ay * s.ay + t.az * s.az;
    a.e += term_1 * p1;
  }
  static constexpr bool kVector = false;
  static __device__ inline double value(const Acc& a) { return a.e; }
};

// ---------------------------------------------------------------------------
// Generic LDS-tiled driver: one target per lane, BLOCK sources per tile.
// ---------------------------------------------------------------------------
template <class Op, bool SELF, bool NEED_NU>
__global__ __launch_bounds__(kPlainBlock) void plain_kernel(PlainArgs a) {
  __shared__ double s_x[kPlainBlock], s_y[kPlainBlock], s_z[kPlainBlock];
  __shared__ double s_ax[kPlainBlock], s_ay[kPlainBlock], s_az[kPlainBlock];
  __shared__ double s_s[kPlainBlock], s_nu[kPlainBlock];
  __shared__ double s_red[kPlainBlock / 64];

  const int64_t i = static_cast<int64_t>(blockIdx.x) * kPlainBlock + threadIdx.x;
  const bool active = i < a.m;

  Target t{};
  t.i = -1;
  if (active) {
    if (SELF) {
      t.x = a.pos[i];
      t.y = a.pos[a.ld + i];
      t.z = a.pos[2 * a.ld + i];
      t.ax = a.str[i];
      t.ay = a.str[a.ld + i];
      t.az = a.str[2 * a.ld + i];
      t.s = a.sig[i];
      t.nu = NEED_NU ? a.nu[i] : 0.0;
      t.i = i;
    } else {
      t.x = a.tpos[i];
      t.y = a.tpos[a.tld + i];
      t.z = a.tpos[2 * a.tld + i];
    }
  }

  typename Op::Acc acc{};
  for (int64_t t0 = 0; t0 < a.n; t0 += kPlainBlock) {
    const int64_t j = t0 + threadIdx.x;
    if (j < a.n) {  // coalesced SoA loads into LDS
      s_x[threadIdx.x] = a.pos[j];
      s_y[threadIdx.x] = a.pos[a.ld + j];
      s_z[threadIdx.x] = a.pos[2 * a.ld + j];
      s_ax[threadIdx.x] = a.str[j];
      s_ay[threadIdx.x] = a.str[a.ld + j];
      s_az[threadIdx.x] = a.str[2 * a.ld + j];
      s_s[threadIdx.x] = a.sig[j];
      s_nu[threadIdx.x] = NEED_NU ? a.nu[j] : 0.0;
    }
    __syncthreads();
    const int cnt = static_cast<int>(a.n - t0 < kPlainBlock ? a.n - t0 : kPlainBlock);
    if (active) {
      for (int k = 0; k < cnt; ++k) {  // every lane reads the same LDS word
        Src s{s_x[k], s_y[k], s_z[k], s_ax[k], s_ay[k], s_az[k], s_s[k], s_nu[k]};
        Op::pair(acc, t, s, t0 + k);
      }
    }
    __syncthreads();
  }

  if constexpr (Op::kVector) {
    if (active) {
      double v[3];
      Op::result(acc, t, v);
      a.out[i] = v[0];
      a.out[a.ld_out + i] = v[1];
      a.out[2 * a.ld_out + i] = v[2];
    }
  } else {
    const double v = active ? Op::value(acc) : 0.0;
    const double bs = block_sum<kPlainBlock>(v, s_red);
    if (threadIdx.x == 0) a.partial[blockIdx.x] = bs;
  }
}

template <class Op, bool SELF, bool NEED_NU>
int launch_plain(const PlainArgs& a, hipStream_t st, const char* name) {
  if (a.m == 0) return VPM_OK;
  const int64_t blocks = ceil_div(a.m, kPlainBlock);
  if (blocks > 0x7fffffff) return fail(VPM_ERR_ARG, "too many targets");
  hipLaunchKernelGGL((plain_kernel<Op, SELF, NEED_NU>), dim3(static_cast<unsigned>(blocks)),
                     dim3(kPlainBlock), 0, st, a);
  VPM_LAUNCH_CHECK(name);
  return VPM_OK;
}

PlainArgs args_from(const vpm_particles_t* p) {
  PlainArgs a{};
  a.m = p->n;
  a.n = p->n;
  a.pos = p->pos;
  a.str = p->str;
  a.ld = p->ld;
  a.sig = p->sigma;
  a.nu = p->nu;
  return a;
}

template <class OpCuda, class OpCpu>
int reduce_pair(const vpm_particles_t* p, int variant, double scale_cuda, double scale_cpu,
                double* result, void* ws, size_t ws_bytes, void* stream, const char* who) {
  int rc = check_particles(p, false, who);
  if (rc) return rc;
  if (!result) return fail(VPM_ERR_ARG, "result pointer is null");
  if (variant != VPM_VARIANT_CUDA && variant != VPM_VARIANT_CPU)
    return fail(VPM_ERR_ARG, "unknown variant");
  const size_t need = vpm_reduction_workspace_bytes(p->n);
  if (ws_bytes < need || (need && !ws)) return fail(VPM_ERR_WORKSPACE, "reduction workspace too small");
  PlainArgs a = args_from(p);
  a.partial = static_cast<double*>(ws);
  hipStream_t st = as_stream(stream);
  const int64_t blocks = ceil_div(p->n, kPlainBlock);
  if (variant == VPM_VARIANT_CUDA)
    rc = launch_plain<OpCuda, true, false>(a, st, who);
  else
    rc = launch_plain<OpCpu, true, false>(a, st, who);
  if (rc) return rc;
  hipLaunchKernelGGL(final_sum_kernel, dim3(1), dim3(1024), 0, st, a.partial, blocks,
                     variant == VPM_VARIANT_CUDA ? scale_cuda : scale_cpu, result);
  VPM_LAUNCH_CHECK("final_sum_kernel");
  return VPM_OK;
}

}  // namespace

}  // namespace vpm

using namespace vpm;

extern "C" {

size_t vpm_reduction_workspace_bytes(int64_t n) {
  if (n <= 0) return sizeof(double);
  return static_cast<size_t>(ceil_div(n, kPlainBlock)) * sizeof(double);
}

int vpm_get_induced_velocity_at(const vpm_points_t* targets, const vpm_particles_t* src,
                                double* out, int64_t ld_out, void* stream) {
  clear_error();
  int rc = check_points(targets, "vpm_get_induced_velocity_at");
  if (rc) return rc;
  if ((rc = check_particles(src, false, "vpm_get_induced_velocity_at"))) return rc;
  if ((rc = check_out(out, ld_out, targets->m, "vpm_get_induced_velocity_at"))) return rc;
  PlainArgs a = args_from(src);
  a.m = targets->m;
  a.tpos = targets->pos;
  a.tld = targets->ld;
  a.out = out;
  a.ld_out = ld_out;
  return launch_plain<OpVelAt, false, false>(a, as_stream(stream), "vel_at");
}

int vpm_get_selfinduced_velocity(const vpm_particles_t* p, int variant, double* out,
                                 int64_t ld_out, void* stream) {
  clear_error();
  int rc = check_particles(p, false, "vpm_get_selfinduced_velocity");
  if (rc) return rc;
  if ((rc = check_out(out, ld_out, p->n, "vpm_get_selfinduced_velocity"))) return rc;
  PlainArgs a = args_from(p);
  a.out = out;
  a.ld_out = ld_out;
  if (variant == VPM_VARIANT_CUDA)
    return launch_plain<OpVelSelf<false>, true, false>(a, as_stream(stream), "vel_self");
  if (variant == VPM_VARIANT_CPU)
    return launch_plain<OpVelSelf<true>, true, false>(a, as_stream(stream), "vel_self_cpu");
  return fail(VPM_ERR_ARG, "unknown variant");
}

int vpm_get_vorticity_field_at(const vpm_points_t* targets, const vpm_particles_t* src,
                               int variant, double* out, int64_t ld_out, void* stream) {
  clear_error();
  int rc = check_points(targets, "vpm_get_vorticity_field_at");
  if (rc) return rc;
  if ((rc = check_particles(src, false, "vpm_get_vorticity_field_at"))) return rc;
  if ((rc = check_out(out, ld_out, targets->m, "vpm_get_vorticity_field_at"))) return rc;
  PlainArgs a = args_from(src);
  a.m = targets->m;
  a.tpos = targets->pos;
  a.tld = targets->ld;
  a.out = out;
  a.ld_out = ld_out;
  if (variant == VPM_VARIANT_CUDA)
    return launch_plain<OpVortAt<false>, false, false>(a, as_stream(stream), "vort_at");
  if (variant == VPM_VARIANT_CPU)
    return launch_plain<OpVortAt<true>, false, false>(a, as_stream(stream), "vort_at_cpu");
  return fail(VPM_ERR_ARG, "unknown variant");
}

int vpm_get_strength_gradients(const vpm_particles_t* p, double* out, int64_t ld_out,
                               void* stream) {
  clear_error();
  int rc = check_particles(p, false, "vpm_get_strength_gradients");
  if (rc) return rc;
  if ((rc = check_out(out, ld_out, p->n, "vpm_get_strength_gradients"))) return rc;
  PlainArgs a = args_from(p);
  a.out = out;
  a.ld_out = ld_out;
  return launch_plain<OpStretch, true, false>(a, as_stream(stream), "stretch");
}

int vpm_get_strength_gradients_pse(const vpm_particles_t* p, int variant, double* out,
                                   int64_t ld_out, void* stream) {
  clear_error();
  int rc = check_particles(p, true, "vpm_get_strength_gradients_pse");
  if (rc) return rc;
  if ((rc = check_out(out, ld_out, p->n, "vpm_get_strength_gradients_pse"))) return rc;
  PlainArgs a = args_from(p);
  a.out = out;
  a.ld_out = ld_out;
  if (variant == VPM_VARIANT_CUDA)
    return launch_plain<OpPse<false>, true, true>(a, as_stream(stream), "pse");
  if (variant == VPM_VARIANT_CPU)
    return launch_plain<OpPse<true>, true, true>(a, as_stream(stream), "pse_cpu");
  return fail(VPM_ERR_ARG, "unknown variant");
}

int vpm_get_total_kinetic_energy(const vpm_particles_t* p, int variant, double* result,
                                 void* workspace, size_t ws_bytes, void* stream) {
  clear_error();
  const double k = 1.0 / (16.0 * kPi);
  return reduce_pair<OpEnergy<false>, OpEnergy<true>>(p, variant, k, k, result, workspace,
                                                      ws_bytes, stream, "vpm_get_total_kinetic_energy");
}

int vpm_get_total_helicity(const vpm_particles_t* p, int variant, double* result,
                           void* workspace, size_t ws_bytes, void* stream) {
  clear_error();
  return reduce_pair<OpHelicity<false>, OpHelicity<true>>(
      p, variant, kOneOverFourPi * kOneOverFourPi, kOneOverFourPi, result, workspace, ws_bytes,
      stream, "vpm_get_total_helicity");
}

int vpm_get_total_enstrophy(const vpm_particles_t* p, int variant, double* result,
                            void* workspace, size_t ws_bytes, void* stream) {
  clear_error();
  const double k8 = 1.0 / (8.0 * kPi);
  return reduce_pair<OpEnstrophy, OpEnstrophy>(p, variant, k8 * kOneOverFourPi, k8, result,
                                               workspace, ws_bytes, stream,
                                               "vpm_get_total_enstrophy");
}

}  // extern "C"
