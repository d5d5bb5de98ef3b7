// Host-side checks of libvpm_hip's C ABI, built with AddressSanitizer and
// UndefinedBehaviorSanitizer on the host code only (device code is not
// compiled: --cuda-host-only), run on a machine without a GPU.  Every entry
// point is called with malformed arguments (null pointers, negative sizes,
// strides below n, unknown options, workspaces too small) and must return
// its documented status without touching memory it does not own; calls that
// get past validation on a machine without a device must fail with a status
// too.  Prints one line per check and "ok N" at the end; exits non-zero on a
// wrong status.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "vpm_hip.h"

static int fails = 0, checks = 0;

static void expect(const char* what, int rc, int want) {
  ++checks;
  if (rc != want) {
    ++fails;
    std::printf("FAIL %s: status %d, expected %d (%s)\n", what, rc, want, vpm_last_error());
  } else {
    std::printf("ok   %s: %d %s\n", what, rc, rc ? vpm_last_error() : "");
  }
}

static void expect_error(const char* what, int rc) {
  ++checks;
  if (rc >= 0) {
    ++fails;
    std::printf("FAIL %s: status %d, expected an error\n", what, rc);
  } else {
    std::printf("ok   %s: %d %s\n", what, rc, vpm_last_error());
  }
}

int main() {
  expect("abi version", vpm_abi_version(), VPM_ABI_VERSION);
  std::vector<double> host(3 * 8, 1.0);  // host memory: never handed to a launch below
  const double* h = host.data();
  vpm_particles_t good{8, 8, h, h, h, h};
  vpm_particles_t neg = good;
  neg.n = -1;
  vpm_particles_t big = good;
  big.n = int64_t(1) << 31;
  big.ld = big.n;
  vpm_particles_t ld = good;
  ld.ld = 4;
  vpm_particles_t nul = good;
  nul.pos = nullptr;
  vpm_particles_t nonu = good;
  nonu.nu = nullptr;
  vpm_points_t tg{8, 8, h};
  vpm_points_t tneg{-2, 8, h};
  double out[24];

  expect("selfinduced null set", vpm_get_selfinduced_velocity(nullptr, 0, out, 8, nullptr), VPM_ERR_ARG);
  expect("selfinduced n<0", vpm_get_selfinduced_velocity(&neg, 0, out, 8, nullptr), VPM_ERR_ARG);
  expect("selfinduced n>=2^31", vpm_get_selfinduced_velocity(&big, 0, out, big.n, nullptr), VPM_ERR_ARG);
  expect("selfinduced ld<n", vpm_get_selfinduced_velocity(&ld, 0, out, 8, nullptr), VPM_ERR_ARG);
  expect("selfinduced null pos", vpm_get_selfinduced_velocity(&nul, 0, out, 8, nullptr), VPM_ERR_ARG);
  expect("selfinduced bad variant", vpm_get_selfinduced_velocity(&good, 7, out, 8, nullptr), VPM_ERR_ARG);
  expect("selfinduced out ld<n", vpm_get_selfinduced_velocity(&good, 0, out, 4, nullptr), VPM_ERR_ARG);
  expect("selfinduced null out", vpm_get_selfinduced_velocity(&good, 0, nullptr, 8, nullptr), VPM_ERR_ARG);
  expect("induced_at null targets", vpm_get_induced_velocity_at(nullptr, &good, out, 8, nullptr), VPM_ERR_ARG);
  expect("induced_at m<0", vpm_get_induced_velocity_at(&tneg, &good, out, 8, nullptr), VPM_ERR_ARG);
  expect("vorticity_at bad variant", vpm_get_vorticity_field_at(&tg, &good, 3, out, 8, nullptr), VPM_ERR_ARG);
  expect("stretching ld<n", vpm_get_strength_gradients(&ld, out, 8, nullptr), VPM_ERR_ARG);
  expect("pse without nu", vpm_get_strength_gradients_pse(&nonu, 0, out, 8, nullptr), VPM_ERR_ARG);
  double res = 0.0;
  expect("energy small workspace",
         vpm_get_total_kinetic_energy(&good, 0, &res, out, 4, nullptr), VPM_ERR_WORKSPACE);
  expect("enstrophy null result",
         vpm_get_total_enstrophy(&good, 0, nullptr, out, sizeof(out), nullptr), VPM_ERR_ARG);

  vpm_forward_opts_t o{VPM_OUT_VEL | VPM_OUT_STRETCH, VPM_PREC_F32C, VPM_SIGMA_UNIFORM, 0, -1, 0,
                       VPM_G64_AUTO};
  vpm_forward_opts_t bad = o;
  vpm_forward_out_t fo{8, out, out, nullptr, nullptr, nullptr};
  expect("forward null opts", vpm_forward(&good, nullptr, &fo, out, sizeof(out), nullptr), VPM_ERR_ARG);
  bad.outputs = 0;
  expect("forward no outputs", vpm_forward(&good, &bad, &fo, out, sizeof(out), nullptr), VPM_ERR_ARG);
  bad = o;
  bad.outputs = 64;
  expect("forward unknown output", vpm_forward(&good, &bad, &fo, out, sizeof(out), nullptr), VPM_ERR_ARG);
  bad = o;
  bad.precision = 9;
  expect("forward bad precision", vpm_forward(&good, &bad, &fo, out, sizeof(out), nullptr), VPM_ERR_ARG);
  bad = o;
  bad.tile_config = 1000;
  expect("forward bad tile config", vpm_forward(&good, &bad, &fo, out, sizeof(out), nullptr), VPM_ERR_ARG);
  bad = o;
  bad.g64_mode = 5;
  expect("forward bad g64 mode", vpm_forward(&good, &bad, &fo, out, sizeof(out), nullptr), VPM_ERR_ARG);
  bad = o;
  bad.max_splits = -1;
  expect("forward negative max_splits", vpm_forward(&good, &bad, &fo, out, sizeof(out), nullptr), VPM_ERR_ARG);
  bad = o;
  bad.outputs = VPM_OUT_PSE;
  expect("forward pse without nu", vpm_forward(&nonu, &bad, &fo, out, sizeof(out), nullptr), VPM_ERR_ARG);
  // the workspace-size query needs the device's sort library: a status or 0, never a crash
  (void)vpm_forward_workspace_bytes(8, &o);
  expect_error("forward workspace too small / no device",
               vpm_forward(&good, &o, &fo, out, 16, nullptr));
  vpm_forward_out_t fo_ld{4, out, out, nullptr, nullptr, nullptr};
  expect("forward_run out ld<n", vpm_forward_run(&good, &o, &fo_ld, out, sizeof(out), nullptr), VPM_ERR_ARG);
  vpm_forward_out_t fo_nov{8, nullptr, out, nullptr, nullptr, nullptr};
  expect("forward_run vel missing", vpm_forward_run(&good, &o, &fo_nov, out, sizeof(out), nullptr), VPM_ERR_ARG);
  expect("forward_run null out", vpm_forward_run(&good, &o, nullptr, out, sizeof(out), nullptr), VPM_ERR_ARG);
  expect("noalloc null out", vpm_forward_noalloc(&good, &o, nullptr, nullptr), VPM_ERR_ARG);
  expect("noalloc ld<n", vpm_forward_noalloc(&ld, &o, &fo, nullptr), VPM_ERR_ARG);
  if (vpm_forward_noalloc_capacity() != (size_t(256) << 20)) {
    ++fails;
    std::printf("FAIL noalloc capacity\n");
  }
  int b = 0, r = 0, t = 0, s = 0, c = 0;
  expect("tile config info out of range", vpm_tile_config_info(-1, &b, &r, &t, &s, &c), VPM_ERR_ARG);
  expect("tile config info 0", vpm_tile_config_info(0, &b, &r, &t, &s, &c), VPM_OK);
  for (int i = 0; i < vpm_tile_config_count(); ++i) {
    if (vpm_tile_config_info(i, &b, &r, &t, &s, &c) != VPM_OK || b <= 0 || r <= 0 || t % s != 0) {
      ++fails;
      std::printf("FAIL tile config %d\n", i);
    }
  }
  int best = -1;
  float ms[32];
  expect("autotune reps<1", vpm_autotune(&good, &o, &fo, out, sizeof(out), nullptr, 0, &best, ms), VPM_ERR_ARG);
  expect("autotune null best", vpm_autotune(&good, &o, &fo, out, sizeof(out), nullptr, 1, nullptr, ms), VPM_ERR_ARG);

  vpm_batch_desc_t d{};
  d.nsets = -1;
  expect("batch negative sets", vpm_batch_prepare(&good, &d, 1, out, sizeof(out), nullptr), VPM_ERR_ARG);
  d.nsets = 2;
  expect("batch null arrays", vpm_batch_prepare(&good, &d, 1, out, sizeof(out), nullptr), VPM_ERR_ARG);
  d.nsets = 0;
  vpm_forward_opts_t ob = o;
  ob.sigma_mode = VPM_SIGMA_GENERAL;
  expect("batch general sigma", vpm_batch_run(&good, &d, &ob, &fo, out, sizeof(out), nullptr), VPM_ERR_UNSUPPORTED);
  expect("batch null opts", vpm_batch_run(&good, &d, nullptr, &fo, out, sizeof(out), nullptr), VPM_ERR_ARG);
  expect("batch les bad ld", vpm_batch_les_update(&d, 2, h, h, h, out, h, out, out, out, sizeof(out), 8, nullptr), VPM_ERR_ARG);

  vpm_adjoint_in_t ai{8, h, nullptr, nullptr};
  vpm_adjoint_out_t ao{8, out, out};
  vpm_adjoint_in_t ai_ld{2, h, nullptr, nullptr};
  expect("adjoint null in", vpm_adjoint(&good, nullptr, &o, &ao, out, sizeof(out), nullptr), VPM_ERR_ARG);
  expect("adjoint in ld<n", vpm_adjoint(&good, &ai_ld, &o, &ao, out, sizeof(out), nullptr), VPM_ERR_ARG);
  expect("velocity_at adjoint null lam",
         vpm_induced_velocity_at_adjoint(&tg, &good, nullptr, 8, out, out, 8, nullptr), VPM_ERR_ARG);

  expect("axpy3 n<0", vpm_axpy3(-1, 8, h, h, 1.0, out, nullptr), VPM_ERR_ARG);
  expect("axpy3 ld<n", vpm_axpy3(8, 4, h, h, 1.0, out, nullptr), VPM_ERR_ARG);
  expect("axpy3 null", vpm_axpy3(8, 8, nullptr, h, 1.0, out, nullptr), VPM_ERR_ARG);
  expect("axpy3 empty", vpm_axpy3(0, 0, nullptr, nullptr, 1.0, nullptr, nullptr), VPM_OK);
  expect("rk4 null", vpm_rk4_combine(8, 8, h, h, h, nullptr, h, 0.1, out, nullptr), VPM_ERR_ARG);
  expect("strength update step<0", vpm_strength_update(8, 8, -1, 0.1, out, out, out, out, h, nullptr), VPM_ERR_ARG);
  expect("core spreading null", vpm_core_spreading(8, nullptr, h, nullptr, 0.1, nullptr), VPM_ERR_ARG);
  expect("relax null", vpm_relax_strengths(8, 8, out, nullptr, nullptr), VPM_ERR_ARG);
  expect("scan null flags", vpm_scan_particles(&good, nullptr, nullptr), VPM_ERR_ARG);
  expect("scan n<0", vpm_scan_particles(&neg, nullptr, nullptr), VPM_ERR_ARG);
  expect("diagnostics small workspace", vpm_diagnostics(8, 8, h, h, out, out, 8, nullptr), VPM_ERR_WORKSPACE);
  expect("diagnostics null out", vpm_diagnostics(8, 8, h, h, nullptr, out, sizeof(out), nullptr), VPM_ERR_ARG);
  expect("les small workspace", vpm_les_update(8, 8, h, h, h, out, out, 1, out, 8, nullptr), VPM_ERR_WORKSPACE);
  expect_error("graph end without begin", vpm_graph_end(nullptr, nullptr));
  expect_error("graph launch null", vpm_graph_launch(nullptr, nullptr));

  std::printf("%s %d checks, %d failed\n", fails ? "FAILED" : "ok", checks, fails);
  return fails ? 1 : 0;
}
