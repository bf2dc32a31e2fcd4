// time_gpu: the reference GPU harness's timing legs over librbdgrad.so, with
// its statistics, printed as one JSON object on stdout.
//
// For each trajectory length N of the reference's harness (1, 10, 16, 32, 64,
// 128, 256; time_GPU.cu:344-351) and each leg, calls are timed one at a time
// with CLOCK_MONOTONIC around the call and its synchronisation, as
// time_GPU.cu:311-330 does:
//   refmode_compute     the reference's timed configuration -- q, qd, u, qdd
//                       and the dense Minv in (77 floats per knot), 98 out --
//                       one launch + hipDeviceSynchronize
//                       ("Completely Fused - GPU Compute", time_GPU.cu:323-330)
//   refmode_compute_io  H2D copy of the 77 floats per knot, the launch, the
//                       synchronisation, D2H copy of the 98
//                       ("Completely Fused - GPU Compute + I/O", :311-320)
//   full_compute        the full path (q, qd, u in; qdd, dqdd/dq, dqdd/dqd
//                       out) on device buffers, launch + synchronisation
//   full_compute_io     the full path from and to host memory
//                       (rbd_fdgrad_traj_host_f32: copies, launch, wait)
// Each leg's times get printStats' statistics (utils/experiment_helpers.h:41-100):
// mean, standard deviation, min, max, the share of calls within each band of
// standard deviations around the mean, the 50/60/70/75/80/85/90/95/99th
// percentiles with the same index rule, and the percentile from which calls
// are 1, 2, 5 and 10% above the mean.  At N = 1 the single-call latency of one
// knot is also measured inside one kernel, as kern_single_full does
// (time_GPU.cu:21-51, :180-186): the computation repeated `iters` times by one
// workgroup, the launch's time divided by `iters`.
//
//   time_gpu [--seconds S] [--max-iters M]      S per leg (default 0.25 s), at most M calls (100000)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "../../include/rbdgrad.h"
#include "../csrc/two_wave.hpp"

namespace {

void die(const char* what, int code) {
  std::fprintf(stderr, "time_gpu: %s failed (%d): %s\n", what, code, rbd_last_error());
  std::exit(1);
}
#define CHECK_RBD(x) do { rbd_status s_ = (x); if (s_ != RBD_OK) die(#x, int(s_)); } while (0)
#define CHECK_HIP(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) die(#x, int(e_)); } while (0)

double now_us() {
  timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return 1e6 * double(t.tv_sec) + 1e-3 * double(t.tv_nsec);
}

// printStats' statistics of one leg as a JSON object
std::string stats_json(std::vector<double> t) {
  const double n = double(t.size());
  const double mean = std::accumulate(t.begin(), t.end(), 0.0) / n;
  double sq = 0;
  for (double x : t) sq += (x - mean) * (x - mean);
  const double sd = std::sqrt(sq / n);
  double hist[7] = {};
  for (double x : t) {
    int b = 3;
    if (x < mean - sd) b = x < mean - 3 * sd ? 0 : (x < mean - 2 * sd ? 1 : 2);
    else if (x > mean + sd) b = x > mean + 3 * sd ? 6 : (x > mean + 2 * sd ? 5 : 4);
    hist[b] += 100.0 / n;
  }
  std::sort(t.begin(), t.end());
  const size_t s = t.size();
  // the reference's index rule: size/2, size/5*3, size/10*7, size/4*3, ...
  const size_t idx[9] = {s / 2, s / 5 * 3, s / 10 * 7, s / 4 * 3, s / 5 * 4, s / 20 * 17, s / 10 * 9, s / 20 * 19,
                         s / 100 * 99};
  const char* names[9] = {"p50", "p60", "p70", "p75", "p80", "p85", "p90", "p95", "p99"};
  char buf[2048];
  int k = std::snprintf(buf, sizeof buf, "{\"calls\": %zu, \"mean\": %.3f, \"std\": %.3f, \"min\": %.3f, \"max\": %.3f",
                        s, mean, sd, t.front(), t.back());
  for (int i = 0; i < 9; ++i) k += std::snprintf(buf + k, sizeof buf - size_t(k), ", \"%s\": %.3f", names[i], t[idx[i]]);
  k += std::snprintf(buf + k, sizeof buf - size_t(k), ", \"sigma_bands_pct\": [%.2f, %.2f, %.2f, %.2f, %.2f, %.2f, %.2f]",
                     hist[0], hist[1], hist[2], hist[3], hist[4], hist[5], hist[6]);
  const double above[4] = {1.01, 1.02, 1.05, 1.10};
  const char* an[4] = {"1", "2", "5", "10"};
  k += std::snprintf(buf + k, sizeof buf - size_t(k), ", \"above_mean_from_percentile\": {");
  for (int a = 0; a < 4; ++a) {
    double pct = -1;
    for (size_t i = 0; i < s; ++i)
      if (t[i] >= mean * above[a]) { pct = 100.0 * double(i) / double(s); break; }
    k += std::snprintf(buf + k, sizeof buf - size_t(k), "%s\"%s\": %.2f", a ? ", " : "", an[a], pct);
  }
  std::snprintf(buf + k, sizeof buf - size_t(k), "}}");
  return buf;
}

template <class F>
std::vector<double> time_calls(F&& call, double seconds, long max_iters) {
  for (int i = 0; i < 20; ++i) call();      // warm
  std::vector<double> t;
  t.reserve(size_t(std::min<long>(max_iters, 200000)));
  const double start = now_us();
  while (long(t.size()) < max_iters) {
    const double a = now_us();
    call();
    t.push_back(now_us() - a);
    if (now_us() - start > seconds * 1e6) break;
  }
  return t;
}

// The single-knot latency kernels: one workgroup repeats one knot's
// computation `iters` times.  The handoff flags are reset between rounds, and
// every round re-reads the inputs and stores the outputs, so no round can be
// folded into another.
__global__ __launch_bounds__(128, 1) void single_full_two_wave(const float* in, float* out, int iters,
                                                              rbd::Options opt) {
  __shared__ rbd::TwoWaveShared<rbd::PrecFP32> sh[rbd::kTwoWaveKnots];
  __shared__ int flags[2];
  const int g = (threadIdx.x % 64) / rbd::kLanesPerKnotGroup;
  for (int it = 0; it < iters; ++it) {
    __syncthreads();
    if (threadIdx.x == 0) { flags[0] = 0; flags[1] = 0; }
    __syncthreads();
    rbd::InAoS i{in};
    rbd::OutMasked<rbd::OutAoS> o{rbd::OutAoS{out}, g == 0};
    rbd::two_wave_group<rbd::PrecFP32>(i, o, opt, sh[g], flags, int(threadIdx.x) / 64);
  }
}

__global__ __launch_bounds__(64, 1) void single_refmode_group(const float* vol, float* dqdd, int iters,
                                                             rbd::Options opt) {
  const int role = threadIdx.x % rbd::kLanesPerKnotGroup;
  const int g = threadIdx.x / rbd::kLanesPerKnotGroup;
  for (int it = 0; it < iters; ++it) {
    rbd::InAoS i{vol};
    rbd::OutMasked<rbd::OutDqddAoS> o{rbd::OutDqddAoS{dqdd}, g == 0};
    rbd::fdgrad_knot_given_qdd_minv_lanes<rbd::PrecFP32>(i, o, opt, role);
    __syncthreads();
  }
}

}  // namespace

int main(int argc, char** argv) {
  double seconds = 0.25;
  long max_iters = 100000;
  for (int a = 1; a + 1 < argc; a += 2) {
    if (!std::strcmp(argv[a], "--seconds")) seconds = std::atof(argv[a + 1]);
    else if (!std::strcmp(argv[a], "--max-iters")) max_iters = std::atol(argv[a + 1]);
  }
  rbd_config cfg;
  rbd_config_default(&cfg);
  rbd_context* ctx = nullptr;
  CHECK_RBD(rbd_create(&cfg, &ctx));
  static const int kN[] = {1, 10, 16, 32, 64, 128, 256};
  std::mt19937 rng(20210101);
  std::uniform_real_distribution<float> uq(-3.14159265f, 3.14159265f), uqd(-2.f, 2.f), uu(-50.f, 50.f);
  std::string legs[4];
  const char* leg_names[4] = {"refmode_compute", "refmode_compute_io", "full_compute", "full_compute_io"};
  std::string single;
  for (int ni = 0; ni < 7; ++ni) {
    const int n = kN[ni];
    std::vector<float> in(size_t(n) * 21), vol(size_t(n) * 77), qdd(size_t(n) * 7), minv(size_t(n) * 49);
    std::vector<float> dqdd(size_t(n) * 98), out(size_t(n) * 105);
    for (int k = 0; k < n; ++k) {
      for (int j = 0; j < 7; ++j) {
        in[21 * k + j] = uq(rng);
        in[21 * k + 7 + j] = uqd(rng);
        in[21 * k + 14 + j] = uu(rng);
      }
    }
    float *d_in, *d_out, *d_vol, *d_dqdd, *d_qdd, *d_minv;
    CHECK_HIP(hipMalloc(&d_in, in.size() * 4));
    CHECK_HIP(hipMalloc(&d_out, out.size() * 4));
    CHECK_HIP(hipMalloc(&d_vol, vol.size() * 4));
    CHECK_HIP(hipMalloc(&d_dqdd, dqdd.size() * 4));
    CHECK_HIP(hipMalloc(&d_qdd, qdd.size() * 4));
    CHECK_HIP(hipMalloc(&d_minv, minv.size() * 4));
    CHECK_HIP(hipMemcpy(d_in, in.data(), in.size() * 4, hipMemcpyHostToDevice));
    // the reference's inputs: qdd and the dense Minv from forward dynamics, packed 77 per knot
    CHECK_RBD(rbd_forward_dynamics_f32(ctx, n, d_in, d_qdd, d_minv, nullptr));
    CHECK_HIP(hipMemcpy(qdd.data(), d_qdd, qdd.size() * 4, hipMemcpyDeviceToHost));
    CHECK_HIP(hipMemcpy(minv.data(), d_minv, minv.size() * 4, hipMemcpyDeviceToHost));
    for (int k = 0; k < n; ++k) {
      std::memcpy(&vol[77 * k], &in[21 * k], 21 * 4);
      std::memcpy(&vol[77 * k + 21], &qdd[7 * k], 7 * 4);
      std::memcpy(&vol[77 * k + 28], &minv[49 * k], 49 * 4);
    }
    CHECK_HIP(hipMemcpy(d_vol, vol.data(), vol.size() * 4, hipMemcpyHostToDevice));
    std::vector<double> t[4];
    t[0] = time_calls([&] {
      CHECK_RBD(rbd_dynamics_gradient_kernel_f32(ctx, n, d_dqdd, d_vol, nullptr, nullptr));
      CHECK_HIP(hipDeviceSynchronize());
    }, seconds, max_iters);
    t[1] = time_calls([&] {
      CHECK_HIP(hipMemcpy(d_vol, vol.data(), vol.size() * 4, hipMemcpyHostToDevice));
      CHECK_RBD(rbd_dynamics_gradient_kernel_f32(ctx, n, d_dqdd, d_vol, nullptr, nullptr));
      CHECK_HIP(hipDeviceSynchronize());
      CHECK_HIP(hipMemcpy(dqdd.data(), d_dqdd, dqdd.size() * 4, hipMemcpyDeviceToHost));
    }, seconds, max_iters);
    t[2] = time_calls([&] {
      CHECK_RBD(rbd_fdgrad_traj_f32(ctx, n, d_in, d_out, nullptr));
      CHECK_HIP(hipDeviceSynchronize());
    }, seconds, max_iters);
    t[3] = time_calls([&] { CHECK_RBD(rbd_fdgrad_traj_host_f32(ctx, n, in.data(), out.data())); }, seconds,
                      max_iters);
    for (int l = 0; l < 4; ++l)
      legs[l] += std::string(ni ? ", " : "") + "\"" + std::to_string(n) + "\": " + stats_json(t[l]);
    if (n == 1) {
      // single-call latency inside one kernel (kern_single_full analogue)
      const int iters = 20000;
      rbd::Options opt;   // gravity on, no damping, no failure word: the default bound
      hipEvent_t e0, e1;
      CHECK_HIP(hipEventCreate(&e0));
      CHECK_HIP(hipEventCreate(&e1));
      float ms_full = 0, ms_ref = 0;
      hipLaunchKernelGGL(single_full_two_wave, dim3(1), dim3(128), 0, nullptr, d_in, d_out, 100, opt);
      CHECK_HIP(hipEventRecord(e0, nullptr));
      hipLaunchKernelGGL(single_full_two_wave, dim3(1), dim3(128), 0, nullptr, d_in, d_out, iters, opt);
      CHECK_HIP(hipEventRecord(e1, nullptr));
      CHECK_HIP(hipEventSynchronize(e1));
      CHECK_HIP(hipEventElapsedTime(&ms_full, e0, e1));
      hipLaunchKernelGGL(single_refmode_group, dim3(1), dim3(64), 0, nullptr, d_vol, d_dqdd, 100, opt);
      CHECK_HIP(hipEventRecord(e0, nullptr));
      hipLaunchKernelGGL(single_refmode_group, dim3(1), dim3(64), 0, nullptr, d_vol, d_dqdd, iters, opt);
      CHECK_HIP(hipEventRecord(e1, nullptr));
      CHECK_HIP(hipEventSynchronize(e1));
      CHECK_HIP(hipEventElapsedTime(&ms_ref, e0, e1));
      CHECK_HIP(hipGetLastError());
      char buf[512];
      std::snprintf(buf, sizeof buf,
                    "{\"iters\": %d, \"full_path_two_wave_us\": %.4f, \"reference_mode_lane_group_us\": %.4f}", iters,
                    1e3 * ms_full / iters, 1e3 * ms_ref / iters);
      single = buf;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
    }
    (void)hipFree(d_in); (void)hipFree(d_out); (void)hipFree(d_vol);
    (void)hipFree(d_dqdd); (void)hipFree(d_qdd); (void)hipFree(d_minv);
  }
  std::printf("{\"unit\": \"us\", \"seconds_per_leg\": %.3f, \"max_calls_per_leg\": %ld", seconds, max_iters);
  for (int l = 0; l < 4; ++l) std::printf(", \"%s\": {%s}", leg_names[l], legs[l].c_str());
  std::printf(", \"single_call_in_kernel\": %s}\n", single.c_str());
  rbd_destroy(ctx);
  return 0;
}
