// C++ parity test through the reference-shaped interface (rbdgrad.hpp), laid
// out like the reference's GPU harness (time_GPU.cu test(): random knots,
// forwardDynamics with DENSE_MINV, pack COMPLETELY_FUSED, dynamicsGradientKernel,
// copy back).  Checks, on the GPU:
//   1. dynamicsGradientKernel on (q, qd, u, qdd, Minv) equals the full path's
//      dqdd blocks (same knots) to FP32 rounding;
//   2. forwardDynamics' qdd equals the full path's qdd to FP32 rounding (the
//      two are separate kernels, so the compiler may contract different
//      products into FMAs);
//   3. with external link forces, forwardDynamics(traj, fext) and
//      forwardDynamicsGradient(out, in, fext) agree on qdd, and the forces
//      change it;
//   4. the staged pipelines (fused kernel + finish; start + split kernel +
//      finish) give the same dqdd as dynamicsGradientKernel;
//   5. the TEST_FOR_EQUIVALENCE inputs (all 0.1, time_GPU.cu:124-127) match
//      the values the FP64 oracle gives for them (passed on the command line).
// Exit code 0 on success; prints what it checked.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../fast-rbd-gradients_amd/rbdgrad.hpp"

using namespace rbdgrad;

template <int N, bool VD>
int run(std::mt19937& rng) {
  std::uniform_real_distribution<float> uq(-3.14159265f, 3.14159265f), uqd(-2.f, 2.f), uu(-50.f, 50.f);
  auto* t = new traj<float, N>;
  for (int k = 0; k < N; ++k)
    for (int i = 0; i < NUM_POS; ++i) {
      t->knots[k].q[i] = uq(rng);
      t->knots[k].qd[i] = uqd(rng);
      t->knots[k].u[i] = uu(rng);
    }
  rbdErrchk((forwardDynamics<float, N, false, VD>(t)));
  std::vector<float> vol(size_t(COMPLETELY_FUSED_VALS) * N), in(21 * size_t(N));
  packCompletelyFused<float, N>(*t, vol.data());
  for (int k = 0; k < N; ++k)
    for (int i = 0; i < 21; ++i) in[21 * k + i] = vol[COMPLETELY_FUSED_VALS * k + i];
  float *d_vol, *d_dqdd, *d_in, *d_out;
  if (hipMalloc(&d_vol, vol.size() * 4) || hipMalloc(&d_dqdd, size_t(N) * DQDD_VALS * 4) ||
      hipMalloc(&d_in, in.size() * 4) || hipMalloc(&d_out, size_t(N) * 105 * 4)) {
    std::fprintf(stderr, "hipMalloc failed\n");
    return 1;
  }
  (void)hipMemcpy(d_vol, vol.data(), vol.size() * 4, hipMemcpyHostToDevice);
  (void)hipMemcpy(d_in, in.data(), in.size() * 4, hipMemcpyHostToDevice);
  rbdErrchk((dynamicsGradientKernel<float, N, false, VD>(d_dqdd, d_vol, nullptr)));
  rbdErrchk((forwardDynamicsGradient<float, N, false, VD>(d_out, d_in)));
  (void)hipDeviceSynchronize();
  std::vector<float> dqdd(size_t(N) * DQDD_VALS), out(size_t(N) * 105);
  (void)hipMemcpy(dqdd.data(), d_dqdd, dqdd.size() * 4, hipMemcpyDeviceToHost);
  (void)hipMemcpy(out.data(), d_out, out.size() * 4, hipMemcpyDeviceToHost);
  double worst = 0, qdd_worst = 0;
  for (int k = 0; k < N; ++k) {
    for (int b = 0; b < 2; ++b) {
      double d = 0, m = 0;
      for (int i = 0; i < 49; ++i) {
        const double a = dqdd[DQDD_VALS * k + 49 * b + i], r = out[105 * k + 7 + 49 * b + i];
        d = std::fmax(d, std::fabs(a - r));
        m = std::fmax(m, std::fabs(r));
      }
      worst = std::fmax(worst, d / std::fmax(m, 1e-6));
    }
    double d = 0, m = 0;
    for (int i = 0; i < NUM_POS; ++i) {
      d = std::fmax(d, std::fabs(double(t->knots[k].qdd[i]) - out[105 * k + i]));
      m = std::fmax(m, std::fabs(double(out[105 * k + i])));
    }
    qdd_worst = std::fmax(qdd_worst, d / std::fmax(m, 1e-6));
  }
  // staged pipelines: fused kernel -> finish, and start -> split kernel -> finish
  std::vector<float> fused(size_t(FUSED_VALS) * N), split(size_t(SPLIT_VALS) * N, 0.f), minv(size_t(N) * 49);
  for (int k = 0; k < N; ++k) {
    for (int i = 0; i < FUSED_VALS; ++i) fused[FUSED_VALS * k + i] = split[SPLIT_VALS * k + i] = vol[COMPLETELY_FUSED_VALS * k + i];
    for (int i = 0; i < 49; ++i) minv[49 * k + i] = vol[COMPLETELY_FUSED_VALS * k + FUSED_VALS + i];
  }
  float *d_fused, *d_split, *d_cdu, *d_minv, *d_dqdd2;
  if (hipMalloc(&d_fused, fused.size() * 4) || hipMalloc(&d_split, split.size() * 4) ||
      hipMalloc(&d_cdu, size_t(N) * DQDD_VALS * 4) || hipMalloc(&d_minv, minv.size() * 4) ||
      hipMalloc(&d_dqdd2, size_t(N) * DQDD_VALS * 4)) {
    std::fprintf(stderr, "hipMalloc failed\n");
    return 1;
  }
  (void)hipMemcpy(d_fused, fused.data(), fused.size() * 4, hipMemcpyHostToDevice);
  (void)hipMemcpy(d_split, split.data(), split.size() * 4, hipMemcpyHostToDevice);
  (void)hipMemcpy(d_minv, minv.data(), minv.size() * 4, hipMemcpyHostToDevice);
  double staged_worst = 0;
  for (int variant = 0; variant < 2; ++variant) {
    if (variant == 0) {
      rbdErrchk((dynamicsGradientKernel_fused<float, N, false, VD>(d_cdu, d_fused, nullptr)));
    } else {
      rbdErrchk((dynamicsGradientStart<float, N, false, VD>(d_split)));
      rbdErrchk((dynamicsGradientKernel_split<float, N, VD>(d_cdu, d_split, nullptr)));
    }
    rbdErrchk((dynamicsGradientFinish<float, N>(d_dqdd2, d_cdu, d_minv)));
    std::vector<float> got(size_t(N) * DQDD_VALS);
    (void)hipMemcpy(got.data(), d_dqdd2, got.size() * 4, hipMemcpyDeviceToHost);
    for (int k = 0; k < N; ++k)
      for (int b = 0; b < 2; ++b) {
        double d = 0, m = 0;
        for (int i = 0; i < 49; ++i) {
          const double a = got[DQDD_VALS * k + 49 * b + i], r = dqdd[DQDD_VALS * k + 49 * b + i];
          d = std::fmax(d, std::fabs(a - r));
          m = std::fmax(m, std::fabs(r));
        }
        staged_worst = std::fmax(staged_worst, d / std::fmax(m, 1e-6));
      }
  }
  (void)hipFree(d_fused); (void)hipFree(d_split); (void)hipFree(d_cdu); (void)hipFree(d_minv); (void)hipFree(d_dqdd2);
  std::printf("N=%d VEL_DAMPING=%d: reference-mode vs full path max rel err %.3e, forwardDynamics qdd max rel err %.3e, "
              "staged pipelines vs reference mode %.3e\n", N, int(VD), worst, qdd_worst, staged_worst);
  worst = std::fmax(worst, staged_worst);
  (void)hipFree(d_vol); (void)hipFree(d_dqdd); (void)hipFree(d_in); (void)hipFree(d_out);
  delete t;
  return (worst < 1e-4 && qdd_worst < 1e-5) ? 0 : 1;
}

int run_fext(std::mt19937& rng) {
  constexpr int N = 64;
  std::uniform_real_distribution<float> uq(-3.14159265f, 3.14159265f), uqd(-2.f, 2.f), uu(-50.f, 50.f),
      uf(-20.f, 20.f);
  auto* t = new traj<float, N>;
  for (int k = 0; k < N; ++k)
    for (int i = 0; i < NUM_POS; ++i) {
      t->knots[k].q[i] = uq(rng);
      t->knots[k].qd[i] = uqd(rng);
      t->knots[k].u[i] = uu(rng);
    }
  float fext[FEXT_VALS];
  for (float& f : fext) f = uf(rng);
  std::vector<float> in(21 * size_t(N));
  for (int k = 0; k < N; ++k)
    for (int i = 0; i < NUM_POS; ++i) {
      in[21 * k + i] = t->knots[k].q[i];
      in[21 * k + 7 + i] = t->knots[k].qd[i];
      in[21 * k + 14 + i] = t->knots[k].u[i];
    }
  rbdErrchk((forwardDynamics<float, N>(t, fext)));
  float *d_in, *d_out, *d_f;
  if (hipMalloc(&d_in, in.size() * 4) || hipMalloc(&d_out, size_t(N) * 105 * 4) || hipMalloc(&d_f, sizeof(fext))) {
    std::fprintf(stderr, "hipMalloc failed\n");
    return 1;
  }
  (void)hipMemcpy(d_in, in.data(), in.size() * 4, hipMemcpyHostToDevice);
  (void)hipMemcpy(d_f, fext, sizeof(fext), hipMemcpyHostToDevice);
  std::vector<float> with(size_t(N) * 105), without(size_t(N) * 105);
  rbdErrchk((forwardDynamicsGradient<float, N>(d_out, d_in, d_f, 0)));
  (void)hipMemcpy(with.data(), d_out, with.size() * 4, hipMemcpyDeviceToHost);
  rbdErrchk((forwardDynamicsGradient<float, N>(d_out, d_in)));
  (void)hipMemcpy(without.data(), d_out, without.size() * 4, hipMemcpyDeviceToHost);
  double worst = 0, moved = 0;
  for (int k = 0; k < N; ++k) {
    double d = 0, m = 0;
    for (int i = 0; i < NUM_POS; ++i) {
      d = std::fmax(d, std::fabs(double(t->knots[k].qdd[i]) - with[105 * k + i]));
      m = std::fmax(m, std::fabs(double(with[105 * k + i])));
      moved = std::fmax(moved, std::fabs(double(with[105 * k + i]) - without[105 * k + i]));
    }
    worst = std::fmax(worst, d / std::fmax(m, 1e-6));
  }
  std::printf("external forces: forwardDynamics vs forwardDynamicsGradient qdd max rel err %.3e, "
              "change from the forces %.3e\n", worst, moved);
  (void)hipFree(d_in); (void)hipFree(d_out); (void)hipFree(d_f);
  delete t;
  return (worst < 1e-5 && moved > 1e-2) ? 0 : 1;
}

int main(int argc, char** argv) {
  std::mt19937 rng(1234);
  int bad = 0;
  bad += run<16, false>(rng);
  bad += run<128, false>(rng);
  bad += run<128, true>(rng);
  bad += run_fext(rng);
  // TEST_FOR_EQUIVALENCE inputs against expected values given on the command line:
  // argv[1..] = 105 floats (qdd, dqdd/dq, dqdd/dqd) from the FP64 oracle
  if (argc == 106) {
    float in[21], *d_in, *d_out;
    for (float& x : in) x = 0.1f;
    (void)hipMalloc(&d_in, sizeof(in));
    (void)hipMalloc(&d_out, 105 * 4);
    (void)hipMemcpy(d_in, in, sizeof(in), hipMemcpyHostToDevice);
    rbdErrchk((forwardDynamicsGradient<float, 1, false, false>(d_out, d_in)));
    float out[105];
    (void)hipMemcpy(out, d_out, sizeof(out), hipMemcpyDeviceToHost);
    double worst = 0;
    for (int b = 0; b < 3; ++b) {
      const int lo = b == 0 ? 0 : 7 + 49 * (b - 1), hi = b == 0 ? 7 : lo + 49;
      double d = 0, m = 0;
      for (int i = lo; i < hi; ++i) {
        const double r = std::atof(argv[1 + i]);
        d = std::fmax(d, std::fabs(out[i] - r));
        m = std::fmax(m, std::fabs(r));
      }
      worst = std::fmax(worst, d / std::fmax(m, 1e-6));
    }
    std::printf("TEST_FOR_EQUIVALENCE inputs (all 0.1): max rel err vs FP64 oracle %.3e\n", worst);
    bad += worst < 1e-4 ? 0 : 1;
    (void)hipFree(d_in); (void)hipFree(d_out);
  }
  std::printf(bad ? "FAIL\n" : "OK\n");
  return bad ? 1 : 0;
}
