// librbdgrad.so: HIP kernels for gfx950 and the C ABI declared in include/rbdgrad.h.
//
// Mapping: one lane per knot point.  Each lane runs the whole per-knot solver
// (fdgrad_knot.hpp) out of registers: no LDS, no barriers, no cross-lane
// traffic; the only memory traffic is the knot's 84 input bytes and its 420
// output bytes.  A wave therefore processes 64 knots in lock step with uniform
// control flow (the derivative-column loop runs over the same column index in
// every lane).
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>

#include "../../include/rbdgrad.h"
#include "fdgrad_knot.hpp"

namespace {

thread_local std::string g_last_error;

rbd_status fail(rbd_status s, const std::string& msg) {
  g_last_error = msg;
  return s;
}
rbd_status hip_fail(hipError_t e, const char* what) {
  return fail(RBD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

constexpr int kBlock = 64;   // one wave per workgroup

}  // namespace

struct rbd_context {
  int device = 0;
  unsigned flags = 0;
  int precision = RBD_PREC_FP32;
  rbd::Options opt;
  // device buffer of the host-pointer convenience call
  std::mutex host_mu;
  float* d_scratch = nullptr;
  size_t scratch_bytes = 0;
};

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
namespace rbd {

template <class P>
__global__ __launch_bounds__(kBlock) void fdgrad_aos_kernel(long n, const float* __restrict__ in,
                                                            float* __restrict__ out, Options opt) {
  const long stride = long(gridDim.x) * blockDim.x;
  for (long k = long(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += stride) {
    InAoS i{in + kInPerKnot * k};
    OutAoS o{out + kOutPerKnot * k};
    fdgrad_knot<P>(i, o, opt);
  }
}

template <class P>
__global__ __launch_bounds__(kBlock) void fdgrad_soa_kernel(long n, const float* __restrict__ in, long ld_in,
                                                            float* __restrict__ out, long ld_out, Options opt) {
  const long stride = long(gridDim.x) * blockDim.x;
  for (long k = long(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += stride) {
    InSoA i{in + k, ld_in};
    OutSoA o{out + k, ld_out};
    fdgrad_knot<P>(i, o, opt);
  }
}

template <class P>
__global__ __launch_bounds__(kBlock) void fdgrad_refmode_kernel(long n, const float* __restrict__ vol,
                                                                float* __restrict__ dqdd, Options opt) {
  const long stride = long(gridDim.x) * blockDim.x;
  for (long k = long(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += stride) {
    InAoS i{vol + kVolPerKnot * k};
    OutDqddAoS o{dqdd + kDqddPerKnot * k};
    fdgrad_knot_given_qdd_minv<P>(i, o, opt);
  }
}

}  // namespace rbd

namespace {

// grid for n knots: one lane per knot, capped so huge batches grid-stride
unsigned grid_for(int64_t n) {
  const int64_t blocks = (n + kBlock - 1) / kBlock;
  const int64_t cap = int64_t(1) << 20;
  return unsigned(blocks < cap ? blocks : cap);
}

template <template <class> class Launch, class... A>
hipError_t by_precision(int prec, A&&... a) {
  switch (prec) {
    case RBD_PREC_FP32: return Launch<rbd::PrecFP32>::go(a...);
    case RBD_PREC_FP64_PIVOT: return Launch<rbd::PrecFP64Pivot>::go(a...);
    case RBD_PREC_FP64_COLUMNS: return Launch<rbd::PrecFP64Col>::go(a...);
    default: return Launch<rbd::PrecFP64>::go(a...);
  }
}

template <class P>
struct LaunchAoS {
  static hipError_t go(int64_t n, const float* in, float* out, const rbd::Options& opt, hipStream_t s) {
    hipLaunchKernelGGL(rbd::fdgrad_aos_kernel<P>, dim3(grid_for(n)), dim3(kBlock), 0, s, long(n), in, out, opt);
    return hipGetLastError();
  }
};
template <class P>
struct LaunchSoA {
  static hipError_t go(int64_t n, const float* in, int64_t ld_in, float* out, int64_t ld_out,
                       const rbd::Options& opt, hipStream_t s) {
    hipLaunchKernelGGL(rbd::fdgrad_soa_kernel<P>, dim3(grid_for(n)), dim3(kBlock), 0, s, long(n), in,
                       long(ld_in), out, long(ld_out), opt);
    return hipGetLastError();
  }
};
template <class P>
struct LaunchRef {
  static hipError_t go(int64_t n, const float* vol, float* dqdd, const rbd::Options& opt, hipStream_t s) {
    hipLaunchKernelGGL(rbd::fdgrad_refmode_kernel<P>, dim3(grid_for(n)), dim3(kBlock), 0, s, long(n), vol, dqdd,
                       opt);
    return hipGetLastError();
  }
};

rbd_status check_ctx(rbd_context* ctx) {
  if (!ctx) return fail(RBD_ERR_INVALID_ARGUMENT, "null context");
  return RBD_OK;
}

// the context's device must be current for a launch on the caller's thread
rbd_status use_device(rbd_context* ctx) {
  int cur = -1;
  hipError_t e = hipGetDevice(&cur);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (cur != ctx->device) {
    e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  }
  return RBD_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int rbd_abi_version(void) { return 1; }

void rbd_config_default(rbd_config* cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->device = 0;
  cfg->flags = 0;
  cfg->precision = RBD_PREC_FP32;
}

const char* rbd_status_string(rbd_status s) {
  switch (s) {
    case RBD_OK: return "ok";
    case RBD_ERR_INVALID_ARGUMENT: return "invalid argument";
    case RBD_ERR_HIP: return "HIP runtime error";
    case RBD_ERR_NO_DEVICE: return "no HIP device";
    case RBD_ERR_OUT_OF_MEMORY: return "out of device memory";
    case RBD_ERR_UNSUPPORTED: return "unsupported";
    case RBD_ERR_CAPACITY: return "capacity exceeded";
  }
  return "unknown status";
}

const char* rbd_last_error(void) { return g_last_error.c_str(); }

rbd_status rbd_create(const rbd_config* cfg, rbd_context** out) {
  if (!out) return fail(RBD_ERR_INVALID_ARGUMENT, "null output pointer");
  *out = nullptr;
  rbd_config c;
  if (cfg) c = *cfg; else rbd_config_default(&c);
  if (c.precision < RBD_PREC_FP32 || c.precision > RBD_PREC_FP64)
    return fail(RBD_ERR_INVALID_ARGUMENT, "unknown precision mode");
  if (c.flags & ~(RBD_FLAG_VEL_DAMPING | RBD_FLAG_NO_GRAVITY))
    return fail(RBD_ERR_INVALID_ARGUMENT, "unknown flag bits");
  for (int r : c.reserved)
    if (r != 0) return fail(RBD_ERR_INVALID_ARGUMENT, "reserved config fields must be zero");
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) return fail(RBD_ERR_NO_DEVICE, "no HIP device visible");
  if (c.device < 0 || c.device >= ndev) return fail(RBD_ERR_INVALID_ARGUMENT, "device ordinal out of range");
  rbd_context* ctx = new (std::nothrow) rbd_context();
  if (!ctx) return fail(RBD_ERR_OUT_OF_MEMORY, "host allocation failed");
  ctx->device = c.device;
  ctx->flags = c.flags;
  ctx->precision = c.precision;
  ctx->opt.damping = (c.flags & RBD_FLAG_VEL_DAMPING) ? 0.5f : 0.0f;
  ctx->opt.gravity = (c.flags & RBD_FLAG_NO_GRAVITY) ? 0.0f : float(rbd::kGravity);
  *out = ctx;
  return RBD_OK;
}

rbd_status rbd_destroy(rbd_context* ctx) {
  if (!ctx) return RBD_OK;
  if (ctx->d_scratch) {
    int cur = -1;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(ctx->device);
    (void)hipFree(ctx->d_scratch);
    if (cur >= 0) (void)hipSetDevice(cur);
  }
  delete ctx;
  return RBD_OK;
}

rbd_status rbd_fdgrad_traj_f32(rbd_context* ctx, int64_t n, const float* d_in, float* d_out, rbd_stream_t stream) {
  if (rbd_status s = check_ctx(ctx)) return s;
  if (n < 0) return fail(RBD_ERR_INVALID_ARGUMENT, "negative knot count");
  if (n == 0) return RBD_OK;
  if (!d_in || !d_out) return fail(RBD_ERR_INVALID_ARGUMENT, "null buffer");
  if (rbd_status s = use_device(ctx)) return s;
  hipError_t e = by_precision<LaunchAoS>(ctx->precision, n, d_in, d_out, ctx->opt, (hipStream_t)stream);
  return e == hipSuccess ? RBD_OK : hip_fail(e, "fdgrad_aos_kernel launch");
}

rbd_status rbd_fdgrad_traj_soa_f32(rbd_context* ctx, int64_t n, const float* d_in, int64_t ld_in, float* d_out,
                                   int64_t ld_out, rbd_stream_t stream) {
  if (rbd_status s = check_ctx(ctx)) return s;
  if (n < 0) return fail(RBD_ERR_INVALID_ARGUMENT, "negative knot count");
  if (n == 0) return RBD_OK;
  if (!d_in || !d_out) return fail(RBD_ERR_INVALID_ARGUMENT, "null buffer");
  if (ld_in < n || ld_out < n) return fail(RBD_ERR_INVALID_ARGUMENT, "leading dimension smaller than knot count");
  if (rbd_status s = use_device(ctx)) return s;
  hipError_t e = by_precision<LaunchSoA>(ctx->precision, n, d_in, ld_in, d_out, ld_out, ctx->opt,
                                         (hipStream_t)stream);
  return e == hipSuccess ? RBD_OK : hip_fail(e, "fdgrad_soa_kernel launch");
}

rbd_status rbd_dynamics_gradient_kernel_f32(rbd_context* ctx, int64_t n, float* d_dqdd, const float* d_mem_vol,
                                            const float* d_mem_const, rbd_stream_t stream) {
  (void)d_mem_const;  // the constants are compiled into the kernel
  if (rbd_status s = check_ctx(ctx)) return s;
  if (n < 0) return fail(RBD_ERR_INVALID_ARGUMENT, "negative knot count");
  if (n == 0) return RBD_OK;
  if (!d_dqdd || !d_mem_vol) return fail(RBD_ERR_INVALID_ARGUMENT, "null buffer");
  if (rbd_status s = use_device(ctx)) return s;
  hipError_t e = by_precision<LaunchRef>(ctx->precision, n, d_mem_vol, d_dqdd, ctx->opt, (hipStream_t)stream);
  return e == hipSuccess ? RBD_OK : hip_fail(e, "fdgrad_refmode_kernel launch");
}

rbd_status rbd_fdgrad_traj_host_f32(rbd_context* ctx, int64_t n, const float* h_in, float* h_out) {
  if (rbd_status s = check_ctx(ctx)) return s;
  if (n < 0) return fail(RBD_ERR_INVALID_ARGUMENT, "negative knot count");
  if (n == 0) return RBD_OK;
  if (!h_in || !h_out) return fail(RBD_ERR_INVALID_ARGUMENT, "null buffer");
  if (rbd_status s = use_device(ctx)) return s;
  std::lock_guard<std::mutex> lock(ctx->host_mu);
  const size_t in_bytes = size_t(n) * rbd::kInPerKnot * sizeof(float);
  const size_t out_bytes = size_t(n) * rbd::kOutPerKnot * sizeof(float);
  if (ctx->scratch_bytes < in_bytes + out_bytes) {
    if (ctx->d_scratch) (void)hipFree(ctx->d_scratch);
    ctx->d_scratch = nullptr;
    ctx->scratch_bytes = 0;
    hipError_t e = hipMalloc(&ctx->d_scratch, in_bytes + out_bytes);
    if (e != hipSuccess) return fail(RBD_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
    ctx->scratch_bytes = in_bytes + out_bytes;
  }
  float* d_in = ctx->d_scratch;
  float* d_out = ctx->d_scratch + size_t(n) * rbd::kInPerKnot;
  hipError_t e = hipMemcpy(d_in, h_in, in_bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpy H2D");
  e = by_precision<LaunchAoS>(ctx->precision, n, d_in, d_out, ctx->opt, (hipStream_t)0);
  if (e != hipSuccess) return hip_fail(e, "fdgrad_aos_kernel launch");
  e = hipMemcpy(h_out, d_out, out_bytes, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpy D2H");
  return RBD_OK;
}

const char* rbd_kernel_name(rbd_context* ctx, int reference_mode) {
  if (!ctx) return "";
  static const char* aos[] = {"fdgrad_aos_kernel<PrecFP32>", "fdgrad_aos_kernel<PrecFP64Pivot>",
                              "fdgrad_aos_kernel<PrecFP64Col>", "fdgrad_aos_kernel<PrecFP64>"};
  static const char* ref[] = {"fdgrad_refmode_kernel<PrecFP32>", "fdgrad_refmode_kernel<PrecFP64Pivot>",
                              "fdgrad_refmode_kernel<PrecFP64Col>", "fdgrad_refmode_kernel<PrecFP64>"};
  return reference_mode ? ref[ctx->precision] : aos[ctx->precision];
}

}  // extern "C"
