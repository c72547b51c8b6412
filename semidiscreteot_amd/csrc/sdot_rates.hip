// sdot_rates.hip -- instruction-rate microbenchmarks of the units the pair
// sweeps run on (the measured peaks of the roofline that bench.py reports):
//   exp     v_exp_f32 (the FP32 sweeps' transcendental)
//   fma64   v_fma_f64 (the FP64 exponent arithmetic of the VALU sweeps)
//   mfma64  v_mfma_f64_16x16x4_f64 (the matrix-core exponent tiles)
//   mix     two MFMA tiles, the VALU work of an FP32 sum sweep on 8 pairs
//           per lane, and both together: whether the two units overlap
//   op      v_add_f64, v_max_f64, the two FP64/FP32 conversions, v_fma_f32:
//           SIMD cycles per wave instruction of every operation of the sweeps
// Every kernel runs independent dependency chains per lane (enough to cover
// the units' latencies), over a grid of 8 waves per SIMD, and stores its
// chains' values so the work is not removed.  Prints one JSON object:
// lane-operations (or MFMA instructions) per second and per CU cycle.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CK(x)                                                                    \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess) {                                                      \
      std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      std::exit(2);                                                              \
    }                                                                            \
  } while (0)

constexpr int kChains = 8;
constexpr int kIters = 4096;

__global__ __launch_bounds__(256) void k_exp(float *out, float seed) {
  float x[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) x[c] = -seed * (c + 1) - 1e-3f * threadIdx.x;
  for (int it = 0; it < kIters; ++it)
#pragma unroll
    for (int c = 0; c < kChains; ++c) x[c] = -__builtin_amdgcn_exp2f(x[c]);
  float s = 0.0f;
#pragma unroll
  for (int c = 0; c < kChains; ++c) s += x[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_fma64(double *out, double seed) {
  double x[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) x[c] = seed * (c + 1) + 1e-9 * threadIdx.x;
  const double m = 0.999999, b = 1e-7;
  for (int it = 0; it < kIters; ++it)
#pragma unroll
    for (int c = 0; c < kChains; ++c) x[c] = fma(x[c], m, b);
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < kChains; ++c) s += x[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// Issue cost of the other instructions of the sweeps' pair arithmetic, one
// operation kind per kernel (8 chains per lane, as above): K = 0 v_add_f64,
// 1 v_max_f64 + v_add_f64, 2 v_cvt_f32_f64 + v_cvt_f64_f32 + v_add_f64,
// 3 v_fma_f32.
template <int K>
__global__ __launch_bounds__(256) void k_op(double *out, double s, double m, float sf) {
  double x[kChains];
  float f[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) {
    x[c] = s * (c + 1) + 1e-9 * threadIdx.x;
    f[c] = sf * (c + 1) + 1e-6f * threadIdx.x;
  }
  for (int it = 0; it < kIters; ++it)
#pragma unroll
    for (int c = 0; c < kChains; ++c) {
      if (K == 0) x[c] = x[c] - s;
      if (K == 1) {
        double r;
        asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(x[c]), "v"(m));
        x[c] = r - s;
      }
      if (K == 2) x[c] = (double)(float)x[c] - s;
      if (K == 3) f[c] = fmaf(f[c], sf, sf);
    }
  double r = 0.0;
#pragma unroll
  for (int c = 0; c < kChains; ++c) r += x[c] + f[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = r;
}

typedef double d4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_mfma64(double *out, double seed) {
  constexpr int kAcc = 4;
  d4 acc[kAcc];
#pragma unroll
  for (int c = 0; c < kAcc; ++c) acc[c] = d4{seed, seed, seed, seed};
  const double a = 1e-3 * (threadIdx.x & 63), b = seed;
  for (int it = 0; it < kIters / 8; ++it)
#pragma unroll
    for (int c = 0; c < kAcc; ++c) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < kAcc; ++c) s += acc[c][0] + acc[c][1] + acc[c][2] + acc[c][3];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// Co-execution of FP64 matrix and vector work: per step, two independent
// MFMA tiles (mode 1), the VALU work an FP32 sum sweep does on 8 pairs per
// lane (8 v_cvt_f32_f64, 8 v_exp_f32, 8 v_cvt_f64_f32, 8 v_fma_f64, 4 v_max_u32;
// mode 2), or both, independent of each other (mode 3).  If the matrix cores
// ran beside the vector ALU, mode 3 would take max(mode 1, mode 2).
__global__ __launch_bounds__(256) void k_mix(double *out, double seed, int mode) {
  const double a = 1e-3 * (threadIdx.x & 63), b = -seed;
  d4 m0 = {-1.0, -2.0, -3.0, -4.0}, m1 = m0;
  double S[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  unsigned E[4] = {0, 0, 0, 0};
  double v[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) v[i] = -1e-3 * (i + 1) - 1e-6 * threadIdx.x;
  const double nu = 1e-3;
  for (int it = 0; it < kIters / 8; ++it) {
    if (mode & 1) {
      m0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, m0, 0, 0, 0);
      m1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, m1, 0, 0, 0);
    }
    if (mode & 2) {
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const float e = __builtin_amdgcn_exp2f((float)v[i]);
        E[i & 3] = max(E[i & 3], __float_as_uint(e));
        S[i] = fma(nu, (double)e, S[i]);
        v[i] = S[i] - 1.0;
      }
    }
  }
  double r = (double)(E[0] ^ E[1] ^ E[2] ^ E[3]);
#pragma unroll
  for (int i = 0; i < 8; ++i) r += S[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = r + m0[0] + m1[1];
}

template <class F>
static double time_ms(F &&launch) {
  hipEvent_t a, b;
  CK(hipEventCreate(&a));
  CK(hipEventCreate(&b));
  launch();  // warm-up
  CK(hipDeviceSynchronize());
  CK(hipEventRecord(a, 0));
  for (int r = 0; r < 5; ++r) launch();
  CK(hipEventRecord(b, 0));
  CK(hipEventSynchronize(b));
  float ms = 0;
  CK(hipEventElapsedTime(&ms, a, b));
  CK(hipEventDestroy(a));
  CK(hipEventDestroy(b));
  return ms / 5.0;
}

int main() {
  hipDeviceProp_t prop;
  CK(hipGetDeviceProperties(&prop, 0));
  const int cus = prop.multiProcessorCount;
  const double clk_hz = prop.clockRate * 1e3;  // kHz -> Hz
  const int blocks = cus * 8;                  // 8 waves per SIMD (256 threads per block)
  const int threads = 256;
  const double lanes = (double)blocks * threads;
  double *dd = nullptr;
  float *df = nullptr;
  CK(hipMalloc((void **)&dd, sizeof(double) * blocks * threads));
  CK(hipMalloc((void **)&df, sizeof(float) * blocks * threads));
  const double t_exp = time_ms([&] { hipLaunchKernelGGL(k_exp, dim3(blocks), dim3(threads), 0, 0, df, 0.5f); });
  const double t_fma = time_ms([&] { hipLaunchKernelGGL(k_fma64, dim3(blocks), dim3(threads), 0, 0, dd, 0.5); });
  const double t_mfma = time_ms([&] { hipLaunchKernelGGL(k_mfma64, dim3(blocks), dim3(threads), 0, 0, dd, 0.5); });
  double t_mix[4] = {0, 0, 0, 0};
  for (int mode = 1; mode <= 3; ++mode)
    t_mix[mode] = time_ms([&] { hipLaunchKernelGGL(k_mix, dim3(blocks), dim3(threads), 0, 0, dd, 0.5, mode); });
  double t_op[4];
  t_op[0] = time_ms([&] { hipLaunchKernelGGL(k_op<0>, dim3(blocks), dim3(threads), 0, 0, dd, 0.5, 0.999999, 0.5f); });
  t_op[1] = time_ms([&] { hipLaunchKernelGGL(k_op<1>, dim3(blocks), dim3(threads), 0, 0, dd, 0.5, 0.999999, 0.5f); });
  t_op[2] = time_ms([&] { hipLaunchKernelGGL(k_op<2>, dim3(blocks), dim3(threads), 0, 0, dd, 0.5, 0.999999, 0.5f); });
  t_op[3] = time_ms([&] { hipLaunchKernelGGL(k_op<3>, dim3(blocks), dim3(threads), 0, 0, dd, 0.5, 0.999999, 0.5f); });
  CK(hipGetLastError());
  const double exp_ops = lanes * kIters * kChains;          // lane-operations
  const double fma_ops = lanes * kIters * kChains;
  const double waves = lanes / 64.0;
  const double mfma_ins = waves * (kIters / 8) * 4;          // wave instructions
  const double steps = waves * (kIters / 8);                 // k_mix steps
  const double cyc = clk_hz * cus;                           // CU cycles per second
  // SIMD cycles per k_mix step (4 SIMDs per CU)
  auto per_step = [&](double ms) { return cyc * 4.0 * (ms * 1e-3) / steps; };
  // SIMD cycles per wave instruction (chain step) of the k_op kernels and of
  // k_exp / k_fma64: cycles of all SIMDs over wave-level chain steps
  const double wsteps = waves * kIters * kChains;
  auto per_ins = [&](double ms) { return cyc * 4.0 * (ms * 1e-3) / wsteps; };
  const double c_add = per_ins(t_op[0]);
  const double c_max = per_ins(t_op[1]) - c_add;
  const double c_cvt = 0.5 * (per_ins(t_op[2]) - c_add);
  std::printf(
      "{\"device\": \"%s\", \"cus\": %d, \"clock_mhz\": %.0f, "
      "\"exp_f32_per_s\": %.4e, \"exp_f32_per_cu_cycle\": %.2f, "
      "\"fma_f64_per_s\": %.4e, \"fma_f64_per_cu_cycle\": %.2f, "
      "\"mfma_f64_16x16x4_per_s\": %.4e, \"mfma_f64_cycles_per_simd\": %.1f, "
      "\"coexec_simd_cycles_per_step\": {\"mfma_only\": %.1f, \"valu_only\": %.1f, \"both\": %.1f}, "
      "\"simd_cycles_per_wave_instruction\": {\"v_fma_f64\": %.2f, \"v_add_f64\": %.2f, "
      "\"v_max_f64\": %.2f, \"v_cvt_f32_f64_or_f64_f32\": %.2f, \"v_fma_f32\": %.2f, "
      "\"v_exp_f32\": %.2f}}\n",
      prop.name, cus, clk_hz / 1e6, exp_ops / (t_exp * 1e-3), exp_ops / (t_exp * 1e-3) / cyc,
      fma_ops / (t_fma * 1e-3), fma_ops / (t_fma * 1e-3) / cyc, mfma_ins / (t_mfma * 1e-3),
      cyc * 4.0 / (mfma_ins / (t_mfma * 1e-3)), per_step(t_mix[1]), per_step(t_mix[2]),
      per_step(t_mix[3]), per_ins(t_fma), c_add, c_max, c_cvt, per_ins(t_op[3]), per_ins(t_exp));
  CK(hipFree(dd));
  CK(hipFree(df));
  return 0;
}
