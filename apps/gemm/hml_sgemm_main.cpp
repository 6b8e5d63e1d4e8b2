// hml_sgemm_main.cpp -- command-line SGEMM driver on libhml.so.
//
// Counterpart of the reference's apps/gemm/hml_sgemm_main.cu:307-515: same
// flags (-c -h -K -M -N -o -r -s -t -u -y; -t is parsed here, the reference
// lists it in its help but leaves it out of getopt), same inputs
// (rand()/RAND_MAX in the order A, B, C, seeded by -r only), same outputs
// ("GPU Csum = ...", "CPU Csum = ...", "Csum absolute/relative difference"
// on stdout; median GFLOPS lines on stderr), same 1e-8 checksum gate for -c.
// Timing follows hmlSgemmTestBasic / hmlSgemmTest (:55-252): wall clock
// around launch + device synchronize, median of the trials.
#include <getopt.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <sys/time.h>

#include <algorithm>
#include <vector>

#include "hml.h"

static const double kRelErrMax = 1.0e-8;  // apps/gemm/hml_sgemm_main.cu:33

static double wall_secs() {
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec + tv.tv_usec * 1e-6;
}

static void die(const char *what, hmlStatus s) {
  fprintf(stderr, "; Error: %s: %s (%s)\n", what, hmlStatusString(s), hmlLastError());
  exit(1);
}

// Reference CPU path (apps/gemm/hml_sgemm_main.cu:260-292): serial loops,
// float accumulator, C = alpha*sum + beta*C.
static void sgemm_cpu(float *C, const float *A, const float *B, int M, int N,
                      int K, float alpha, float beta) {
  const double gflop = 2.0 * M * N * K / 1e9;
  const double t0 = wall_secs();
  for (int i = 0; i < M; ++i, A += K, C += N) {
    for (int j = 0; j < N; ++j) {
      const float *colB = B + j;
      volatile float sum = 0.0f;  // keep the reference's separate roundings
      for (int e = 0; e < K; ++e) {
        const float p = A[e] * colB[(size_t)e * N];
        sum = sum + p;
      }
      C[j] = alpha * sum + beta * C[j];
    }
  }
  const double t1 = wall_secs();
  fprintf(stderr, "CPU time = %lf, wall time = %lf\n", t1 - t0, t1 - t0);
  fprintf(stderr, "GFLOPS = %lf\n", gflop / (t1 - t0));
}

static double csum(const float *C, int M, int N) {
  double s = 0.0;
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < N; ++j) s += (double)C[(size_t)i * N + j];
  return s;
}

// Copies in, runs `trials` times (median GFLOPS), copies C back.
static double run_test(float *C, const float *A, const float *B, int M, int N,
                       int K, float alpha, float beta, int trials,
                       bool basic) {
  float *dA, *dB, *dC;
  hmlStatus s;
  if ((s = hmlDeviceMalloc((void **)&dA, sizeof(float) * (size_t)M * K))) die("alloc A", s);
  if ((s = hmlDeviceMalloc((void **)&dB, sizeof(float) * (size_t)K * N))) die("alloc B", s);
  if ((s = hmlDeviceMalloc((void **)&dC, sizeof(float) * (size_t)M * N))) die("alloc C", s);
  if ((s = hmlMemcpyHtoD(dA, A, sizeof(float) * (size_t)M * K))) die("copy A", s);
  if ((s = hmlMemcpyHtoD(dB, B, sizeof(float) * (size_t)K * N))) die("copy B", s);
  if ((s = hmlMemcpyHtoD(dC, C, sizeof(float) * (size_t)M * N))) die("copy C", s);
  hmlSgemmVariant basic_v = {HML_SGEMM_FAMILY_DOT, 1, 4, 1, 4};
  int32_t src = 0;
  if (!basic) {
    hmlSgemmSelect(HML_ROW_MAJOR, HML_OP_N, HML_OP_N, M, N, K, beta, &basic_v, &src);
    fprintf(stderr, "Info: kernel type = %d, colStops = %d, rowStops = %d\n",
            basic_v.family, basic_v.ct, basic_v.rt);
  } else if (N > M) {
    basic_v.family = HML_SGEMM_FAMILY_AXPY;  // the basic tile of the transposed problem
  }
  const double gflop = 2.0 * M * N * K / 1e9;
  std::vector<double> g;
  for (int t = 0; t < trials; ++t) {
    const double t0 = wall_secs();
    s = hmlSgemmExVariant(HML_ROW_MAJOR, HML_OP_N, HML_OP_N, M, N, K, alpha, dA, K, dB, N,
                          beta, dC, N, HML_FLAG_READ_C_ALWAYS, &basic_v, nullptr, 0, nullptr);
    if (s == HML_STATUS_SUCCESS) s = hmlDeviceSynchronize();
    const double t1 = wall_secs();
    if (s) {
      fprintf(stderr, "Run kernel: %s\n", hmlLastError());
      exit(1);
    }
    g.push_back(gflop / (t1 - t0));
  }
  std::sort(g.begin(), g.end());
  if ((s = hmlMemcpyDtoH(C, dC, sizeof(float) * (size_t)M * N))) {
    fprintf(stderr, "Copy C off of device: %s\n", hmlLastError());
    exit(1);
  }
  hmlDeviceFree(dA);
  hmlDeviceFree(dB);
  hmlDeviceFree(dC);
  return g[trials / 2];
}

int main(int argc, char *argv[]) {
  const char help[] =
      "This program does GPU-based Single-precision GEneral Matrix-Matrix multiplication (SGEMM):\n"
      "C = alpha * A * B + beta * C\n"
      "where A is M x K, B is K x N, and C is M x N\n"
      "Options:\n"
      "\t-c Cross check with CPU-based SGEMM results\n"
      "\t-h Print this help message\n"
      "\t-K <K>\n\t-M <M>\n\t-N <N>\n"
      "\t-o <output smart-kernel config file name>\n"
      "\t-r <random seed to generate the matrices>\n"
      "\t-s <input smart-kernel config file name>\n"
      "\t-t <number of tests for each kernel>\n"
      "\t-u <GPU ID in [0, #GPUs - 1]>\n"
      "\t-y <verbosity level in [0,2]>\n";
  int M = 0, N = 0, K = 0, trials = 5, verbosity = 1, gpu = 0, option;
  bool run_cpu = false;
  char *cfg_in = nullptr, *cfg_out = nullptr;
  while ((option = getopt(argc, argv, ":chK:M:N:o:r:s:t:u:y:")) != -1) {
    switch (option) {
      case 'c': run_cpu = true; break;
      case 'h': fprintf(stderr, "Help:\n%s\n", help); exit(EXIT_FAILURE);
      case 'K': K = atoi(optarg); break;
      case 'M': M = atoi(optarg); break;
      case 'N': N = atoi(optarg); break;
      case 'o': cfg_out = optarg; break;
      case 'r': {
        const unsigned seed = (unsigned)atoi(optarg);
        srand(seed);
        if (verbosity >= 2) fprintf(stderr, "Info: random seed = %u\n", seed);
        break;
      }
      case 's': cfg_in = optarg; break;
      case 't': trials = atoi(optarg); break;
      case 'u': gpu = atoi(optarg); break;
      case 'y': verbosity = atoi(optarg); break;
      case ':': fprintf(stderr, "Option -%c requires an argument\n", optopt); exit(EXIT_FAILURE);
      case '?': fprintf(stderr, "Unknown option character '%c'.\n", optopt); exit(EXIT_FAILURE);
    }
  }
  if (M == 0 || N == 0 || K == 0) {
    fprintf(stderr, "M, N, and K cannot be zero\n");
    exit(EXIT_FAILURE);
  }
  if (trials < 1) trials = 1;
  hmlStatus s = hmlInit();
  if (s) die("init", s);
  if (gpu >= hmlDeviceCount()) {
    fprintf(stderr, "Warning: Invalid GPU card #%d, resetting to default (0)\n", gpu);
    gpu = 0;
  }
  if ((s = hmlSetDevice(gpu))) die("set device", s);
  if (cfg_in) {
    hmlSgemmKernelConfig(cfg_in);
  } else {
    hmlSgemmKernelConfig(nullptr);  // online autotune, as the reference app does
  }
  if (cfg_out) {
    if (cfg_in) fprintf(stderr, "Warning: Both input and output config files are specified\n");
    if ((s = hmlSgemmConfigSave(cfg_out))) die("write config", s);
  }
  std::vector<float> A((size_t)M * K), B((size_t)K * N), C((size_t)M * N);
  for (auto &v : A) v = ((float)rand()) / RAND_MAX;  // src/hml_utils.cu:60
  for (auto &v : B) v = ((float)rand()) / RAND_MAX;
  for (auto &v : C) v = ((float)rand()) / RAND_MAX;
  fprintf(stderr, "alloc done\n");
  double g = run_test(C.data(), A.data(), B.data(), M, N, K, 1.0f, 0.0f, trials, true);
  fprintf(stderr, "Basic kernel median GFLOPS = %.2lf\n", g);
  g = run_test(C.data(), A.data(), B.data(), M, N, K, 1.0f, 0.0f, trials, false);
  fprintf(stderr, "Smart kernel median GFLOPS = %.2lf\n", g);
  const double gpu_sum = csum(C.data(), M, N);
  fprintf(stdout, "GPU Csum = %.20lf\n", gpu_sum);
  if (run_cpu) {
    sgemm_cpu(C.data(), A.data(), B.data(), M, N, K, 1.0f, 0.0f);
    const double cpu_sum = csum(C.data(), M, N);
    fprintf(stdout, "CPU Csum = %.20lf\n", cpu_sum);
    const double absd = fabs(gpu_sum - cpu_sum);
    fprintf(stdout, "Csum absolute difference: %.20lf\n", absd);
    const double reld = absd / cpu_sum;
    fprintf(stdout, "Csum relative difference: %.20lf\n", reld);
    if (reld > kRelErrMax) {
      fprintf(stderr, "; Error: CPU and GPU results differ\n");
      return 1;
    }
  }
  return 0;
}
