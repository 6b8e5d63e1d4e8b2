// hml_sgemv_main.cpp -- command-line SGEMV driver on libhml.so.
//
// Counterpart of the reference's apps/gemv/hml_sgemv_main.cu:294-487: same
// flags (-c -h -M -N -o -r -s -t -u -y), inputs (rand()/RAND_MAX in the
// order A, x, y), outputs ("GPU ysum = ...", "CPU ysum = ...", "ysum
// absolute/relative difference"; median GFLOPS lines on stderr) and -c
// gate.  Timing as hmlSgemvBasicTest / hmlSgemvTest (:46-247).
#include <getopt.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <sys/time.h>

#include <algorithm>
#include <vector>

#include "hml.h"

static const double kRelErrMax = 1.0e-8;

static double wall_secs() {
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec + tv.tv_usec * 1e-6;
}

static void die(const char *what, hmlStatus s) {
  fprintf(stderr, "; Error: %s: %s (%s)\n", what, hmlStatusString(s), hmlLastError());
  exit(1);
}

// Reference CPU path (apps/gemv/hml_sgemv_main.cu:254-283).
static void sgemv_cpu(float *y, const float *A, const float *x, int M, int N,
                      float alpha, float beta) {
  const double gflop = (2.0 * M * N + 2.0 * M) / 1e9;
  const double t0 = wall_secs();
  for (int i = 0; i < M; ++i, A += N) {
    volatile float ax = 0.0f;
    for (int j = 0; j < N; ++j) {
      const float p = A[j] * x[j];
      ax = ax + p;
    }
    y[i] = alpha * ax + beta * y[i];
  }
  const double t1 = wall_secs();
  fprintf(stderr, "CPU time = %lf, wall time = %lf\n", t1 - t0, t1 - t0);
  fprintf(stderr, "GFLOPS = %lf\n", gflop / (t1 - t0));
}

static double run_test(float *y, const float *A, const float *x, int M, int N,
                       float alpha, float beta, int trials, bool basic) {
  float *dA, *dx, *dy;
  hmlStatus s;
  if ((s = hmlDeviceMalloc((void **)&dA, sizeof(float) * (size_t)M * N))) die("alloc A", s);
  if ((s = hmlDeviceMalloc((void **)&dx, sizeof(float) * (size_t)N))) die("alloc x", s);
  if ((s = hmlDeviceMalloc((void **)&dy, sizeof(float) * (size_t)M))) die("alloc y", s);
  if ((s = hmlMemcpyHtoD(dA, A, sizeof(float) * (size_t)M * N))) die("copy A", s);
  if ((s = hmlMemcpyHtoD(dx, x, sizeof(float) * (size_t)N))) die("copy x", s);
  if ((s = hmlMemcpyHtoD(dy, y, sizeof(float) * (size_t)M))) die("copy y", s);
  hmlSgemvVariant v = {HML_SGEMV_FAMILY_DOT, 64, 1, 1, 4};  // the basic variant
  int32_t src = 0;
  if (!basic) {
    hmlSgemvSelect(HML_ROW_MAJOR, HML_OP_N, M, N, beta, &v, &src);
    fprintf(stderr, "Info: kernel type = %d, blockStops = %d, rowStops = %d\n",
            v.family, v.lpr / 16 > 0 ? v.lpr / 16 : 1, v.rb);
  }
  const double gflop = (2.0 * M * N + 2.0 * M) / 1e9;
  std::vector<double> g;
  for (int t = 0; t < trials; ++t) {
    const double t0 = wall_secs();
    s = hmlSgemvExVariant(HML_ROW_MAJOR, HML_OP_N, M, N, alpha, dA, N, dx, 1, beta, dy, 1,
                          HML_FLAG_READ_C_ALWAYS, &v, nullptr, 0, nullptr);
    if (s == HML_STATUS_SUCCESS) s = hmlDeviceSynchronize();
    const double t1 = wall_secs();
    if (s) {
      fprintf(stderr, "Run kernel: %s\n", hmlLastError());
      exit(1);
    }
    g.push_back(gflop / (t1 - t0));
  }
  std::sort(g.begin(), g.end());
  if ((s = hmlMemcpyDtoH(y, dy, sizeof(float) * (size_t)M))) {
    fprintf(stderr, "Copy y off of device: %s\n", hmlLastError());
    exit(1);
  }
  hmlDeviceFree(dA);
  hmlDeviceFree(dx);
  hmlDeviceFree(dy);
  return g[trials / 2];
}

int main(int argc, char *argv[]) {
  const char help[] =
      "This program does GPU-based Single-precision GEneral Matrix-Vector multiplication (SGEMV):\n"
      "y = alpha * A * x + beta * y\n"
      "where A is M x N, x is N x 1, and y is M x 1\n"
      "Options:\n"
      "\t-c Cross check with CPU-based SGEMM results\n"
      "\t-h Print this help message\n"
      "\t-M <M>\n\t-N <N>\n"
      "\t-o <output smart-kernel config file name>\n"
      "\t-r <random seed to generate the matrices>\n"
      "\t-s <input smart-kernel config file name>\n"
      "\t-t <number of tests for each kernel>\n"
      "\t-u <GPU ID in [0, #GPUs - 1]>\n"
      "\t-y <verbosity level in [0,2]>\n";
  int M = 0, N = 0, trials = 5, verbosity = 1, gpu = 0, option;
  bool run_cpu = false;
  char *cfg_in = nullptr, *cfg_out = nullptr;
  while ((option = getopt(argc, argv, ":chM:N:o:r:s:t:u:y:")) != -1) {
    switch (option) {
      case 'c': run_cpu = true; break;
      case 'h': fprintf(stderr, "Help:\n%s\n", help); exit(EXIT_FAILURE);
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
  if (M == 0 || N == 0) {
    fprintf(stderr, "M and N cannot be zero\n");
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
  if (cfg_in) hmlSgemvKernelConfig(cfg_in);
  else hmlSgemvKernelConfig(nullptr);
  if (cfg_out) {
    if (cfg_in) fprintf(stderr, "Warning: Both input and output config files are specified\n");
    if ((s = hmlSgemvConfigSave(cfg_out))) die("write config", s);
  }
  std::vector<float> A((size_t)M * N), x(N), y(M);
  for (auto &v : A) v = ((float)rand()) / RAND_MAX;
  for (auto &v : x) v = ((float)rand()) / RAND_MAX;
  for (auto &v : y) v = ((float)rand()) / RAND_MAX;
  double g = run_test(y.data(), A.data(), x.data(), M, N, 1.0f, 0.0f, trials, true);
  fprintf(stderr, "Basic kernel median GFLOPS = %.2lf\n", g);
  g = run_test(y.data(), A.data(), x.data(), M, N, 1.0f, 0.0f, trials, false);
  fprintf(stderr, "Smart kernel median GFLOPS = %.2lf\n", g);
  double ysum_gpu = 0.0;
  for (int i = 0; i < M; ++i) ysum_gpu += y[i];
  fprintf(stdout, "GPU ysum = %.20lf\n", ysum_gpu);
  if (run_cpu) {
    sgemv_cpu(y.data(), A.data(), x.data(), M, N, 1.0f, 0.0f);
    double ysum_cpu = 0.0;
    for (int i = 0; i < M; ++i) ysum_cpu += y[i];
    fprintf(stdout, "CPU ysum = %.20lf\n", ysum_cpu);
    const double absd = fabs(ysum_gpu - ysum_cpu);
    fprintf(stdout, "ysum absolute difference: %.20lf\n", absd);
    const double reld = absd / ysum_cpu;
    fprintf(stdout, "ysum relative difference: %.20lf\n", reld);
    if (reld > kRelErrMax) {
      fprintf(stderr, "; Error: CPU and GPU results differ\n");
      return 1;
    }
  }
  return 0;
}
