// Drop-in example: the reference's examples/pcg_solve.cu (n=2, N=3 system, Pinv
// the element-wise reciprocal of the non-zero S entries, default pcg_config)
// through the C++ facade, in FP32 and FP64; then the documented pcg_solve with
// no Pinv (symmetric stair built on the device) on a larger system.
// Prints one line per solve: "<name> iters=<k> lambda=<v0> <v1> ...".
#include <math.h>
#include <stdio.h>

#include <vector>

#include "gpu_pcg.hpp"

using namespace gbdpcg;  // the reference declares these names globally

template <typename T>
static int readme_example(const char *name) {
  const uint32_t state_size = 2, knot_points = 3;
  T h_S[36] = {0,     0,      0,      0,       -.999, 0,     0,     -.999, .999,   .0999,  -.98, .999,
               .999,  -.98,   .0999,  .999,    -2.008, .8801, .8801, -3.0584, .999, .0999, -.98, .999,
               .999,  -.98,   .0999,  .999,    -1.019, .8801, .8801, -2.0694, 0,    0,      0,    0};
  T h_Pinv[36];
  for (int i = 0; i < 36; i++) h_Pinv[i] = h_S[i] == 0 ? T(0) : T(1) / h_S[i];
  T h_gamma[6] = {T(3.1385), 0, 0, T(3.0788), T(.0031), T(3.0788)};
  T h_lambda[6] = {0, 0, 0, 0, 0, 0};
  struct pcg_config<T> config;  // the reference example, examples/pcg_solve.cu:34-41
  const int iters = solvePCG<T>(h_S, h_Pinv, h_gamma, h_lambda, state_size, knot_points, &config);
  if (iters < 0) {
    printf("%s error %d: %s\n", name, iters, gbdpcg_last_error());
    return 1;
  }
  printf("%s iters=%d lambda=", name, iters);
  for (int i = 0; i < 6; i++) printf("%.9g ", (double)h_lambda[i]);
  printf("\n");
  return 0;
}

static int lqr_example() {
  const uint32_t n = 14, m = 7, N = 128;
  std::vector<double> S(3 * n * n * N), gamma(n * N), lambda(n * N, 0.0);
  if (gbdpcg_lqr_generate(GBDPCG_F64, n, m, N, 2128, 0.1, S.data(), gamma.data(), nullptr, nullptr, nullptr,
                          nullptr) != GBDPCG_OK)
    return 1;
  cbtd_t<double> cb{S.data(), nullptr};
  pcg_config<double> config(1e-10, 500);
  config.exit_mode = GBDPCG_EXIT_REL_RESID;
  const int iters = pcg_solve(&cb, gamma.data(), lambda.data(), n, N, &config);
  if (iters < 0) {
    printf("lqr error %d: %s\n", iters, gbdpcg_last_error());
    return 1;
  }
  double nrm = 0;
  for (double v : lambda) nrm += v * v;
  printf("lqr_n14_N128_fp64 iters=%d lambda_norm=%.12g lambda=%.12g %.12g\n", iters, sqrt(nrm), lambda[0],
         lambda[n * N - 1]);
  return 0;
}

int main() {
  int rc = 0;
  rc |= readme_example<float>("readme_fp32");
  rc |= readme_example<double>("readme_fp64");
  rc |= lqr_example();
  return rc;
}
