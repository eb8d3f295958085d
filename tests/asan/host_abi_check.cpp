// Host-side exercise of the C ABI for an AddressSanitizer build (no GPU
// needed, no kernel launched): the set-up, transitions, initial trajectories,
// vector fields, size queries and every argument check of include/pof.h that
// runs on the host, over every field kind and order the kernels cover, with
// boundary sizes.  Exit status 0 when every call returns what the header says;
// ASan aborts the process on any out-of-bounds access, use after free or leak.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/pof.h"

static int failures = 0;
#define CHECK(cond)                                                      \
  do {                                                                   \
    if (!(cond)) {                                                       \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #cond, __LINE__); \
      ++failures;                                                        \
    }                                                                    \
  } while (0)

static const int kKinds[] = {POF_LOGISTIC, POF_LOTKAVOLTERRA, POF_VANDERPOL, POF_FITZHUGHNAGUMO, POF_FHN_CLASSIC,
                             POF_ROBER, POF_RIGID_BODY, POF_SEIR, POF_THREEBODY, POF_HENONHEILES, POF_LORENZ63,
                             POF_LORENZ96};
static int dim_of(int kind) {
  switch (kind) {
    case POF_LOGISTIC: return 1;
    case POF_LOTKAVOLTERRA: case POF_VANDERPOL: case POF_FITZHUGHNAGUMO: case POF_FHN_CLASSIC: return 2;
    case POF_ROBER: case POF_RIGID_BODY: case POF_LORENZ63: return 3;
    case POF_SEIR: case POF_THREEBODY: case POF_HENONHEILES: return 4;
    default: return 4;  // Lorenz-96 with d = 4
  }
}

int main() {
  // transitions of every order and dimension the kernels cover
  for (int d = 1; d <= 4; ++d)
    for (int q = 0; q <= 5; ++q) {
      const int D = d * (q + 1);
      std::vector<double> F(D * D), QL(D * D), P(D), PI(D);
      CHECK(pof_transitions(d, q, 0.01, F.data(), QL.data(), P.data(), PI.data()) == POF_OK);
      CHECK(pof_transitions(d, q, 0.01, nullptr, nullptr, nullptr, nullptr) == POF_OK);
      for (int i = 0; i < D; ++i) CHECK(std::fabs(P[i] * PI[i] - 1.0) < 1e-12);
    }
  CHECK(pof_transitions(0, 1, 0.1, nullptr, nullptr, nullptr, nullptr) == POF_E_ARG);
  // set-up and initial trajectories for every field, orders 1..4, grids of 2 .. 300 points
  for (int kind : kKinds) {
    const int d = dim_of(kind);
    for (int q = 1; q <= 4; ++q)
      for (int n : {2, 3, 17, 300}) {
        std::vector<double> ts(n), y0(d, 0.5);
        for (int k = 0; k < n; ++k) ts[k] = 0.1 * k;
        pof_problem p{};
        p.kind = kind;
        p.d = d;
        p.q = q;
        p.n_times = n;
        p.ts_host = ts.data();
        p.y0_host = y0.data();
        if (kind == POF_LOGISTIC) p.params[0] = p.params[1] = 1.0;
        if (kind == POF_LORENZ96) p.params[0] = 8.0;
        const int D = d * (q + 1);
        std::vector<double> F(D * D), QL(D * D), P(D), PI(D), m0(D), L0(D * D), means((size_t)n * D);
        const int rc = pof_setup(&p, F.data(), QL.data(), P.data(), PI.data(), m0.data(), L0.data());
        if (rc != POF_OK) continue;  // a field whose parameters this check does not set
        for (int init : {POF_INIT_CONSTANT, POF_INIT_PRIOR}) {
          p.init = init;
          CHECK(pof_initial_means(&p, means.data()) == POF_OK);
        }
        p.init = POF_INIT_COARSE;
        CHECK(pof_initial_means(&p, means.data()) == POF_E_UNSUPPORTED);
        p.init = 7;
        CHECK(pof_initial_means(&p, means.data()) == POF_E_INIT);
        std::vector<double> f(d);
        CHECK(pof_eval_field(kind, d, p.params, y0.data(), f.data()) == POF_OK);
      }
  }
  // argument checks that run on the host
  pof_problem bad{};
  bad.kind = POF_LOGISTIC;
  bad.d = 1;
  bad.q = 1;
  bad.n_times = 1;  // fewer than two time points
  double ts1[1] = {0.0}, y1[1] = {0.1};
  bad.ts_host = ts1;
  bad.y0_host = y1;
  double m[4];
  CHECK(pof_initial_means(&bad, m) == POF_E_ARG);
  CHECK(pof_eval_field(99, 1, y1, y1, m) == POF_E_ARG);
  CHECK(pof_eval_field(POF_LOGISTIC, 1, nullptr, y1, m) == POF_E_ARG);
  CHECK(pof_supported(12, 3) == 1 && pof_supported(13, 3) == 0);
  CHECK(pof_batch_entries_for(100) == 200);
  CHECK(pof_batch_bytes(12, 4096, 8) > 0);
  CHECK(std::strlen(pof_last_error(nullptr)) > 0);
  CHECK(pof_ctx_create(0, nullptr) == POF_E_ARG);
  CHECK(pof_solve(nullptr, nullptr, nullptr, nullptr, nullptr) != POF_OK);
  std::printf("host ABI checks: %d failures\n", failures);
  return failures == 0 ? 0 : 1;
}
