// Host build of the engine's device orientation predicates
// (visacd_amd/csrc/predicates.hpp), for tests/test_predicates_host.py: each
// predicate evaluated by its complete form (filter, integer grid, expansion)
// and by its integer and expansion paths alone, so that every path can be
// checked against the oracle's exact predicates on the same float32 inputs.
#include "../visacd_amd/csrc/predicates.hpp"

#include <cstdint>

extern "C" {

// kind 0: orient3d(a, b, c, d)      inputs 12 floats per case
// kind 1: orient3d(a, b, c, (p+q)/2) inputs 15 floats
// kind 2: orient2d(a, b, c) in projection k (0: y,z  1: z,x  2: x,y), 9 floats
// kind 3: orient2d(a, b, (p+q)/2) in projection k, 12 floats
// out[3 i] = complete form, out[3 i + 1] = integer path (2: not applicable,
// exponents too far apart), out[3 i + 2] = expansion path.
void ph_eval(int kind, int64_t n, const float *in, int k, int *out) {
  using namespace vacd;
  const int stride = kind == 0 ? 12 : (kind == 1 ? 15 : (kind == 2 ? 9 : 12));
  for (int64_t i = 0; i < n; ++i) {
    const float *x = in + stride * i;
    int s = 2, full = 0, ex = 0;
    switch (kind) {
      case 0:
        full = orient3d_exact(x, x + 3, x + 6, x + 9);
        if (!orient3d_int(x, x + 3, x + 6, x + 9, &s)) s = 2;
        ex = orient3d_expansion(x, x + 3, x + 6, x + 9);
        break;
      case 1:
        full = orient3d_mid_exact(x, x + 3, x + 6, x + 9, x + 12);
        if (!orient3d_mid_int(x, x + 3, x + 6, x + 9, x + 12, &s)) s = 2;
        ex = orient3d_mid_expansion(x, x + 3, x + 6, x + 9, x + 12);
        break;
      case 2:
        full = orient2d_exact(x, x + 3, x + 6, k);
        if (!orient2d_int(x, x + 3, x + 6, k, &s)) s = 2;
        ex = orient2d_expansion(x, x + 3, x + 6, k);
        break;
      default:
        full = orient2d_mid_exact(x, x + 3, x + 6, x + 9, k);
        if (!orient2d_mid_int(x, x + 3, x + 6, x + 9, k, &s)) s = 2;
        ex = orient2d_mid_expansion(x, x + 3, x + 6, x + 9, k);
        break;
    }
    out[3 * i] = full;
    out[3 * i + 1] = s;
    out[3 * i + 2] = ex;
  }
}

}  // extern "C"

#ifdef PH_MAIN
// Standalone check (built with the oracle and run under AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/test_predicates_host.py): random
// inputs of every class, each path of each predicate against the oracle.
#include <cstdio>
#include <random>
#include <vector>

extern "C" void orc_predicates_batch(int kind, int64_t n, const float *in, int k, int *out);

int main() {
  std::mt19937 rng(7);
  std::uniform_real_distribution<float> u(-1.0f, 1.0f);
  std::uniform_int_distribution<int> small(-8, 8), ex(-40, 12);
  long bad = 0, n_int = 0, n_exp_only = 0;
  for (int kind = 0; kind < 4; ++kind) {
    const int stride = kind == 0 ? 12 : (kind == 1 ? 15 : (kind == 2 ? 9 : 12));
    const int n = 30000;
    std::vector<float> in((size_t)n * stride);
    for (int i = 0; i < n; ++i) {
      float *x = &in[(size_t)i * stride];
      const int cls = i % 3;
      for (int t = 0; t < stride; ++t) {
        if (cls == 0) x[t] = u(rng);                                 // generic
        else if (cls == 1) x[t] = std::ldexp((float)small(rng), -3);  // lattice: exact zeros
        else x[t] = std::ldexp(u(rng), ex(rng));                      // exponents far apart
      }
    }
    for (int k = 0; k < (kind >= 2 ? 3 : 1); ++k) {
      std::vector<int> got((size_t)3 * n), want(n);
      ph_eval(kind, n, in.data(), k, got.data());
      orc_predicates_batch(kind, n, in.data(), k, want.data());
      for (int i = 0; i < n; ++i) {
        bad += got[3 * i] != want[i] || got[3 * i + 2] != want[i];
        if (got[3 * i + 1] != 2) ++n_int, bad += got[3 * i + 1] != want[i];
        else ++n_exp_only;
      }
    }
  }
  std::printf("mismatches %ld integer-path cases %ld expansion-only cases %ld\n", bad, n_int, n_exp_only);
  return bad == 0 && n_int > 0 && n_exp_only > 0 ? 0 : 1;
}
#endif
