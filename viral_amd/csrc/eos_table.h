// Host build of the tabulated s95p EOS read by the fast-mode kernels (the
// EosTab of physics.h).  Shared by the library and the host test harness.
#ifndef VIRAL_EOS_TABLE_H
#define VIRAL_EOS_TABLE_H

#include <cmath>
#include <vector>

#include "physics.h"

namespace viral {

// Host build of the tabulated s95p EOS (see EosTab in physics.h): octaves
// 2^-13 .. 2^13 fm^-4 (1.2e-4 .. 8192 fm^-4, i.e. up to ~1600 GeV/fm^3;
// below, the closed form's own exp(x)-1 cancellation is already ~1e-10, so
// those energies keep it), 1024 intervals per octave, 4-point Lagrange
// stencils (interpolation error ~1e-11 on the steepest branch); intervals whose
// stencil [e_{q-1}, e_{q+2}] contains a branch point of s95p_p/a/s/T (or the
// 1e-5 fm^-4 low-density switch) are flagged to use the closed form.
constexpr int kTabEmin = -13, kTabNoct = 26, kTabK = 1024;

inline void build_eos_table(std::vector<double>& v, std::vector<unsigned char>& closed)
{
  const int K = kTabK, W = K + 3;
  v.assign((size_t)kTabNoct * W * 4, 0.0);
  closed.assign((size_t)kTabNoct * K, 0);
  const double brk[] = {1e-5,
                        S95_E1 / GEVFM, S95_E2 / GEVFM, S95_E3 / GEVFM, S95_E4 / GEVFM,
                        S95_ES1 / GEVFM, S95_ES2 / GEVFM, S95_ES3 / GEVFM, S95_ES4 / GEVFM,
                        0.5143939846236409 / GEVFM};
  for (int o = 0; o < kTabNoct; o++) {
    const double base = ldexp(1.0, kTabEmin + o);
    for (int q = -1; q <= K + 1; q++) {
      const double e = base * (1.0 + (double)q / K);
      double out[4];
      s95p_all(e, out);
      for (int c = 0; c < 4; c++) v[((size_t)o * W + (q + 1)) * 4 + c] = out[c];
    }
    for (int q = 0; q < K; q++) {
      const double lo = base * (1.0 + (double)(q - 1) / K), hi = base * (1.0 + (double)(q + 2) / K);
      for (double b : brk)
        if (b >= lo && b <= hi) closed[(size_t)o * K + q] = 1;
    }
  }
}

}  // namespace viral

#endif
