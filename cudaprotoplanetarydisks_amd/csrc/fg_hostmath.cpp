// Host-side transcendental tables of the device kernels.  Built with
// -fno-builtin-sin -fno-builtin-cos (as is the oracle) so that cos and sin
// are separate libm calls: a fused sincos can differ from sin in the last
// bit, and the tables must hold the same values on both sides.
#include <cmath>

#include "fargo_hip.h"

// twiddles of one level of the mixed-radix FFT: cs[2q], cs[2q+1] =
// cos, sin of sign * 2 pi q / L for q < L (the angle expression of the
// oracle's FFT, with pi = 3.14159265358979323844, Fondam.cuh:3)
extern "C" void fg_fft_twiddles(int L, int sign, double *cs) {
  const double PI = 3.14159265358979323844;
  for (int q = 0; q < L; q++) {
    double ang = sign * 2.0 * PI * (double)q / (double)L;
    cs[2 * q] = cos(ang);
    cs[2 * q + 1] = sin(ang);
  }
}
