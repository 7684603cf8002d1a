// TEST INFRASTRUCTURE ONLY.  Replacement definition of the reference's
// Rys::root2 (Rys.cpp:413-631) that returns exact two-point Rys roots/weights from
// the oracle's quad-precision method.  Linked ahead of the reference's Rys.o (with
// --allow-multiple-definition) into libref_harness_x2.so, so every other step of the
// reference ERI code is its own while the 2-root quadrature is exact.  Used to pin
// the oracle's ERIs at 1e-12 in the 2-root classes, where the reference fit for
// 15 < x <= 40 (Rys.cpp:615-624) deviates by up to ~3.5e-4 relative.
#include "Rys.hpp"

extern "C" int orc_rys(int n, double x, double* u, double* w);

void unomol::Rys::root2(double x) noexcept {
    double u[2], w[2];
    orc_rys(2, x, u, w);
    for (int i = 0; i < 2; ++i) {
        roots[i] = u[i] / (1.0 - u[i]);   // the reference stores t^2/(1-t^2)
        weights[i] = w[i];
    }
}
