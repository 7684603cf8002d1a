// TEST INFRASTRUCTURE ONLY.  Exposes the reference's own Rys root/weight evaluation
// (Rys::calculate_roots, Rys.hpp:145-164, with root1..root5 of Rys.cpp:314-2197) to the
// fixture generators.  The reference keeps the results in private pointer members
// (Rys.hpp:44-49: Gx, Gy, Gz, binomial, roots, weights, in that order, in a class with no
// virtual functions), so this file reads them through a mirror of that member layout.
// Linked into oracle/_ref/libref_harness.so; never part of the product.
#include "Rys.hpp"

namespace {
struct RysLayout {
    double*** Gx;
    double*** Gy;
    double*** Gz;
    double** binomial;
    double* roots;
    double* weights;
};
}

extern "C" int ref_rys_roots(int n, double x, double* roots, double* weights) {
    static unomol::Rys* rys = nullptr;
    if (n < 1 || n > 5 || !(x >= 0)) return 1;
    if (!rys) rys = new unomol::Rys(2);   // maxr = 5
    rys->calculate_roots(x, n);
    const RysLayout* m = reinterpret_cast<const RysLayout*>(rys);
    for (int i = 0; i < n; ++i) {
        roots[i] = m->roots[i];       // the reference's u = t^2 / (1 - t^2)
        weights[i] = m->weights[i];
    }
    return 0;
}
