// Drop-in replacement of the reference's TwoElectronInts.hpp (TwoElectronInts.hpp:79-110,
// with TwoElectronInts.cpp): same class name, constructor and methods, so the reference's
// drivers (Unomol.cc, RHF.hpp, UHF.hpp) compile unchanged against it.  The integral store
// becomes a device context of libunomol_hip.so; formGmatrix is a direct Fock build on the
// GPU that accumulates into the caller's packed G as the reference's does.  Nonzero status
// codes end the process through the reference's fatal_error (Util.cpp:6-9).
#ifndef UNOMOL_TwoElectronInts_HPP
#define UNOMOL_TwoElectronInts_HPP
#include <string>
#include <vector>
#include "Basis.hpp"
#include "Util.hpp"
#include "Stopwatch.hpp"
#include "unomol.h"

namespace unomol {

inline void uno_check(int rc) {
    if (rc != UNO_OK) fatal_error(uno_last_error());
}

class TwoElectronInts {
  public:
    TwoElectronInts() = delete;
    TwoElectronInts(const Basis& basis, int start_shell, const std::string& /*label*/) {
        view(basis);
        uno_check(uno_ctx_create(&b_, start_shell, &ctx_));
    }
    TwoElectronInts(const TwoElectronInts&) = delete;
    ~TwoElectronInts() { uno_ctx_destroy(ctx_); }
    void calculate(const Basis& basis) {
        view(basis);
        uno_check(uno_ctx_rebuild(ctx_, &b_));
    }
    void recalculate(const Basis& basis) { calculate(basis); }
    // G += 2J[P] - K[P] on packed arrays (TwoElectronInts.cpp:822-844)
    void formGmatrix(const double* P, double* G) { uno_check(uno_fock_rhf(ctx_, P, G)); }
    // Gs += J[PA+PB] - K[Ps] (TwoElectronInts.cpp:846-869)
    void formGmatrix(const double* PA, const double* PB, double* GA, double* GB) {
        uno_check(uno_fock_uhf(ctx_, PA, PB, GA, GB));
    }
    void directFormGMatrix(const double* P, double* G, const Basis&) { formGmatrix(P, G); }

  private:
    // uno_basis view of a reference Basis; its contraction coefficients are already
    // normalised (Basis.hpp:246), hence normalized = 1
    void view(const Basis& bas) {
        const int ns = bas.number_of_shells(), nc = bas.number_of_centers();
        l_.resize(ns);
        npr_.resize(ns);
        cen_.resize(ns);
        poff_.resize(ns);
        al_.clear();
        co_.clear();
        xyz_.resize(3 * nc);
        q_.resize(nc);
        const Shell* sh = bas.shell_ptr();
        for (int s = 0; s < ns; ++s) {
            l_[s] = sh[s].Lvalue();
            npr_[s] = sh[s].number_of_prims();
            cen_[s] = sh[s].center();
            poff_[s] = (int)al_.size();
            for (int p = 0; p < npr_[s]; ++p) {
                al_.push_back(sh[s].alf(p));
                co_.push_back(sh[s].cof(p));
            }
        }
        const Center* c = bas.center_ptr();
        for (int i = 0; i < nc; ++i) {
            q_[i] = c[i].charge();
            for (int k = 0; k < 3; ++k) xyz_[3 * i + k] = c[i].position(k);
        }
        b_ = uno_basis{ns, nc, l_.data(), npr_.data(), cen_.data(), poff_.data(), al_.data(), co_.data(),
                       xyz_.data(), q_.data(), 1};
    }
    uno_ctx* ctx_ = nullptr;
    uno_basis b_{};
    std::vector<int> l_, npr_, cen_, poff_;
    std::vector<double> al_, co_, xyz_, q_;
};

}  // namespace unomol
#endif
