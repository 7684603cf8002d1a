// TEST INFRASTRUCTURE ONLY -- a harness of this project's own, linked against the
// reference's objects compiled from its sources (see Makefile, target "ref").
// It exposes the reference's per-quartet Rys ERIs and its one-electron matrices so
// the oracle and the GPU path can be checked against the reference itself, and a
// timed sample of the reference's ERI loop for the CPU baseline.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <random>
#include <vector>
#include "Basis.hpp"
#include "TwoElectronInts.hpp"
#include "OneElectronInts.hpp"

namespace unomol {
// defined in TwoElectronInts.cpp (no header declares it)
void calc_two_electron_ints_rys(const ShellQuartet& sq, const AuxFunctions& aux, Rys& rys,
                                TwoInts* sints);
}

using namespace unomol;

namespace {

struct Ctx {
    const Basis& B;
    const Shell* sh;
    const Center* cen;
    const AuxFunctions& aux;
    Rys rys;
    ShellQuartet sq;
    std::vector<TwoInts> buf;
    explicit Ctx(const Basis& b)
        : B(b), sh(b.shell_ptr()), cen(b.center_ptr()), aux(*b.auxfun_ptr()), rys(b.maxLvalue()),
          sq(b.maxLvalue()) {
        int m = aux.maxLstates();
        buf.resize((size_t)m * m * m * m);
    }
    // fills buf with the canonical function quartets of shell quartet (i,j,k,l) and
    // evaluates them with the reference kernel; returns the count
    int quartet(int ish, int jsh, int ksh, int lsh) {
        const Shell* S[4] = {sh + ish, sh + jsh, sh + ksh, sh + lsh};
        int o[4] = {B.offset(ish), B.offset(jsh), B.offset(ksh), B.offset(lsh)};
        int l[4], n[4];
        for (int x = 0; x < 4; ++x) { l[x] = S[x]->Lvalue(); n[x] = aux.number_of_lstates(l[x]); }
        bool sw12 = l[0] < l[1], sw34 = l[2] < l[3];
        const Shell* s1 = sw12 ? S[1] : S[0];
        const Shell* s2 = sw12 ? S[0] : S[1];
        const Shell* s3 = sw34 ? S[3] : S[2];
        const Shell* s4 = sw34 ? S[2] : S[3];
        sq.npr1 = s1->number_of_prims(); sq.lv1 = s1->Lvalue(); sq.al1 = s1->alf_ptr(); sq.co1 = s1->cof_ptr();
        sq.npr2 = s2->number_of_prims(); sq.lv2 = s2->Lvalue(); sq.al2 = s2->alf_ptr(); sq.co2 = s2->cof_ptr();
        sq.npr3 = s3->number_of_prims(); sq.lv3 = s3->Lvalue(); sq.al3 = s3->alf_ptr(); sq.co3 = s3->cof_ptr();
        sq.npr4 = s4->number_of_prims(); sq.lv4 = s4->Lvalue(); sq.al4 = s4->alf_ptr(); sq.co4 = s4->cof_ptr();
        sq.a = cen[s1->center()].r_vec(); sq.b = cen[s2->center()].r_vec();
        sq.c = cen[s3->center()].r_vec(); sq.d = cen[s4->center()].r_vec();
        sq.ab2 = dist_sqr(sq.a, sq.b);
        sq.cd2 = dist_sqr(sq.c, sq.d);
        int knt = 0;
        for (int a = 0; a < n[0]; ++a) {
            int ir = o[0] + a;
            for (int b = 0; b < n[1]; ++b) {
                int jr = o[1] + b;
                if (jr > ir) break;
                for (int c = 0; c < n[2]; ++c) {
                    int kr = o[2] + c;
                    if (kr > ir) break;
                    for (int d = 0; d < n[3]; ++d) {
                        int lr = o[3] + d;
                        if (lr > kr || (ir == kr && lr > jr)) break;
                        TwoInts& t = buf[knt];
                        t.val = 0.0; t.i = ir; t.j = jr; t.k = kr; t.l = lr;
                        unsigned l12 = sw12 ? ((unsigned)b << 4) + a : ((unsigned)a << 4) + b;
                        unsigned l34 = sw34 ? ((unsigned)d << 4) + c : ((unsigned)c << 4) + d;
                        sq.lstates[knt] = (l12 << 8) + l34;
                        sq.norms[knt] = aux.normalization_factor(l[0], a) * aux.normalization_factor(l[1], b) *
                                        aux.normalization_factor(l[2], c) * aux.normalization_factor(l[3], d);
                        ++knt;
                    }
                }
            }
        }
        sq.len = knt;
        if (knt) calc_two_electron_ints_rys(sq, aux, rys, buf.data());
        return knt;
    }
};

} // namespace

extern "C" {

// every canonical ERI of the deck (no |val| drop), reference order
int ref_eri_all(const char* deck, long cap, double* val, int* ijkl, long* count) {
    Basis B{std::string(deck)};
    Ctx X(B);
    const int ns = B.number_of_shells();
    long n = 0;
    for (int i = 0; i < ns; ++i)
        for (int j = 0; j <= i; ++j)
            for (int k = 0; k <= i; ++k)
                for (int l = 0; l <= k; ++l) {
                    int m = X.quartet(i, j, k, l);
                    for (int c = 0; c < m; ++c, ++n)
                        if (n < cap) {
                            val[n] = X.buf[c].val;
                            ijkl[4 * n] = X.buf[c].i; ijkl[4 * n + 1] = X.buf[c].j;
                            ijkl[4 * n + 2] = X.buf[c].k; ijkl[4 * n + 3] = X.buf[c].l;
                        }
                }
    *count = n;
    return 0;
}

// one-electron matrices (packed) exactly as the reference computes them
int ref_oneint(const char* deck, double* S, double* T, double* H) {
    Basis B{std::string(deck)};
    OneElectronInts(B, S, T, H);
    return 0;
}

// Times the reference ERI kernel over a random sample of shell quartets of the deck
// (canonical quartets drawn uniformly with replacement by the given seed).
// Returns the seconds spent, the number of function quartets evaluated and the
// number of canonical shell quartets in the whole deck.
int ref_eri_sample(const char* deck, long nsample, unsigned long seed, double* seconds,
                   long* nfunc, double* total_shell_quartets) {
    Basis B{std::string(deck)};
    Ctx X(B);
    const int ns = B.number_of_shells();
    std::mt19937_64 rng(seed);
    std::uniform_int_distribution<int> pick(0, ns - 1);
    long nf = 0;
    auto t0 = std::chrono::steady_clock::now();
    for (long s = 0; s < nsample; ++s) {
        int i = pick(rng), j = pick(rng), k = pick(rng), l = pick(rng);
        if (j > i) std::swap(i, j);
        if (l > k) std::swap(k, l);
        if (k > i) { std::swap(i, k); std::swap(j, l); }
        if (j > i) std::swap(i, j);
        nf += X.quartet(i, j, k, l);
    }
    auto t1 = std::chrono::steady_clock::now();
    *seconds = std::chrono::duration<double>(t1 - t0).count();
    *nfunc = nf;
    double m = (double)ns * (ns + 1) / 2;
    *total_shell_quartets = (double)ns * (ns + 1) / 2 * ((double)ns * (ns + 1) / 2 + 1) / 2;
    (void)m;
    return 0;
}

} // extern "C"
