// TEST INFRASTRUCTURE ONLY -- a harness of this project's own, linked against the
// reference's objects compiled from its sources (see Makefile, target "ref").
// It exposes the reference's per-quartet Rys ERIs and its one-electron matrices so
// the oracle and the GPU path can be checked against the reference itself, and a
// timed sample of the reference's ERI loop for the CPU baseline.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <cmath>
#include <algorithm>
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
// Returns the seconds spent, the number of function quartets evaluated, how many of
// them the reference would store, and the number of canonical shell quartets in the
// whole deck.
int ref_eri_sample(const char* deck, long nsample, unsigned long seed, double* seconds,
                   long* nfunc, double* total_shell_quartets, long* nstored) {
    Basis B{std::string(deck)};
    Ctx X(B);
    const int ns = B.number_of_shells();
    std::mt19937_64 rng(seed);
    std::uniform_int_distribution<int> pick(0, ns - 1);
    long nf = 0, nk = 0;
    auto t0 = std::chrono::steady_clock::now();
    for (long s = 0; s < nsample; ++s) {
        int i = pick(rng), j = pick(rng), k = pick(rng), l = pick(rng);
        if (j > i) std::swap(i, j);
        if (l > k) std::swap(k, l);
        if (k > i) { std::swap(i, k); std::swap(j, l); }
        if (j > i) std::swap(i, j);
        const int m = X.quartet(i, j, k, l);
        nf += m;
        // the records the reference would store (|val| > 1e-14, TwoElectronInts.cpp:668)
        for (int c = 0; c < m; ++c) nk += std::fabs(X.buf[c].val) > 1e-14 ? 1 : 0;
    }
    auto t1 = std::chrono::steady_clock::now();
    *seconds = std::chrono::duration<double>(t1 - t0).count();
    *nfunc = nf;
    if (nstored) *nstored = nk;
    *total_shell_quartets = (double)ns * (ns + 1) / 2 * ((double)ns * (ns + 1) / 2 + 1) / 2;
    return 0;
}

// The reference's own direct-store integral pass and Fock replay on a deck: builds the
// reference TwoElectronInts (every canonical ERI with |val| > 1e-14 into its in-memory
// store, TwoElectronInts.cpp:511-697) and then times `reps` calls of its
// formGmatrix(P, G) (TwoElectronInts.cpp:822-844) on a fixed symmetric density.  The store
// lives in the reference's 8 GiB in-memory cache (no disk spill for decks that fit it).
// store_label is the cache's file name (the reference removes it at exit).  Returns the
// seconds of the integral pass, the average seconds per G build and the stored ERI count.
int ref_fock_sample(const char* deck, const char* store_label, int reps, double* eri_seconds, double* g_seconds,
                    double* nstored) {
    Basis B{std::string(deck)};
    const int n = B.number_of_orbitals();
    const long n2 = (long)n * (n + 1) / 2;
    auto t0 = std::chrono::steady_clock::now();
    TwoElectronInts ti(B, 0, std::string(store_label));
    auto t1 = std::chrono::steady_clock::now();
    std::vector<double> P(n2), G(n2);
    for (long i = 0; i < n2; ++i) P[i] = 0.01 * (double)((i * 7919) % 101) / 101.0;
    auto t2 = std::chrono::steady_clock::now();
    for (int r = 0; r < reps; ++r) {
        std::fill(G.begin(), G.end(), 0.0);
        ti.formGmatrix(P.data(), G.data());
    }
    auto t3 = std::chrono::steady_clock::now();
    *eri_seconds = std::chrono::duration<double>(t1 - t0).count();
    *g_seconds = std::chrono::duration<double>(t3 - t2).count() / (reps > 0 ? reps : 1);
    // stored records: count the ERIs above the reference's 1e-14 drop
    Ctx X(B);
    const int ns = B.number_of_shells();
    double cnt = 0;
    for (int i = 0; i < ns; ++i)
        for (int j = 0; j <= i; ++j)
            for (int k = 0; k <= i; ++k)
                for (int l = 0; l <= k; ++l) {
                    int m = X.quartet(i, j, k, l);
                    for (int c = 0; c < m; ++c) cnt += std::fabs(X.buf[c].val) > 1e-14 ? 1 : 0;
                }
    *nstored = cnt;
    return 0;
}

} // extern "C"
