// Context build (shell pairs, primitive pairs, Schwarz bounds, quartet spaces) and
// the direct Fock build.  Replaces the reference's store-then-replay design
// (TwoElectronInts.cpp:511-697 computes every unique ERI once into an 8 GiB buffer
// spilling to MINTS.DAT; formGmatrix replays it each iteration, :822-869) with a direct
// build: every iteration recomputes the significant shell quartets on the GPU and
// contracts them with the density on the fly.
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <cmath>
#include <numeric>
#include "engine.hpp"
#include "../hip/kernels.hpp"

namespace uno {

static thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }
const std::string& get_error() { return g_err; }

static constexpr double kTwoPi25 = 34.986836655249725;

static int ensure_device() {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        set_error("no HIP device is visible (the engine needs a gfx950 GPU)");
        return UNO_EDEVICE;
    }
    return UNO_OK;
}

// Quartet spaces of one launch.  Both runs are sorted by Schwarz bound (descending), so
// for bra i the kets passing Qi Qj dmax >= tau form a prefix of the ket run, and within
// it the kets with Qi Qj >= fp32_thr (FP64 in a mixed build) form a shorter prefix.
void build_prefix(uno_ctx* c, QuartetClass& q, double dmax, double fp32_thr, double tau) {
    const PairClassList& Lb = c->cls[q.X];
    const PairClassList& Lk = c->cls[q.Y];
    const int nb = (int)Lb.Q.size();
    const bool mixed = fp32_thr > 0.0;
    q.s64.prefix.assign(nb + 1, 0);
    q.s32.prefix.assign(mixed ? nb + 1 : 1, 0);
    q.s32.kstart.assign(mixed ? nb : 0, 0);
    // both runs are sorted by bound, descending: the thresholds grow with the bra index, so
    // the counts of kets at or above them only shrink, and two pointers walk the ket run
    // once (the same counts as a binary search per bra)
    const long long nk = (long long)Lk.Q.size();
    long long p_all = nk, p32 = nk;
    for (int i = 0; i < nb; ++i) {
        long long cnt = 0, c64 = 0;
        const double qd = Lb.Q[i] * dmax;
        if (qd > 0.0) {
            const double thr = tau / qd;
            while (p_all > 0 && !(Lk.Q[p_all - 1] >= thr)) --p_all;
            cnt = p_all;
            if (q.X == q.Y) cnt = std::min<long long>(cnt, i + 1);
        } else {
            p_all = 0;
        }
        c64 = cnt;
        if (mixed && cnt > 0) {
            const double t32 = fp32_thr / Lb.Q[i];
            while (p32 > 0 && !(Lk.Q[p32 - 1] >= t32)) --p32;
            c64 = std::min(p32, cnt);
        }
        q.s64.prefix[i + 1] = q.s64.prefix[i] + c64;
        if (mixed) {
            q.s32.prefix[i + 1] = q.s32.prefix[i] + (cnt - c64);
            q.s32.kstart[i] = (int)c64;
        }
    }
    q.s64.nq = q.s64.prefix[nb];
    q.s32.nq = mixed ? q.s32.prefix[nb] : 0;
    // drop the empty bras at either end of each space (with a small density most bras of
    // an incremental build have no ket left): the arrays shrink, b0 records the offset
    auto trim = [](QuartetSpace& sp, bool has_kstart) {
        sp.b0 = 0;
        const int nbf = (int)sp.prefix.size() - 1;
        int lo = 0, hi = nbf;
        while (lo < hi && sp.prefix[lo + 1] == sp.prefix[lo]) ++lo;
        while (hi > lo && sp.prefix[hi] == sp.prefix[hi - 1]) --hi;
        if (lo == 0 && hi == nbf) return;
        std::vector<long long> np(sp.prefix.begin() + lo, sp.prefix.begin() + hi + 1);
        const long long base = np[0];
        for (long long& v : np) v -= base;
        sp.prefix.swap(np);
        if (has_kstart) sp.kstart = std::vector<int>(sp.kstart.begin() + lo, sp.kstart.begin() + hi);
        sp.b0 = lo;
    };
    trim(q.s64, false);
    if (mixed) trim(q.s32, true);
}

// wave tasks: bra i owns ceil(count_i / 64) consecutive tasks from wpre[i]; tblk[b] is
// the bra of task 64 b, bracketing each wave's bra search
// row-merged block tasks for the bras with at least min_kets kets (none if min_kets <= 0):
// rbra lists those bras; a task is a group of `group` consecutive ones (the last group may
// be smaller) and one segment of seg kets of the group's largest ket count, rpre[g] the
// first task of group g
void build_rows(QuartetSpace& s, long long min_kets, int seg, int group) {
    const int nb = (int)s.prefix.size() - 1;
    s.rbra.clear();
    s.rpre.assign(1, 0);
    if (min_kets > 0) {
        for (int i = 0; i < nb; ++i)
            if (s.prefix[i + 1] - s.prefix[i] >= min_kets) s.rbra.push_back(i);
        for (size_t g = 0; g < s.rbra.size(); g += group) {
            long long cmax = 0;
            for (size_t k = g; k < std::min(s.rbra.size(), g + group); ++k)
                cmax = std::max(cmax, s.prefix[s.rbra[k] + 1] - s.prefix[s.rbra[k]]);
            s.rpre.push_back(s.rpre.back() + (cmax + seg - 1) / seg);
        }
    }
    s.nrtask = s.rpre.back();
}

// wave tasks (bra, chunk of 64 kets) for the bras not run as row tasks
void build_tasks(QuartetSpace& s, long long row_min_kets) {
    const int nb = (int)s.prefix.size() - 1;
    s.wpre.assign(nb + 1, 0);
    for (int i = 0; i < nb; ++i) {
        const long long cnt = s.prefix[i + 1] - s.prefix[i];
        s.wpre[i + 1] = s.wpre[i] + (row_min_kets > 0 && cnt >= row_min_kets ? 0 : (cnt + 63) / 64);
    }
    s.ntask = s.wpre[nb];
    const long long nblk = (s.ntask + 63) / 64;
    s.tblk.resize(nblk + 1);
    int i = 0;
    for (long long b = 0; b <= nblk; ++b) {
        const long long t = std::min(b * 64, s.ntask - 1);
        while (s.wpre[i + 1] <= t) ++i;
        s.tblk[b] = i;
    }
}

// block table: bra holding the first quartet of every block, so that each lane's bra
// search covers a short range
void build_blocks(QuartetSpace& s, int qpb) {
    const long long nblk = (s.nq + qpb - 1) / qpb;
    s.blk.resize(nblk + 1);
    int i = 0;
    for (long long b = 0; b <= nblk; ++b) {
        const long long g = std::min(b * qpb, s.nq - 1);
        while (s.prefix[i + 1] <= g) ++i;
        s.blk[b] = i;
    }
}

// tuned row-merge configurations, shared by every context of the same layout (basis size,
// shells, pairs and quartet spaces) in this process
static std::map<std::tuple<int, int, int, int>, std::map<std::tuple<int, int, int, int>, TuneChoice>> g_tuning;

static std::tuple<int, int, int, int> tuning_key(const uno_ctx* c) {
    return std::make_tuple(c->n, c->nsh, (int)c->pairs.size(), (int)c->qcls.size());
}

void remember_tuning(uno_ctx* c) {
    auto& m = g_tuning[tuning_key(c)];
    for (const auto& kv : c->tune) m[kv.first] = kv.second;
}

static void recall_tuning(uno_ctx* c) {
    auto it = g_tuning.find(tuning_key(c));
    c->tune.clear();
    if (it != g_tuning.end()) c->tune = it->second;
}

EriArgs base_args(uno_ctx* c) {
    EriArgs a{};
    a.pairs = c->d_pairs.p;
    a.prims = c->d_prims.p;
    a.tau = c->tau;
    a.primcut = c->primcut;
    a.n = c->n;
    a.nsh = c->nsh;
    a.rys2_exact = c->rys2_exact;
    return a;
}

int ctx_build(uno_ctx* c, const uno_basis* b, int start_shell) {
    using bclk = std::chrono::steady_clock;
    auto since = [](bclk::time_point t) { return std::chrono::duration<double, std::milli>(bclk::now() - t).count(); };
    const auto b0 = bclk::now();
    if (int rc = ensure_device()) return rc;
    std::string err;
    if (!make_basis(b, c->B, err)) {
        set_error(err);
        return UNO_EINVAL;
    }
    if (c->B.maxl > kMaxL) {
        set_error("the device ERI kernels cover s, p and d shells (l <= 2); this basis has l = " +
                  std::to_string(c->B.maxl));
        return UNO_EINVAL;
    }
    const Basis& B = c->B;
    c->n = B.norb;
    c->nsh = (int)B.sh.size();
    c->start_shell = start_shell;
    if (!c->st) UNO_HIP(hipStreamCreate(&c->st));
    hipStream_t st = c->st;
    const int nsh = c->nsh;
    // smallest exponent sum a ket pair can have: bounds the bare prefactor from above
    double amin = 1e300;
    for (const Shell& s : B.sh)
        for (double a : s.a) amin = std::min(amin, a);
    const double qmin = 2.0 * amin;
    c->pairs.clear();
    c->prims.clear();
    c->pair_of.assign((size_t)nsh * nsh, -1);
    // shell pairs (a, b <= a), built per first shell a in parallel into lists of their own,
    // then laid out in the serial order (a, then b, ascending) at prefix-sum offsets
    std::vector<std::vector<PairHdr>> hpa(nsh);
    std::vector<std::vector<PrimPair>> ppa(nsh);
    std::vector<std::vector<int>> bpa(nsh);
#pragma omp parallel for schedule(dynamic, 4)
    for (int a = 0; a < nsh; ++a) {
        std::vector<PairHdr>& hv = hpa[a];
        std::vector<PrimPair>& pv = ppa[a];
        for (int bsh = 0; bsh <= a; ++bsh) {
            const int s1 = (B.sh[a].l >= B.sh[bsh].l) ? a : bsh;
            const int s2 = (s1 == a) ? bsh : a;
            const Shell& S1 = B.sh[s1];
            const Shell& S2 = B.sh[s2];
            const double* A = &B.xyz[3 * S1.cen];
            const double* Bc = &B.xyz[3 * S2.cen];
            PairHdr h{};
            double ab2 = 0.0;
            for (int x = 0; x < 3; ++x) {
                h.A[x] = A[x];
                h.AB[x] = A[x] - Bc[x];
                ab2 += h.AB[x] * h.AB[x];
            }
            h.f1 = B.off[s1];
            h.f2 = B.off[s2];
            h.prim0 = (int)pv.size();   // within this shell's list; made global below
            h.same = (a == bsh);
            h.sh1 = s1;
            h.sh2 = s2;
            for (int i = 0; i < S1.npr; ++i) {
                // same shell: lower triangle of primitive pairs, off-diagonal doubled
                // (TwoElectronInts.cpp:444-449)
                const int jend = h.same ? i + 1 : S2.npr;
                for (int j = 0; j < jend; ++j) {
                    const double al = S1.a[i], be = S2.a[j], p = al + be, ip = 1.0 / p;
                    const double s = std::exp(-al * be * ip * ab2);
                    // every quartet with this primitive pair has a bare prefactor below
                    // this bound; drop it when the reference cut would reject them all
                    const double bound = kTwoPi25 * s * ip / (qmin * std::sqrt(p + qmin));
                    if (bound < c->primcut) continue;
                    PrimPair pp{};
                    pp.p = p;
                    pp.ip = ip;
                    for (int x = 0; x < 3; ++x) pp.P[x] = (al * A[x] + be * Bc[x]) * ip;
                    pp.K = S1.c[i] * S2.c[j] * ((h.same && i != j) ? 2.0 : 1.0) * s;
                    pp.s = s;
                    pv.push_back(pp);
                }
            }
            h.nprim = (int)pv.size() - h.prim0;
            if (h.nprim == 0) continue;
            hv.push_back(h);
            bpa[a].push_back(bsh);
        }
    }
    {
        std::vector<long long> hoff(nsh + 1, 0), poff(nsh + 1, 0);
        for (int a = 0; a < nsh; ++a) {
            hoff[a + 1] = hoff[a] + (long long)hpa[a].size();
            poff[a + 1] = poff[a] + (long long)ppa[a].size();
        }
        if (poff[nsh] > (long long)INT32_MAX) {
            set_error("too many primitive pairs for 32-bit offsets");
            return UNO_EINVAL;
        }
        c->pairs.resize((size_t)hoff[nsh]);
        c->prims.resize((size_t)poff[nsh]);
#pragma omp parallel for schedule(dynamic, 4)
        for (int a = 0; a < nsh; ++a) {
            for (size_t k = 0; k < hpa[a].size(); ++k) {
                PairHdr h = hpa[a][k];
                h.prim0 += (int)poff[a];
                c->pairs[(size_t)hoff[a] + k] = h;
                c->pair_of[(size_t)a * nsh + bpa[a][k]] = (int)(hoff[a] + (long long)k);
            }
            std::copy(ppa[a].begin(), ppa[a].end(), c->prims.begin() + poff[a]);
        }
    }
    const int np = (int)c->pairs.size();
    UNO_HIP(c->d_pairs.upload(c->pairs.data(), np, st));
    UNO_HIP(c->d_prims.upload(c->prims.data(), c->prims.size(), st));
    c->ctx_ms[0] = since(b0);
    const auto b1 = bclk::now();
    // Schwarz bounds Q_ab = sqrt(max |(ij|ij)|) from diagonal quartets, per pair class
    c->pairQ.assign(np, 0.0);
    DBuf<double> dq;
    UNO_HIP(dq.alloc(np));
    std::vector<int> pcls(np);
    for (int p = 0; p < np; ++p) pcls[p] = pair_class(B.sh[c->pairs[p].sh1].l, B.sh[c->pairs[p].sh2].l);
    for (int X = 0; X < kNumPairClasses; ++X) {
        std::vector<int2> ql;
        for (int p = 0; p < np; ++p)
            if (pcls[p] == X) ql.push_back(make_int2(p, p));
        if (ql.empty()) continue;
        DBuf<int2> dql;
        UNO_HIP(dql.upload(ql.data(), ql.size(), st));
        EriArgs a = base_args(c);
        a.explicit_q = dql.p;
        a.nquart = (long long)ql.size();
        a.out = dq.p;
        // bounds from the uncut integrals: with the primitive cut a pair of distant
        // tight primitives can lose its whole diagonal quartet while its quartets with
        // diffuse partners survive the cut, so a cut diagonal would under-bound them
        a.primcut = 0.0;
        UNO_HIP(eri_launch(kDigestSchwarz, X, X, a, st));
        UNO_HIP(hipStreamSynchronize(st));
    }
    UNO_HIP(hipMemcpyAsync(c->pairQ.data(), dq.p, np * sizeof(double), hipMemcpyDeviceToHost, st));
    UNO_HIP(hipStreamSynchronize(st));
    // Reorder the pair and primitive arrays into runs of (angular class, primitive-count
    // bucket), each sorted by Schwarz bound: a launch then walks contiguous memory, and the
    // lanes of a wave see alike primitive loops.
    c->ctx_ms[1] = since(b1);
    const auto b2 = bclk::now();
    // run key of every pair, taken before the arrays are permuted
    std::vector<int> runkey(np);
    for (int p = 0; p < np; ++p) runkey[p] = pcls[p] * kNumBuckets + prim_bucket(c->pairs[p].nprim);
    std::vector<int> order(np);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        if (runkey[x] != runkey[y]) return runkey[x] < runkey[y];
        return c->pairQ[x] > c->pairQ[y];
    });
    // new primitive offsets and pair-major Coulomb offsets (each pair's function block), in
    // run order
    static const int kPairComp[kNumPairClasses] = {1, 3, 9, 6, 18, 36};
    std::vector<int> nprim0(np), njo(np);
    {
        long long tp = 0, jo = 0;
        for (int k = 0; k < np; ++k) {
            nprim0[k] = (int)tp;
            tp += c->pairs[order[k]].nprim;
            njo[k] = (int)jo;
            jo += kPairComp[runkey[order[k]] / kNumBuckets];
        }
        c->njp = jo;
        UNO_HIP(c->JP.alloc((size_t)std::max(jo, 1LL)));
    }
    std::vector<PairHdr> pairs2(np);
    std::vector<PrimPair> prims2(c->prims.size());
    std::vector<double> q2(np);
    std::vector<int> newidx(np);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < np; ++k) {
        PairHdr h = c->pairs[order[k]];
        const int p0 = h.prim0;
        h.prim0 = nprim0[k];
        h.jo = njo[k];
        std::copy(c->prims.begin() + p0, c->prims.begin() + p0 + h.nprim, prims2.begin() + h.prim0);
        pairs2[k] = h;
        q2[k] = c->pairQ[order[k]];
        newidx[order[k]] = k;
    }
    for (int& p : c->pair_of)
        if (p >= 0) p = newidx[p];
    c->pairs.swap(pairs2);
    c->prims.swap(prims2);
    c->pairQ.swap(q2);
    // the device copies: the same permutation applied to the arrays uploaded above (the
    // three index arrays go up instead of the records themselves)
    {
        DBuf<int> dorder, dp0, djo;
        DBuf<PairHdr> np_d;
        DBuf<PrimPair> nq_d;
        UNO_HIP(dorder.upload(order.data(), np, st));
        UNO_HIP(dp0.upload(nprim0.data(), np, st));
        UNO_HIP(djo.upload(njo.data(), np, st));
        UNO_HIP(np_d.alloc(np));
        UNO_HIP(nq_d.alloc(c->prims.size()));
        UNO_HIP(permute_pairs(np, dorder.p, dp0.p, djo.p, c->d_pairs.p, c->d_prims.p, np_d.p, nq_d.p, st));
        UNO_HIP(hipStreamSynchronize(st));
        c->d_pairs = std::move(np_d);
        c->d_prims = std::move(nq_d);
    }
    UNO_HIP(c->d_pairQ.upload(c->pairQ.data(), np, st));
    c->cls.clear();
    for (int k = 0; k < np;) {
        const int r = runkey[order[k]];
        PairClassList L;
        L.lcls = r / kNumBuckets;
        L.bucket = r % kNumBuckets;
        L.off = k;
        while (k < np && runkey[order[k]] == r) L.Q.push_back(c->pairQ[k++]);
        L.count = k - L.off;
        c->cls.push_back(std::move(L));
    }
    // one run per (angular class, bucket) at most, each sorted by bound
    if ((int)c->cls.size() > kNumPairClasses * kNumBuckets) {
        set_error("internal error: pair runs are not grouped (" + std::to_string(c->cls.size()) + " runs)");
        return UNO_EINVAL;
    }
    for (const PairClassList& L : c->cls)
        for (size_t i = 1; i < L.Q.size(); ++i)
            if (L.Q[i] > L.Q[i - 1]) {
                set_error("internal error: pair run is not sorted by Schwarz bound");
                return UNO_EINVAL;
            }
    c->ctx_ms[2] = since(b2);
    const auto b3 = bclk::now();
    // columns of the row-merged launches' LDS rows: per ket run, the functions its pairs
    // touch (whole shells), in function order
    {
        static const int kPairL[kNumPairClasses][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}};
        const int nr = (int)c->cls.size(), n = c->n;
        std::vector<int> cmap((size_t)nr * n, -1), cinv((size_t)nr * n, 0);
        c->run_ncols.assign(nr, 0);
        for (int r = 0; r < nr; ++r) {
            const PairClassList& L = c->cls[r];
            const int n1 = (kPairL[L.lcls][0] + 1) * (kPairL[L.lcls][0] + 2) / 2;
            const int n2 = (kPairL[L.lcls][1] + 1) * (kPairL[L.lcls][1] + 2) / 2;
            std::vector<char> used(n, 0);
            for (int k = L.off; k < L.off + L.count; ++k) {
                for (int f = 0; f < n1; ++f) used[c->pairs[k].f1 + f] = 1;
                for (int f = 0; f < n2; ++f) used[c->pairs[k].f2 + f] = 1;
            }
            int m = 0;
            for (int f = 0; f < n; ++f)
                if (used[f]) {
                    cmap[(size_t)r * n + f] = m;
                    cinv[(size_t)r * n + m] = f;
                    ++m;
                }
            c->run_ncols[r] = m;
        }
        UNO_HIP(c->d_colmap.upload(cmap.data(), cmap.size(), st));
        UNO_HIP(c->d_colinv.upload(cinv.data(), cinv.size(), st));
    }
    c->qcls.clear();
    for (int X = 0; X < (int)c->cls.size(); ++X)
        for (int Y = 0; Y <= X; ++Y) {
            QuartetClass q;
            q.X = X;
            q.Y = Y;
            c->qcls.push_back(std::move(q));
        }
    recall_tuning(c);
    std::vector<int> shoff(nsh), shn(nsh);
    for (int s = 0; s < nsh; ++s) {
        shoff[s] = B.off[s];
        shn[s] = ncart(B.sh[s].l);
    }
    UNO_HIP(c->d_shoff.upload(shoff.data(), nsh, st));
    UNO_HIP(c->d_shn.upload(shn.data(), nsh, st));
    // one-electron data: shells with their normalised primitives, centres, and the shell
    // pairs of each class with the higher-l shell first
    {
        std::vector<OneIntShell> hs(nsh);
        std::vector<double> al, co;
        for (int s = 0; s < nsh; ++s) {
            hs[s] = OneIntShell{B.sh[s].l, B.sh[s].npr, B.sh[s].cen, (int)al.size(), B.off[s]};
            al.insert(al.end(), B.sh[s].a.begin(), B.sh[s].a.end());
            co.insert(co.end(), B.sh[s].c.begin(), B.sh[s].c.end());
        }
        std::vector<std::vector<int2>> pl(kNumPairClasses);
        for (int a = 0; a < nsh; ++a)
            for (int bs = 0; bs <= a; ++bs) {
                const int s1 = B.sh[a].l >= B.sh[bs].l ? a : bs, s2 = s1 == a ? bs : a;
                pl[pair_class(B.sh[s1].l, B.sh[s2].l)].push_back(int2{s1, s2});
            }
        UNO_HIP(c->oi_shells.upload(hs.data(), hs.size(), st));
        UNO_HIP(c->oi_alpha.upload(al.data(), al.size(), st));
        UNO_HIP(c->oi_coef.upload(co.data(), co.size(), st));
        UNO_HIP(c->oi_xyz.upload(B.xyz.data(), B.xyz.size(), st));
        UNO_HIP(c->oi_charge.upload(B.q.data(), B.q.size(), st));
        for (int k = 0; k < kNumPairClasses; ++k) {
            c->oi_npair[k] = (int)pl[k].size();
            UNO_HIP(c->oi_pairs[k].upload(pl[k].data(), pl[k].size(), st));
        }
    }
    const size_t nn = (size_t)c->n * c->n;
    UNO_HIP(c->D0.alloc(nn));
    UNO_HIP(c->D1.alloc(nn));
    UNO_HIP(c->D2.alloc(nn));
    UNO_HIP(c->W0.alloc(nn));
    UNO_HIP(c->W1.alloc(nn));
    UNO_HIP(c->Dsh.alloc((size_t)nsh * nsh));
    UNO_HIP(c->pk.alloc((size_t)c->n * (c->n + 1)));
    UNO_HIP(c->scal.alloc(16));
    UNO_HIP(c->red.alloc(kRedBlocks));
    UNO_HIP(hipStreamSynchronize(st));
    c->ctx_ms[3] = since(b3);
    return UNO_OK;
}

int oneint_device(uno_ctx* c, double* S, double* T, double* H) {
    OneIntArgs a{};
    a.shells = c->oi_shells.p;
    a.alpha = c->oi_alpha.p;
    a.coef = c->oi_coef.p;
    a.xyz = c->oi_xyz.p;
    a.charge = c->oi_charge.p;
    a.ncen = c->B.ncen;
    a.n = c->n;
    a.skip = c->skip_center;
    a.S = S;
    a.T = T;
    a.H = H;
    static const int kl[kNumPairClasses][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}};
    for (int k = 0; k < kNumPairClasses; ++k)
        UNO_HIP(oneint_launch(kl[k][0], kl[k][1], c->oi_npair[k], c->oi_pairs[k].p, a, c->st));
    return UNO_OK;
}

int fock_build(uno_ctx* c, bool uhf, double dmax_hint, bool mixed) {
    using hclk = std::chrono::steady_clock;
    const auto h0 = hclk::now();
    auto ms_since = [](hclk::time_point t) { return std::chrono::duration<double, std::milli>(hclk::now() - t).count(); };
    hipStream_t st = c->st;
    const long nn = (long)c->n * c->n;
    // deterministic builds: W holds two int64 words per entry (the second nn entries on),
    // every bra runs as wave tasks and the kets' Coulomb outputs go straight to W
    const bool det = c->deterministic != 0;
    const bool rows_on = c->row_merge && !det;
    const long wlen = det ? 2 * nn : nn;
    if (det) {
        UNO_HIP(c->W0.alloc(wlen));
        UNO_HIP(c->W1.alloc(wlen));
    }
    UNO_HIP(k_fill(wlen, 0.0, c->W0.p, st));
    if (uhf) UNO_HIP(k_fill(wlen, 0.0, c->W1.p, st));
    UNO_HIP(k_fill(c->njp, 0.0, c->JP.p, st));
    // shell-block density maxima for the per-quartet screen, and the global bound
    UNO_HIP(k_shell_max(c->n, c->nsh, c->d_shoff.p, c->d_shn.p, c->D0.p, uhf ? c->D1.p : nullptr,
                        uhf ? c->D2.p : nullptr, c->Dsh.p, st));
    double dmax = dmax_hint;
    if (dmax < 0.0) {
        UNO_HIP(k_absmax((long)c->nsh * c->nsh, c->Dsh.p, c->scal.p, st));
        UNO_HIP(hipMemcpyAsync(&dmax, c->scal.p, sizeof(double), hipMemcpyDeviceToHost, st));
        UNO_HIP(hipStreamSynchronize(st));
    }
    c->last_ms[0] = ms_since(h0);
    const auto h1 = hclk::now();
    double total = 0.0, total32 = 0.0;
    const int mode = uhf ? kDigestUHF : kDigestRHF;

    // Pass 1 (host): every non-empty space's configuration and index arrays, appended to one
    // staging buffer (uploaded once per build; the tuner uploads per candidate instead)
    struct SpaceRec {
        QuartetClass* q;
        int prec, variant, dlds;
        bool tasks;
        size_t o_prefix, o_wpre, o_tblk, o_rbra, o_rpre, o_blk, o_kstart;
    };
    std::vector<SpaceRec> recs;
    std::vector<char>& hs = c->stage_h;
    hs.clear();
    auto put = [&](const auto& v) -> size_t {
        const size_t off = (hs.size() + 15) / 16 * 16;
        const size_t bytes = v.size() * sizeof(v[0]);
        hs.resize(off + bytes);
        if (bytes) std::memcpy(hs.data() + off, v.data(), bytes);
        return off;
    };
    // the spaces are independent: their index arrays are built in parallel, then appended
    // to the staging buffer in class order
    const int nqc = (int)c->qcls.size();
    std::vector<SpaceRec> built(2 * (size_t)nqc);
    std::vector<char> used(2 * (size_t)nqc, 0);
#pragma omp parallel for schedule(dynamic, 4)
    for (int qi = 0; qi < nqc; ++qi) {
        QuartetClass& q = c->qcls[qi];
        build_prefix(c, q, 4.0 * dmax, mixed ? c->fp32_thr : 0.0, c->tau);
        const PairClassList& LX = c->cls[q.X];
        const PairClassList& LY = c->cls[q.Y];
        for (int prec = 0; prec < 2; ++prec) {
            QuartetSpace& sp = prec ? q.s32 : q.s64;
            if (sp.nq == 0) continue;
            SpaceRec r{&q, prec, 0, 0, eri_uses_wave_tasks(LX.lcls, LY.lcls), 0, 0, 0, 0, 0, 0, 0};
            // row-merge configuration: the tuned one for this space if any, else the defaults
            int variant = r.tasks && rows_on
                              ? eri_rows_variant(LX.lcls, LY.lcls, mode, c->run_ncols[q.Y], prec == 1, c->row_variant)
                              : 0;
            long long min_kets = std::max(1LL, c->row_min_kets);
            int dlds = variant && c->row_dens != 0 &&
                       eri_rows_dlds_fits(LX.lcls, LY.lcls, mode, c->run_ncols[q.Y], prec == 1, variant);
            if (r.tasks && rows_on && !c->tuning) {
                auto it = c->tune.find(std::make_tuple(qi, prec, uhf ? 1 : 0, mixed ? 1 : 0));
                if (it != c->tune.end()) {
                    const int v = it->second.variant;
                    if (v == 0 || eri_rows_variant(LX.lcls, LY.lcls, mode, c->run_ncols[q.Y], prec == 1, v) == v) {
                        variant = v;
                        min_kets = it->second.min_kets;
                        dlds = v && it->second.dlds &&
                               eri_rows_dlds_fits(LX.lcls, LY.lcls, mode, c->run_ncols[q.Y], prec == 1, v);
                    }
                }
            }
            r.variant = variant;
            r.dlds = dlds;
            if (r.tasks) {
                build_rows(sp, variant ? min_kets : 0, kRowSeg, eri_rows_group(variant));
                build_tasks(sp, variant ? min_kets : 0);
            } else {
                build_blocks(sp, eri_quartets_per_block(LX.lcls, LY.lcls));
            }
            built[2 * (size_t)qi + prec] = r;
            used[2 * (size_t)qi + prec] = 1;
        }
    }
    for (size_t k = 0; k < built.size(); ++k) {
        if (!used[k]) continue;
        SpaceRec r = built[k];
        const QuartetSpace& sp = r.prec ? r.q->s32 : r.q->s64;
        if (!c->tuning) {
            r.o_prefix = put(sp.prefix);
            if (r.tasks) {
                r.o_wpre = put(sp.wpre);
                r.o_tblk = put(sp.tblk);
                r.o_rbra = put(sp.rbra);
                r.o_rpre = put(sp.rpre);
            } else {
                r.o_blk = put(sp.blk);
            }
            if (r.prec) r.o_kstart = put(sp.kstart);
        }
        recs.push_back(r);
    }
    c->last_ms[1] = ms_since(h1);
    const auto h2 = hclk::now();
    char* ds = nullptr;
    if (!c->tuning && !hs.empty()) {
        // the previous build's upload must have read the pinned copy before it is reused
        // (the device buffer is reused in stream order)
        if (c->stage_ev) UNO_HIP(hipEventSynchronize(c->stage_ev));
        else UNO_HIP(hipEventCreateWithFlags(&c->stage_ev, hipEventDisableTiming));
        if (c->stage_pin_cap < hs.size()) {
            // grown to twice the need: a pinned allocation costs milliseconds, and the
            // full builds' arrays grow as the density bound loosens
            if (c->stage_pin) UNO_HIP(hipHostFree(c->stage_pin));
            c->stage_pin = nullptr;
            c->stage_pin_cap = 0;
            UNO_HIP(hipHostMalloc((void**)&c->stage_pin, 2 * hs.size()));
            c->stage_pin_cap = 2 * hs.size();
        }
        std::memcpy(c->stage_pin, hs.data(), hs.size());
        if (c->stage_d.n < hs.size()) UNO_HIP(c->stage_d.alloc(2 * hs.size()));
        UNO_HIP(hipMemcpyAsync(c->stage_d.p, c->stage_pin, hs.size(), hipMemcpyHostToDevice, st));
        UNO_HIP(hipEventRecord(c->stage_ev, st));
        ds = c->stage_d.p;
    }

    c->last_ms[2] = ms_since(h2);
    const auto h3 = hclk::now();
    // worker streams (not while tuning: the tuner times each launch on the context's
    // stream); they start after everything enqueued so far on the context's stream
    const int nls = c->tuning ? 1 : std::max(1, std::min(kMaxLaunchStreams, c->launch_streams));
    for (int w = 0; w + 1 < nls; ++w)
        if (!c->wst[w]) {
            UNO_HIP(hipStreamCreate(&c->wst[w]));
            UNO_HIP(hipEventCreateWithFlags(&c->wjoin[w], hipEventDisableTiming));
        }
    if (nls > 1) {
        if (!c->wfork) UNO_HIP(hipEventCreateWithFlags(&c->wfork, hipEventDisableTiming));
        UNO_HIP(hipEventRecord(c->wfork, st));
        for (int w = 0; w + 1 < nls; ++w) UNO_HIP(hipStreamWaitEvent(c->wst[w], c->wfork, 0));
    }
    int nlaunch = 0;
    // Pass 2: the launches
    for (const SpaceRec& r : recs) {
        QuartetClass& q = *r.q;
        const int prec = r.prec;
        QuartetSpace& sp = prec ? q.s32 : q.s64;
        const PairClassList& LX = c->cls[q.X];
        const PairClassList& LY = c->cls[q.Y];
        const int qi = (int)(&q - c->qcls.data());
        const bool tasks = r.tasks;
        EriArgs a = base_args(c);
        a.bra_off = LX.off + sp.b0;
        a.ket_off = LY.off;
        a.pairQ = c->d_pairQ.p;
        a.rseg = kRowSeg;
        a.colmap = c->d_colmap.p + (size_t)q.Y * c->n;
        a.colinv = c->d_colinv.p + (size_t)q.Y * c->n;
        a.ncols = c->run_ncols[q.Y];
        a.dlds = r.dlds;
        a.qpb = eri_quartets_per_block(LX.lcls, LY.lcls);
        a.nbra = (int)sp.prefix.size() - 1;
        a.nquart = sp.nq;
        a.same_class = q.X == q.Y;
        a.dsh = c->Dsh.p;
        a.jp = det ? nullptr : c->JP.p;
        a.fix_lo = det ? nn : 0;
        a.dmax_global = 4.0 * dmax;
        a.D0 = c->D0.p;
        a.W0 = c->W0.p;
        if (uhf) {
            a.D1 = c->D1.p;
            a.D2 = c->D2.p;
            a.W1 = c->W1.p;
        }
        // device index arrays: the staging buffer's slices, or (tuning) the space's own
        auto set_index = [&]() -> int {
            if (ds) {
                a.prefix = reinterpret_cast<const long long*>(ds + r.o_prefix);
                a.blk = tasks ? nullptr : reinterpret_cast<const int*>(ds + r.o_blk);
                a.wpre = tasks ? reinterpret_cast<const long long*>(ds + r.o_wpre) : nullptr;
                a.tblk = tasks ? reinterpret_cast<const int*>(ds + r.o_tblk) : nullptr;
                a.rbra = tasks && sp.nrtask > 0 ? reinterpret_cast<const int*>(ds + r.o_rbra) : nullptr;
                a.rpre = tasks && sp.nrtask > 0 ? reinterpret_cast<const long long*>(ds + r.o_rpre) : nullptr;
                a.kstart = prec ? reinterpret_cast<const int*>(ds + r.o_kstart) : nullptr;
            } else {
                UNO_HIP(sp.d_prefix.upload(sp.prefix.data(), sp.prefix.size(), st));
                if (tasks) {
                    UNO_HIP(sp.d_wpre.upload(sp.wpre.data(), sp.wpre.size(), st));
                    UNO_HIP(sp.d_tblk.upload(sp.tblk.data(), sp.tblk.size(), st));
                    if (sp.nrtask > 0) {
                        UNO_HIP(sp.d_rbra.upload(sp.rbra.data(), sp.rbra.size(), st));
                        UNO_HIP(sp.d_rpre.upload(sp.rpre.data(), sp.rpre.size(), st));
                    }
                } else {
                    UNO_HIP(sp.d_blk.upload(sp.blk.data(), sp.blk.size(), st));
                }
                if (prec) UNO_HIP(sp.d_kstart.upload(sp.kstart.data(), sp.kstart.size(), st));
                a.prefix = sp.d_prefix.p;
                a.blk = tasks ? nullptr : sp.d_blk.p;
                a.wpre = tasks ? sp.d_wpre.p : nullptr;
                a.tblk = tasks ? sp.d_tblk.p : nullptr;
                a.rbra = tasks && sp.nrtask > 0 ? sp.d_rbra.p : nullptr;
                a.rpre = tasks && sp.nrtask > 0 ? sp.d_rpre.p : nullptr;
                a.kstart = prec ? sp.d_kstart.p : nullptr;
            }
            a.ntask = tasks ? sp.ntask : 0;
            a.nrbra = (int)sp.rbra.size();
            a.nrtask = tasks ? sp.nrtask : 0;
            return UNO_OK;
        };
        const int ci = prec * kNumPairClasses * kNumPairClasses + LX.lcls * kNumPairClasses + LY.lcls;
        // one launch per kernel kind, each with its own profiling slot
        auto launch = [&](int kind) -> int {
            const int slot = kind * 2 * kNumPairClasses * kNumPairClasses + ci;
            const bool prof = c->profile && !c->tuning && c->d_counter.p;
            const int li = nlaunch++ % nls;
            const hipStream_t ls = li == 0 ? st : c->wst[li - 1];
            hipEvent_t e1 = nullptr;
            if (prof) {
                a.counter = c->d_counter.p + 2 * kCounterSlots * slot;
                const size_t k = c->prof_pending++;
                while (c->prof_ev.size() < 2 * (k + 1)) {
                    hipEvent_t e = nullptr;
                    UNO_HIP(hipEventCreate(&e));
                    c->prof_ev.push_back(e);
                }
                if (c->prof_evslot.size() < k + 1) c->prof_evslot.resize(k + 1);
                c->prof_evslot[k] = slot;
                e1 = c->prof_ev[2 * k + 1];
                UNO_HIP(hipEventRecord(c->prof_ev[2 * k], ls));
            } else {
                a.counter = nullptr;
            }
            if (kind) UNO_HIP(eri_launch_rows(mode, LX.lcls, LY.lcls, a, kind, ls, prec == 1));
            else UNO_HIP(eri_launch(mode, LX.lcls, LY.lcls, a, ls, prec == 1));
            if (prof) UNO_HIP(hipEventRecord(e1, ls));
            return UNO_OK;
        };
        // the space's launches: row-merged bras with LDS variant v, then the wave tasks
        auto run = [&](int v) -> int {
            if (a.nrtask > 0)
                if (int rc = launch(v)) return rc;
            if (!tasks || a.ntask > 0)
                if (int rc = launch(0)) return rc;
            return UNO_OK;
        };
        if (c->tuning && tasks && rows_on) {
            // time every candidate on this space (each one run three times, the fastest counts)
            std::vector<TuneChoice> cand{{0, 0, 0}};
            for (int v = 1; v <= 5; ++v)
                if (eri_rows_variant(LX.lcls, LY.lcls, mode, c->run_ncols[q.Y], prec == 1, v) == v) {
                    const bool df = c->row_dens != 0 &&
                                    eri_rows_dlds_fits(LX.lcls, LY.lcls, mode, c->run_ncols[q.Y], prec == 1, v);
                    for (long long mk : {64LL, 192LL, 512LL}) {
                        if (c->row_dens != 1 || !df) cand.push_back({v, mk, 0});
                        if (df) cand.push_back({v, mk, 1});
                    }
                }
            TuneChoice best{r.variant, std::max(1LL, c->row_min_kets), r.dlds};
            float best_ms = 1e30f;
            hipEvent_t t0, t1;
            UNO_HIP(hipEventCreate(&t0));
            UNO_HIP(hipEventCreate(&t1));
            for (const TuneChoice& tc : cand) {
                build_rows(sp, tc.variant ? tc.min_kets : 0, kRowSeg, eri_rows_group(tc.variant));
                build_tasks(sp, tc.variant ? tc.min_kets : 0);
                a.dlds = tc.dlds;
                if (int rc = set_index()) return rc;
                float ms_min = 1e30f;
                for (int rep = 0; rep < 3; ++rep) {
                    UNO_HIP(hipEventRecord(t0, st));
                    if (int rc = run(tc.variant)) return rc;
                    UNO_HIP(hipEventRecord(t1, st));
                    UNO_HIP(hipEventSynchronize(t1));
                    float ms = 0.f;
                    UNO_HIP(hipEventElapsedTime(&ms, t0, t1));
                    ms_min = std::min(ms_min, ms);
                }
                if (ms_min < best_ms) {
                    best_ms = ms_min;
                    best = tc;
                }
            }
            (void)hipEventDestroy(t0);
            (void)hipEventDestroy(t1);
            c->tune[std::make_tuple(qi, prec, uhf ? 1 : 0, mixed ? 1 : 0)] = best;
            c->tune_ms += best_ms;
        } else {
            if (int rc = set_index()) return rc;
            if (int rc = run(r.variant)) return rc;
        }
        (prec ? total32 : total) += (double)sp.nq;
    }
    // the context's stream continues once every worker's launches are done
    for (int w = 0; w + 1 < nls; ++w) {
        UNO_HIP(hipEventRecord(c->wjoin[w], c->wst[w]));
        UNO_HIP(hipStreamWaitEvent(st, c->wjoin[w], 0));
    }
    c->last_ms[3] = ms_since(h3);
    if (det) {
        UNO_HIP(fixed_to_f64(nn, c->W0.p, nn, st));
        if (uhf) UNO_HIP(fixed_to_f64(nn, c->W1.p, nn, st));
    }
    // the kets' Coulomb outputs, accumulated per pair, into W
    if (!det) UNO_HIP(jp_scatter((int)c->pairs.size(), c->d_pairs.p, c->njp, c->JP.p, c->W0.p, uhf ? c->W1.p : nullptr, c->n, st));
    c->last_quartets = total + total32;
    c->last_quartets32 = total32;
    return UNO_OK;
}

int profile_reset(uno_ctx* c, bool enable) {
    for (int i = 0; i < kProfSlots; ++i) c->prof_ms[i] = c->prof_launch[i] = c->prof_prim[i] = c->prof_shell[i] = 0.0;
    c->prof_pending = 0;   // pairs of earlier builds are dropped (their events are re-recorded)
    if (enable) {
        // the counters accumulate on the device from here until the next reset
        UNO_HIP(c->d_counter.alloc(2 * kCounterSlots * kProfSlots));
        // (all-zero bits: a fill of 0.0 clears the 64-bit counters)
        UNO_HIP(k_fill(2 * kCounterSlots * kProfSlots, 0.0, reinterpret_cast<double*>(c->d_counter.p), c->st));
    }
    return UNO_OK;
}

int profile_collect(uno_ctx* c) {
    if (!c->d_counter.p) return UNO_OK;
    UNO_HIP(hipStreamSynchronize(c->st));
    for (size_t k = 0; k < c->prof_pending; ++k) {
        float ms = 0.f;
        UNO_HIP(hipEventElapsedTime(&ms, c->prof_ev[2 * k], c->prof_ev[2 * k + 1]));
        c->prof_ms[c->prof_evslot[k]] += ms;
        c->prof_launch[c->prof_evslot[k]] += 1.0;
    }
    c->prof_pending = 0;
    std::vector<unsigned long long> cnt(2 * kCounterSlots * kProfSlots);
    UNO_HIP(hipMemcpy(cnt.data(), c->d_counter.p, cnt.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    // device totals since the last reset
    for (int ci = 0; ci < kProfSlots; ++ci) {
        double pr = 0.0, sh = 0.0;
        for (int sl = 0; sl < kCounterSlots; ++sl) {
            pr += (double)cnt[2 * (kCounterSlots * ci + sl)];
            sh += (double)cnt[2 * (kCounterSlots * ci + sl) + 1];
        }
        c->prof_prim[ci] = pr;
        c->prof_shell[ci] = sh;
    }
    return UNO_OK;
}

// Positron model potential (GDPMInts.cpp:6-146): the positron's charge as six s Gaussians
// c_k exp(-a_k |r - R|^2) at R; every shell pair's (ab|g_k) from the FP64 ERI kernels in
// block mode with the Gaussians as a six-primitive ket pair, no primitive cut (the
// reference applies none there).  The expansion is the reference's fitted table.
int gdpm_matrix(uno_ctx* c, const double* R, std::vector<double>& dh) {
    static const double kAlf[6] = {6.8505018000, 4.0491646000, 3.5941062989, 1.2478274000, 0.7927690989, 0.3377107978};
    static const double kCof[6] = {0.0766926784, 0.1483475912, 0.0462334641, 0.0717376427, 0.0447149792, 0.0069678529};
    const int n = c->n;
    dh.assign((size_t)n * (n + 1) / 2, 0.0);
    std::vector<PairHdr> pairs(c->pairs);
    std::vector<PrimPair> prims(c->prims);
    PairHdr g{};
    for (int x = 0; x < 3; ++x) g.A[x] = R[x];
    g.prim0 = (int)prims.size();
    g.nprim = 6;
    g.sh1 = g.sh2 = -1;
    const int gi = (int)pairs.size();
    pairs.push_back(g);
    for (int k = 0; k < 6; ++k) {
        PrimPair q{};
        q.p = kAlf[k];
        q.ip = 1.0 / kAlf[k];
        for (int x = 0; x < 3; ++x) q.P[x] = R[x];
        q.K = kCof[k];
        q.s = 1.0;
        prims.push_back(q);
    }
    hipStream_t st = c->st;
    DBuf<PairHdr> dp;
    DBuf<PrimPair> dpr;
    UNO_HIP(dp.upload(pairs.data(), pairs.size(), st));
    UNO_HIP(dpr.upload(prims.data(), prims.size(), st));
    std::vector<std::vector<int2>> lists(kNumPairClasses);
    for (int p = 0; p < gi; ++p) {
        const PairHdr& h = c->pairs[p];
        lists[pair_class(c->B.sh[h.sh1].l, c->B.sh[h.sh2].l)].push_back(make_int2(p, gi));
    }
    for (int X = 0; X < kNumPairClasses; ++X) {
        if (lists[X].empty()) continue;
        const int nc = eri_class_ncomp(X, 0);
        DBuf<int2> dq;
        DBuf<double> dout;
        UNO_HIP(dq.upload(lists[X].data(), lists[X].size(), st));
        UNO_HIP(dout.alloc(lists[X].size() * (size_t)nc));
        EriArgs a = base_args(c);
        a.pairs = dp.p;
        a.prims = dpr.p;
        a.primcut = 0.0;
        a.explicit_q = dq.p;
        a.nquart = (long long)lists[X].size();
        a.out = dout.p;
        UNO_HIP(eri_launch(kDigestBlock, X, 0, a, st));
        std::vector<double> blk(lists[X].size() * (size_t)nc);
        UNO_HIP(hipMemcpyAsync(blk.data(), dout.p, blk.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        UNO_HIP(hipStreamSynchronize(st));
        for (size_t q = 0; q < lists[X].size(); ++q) {
            const PairHdr& h = c->pairs[lists[X][q].x];
            const int na = ncart(c->B.sh[h.sh1].l), nb = ncart(c->B.sh[h.sh2].l);
            for (int i = 0; i < na; ++i)
                for (int j = 0; j < nb; ++j) {
                    const int r = h.f1 + i, col = h.f2 + j;
                    if (h.sh1 == h.sh2 && col > r) continue;   // same shell: each element once
                    dh[pidx(r, col)] -= blk[q * nc + i * nb + j];
                }
        }
    }
    return UNO_OK;
}

// Every canonical function quartet in the reference's order, evaluated on the device
// shell-quartet blocks.
int eri_emit(uno_ctx* c, long cap, double* val, int* ijkl, long* count) {
    const Basis& B = c->B;
    const int nsh = c->nsh;
    struct Rec {
        int cls;        // index into groups
        long q;         // quartet within the group
        int comp;       // component index within the block
    };
    std::vector<Rec> recs;
    std::vector<int> fq;   // 4 per record
    std::vector<std::vector<int2>> groups(kNumPairClasses * kNumPairClasses);
    long n = 0;
    for (int ish = c->start_shell; ish < nsh; ++ish)
        for (int jsh = 0; jsh <= ish; ++jsh)
            for (int ksh = 0; ksh <= ish; ++ksh)
                for (int lsh = 0; lsh <= ksh; ++lsh) {
                    const int sh[4] = {ish, jsh, ksh, lsh};
                    int nf[4], o[4];
                    for (int x = 0; x < 4; ++x) {
                        nf[x] = ncart(B.sh[sh[x]].l);
                        o[x] = B.off[sh[x]];
                    }
                    const int p1 = c->pair_of[(size_t)ish * nsh + jsh];
                    const int p2 = c->pair_of[(size_t)ksh * nsh + lsh];
                    long qidx = -1;
                    int gidx = -1;
                    bool swap = false;
                    int NBb = 0, NCb = 0, NDb = 0;
                    for (int a = 0; a < nf[0]; ++a) {
                        const int ir = o[0] + a;
                        for (int b = 0; b < nf[1]; ++b) {
                            const int jr = o[1] + b;
                            if (jr > ir) break;
                            for (int cc = 0; cc < nf[2]; ++cc) {
                                const int kr = o[2] + cc;
                                if (kr > ir) break;
                                for (int d = 0; d < nf[3]; ++d) {
                                    const int lr = o[3] + d;
                                    if (lr > kr || (ir == kr && lr > jr)) break;
                                    Rec r{-1, -1, 0};
                                    if (p1 >= 0 && p2 >= 0) {
                                        if (qidx < 0) {
                                            const PairHdr& h1 = c->pairs[p1];
                                            const PairHdr& h2 = c->pairs[p2];
                                            const int X = pair_class(B.sh[h1.sh1].l, B.sh[h1.sh2].l);
                                            const int Y = pair_class(B.sh[h2.sh1].l, B.sh[h2.sh2].l);
                                            swap = X < Y;
                                            const int gb = swap ? Y : X, gk = swap ? X : Y;
                                            gidx = gb * kNumPairClasses + gk;
                                            const PairHdr& hb = swap ? h2 : h1;
                                            const PairHdr& hk = swap ? h1 : h2;
                                            NBb = ncart(B.sh[hb.sh2].l);
                                            NCb = ncart(B.sh[hk.sh1].l);
                                            NDb = ncart(B.sh[hk.sh2].l);
                                            groups[gidx].push_back(swap ? make_int2(p2, p1) : make_int2(p1, p2));
                                            qidx = (long)groups[gidx].size() - 1;
                                        }
                                        const PairHdr& h1 = c->pairs[p1];
                                        const PairHdr& h2 = c->pairs[p2];
                                        // component indices in each pair's stored orientation
                                        const int i1 = (h1.sh1 == ish) ? a : b, i2 = (h1.sh1 == ish) ? b : a;
                                        const int k1 = (h2.sh1 == ksh) ? cc : d, k2 = (h2.sh1 == ksh) ? d : cc;
                                        int comp;
                                        if (!swap) comp = ((i1 * NBb + i2) * NCb + k1) * NDb + k2;
                                        else comp = ((k1 * NBb + k2) * NCb + i1) * NDb + i2;
                                        r = Rec{gidx, qidx, comp};
                                    }
                                    recs.push_back(r);
                                    fq.push_back(ir); fq.push_back(jr); fq.push_back(kr); fq.push_back(lr);
                                    ++n;
                                }
                            }
                        }
                    }
                }
    *count = n;
    if (cap <= 0) return UNO_OK;
    hipStream_t st = c->st;
    std::vector<std::vector<double>> blocks(groups.size());
    for (int g = 0; g < (int)groups.size(); ++g) {
        if (groups[g].empty()) continue;
        const int X = g / kNumPairClasses, Y = g % kNumPairClasses;
        const int nc = eri_class_ncomp(X, Y);
        DBuf<int2> dq;
        DBuf<double> dout;
        UNO_HIP(dq.upload(groups[g].data(), groups[g].size(), st));
        UNO_HIP(dout.alloc(groups[g].size() * (size_t)nc));
        EriArgs a = base_args(c);
        a.explicit_q = dq.p;
        a.nquart = (long long)groups[g].size();
        a.out = dout.p;
        UNO_HIP(eri_launch(kDigestBlock, X, Y, a, st));
        blocks[g].resize(groups[g].size() * (size_t)nc);
        UNO_HIP(hipMemcpyAsync(blocks[g].data(), dout.p, blocks[g].size() * sizeof(double), hipMemcpyDeviceToHost, st));
        UNO_HIP(hipStreamSynchronize(st));
    }
    for (long e = 0; e < n && e < cap; ++e) {
        const Rec& r = recs[e];
        double v = 0.0;
        if (r.cls >= 0) {
            const int X = r.cls / kNumPairClasses, Y = r.cls % kNumPairClasses;
            v = blocks[r.cls][(size_t)r.q * eri_class_ncomp(X, Y) + r.comp];
        }
        val[e] = v;
        for (int x = 0; x < 4; ++x) ijkl[4 * e + x] = fq[4 * e + x];
    }
    return UNO_OK;
}

}  // namespace uno
