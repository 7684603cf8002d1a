// Batched SCF: many molecules of uneven size (each N <= 256 functions), RHF and UHF mixed,
// iterated in lockstep on one device (no reference counterpart: the reference runs one
// molecule per process).
//
// Every molecule keeps its own context for the pair data (shell pairs, Schwarz bounds,
// runs sorted by bound); their pair and primitive arrays are concatenated into one device
// array.  A Fock build is one launch per (bra class, ket class, RHF/UHF) over all
// molecules: the quartet space is a list of entries (bra pair, first ket pair, ket count,
// molecule) built from the same screened prefixes as a single-molecule build, and the
// kernels read and accumulate each molecule's slice of the density / accumulator stores.
// Dense steps run as grouped GEMMs over every (molecule, spin) slot and a batched Jacobi
// eigensolver (one block per matrix).  The iteration is the single-molecule one with Pulay
// DIIS (8 vectors, error X^T(FPS - SPF)X); molecules stop updating once converged.
#include <algorithm>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <tuple>
#include "engine.hpp"
#include "../hip/batch_kernels.hpp"
#include "../hip/kernels.hpp"

using namespace uno;

namespace {

struct BatchMol {
    uno_ctx* ctx = nullptr;
    int nelec = 0, n = 0, nsh = 0, nocc[2] = {0, 0};
    bool uhf = false;
    long long moff = 0, dshoff = 0;
    int shoff = 0, pair_off = 0;
    // state
    bool active = true, converged = false;
    int iterations = 0;
    double energy = 0.0, eold = 0.0, ediff = 0.0, pdiff = 0.0, nuc = 0.0, init_energy = 0.0;
};

constexpr int kBatchDiis = 8;

using clk = std::chrono::steady_clock;
double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }


// Device copies of the small per-iteration arrays (GEMM groups and tiles, matrix lists,
// quartet entries, DIIS coefficients): the k-th upload of an iteration reuses slot k,
// grown when needed, so the steady state allocates nothing.  Every iteration ends with a
// stream synchronisation before reset() hands the slots out again.
struct Scratch {
    std::vector<DBuf<char>> slot;
    size_t next = 0;
    void reset() { next = 0; }
    template <class T>
    hipError_t upload(const T* h, size_t n, hipStream_t st, T** out) {
        if (next == slot.size()) slot.emplace_back();
        DBuf<char>& s = slot[next++];
        hipError_t e = s.alloc(std::max<size_t>(n, 1) * sizeof(T));
        if (e != hipSuccess) return e;
        *out = reinterpret_cast<T*>(s.p);
        return n ? hipMemcpyAsync(s.p, h, n * sizeof(T), hipMemcpyHostToDevice, st) : hipSuccess;
    }
};

}  // namespace

struct uno_batch {
    std::vector<BatchMol> mol;
    hipStream_t st = nullptr;
    long long store = 0, dsh_total = 0;
    DBuf<PairHdr> pairs;
    DBuf<PrimPair> prims;
    DBuf<double> pairQ;
    DBuf<int> sh_off, sh_n;
    DBuf<MolDesc> mdesc;
    // stores (each sum_m N_m^2 doubles); index 0 alpha / closed shell, 1 beta
    DBuf<double> H, S, X, P[2], Pold[2], F[2], G[2], C[2], Fp, Tm, Tm2, V, Z, D0, D1, D2, W0, W1, Dsh, ev[2];
    // per spin: the last eigenvectors of X^T F X (orthogonal basis), the next eigensolve's
    // starting basis once warm[spin] is set
    DBuf<double> Wp[2];
    bool warm[2] = {false, false};
    DBuf<double> dF[kBatchDiis][2], dE[kBatchDiis][2];
    DBuf<double> red;
    Scratch scratch;
    // the Fock build's static quartet spaces (one per bra class, ket class, open shell),
    // their work-queue counters, the per-molecule active flags, and the build captured as
    // a HIP graph (replayed every iteration while the spaces stay)
    struct Group {
        int cx, cy, uhf, tasks, qpb, ne;
        long long nq, ntask;
        int *bra, *ket, *mol, *tblk, *blk;
        long long *prefix, *wpre;
    };
    std::vector<Group> groups;
    DBuf<char> gstore;
    DBuf<unsigned long long> queues;
    DBuf<int> active_d;
    std::vector<double> dbound;     // per molecule: the density bound the spaces were built with
    std::vector<double> mol_quartets;   // per molecule: quartets in the static spaces
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int use_graph = 1, use_queue = 1;
    // the groups' launches go round-robin over st and launch_streams - 1 worker streams
    // (forked from st after the fills and joined back; under capture the graph then holds
    // parallel branches)
    int launch_streams = 3;
    hipStream_t wst[kMaxLaunchStreams - 1] = {};
    hipEvent_t wfork = nullptr, wjoin[kMaxLaunchStreams - 1] = {};
    int graph_builds = 0, graph_replays = 0, space_builds = 0;
    void drop_graph() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
    }
    ~uno_batch() {
        drop_graph();
        for (BatchMol& m : mol)
            if (m.ctx) uno_ctx_destroy(m.ctx);
        if (st) (void)hipStreamSynchronize(st);
        for (int w = 0; w + 1 < kMaxLaunchStreams; ++w) {
            if (wst[w]) (void)hipStreamDestroy(wst[w]);
            if (wjoin[w]) (void)hipEventDestroy(wjoin[w]);
        }
        if (wfork) (void)hipEventDestroy(wfork);
        if (st) (void)hipStreamDestroy(st);
    }
};

namespace {

// slots: per (molecule, spin) matrix; spin 1 only for UHF molecules
void make_slots(const uno_batch* b, bool only_active, int spin, std::vector<MatSlot>& out, std::vector<int>& who) {
    out.clear();
    who.clear();
    for (int m = 0; m < (int)b->mol.size(); ++m) {
        const BatchMol& x = b->mol[m];
        if (only_active && !x.active) continue;
        if (spin == 1 && !x.uhf) continue;
        out.push_back(MatSlot{x.moff, x.n, x.nsh, x.shoff, x.dshoff});
        who.push_back(m);
    }
}

struct GemmBatch {
    std::vector<GemmGroup> g;
    std::vector<GemmTile> t;
    Scratch* sc;
    explicit GemmBatch(Scratch& s) : sc(&s) {}
    void add(int m, int n, int k, bool ta, bool tb, const double* A, const double* B, double* C, int lda, int ldb,
             int ldc) {
        if (m <= 0 || n <= 0 || k <= 0) return;
        const int gi = (int)g.size();
        g.push_back(GemmGroup{m, n, k, ta ? 1 : 0, tb ? 1 : 0, A, B, C, lda, ldb, ldc});
        for (int i = 0; i < (m + 31) / 32; ++i)
            for (int j = 0; j < (n + 31) / 32; ++j) t.push_back(GemmTile{gi, i, j});
    }
    int run(double alpha, double beta, hipStream_t st) {
        if (t.empty()) return UNO_OK;
        GemmGroup* dg = nullptr;
        GemmTile* dt = nullptr;
        UNO_HIP(sc->upload(g.data(), g.size(), st, &dg));
        UNO_HIP(sc->upload(t.data(), t.size(), st, &dt));
        UNO_HIP(grouped_gemm((int)t.size(), dt, dg, alpha, beta, st));
        // the group / tile arrays are pageable host memory: keep them alive until done
        UNO_HIP(hipStreamSynchronize(st));
        return UNO_OK;
    }
};

// Z, evals of every slot's symmetric matrix A (destroyed), A given per slot
int batch_eigh(uno_batch* b, const std::vector<MatSlot>& sl, double* A, double* evals_base,
               const std::vector<long long>& evoff) {
    std::vector<BatchMat> bm(sl.size());
    for (size_t k = 0; k < sl.size(); ++k)
        bm[k] = BatchMat{sl[k].n, A + sl[k].off, b->V.p + sl[k].off, evals_base + evoff[k], b->Z.p + sl[k].off};
    BatchMat* d = nullptr;
    UNO_HIP(b->scratch.upload(bm.data(), bm.size(), b->st, &d));
    UNO_HIP(batched_jacobi((int)bm.size(), d, 50, nullptr, b->st));
    UNO_HIP(hipStreamSynchronize(b->st));
    return UNO_OK;
}

}  // namespace

namespace uno {

// Static quartet spaces of the batched Fock build: every molecule's screened prefixes built
// with a density bound (twice the largest |D| block maximum seen so far, per molecule)
// instead of the current density, so that the launches -- their index arrays, counts and
// pointers -- stay the same from build to build.  The kernels apply the exact per-shell-block
// density screen, which selects the same quartets prefixes built at the current density
// would; a density that exceeds its bound rebuilds the spaces.  Entries are sorted by
// their ket count (largest first) so that the work queue hands out the costly bras first.
static int build_static(uno_batch* b, const std::vector<double>& dmax) {
    b->drop_graph();
    const int nm = (int)b->mol.size();
    if ((int)b->dbound.size() != nm) b->dbound.assign(nm, 0.0);
    for (int m = 0; m < nm; ++m) b->dbound[m] = std::max(2.0 * dmax[m], b->dbound[m]);
    struct Ent { std::vector<int> bra, ket, mol; std::vector<long long> cnt; };
    std::map<std::tuple<int, int, int>, Ent> groups;
    b->mol_quartets.assign(nm, 0.0);
    for (int m = 0; m < nm; ++m) {
        BatchMol& x = b->mol[m];
        uno_ctx* c = x.ctx;
        for (QuartetClass& q : c->qcls) {
            build_prefix(c, q, 4.0 * b->dbound[m], 0.0, c->tau);
            if (q.s64.nq == 0) continue;
            b->mol_quartets[m] += (double)q.s64.nq;
            const PairClassList& LX = c->cls[q.X];
            const PairClassList& LY = c->cls[q.Y];
            Ent& e = groups[std::make_tuple(LX.lcls, LY.lcls, x.uhf ? 1 : 0)];
            for (int i = 0; i + 1 < (int)q.s64.prefix.size(); ++i) {
                const long long cnt = q.s64.prefix[i + 1] - q.s64.prefix[i];
                if (cnt == 0) continue;
                e.bra.push_back(x.pair_off + LX.off + q.s64.b0 + i);
                e.ket.push_back(x.pair_off + LY.off);
                e.mol.push_back(m);
                e.cnt.push_back(cnt);
            }
        }
    }
    // host layout of every group's arrays in one buffer (16-byte aligned pieces)
    std::vector<char> hs;
    auto put = [&](const auto& v) -> size_t {
        const size_t off = (hs.size() + 15) / 16 * 16;
        const size_t bytes = v.size() * sizeof(v[0]);
        hs.resize(off + std::max<size_t>(bytes, 16));
        if (bytes) std::memcpy(hs.data() + off, v.data(), bytes);
        return off;
    };
    struct Offs { size_t bra, ket, mol, prefix, wpre, tblk, blk; };
    std::vector<Offs> offs;
    b->groups.clear();
    for (auto& kv : groups) {
        Ent& e = kv.second;
        const int ne = (int)e.bra.size();
        std::vector<int> order(ne);
        for (int i = 0; i < ne; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int u, int v) { return e.cnt[u] > e.cnt[v]; });
        std::vector<int> bra(ne), ket(ne), mol(ne);
        QuartetSpace sp;
        sp.prefix.assign(ne + 1, 0);
        for (int i = 0; i < ne; ++i) {
            bra[i] = e.bra[order[i]];
            ket[i] = e.ket[order[i]];
            mol[i] = e.mol[order[i]];
            sp.prefix[i + 1] = sp.prefix[i] + e.cnt[order[i]];
        }
        sp.nq = sp.prefix[ne];
        uno_batch::Group g{};
        g.cx = std::get<0>(kv.first);
        g.cy = std::get<1>(kv.first);
        g.uhf = std::get<2>(kv.first);
        g.tasks = eri_uses_wave_tasks(g.cx, g.cy) ? 1 : 0;
        g.qpb = eri_quartets_per_block(g.cx, g.cy);
        g.ne = ne;
        g.nq = sp.nq;
        Offs o{};
        o.bra = put(bra);
        o.ket = put(ket);
        o.mol = put(mol);
        o.prefix = put(sp.prefix);
        if (g.tasks) {
            build_tasks(sp);
            g.ntask = sp.ntask;
            o.wpre = put(sp.wpre);
            o.tblk = put(sp.tblk);
        } else {
            build_blocks(sp, g.qpb);
            o.blk = put(sp.blk);
        }
        b->groups.push_back(g);
        offs.push_back(o);
    }
    UNO_HIP(b->gstore.alloc(std::max<size_t>(hs.size(), 16)));
    UNO_HIP(hipMemcpyAsync(b->gstore.p, hs.data(), hs.size(), hipMemcpyHostToDevice, b->st));
    UNO_HIP(b->queues.alloc(std::max<size_t>(b->groups.size(), 1)));
    UNO_HIP(b->active_d.alloc(std::max(nm, 1)));
    for (size_t k = 0; k < b->groups.size(); ++k) {
        uno_batch::Group& g = b->groups[k];
        const Offs& o = offs[k];
        char* base = b->gstore.p;
        g.bra = reinterpret_cast<int*>(base + o.bra);
        g.ket = reinterpret_cast<int*>(base + o.ket);
        g.mol = reinterpret_cast<int*>(base + o.mol);
        g.prefix = reinterpret_cast<long long*>(base + o.prefix);
        g.wpre = g.tasks ? reinterpret_cast<long long*>(base + o.wpre) : nullptr;
        g.tblk = g.tasks ? reinterpret_cast<int*>(base + o.tblk) : nullptr;
        g.blk = g.tasks ? nullptr : reinterpret_cast<int*>(base + o.blk);
    }
    UNO_HIP(hipStreamSynchronize(b->st));   // hs is freed on return
    ++b->space_builds;
    return UNO_OK;
}

// the launches of one build: accumulators and queue counters cleared by fill kernels (no
// memset nodes when captured), then one launch per static space
// the worker streams and events, made before any capture
static int ensure_workers(uno_batch* b) {
    const int nls = std::max(1, std::min(kMaxLaunchStreams, b->launch_streams));
    for (int w = 0; w + 1 < nls; ++w)
        if (!b->wst[w]) {
            UNO_HIP(hipStreamCreate(&b->wst[w]));
            UNO_HIP(hipEventCreateWithFlags(&b->wjoin[w], hipEventDisableTiming));
        }
    if (nls > 1 && !b->wfork) UNO_HIP(hipEventCreateWithFlags(&b->wfork, hipEventDisableTiming));
    return UNO_OK;
}

static int enqueue_fock(uno_batch* b) {
    hipStream_t st = b->st;
    UNO_HIP(k_fill(b->store, 0.0, b->W0.p, st));
    UNO_HIP(k_fill(b->store, 0.0, b->W1.p, st));
    if (b->use_queue && !b->groups.empty())
        UNO_HIP(k_fill((long)b->groups.size(), 0.0, reinterpret_cast<double*>(b->queues.p), st));
    const int nls = std::max(1, std::min(kMaxLaunchStreams, b->launch_streams));
    if (nls > 1) {
        UNO_HIP(hipEventRecord(b->wfork, st));
        for (int w = 0; w + 1 < nls; ++w) UNO_HIP(hipStreamWaitEvent(b->wst[w], b->wfork, 0));
    }
    for (size_t k = 0; k < b->groups.size(); ++k) {
        const int li = (int)(k % (size_t)nls);
        const hipStream_t ls = li == 0 ? st : b->wst[li - 1];
        const uno_batch::Group& g = b->groups[k];
        EriArgs a{};
        a.pairs = b->pairs.p;
        a.prims = b->prims.p;
        a.pairQ = b->pairQ.p;
        a.tau = b->mol[0].ctx->tau;
        a.primcut = b->mol[0].ctx->primcut;
        a.rys2_exact = b->mol[0].ctx->rys2_exact;
        a.prefix = g.prefix;
        a.blk = g.blk;
        a.wpre = g.wpre;
        a.tblk = g.tblk;
        a.ntask = g.tasks ? g.ntask : 0;
        a.qpb = g.qpb;
        a.nquart = g.nq;
        a.e_bra = g.bra;
        a.e_ket = g.ket;
        a.e_mol = g.mol;
        a.mols = b->mdesc.p;
        a.dsh = b->Dsh.p;
        a.dmax_global = 0.0;
        a.D0 = b->D0.p;
        a.W0 = b->W0.p;
        if (g.uhf) {
            a.D1 = b->D1.p;
            a.D2 = b->D2.p;
            a.W1 = b->W1.p;
        }
        // queue 1: the wave-per-quartet classes take their blocks from the queue; queue 2:
        // the lane classes' 64-ket tasks too (one atomic per task)
        const bool q = b->use_queue >= 2 || (b->use_queue == 1 && !g.tasks);
        a.queue = q ? b->queues.p + k : nullptr;
        a.mol_active = b->active_d.p;
        UNO_HIP(eri_launch(g.uhf ? kDigestUHF : kDigestRHF, g.cx, g.cy, a, ls));
    }
    for (int w = 0; w + 1 < nls; ++w) {
        UNO_HIP(hipEventRecord(b->wjoin[w], b->wst[w]));
        UNO_HIP(hipStreamWaitEvent(st, b->wjoin[w], 0));
    }
    return UNO_OK;
}


// batched Fock build on the densities in D0 (D1, D2): W0 (W1) accumulators.  dmax: each
// molecule's largest shell-block density maximum.
static int batch_fock(uno_batch* b, const std::vector<double>& dmax, double& quartets) {
    hipStream_t st = b->st;
    const int nm = (int)b->mol.size();
    bool rebuild = b->groups.empty() && b->space_builds == 0;
    for (int m = 0; m < nm && !rebuild; ++m)
        if (b->mol[m].active && dmax[m] > b->dbound[m]) rebuild = true;
    if (rebuild)
        if (int rc = build_static(b, dmax)) return rc;
    std::vector<int> act(nm);
    for (int m = 0; m < nm; ++m) {
        act[m] = b->mol[m].active ? 1 : 0;
        if (act[m]) quartets += b->mol_quartets[m];   // before the kernels' density screen
    }
    UNO_HIP(hipMemcpyAsync(b->active_d.p, act.data(), nm * sizeof(int), hipMemcpyHostToDevice, st));
    if (int rc = ensure_workers(b)) return rc;
    if (b->use_graph) {
        if (!b->gexec) {
            UNO_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue_fock(b);
            hipGraph_t g = nullptr;
            const hipError_t ee = hipStreamEndCapture(st, &g);
            if (rc || ee != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                if (rc) return rc;
                UNO_HIP(ee);
            }
            b->graph = g;
            UNO_HIP(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
            ++b->graph_builds;
        }
        UNO_HIP(hipGraphLaunch(b->gexec, st));
        ++b->graph_replays;
    } else {
        if (int rc = enqueue_fock(b)) return rc;
    }
    UNO_HIP(hipStreamSynchronize(st));
    return UNO_OK;
}

}  // namespace uno

extern "C" {

int uno_batch_create(int nmol, const uno_basis* bases, const int* nelec, uno_batch** out) {
    if (nmol <= 0 || !bases || !nelec || !out) { set_error("bad argument"); return UNO_EINVAL; }
    auto* b = new uno_batch;
    b->mol.resize(nmol);
    std::vector<PairHdr> pairs;
    std::vector<PrimPair> prims;
    std::vector<double> q;
    std::vector<int> shoff, shn;
    std::vector<MolDesc> md;
    for (int m = 0; m < nmol; ++m) {
        BatchMol& x = b->mol[m];
        if (int rc = uno_ctx_create(&bases[m], 0, &x.ctx)) { delete b; return rc; }
        uno_ctx* c = x.ctx;
        x.n = c->n;
        x.nsh = c->nsh;
        x.nelec = nelec[m];
        x.uhf = nelec[m] % 2 != 0;
        x.nocc[0] = x.uhf ? nelec[m] - nelec[m] / 2 : nelec[m] / 2;
        x.nocc[1] = nelec[m] / 2;
        if (x.n > 256 || nelec[m] <= 0 || x.nocc[0] > x.n) {
            set_error("batched SCF: molecule " + std::to_string(m) + " has more than 256 functions or a bad "
                      "electron count");
            delete b;
            return UNO_EINVAL;
        }
        x.moff = b->store;
        b->store += (long long)x.n * x.n;
        x.dshoff = b->dsh_total;
        b->dsh_total += (long long)x.nsh * x.nsh;
        x.shoff = (int)shoff.size();
        for (int s = 0; s < x.nsh; ++s) {
            shoff.push_back(c->B.off[s]);
            shn.push_back(ncart(c->B.sh[s].l));
        }
        x.pair_off = (int)pairs.size();
        const int p0 = (int)prims.size();
        for (PairHdr h : c->pairs) {
            h.prim0 += p0;
            pairs.push_back(h);
        }
        prims.insert(prims.end(), c->prims.begin(), c->prims.end());
        q.insert(q.end(), c->pairQ.begin(), c->pairQ.end());
        md.push_back(MolDesc{x.moff, x.dshoff, x.n, x.nsh});
        for (int i = 0; i < c->B.ncen; ++i)
            for (int j = i + 1; j < c->B.ncen; ++j) {
                const double* r1 = &c->B.xyz[3 * i];
                const double* r2 = &c->B.xyz[3 * j];
                const double dx = r1[0] - r2[0], dy = r1[1] - r2[1], dz = r1[2] - r2[2];
                x.nuc += c->B.q[i] * c->B.q[j] / std::sqrt(dx * dx + dy * dy + dz * dz);
            }
    }
    int ndev = 0;
    (void)hipGetDeviceCount(&ndev);
    hipError_t e = hipStreamCreate(&b->st);
    if (e != hipSuccess) { set_error("HIP stream creation failed"); delete b; return UNO_EDEVICE; }
    hipStream_t st = b->st;
    auto up = [&](auto& buf, const auto& v) { return buf.upload(v.data(), v.size(), st); };
    if (up(b->pairs, pairs) != hipSuccess || up(b->prims, prims) != hipSuccess || up(b->pairQ, q) != hipSuccess ||
        up(b->sh_off, shoff) != hipSuccess || up(b->sh_n, shn) != hipSuccess || up(b->mdesc, md) != hipSuccess) {
        set_error("batched SCF: device upload failed");
        delete b;
        return UNO_ENOMEM;
    }
    const size_t S = (size_t)b->store;
    DBuf<double>* stores[] = {&b->H, &b->S, &b->X, &b->P[0], &b->P[1], &b->Pold[0], &b->Pold[1], &b->F[0], &b->F[1],
                              &b->G[0], &b->G[1], &b->C[0], &b->C[1], &b->Fp, &b->Tm, &b->Tm2, &b->V, &b->Z,
                              &b->D0, &b->D1, &b->D2, &b->W0, &b->W1, &b->Wp[0], &b->Wp[1]};
    for (DBuf<double>* p : stores)
        if (p->alloc(S) != hipSuccess) { set_error("batched SCF: device allocation failed"); delete b; return UNO_ENOMEM; }
    for (int k = 0; k < kBatchDiis; ++k)
        for (int s = 0; s < 2; ++s)
            if (b->dF[k][s].alloc(S) != hipSuccess || b->dE[k][s].alloc(S) != hipSuccess) {
                set_error("batched SCF: device allocation failed");
                delete b;
                return UNO_ENOMEM;
            }
    if (b->Dsh.alloc(std::max<long long>(1, b->dsh_total)) != hipSuccess || b->ev[0].alloc(S) != hipSuccess ||
        b->ev[1].alloc(S) != hipSuccess || b->red.alloc(8 * (size_t)nmol * (kBatchDiis + 4)) != hipSuccess) {
        set_error("batched SCF: device allocation failed");
        delete b;
        return UNO_ENOMEM;
    }
    *out = b;
    return UNO_OK;
}

int uno_batch_destroy(uno_batch* b) {
    delete b;
    return UNO_OK;
}

int uno_batch_nmol(const uno_batch* b) { return b ? (int)b->mol.size() : 0; }

int uno_batch_set(uno_batch* b, const char* key, double v) {
    if (!b || !key) { set_error("null argument"); return UNO_EINVAL; }
    const std::string k(key);
    if (k == "graph") b->use_graph = (int)v;
    else if (k == "queue") b->use_queue = (int)v;
    else if (k == "launch_streams") {
        b->launch_streams = std::max(1, std::min(kMaxLaunchStreams, (int)v));
        b->drop_graph();   // a captured build keeps the stream layout it was captured with
    }
    else { set_error("unknown batch option " + k); return UNO_EINVAL; }
    b->drop_graph();
    return UNO_OK;
}

int uno_batch_get(const uno_batch* b, const char* key, double* v) {
    if (!b || !key || !v) { set_error("null argument"); return UNO_EINVAL; }
    const std::string k(key);
    if (k == "graph") *v = b->use_graph;
    else if (k == "queue") *v = b->use_queue;
    else if (k == "launch_streams") *v = b->launch_streams;
    else if (k == "space_builds") *v = b->space_builds;
    else if (k == "graph_builds") *v = b->graph_builds;
    else if (k == "graph_replays") *v = b->graph_replays;
    else if (k == "launches_per_build") *v = (double)b->groups.size();
    else { set_error("unknown batch option " + k); return UNO_EINVAL; }
    return UNO_OK;
}

int uno_batch_density(const uno_batch* b, int m, double* P) {
    if (!b || !P || m < 0 || m >= (int)b->mol.size()) { set_error("bad argument"); return UNO_EINVAL; }
    const BatchMol& x = b->mol[m];
    const size_t nn = (size_t)x.n * x.n;
    if (!b->P[0].p) { set_error("the batch has not run"); return UNO_EINVAL; }
    UNO_HIP(hipMemcpyAsync(P, b->P[0].p + x.moff, nn * sizeof(double), hipMemcpyDeviceToHost, b->st));
    if (x.uhf) UNO_HIP(hipMemcpyAsync(P + nn, b->P[1].p + x.moff, nn * sizeof(double), hipMemcpyDeviceToHost, b->st));
    UNO_HIP(hipStreamSynchronize(b->st));
    return UNO_OK;
}

int uno_batch_run(uno_batch* b, const uno_scf_opts* opts, uno_scf_result* res, double* energies) {
    if (!b || !res) { set_error("null argument"); return UNO_EINVAL; }
    uno_scf_opts o;
    if (opts) o = *opts; else uno_scf_defaults(&o);
    if (o.precision || o.incremental || o.guess != 0) {
        set_error("batched SCF: FP64 full builds from the core-Hamiltonian guess only");
        return UNO_EINVAL;
    }
    const auto t0 = clk::now();
    hipStream_t st = b->st;
    b->scratch.reset();
    b->warm[0] = b->warm[1] = false;
    // static spaces and the captured build belong to one run (its density bounds)
    b->drop_graph();
    b->groups.clear();
    b->dbound.clear();
    b->space_builds = b->graph_builds = b->graph_replays = 0;
    const int nm = (int)b->mol.size();
    const long long S = b->store;
    // per-molecule eigenvalue offsets into ev stores: reuse the matrix offsets (n <= n^2)
    // one-electron matrices of every molecule
    for (BatchMol& x : b->mol) {
        x.active = true;
        x.converged = false;
        x.iterations = 0;
        x.eold = x.energy = x.ediff = x.pdiff = 0.0;
        if (int rc = oneint_device(x.ctx, b->S.p + x.moff, b->Tm.p + x.moff, b->H.p + x.moff)) return rc;
        UNO_HIP(hipStreamSynchronize(x.ctx->st));
    }
    for (int s = 0; s < 2; ++s) {
        UNO_HIP(k_fill(S, 0.0, b->P[s].p, st));
        UNO_HIP(k_fill(S, 0.0, b->G[s].p, st));
    }
    std::vector<MatSlot> sl[2];
    std::vector<int> who[2];
    std::vector<DBuf<MatSlot>> dsl(2);
    auto refresh_slots = [&]() -> int {
        for (int s = 0; s < 2; ++s) {
            make_slots(b, true, s, sl[s], who[s]);
            if (!sl[s].empty()) UNO_HIP(dsl[s].upload(sl[s].data(), sl[s].size(), st));
        }
        return UNO_OK;
    };
    if (int rc = refresh_slots()) return rc;
    auto evoff = [&](int s) {
        std::vector<long long> o;
        for (const MatSlot& m : sl[s]) o.push_back(m.off);
        return o;
    };
    // X = U s^-1/2 (RHF.hpp:214-233): Jacobi of S, column scaling on the host read-back
    {
        UNO_HIP(hipMemcpyAsync(b->Tm.p, b->S.p, S * sizeof(double), hipMemcpyDeviceToDevice, st));
        if (int rc = batch_eigh(b, sl[0], b->Tm.p, b->ev[0].p, evoff(0))) return rc;
        std::vector<double> ev(S), Z(S);
        UNO_HIP(hipMemcpyAsync(ev.data(), b->ev[0].p, S * sizeof(double), hipMemcpyDeviceToHost, st));
        UNO_HIP(hipMemcpyAsync(Z.data(), b->Z.p, S * sizeof(double), hipMemcpyDeviceToHost, st));
        UNO_HIP(hipStreamSynchronize(st));
        for (const BatchMol& x : b->mol) {
            for (int j = 0; j < x.n; ++j) {
                double f = ev[x.moff + j];
                if (std::fabs(f) < 1e-7 && f < 0.0) f = -f;
                if (!(f > 0.0)) { set_error("Zero or negative eigenvalue in an overlap matrix"); return UNO_ESINGULAR; }
                const double sc = 1.0 / std::sqrt(f);
                for (int i = 0; i < x.n; ++i) Z[x.moff + (long long)i * x.n + j] *= sc;
            }
        }
        UNO_HIP(b->X.upload(Z.data(), S, st));
    }
    // F (store, per spin) -> P per spin slot: X^T F X, Jacobi, C = X W, P = C_occ C_occ^T
    double diag_s = 0.0;
    auto diag_to_P = [&](int s, const double* Fs) -> int {
        if (sl[s].empty()) return UNO_OK;
        const auto td = clk::now();
        struct Acc { double& t; clk::time_point t0; ~Acc() { t += secs(t0, clk::now()); } } acc{diag_s, td};
        GemmBatch g1(b->scratch), g2(b->scratch);
        for (const MatSlot& m : sl[s]) g1.add(m.n, m.n, m.n, false, false, Fs + m.off, b->X.p + m.off, b->Tm.p + m.off, m.n, m.n, m.n);
        if (int rc = g1.run(1.0, 0.0, st)) return rc;
        for (const MatSlot& m : sl[s]) g2.add(m.n, m.n, m.n, true, false, b->X.p + m.off, b->Tm.p + m.off, b->Fp.p + m.off, m.n, m.n, m.n);
        if (int rc = g2.run(1.0, 0.0, st)) return rc;
        // after their first eigensolve the molecules start from their last eigenvectors Wp:
        // in that basis the matrix is nearly diagonal and the Jacobi sweeps settle in a few
        // rounds (W = Wp Z after)
        std::vector<char> warm(sl[s].size(), 0);
        GemmBatch w1(b->scratch), w2(b->scratch), w3(b->scratch);
        for (size_t k = 0; k < sl[s].size(); ++k) {
            const MatSlot& m = sl[s][k];
            warm[k] = b->warm[s];
            if (warm[k]) {
                w1.add(m.n, m.n, m.n, false, false, b->Fp.p + m.off, b->Wp[s].p + m.off, b->Tm.p + m.off, m.n, m.n, m.n);
                w2.add(m.n, m.n, m.n, true, false, b->Wp[s].p + m.off, b->Tm.p + m.off, b->Tm2.p + m.off, m.n, m.n, m.n);
                w3.add(m.n, m.n, m.n, false, false, b->Wp[s].p + m.off, b->Z.p + m.off, b->Tm.p + m.off, m.n, m.n, m.n);
            }
        }
        if (int rc = w1.run(1.0, 0.0, st)) return rc;
        if (int rc = w2.run(1.0, 0.0, st)) return rc;
        {
            std::vector<BatchMat> bm(sl[s].size());
            for (size_t k = 0; k < sl[s].size(); ++k) {
                const MatSlot& m = sl[s][k];
                bm[k] = BatchMat{m.n, (warm[k] ? b->Tm2.p : b->Fp.p) + m.off, b->V.p + m.off, b->ev[s].p + m.off,
                                 b->Z.p + m.off};
            }
            BatchMat* d = nullptr;
            UNO_HIP(b->scratch.upload(bm.data(), bm.size(), st, &d));
            UNO_HIP(batched_jacobi((int)bm.size(), d, 50, nullptr, st));
            // bm is pageable host memory: its upload must be done before it goes
            UNO_HIP(hipStreamSynchronize(st));
        }
        if (int rc = w3.run(1.0, 0.0, st)) return rc;
        // the new starting bases in one launch: Wp Z (warm, in Tm) or Z (cold)
        {
            std::vector<double> coef(2 * sl[s].size());
            for (size_t k = 0; k < sl[s].size(); ++k) {
                coef[2 * k] = warm[k] ? 1.0 : 0.0;
                coef[2 * k + 1] = warm[k] ? 0.0 : 1.0;
            }
            const double* src[2] = {b->Tm.p, b->Z.p};
            double* dcoef = nullptr;
            const double** dsrc = nullptr;
            UNO_HIP(b->scratch.upload(coef.data(), coef.size(), st, &dcoef));
            UNO_HIP(b->scratch.upload(src, 2, st, &dsrc));
            UNO_HIP(batch_lincomb((int)sl[s].size(), dsl[s].p, 2, dcoef, dsrc, b->Wp[s].p, st));
            UNO_HIP(hipStreamSynchronize(st));
        }
        b->warm[s] = true;
        GemmBatch g3(b->scratch), g4(b->scratch);
        for (const MatSlot& m : sl[s]) g3.add(m.n, m.n, m.n, false, false, b->X.p + m.off, b->Wp[s].p + m.off, b->C[s].p + m.off, m.n, m.n, m.n);
        if (int rc = g3.run(1.0, 0.0, st)) return rc;
        for (size_t k = 0; k < sl[s].size(); ++k) {
            const MatSlot& m = sl[s][k];
            g4.add(m.n, m.n, b->mol[who[s][k]].nocc[s], false, true, b->C[s].p + m.off, b->C[s].p + m.off,
                   b->P[s].p + m.off, m.n, m.n, m.n);
        }
        return g4.run(1.0, 0.0, st);
    };
    // core-Hamiltonian guess (UHF beta reuses the alpha orbitals, UHF.hpp:206-215)
    if (int rc = diag_to_P(0, b->H.p)) return rc;
    if (!sl[1].empty()) {
        GemmBatch g(b->scratch);
        for (size_t k = 0; k < sl[1].size(); ++k) {
            const MatSlot& m = sl[1][k];
            g.add(m.n, m.n, b->mol[who[1][k]].nocc[1], false, true, b->C[0].p + m.off, b->C[0].p + m.off,
                  b->P[1].p + m.off, m.n, m.n, m.n);
        }
        if (int rc = g.run(1.0, 0.0, st)) return rc;
    }
    double eps = o.eps;
    std::vector<double> hs(4 * (size_t)nm + 8);
    int ndiis = 0, slot_next = 0;
    std::vector<double> Bm((size_t)nm * kBatchDiis * kBatchDiis, 0.0);
    int iteration = 0;
    double fock_s = 0.0;
    double quartets = 0.0;
    int nstopped = 0;
    double wall99 = -1.0;
    while (true) {
        b->scratch.reset();   // the previous iteration ended synchronised
        // densities of the build: D1 = Pa, D2 = Pb (zero for closed shells), D0 = D1 + D2
        const auto tf = clk::now();
        UNO_HIP(hipMemcpyAsync(b->D1.p, b->P[0].p, S * sizeof(double), hipMemcpyDeviceToDevice, st));
        UNO_HIP(hipMemcpyAsync(b->D2.p, b->P[1].p, S * sizeof(double), hipMemcpyDeviceToDevice, st));
        UNO_HIP(k_axpby(S, 1.0, b->D1.p, 1.0, b->D2.p, b->D0.p, st));
        UNO_HIP(batch_shell_max((int)sl[0].size(), dsl[0].p, b->sh_off.p, b->sh_n.p, b->D0.p, b->D1.p, b->D2.p,
                                b->Dsh.p, st));
        std::vector<double> dsh(b->dsh_total);
        UNO_HIP(hipMemcpyAsync(dsh.data(), b->Dsh.p, dsh.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        UNO_HIP(hipStreamSynchronize(st));
        std::vector<double> dmax(nm, 0.0);
        for (int m = 0; m < nm; ++m)
            if (b->mol[m].active)
                for (long long k = 0; k < (long long)b->mol[m].nsh * b->mol[m].nsh; ++k)
                    dmax[m] = std::max(dmax[m], dsh[b->mol[m].dshoff + k]);
        if (int rc = batch_fock(b, dmax, quartets)) return rc;
        UNO_HIP(batch_sym((int)sl[0].size(), dsl[0].p, b->W0.p, b->G[0].p, false, st));
        if (!sl[1].empty()) UNO_HIP(batch_sym((int)sl[1].size(), dsl[1].p, b->W1.p, b->G[1].p, false, st));
        fock_s += secs(tf, clk::now());
        // energies: e1 = tr(Pa H) + tr(Pb H), e2 = tr(Pa Ga) + tr(Pb Gb)
        double* r = b->red.p;
        const int n0 = (int)sl[0].size();
        UNO_HIP(batch_dot(n0, dsl[0].p, b->P[0].p, b->H.p, r, false, st));
        UNO_HIP(batch_dot(n0, dsl[0].p, b->P[1].p, b->H.p, r, true, st));
        UNO_HIP(batch_dot(n0, dsl[0].p, b->P[0].p, b->G[0].p, r + nm, false, st));
        UNO_HIP(batch_dot(n0, dsl[0].p, b->P[1].p, b->G[1].p, r + nm, true, st));
        UNO_HIP(hipMemcpyAsync(hs.data(), r, 2 * (size_t)nm * sizeof(double), hipMemcpyDeviceToHost, st));
        UNO_HIP(hipStreamSynchronize(st));
        for (int k = 0; k < n0; ++k) {
            BatchMol& x = b->mol[who[0][k]];
            const double e1 = hs[k], e2 = hs[nm + k];
            x.energy = x.uhf ? e1 + 0.5 * e2 : 2.0 * e1 + e2;
            x.ediff = x.energy - x.eold;
            x.eold = x.energy;
            if (iteration == 0) x.init_energy = x.energy + x.nuc;
        }
        // F = H + G; DIIS on every active slot (one shared history position per iteration)
        for (int s = 0; s < 2; ++s) UNO_HIP(k_axpby(S, 1.0, b->H.p, 1.0, b->G[s].p, b->F[s].p, st));
        const int slot = slot_next;
        slot_next = (slot_next + 1) % kBatchDiis;
        if (ndiis < kBatchDiis) ++ndiis;
        for (int s = 0; s < 2; ++s) {
            if (sl[s].empty()) continue;
            UNO_HIP(hipMemcpyAsync(b->dF[slot][s].p, b->F[s].p, S * sizeof(double), hipMemcpyDeviceToDevice, st));
            GemmBatch g1(b->scratch), g2(b->scratch), g3(b->scratch), g4(b->scratch);
            for (const MatSlot& m : sl[s]) g1.add(m.n, m.n, m.n, false, false, b->F[s].p + m.off, b->P[s].p + m.off, b->Tm.p + m.off, m.n, m.n, m.n);
            if (int rc = g1.run(1.0, 0.0, st)) return rc;
            for (const MatSlot& m : sl[s]) g2.add(m.n, m.n, m.n, false, false, b->Tm.p + m.off, b->S.p + m.off, b->Fp.p + m.off, m.n, m.n, m.n);
            if (int rc = g2.run(1.0, 0.0, st)) return rc;
            // FPS - SPF = FPS - (FPS)^T
            UNO_HIP(batch_sym((int)sl[s].size(), dsl[s].p, b->Fp.p, b->Tm2.p, false, st));
            UNO_HIP(k_axpby(S, 2.0, b->Fp.p, -1.0, b->Tm2.p, b->Tm2.p, st));
            for (const MatSlot& m : sl[s]) g3.add(m.n, m.n, m.n, false, false, b->Tm2.p + m.off, b->X.p + m.off, b->Tm.p + m.off, m.n, m.n, m.n);
            if (int rc = g3.run(1.0, 0.0, st)) return rc;
            for (const MatSlot& m : sl[s]) g4.add(m.n, m.n, m.n, true, false, b->X.p + m.off, b->Tm.p + m.off, b->dE[slot][s].p + m.off, m.n, m.n, m.n);
            if (int rc = g4.run(1.0, 0.0, st)) return rc;
        }
        // B-matrix row of the new vector: sum over spins of <E_new, E_j>
        std::vector<double> row((size_t)kBatchDiis * nm, 0.0);
        for (int j = 0; j < ndiis; ++j) {
            UNO_HIP(batch_dot(n0, dsl[0].p, b->dE[slot][0].p, b->dE[j][0].p, r, false, st));
            std::vector<double> t0v(n0);
            UNO_HIP(hipMemcpyAsync(t0v.data(), r, n0 * sizeof(double), hipMemcpyDeviceToHost, st));
            std::vector<double> t1v(sl[1].size());
            if (!sl[1].empty()) {
                UNO_HIP(batch_dot((int)sl[1].size(), dsl[1].p, b->dE[slot][1].p, b->dE[j][1].p, r + nm, false, st));
                UNO_HIP(hipMemcpyAsync(t1v.data(), r + nm, t1v.size() * sizeof(double), hipMemcpyDeviceToHost, st));
            }
            UNO_HIP(hipStreamSynchronize(st));
            for (int k = 0; k < n0; ++k) row[(size_t)who[0][k] * kBatchDiis + j] = t0v[k];
            for (size_t k = 0; k < t1v.size(); ++k) row[(size_t)who[1][k] * kBatchDiis + j] += t1v[k];
        }
        std::vector<double> coef((size_t)n0 * kBatchDiis, 0.0);
        bool any = false;
        for (int k = 0; k < n0; ++k) {
            const int m = who[0][k];
            double* Bmm = &Bm[(size_t)m * kBatchDiis * kBatchDiis];
            for (int j = 0; j < ndiis; ++j) Bmm[slot * kBatchDiis + j] = Bmm[j * kBatchDiis + slot] = row[(size_t)m * kBatchDiis + j];
            coef[(size_t)k * kBatchDiis + slot] = 1.0;
            if (ndiis >= 2 && diis_coefficients(Bmm, kBatchDiis, ndiis, slot, &coef[(size_t)k * kBatchDiis]) >= 2)
                any = true;
        }
        if (any) {
            for (int s = 0; s < 2; ++s) {
                if (sl[s].empty()) continue;
                std::vector<const double*> src(kBatchDiis);
                for (int j = 0; j < kBatchDiis; ++j) src[j] = b->dF[j][s].p;
                const double** dsrc = nullptr;
                UNO_HIP(b->scratch.upload(src.data(), src.size(), st, &dsrc));
                // coefficients of slot k of spin s: the molecule's row (spin 1 slots are a subset)
                std::vector<double> cs((size_t)sl[s].size() * kBatchDiis);
                for (size_t k = 0; k < sl[s].size(); ++k) {
                    const int m = who[s][k];
                    const size_t k0 = std::find(who[0].begin(), who[0].end(), m) - who[0].begin();
                    std::copy(&coef[k0 * kBatchDiis], &coef[k0 * kBatchDiis] + kBatchDiis, &cs[k * kBatchDiis]);
                }
                double* dcs = nullptr;
                UNO_HIP(b->scratch.upload(cs.data(), cs.size(), st, &dcs));
                UNO_HIP(batch_lincomb((int)sl[s].size(), dsl[s].p, kBatchDiis, dcs, dsrc, b->F[s].p, st));
                UNO_HIP(hipStreamSynchronize(st));
            }
        }
        // new densities
        for (int s = 0; s < 2; ++s)
            UNO_HIP(hipMemcpyAsync(b->Pold[s].p, b->P[s].p, S * sizeof(double), hipMemcpyDeviceToDevice, st));
        for (int s = 0; s < 2; ++s)
            if (int rc = diag_to_P(s, b->F[s].p)) return rc;
        UNO_HIP(batch_diffnorm(n0, dsl[0].p, b->P[0].p, b->Pold[0].p, r, st));
        if (!sl[1].empty()) UNO_HIP(batch_diffnorm((int)sl[1].size(), dsl[1].p, b->P[1].p, b->Pold[1].p, r + nm, st));
        std::vector<double> pd(2 * (size_t)nm, 0.0);
        UNO_HIP(hipMemcpyAsync(pd.data(), r, 2 * (size_t)nm * sizeof(double), hipMemcpyDeviceToHost, st));
        UNO_HIP(hipStreamSynchronize(st));
        ++iteration;
        bool changed = false;
        for (int k = 0; k < n0; ++k) {
            BatchMol& x = b->mol[who[0][k]];
            x.pdiff = pd[k];
            x.iterations = iteration;
        }
        for (size_t k = 0; k < sl[1].size(); ++k) b->mol[who[1][k]].pdiff += pd[nm + k];
        for (int k = 0; k < n0; ++k) {
            BatchMol& x = b->mol[who[0][k]];
            bool conv = false;
            if (iteration >= 2) {
                if (o.cflag == 0) conv = x.ediff < eps;
                else if (o.cflag == 1) conv = x.pdiff < eps;
                else if (o.cflag == 3) conv = std::fabs(x.ediff) < eps;
                else conv = x.pdiff < eps && x.ediff < eps;
            }
            if (conv || iteration >= o.maxits) {
                x.converged = conv;
                x.active = false;
                changed = true;
                ++nstopped;
            }
        }
        if (wall99 < 0.0 && 100 * nstopped >= 99 * nm) wall99 = secs(t0, clk::now());
        if (o.verbose)
            fprintf(stderr, "[batch] iteration %d: %d molecules were active, %d stopped; Fock %.3f s, "
                            "eigen/GEMM %.3f s, wall %.3f s\n", iteration, n0, nstopped, fock_s, diag_s,
                    secs(t0, clk::now()));
        bool any_active = false;
        for (const BatchMol& x : b->mol) any_active = any_active || x.active;
        if (!any_active) break;
        if (changed) {
            if (int rc = refresh_slots()) return rc;
        }
    }
    const double wall = secs(t0, clk::now());
    int rc = UNO_OK;
    for (int m = 0; m < nm; ++m) {
        const BatchMol& x = b->mol[m];
        uno_scf_result& r = res[m];
        std::memset(&r, 0, sizeof(r));
        r.energy = x.energy + x.nuc;
        r.init_energy = x.init_energy;
        r.ediff = x.ediff;
        r.pdiff = x.pdiff;
        r.nucrep = x.nuc;
        r.electronic = x.energy;
        r.iterations = x.iterations;
        r.converged = x.converged ? 1 : 0;
        r.uhf = x.uhf ? 1 : 0;
        r.nfock = x.iterations;
        r.wall_s = wall;
        r.wall99_s = wall99 < 0.0 ? wall : wall99;
        r.fock_s = fock_s;
        r.diag_s = diag_s;
        r.quartets = quartets;
        if (energies) energies[m] = r.energy;
        if (!x.converged) rc = UNO_ENOCONV;
    }
    return rc;
}

}  // extern "C"
