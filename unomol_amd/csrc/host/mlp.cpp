// The density-guess network: a multilayer perceptron (SiLU hidden layers, linear output)
// trained and evaluated on the GPU with the BF16 MFMA kernels of mlp.hip.  Master
// parameters, gradients and Adam moments are FP32; GEMM operands (inputs, activations,
// weights, back-propagated errors) are BF16; every product accumulates in FP32.
//
// Parameter order (also the file format after the header): for each layer l,
// W_l [dims[l] x dims[l+1]] row-major, then b_l [dims[l+1]].
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>
#include "engine.hpp"
#include "../hip/mlp_kernels.hpp"

using namespace uno;

struct uno_mlp {
    std::vector<int> dims;               // in, hidden..., out
    std::vector<long> woff, boff;
    long np = 0;
    hipStream_t st = nullptr;
    DBuf<float> p, g, m, v;
    DBuf<__bf16> pb;
    int step = 0;
    // batch workspaces
    int bcap = 0;
    DBuf<__bf16> xb, dzb[2];
    std::vector<DBuf<float>> z;
    std::vector<DBuf<__bf16>> h;
    DBuf<float> y, dy, dz[2], part, tb, mb;
    DBuf<float> xin;   // inference inputs, kept between calls (no allocation per call)
    DBuf<int> idx;
    ~uno_mlp() {
        if (st) (void)hipStreamDestroy(st);
    }
};

namespace {

constexpr uint32_t kMagic = 0x504c4d55u;   // "UMLP"

int nlayers(const uno_mlp* m) { return (int)m->dims.size() - 1; }

int init_layout(uno_mlp* m) {
    const int L = nlayers(m);
    m->woff.assign(L, 0);
    m->boff.assign(L, 0);
    long o = 0;
    for (int l = 0; l < L; ++l) {
        m->woff[l] = o;
        o += (long)m->dims[l] * m->dims[l + 1];
        m->boff[l] = o;
        o += m->dims[l + 1];
    }
    m->np = o;
    UNO_HIP(hipStreamCreate(&m->st));
    UNO_HIP(m->p.alloc(o));
    UNO_HIP(m->g.alloc(o));
    UNO_HIP(m->m.alloc(o));
    UNO_HIP(m->v.alloc(o));
    UNO_HIP(m->pb.alloc(o));
    UNO_HIP(hipMemsetAsync(m->m.p, 0, o * sizeof(float), m->st));
    UNO_HIP(hipMemsetAsync(m->v.p, 0, o * sizeof(float), m->st));
    return UNO_OK;
}

int upload_params(uno_mlp* m, const float* hp) {
    UNO_HIP(m->p.upload(hp, m->np, m->st));
    UNO_HIP(to_bf16(m->np, m->p.p, m->pb.p, m->st));
    UNO_HIP(hipStreamSynchronize(m->st));
    return UNO_OK;
}

int ensure_batch(uno_mlp* m, int B) {
    if (B <= m->bcap) return UNO_OK;
    const int L = nlayers(m);
    int wmax = 0;
    for (int d : m->dims) wmax = std::max(wmax, d);
    m->z.resize(L);
    m->h.resize(L);
    for (int l = 0; l < L; ++l) {
        UNO_HIP(m->z[l].alloc((size_t)B * m->dims[l + 1]));
        UNO_HIP(m->h[l].alloc((size_t)B * m->dims[l + 1]));
    }
    UNO_HIP(m->xb.alloc((size_t)B * m->dims[0]));
    for (int k = 0; k < 2; ++k) {
        UNO_HIP(m->dz[k].alloc((size_t)B * wmax));
        UNO_HIP(m->dzb[k].alloc((size_t)B * wmax));
    }
    const int out = m->dims[L];
    UNO_HIP(m->y.alloc((size_t)B * out));
    UNO_HIP(m->dy.alloc((size_t)B * out));
    UNO_HIP(m->tb.alloc((size_t)B * out));
    UNO_HIP(m->mb.alloc((size_t)B * out));
    UNO_HIP(m->part.alloc(((size_t)B * out + 255) / 256));
    UNO_HIP(m->idx.alloc(B));
    m->bcap = B;
    return UNO_OK;
}

// forward over the B rows staged in xb; outputs in y
int forward(uno_mlp* m, int B) {
    const int L = nlayers(m);
    const __bf16* in = m->xb.p;
    for (int l = 0; l < L; ++l) {
        MlpEpilogue ep{};
        ep.ldc = m->dims[l + 1];
        ep.bias = m->p.p + m->boff[l];
        const bool last = l == L - 1;
        if (!last) {
            ep.out_f32 = m->z[l].p;
            ep.out_bf16 = m->h[l].p;
        } else {
            ep.out_f32 = m->y.p;
        }
        UNO_HIP(gemm_bf16(false, false, last ? kEpiStore : kEpiBiasSilu, B, m->dims[l + 1], m->dims[l], in, m->dims[l],
                          m->pb.p + m->woff[l], m->dims[l + 1], ep, m->st));
        in = m->h[l].p;
    }
    return UNO_OK;
}

// loss on the staged batch (targets tb, mask mb, norm = sum of the mask) and gradients in g
int backward(uno_mlp* m, int B, float inv_norm, int* nblk) {
    const int L = nlayers(m);
    const int out = m->dims[L];
    UNO_HIP(mse_loss((long)B * out, m->y.p, m->tb.p, m->mb.p, inv_norm, m->dz[0].p, m->dzb[0].p, m->part.p, nblk,
                     m->st));
    int cur = 0;
    for (int l = L - 1; l >= 0; --l) {
        const __bf16* a_in = l == 0 ? m->xb.p : m->h[l - 1].p;
        MlpEpilogue ew{};
        ew.ldc = m->dims[l + 1];
        ew.out_f32 = m->g.p + m->woff[l];
        UNO_HIP(gemm_bf16(true, false, kEpiStore, m->dims[l], m->dims[l + 1], B, a_in, m->dims[l], m->dzb[cur].p,
                          m->dims[l + 1], ew, m->st));
        UNO_HIP(col_sum(B, m->dims[l + 1], m->dz[cur].p, m->g.p + m->boff[l], m->st));
        if (l > 0) {
            MlpEpilogue eh{};
            eh.ldc = m->dims[l];
            eh.aux_f32 = m->z[l - 1].p;
            eh.out_f32 = m->dz[1 - cur].p;
            eh.out_bf16 = m->dzb[1 - cur].p;
            UNO_HIP(gemm_bf16(false, true, kEpiSiluGrad, B, m->dims[l], m->dims[l + 1], m->dzb[cur].p, m->dims[l + 1],
                              m->pb.p + m->woff[l], m->dims[l + 1], eh, m->st));
            cur = 1 - cur;
        }
    }
    return UNO_OK;
}

float mask_sum(const float* M, const std::vector<int>& rows, int out) {
    double s = 0.0;
    for (int r : rows)
        if (r >= 0)
            for (int c = 0; c < out; ++c) s += M[(long)r * out + c];
    return (float)s;
}

// stage rows of host arrays (X as BF16; T, M as f32) through a device copy of the rows
int stage_rows(uno_mlp* m, const std::vector<int>& rows, const DBuf<float>& X, const DBuf<float>* T,
               const DBuf<float>* M) {
    const int B = (int)rows.size();
    UNO_HIP(m->idx.upload(rows.data(), B, m->st));
    UNO_HIP(gather_rows(B, m->dims[0], m->idx.p, X.p, m->xb.p, nullptr, m->st));
    const int out = m->dims.back();
    if (T) UNO_HIP(gather_rows(B, out, m->idx.p, T->p, nullptr, m->tb.p, m->st));
    if (M) UNO_HIP(gather_rows(B, out, m->idx.p, M->p, nullptr, m->mb.p, m->st));
    return UNO_OK;
}

int round_up(int x, int a) { return (x + a - 1) / a * a; }

}  // namespace

extern "C" {

int uno_mlp_create(int in_dim, int width, int hidden_layers, int out_dim, unsigned long long seed, uno_mlp** out) {
    if (!out || in_dim <= 0 || width <= 0 || hidden_layers < 0 || out_dim <= 0 || in_dim % 64 || width % 64 ||
        out_dim % 64) {
        set_error("MLP dimensions must be positive multiples of 64");
        return UNO_EINVAL;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        set_error("no HIP device is visible (the MLP kernels need a gfx950 GPU)");
        return UNO_EDEVICE;
    }
    auto* m = new uno_mlp;
    m->dims.push_back(in_dim);
    for (int l = 0; l < hidden_layers; ++l) m->dims.push_back(width);
    m->dims.push_back(out_dim);
    if (int rc = init_layout(m)) { delete m; return rc; }
    // LeCun-normal weights, zero biases
    std::vector<float> hp(m->np, 0.0f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < nlayers(m); ++l) {
        std::normal_distribution<float> nd(0.0f, 1.0f / std::sqrt((float)m->dims[l]));
        for (long k = 0; k < (long)m->dims[l] * m->dims[l + 1]; ++k) hp[m->woff[l] + k] = nd(rng);
    }
    if (int rc = upload_params(m, hp.data())) { delete m; return rc; }
    *out = m;
    return UNO_OK;
}

int uno_mlp_destroy(uno_mlp* m) {
    delete m;
    return UNO_OK;
}

int uno_mlp_nparams(const uno_mlp* m, long* n) {
    if (!m || !n) { set_error("null argument"); return UNO_EINVAL; }
    *n = m->np;
    return UNO_OK;
}

int uno_mlp_dims(const uno_mlp* m, int* ndims, int* dims) {
    if (!m || !ndims) { set_error("null argument"); return UNO_EINVAL; }
    if (dims)
        for (size_t i = 0; i < m->dims.size() && (int)i < *ndims; ++i) dims[i] = m->dims[i];
    *ndims = (int)m->dims.size();
    return UNO_OK;
}

int uno_mlp_get_params(uno_mlp* m, float* p) {
    if (!m || !p) { set_error("null argument"); return UNO_EINVAL; }
    UNO_HIP(hipMemcpyAsync(p, m->p.p, m->np * sizeof(float), hipMemcpyDeviceToHost, m->st));
    UNO_HIP(hipStreamSynchronize(m->st));
    return UNO_OK;
}

int uno_mlp_set_params(uno_mlp* m, const float* p) {
    if (!m || !p) { set_error("null argument"); return UNO_EINVAL; }
    return upload_params(m, p);
}

int uno_mlp_forward(uno_mlp* m, long n, const float* X, float* Y) {
    if (!m || !X || !Y || n <= 0) { set_error("bad argument"); return UNO_EINVAL; }
    const int in = m->dims[0], out = m->dims.back();
    const int chunk = 4096;
    if (int rc = ensure_batch(m, chunk)) return rc;
    DBuf<float>& dX = m->xin;
    UNO_HIP(dX.upload(X, (size_t)n * in, m->st));
    for (long r0 = 0; r0 < n; r0 += chunk) {
        const int cnt = (int)std::min<long>(chunk, n - r0);
        const int B = round_up(cnt, 64);
        std::vector<int> rows(B, -1);
        for (int r = 0; r < cnt; ++r) rows[r] = (int)(r0 + r);
        if (int rc = stage_rows(m, rows, dX, nullptr, nullptr)) return rc;
        if (int rc = forward(m, B)) return rc;
        UNO_HIP(hipMemcpyAsync(Y + r0 * out, m->y.p, (size_t)cnt * out * sizeof(float), hipMemcpyDeviceToHost, m->st));
        UNO_HIP(hipStreamSynchronize(m->st));
    }
    return UNO_OK;
}

int uno_mlp_loss_grad(uno_mlp* m, long n, const float* X, const float* T, const float* M, float* loss,
                      float* grad) {
    if (!m || !X || !T || !M || n <= 0 || n > 65536) { set_error("bad argument (1 <= n <= 65536)"); return UNO_EINVAL; }
    const int in = m->dims[0], out = m->dims.back();
    const int B = round_up((int)n, 64);
    if (int rc = ensure_batch(m, B)) return rc;
    DBuf<float> dX, dT, dM;
    UNO_HIP(dX.upload(X, (size_t)n * in, m->st));
    UNO_HIP(dT.upload(T, (size_t)n * out, m->st));
    UNO_HIP(dM.upload(M, (size_t)n * out, m->st));
    std::vector<int> rows(B, -1);
    std::iota(rows.begin(), rows.begin() + n, 0);
    if (int rc = stage_rows(m, rows, dX, &dT, &dM)) return rc;
    const float norm = mask_sum(M, rows, out);
    if (!(norm > 0.0f)) { set_error("mask is empty"); return UNO_EINVAL; }
    if (int rc = forward(m, B)) return rc;
    int nblk = 0;
    if (int rc = backward(m, B, 1.0f / norm, &nblk)) return rc;
    std::vector<float> part(nblk);
    UNO_HIP(hipMemcpyAsync(part.data(), m->part.p, nblk * sizeof(float), hipMemcpyDeviceToHost, m->st));
    if (grad) UNO_HIP(hipMemcpyAsync(grad, m->g.p, m->np * sizeof(float), hipMemcpyDeviceToHost, m->st));
    UNO_HIP(hipStreamSynchronize(m->st));
    double s = 0.0;
    for (float v : part) s += v;
    if (loss) *loss = (float)(s / norm);
    return UNO_OK;
}

int uno_mlp_train(uno_mlp* m, long n, const float* X, const float* T, const float* M, int steps, int batch,
                  float lr, unsigned long long seed, float* loss) {
    if (!m || !X || !T || !M || n <= 0 || steps <= 0 || batch <= 0 || batch % 64 || !(lr > 0.0f) || n > (1L << 30)) {
        set_error("bad argument (batch must be a positive multiple of 64)");
        return UNO_EINVAL;
    }
    const int in = m->dims[0], out = m->dims.back();
    if (int rc = ensure_batch(m, batch)) return rc;
    DBuf<float> dX, dT, dM;
    UNO_HIP(dX.upload(X, (size_t)n * in, m->st));
    UNO_HIP(dT.upload(T, (size_t)n * out, m->st));
    UNO_HIP(dM.upload(M, (size_t)n * out, m->st));
    std::vector<float> cnt(n);
    for (long r = 0; r < n; ++r) {
        double s = 0.0;
        for (int c = 0; c < out; ++c) s += M[r * out + c];
        cnt[r] = (float)s;
    }
    std::mt19937_64 rng(seed);
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    long pos = n;
    std::vector<int> rows(batch);
    for (int it = 0; it < steps; ++it) {
        double norm = 0.0;
        for (int r = 0; r < batch; ++r) {
            if (pos >= n) {
                std::shuffle(perm.begin(), perm.end(), rng);
                pos = 0;
            }
            rows[r] = perm[pos++];
            norm += cnt[rows[r]];
        }
        if (int rc = stage_rows(m, rows, dX, &dT, &dM)) return rc;
        if (int rc = forward(m, batch)) return rc;
        int nblk = 0;
        if (int rc = backward(m, batch, norm > 0.0 ? (float)(1.0 / norm) : 0.0f, &nblk)) return rc;
        // cosine-decayed learning rate
        const float lr_t = lr * 0.5f * (1.0f + std::cos(3.14159265f * (float)it / (float)steps));
        ++m->step;
        UNO_HIP(adam_step(m->np, m->p.p, m->m.p, m->v.p, m->g.p, lr_t, 0.9f, 0.999f, m->step, 1e-8f, m->pb.p, m->st));
    }
    UNO_HIP(hipStreamSynchronize(m->st));
    if (loss) {
        // masked MSE over the whole training set with the final parameters
        double se = 0.0, ns = 0.0;
        const int chunk = batch;
        for (long r0 = 0; r0 < n; r0 += chunk) {
            const int c = (int)std::min<long>(chunk, n - r0);
            std::vector<int> rr(chunk, -1);
            for (int r = 0; r < c; ++r) rr[r] = (int)(r0 + r);
            if (int rc = stage_rows(m, rr, dX, &dT, &dM)) return rc;
            if (int rc = forward(m, chunk)) return rc;
            int nblk = 0;
            UNO_HIP(mse_loss((long)chunk * out, m->y.p, m->tb.p, m->mb.p, 1.0f, m->dz[0].p, m->dzb[0].p, m->part.p,
                             &nblk, m->st));
            std::vector<float> part(nblk);
            UNO_HIP(hipMemcpyAsync(part.data(), m->part.p, nblk * sizeof(float), hipMemcpyDeviceToHost, m->st));
            UNO_HIP(hipStreamSynchronize(m->st));
            for (float v : part) se += v;
            for (int r = 0; r < c; ++r) ns += cnt[r0 + r];
        }
        *loss = (float)(ns > 0.0 ? se / ns : 0.0);
    }
    return UNO_OK;
}

int uno_mlp_save(uno_mlp* m, const char* path) {
    if (!m || !path) { set_error("null argument"); return UNO_EINVAL; }
    std::vector<float> hp(m->np);
    if (int rc = uno_mlp_get_params(m, hp.data())) return rc;
    FILE* f = std::fopen(path, "wb");
    if (!f) { set_error(std::string("could not create ") + path); return UNO_EIO; }
    const uint32_t hdr[2] = {kMagic, (uint32_t)m->dims.size()};
    std::fwrite(hdr, sizeof hdr, 1, f);
    std::fwrite(m->dims.data(), sizeof(int), m->dims.size(), f);
    std::fwrite(hp.data(), sizeof(float), hp.size(), f);
    std::fclose(f);
    return UNO_OK;
}

int uno_mlp_load(const char* path, uno_mlp** out) {
    if (!path || !out) { set_error("null argument"); return UNO_EINVAL; }
    FILE* f = std::fopen(path, "rb");
    if (!f) { set_error(std::string("could not open ") + path); return UNO_EIO; }
    uint32_t hdr[2];
    std::vector<int> dims;
    bool ok = std::fread(hdr, sizeof hdr, 1, f) == 1 && hdr[0] == kMagic && hdr[1] >= 2 && hdr[1] <= 16;
    if (ok) {
        dims.resize(hdr[1]);
        ok = std::fread(dims.data(), sizeof(int), dims.size(), f) == dims.size();
    }
    if (!ok) { std::fclose(f); set_error(std::string("not an MLP file: ") + path); return UNO_EIO; }
    for (size_t i = 1; i + 1 < dims.size(); ++i)
        if (dims[i] != dims[1]) { std::fclose(f); set_error("MLP file has unequal hidden widths"); return UNO_EIO; }
    uno_mlp* m = nullptr;
    if (int rc = uno_mlp_create(dims.front(), dims.size() > 2 ? dims[1] : 64, (int)dims.size() - 2, dims.back(), 0,
                                &m)) {
        std::fclose(f);
        return rc;
    }
    std::vector<float> hp(m->np);
    ok = std::fread(hp.data(), sizeof(float), hp.size(), f) == hp.size();
    std::fclose(f);
    if (!ok) { delete m; set_error(std::string("truncated MLP file: ") + path); return UNO_EIO; }
    if (int rc = upload_params(m, hp.data())) { delete m; return rc; }
    *out = m;
    return UNO_OK;
}

}  // extern "C"
