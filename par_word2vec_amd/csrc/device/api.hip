// api.hip -- C-ABI of libpar_sgns.so (declared in include/par_sgns.h).
//
// A handle owns every device buffer, one HIP stream (default priority) and
// two timing events.  All uploads validate indices on the host before any
// kernel can see them, so a bad token id or offset becomes PAR_SGNS_EINVAL
// instead of an out-of-bounds access on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "../../../include/par_sgns.h"
#include "sgns_device.h"
#include "sgns_internal.h"

namespace {

struct GraphKey {
    int64_t b, e;
    int32_t epoch, loss, wpb, blocks, iter, grab;
    uint32_t alpha_bits;
    uint64_t seed, T;
    uint32_t hot_bytes;
    bool operator<(const GraphKey &o) const {
        return std::tie(b, e, epoch, loss, wpb, blocks, iter, grab, alpha_bits, seed, T, hot_bytes) <
               std::tie(o.b, o.e, o.epoch, o.loss, o.wpb, o.blocks, o.iter, o.grab, o.alpha_bits, o.seed, o.T,
                        o.hot_bytes);
    }
};

// Device scratch of the allocation-free window loss (par_sgns_loss_forward_noalloc):
// the FP64 loss accumulators in the first 256 bytes, then the staged window
// indices.  Each device holds its own copy until the process ends; host code
// reaches it only through the address hipGetSymbolAddress returns for the
// current device (the array's name is a host-side stand-in).
constexpr size_t kScratchBytes = size_t(8) << 20;
constexpr size_t kScratchIdx = 256;  // byte offset of the staged indices

std::mutex g_scratch_mu;
std::vector<char *> g_scratch_ptr;  // per device ordinal, looked up once

}  // namespace

static __device__ __align__(256) char scratch[kScratchBytes];

struct par_sgns_handle_s {
    par_sgns_config cfg{};
    int ld = 0, nq = 0, np = 0, bf16 = 0, kmax = 0, R = 0;
    int cu_count = 0;
    // launch shape chosen by par_sgns_autotune (0: not tuned)
    int tuned_wpb = 0, tuned_wpcu = 0, tuned_grab = 0;
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    void *w_in = nullptr, *w_out = nullptr;
    float *exp_table = nullptr;
    uint2 *alias_tab = nullptr;  // per bucket: {keep probability (float bits), alias}
    bool have_exp = false, have_sampler = false, have_weights = false;
    // fast-mode corpus
    int32_t *tokens = nullptr;
    int64_t *sent_off = nullptr, *sent_raw = nullptr;
    int32_t *order = nullptr;
    int64_t *queue_raw = nullptr;
    std::vector<int64_t> h_sent_raw;
    int64_t n_tokens = 0, n_sent = 0;
    std::vector<int64_t> h_sent_off;
    std::set<std::pair<int64_t, int64_t>> ordered;
    // deterministic plan
    int32_t *p_tokens = nullptr, *p_negs = nullptr;
    uint8_t *p_b = nullptr;
    int64_t *p_sent_off = nullptr, *p_batch_off = nullptr, *p_neg_off = nullptr;
    float *p_alpha = nullptr;
    float *det_scratch = nullptr;
    int64_t p_n_sent = 0;
    // counters
    unsigned long long *counters = nullptr, *queue = nullptr;
    double *loss = nullptr;
    // window kernels
    float *g_in = nullptr, *g_out = nullptr;
    char *scratch = nullptr;  // this device's static scratch (device address)
    int32_t *win_buf = nullptr;
    size_t win_cap = 0;
    std::map<GraphKey, hipGraphExec_t> graphs;
    double last_ms[3] = {0, 0, 0};
    int last_launches[3] = {0, 0, 0};
    std::string err;
};

namespace {

int seterr(par_sgns_handle h, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (h) h->err = buf;
    return code;
}

#define HIPCHK(h, call)                                                                          \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return seterr(h, e_ == hipErrorOutOfMemory ? PAR_SGNS_ENOMEM : PAR_SGNS_EDEVICE,     \
                          "%s failed: %s", #call, hipGetErrorString(e_));                        \
    } while (0)

template <typename P>
void dfree(P *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

template <typename P>
int dalloc(par_sgns_handle h, P *&p, size_t n) {
    dfree(p);
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)&p, n * sizeof(P));
    if (e != hipSuccess) {
        p = nullptr;
        return seterr(h, PAR_SGNS_ENOMEM, "hipMalloc of %zu bytes failed: %s", n * sizeof(P), hipGetErrorString(e));
    }
    return 0;
}

template <typename P>
int upload(par_sgns_handle h, P *&p, const P *src, size_t n) {
    int rc = dalloc(h, p, n);
    if (rc) return rc;
    if (n) HIPCHK(h, hipMemcpyAsync(p, src, n * sizeof(P), hipMemcpyHostToDevice, h->st));
    return 0;
}

uint16_t f2bf_rne(float x) {
    uint32_t b;
    std::memcpy(&b, &x, 4);
    return (uint16_t)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}
float bf2f(uint16_t v) {
    uint32_t b = (uint32_t)v << 16;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

// Captured graphs hold every device pointer of their launch by value, so a
// call that frees or rewrites a buffer they read drops them first.
void drop_graphs(par_sgns_handle h) {
    for (auto &kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
    h->graphs.clear();
}

// Device address of `scratch` on the current device `dev` (looked up once per
// device, with that device current).
int scratch_address(par_sgns_handle h, int dev, char **out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((size_t)dev >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)dev + 1, nullptr);
    if (!g_scratch_ptr[(size_t)dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return seterr(h, PAR_SGNS_EDEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                          hipGetErrorString(e));
        g_scratch_ptr[(size_t)dev] = static_cast<char *>(p);
    }
    *out = g_scratch_ptr[(size_t)dev];
    return 0;
}

int check_ids(par_sgns_handle h, const int32_t *ids, int64_t n, int32_t lo, int32_t hi, const char *what) {
    int64_t bad = -1;
#pragma omp parallel for reduction(max : bad)
    for (int64_t i = 0; i < n; i++)
        if (ids[i] < lo || ids[i] >= hi) bad = std::max(bad, i);
    if (bad >= 0) return seterr(h, PAR_SGNS_EINVAL, "%s[%lld] = %d outside [%d, %d)", what, (long long)bad, ids[bad], lo, hi);
    return 0;
}

int check_offsets(par_sgns_handle h, const int64_t *off, int64_t n_sent, int64_t n_tokens, int max_len, const char *what) {
    if (off[0] != 0) return seterr(h, PAR_SGNS_EINVAL, "%s[0] must be 0", what);
    for (int64_t s = 0; s < n_sent; s++) {
        int64_t L = off[s + 1] - off[s];
        if (L < 0 || L > max_len)
            return seterr(h, PAR_SGNS_EINVAL, "%s: sentence %lld has length %lld (allowed 0..%d)", what, (long long)s,
                          (long long)L, max_len);
    }
    if (off[n_sent] != n_tokens) return seterr(h, PAR_SGNS_EINVAL, "%s[n_sent] != n_tokens", what);
    return 0;
}

}  // namespace

extern "C" {

int par_sgns_create(const par_sgns_config *cfg, par_sgns_handle *out) {
    if (!cfg || !out) return PAR_SGNS_EINVAL;
    *out = nullptr;
    if (cfg->vocab_size < 2 || cfg->dim < 1 || cfg->dim > 512 || cfg->window < 1 || cfg->window > 31 ||
        cfg->negative < 0 || cfg->negative > 64 || cfg->batch_size < 1 || cfg->batch_size > 64 ||
        (cfg->precision != PAR_SGNS_FP32 && cfg->precision != PAR_SGNS_BF16) ||
        cfg->rounding < PAR_SGNS_ROUND_STOCHASTIC || cfg->rounding > PAR_SGNS_ROUND_NEAREST_IN)
        return PAR_SGNS_EINVAL;
    par_sgns_handle h = new par_sgns_handle_s;
    h->cfg = *cfg;
    h->bf16 = cfg->precision == PAR_SGNS_BF16;
    h->ld = h->bf16 ? (cfg->dim + 63) / 64 * 64 : (cfg->dim + 31) / 32 * 32;
    h->nq = (cfg->dim + 1) / 2;
    h->np = (h->nq + 63) / 64;
    h->R = 2 * cfg->window + 1;
    h->kmax = (2 * cfg->window - 1) * cfg->negative;
    if (h->kmax < 1) h->kmax = 1;
    *out = h;
    hipError_t e = hipSetDevice(cfg->device);
    if (e != hipSuccess) return seterr(h, PAR_SGNS_EDEVICE, "hipSetDevice(%d): %s", cfg->device, hipGetErrorString(e));
    HIPCHK(h, hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking));
    HIPCHK(h, hipEventCreate(&h->ev0));
    HIPCHK(h, hipEventCreate(&h->ev1));
    HIPCHK(h, hipDeviceGetAttribute(&h->cu_count, hipDeviceAttributeMultiprocessorCount, cfg->device));
    int rc0 = scratch_address(h, cfg->device, &h->scratch);
    if (rc0) return rc0;
    const size_t rows = (size_t)cfg->vocab_size * (size_t)h->ld;
    const size_t esz = h->bf16 ? 2 : 4;
    int rc;
    if ((rc = dalloc(h, reinterpret_cast<char *&>(h->w_in), rows * esz))) return rc;
    if ((rc = dalloc(h, reinterpret_cast<char *&>(h->w_out), rows * esz))) return rc;
    HIPCHK(h, hipMemsetAsync(h->w_in, 0, rows * esz, h->st));
    HIPCHK(h, hipMemsetAsync(h->w_out, 0, rows * esz, h->st));
    if ((rc = dalloc(h, h->counters, sgns::kCntSlots))) return rc;
    if ((rc = dalloc(h, h->queue, 1))) return rc;
    if ((rc = dalloc(h, h->loss, 2))) return rc;
    HIPCHK(h, hipStreamSynchronize(h->st));
    return 0;
}

int par_sgns_destroy(par_sgns_handle h) {
    if (!h) return PAR_SGNS_EINVAL;
    (void)hipSetDevice(h->cfg.device);
    if (h->st) (void)hipStreamSynchronize(h->st);
    for (auto &kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
    dfree(h->w_in); dfree(h->w_out); dfree(h->exp_table); dfree(h->alias_tab);
    dfree(h->tokens); dfree(h->sent_off); dfree(h->sent_raw); dfree(h->order); dfree(h->queue_raw);
    dfree(h->p_tokens); dfree(h->p_negs); dfree(h->p_b); dfree(h->p_sent_off); dfree(h->p_batch_off);
    dfree(h->p_neg_off); dfree(h->p_alpha); dfree(h->det_scratch);
    dfree(h->counters); dfree(h->queue); dfree(h->loss); dfree(h->g_in); dfree(h->g_out); dfree(h->win_buf);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->st) (void)hipStreamDestroy(h->st);
    delete h;
    return 0;
}

const char *par_sgns_last_error(par_sgns_handle h) { return h ? h->err.c_str() : "null handle"; }

int32_t par_sgns_row_stride(par_sgns_handle h) { return h ? h->ld : PAR_SGNS_EINVAL; }

int par_sgns_set_exp_table(par_sgns_handle h, const float *t, int32_t n) {
    if (!h) return PAR_SGNS_EINVAL;
    if (!t || n != sgns::kExpTable + 1) return seterr(h, PAR_SGNS_EINVAL, "exp table must have 1001 entries");
    drop_graphs(h);
    int rc = upload(h, h->exp_table, t, (size_t)n);
    if (rc) return rc;
    HIPCHK(h, hipStreamSynchronize(h->st));
    h->have_exp = true;
    return 0;
}

int par_sgns_set_sampler(par_sgns_handle h, const int64_t *entries, int32_t V) {
    if (!h) return PAR_SGNS_EINVAL;
    if (!entries || V != h->cfg.vocab_size) return seterr(h, PAR_SGNS_EINVAL, "sampler needs vocab_size entries");
    // distribution of one negative draw before in-sentence rejection:
    // slot word i (i >= 1), or a "</s>" slot replaced by uniform [1, V-1]
    double total = 0;
    for (int32_t i = 0; i < V; i++) {
        if (entries[i] < 0) return seterr(h, PAR_SGNS_EINVAL, "negative table entry count");
        total += (double)entries[i];
    }
    if (total <= 0) return seterr(h, PAR_SGNS_EINVAL, "empty unigram table");
    std::vector<double> w((size_t)V);
    const double uni = (double)entries[0] / total / (double)(V - 1);
    w[0] = 0.0;
    for (int32_t i = 1; i < V; i++) w[(size_t)i] = (double)entries[i] / total + uni;
    // Vose alias construction
    std::vector<double> q((size_t)V);
    std::vector<float> prob((size_t)V, 1.0f);
    std::vector<int32_t> alias((size_t)V);
    std::vector<int32_t> small, large;
    for (int32_t i = 0; i < V; i++) {
        q[(size_t)i] = w[(size_t)i] * V;
        alias[(size_t)i] = i;
        (q[(size_t)i] < 1.0 ? small : large).push_back(i);
    }
    while (!small.empty() && !large.empty()) {
        int32_t s = small.back(), l = large.back();
        small.pop_back();
        prob[(size_t)s] = (float)q[(size_t)s];
        alias[(size_t)s] = l;
        q[(size_t)l] = (q[(size_t)l] + q[(size_t)s]) - 1.0;
        if (q[(size_t)l] < 1.0) {
            large.pop_back();
            small.push_back(l);
        }
    }
    for (int32_t i : large) { prob[(size_t)i] = 1.0f; alias[(size_t)i] = i; }
    for (int32_t i : small) { prob[(size_t)i] = 1.0f; alias[(size_t)i] = i; }
    // a word with zero mass must never be returned by its own bucket
    if (prob[0] > 0.0f && w[0] == 0.0) prob[0] = 0.0f;
    if (alias[0] == 0) alias[0] = 1;
    // one 8-byte entry per bucket: a draw is one load
    std::vector<uint2> tab((size_t)V);
    for (int32_t i = 0; i < V; i++) {
        uint32_t pb;
        std::memcpy(&pb, &prob[(size_t)i], 4);
        tab[(size_t)i] = make_uint2(pb, (uint32_t)alias[(size_t)i]);
    }
    drop_graphs(h);
    int rc = upload(h, h->alias_tab, tab.data(), (size_t)V);
    if (rc) return rc;
    HIPCHK(h, hipStreamSynchronize(h->st));
    h->have_sampler = true;
    return 0;
}

int par_sgns_sample_negatives(par_sgns_handle h, uint64_t seed, int64_t sent, int64_t n, int32_t *out) {
    if (!h) return PAR_SGNS_EINVAL;
    if (!out || n < 0 || sent < 0) return seterr(h, PAR_SGNS_EINVAL, "bad sample arguments");
    if (!h->have_sampler) return seterr(h, PAR_SGNS_ESTATE, "sampler not set");
    HIPCHK(h, hipSetDevice(h->cfg.device));
    int32_t *d = nullptr;
    int rc = dalloc(h, d, (size_t)n);
    if (rc) return rc;
    hipError_t e = sgns::launch_sample(h->alias_tab, h->cfg.vocab_size, seed, sent, n, d, h->st);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, (size_t)n * 4, hipMemcpyDeviceToHost, h->st);
    if (e == hipSuccess) e = hipStreamSynchronize(h->st);
    dfree(d);
    if (e != hipSuccess) return seterr(h, PAR_SGNS_EDEVICE, "sampling failed: %s", hipGetErrorString(e));
    return 0;
}

int par_sgns_set_weights(par_sgns_handle h, const float *w_in, const float *w_out) {
    if (!h) return PAR_SGNS_EINVAL;
    const int64_t V = h->cfg.vocab_size, d = h->cfg.dim, ld = h->ld;
    const size_t esz = h->bf16 ? 2 : 4;
    std::vector<char> buf((size_t)V * ld * esz);
    const float *srcs[2] = {w_in, w_out};
    void *dsts[2] = {h->w_in, h->w_out};
    for (int k = 0; k < 2; k++) {
        const float *src = srcs[k];
        if (!src) continue;
        std::memset(buf.data(), 0, buf.size());
        if (h->bf16) {
            uint16_t *o = reinterpret_cast<uint16_t *>(buf.data());
#pragma omp parallel for
            for (int64_t r = 0; r < V; r++)
                for (int64_t j = 0; j < d; j++) o[r * ld + j] = f2bf_rne(src[r * d + j]);
        } else {
            float *o = reinterpret_cast<float *>(buf.data());
#pragma omp parallel for
            for (int64_t r = 0; r < V; r++) std::memcpy(o + r * ld, src + r * d, (size_t)d * 4);
        }
        HIPCHK(h, hipMemcpy(dsts[k], buf.data(), buf.size(), hipMemcpyHostToDevice));
    }
    h->have_weights = true;
    return 0;
}

int par_sgns_get_weights(par_sgns_handle h, float *w_in, float *w_out) {
    if (!h) return PAR_SGNS_EINVAL;
    const int64_t V = h->cfg.vocab_size, d = h->cfg.dim, ld = h->ld;
    const size_t esz = h->bf16 ? 2 : 4;
    HIPCHK(h, hipStreamSynchronize(h->st));
    std::vector<char> buf((size_t)V * ld * esz);
    float *dsts[2] = {w_in, w_out};
    void *srcs[2] = {h->w_in, h->w_out};
    for (int k = 0; k < 2; k++) {
        float *dst = dsts[k];
        if (!dst) continue;
        HIPCHK(h, hipMemcpy(buf.data(), srcs[k], buf.size(), hipMemcpyDeviceToHost));
        if (h->bf16) {
            const uint16_t *s = reinterpret_cast<const uint16_t *>(buf.data());
#pragma omp parallel for
            for (int64_t r = 0; r < V; r++)
                for (int64_t j = 0; j < d; j++) dst[r * d + j] = bf2f(s[r * ld + j]);
        } else {
            const float *s = reinterpret_cast<const float *>(buf.data());
#pragma omp parallel for
            for (int64_t r = 0; r < V; r++) std::memcpy(dst + r * d, s + r * ld, (size_t)d * 4);
        }
    }
    return 0;
}

int par_sgns_set_corpus(par_sgns_handle h, const int32_t *tokens, int64_t n_tokens, const int64_t *sent_off,
                        const int64_t *sent_raw, int64_t n_sent) {
    if (!h) return PAR_SGNS_EINVAL;
    if (!tokens || !sent_off || !sent_raw || n_tokens < 0 || n_sent < 0)
        return seterr(h, PAR_SGNS_EINVAL, "null corpus array");
    int rc = check_offsets(h, sent_off, n_sent, n_tokens, sgns::kMaxSentence, "sent_off");
    if (rc) return rc;
    if ((rc = check_ids(h, tokens, n_tokens, 0, h->cfg.vocab_size, "tokens"))) return rc;
    for (int64_t s = 0; s < n_sent; s++)
        if (sent_raw[s] < 0) return seterr(h, PAR_SGNS_EINVAL, "sent_raw[%lld] < 0", (long long)s);
    if ((rc = upload(h, h->tokens, tokens, (size_t)n_tokens))) return rc;
    if ((rc = upload(h, h->sent_off, sent_off, (size_t)n_sent + 1))) return rc;
    if ((rc = upload(h, h->sent_raw, sent_raw, (size_t)n_sent))) return rc;
    std::vector<int32_t> ident((size_t)std::max<int64_t>(n_sent, 1));
    std::iota(ident.begin(), ident.end(), 0);
    if ((rc = upload(h, h->order, ident.data(), ident.size()))) return rc;
    if ((rc = upload(h, h->queue_raw, sent_raw, (size_t)std::max<int64_t>(n_sent, 1)))) return rc;
    HIPCHK(h, hipStreamSynchronize(h->st));
    h->h_sent_off.assign(sent_off, sent_off + n_sent + 1);
    h->h_sent_raw.assign(sent_raw, sent_raw + n_sent);
    h->n_tokens = n_tokens;
    h->n_sent = n_sent;
    h->ordered.clear();
    drop_graphs(h);
    return 0;
}

int par_sgns_set_plan(par_sgns_handle h, const int32_t *tokens, const uint8_t *b, int64_t n_tokens,
                      const int64_t *sent_off, const float *sent_alpha, int64_t n_sent, const int64_t *batch_off,
                      const int64_t *neg_off, int64_t n_batches, const int32_t *negs, int64_t n_negs) {
    if (!h) return PAR_SGNS_EINVAL;
    if (!tokens || !b || !sent_off || !sent_alpha || !batch_off || !neg_off || (!negs && n_negs > 0))
        return seterr(h, PAR_SGNS_EINVAL, "null plan array");
    const int W = h->cfg.window, B = h->cfg.batch_size, NEG = h->cfg.negative;
    int rc = check_offsets(h, sent_off, n_sent, n_tokens, sgns::kMaxSentence, "plan sent_off");
    if (rc) return rc;
    if ((rc = check_ids(h, tokens, n_tokens, 0, h->cfg.vocab_size, "plan tokens"))) return rc;
    if ((rc = check_ids(h, negs, n_negs, 0, h->cfg.vocab_size, "plan negs"))) return rc;
    for (int64_t i = 0; i < n_tokens; i++)
        if (b[i] >= W) return seterr(h, PAR_SGNS_EINVAL, "plan b[%lld] >= window", (long long)i);
    if (batch_off[0] != 0 || neg_off[0] != 0) return seterr(h, PAR_SGNS_EINVAL, "plan offsets must start at 0");
    for (int64_t s = 0; s < n_sent; s++) {
        int64_t L = sent_off[s + 1] - sent_off[s];
        if (batch_off[s + 1] - batch_off[s] != (L + B - 1) / B)
            return seterr(h, PAR_SGNS_EINVAL, "plan batch_off inconsistent at sentence %lld", (long long)s);
    }
    if (batch_off[n_sent] != n_batches) return seterr(h, PAR_SGNS_EINVAL, "plan batch count mismatch");
    for (int64_t k = 0; k < n_batches; k++) {
        int64_t K = neg_off[k + 1] - neg_off[k];
        if (K < 0 || K > (int64_t)(2 * W - 1) * NEG || K % (NEG > 0 ? NEG : 1) != 0)
            return seterr(h, PAR_SGNS_EINVAL, "plan batch %lld has %lld negatives", (long long)k, (long long)K);
    }
    if (neg_off[n_batches] != n_negs) return seterr(h, PAR_SGNS_EINVAL, "plan negative count mismatch");
    drop_graphs(h);
    if ((rc = upload(h, h->p_tokens, tokens, (size_t)n_tokens))) return rc;
    if ((rc = upload(h, h->p_b, b, (size_t)n_tokens))) return rc;
    if ((rc = upload(h, h->p_sent_off, sent_off, (size_t)n_sent + 1))) return rc;
    if ((rc = upload(h, h->p_alpha, sent_alpha, (size_t)n_sent * 2))) return rc;
    if ((rc = upload(h, h->p_batch_off, batch_off, (size_t)n_sent + 1))) return rc;
    if ((rc = upload(h, h->p_neg_off, neg_off, (size_t)n_batches + 1))) return rc;
    if ((rc = upload(h, h->p_negs, negs, (size_t)n_negs))) return rc;
    const size_t ld = (size_t)h->ld, kx = (size_t)h->kmax;
    if ((rc = dalloc(h, h->det_scratch, (size_t)B * ld * 2 + kx * ld * 2 + kx * (size_t)B))) return rc;
    HIPCHK(h, hipStreamSynchronize(h->st));
    h->p_n_sent = n_sent;
    return 0;
}

// The queue order of a sentence range [b, e) and the progress at each of its
// slots live in order[b..e) and queue_raw[b..e), shared by every range.
// Ordering a range therefore invalidates each prepared range that overlaps
// it, and the graphs captured for those ranges.
static int ensure_order(par_sgns_handle h, int64_t b, int64_t e) {
    if (e <= b || h->ordered.count({b, e})) return 0;
    for (auto it = h->ordered.begin(); it != h->ordered.end();) {
        if (it->first < e && b < it->second) it = h->ordered.erase(it);
        else ++it;
    }
    for (auto it = h->graphs.begin(); it != h->graphs.end();) {
        if (it->first.b < e && b < it->first.e) {
            (void)hipGraphExecDestroy(it->second);
            it = h->graphs.erase(it);
        } else {
            ++it;
        }
    }
    // longest sentences first within the range (balances uneven lengths)
    std::vector<int32_t> ord((size_t)(e - b));
    std::iota(ord.begin(), ord.end(), (int32_t)b);
    const int64_t *off = h->h_sent_off.data();
    std::stable_sort(ord.begin(), ord.end(), [&](int32_t x, int32_t y) {
        return off[x + 1] - off[x] > off[y + 1] - off[y];
    });
    // training progress at each queue slot: raw words of the sentences dequeued
    // before it (a sentence's raw span runs to the next sentence's start)
    const int64_t *raw = h->h_sent_raw.data();
    std::vector<int64_t> prog(ord.size());
    int64_t acc = raw[b];
    for (size_t k = 0; k < ord.size(); k++) {
        prog[k] = acc;
        const int32_t x = ord[k];
        acc += x + 1 < h->n_sent ? raw[x + 1] - raw[x] : (off[x + 1] - off[x]);
    }
    HIPCHK(h, hipMemcpyAsync(h->order + b, ord.data(), ord.size() * sizeof(int32_t), hipMemcpyHostToDevice, h->st));
    HIPCHK(h, hipMemcpyAsync(h->queue_raw + b, prog.data(), prog.size() * sizeof(int64_t), hipMemcpyHostToDevice, h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    h->ordered.insert({b, e});
    return 0;
}

// Default persistent grid: as many resident waves as the LDS allows, but at
// least kMinSentencesPerWave sentences per wave.  Hogwild quality depends on
// each row seeing a sequence of updates; a small corpus spread over every
// resident wave of the chip turns a pass into a few huge stale steps.  How
// stale a step is depends on the chip's update rate as well as on the number
// of waves: on an epoch of 17,000 sentences of 1,000 words (bf16, 570 kept
// tokens per sentence) 16 sentences per wave kept planted-cluster recall@10
// at 0.99 while the kernel ran at 300 Mwords/s, and fell to 0.96 once it ran
// at 500; 24 keeps 0.998 at 355 Mwords/s.
constexpr int64_t kMinSentencesPerWave = 24;

// Resident blocks per CU for a block of wpb waves: 160 KB of LDS per CU (the
// kernel has no per-block LDS of its own); at most kMaxWavesCu waves per CU
// (two per SIMD at the kernel's register budget).
constexpr int kLdsCu = 160 * 1024, kMaxWavesCu = 8, kMaxWpb = 8;
// log2 slots of a long sentence's token hash set: 256 (sentences up to 192
// tokens) to 2048 (any sentence)
constexpr int kTabLog2Min = 8, kTabLog2Max = 11;
static int blocks_per_cu(int lds_per_wave, int wpb) {
    int b = kLdsCu / (lds_per_wave * wpb);
    if (b * wpb > kMaxWavesCu) b = kMaxWavesCu / wpb;
    return b;
}

// Whether the matrix-core repulsion can run: it addresses W_out with 32-bit
// byte offsets and holds a batch's positions in one group of 32.
static bool mfma_fits(par_sgns_handle h) {
    return (uint64_t)h->cfg.vocab_size * (uint64_t)h->ld * 2u < (1ull << 32) - 4096 &&
           h->cfg.batch_size <= 16 * sgns::kMfI;
}
static bool use_mfma(par_sgns_handle h, int32_t repulsion) {
    return h->bf16 && mfma_fits(h) && repulsion != PAR_SGNS_REP_VECTOR;
}

// floor: sentences per wave (the caller's, or kMinSentencesPerWave)
static void default_shape(par_sgns_handle h, int lds_per_wave, int64_t n_sent, int64_t floor, int &wpb,
                          int &blocks) {
    if (wpb <= 0 && h->tuned_wpb > 0) wpb = h->tuned_wpb;
    if (blocks <= 0 && h->tuned_wpcu > 0) {  // the tuned resident waves per CU, where they fit
        const int w = wpb > 0 ? wpb : 1;
        int bpc = h->tuned_wpcu / w;
        const int fit = blocks_per_cu(lds_per_wave, w);
        if (bpc > fit) bpc = fit;
        if (bpc >= 1) {
            blocks = bpc * (h->cu_count > 0 ? h->cu_count : 256);
            const int64_t cap = n_sent / floor / w;
            if (blocks > cap) blocks = (int)(cap < 1 ? 1 : cap);
        }
    }
    if (wpb <= 0) {  // the block size that keeps the most waves resident (smallest on ties)
        int best = 0;
        wpb = 1;
        for (int k = 1; k <= kMaxWpb; k++) {
            const int waves = blocks_per_cu(lds_per_wave, k) * k;
            if (waves > best) {
                best = waves;
                wpb = k;
            }
        }
    }
    if (blocks <= 0) {
        int bpc = blocks_per_cu(lds_per_wave, wpb);
        if (bpc < 1) bpc = 1;
        blocks = bpc * (h->cu_count > 0 ? h->cu_count : 256);
        const int64_t cap = n_sent / floor / wpb;
        if (blocks > cap) blocks = (int)(cap < 1 ? 1 : cap);
    }
}

static int train_impl(par_sgns_handle h, const par_sgns_train_args *a, par_sgns_stats *stats, bool prepare_only) {
    if (!h) return PAR_SGNS_EINVAL;
    if (!a) return seterr(h, PAR_SGNS_EINVAL, "null args");
    if (!h->have_exp || !h->have_weights) return seterr(h, PAR_SGNS_ESTATE, "exp table and weights must be set first");
    if (a->iter < 1 || a->epoch < 0 || a->sent_begin < 0 || a->sent_end < a->sent_begin)
        return seterr(h, PAR_SGNS_EINVAL, "bad schedule arguments");
    HIPCHK(h, hipSetDevice(h->cfg.device));
    int rc;
    float ms = 0.f;
    int launches = 1;
    int shape_blocks = 1, shape_wpb = 1, shape_grab = 0;  // launch shape, reported in the stats
    if (a->mode == PAR_SGNS_DETERMINISTIC) {
        if (h->bf16) return seterr(h, PAR_SGNS_EINVAL, "deterministic mode runs on fp32 tables");
        if (!h->p_tokens) return seterr(h, PAR_SGNS_ESTATE, "no plan uploaded");
        if (a->sent_end > h->p_n_sent) return seterr(h, PAR_SGNS_EINVAL, "sentence range beyond the plan");
        sgns::DetArgs d{};
        d.tokens = h->p_tokens; d.b = h->p_b; d.sent_off = h->p_sent_off; d.sent_alpha = h->p_alpha;
        d.batch_off = h->p_batch_off; d.neg_off = h->p_neg_off; d.negs = h->p_negs; d.exp_table = h->exp_table;
        d.w_in = (float *)h->w_in; d.w_out = (float *)h->w_out;
        const size_t ld = (size_t)h->ld, B = (size_t)h->cfg.batch_size, kx = (size_t)h->kmax;
        d.m_in = h->det_scratch;
        d.m_in_new = d.m_in + B * ld;
        d.m_out = d.m_in_new + B * ld;
        d.m_out_upd = d.m_out + kx * ld;
        d.m_grad = d.m_out_upd + kx * ld;
        d.ld = h->ld; d.d = h->cfg.dim; d.W = h->cfg.window; d.B = h->cfg.batch_size; d.kmax = h->kmax;
        d.compute_loss = a->compute_loss; d.s_begin = a->sent_begin; d.s_end = a->sent_end;
        d.loss = h->loss; d.counters = h->counters;
        if (prepare_only) return 0;
        HIPCHK(h, sgns::launch_reset(h->counters, sgns::kCntSlots, h->loss, nullptr, h->st));
        HIPCHK(h, hipEventRecord(h->ev0, h->st));
        HIPCHK(h, sgns::launch_det(d, h->st));
        HIPCHK(h, hipEventRecord(h->ev1, h->st));
    } else if (a->mode == PAR_SGNS_PLAIN || a->mode == PAR_SGNS_PERSISTENT) {
        const bool planned = a->schedule == 1;
        if (planned) {
            if (!h->p_tokens) return seterr(h, PAR_SGNS_ESTATE, "no plan uploaded");
            if (a->sent_end > h->p_n_sent) return seterr(h, PAR_SGNS_EINVAL, "sentence range beyond the plan");
        } else {
            if (!h->have_sampler) return seterr(h, PAR_SGNS_ESTATE, "sampler not set");
            if (!h->tokens) return seterr(h, PAR_SGNS_ESTATE, "no corpus uploaded");
            if (a->sent_end > h->n_sent) return seterr(h, PAR_SGNS_EINVAL, "sentence range beyond the corpus");
        }
        const bool pers = a->mode == PAR_SGNS_PERSISTENT;
        sgns::FastArgs f{};
        f.tokens = h->tokens; f.sent_off = h->sent_off; f.sent_raw = h->sent_raw; f.order = pers ? h->order : nullptr;
        f.queue_raw = h->queue_raw;
        if (planned) {
            // exact schedule: plan tokens in plan order (sentence order is kept so a
            // single wave replays the reference's sequence)
            f.tokens = h->p_tokens; f.sent_off = h->p_sent_off; f.sent_raw = nullptr; f.order = nullptr;
            f.plan_b = h->p_b; f.plan_alpha = h->p_alpha; f.plan_batch_off = h->p_batch_off;
            f.plan_neg_off = h->p_neg_off; f.plan_negs = h->p_negs;
        }
        f.alias = h->alias_tab;
        f.w_in = h->w_in; f.w_out = h->w_out;
        f.ld = h->ld; f.d = h->cfg.dim; f.nq = h->nq; f.W = h->cfg.window; f.R = h->R; f.NEG = h->cfg.negative;
        f.B = h->cfg.batch_size; f.V = h->cfg.vocab_size; f.kmax = h->kmax;
        if (a->repulsion < PAR_SGNS_REP_AUTO || a->repulsion > PAR_SGNS_REP_MATRIX)
            return seterr(h, PAR_SGNS_EINVAL, "repulsion must be AUTO, VECTOR or MATRIX");
        if (a->repulsion == PAR_SGNS_REP_MATRIX && (!h->bf16 || !mfma_fits(h)))
            return seterr(h, PAR_SGNS_EINVAL, "matrix-core repulsion runs on bf16 tables below 4 GiB with batch size <= 32");
        f.rep_mfma = use_mfma(h, a->repulsion);
        if (a->write_back < PAR_SGNS_WB_ATOMIC || a->write_back > PAR_SGNS_WB_HOT)
            return seterr(h, PAR_SGNS_EINVAL, "write_back must be ATOMIC, STORE_OUT, STORE_IN, STORE or HOT");
        if (a->write_back != PAR_SGNS_WB_ATOMIC && a->write_back != PAR_SGNS_WB_STORE_OUT && !f.rep_mfma)
            return seterr(h, PAR_SGNS_EINVAL,
                          "write-back STORE_IN, STORE and HOT apply to the matrix-core repulsion on bf16 tables");
        if (a->write_back == PAR_SGNS_WB_HOT && (a->hot_rows < 1 || a->hot_rows > h->cfg.vocab_size))
            return seterr(h, PAR_SGNS_EINVAL, "write_back HOT needs 1 <= hot_rows <= vocab_size");
        f.write_back = a->write_back;
        f.hot_bytes = a->write_back == PAR_SGNS_WB_HOT ? (uint32_t)a->hot_rows * (uint32_t)h->ld * 2u : 0u;
        f.lds_per_wave = sgns::fast_lds_per_wave(h->R, h->nq, h->bf16, h->kmax, h->cfg.batch_size, h->ld, f.rep_mfma);
        f.lds_negs = sgns::fast_lds_negs(h->nq, h->bf16, h->cfg.batch_size, h->ld, f.rep_mfma);
        f.persistent = pers ? 1 : 0; f.rounding = h->cfg.rounding; f.compute_loss = a->compute_loss; f.iter = a->iter; f.epoch = a->epoch;
        f.a0 = a->alpha; f.T = a->train_words; f.seed = a->seed; f.s_begin = a->sent_begin; f.s_end = a->sent_end;
        f.queue = h->queue; f.loss = h->loss; f.counters = h->counters;
        int wpb = pers ? a->waves_per_block : 1, blocks = a->blocks;
        if (a->min_sentences_per_wave < 0) return seterr(h, PAR_SGNS_EINVAL, "min_sentences_per_wave must be >= 0");
        const int64_t floor = a->min_sentences_per_wave > 0 ? a->min_sentences_per_wave : kMinSentencesPerWave;
        default_shape(h, f.lds_per_wave, a->sent_end - a->sent_begin, floor, wpb, blocks);
        if (!pers) wpb = a->waves_per_block > 0 ? a->waves_per_block : 1;
        if (wpb > kMaxWpb) return seterr(h, PAR_SGNS_EINVAL, "at most %d waves per block", kMaxWpb);
        f.grab = a->sentences_per_dequeue > 0 ? a->sentences_per_dequeue : (h->tuned_grab > 0 ? h->tuned_grab : 1);
        if (f.grab > 64) return seterr(h, PAR_SGNS_EINVAL, "sentences_per_dequeue must be 0..64");
        if ((size_t)f.lds_per_wave * wpb > 160 * 1024)
            return seterr(h, PAR_SGNS_EINVAL, "window ring does not fit in LDS (%d bytes per wave)", f.lds_per_wave);
        {   // the token hash set of long sentences (rejection check), after the
            // batch's negatives: in the share's free bytes, or in a larger share
            // when the persistent grid fits the device at once with it; 2^11
            // slots hold any sentence (at most 1000 tokens) at a load below 1/2
            f.tab_off = (f.lds_negs + 12 * h->kmax + 15) & ~15;
            int64_t cap = f.lds_per_wave;
            if (pers) {
                const int cus = h->cu_count > 0 ? h->cu_count : 256;
                const int64_t per_cu = ((int64_t)blocks + cus - 1) / cus * wpb;  // resident waves per CU
                const int64_t fit = (kLdsCu / per_cu) & ~15;
                if (fit > cap) cap = std::min<int64_t>(fit, f.tab_off + (4 << kTabLog2Max));
            }
            f.tab_log2 = 0;
            for (int k = kTabLog2Max; k >= kTabLog2Min && !f.tab_log2; k--)
                if (f.tab_off + (4 << k) <= cap) f.tab_log2 = k;
            if (f.tab_log2 && f.tab_off + (4 << f.tab_log2) > f.lds_per_wave) f.lds_per_wave = f.tab_off + (4 << f.tab_log2);
        }
        shape_wpb = wpb;
        shape_blocks = pers ? blocks : (int)((a->sent_end - a->sent_begin + wpb - 1) / wpb);
        shape_grab = pers ? f.grab : 0;
        if (pers && !planned && (rc = ensure_order(h, a->sent_begin, a->sent_end))) return rc;
        if (pers && a->use_graph) {
            GraphKey k{a->sent_begin, a->sent_end, a->epoch, a->compute_loss | (planned ? 2 : 0) | (f.rep_mfma ? 4 : 0), wpb, blocks, a->iter,
                       f.grab, 0, a->seed, a->train_words, f.hot_bytes | (uint32_t)f.write_back};
            std::memcpy(&k.alpha_bits, &a->alpha, 4);
            auto it = h->graphs.find(k);
            if (it == h->graphs.end()) {
                hipGraph_t g;
                HIPCHK(h, hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
                hipError_t e1 = sgns::launch_reset(h->counters, sgns::kCntSlots, h->loss, h->queue, h->st);
                hipError_t e2 = e1 == hipSuccess ? sgns::launch_fast(f, h->bf16, h->np, wpb, blocks, h->st) : e1;
                hipError_t e3 = hipStreamEndCapture(h->st, &g);
                if (e2 != hipSuccess || e3 != hipSuccess)
                    return seterr(h, PAR_SGNS_EDEVICE, "graph capture failed: %s",
                                  hipGetErrorString(e2 != hipSuccess ? e2 : e3));
                hipGraphExec_t ge;
                hipError_t e4 = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (e4 != hipSuccess) return seterr(h, PAR_SGNS_EDEVICE, "graph instantiate: %s", hipGetErrorString(e4));
                it = h->graphs.emplace(k, ge).first;
                // upload the graph now, so its first launch (the one a caller may
                // time, when a chunk is trained once) carries no set-up work
                HIPCHK(h, hipGraphUpload(ge, h->st));
                HIPCHK(h, hipStreamSynchronize(h->st));
            }
            if (prepare_only) return 0;
            HIPCHK(h, hipEventRecord(h->ev0, h->st));
            HIPCHK(h, hipGraphLaunch(it->second, h->st));
            HIPCHK(h, hipEventRecord(h->ev1, h->st));
        } else {
            if (prepare_only) return 0;
            // Plain mode's optional staleness bound (min_sentences_per_wave >
            // 0): at most n / floor sentences in flight.  Where that is fewer
            // than fit on the device the range runs as consecutive launches
            // of that many sentences -- bulk-synchronous rounds.  Without it
            // every sentence of the range is one block of one launch, like
            // the reference GPU kernel's grid (text.cu:887-912).
            const int64_t n = a->sent_end - a->sent_begin;
            const int64_t resident = (int64_t)blocks_per_cu(f.lds_per_wave, wpb) * wpb * (h->cu_count > 0 ? h->cu_count : 256);
            int64_t part = n;
            if (!pers && a->min_sentences_per_wave > 0) {
                part = (n + a->min_sentences_per_wave - 1) / a->min_sentences_per_wave;
                if (part >= resident || part < 1) part = n;
            }
            HIPCHK(h, sgns::launch_reset(h->counters, sgns::kCntSlots, h->loss, h->queue, h->st));
            HIPCHK(h, hipEventRecord(h->ev0, h->st));
            launches = 0;
            for (int64_t b = a->sent_begin; b < a->sent_end; b += part) {
                f.s_begin = b;
                f.s_end = b + part < a->sent_end ? b + part : a->sent_end;
                HIPCHK(h, sgns::launch_fast(f, h->bf16, h->np, wpb, blocks, h->st));
                launches++;
            }
            HIPCHK(h, hipEventRecord(h->ev1, h->st));
            if (!pers) shape_blocks = (int)((part + wpb - 1) / wpb);
        }
    } else {
        return seterr(h, PAR_SGNS_EINVAL, "unknown mode %d", a->mode);
    }
    HIPCHK(h, hipStreamSynchronize(h->st));
    HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->last_ms[0] = ms;
    h->last_launches[0] = launches;
    if (stats) {
        unsigned long long c[sgns::kCntSlots];
        double l[2];
        HIPCHK(h, hipMemcpy(c, h->counters, sizeof c, hipMemcpyDeviceToHost));
        HIPCHK(h, hipMemcpy(l, h->loss, sizeof l, hipMemcpyDeviceToHost));
        stats->loss_attr = l[0];
        stats->loss_rep = l[1];
        stats->tokens = c[sgns::kCntTokens];
        stats->sentences = c[sgns::kCntSentences];
        stats->pairs = c[sgns::kCntPairs];
        stats->batches = c[sgns::kCntBatches];
        stats->negatives = c[sgns::kCntNegatives];
        stats->rep_pairs = c[sgns::kCntRepPairs];
        stats->kernel_ms = ms;
        const double e = h->bf16 ? 2.0 : 4.0, dim = (double)h->cfg.dim;
        const double tok = (double)stats->tokens, negs = (double)stats->negatives;
        stats->alg_bytes = 4.0 * tok + e * dim * (4.0 * tok + 2.0 * negs);
        stats->alg_flops = 6.0 * dim * ((double)stats->pairs + (double)stats->rep_pairs);
        stats->blocks = shape_blocks;
        stats->waves_per_block = shape_wpb;
        stats->sentences_per_dequeue = shape_grab;
        stats->reserved = 0;
    }
    return 0;
}

int par_sgns_train(par_sgns_handle h, const par_sgns_train_args *a, par_sgns_stats *stats) {
    return train_impl(h, a, stats, false);
}

int par_sgns_prepare(par_sgns_handle h, const par_sgns_train_args *a) { return train_impl(h, a, nullptr, true); }

int par_sgns_autotune(par_sgns_handle h, const par_sgns_train_args *a, int32_t reps, par_sgns_tune_point *points,
                      int32_t cap, int32_t *n_points, par_sgns_tune_point *best) {
    if (!h) return PAR_SGNS_EINVAL;
    if (!a || !best || !n_points || cap < 0 || (cap > 0 && !points)) return seterr(h, PAR_SGNS_EINVAL, "null argument");
    if (a->mode != PAR_SGNS_PERSISTENT) return seterr(h, PAR_SGNS_EINVAL, "the autotuner times the persistent mode");
    if (reps < 1 || reps > 64) return seterr(h, PAR_SGNS_EINVAL, "reps must be 1..64");
    if (!h->have_exp || !h->have_weights) return seterr(h, PAR_SGNS_ESTATE, "exp table and weights must be set first");
    HIPCHK(h, hipSetDevice(h->cfg.device));
    int rc;
    // the shape the first trial takes validates the arguments (and orders the range)
    par_sgns_train_args t = *a;
    t.use_graph = 0;
    t.waves_per_block = 1;
    t.blocks = h->cu_count > 0 ? h->cu_count : 256;
    t.sentences_per_dequeue = 1;
    if ((rc = train_impl(h, &t, nullptr, true))) return rc;
    const int lds = sgns::fast_lds_per_wave(h->R, h->nq, h->bf16, h->kmax, h->cfg.batch_size, h->ld,
                                            use_mfma(h, a->repulsion));
    int max_wpcu = 0;
    for (int k = 1; k <= kMaxWpb; k++) max_wpcu = std::max(max_wpcu, blocks_per_cu(lds, k) * k);
    if (max_wpcu < 1) return seterr(h, PAR_SGNS_EINVAL, "window ring does not fit in LDS (%d bytes per wave)", lds);
    std::vector<int> wpcus;
    for (int w : {4, 6, 8})
        if (w <= max_wpcu) wpcus.push_back(w);
    if (std::find(wpcus.begin(), wpcus.end(), max_wpcu) == wpcus.end()) wpcus.push_back(max_wpcu);
    // both tables aside: every trial starts from them and they come back at the end
    const size_t bytes = (size_t)h->cfg.vocab_size * (size_t)h->ld * (h->bf16 ? 2 : 4);
    char *s_in = nullptr, *s_out = nullptr;
    if ((rc = dalloc(h, s_in, bytes))) return rc;
    if ((rc = dalloc(h, s_out, bytes))) {
        dfree(s_in);
        return rc;
    }
    auto restore = [&]() -> hipError_t {
        hipError_t e = hipMemcpyAsync(h->w_in, s_in, bytes, hipMemcpyDeviceToDevice, h->st);
        if (e == hipSuccess) e = hipMemcpyAsync(h->w_out, s_out, bytes, hipMemcpyDeviceToDevice, h->st);
        return e;
    };
    hipError_t e = hipMemcpyAsync(s_in, h->w_in, bytes, hipMemcpyDeviceToDevice, h->st);
    if (e == hipSuccess) e = hipMemcpyAsync(s_out, h->w_out, bytes, hipMemcpyDeviceToDevice, h->st);
    const int64_t n_sent = a->sent_end - a->sent_begin;
    int32_t n = 0;
    par_sgns_tune_point bp{};
    bp.kernel_ms = -1.0;
    for (int wpcu : wpcus) {
        for (int wpb : {1, 2, 4, 8}) {
            if (e != hipSuccess || rc) break;
            if (wpcu % wpb || blocks_per_cu(lds, wpb) * wpb < wpcu) continue;
            for (int grab : {1, 2, 4}) {
                t.waves_per_block = wpb;
                t.blocks = (wpcu / wpb) * (h->cu_count > 0 ? h->cu_count : 256);
                const int64_t capb = n_sent / (a->min_sentences_per_wave > 0 ? a->min_sentences_per_wave : kMinSentencesPerWave) / wpb;
                if (t.blocks > capb) t.blocks = (int)(capb < 1 ? 1 : capb);
                t.sentences_per_dequeue = grab;
                std::vector<double> ms;
                for (int r = 0; r < reps && e == hipSuccess && !rc; r++) {
                    e = restore();
                    if (e == hipSuccess && !(rc = train_impl(h, &t, nullptr, false))) ms.push_back(h->last_ms[0]);
                }
                if (e != hipSuccess || rc) break;
                std::sort(ms.begin(), ms.end());
                par_sgns_tune_point p{wpb, wpcu, grab, 0, ms[ms.size() / 2]};
                if (n < cap) points[n] = p;
                n++;
                if (bp.kernel_ms < 0 || p.kernel_ms < bp.kernel_ms) bp = p;
            }
        }
    }
    if (e == hipSuccess) e = restore();
    if (e == hipSuccess) e = hipStreamSynchronize(h->st);
    dfree(s_in);
    dfree(s_out);
    if (rc) return rc;
    if (e != hipSuccess) return seterr(h, PAR_SGNS_EDEVICE, "autotune: %s", hipGetErrorString(e));
    h->tuned_wpb = bp.waves_per_block;
    h->tuned_wpcu = bp.waves_per_cu;
    h->tuned_grab = bp.sentences_per_dequeue;
    *best = bp;
    *n_points = n;
    return 0;
}

int par_sgns_set_launch_shape(par_sgns_handle h, int32_t wpb, int32_t wpcu, int32_t grab) {
    if (!h) return PAR_SGNS_EINVAL;
    if (wpb == 0 && wpcu == 0 && grab == 0) {
        h->tuned_wpb = h->tuned_wpcu = h->tuned_grab = 0;
        return 0;
    }
    if (wpb < 1 || wpb > kMaxWpb || wpcu < 1 || wpcu > kMaxWavesCu || wpcu % wpb || grab < 1 || grab > 64)
        return seterr(h, PAR_SGNS_EINVAL, "launch shape: waves per block 1..%d, waves per CU a multiple of it up to %d, "
                      "sentences per dequeue 1..64", kMaxWpb, kMaxWavesCu);
    h->tuned_wpb = wpb;
    h->tuned_wpcu = wpcu;
    h->tuned_grab = grab;
    return 0;
}

// noalloc: stage the indices and accumulate the loss in the device's static
// scratch instead of the handle's buffers (no allocation on this path).
static int window_call(par_sgns_handle h, const int32_t *centre, const int32_t *ctx, const int32_t *neg, int64_t nw,
                       int32_t nctx, int32_t nneg, int32_t flags, double *loss_out, int adjoint, bool noalloc = false) {
    if (!h) return PAR_SGNS_EINVAL;
    if (!centre || !ctx || (nneg > 0 && !neg) || nw < 0 || nctx < 1 || nctx > 32 || nneg < 0 || nneg > 7)
        return seterr(h, PAR_SGNS_EINVAL, "bad window batch (nctx 1..32, nneg 0..7)");
    if (!h->have_weights) return seterr(h, PAR_SGNS_ESTATE, "weights not set");
    if ((flags & PAR_SGNS_WIN_MFMA) && !h->bf16 && nctx + 1 + nneg > 16)
        return seterr(h, PAR_SGNS_EINVAL, "MFMA path on fp32 tables takes windows of at most 16 rows "
                      "(1 + nctx + nneg <= 16); bf16 tables take any window");
    if ((flags & PAR_SGNS_WIN_MFMA) && (flags & PAR_SGNS_WIN_FP64))
        return seterr(h, PAR_SGNS_EINVAL, "MFMA path accumulates in fp32 (no FP64 variant)");
    if ((flags & PAR_SGNS_WIN_MFMA) && h->bf16 && (uint64_t)h->cfg.vocab_size * (uint64_t)h->ld * 2u > 0xFFFFF000ull)
        return seterr(h, PAR_SGNS_EINVAL, "MFMA path addresses each table through a buffer resource: tables below "
                      "4 GiB - 4 KiB");
    const int32_t V = h->cfg.vocab_size;
    int rc;
    if ((rc = check_ids(h, centre, nw, 0, V, "centre"))) return rc;
    if ((rc = check_ids(h, ctx, nw * nctx, 0, V, "ctx"))) return rc;
    if ((rc = check_ids(h, neg, nw * nneg, 0, V, "neg"))) return rc;
    HIPCHK(h, hipSetDevice(h->cfg.device));
    const size_t need = (size_t)nw * (size_t)(1 + nctx + nneg);
    int32_t *dc = nullptr;
    double *dloss = h->loss;
    if (noalloc) {
        if (need > (kScratchBytes - kScratchIdx) / sizeof(int32_t))
            return seterr(h, PAR_SGNS_EINVAL,
                          "window batch needs %zu bytes of indices; the allocation-free entry point stages at most "
                          "%zu (n_windows * (1 + nctx + nneg) <= %zu)",
                          need * sizeof(int32_t), kScratchBytes - kScratchIdx,
                          (kScratchBytes - kScratchIdx) / sizeof(int32_t));
        dc = reinterpret_cast<int32_t *>(h->scratch + kScratchIdx);
        dloss = reinterpret_cast<double *>(h->scratch);
    } else {
        if (need > h->win_cap) {
            if ((rc = dalloc(h, h->win_buf, need))) return rc;
            h->win_cap = need;
        }
        dc = h->win_buf;
    }
    int32_t *dx = dc + nw, *dn = dx + nw * nctx;
    if (nw > 0) {  // (an empty batch copies nothing and launches no window kernel)
        HIPCHK(h, hipMemcpyAsync(dc, centre, (size_t)nw * 4, hipMemcpyHostToDevice, h->st));
        HIPCHK(h, hipMemcpyAsync(dx, ctx, (size_t)nw * nctx * 4, hipMemcpyHostToDevice, h->st));
        if (nneg > 0) HIPCHK(h, hipMemcpyAsync(dn, neg, (size_t)nw * nneg * 4, hipMemcpyHostToDevice, h->st));
    }
    const size_t rows = (size_t)V * (size_t)h->ld;
    if (adjoint) {
        if (!h->g_in) {
            if ((rc = dalloc(h, h->g_in, rows))) return rc;
            if ((rc = dalloc(h, h->g_out, rows))) return rc;
        }
        HIPCHK(h, hipMemsetAsync(h->g_in, 0, rows * 4, h->st));
        HIPCHK(h, hipMemsetAsync(h->g_out, 0, rows * 4, h->st));
    }
    // loss accumulators cleared by a kernel (no memset node if a caller captures this)
    if (noalloc) HIPCHK(h, sgns::launch_reset(nullptr, 0, dloss, nullptr, h->st));
    else HIPCHK(h, sgns::launch_reset(h->counters, sgns::kCntSlots, h->loss, nullptr, h->st));
    sgns::WindowArgs w{};
    w.centre = dc; w.ctx = dx; w.neg = dn; w.n_windows = nw; w.nctx = nctx; w.nneg = nneg;
    w.d = h->cfg.dim; w.ld = h->ld; w.V = V; w.w_in = h->w_in; w.w_out = h->w_out; w.g_in = h->g_in; w.g_out = h->g_out;
    w.loss = dloss;
    HIPCHK(h, hipEventRecord(h->ev0, h->st));
    hipError_t e = nw > 0 ? sgns::launch_window(w, h->bf16, h->np, adjoint, flags, h->st) : hipSuccess;
    if (e == hipErrorNotSupported) return seterr(h, PAR_SGNS_EINVAL, "requested window path is not built");
    HIPCHK(h, e);
    HIPCHK(h, hipEventRecord(h->ev1, h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    float ms = 0.f;
    HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->last_ms[adjoint ? 2 : 1] = ms;
    h->last_launches[adjoint ? 2 : 1] = 1;
    double l[2];
    HIPCHK(h, hipMemcpy(l, dloss, sizeof l, hipMemcpyDeviceToHost));
    if (loss_out) *loss_out = l[0];
    return 0;
}

int par_sgns_loss_forward_noalloc(par_sgns_handle h, const int32_t *centre, const int32_t *ctx, const int32_t *neg,
                                  int64_t n_windows, int32_t nctx, int32_t nneg, int32_t flags, double *loss_out) {
    return window_call(h, centre, ctx, neg, n_windows, nctx, nneg, flags, loss_out, 0, true);
}

int par_sgns_loss_forward(par_sgns_handle h, const int32_t *centre, const int32_t *ctx, const int32_t *neg,
                          int64_t n_windows, int32_t nctx, int32_t nneg, int32_t flags, double *loss_out) {
    return window_call(h, centre, ctx, neg, n_windows, nctx, nneg, flags, loss_out, 0);
}

int par_sgns_loss_adjoint(par_sgns_handle h, const int32_t *centre, const int32_t *ctx, const int32_t *neg,
                          int64_t n_windows, int32_t nctx, int32_t nneg, int32_t flags, double *loss_out) {
    return window_call(h, centre, ctx, neg, n_windows, nctx, nneg, flags, loss_out, 1);
}

int par_sgns_get_grads(par_sgns_handle h, float *g_in, float *g_out) {
    if (!h) return PAR_SGNS_EINVAL;
    if (!h->g_in) return seterr(h, PAR_SGNS_ESTATE, "no adjoint has run");
    const int64_t V = h->cfg.vocab_size, d = h->cfg.dim, ld = h->ld;
    std::vector<float> buf((size_t)V * ld);
    float *dsts[2] = {g_in, g_out};
    float *srcs[2] = {h->g_in, h->g_out};
    for (int k = 0; k < 2; k++) {
        if (!dsts[k]) continue;
        HIPCHK(h, hipMemcpy(buf.data(), srcs[k], buf.size() * 4, hipMemcpyDeviceToHost));
        for (int64_t r = 0; r < V; r++) std::memcpy(dsts[k] + r * d, buf.data() + r * ld, (size_t)d * 4);
    }
    return 0;
}

int par_sgns_last_kernel_time(par_sgns_handle h, int32_t kind, double *ms, int32_t *launches) {
    if (!h || kind < 0 || kind > 2) return PAR_SGNS_EINVAL;
    if (ms) *ms = h->last_ms[kind];
    if (launches) *launches = h->last_launches[kind];
    return 0;
}

}  // extern "C"
