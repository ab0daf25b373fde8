// par_host.cpp -- host side of the MI355X PAR-SGNS trainer: tokenizer,
// vocabulary, id streams, subsampling into a sentence CSR, the exact
// single-thread random schedule (deterministic mode), synthetic corpora and
// the embedding writer.  C-ABI declared in include/par_host.h.
//
// Reference semantics followed (PAR_Word2Vec_cpu.cpp, "cpu:"):
//   tokenizer cpu:120-144, vocabulary cpu:147-335, stream cpu:359-370,
//   subsampling + sentence cut cpu:447-471, training schedule cpu:430-627,
//   unigram table cpu:96-117, expTable cpu:822-826, InitNet cpu:337-357,
//   writer cpu:721-737.
#include "par_host.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include <omp.h>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

// ------------------------------------------------------------ randomness --
inline uint64_t mix64(uint64_t z) {  // splitmix64 finaliser
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct Rng {  // xoshiro256** seeded through splitmix64
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (int i = 0; i < 4; i++) { z = mix64(z + (uint64_t)i * 0x632BE59BD9B4E019ULL); s[i] = z; }
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    uint64_t below(uint64_t n) { return (uint64_t)(uniform() * (double)n) % n; }
    double normal() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

inline uint64_t lcg(uint64_t x) { return x * 25214903917ULL + 11ULL; }  // cpu:461

// Walker/Vose alias table over non-negative weights (double precision).
struct Alias {
    std::vector<double> prob;
    std::vector<int32_t> alias;
    void build(const std::vector<double> &w) {
        size_t n = w.size();
        prob.assign(n, 0.0);
        alias.assign(n, 0);
        double tot = 0;
        for (double x : w) tot += x;
        std::vector<double> q(n);
        std::vector<int32_t> small, large;
        for (size_t i = 0; i < n; i++) {
            q[i] = tot > 0 ? w[i] * (double)n / tot : 1.0;
            (q[i] < 1.0 ? small : large).push_back((int32_t)i);
        }
        while (!small.empty() && !large.empty()) {
            int32_t s = small.back(), l = large.back();
            small.pop_back();
            prob[s] = q[s];
            alias[s] = l;
            q[l] = (q[l] + q[s]) - 1.0;
            if (q[l] < 1.0) { large.pop_back(); small.push_back(l); }
        }
        for (int32_t i : large) { prob[i] = 1.0; alias[i] = i; }
        for (int32_t i : small) { prob[i] = 1.0; alias[i] = i; }
    }
    int32_t sample(Rng &r) const {
        double u = r.uniform() * (double)prob.size();
        size_t k = (size_t)u;
        if (k >= prob.size()) k = prob.size() - 1;
        return (u - (double)k) < prob[k] ? (int32_t)k : alias[k];
    }
};

// ------------------------------------------------------------ tokenizer ----
// Emits the tokens ReadWord (cpu:120-144) would return, in order, with the
// reference's end-of-file rule: a word cut off by EOF sets feof and is
// discarded by every caller (cpu:172, 269).  Words keep their first 98 bytes
// (MAX_STRING 100, cpu:140-143) and end at an embedded NUL like a C string.
// The bytes [begin, end) of the file; a range that starts right after a
// space, tab or newline (or at 0) yields exactly the tokens one pass over the
// whole file yields there.
template <class F>
int for_each_token_range(const char *path, int64_t begin, int64_t end, F &&emit) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(PH_EIO, std::string("cannot open ") + path);
    if (begin > 0 && std::fseek(f, (long)begin, SEEK_SET) != 0) {
        std::fclose(f);
        return fail(PH_EIO, std::string("cannot seek in ") + path);
    }
    const size_t BUF = 1 << 22;
    std::unique_ptr<unsigned char[]> buf(new (std::nothrow) unsigned char[BUF]);
    if (!buf) { std::fclose(f); return fail(PH_ENOMEM, "tokenizer buffer"); }
    char word[100];
    int a = 0;
    static const char EOS[] = "</s>";
    int64_t left = end < 0 ? INT64_MAX : end - begin;
    while (left > 0) {
        size_t got = std::fread(buf.get(), 1, (size_t)std::min<int64_t>((int64_t)BUF, left), f);
        if (got == 0) break;
        left -= (int64_t)got;
        for (size_t k = 0; k < got; k++) {
            int ch = buf[k];
            if (ch == 13) continue;
            if (ch == ' ' || ch == '\t' || ch == '\n') {
                if (a > 0) {
                    word[a] = 0;
                    emit(word, std::strlen(word));
                    a = 0;
                }
                if (ch == '\n') emit(EOS, 4);
                continue;
            }
            word[a++] = (char)ch;
            if (a >= 99) a--;
        }
    }
    std::fclose(f);
    return 0;  // a pending word (a > 0) at EOF is dropped
}

template <class F>
int for_each_token(const char *path, F &&emit) {
    return for_each_token_range(path, 0, -1, emit);
}

// Cuts the file into at most `parts` byte ranges for for_each_token_range:
// each cut lies right after the first space, tab or newline at or past an
// equal share of the bytes.  Returns the boundaries (first 0, last the file
// size), or an empty vector if the file cannot be read.
std::vector<int64_t> token_ranges(const char *path, int parts) {
    std::vector<int64_t> cut;
    FILE *f = std::fopen(path, "rb");
    if (!f) return cut;
    std::fseek(f, 0, SEEK_END);
    const int64_t size = (int64_t)std::ftell(f);
    const int64_t kMinPart = 1 << 20;  // smaller files: one range
    if (parts > size / kMinPart) parts = (int)std::max<int64_t>(1, size / kMinPart);
    cut.push_back(0);
    unsigned char tmp[4096];
    for (int k = 1; k < parts; k++) {
        int64_t p = std::max(size * k / parts, cut.back());
        int64_t at = size;
        std::fseek(f, (long)p, SEEK_SET);
        while (p < size && at == size) {
            size_t got = std::fread(tmp, 1, sizeof tmp, f);
            if (got == 0) break;
            for (size_t i = 0; i < got; i++)
                if (tmp[i] == ' ' || tmp[i] == '\t' || tmp[i] == '\n') { at = p + (int64_t)i + 1; break; }
            p += (int64_t)got;
        }
        if (at > cut.back() && at < size) cut.push_back(at);
    }
    cut.push_back(size);
    std::fclose(f);
    return cut;
}

// Open-addressing string -> id map (FNV-1a); the reference's own hash only
// decides lookup cost, never ids or order.
struct WordMap {
    std::vector<int32_t> slot;
    const std::vector<std::string> *words = nullptr;
    size_t mask = 0;
    static uint64_t h(const char *s, size_t n) {
        uint64_t x = 1469598103934665603ULL;
        for (size_t i = 0; i < n; i++) { x ^= (unsigned char)s[i]; x *= 1099511628211ULL; }
        return x;
    }
    void reset(size_t expect, const std::vector<std::string> *w) {
        size_t cap = 1024;
        while (cap < expect * 2) cap <<= 1;
        slot.assign(cap, -1);
        mask = cap - 1;
        words = w;
    }
    int32_t find(const char *s, size_t n) const {
        if (slot.empty()) return -1;
        size_t i = h(s, n) & mask;
        for (;;) {
            int32_t k = slot[i];
            if (k < 0) return -1;
            const std::string &w = (*words)[(size_t)k];
            if (w.size() == n && std::memcmp(w.data(), s, n) == 0) return k;
            i = (i + 1) & mask;
        }
    }
    void insert(const char *s, size_t n, int32_t id) {
        size_t i = h(s, n) & mask;
        while (slot[i] >= 0) i = (i + 1) & mask;
        slot[i] = id;
    }
};

}  // namespace

struct ph_vocab {
    std::vector<std::string> words;
    std::vector<uint32_t> cn;
    WordMap map;
    uint64_t train_words = 0;
    std::vector<int32_t> raw_map;  // raw id -> vocab id (raw corpora only)

    void rebuild_map() {
        map.reset(words.size() + 16, &words);
        for (size_t i = 0; i < words.size(); i++) map.insert(words[i].data(), words[i].size(), (int32_t)i);
    }
    int32_t add(const char *s, size_t n) {
        if (map.slot.empty() || words.size() * 2 + 2 > map.slot.size()) {
            words.emplace_back(s, n);
            cn.push_back(0);
            rebuild_map();
            return (int32_t)words.size() - 1;
        }
        words.emplace_back(s, n);
        cn.push_back(0);
        map.insert(s, n, (int32_t)words.size() - 1);
        return (int32_t)words.size() - 1;
    }
    // ReduceVocab (cpu:229-251): drop counts <= min_reduce, compact in order.
    void reduce(uint32_t min_reduce) {
        size_t k = 0;
        for (size_t i = 0; i < words.size(); i++)
            if (cn[i] > min_reduce) { words[k] = std::move(words[i]); cn[k] = cn[i]; k++; }
        words.resize(k);
        cn.resize(k);
        rebuild_map();
    }
    // SortVocab (cpu:204-226): stable sort of [1, V) by count descending,
    // prune count < min_count except index 0, train_words = kept counts.
    std::vector<int32_t> sort_prune(int32_t min_count) {
        size_t n = words.size();
        std::vector<int32_t> order(n);
        for (size_t i = 0; i < n; i++) order[i] = (int32_t)i;
        if (n > 1)
            std::stable_sort(order.begin() + 1, order.end(),
                             [&](int32_t x, int32_t y) { return cn[(size_t)x] > cn[(size_t)y]; });
        std::vector<std::string> nw;
        std::vector<uint32_t> nc;
        std::vector<int32_t> old_to_new(n, -1);
        train_words = 0;
        for (size_t r = 0; r < n; r++) {
            int32_t o = order[r];
            if (cn[(size_t)o] < (uint32_t)min_count && r != 0) continue;
            old_to_new[(size_t)o] = (int32_t)nw.size();
            nw.push_back(std::move(words[(size_t)o]));
            nc.push_back(cn[(size_t)o]);
            train_words += cn[(size_t)o];
        }
        words.swap(nw);
        cn.swap(nc);
        rebuild_map();
        return old_to_new;
    }
};

extern "C" {

const char *ph_last_error(void) { return g_err.c_str(); }

ph_vocab *ph_vocab_learn_text(const char *path, int32_t min_count) {
    if (!path) { fail(PH_EINVAL, "null path"); return nullptr; }
    std::unique_ptr<ph_vocab> v(new (std::nothrow) ph_vocab);
    if (!v) { fail(PH_ENOMEM, "vocab"); return nullptr; }
    v->add("</s>", 4);
    uint32_t min_reduce = 1;
    const size_t reduce_at = (size_t)(30000000 * 0.7);  // cpu:74, 282
    // Byte ranges counted in parallel, each with its words in first-appearance
    // order, then merged in file order: the same words, counts and first-
    // appearance order as one pass -- as long as the one pass never reaches
    // ReduceVocab's bound, i.e. the merged vocabulary stays within it;
    // otherwise the one pass below runs.
    const std::vector<int64_t> cut = token_ranges(path, omp_get_max_threads());
    if (cut.size() > 2) {
        const int parts = (int)cut.size() - 1;
        std::vector<ph_vocab> local((size_t)parts);
        std::vector<int> rcs((size_t)parts, 0);
#pragma omp parallel for schedule(dynamic, 1)
        for (int k = 0; k < parts; k++) {
            ph_vocab &lv = local[(size_t)k];
            rcs[(size_t)k] = for_each_token_range(path, cut[(size_t)k], cut[(size_t)k + 1], [&](const char *s, size_t n) {
                if (lv.words.size() > reduce_at) return;  // the merged vocabulary is past the bound too
                int32_t i = lv.map.find(s, n);
                if (i < 0) { i = lv.add(s, n); lv.cn[(size_t)i] = 1; }
                else lv.cn[(size_t)i]++;
            });
        }
        for (int k = 0; k < parts; k++)
            if (rcs[(size_t)k] < 0) return nullptr;
        bool over = false;
        for (int k = 0; k < parts && !over; k++) {
            const ph_vocab &lv = local[(size_t)k];
            over = lv.words.size() > reduce_at;
            for (size_t j = 0; j < lv.words.size() && !over; j++) {
                const std::string &w = lv.words[j];
                int32_t i = v->map.find(w.data(), w.size());
                if (i < 0) i = v->add(w.data(), w.size());
                v->cn[(size_t)i] += lv.cn[j];
                over = v->words.size() > reduce_at;
            }
        }
        if (!over) {
            v->sort_prune(min_count);
            return v.release();
        }
        v.reset(new (std::nothrow) ph_vocab);
        if (!v) { fail(PH_ENOMEM, "vocab"); return nullptr; }
        v->add("</s>", 4);
    }
    int rc = for_each_token(path, [&](const char *s, size_t n) {
        int32_t i = v->map.find(s, n);
        if (i < 0) { i = v->add(s, n); v->cn[(size_t)i] = 1; }
        else v->cn[(size_t)i]++;
        if (v->words.size() > reduce_at) v->reduce(min_reduce++);
    });
    if (rc < 0) return nullptr;
    v->sort_prune(min_count);
    return v.release();
}

ph_vocab *ph_vocab_read(const char *vocab_path, int32_t min_count) {
    FILE *f = vocab_path ? std::fopen(vocab_path, "rb") : nullptr;
    if (!f) { fail(PH_EIO, "cannot open vocabulary file"); return nullptr; }
    std::unique_ptr<ph_vocab> v(new ph_vocab);
    char line[4096];
    while (std::fgets(line, sizeof line, f)) {
        char *sp = std::strrchr(line, ' ');
        if (!sp) continue;
        *sp = 0;
        size_t n = std::strlen(line);
        if (n > 98) n = 98;
        int32_t i = v->add(line, n);
        v->cn[(size_t)i] = (uint32_t)std::strtoul(sp + 1, nullptr, 10);
    }
    std::fclose(f);
    v->sort_prune(min_count);
    return v.release();
}

int ph_vocab_save(const ph_vocab *v, const char *path) {
    if (!v || !path) return fail(PH_EINVAL, "null argument");
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(PH_EIO, "cannot create vocabulary file");
    for (size_t i = 0; i < v->words.size(); i++) std::fprintf(f, "%s %d\n", v->words[i].c_str(), (int)v->cn[i]);
    std::fclose(f);
    return 0;
}

ph_vocab *ph_vocab_from_raw(const int32_t *raw, int64_t n, int32_t raw_vocab, int32_t min_count) {
    if (!raw || n < 0 || raw_vocab < 1) { fail(PH_EINVAL, "bad raw corpus"); return nullptr; }
    // The text rendering drops a last word not followed by a newline.
    int64_t end = (n > 0 && raw[n - 1] != 0) ? n - 1 : n;
    std::vector<uint32_t> cnt((size_t)raw_vocab + 1, 0);
    std::vector<int32_t> first_order;
    first_order.reserve((size_t)std::min<int64_t>(raw_vocab, 1 << 20));
    for (int64_t i = 0; i < end; i++) {
        int32_t r = raw[i];
        if (r < 0 || r > raw_vocab) { fail(PH_EINVAL, "raw id out of range"); return nullptr; }
        if (r != 0 && cnt[(size_t)r] == 0) first_order.push_back(r);
        cnt[(size_t)r]++;
    }
    std::unique_ptr<ph_vocab> v(new ph_vocab);
    v->words.reserve(first_order.size() + 1);
    v->words.emplace_back("</s>");
    v->cn.push_back(cnt[0]);
    std::vector<int32_t> pre_id((size_t)raw_vocab + 1, -1);
    pre_id[0] = 0;
    for (int32_t r : first_order) {
        pre_id[(size_t)r] = (int32_t)v->words.size();
        v->words.push_back("w" + std::to_string(r));
        v->cn.push_back(cnt[(size_t)r]);
    }
    std::vector<int32_t> o2n = v->sort_prune(min_count);
    v->raw_map.assign((size_t)raw_vocab + 1, -1);
    for (int32_t r = 0; r <= raw_vocab; r++)
        if (pre_id[(size_t)r] >= 0) v->raw_map[(size_t)r] = o2n[(size_t)pre_id[(size_t)r]];
    return v.release();
}

void ph_vocab_free(ph_vocab *v) { delete v; }
int32_t ph_vocab_size(const ph_vocab *v) { return v ? (int32_t)v->words.size() : PH_EINVAL; }
uint64_t ph_vocab_train_words(const ph_vocab *v) { return v ? v->train_words : 0; }

int ph_vocab_counts(const ph_vocab *v, uint32_t *out) {
    if (!v || !out) return fail(PH_EINVAL, "null argument");
    std::memcpy(out, v->cn.data(), v->cn.size() * sizeof(uint32_t));
    return 0;
}

const char *ph_vocab_word(const ph_vocab *v, int32_t id) {
    if (!v || id < 0 || (size_t)id >= v->words.size()) return nullptr;
    return v->words[(size_t)id].c_str();
}

int32_t ph_vocab_lookup(const ph_vocab *v, const char *word) {
    if (!v || !word) return -1;
    return v->map.find(word, std::strlen(word));
}

int ph_vocab_raw_map(const ph_vocab *v, int32_t *out, int32_t raw_vocab) {
    if (!v || !out) return fail(PH_EINVAL, "null argument");
    if (v->raw_map.empty()) return fail(PH_EINVAL, "vocabulary was not built from a raw corpus");
    if ((size_t)raw_vocab + 1 != v->raw_map.size()) return fail(PH_EINVAL, "raw vocabulary size mismatch");
    std::memcpy(out, v->raw_map.data(), v->raw_map.size() * sizeof(int32_t));
    return 0;
}

int64_t ph_load_text_stream(const ph_vocab *v, const char *path, int32_t *out, int64_t cap) {
    if (!v || !path) return fail(PH_EINVAL, "null argument");
    // byte ranges tokenised in parallel, each into its own id list, then
    // concatenated in file order
    std::vector<int64_t> cut = token_ranges(path, omp_get_max_threads());
    if (cut.empty()) return fail(PH_EIO, std::string("cannot open ") + path);
    const int parts = (int)cut.size() - 1;
    std::vector<std::vector<int32_t>> ids((size_t)parts);
    std::vector<int64_t> cnt((size_t)parts + 1, 0);
    std::vector<int> rcs((size_t)parts, 0);
    std::atomic<int> oom(0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < parts; k++) {
        std::vector<int32_t> &o = ids[(size_t)k];
        int64_t n = 0;
        rcs[(size_t)k] = for_each_token_range(path, cut[(size_t)k], cut[(size_t)k + 1], [&](const char *s, size_t len) {
            int32_t id = v->map.find(s, len);
            if (id < 0) return;
            n++;
            if (out && !oom) {
                try { o.push_back(id); } catch (const std::bad_alloc &) { oom = 1; }
            }
        });
        cnt[(size_t)k + 1] = n;
    }
    for (int k = 0; k < parts; k++)
        if (rcs[(size_t)k] < 0) return rcs[(size_t)k];
    if (oom) return fail(PH_ENOMEM, "stream buffers");
    for (int k = 0; k < parts; k++) cnt[(size_t)k + 1] += cnt[(size_t)k];
    const int64_t n = cnt[(size_t)parts];
    if (!out) return n;
    if (n > cap) return fail(PH_ERANGE, "stream capacity too small");
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < parts; k++)
        if (!ids[(size_t)k].empty())
            std::memcpy(out + cnt[(size_t)k], ids[(size_t)k].data(), ids[(size_t)k].size() * sizeof(int32_t));
    return n;
}

int64_t ph_map_raw_stream(const ph_vocab *v, const int32_t *raw, int64_t n, int32_t *out, int64_t cap) {
    if (!v || !raw || n < 0) return fail(PH_EINVAL, "null argument");
    if (v->raw_map.empty()) return fail(PH_EINVAL, "vocabulary was not built from a raw corpus");
    int64_t end = (n > 0 && raw[n - 1] != 0) ? n - 1 : n;
    // parallel two-pass compaction
    int nt = omp_get_max_threads();
    int64_t blk = (end + nt - 1) / std::max(nt, 1);
    std::vector<int64_t> cnt((size_t)nt + 1, 0);
    const int32_t *map = v->raw_map.data();
    int32_t rv = (int32_t)v->raw_map.size() - 1;
    std::atomic<int> bad(0);
#pragma omp parallel num_threads(nt)
    {
        int t = omp_get_thread_num();
        int64_t lo = std::min(end, blk * t), hi = std::min(end, lo + blk);
        int64_t c = 0;
        for (int64_t i = lo; i < hi; i++) {
            int32_t r = raw[i];
            if (r < 0 || r > rv) { bad = 1; continue; }
            c += map[r] >= 0;
        }
        cnt[(size_t)t + 1] = c;
#pragma omp barrier
#pragma omp single
        for (int k = 0; k < nt; k++) cnt[(size_t)k + 1] += cnt[(size_t)k];
        int64_t o = cnt[(size_t)t];
        if (out && cnt[(size_t)nt] <= cap)
            for (int64_t i = lo; i < hi; i++) {
                int32_t r = raw[i];
                if (r < 0 || r > rv) continue;
                int32_t id = map[r];
                if (id >= 0) out[o++] = id;
            }
    }
    if (bad) return fail(PH_EINVAL, "raw id out of range");
    if (out && cnt[(size_t)nt] > cap) return fail(PH_ERANGE, "stream capacity too small");
    return cnt[(size_t)nt];
}

// ------------------------------------------------------------- subsampling
int64_t ph_subsample(const int32_t *stream, int64_t n, const uint32_t *counts, int32_t V,
                     uint64_t train_words, float sample, uint64_t seed, int32_t max_len,
                     int32_t *tokens, int64_t cap_tokens, int64_t *sent_off, int64_t *sent_raw,
                     int64_t cap_sent, int64_t *n_tokens_out) {
    if (!stream || n < 0 || !counts || V < 1 || max_len < 1) return fail(PH_EINVAL, "bad subsample arguments");
    // keep threshold per word, float arithmetic of cpu:459-460
    std::vector<float> ran((size_t)V, 2.0f);
    if (sample > 0)
        for (int32_t w = 1; w < V; w++) {
            float ratio = (sample * (float)train_words) / (float)counts[w];
            ran[(size_t)w] = std::sqrt(ratio) + ratio;
        }
    // blocks aligned to line starts so every block starts a sentence
    int nt = omp_get_max_threads();
    int64_t nb = std::max<int64_t>(1, std::min<int64_t>((int64_t)nt * 8, n / 4096 + 1));
    std::vector<int64_t> bstart((size_t)nb + 1, n);
    bstart[0] = 0;
    for (int64_t b = 1; b < nb; b++) {
        int64_t p = n * b / nb;
        while (p < n && p > 0 && stream[p - 1] != 0) p++;
        bstart[(size_t)b] = p;
    }
    for (int64_t b = nb - 1; b >= 1; b--) bstart[(size_t)b] = std::min(bstart[(size_t)b], bstart[(size_t)b + 1]);
    std::vector<int64_t> bs((size_t)nb + 1, 0), bt((size_t)nb + 1, 0);
    std::atomic<int> bad(0);
    auto run = [&](int64_t b, bool write, int64_t so, int64_t to) {
        int64_t s = 0, t = 0, len = 0, start = bstart[(size_t)b];
        int64_t lo = bstart[(size_t)b], hi = bstart[(size_t)b + 1];
        for (int64_t i = lo; i < hi; i++) {
            int32_t w = stream[i];
            if (w < 0 || w >= V) { bad = 1; continue; }
            bool end_line = (w == 0);
            if (!end_line) {
                bool keep = true;
                if (sample > 0) {
                    uint64_t r = mix64(seed ^ mix64((uint64_t)i));
                    keep = !(ran[(size_t)w] < (float)(r & 0xFFFF) / 65536.f);
                }
                if (keep) {
                    if (write) tokens[to + t] = w;
                    t++;
                    len++;
                }
            }
            if (end_line || len >= max_len) {
                if (len > 0) {
                    if (write) { sent_off[so + s] = to + t - len; sent_raw[so + s] = start; }
                    s++;
                }
                len = 0;
                start = i + 1;
            }
        }
        // a trailing partial sentence without a terminating 0 is dropped
        // (it would read the sentinel, cpu:472)
        t -= len;
        return std::make_pair(s, t);
    };
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int64_t b = 0; b < nb; b++) {
        auto r = run(b, false, 0, 0);
        bs[(size_t)b + 1] = r.first;
        bt[(size_t)b + 1] = r.second;
    }
    if (bad) return fail(PH_EINVAL, "token id out of range");
    // only the final block may end without a terminator; earlier blocks end on
    // a line boundary by construction
    for (int64_t b = 0; b < nb; b++) { bs[(size_t)b + 1] += bs[(size_t)b]; bt[(size_t)b + 1] += bt[(size_t)b]; }
    int64_t S = bs[(size_t)nb], T = bt[(size_t)nb];
    if (n_tokens_out) *n_tokens_out = T;
    if (!tokens || !sent_off || !sent_raw) return S;
    if (S > cap_sent || T > cap_tokens) return fail(PH_ERANGE, "subsample capacity too small");
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int64_t b = 0; b < nb; b++) run(b, true, bs[(size_t)b], bt[(size_t)b]);
    sent_off[S] = T;
    return S;
}

// ----------------------------------------------------------- unigram table
int ph_unigram_entries(const uint32_t *cn, int32_t V, int64_t table_size, int64_t *entries) {
    if (!cn || V < 1 || table_size < 1 || !entries) return fail(PH_EINVAL, "bad unigram arguments");
    const double power = (double)0.75f;
    double z = 0.0;
    for (int32_t i = 0; i < V; i++) z += std::pow((double)cn[i], power);
    for (int32_t i = 0; i < V; i++) entries[i] = 0;
    const float fs = (float)table_size;
    // predicate of cpu:110, monotone in a
    auto past = [&](int64_t a, float d1) { return (float)a / fs > d1; };
    int32_t i = 0;
    float d1 = (float)(std::pow((double)cn[0], power) / z);
    int64_t a = 0;
    while (a < table_size) {
        if (i == V - 1) { entries[i] += table_size - a; break; }
        // first a' >= a with past(a', d1)
        int64_t lo = a, hi = table_size;  // answer in [lo, hi], hi = none
        if (!past(lo, d1)) {
            int64_t step = 1;
            int64_t probe = lo + 1;
            while (probe < table_size && !past(probe, d1)) { lo = probe; step <<= 1; probe = lo + step; }
            hi = std::min(probe, table_size);
            // invariant: !past(lo); answer in (lo, hi]
            while (hi - lo > 1) {
                int64_t mid = lo + (hi - lo) / 2;
                if (past(mid, d1)) hi = mid; else lo = mid;
            }
        } else {
            hi = lo;
        }
        if (hi >= table_size) { entries[i] += table_size - a; break; }
        entries[i] += hi - a + 1;
        a = hi + 1;
        i++;
        if (i < V) d1 = (float)((double)d1 + std::pow((double)cn[i], power) / z);
        if (i >= V) i = V - 1;
    }
    return 0;
}

// ------------------------------------------------------ deterministic plan
}  // extern "C"

struct ph_plan {
    std::vector<int32_t> tokens;
    std::vector<uint8_t> b;
    std::vector<int64_t> sent_off{0};
    std::vector<float> sent_alpha;
    std::vector<int64_t> batch_off{0};
    std::vector<int64_t> neg_off{0};
    std::vector<int32_t> negs;
    ph_plan_info info{};
};

extern "C" {

ph_plan *ph_plan_build(const int32_t *stream, int64_t n_words, const uint32_t *counts, int32_t V,
                       uint64_t train_words, int32_t W, int32_t NEG, int32_t B, int32_t iter,
                       float alpha0, float sample) {
    if (!stream || n_words < 0 || !counts || V < 2 || W < 1 || W > 255 || NEG < 0 || B < 1 || iter < 1) {
        fail(PH_EINVAL, "bad plan arguments");
        return nullptr;
    }
    const int64_t TABLE = 100000000;  // cpu:75
    std::vector<int64_t> ent((size_t)V);
    if (ph_unigram_entries(counts, V, TABLE, ent.data()) < 0) return nullptr;
    std::vector<int64_t> cum((size_t)V + 1, 0);
    for (int32_t i = 0; i < V; i++) cum[(size_t)i + 1] = cum[(size_t)i] + ent[(size_t)i];
    auto table_at = [&](int64_t idx) {
        return (int32_t)(std::upper_bound(cum.begin(), cum.end(), idx) - cum.begin() - 1);
    };
    std::vector<float> ran((size_t)V, 2.0f);
    if (sample > 0)
        for (int32_t w = 1; w < V; w++) {
            float ratio = (sample * (float)train_words) / (float)counts[w];
            ran[(size_t)w] = std::sqrt(ratio) + ratio;
        }
    std::unique_ptr<ph_plan> P(new ph_plan);
    std::vector<int64_t> stamp((size_t)V, -1);  // membership of the current sentence
    uint64_t x = 0;                            // thread 0 seed, cpu:401
    uint64_t draws = 0;
    float alpha = alpha0;
    const float floor_a = alpha0 * 0.0001f;
    const float denom = (float)((uint64_t)iter * train_words + 1);
    uint64_t wc = 0, last = 0, wca = 0, raw = 0;
    int local_iter = iter;
    int32_t sen[1000];
    int64_t sid = 0;
    for (;;) {
        // one loop iteration of cpu:430 per centre; a sentence spans L of them
        auto check_alpha = [&]() {
            if (wc - last > 10000) {
                wca += wc - last;
                last = wc;
                alpha = alpha0 * (1 - (float)wca / denom);
                if (alpha < floor_a) alpha = floor_a;
            }
        };
        check_alpha();
        int L = 0;
        for (;;) {
            int32_t w = wc <= (uint64_t)n_words ? (wc < (uint64_t)n_words ? stream[wc] : 0) : 0;
            wc++;
            if (w == 0) break;
            if (sample > 0) {
                x = lcg(x);
                draws++;
                if (ran[(size_t)w] < (float)(x & 0xFFFF) / 65536.f) continue;
            }
            sen[L++] = w;
            if (L >= 1000) break;
        }
        if (wc > (uint64_t)n_words) {  // epoch end, cpu:472-527
            wca += wc - last;
            raw += (uint64_t)n_words;
            if (--local_iter == 0) break;
            wc = 0;
            last = 0;
            continue;
        }
        if (L == 0) {  // an empty sentence still draws one centre window
            x = lcg(x);
            draws++;
            continue;
        }
        ++sid;
        for (int p = 0; p < L; p++) stamp[(size_t)sen[p]] = sid;
        float a_first = alpha, a_rest = alpha;
        for (int p = 0; p < L; p++) {
            if (p == 1) { check_alpha(); a_rest = alpha; }
            x = lcg(x);
            draws++;
            P->tokens.push_back(sen[p]);
            P->b.push_back((uint8_t)(x % (uint64_t)W));
        }
        P->sent_off.push_back((int64_t)P->tokens.size());
        P->sent_alpha.push_back(a_first);
        P->sent_alpha.push_back(a_rest);
        int nbat = (L + B - 1) / B;
        for (int nb = 0; nb < nbat; nb++) {
            int lo = nb * B;
            x = lcg(x);
            draws++;
            int rep_b = (lo <= W - 1 || (L - 1) - lo <= W - 1) ? (int)(x % (uint64_t)W) : (int)(x % (uint64_t)(2 * W));
            int K = rep_b * NEG;
            for (int k = 0; k < K; k++) {
                x = lcg(x);
                draws++;
                int32_t s = table_at((int64_t)((x >> 16) % (uint64_t)TABLE));
                if (!s) s = (int32_t)(x % (uint64_t)(V - 1) + 1);
                while (stamp[(size_t)s] == sid) {
                    x = lcg(x);
                    draws++;
                    s = (int32_t)(x % (uint64_t)(V - 1) + 1);
                }
                P->negs.push_back(s);
            }
            P->neg_off.push_back((int64_t)P->negs.size());
        }
        P->batch_off.push_back((int64_t)P->neg_off.size() - 1);
    }
    P->info.n_sent = (int64_t)P->sent_off.size() - 1;
    P->info.n_tokens = (int64_t)P->tokens.size();
    P->info.n_batches = (int64_t)P->neg_off.size() - 1;
    P->info.n_negs = (int64_t)P->negs.size();
    P->info.draws = draws;
    P->info.final_rng = x;
    P->info.raw_words = raw;
    P->info.final_alpha = alpha;
    return P.release();
}

int ph_plan_info_get(const ph_plan *p, ph_plan_info *info) {
    if (!p || !info) return fail(PH_EINVAL, "null argument");
    *info = p->info;
    return 0;
}

int ph_plan_copy(const ph_plan *p, int32_t *tokens, uint8_t *b, int64_t *sent_off, float *sent_alpha,
                 int64_t *batch_off, int64_t *neg_off, int32_t *negs) {
    if (!p) return fail(PH_EINVAL, "null plan");
    if (tokens) std::memcpy(tokens, p->tokens.data(), p->tokens.size() * sizeof(int32_t));
    if (b) std::memcpy(b, p->b.data(), p->b.size());
    if (sent_off) std::memcpy(sent_off, p->sent_off.data(), p->sent_off.size() * sizeof(int64_t));
    if (sent_alpha) std::memcpy(sent_alpha, p->sent_alpha.data(), p->sent_alpha.size() * sizeof(float));
    if (batch_off) std::memcpy(batch_off, p->batch_off.data(), p->batch_off.size() * sizeof(int64_t));
    if (neg_off) std::memcpy(neg_off, p->neg_off.data(), p->neg_off.size() * sizeof(int64_t));
    if (negs) std::memcpy(negs, p->negs.data(), p->negs.size() * sizeof(int32_t));
    return 0;
}

void ph_plan_free(ph_plan *p) { delete p; }

// ------------------------------------------------------- reference tables
void ph_exp_table(float *t) {
    for (int i = 0; i <= 1000; i++) {
        float e = std::exp(((float)i / 1000.0f * 2.0f - 1.0f) * 6.0f);  // float exp, cpu:824
        t[i] = e / (e + 1.0f);
    }
}

void ph_init_net(int64_t V, int32_t d, float *w_in, float *w_out) {
    uint64_t x = 1;  // cpu:352
    int64_t n = V * (int64_t)d;
    if (w_out) std::memset(w_out, 0, (size_t)n * sizeof(float));
    if (!w_in) return;
    const float fd = (float)d;
    for (int64_t i = 0; i < n; i++) {
        x = lcg(x);
        w_in[i] = (((float)(x & 0xFFFF) / 65536.f) - 0.5f) / fd;
    }
}

// --------------------------------------------------------- synthetic data
int64_t ph_synth_zipf(int64_t n_tokens, int32_t raw_vocab, double zipf_s, int32_t line_len,
                      double mu, double sigma, int32_t min_len, int32_t max_len,
                      int32_t n_clusters, int32_t cluster_size, int32_t band_lo, int32_t band_hi,
                      int32_t span, double p_replace, uint64_t seed,
                      int32_t *out, int64_t cap, int32_t *cluster_words) {
    if (n_tokens < 1 || raw_vocab < 2 || !out) return fail(PH_EINVAL, "bad zipf arguments");
    int64_t n_cw = (int64_t)n_clusters * cluster_size;
    if (n_cw > 0 && (band_lo < 1 || band_hi > raw_vocab + 1 || band_hi - band_lo < n_cw))
        return fail(PH_EINVAL, "cluster band too small");
    // 1. line lengths
    Rng lr(mix64(seed ^ 0x4C494E45ULL));
    std::vector<int32_t> lens;
    int64_t tot = 0;
    while (tot < n_tokens) {
        int64_t L;
        if (sigma <= 0) L = line_len;
        else {
            L = (int64_t)std::llround(std::exp(mu + sigma * lr.normal()));
            L = std::min<int64_t>(std::max<int64_t>(L, min_len), max_len);
        }
        L = std::min<int64_t>(L, n_tokens - tot);
        lens.push_back((int32_t)L);
        tot += L;
    }
    int64_t nlines = (int64_t)lens.size();
    if (n_tokens + nlines > cap) return fail(PH_ERANGE, "zipf output capacity too small");
    std::vector<int64_t> off((size_t)nlines + 1, 0);
    for (int64_t l = 0; l < nlines; l++) off[(size_t)l + 1] = off[(size_t)l] + lens[(size_t)l] + 1;
    // 2. clusters from the rank band
    std::vector<int32_t> cl_of((size_t)raw_vocab + 1, -1);
    if (n_cw > 0) {
        Rng cr(mix64(seed ^ 0x434C5553ULL));
        std::vector<int32_t> band;
        for (int32_t r = band_lo; r < band_hi; r++) band.push_back(r);
        for (int64_t k = 0; k < n_cw; k++) {
            int64_t j = k + (int64_t)cr.below((uint64_t)((int64_t)band.size() - k));
            std::swap(band[(size_t)k], band[(size_t)j]);
            int32_t w = band[(size_t)k];
            if (cluster_words) cluster_words[k] = w;
            cl_of[(size_t)w] = (int32_t)(k / cluster_size);
        }
        if (!cluster_words) return fail(PH_EINVAL, "cluster_words required");
    }
    // 3. Zipf alias over ranks 1..raw_vocab
    std::vector<double> wt((size_t)raw_vocab);
    for (int32_t r = 1; r <= raw_vocab; r++) wt[(size_t)r - 1] = std::pow((double)r, -zipf_s);
    Alias al;
    al.build(wt);
    // 4. lines in parallel, one RNG stream per line
#pragma omp parallel
    {
        std::vector<int32_t> orig;
#pragma omp for schedule(dynamic, 256)
        for (int64_t l = 0; l < nlines; l++) {
            Rng r(mix64(seed ^ mix64((uint64_t)l + 0x5A49ULL)));
            int32_t L = lens[(size_t)l];
            int32_t *dst = out + off[(size_t)l];
            orig.resize((size_t)L);
            for (int32_t i = 0; i < L; i++) orig[(size_t)i] = dst[i] = al.sample(r) + 1;
            if (n_cw > 0 && p_replace > 0)
                for (int32_t i = 0; i < L; i++) {
                    int32_t c = cl_of[(size_t)orig[(size_t)i]];
                    if (c < 0) continue;
                    for (int32_t o = -span; o <= span; o++) {
                        int32_t j = i + o;
                        if (o == 0 || j < 0 || j >= L) continue;
                        if (r.uniform() < p_replace)
                            dst[j] = cluster_words[(int64_t)c * cluster_size + (int64_t)r.below((uint64_t)cluster_size)];
                    }
                }
            dst[L] = 0;
        }
    }
    return off[(size_t)nlines];
}

namespace {
// The degree-corrected SBM of ph_synth_sbm_walks / ph_synth_sbm_graph as an
// adjacency list (ro: CSR offsets, adj: both ends of every edge).
struct SbmGraph {
    std::vector<int64_t> ro;
    std::vector<int32_t> adj;
    int64_t intra = 0;  // edges inside a community
};
SbmGraph sbm_build(int32_t N, int32_t C, double expo, double mean_deg, double p_intra, uint64_t seed,
                   int32_t *comm) {
    Rng r(mix64(seed ^ 0x53424DULL));
    // target degrees: Pareto(expo) scaled to the requested mean
    std::vector<double> deg((size_t)N);
    double sum = 0;
    for (int32_t u = 0; u < N; u++) {
        double k = std::pow(1.0 - r.uniform(), -1.0 / (expo - 1.0));
        k = std::min(k, (double)N / 4);
        deg[(size_t)u] = k;
        sum += k;
    }
    for (double &k : deg) k *= mean_deg * N / sum;
    std::vector<std::vector<double>> cw((size_t)C);
    std::vector<std::vector<int32_t>> members((size_t)C);
    for (int32_t u = 0; u < N; u++) {
        members[(size_t)(u % C)].push_back(u);
        cw[(size_t)(u % C)].push_back(deg[(size_t)u]);
        if (comm) comm[u] = u % C;
    }
    Alias all;
    all.build(deg);
    std::vector<Alias> per((size_t)C);
    for (int32_t c = 0; c < C; c++) per[(size_t)c].build(cw[(size_t)c]);
    int64_t M = (int64_t)std::llround(N * mean_deg / 2.0);
    std::vector<std::pair<int32_t, int32_t>> edges;
    edges.reserve((size_t)M);
    SbmGraph g;
    while ((int64_t)edges.size() < M) {
        int32_t u = all.sample(r), v;
        int32_t cu = u % C;
        if (r.uniform() < p_intra) v = members[(size_t)cu][(size_t)per[(size_t)cu].sample(r)];
        else {
            do { v = all.sample(r); } while (v % C == cu && C > 1);
        }
        if (u == v) continue;
        edges.emplace_back(u, v);
        g.intra += v % C == cu;
    }
    g.ro.assign((size_t)N + 1, 0);
    for (auto &e : edges) { g.ro[(size_t)e.first + 1]++; g.ro[(size_t)e.second + 1]++; }
    for (int32_t u = 0; u < N; u++) g.ro[(size_t)u + 1] += g.ro[(size_t)u];
    g.adj.resize((size_t)g.ro[(size_t)N]);
    std::vector<int64_t> fill(g.ro.begin(), g.ro.end() - 1);
    for (auto &e : edges) {
        g.adj[(size_t)fill[(size_t)e.first]++] = e.second;
        g.adj[(size_t)fill[(size_t)e.second]++] = e.first;
    }
    return g;
}
}  // namespace

int64_t ph_synth_sbm_graph(int32_t N, int32_t C, double expo, double mean_deg, double p_intra, uint64_t seed,
                           int64_t *degree, int64_t *intra_edges) {
    if (N < 2 || C < 1 || C > N || expo <= 2.0 || mean_deg <= 0 || !degree || !intra_edges)
        return fail(PH_EINVAL, "bad sbm arguments");
    SbmGraph g = sbm_build(N, C, expo, mean_deg, p_intra, seed, nullptr);
    for (int32_t u = 0; u < N; u++) degree[u] = g.ro[(size_t)u + 1] - g.ro[(size_t)u];
    *intra_edges = g.intra;
    return g.ro[(size_t)N] / 2;
}

int64_t ph_synth_sbm_walks(int32_t N, int32_t C, double expo, double mean_deg, double p_intra,
                           int32_t walks_per_node, int32_t walk_len, uint64_t seed,
                           int32_t *out, int64_t cap, int32_t *comm) {
    if (N < 2 || C < 1 || C > N || expo <= 2.0 || mean_deg <= 0 || walk_len < 1 || walks_per_node < 1 || !out)
        return fail(PH_EINVAL, "bad sbm arguments");
    SbmGraph g = sbm_build(N, C, expo, mean_deg, p_intra, seed, comm);
    const std::vector<int64_t> &ro = g.ro;
    const std::vector<int32_t> &adj = g.adj;
    int64_t nwalks = (int64_t)N * walks_per_node;
    if (nwalks * (walk_len + 1) > cap) return fail(PH_ERANGE, "walk output capacity too small");
    std::vector<int32_t> wl((size_t)nwalks);
#pragma omp parallel for schedule(static)
    for (int64_t w = 0; w < nwalks; w++) {
        Rng wr(mix64(seed ^ mix64((uint64_t)w + 0x57414CULL)));
        int32_t u = (int32_t)(w % N);
        int32_t *dst = out + w * (walk_len + 1);
        int32_t len = 0;
        dst[len++] = u + 1;
        while (len < walk_len) {
            int64_t a = ro[(size_t)u], b = ro[(size_t)u + 1];
            if (b == a) break;
            u = adj[(size_t)(a + (int64_t)wr.below((uint64_t)(b - a)))];
            dst[len++] = u + 1;
        }
        wl[(size_t)w] = len;
    }
    // compact walks (isolated nodes end early)
    int64_t o = 0;
    for (int64_t w = 0; w < nwalks; w++) {
        const int32_t *src = out + w * (walk_len + 1);
        int32_t len = wl[(size_t)w];
        if (o != w * (walk_len + 1)) std::memmove(out + o, src, (size_t)len * sizeof(int32_t));
        o += len;
        out[o++] = 0;
    }
    return o;
}

// ------------------------------------------------------- binary corpus --
// Layout (little-endian): a 64-byte header, counts[V] uint32, the V words
// NUL-terminated, zero padding to a multiple of 8, ids[n_ids] int32.
namespace {
constexpr char kBinMagic[8] = {'P', 'S', 'G', 'N', 'S', 'B', 'C', '1'};
struct BinHeader {
    char magic[8];
    uint32_t header_bytes;   // 64
    int32_t vocab_size;
    uint64_t train_words;
    int64_t n_ids;
    uint64_t words_bytes;    // the NUL-terminated words, without padding
    uint64_t ids_offset;     // byte offset of ids[0], a multiple of 8
    uint64_t reserved[2];
};
static_assert(sizeof(BinHeader) == 64, "binary corpus header is 64 bytes");

int read_bin_header(FILE *f, BinHeader &h) {
    if (std::fread(&h, sizeof h, 1, f) != 1) return fail(PH_EINVAL, "binary corpus: truncated header");
    if (std::memcmp(h.magic, kBinMagic, 8) != 0 || h.header_bytes != sizeof(BinHeader))
        return fail(PH_EINVAL, "binary corpus: bad magic (not a par_word2vec binary corpus)");
    if (h.vocab_size < 1 || h.n_ids < 0 || h.ids_offset % 8 != 0 ||
        h.ids_offset < sizeof(BinHeader) + 4 * (uint64_t)h.vocab_size + h.words_bytes)
        return fail(PH_EINVAL, "binary corpus: inconsistent header");
    return 0;
}
}  // namespace

int ph_corpus_save(const char *path, const ph_vocab *v, const int32_t *stream, int64_t n) {
    if (!path || !v || (!stream && n > 0) || n < 0) return fail(PH_EINVAL, "null argument");
    const int32_t V = (int32_t)v->words.size();
    for (int64_t i = 0; i < n; i++)
        if (stream[i] < 0 || stream[i] >= V) return fail(PH_EINVAL, "stream id out of vocabulary range");
    BinHeader h{};
    std::memcpy(h.magic, kBinMagic, 8);
    h.header_bytes = sizeof(BinHeader);
    h.vocab_size = V;
    h.train_words = v->train_words;
    h.n_ids = n;
    for (const std::string &w : v->words) h.words_bytes += w.size() + 1;
    h.ids_offset = (sizeof(BinHeader) + 4 * (uint64_t)V + h.words_bytes + 7) / 8 * 8;
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(PH_EIO, std::string("cannot create ") + path);
    bool ok = std::fwrite(&h, sizeof h, 1, f) == 1 &&
              std::fwrite(v->cn.data(), 4, (size_t)V, f) == (size_t)V;
    for (size_t i = 0; ok && i < v->words.size(); i++)
        ok = std::fwrite(v->words[i].c_str(), 1, v->words[i].size() + 1, f) == v->words[i].size() + 1;
    static const char zeros[8] = {0};
    size_t pad = (size_t)(h.ids_offset - sizeof(BinHeader) - 4 * (uint64_t)V - h.words_bytes);
    ok = ok && std::fwrite(zeros, 1, pad, f) == pad;
    ok = ok && (n == 0 || std::fwrite(stream, 4, (size_t)n, f) == (size_t)n);
    ok = (std::fclose(f) == 0) && ok;
    return ok ? 0 : fail(PH_EIO, std::string("write failed: ") + path);
}

ph_vocab *ph_corpus_vocab(const char *path, int64_t *n_ids) {
    FILE *f = path ? std::fopen(path, "rb") : nullptr;
    if (!f) { fail(PH_EIO, std::string("cannot open ") + (path ? path : "(null)")); return nullptr; }
    BinHeader h{};
    if (read_bin_header(f, h) < 0) { std::fclose(f); return nullptr; }
    std::unique_ptr<ph_vocab> v(new (std::nothrow) ph_vocab);
    std::vector<char> words((size_t)h.words_bytes);
    if (v) v->cn.resize((size_t)h.vocab_size);
    bool ok = v && std::fread(v->cn.data(), 4, (size_t)h.vocab_size, f) == (size_t)h.vocab_size &&
              std::fread(words.data(), 1, words.size(), f) == words.size();
    std::fclose(f);
    if (!ok) { fail(PH_EINVAL, "binary corpus: truncated vocabulary"); return nullptr; }
    uint64_t total = 0;
    size_t p = 0;
    for (int32_t i = 0; i < h.vocab_size; i++) {
        const char *s = words.data() + p;
        size_t len = strnlen(s, words.size() - p);
        if (p + len >= words.size() || len == 0 || len > 98) {
            fail(PH_EINVAL, "binary corpus: malformed word table");
            return nullptr;
        }
        v->words.emplace_back(s, len);
        p += len + 1;
        total += v->cn[(size_t)i];
    }
    if (p != words.size() || total != h.train_words) {
        fail(PH_EINVAL, "binary corpus: word table or counts do not match the header");
        return nullptr;
    }
    v->train_words = h.train_words;
    v->rebuild_map();
    if (n_ids) *n_ids = h.n_ids;
    return v.release();
}

int64_t ph_corpus_ids(const char *path, int32_t *out, int64_t cap) {
    FILE *f = path ? std::fopen(path, "rb") : nullptr;
    if (!f) return fail(PH_EIO, std::string("cannot open ") + (path ? path : "(null)"));
    BinHeader h{};
    if (read_bin_header(f, h) < 0) { std::fclose(f); return PH_EINVAL; }
    if (!out) { std::fclose(f); return h.n_ids; }
    if (h.n_ids > cap) { std::fclose(f); return fail(PH_ERANGE, "stream capacity too small"); }
    bool ok = std::fseek(f, (long)h.ids_offset, SEEK_SET) == 0 &&
              std::fread(out, 4, (size_t)h.n_ids, f) == (size_t)h.n_ids;
    std::fclose(f);
    if (!ok) return fail(PH_EINVAL, "binary corpus: truncated id stream");
    std::atomic<int> bad(0);
    const int32_t V = h.vocab_size;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < h.n_ids; i++)
        if (out[i] < 0 || out[i] >= V) bad = 1;
    if (bad) return fail(PH_EINVAL, "binary corpus: id out of vocabulary range");
    return h.n_ids;
}

int ph_write_raw_text(const char *path, const int32_t *raw, int64_t n) {
    if (!path || !raw || n < 0) return fail(PH_EINVAL, "null argument");
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(PH_EIO, "cannot create corpus file");
    std::string buf;
    buf.reserve(1 << 22);
    bool first = true;
    char tmp[24];
    for (int64_t i = 0; i < n; i++) {
        if (raw[i] == 0) { buf.push_back('\n'); first = true; }
        else {
            if (!first) buf.push_back(' ');
            int k = std::snprintf(tmp, sizeof tmp, "w%d", raw[i]);
            buf.append(tmp, (size_t)k);
            first = false;
        }
        if (buf.size() > (1 << 22) - 64) { std::fwrite(buf.data(), 1, buf.size(), f); buf.clear(); }
    }
    std::fwrite(buf.data(), 1, buf.size(), f);
    std::fclose(f);
    return 0;
}

int ph_save_embeddings(const char *path, const ph_vocab *v, const float *w_in, int32_t d, int32_t binary) {
    if (!path || !v || !w_in || d < 1) return fail(PH_EINVAL, "null argument");
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(PH_EIO, "cannot create output file");
    int64_t V = (int64_t)v->words.size();
    std::fprintf(f, "%d %d\n", (int)V, d);
    const int64_t CH = 4096;
    for (int64_t lo = 0; lo < V; lo += CH * 16) {
        int64_t hi = std::min(V, lo + CH * 16);
        int64_t nch = (hi - lo + CH - 1) / CH;
        std::vector<std::string> parts((size_t)nch);
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t c = 0; c < nch; c++) {
            std::string &s = parts[(size_t)c];
            char tmp[64];
            for (int64_t a = lo + c * CH; a < std::min(hi, lo + (c + 1) * CH); a++) {
                s += v->words[(size_t)a];
                s.push_back(' ');
                const float *row = w_in + a * d;
                if (binary) s.append((const char *)row, (size_t)d * sizeof(float));
                else
                    for (int32_t b = 0; b < d; b++) {
                        int k = std::snprintf(tmp, sizeof tmp, "%f ", row[b]);
                        s.append(tmp, (size_t)k);
                    }
                s.push_back('\n');
            }
        }
        for (auto &s : parts) std::fwrite(s.data(), 1, s.size(), f);
    }
    std::fclose(f);
    return 0;
}

}  // extern "C"
