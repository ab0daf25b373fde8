/*
 * par_sgns_oracle.c -- CPU restatement of PAR_Word2Vec_cpu.cpp.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker or
 * the CPU baseline.  The product (par_word2vec_amd) never links or calls it.
 *
 * Build (oracle/Makefile): gcc -O2 -ffp-contract=off -fno-fast-math, no -march,
 * so x86-64 SSE2 float arithmetic rounds every operation exactly as written.
 * A second flavour (-O3 -mavx2 -mfma -fopenmp) is built for throughput timing
 * only; it is never used for parity.
 *
 * Reference: PAR_Word2Vec_cpu.cpp ("cpu:" below).  The reference is not
 * buildable in this image (it includes "mkl.h" unconditionally at cpu:27 and
 * the MKL headers are absent), so this restatement is pinned by known answers,
 * closed forms and an independent transcription instead (see the header).
 */
#include "par_sgns_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_HASH_SIZE 30000000 /* cpu:74 */

/* ---------------------------------------------------- planted misreadings */

/* 0 in every real use.  The pin tests set one ORC_MUT_* value at a time to
 * show that the oracle's known-answer checks fail under that misreading. */
static int g_mutation = 0;
void orc_set_mutation(int32_t m) { g_mutation = m; }
int32_t orc_mutation(void) { return g_mutation; }
#define MUT(k) (g_mutation == (k))

/* ---------------------------------------------------------------- LCG ---- */

/* next_random = next_random * 25214903917 + 11 (mod 2^64), cpu:461, 532, 591 */
uint64_t orc_lcg(uint64_t x) { return x * (uint64_t)25214903917ULL + 11ULL; }

/* EXP_RESOLUTION = EXP_TABLE_SIZE / (MAX_EXP * 2.0f), cpu:70 */
float orc_exp_resolution(void) { return (float)ORC_EXP_TABLE_SIZE / ((float)ORC_MAX_EXP * 2.0f); }

/* cpu:822-826: x = (i/1000f*2-1)*6 in float, exp in float (std::exp(float)),
 * then e/(e+1) in float. */
void orc_exp_table(float *t) {
    for (int i = 0; i < ORC_EXP_TABLE_SIZE + 1; i++) {
        float x = ((float)i / (float)ORC_EXP_TABLE_SIZE * 2.0f - 1.0f) * (float)ORC_MAX_EXP;
        float e = expf(x);
        t[i] = e / (e + 1.0f);
    }
}

/* cpu:337-357: W_out = 0; W_in[i] = (((x & 0xFFFF)/65536f) - 0.5f)/d with the
 * LCG started at 1 and advanced before every element, row-major over V*d. */
void orc_init_net(int64_t vocab_size, int32_t dim, float *w_in, float *w_out) {
    uint64_t x = 1;
    int64_t n = vocab_size * (int64_t)dim;
    if (w_out) memset(w_out, 0, (size_t)n * sizeof(float));
    for (int64_t i = 0; i < n; i++) {
        x = orc_lcg(x);
        w_in[i] = (((float)(x & 0xFFFF) / 65536.0f) - 0.5f) / (float)dim;
    }
}

/* cpu:96-117.  Z = sum cn^0.75 in double (sequential with one thread);
 * d1 is a float that accumulates in double and is rounded back each step; the
 * comparison uses a/(float)table_size in float.  Once i runs past the last
 * word the reference reads vocab[vocab_size] (out of bounds) before clamping;
 * that read never influences the table, so it is skipped here. */
int orc_unigram_table_n(const uint32_t *cn, int32_t vocab_size, int32_t *table, int64_t table_size) {
    if (vocab_size <= 0) return -1;
    const float power = MUT(ORC_MUT_UNIGRAM_LINEAR) ? 1.0f : 0.75f;
    double z = 0.0;
    for (int32_t i = 0; i < vocab_size; i++) z += pow((double)cn[i], (double)power);
    int32_t i = 0;
    float d1 = (float)(pow((double)cn[0], (double)power) / z);
    for (int64_t a = 0; a < table_size; a++) {
        table[a] = i;
        if ((float)a / (float)table_size > d1) {
            i++;
            if (i < vocab_size) d1 = (float)((double)d1 + pow((double)cn[i], (double)power) / z);
        }
        if (i >= vocab_size) i = vocab_size - 1;
    }
    return 0;
}

int orc_unigram_table(const uint32_t *cn, int32_t vocab_size, int32_t *table) {
    return orc_unigram_table_n(cn, vocab_size, table, ORC_TABLE_SIZE);
}

/* -------------------------------------------------------- vocabulary ---- */

struct orc_vocab {
    int32_t size, cap;
    uint32_t *cn;
    char **word;
    int32_t *hash;
    uint64_t train_words;
    uint64_t file_size;
    int32_t min_reduce;
};

/* cpu:120-144: '\r' skipped; space/tab/newline delimit; a newline seen with an
 * empty buffer yields "</s>", otherwise it is pushed back; words are cut to
 * MAX_STRING-1 bytes (the last byte keeps being overwritten). */
static void orc_read_word(char *word, FILE *fin) {
    int a = 0, ch;
    while (!feof(fin)) {
        ch = fgetc(fin);
        if (ch == 13) continue;
        if (ch == ' ' || ch == '\t' || ch == '\n') {
            if (a > 0) {
                if (ch == '\n') ungetc(ch, fin);
                break;
            }
            if (ch == '\n') {
                strcpy(word, "</s>");
                return;
            }
            continue;
        }
        word[a] = (char)ch;
        a++;
        if (a >= ORC_MAX_STRING - 1) a--;
    }
    word[a] = 0;
}

/* cpu:147-153: hash = hash*257 + (signed char), uint32 wrap, mod table size */
static int32_t orc_hash(const char *w) {
    uint32_t h = 0;
    size_t n = strlen(w);
    for (size_t i = 0; i < n; i++) h = h * 257u + (uint32_t)(int32_t)(signed char)w[i];
    return (int32_t)(h % ORC_HASH_SIZE);
}

int32_t orc_vocab_search(const orc_vocab *v, const char *word) { /* cpu:156-166 */
    int32_t h = orc_hash(word);
    for (;;) {
        int32_t k = v->hash[h];
        if (k == -1) return -1;
        if (!strcmp(word, v->word[k])) return k;
        h = (h + 1) % ORC_HASH_SIZE;
    }
}

static int32_t orc_vocab_add(orc_vocab *v, const char *word) { /* cpu:178-196 */
    size_t len = strlen(word) + 1;
    if (len > ORC_MAX_STRING) len = ORC_MAX_STRING;
    v->word[v->size] = (char *)calloc(len, 1);
    memcpy(v->word[v->size], word, len - 1);
    v->cn[v->size] = 0;
    v->size++;
    if (v->size + 2 >= v->cap) {
        v->cap += 1000;
        v->cn = (uint32_t *)realloc(v->cn, (size_t)v->cap * sizeof(uint32_t));
        v->word = (char **)realloc(v->word, (size_t)v->cap * sizeof(char *));
    }
    int32_t h = orc_hash(word);
    while (v->hash[h] != -1) h = (h + 1) % ORC_HASH_SIZE;
    v->hash[h] = v->size - 1;
    return v->size - 1;
}

static void orc_rehash(orc_vocab *v) {
    for (int64_t i = 0; i < ORC_HASH_SIZE; i++) v->hash[i] = -1;
    for (int32_t i = 0; i < v->size; i++) {
        int32_t h = orc_hash(v->word[i]);
        while (v->hash[h] != -1) h = (h + 1) % ORC_HASH_SIZE;
        v->hash[h] = i;
    }
}

/* Stable merge sort of entries [lo, hi) by count descending.  glibc 2.35
 * qsort (used at cpu:206) is a merge sort and hence stable, so ties keep
 * first-appearance order. */
static void orc_sort_desc(uint32_t *cn, char **w, int32_t n) {
    if (n < 2) return;
    uint32_t *tc = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
    char **tw = (char **)malloc((size_t)n * sizeof(char *));
    for (int32_t width = 1; width < n; width *= 2) {
        for (int32_t lo = 0; lo < n; lo += 2 * width) {
            int32_t mid = lo + width < n ? lo + width : n;
            int32_t hi = lo + 2 * width < n ? lo + 2 * width : n;
            int32_t a = lo, b = mid, o = lo;
            while (a < mid && b < hi) {
                if (cn[b] > cn[a]) { tc[o] = cn[b]; tw[o++] = w[b++]; }
                else { tc[o] = cn[a]; tw[o++] = w[a++]; }
            }
            while (a < mid) { tc[o] = cn[a]; tw[o++] = w[a++]; }
            while (b < hi) { tc[o] = cn[b]; tw[o++] = w[b++]; }
        }
        memcpy(cn, tc, (size_t)n * sizeof(uint32_t));
        memcpy(w, tw, (size_t)n * sizeof(char *));
    }
    free(tc);
    free(tw);
}

/* cpu:204-226: sort [1, V) by count desc, drop cn < min_count (never index 0),
 * rebuild the hash, train_words = sum of the kept counts. */
static void orc_sort_vocab(orc_vocab *v, int32_t min_count) {
    orc_sort_desc(v->cn + 1, v->word + 1, v->size - 1);
    int32_t keep = 0;
    v->train_words = 0;
    for (int32_t i = 0; i < v->size; i++) {
        if (v->cn[i] < (uint32_t)min_count && i != 0) {
            free(v->word[i]);
            v->word[i] = NULL;
        } else {
            v->cn[keep] = v->cn[i];
            v->word[keep] = v->word[i];
            v->train_words += v->cn[i];
            keep++;
        }
    }
    v->size = keep;
    orc_rehash(v);
}

/* cpu:229-251 */
static void orc_reduce_vocab(orc_vocab *v) {
    int32_t count = 0;
    for (int32_t i = 0; i < v->size; i++) {
        if (v->cn[i] > (uint32_t)v->min_reduce) {
            v->cn[count] = v->cn[i];
            v->word[count] = v->word[i];
            count++;
        } else {
            free(v->word[i]);
        }
    }
    v->size = count;
    orc_rehash(v);
    v->min_reduce++;
}

static orc_vocab *orc_vocab_new(void) {
    orc_vocab *v = (orc_vocab *)calloc(1, sizeof(orc_vocab));
    v->cap = 1000;
    v->cn = (uint32_t *)calloc((size_t)v->cap, sizeof(uint32_t));
    v->word = (char **)calloc((size_t)v->cap, sizeof(char *));
    v->hash = (int32_t *)malloc((size_t)ORC_HASH_SIZE * sizeof(int32_t));
    for (int64_t i = 0; i < ORC_HASH_SIZE; i++) v->hash[i] = -1;
    v->min_reduce = 1;
    return v;
}

/* cpu:253-292 */
orc_vocab *orc_vocab_learn(const char *path, int32_t min_count) {
    FILE *fin = fopen(path, "rb");
    if (!fin) return NULL;
    orc_vocab *v = orc_vocab_new();
    char word[ORC_MAX_STRING];
    orc_vocab_add(v, "</s>");
    for (;;) {
        orc_read_word(word, fin);
        if (feof(fin)) break;
        int32_t i = orc_vocab_search(v, word);
        if (i == -1) {
            int32_t a = orc_vocab_add(v, word);
            v->cn[a] = 1;
        } else {
            v->cn[i]++;
        }
        if (v->size > ORC_HASH_SIZE * 0.7) orc_reduce_vocab(v);
    }
    orc_sort_vocab(v, min_count);
    v->file_size = (uint64_t)ftell(fin);
    fclose(fin);
    return v;
}

/* cpu:301-335: "<word> <count>" lines, then the same sort/prune. */
orc_vocab *orc_vocab_read(const char *vocab_path, const char *train_path, int32_t min_count) {
    FILE *fin = fopen(vocab_path, "rb");
    if (!fin) return NULL;
    orc_vocab *v = orc_vocab_new();
    char word[ORC_MAX_STRING];
    char c;
    for (;;) {
        orc_read_word(word, fin);
        if (feof(fin)) break;
        int32_t i = orc_vocab_add(v, word);
        int cnt = 0;
        if (fscanf(fin, "%d%c", &cnt, &c) < 1) cnt = 0;
        v->cn[i] = (uint32_t)cnt;
    }
    fclose(fin);
    orc_sort_vocab(v, min_count);
    FILE *f2 = fopen(train_path, "rb");
    if (f2) {
        fseek(f2, 0, SEEK_END);
        v->file_size = (uint64_t)ftell(f2);
        fclose(f2);
    }
    return v;
}

int orc_vocab_save(const orc_vocab *v, const char *path) { /* cpu:294-299 */
    FILE *fo = fopen(path, "wb");
    if (!fo) return -1;
    for (int32_t i = 0; i < v->size; i++) fprintf(fo, "%s %d\n", v->word[i], (int)v->cn[i]);
    fclose(fo);
    return 0;
}

int32_t orc_vocab_size(const orc_vocab *v) { return v->size; }
uint64_t orc_vocab_train_words(const orc_vocab *v) { return v->train_words; }
uint64_t orc_vocab_file_size(const orc_vocab *v) { return v->file_size; }
uint32_t orc_vocab_count(const orc_vocab *v, int32_t i) { return v->cn[i]; }
const char *orc_vocab_word(const orc_vocab *v, int32_t i) { return v->word[i]; }
void orc_vocab_counts(const orc_vocab *v, uint32_t *out) { memcpy(out, v->cn, (size_t)v->size * sizeof(uint32_t)); }

void orc_vocab_free(orc_vocab *v) {
    if (!v) return;
    for (int32_t i = 0; i < v->size; i++) free(v->word[i]);
    free(v->word);
    free(v->cn);
    free(v->hash);
    free(v);
}

/* cpu:169-175 + cpu:359-370 */
int64_t orc_load_stream(const orc_vocab *v, const char *path, uint64_t offset,
                        uint64_t max_words, int32_t *stream) {
    FILE *fin = fopen(path, "rb");
    if (!fin) return -1;
    fseek(fin, (long)offset, SEEK_SET);
    char word[ORC_MAX_STRING];
    uint64_t n = 0;
    while (!feof(fin) && n < max_words) {
        orc_read_word(word, fin);
        if (feof(fin)) continue; /* ReadWordIndex returns -1 at EOF */
        int32_t w = orc_vocab_search(v, word);
        if (w == -1) continue;
        stream[n++] = w;
    }
    stream[n] = 0;
    fclose(fin);
    return (int64_t)n;
}

/* ------------------------------------------------------------- GEMMs ---- */
/* Row-major, alpha = 1, beta = 0, ascending reduction index, acc from 0.0f. */

/* C[m x n] = A[m x k] * B[n x k]^T   (cpu:633 shape: NoTrans, Trans) */
void orc_gemm_nt(int m, int n, int k, const float *a, int lda, const float *b, int ldb, float *c, int ldc) {
    for (int r = 0; r < m; r++)
        for (int s = 0; s < n; s++) {
            float acc = 0.0f;
            for (int p = 0; p < k; p++) acc += a[(int64_t)r * lda + p] * b[(int64_t)s * ldb + p];
            c[(int64_t)r * ldc + s] = acc;
        }
}

/* C[m x n] = A[m x k] * B[k x n]   (cpu:652 shape: NoTrans, NoTrans) */
void orc_gemm_nn(int m, int n, int k, const float *a, int lda, const float *b, int ldb, float *c, int ldc) {
    for (int r = 0; r < m; r++)
        for (int s = 0; s < n; s++) {
            float acc = 0.0f;
            for (int p = 0; p < k; p++) acc += a[(int64_t)r * lda + p] * b[(int64_t)p * ldb + s];
            c[(int64_t)r * ldc + s] = acc;
        }
}

/* C[m x n] = A[k x m]^T * B[k x n]   (cpu:655 shape: Trans, NoTrans) */
void orc_gemm_tn(int m, int n, int k, const float *a, int lda, const float *b, int ldb, float *c, int ldc) {
    for (int r = 0; r < m; r++)
        for (int s = 0; s < n; s++) {
            float acc = 0.0f;
            for (int p = 0; p < k; p++) acc += a[(int64_t)p * lda + r] * b[(int64_t)p * ldb + s];
            c[(int64_t)r * ldc + s] = acc;
        }
}

/* ------------------------------------------------------------ trainer ---- */

static inline double softplus(double x) { /* log(1 + e^x), stable */
    return x > 0 ? x + log1p(exp(-x)) : log1p(exp(x));
}

typedef struct orc_thread {
    const orc_params *p;
    const int32_t *stream;
    uint64_t local_words;
    const uint32_t *cn;
    int32_t V;
    uint64_t T;
    const int32_t *table;
    const float *et;
    float *w_in, *w_out;
    uint64_t rng;
    orc_stats st;
    orc_trace *tr;
    orc_curve *cv;
    /* shared progress (Hogwild) */
    uint64_t *wc_actual;
    float *alpha_shared;
    int shared;
} orc_thread;

static inline void trace_sentence(orc_trace *tr, const int32_t *sen, int L, float a0) {
    if (!tr || tr->overflow) return;
    if (tr->n_sent >= tr->cap_sent || tr->n_tokens + L > tr->cap_tokens) { tr->overflow = 1; return; }
    tr->sent_len[tr->n_sent] = L;
    tr->sent_alpha[2 * tr->n_sent] = a0;
    tr->sent_alpha[2 * tr->n_sent + 1] = a0;
    memcpy(tr->tokens + tr->n_tokens, sen, (size_t)L * sizeof(int32_t));
    tr->n_sent++;
}

/* Test-only replacement of the LCG (orc_params.strong_rng): the SplitMix64
 * finalizer (Steele, Lea, Flood, "Fast splittable pseudorandom number
 * generators", OOPSLA 2014) of x + the golden-ratio increment -- every bit of
 * a draw is well mixed, so x % 2W has no parity pattern. */
static inline uint64_t strong_step(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* bf16 emulation (orc_params.bf16): stochastic rounding of a row to bf16
 * (16 uniform bits added below the kept mantissa, then truncation). */
static void round_row_bf16(float *row, int d, uint32_t *rs) {
    for (int j = 0; j < d; j++) {
        uint32_t x = *rs;
        x ^= x << 13;
        x ^= x >> 17;
        x ^= x << 5;
        *rs = x;
        uint32_t b;
        memcpy(&b, &row[j], 4);
        b = (b + (x & 0xFFFFu)) & 0xFFFF0000u;
        memcpy(&row[j], &b, 4);
    }
}

/* The body of the per-thread loop at cpu:397-700.  With one thread and the
 * thread-0 seed this is the reference's deterministic schedule. */
static void orc_thread_run(orc_thread *t) {
    const orc_params *p = t->p;
    const int d = p->dim, W = p->window, NEG = p->negative, B = p->batch_size;
    const float EXP_RES = orc_exp_resolution();
    const float starting_alpha = p->alpha;
    float alpha = starting_alpha;
    int local_iter = p->iter;
    uint64_t x = t->rng;
    uint64_t word_count = 0, last_word_count = 0, epoch_base = 0;
    int sentence_length = 0, sentence_position = 0, rep_sentence_position = 0;
    int sen[ORC_MAX_SENTENCE];
    memset(sen, 0, sizeof(sen));
    int64_t sent_trace_idx = -1;
    uint64_t sent_key = 0; /* raw offset used to bucket the loss curve */
    float *M_in = (float *)malloc((size_t)B * d * sizeof(float));
    float *M_out = (float *)malloc((size_t)(W * 2 * NEG + 1) * d * sizeof(float));
    float *M_out_upd = (float *)malloc((size_t)(W * 2 * NEG + 1) * d * sizeof(float));
    float *M_grad = (float *)malloc((size_t)(W * 2 * NEG + 1) * B * sizeof(float));
    int32_t *shared_negative = (int32_t *)malloc((size_t)(W * 2 * NEG + 1) * sizeof(int32_t));
    const float *et = t->et;
    float *w_in = t->w_in, *w_out = t->w_out;
    const uint64_t denom_words = (uint64_t)p->iter * t->T + 1;
    uint32_t rs = (uint32_t)(p->seed * 2654435761u) ^ 0x9E3779B9u;  /* bf16 emulation noise */
    if (!rs) rs = 1;

    for (;;) {
        /* cpu:431-446 learning-rate update, checked before loading */
        if (!MUT(ORC_MUT_ALPHA_AFTER_LOAD) &&
            (word_count - last_word_count > 10000 || (MUT(ORC_MUT_ALPHA_GE) && word_count - last_word_count >= 10000))) {
            uint64_t diff = word_count - last_word_count;
            uint64_t wca;
            if (t->shared) {
#pragma omp atomic capture
                { *t->wc_actual += diff; wca = *t->wc_actual; }
            } else {
                *t->wc_actual += diff;
                wca = *t->wc_actual;
            }
            last_word_count = word_count;
            alpha = starting_alpha * (1 - (float)wca / (float)denom_words);
            if (alpha < starting_alpha * 0.0001f) alpha = starting_alpha * 0.0001f;
        }
        /* cpu:447-471 sentence build with subsampling */
        if (sentence_length == 0) {
            sent_key = epoch_base + word_count;
            if (p->alpha_per_sentence) { /* test-only: the rate of the words before this sentence */
                alpha = starting_alpha * (1 - (float)(epoch_base + word_count) / (float)denom_words);
                if (alpha < starting_alpha * 0.0001f) alpha = starting_alpha * 0.0001f;
            }
            for (;;) {
                int w = t->stream[word_count];
                word_count++;
                if (w == 0) break;
                if (p->sample > 0) {
                    float ratio = (p->sample * (float)t->T) / (float)t->cn[w];
                    float ran = sqrtf(ratio) + ratio;
                    x = p->strong_rng ? strong_step(x) : orc_lcg(x);
                    t->st.draws++;
                    if (MUT(ORC_MUT_SUBSAMPLE_FLIP) ? !(ran < (float)(x & 0xFFFF) / 65536.f)
                                                    : ran < (float)(x & 0xFFFF) / 65536.f)
                        continue;
                }
                sen[sentence_length] = w;
                sentence_length++;
                if (sentence_length >= ORC_MAX_SENTENCE) break;
            }
            sentence_position = 0;
            rep_sentence_position = 0;
            sent_trace_idx = -1;
        }
        if (MUT(ORC_MUT_ALPHA_AFTER_LOAD) && word_count - last_word_count > 10000) {
            *t->wc_actual += word_count - last_word_count;
            last_word_count = word_count;
            alpha = starting_alpha * (1 - (float)*t->wc_actual / (float)denom_words);
            if (alpha < starting_alpha * 0.0001f) alpha = starting_alpha * 0.0001f;
        }
        /* cpu:472-527 epoch boundary: the sentence just built is dropped */
        if (word_count > t->local_words) {
            uint64_t diff = (MUT(ORC_MUT_WCA_LOCAL) ? t->local_words : word_count) - last_word_count;
            if (t->shared) {
#pragma omp atomic
                *t->wc_actual += diff;
            } else {
                *t->wc_actual += diff;
            }
            t->st.raw_words += t->local_words;
            local_iter--;
            if (local_iter == 0) break;
            epoch_base += t->local_words;
            word_count = 0;
            last_word_count = 0;
            sentence_length = 0;
            continue;
        }
        if (sentence_position == 0 && sentence_length > 0) {
            t->st.sentences++;
            if (t->tr) {
                trace_sentence(t->tr, sen, sentence_length, alpha);
                sent_trace_idx = t->tr->overflow ? -2 : t->tr->n_sent - 1;
            }
        }
        if (sentence_position == 1 && sent_trace_idx >= 0) t->tr->sent_alpha[2 * sent_trace_idx + 1] = alpha;

        /* cpu:529-569 attraction for one centre */
        int target = sen[sentence_position < ORC_MAX_SENTENCE ? sentence_position : 0];
        x = p->strong_rng ? strong_step(x) : orc_lcg(x);
        t->st.draws++;
        int b = (int)(x % (uint64_t)W);
        if (sentence_length > 0) {
            t->st.trained_tokens++;
            if (t->tr && sent_trace_idx >= 0) t->tr->b[t->tr->n_tokens + sentence_position] = (uint8_t)b;
        }
        for (int i = b; i < 2 * W + 1 - b + (MUT(ORC_MUT_WINDOW_INCL) ? 1 : 0); i++) {
            if (i == W) continue;
            int c = sentence_position - W + i;
            if (c < 0) continue;
            if (c >= sentence_length) break;
            int input_word = sen[c];
            if (input_word == -1) continue;
            int64_t l1 = (int64_t)input_word * d, l2 = (int64_t)target * d;
            float f = 0.f, g;
            for (int j = 0; j < d; j++) f += w_in[l1 + j] * w_out[l2 + j];
            if (f > ORC_MAX_EXP) g = (1 - 1) * alpha;
            else if (f < -ORC_MAX_EXP) g = 1 * alpha;
            else g = (1 - et[(int)((f + ORC_MAX_EXP) * EXP_RES)]) * alpha;
            for (int j = 0; j < d; j++) {
                float tmp = g * w_out[l2 + j];
                w_out[l2 + j] += g * w_in[l1 + j];
                if (MUT(ORC_MUT_PAIR_ORDER)) tmp = g * w_out[l2 + j];
                w_in[l1 + j] += tmp;
            }
            t->st.pairs++;
            double lp = softplus(-(double)f);
            t->st.loss_attr += lp;
            if (t->cv && t->cv->n_buckets > 0) {
                int64_t k = (int64_t)(sent_key / (uint64_t)t->cv->bucket_words);
                if (k < t->cv->n_buckets) {
                    t->cv->loss[k] += lp;
                    t->cv->pairs[k] += 1;
                    if (t->cv->loss_attr) { t->cv->loss_attr[k] += lp; t->cv->pairs_attr[k] += 1; }
                }
            }
        }
        sentence_position++;
        if (p->bf16 && sentence_length > 0) {
            /* the GPU trainer's write-back points: the centre's W_out row after its
             * pairs (kept in registers while the next centre is the same word), and
             * the W_in row of the position leaving the window unless its word is
             * still referenced by the window of the next centre */
            const int pc = sentence_position - 1;
            if (!(sentence_position < sentence_length && sen[sentence_position] == target))
                round_row_bf16(w_out + (int64_t)target * d, d, &rs);
            const int r = pc - W;
            if (r >= 0) {
                int live = 0;
                for (int q = r + 1; q <= pc + 1 + W && q < sentence_length; q++) live |= sen[q] == sen[r];
                if (!live) round_row_bf16(w_in + (int64_t)sen[r] * d, d, &rs);
            }
            if (sentence_position >= sentence_length) /* flush: rounding a bf16 value changes nothing */
                for (int q = 0; q < sentence_length; q++) round_row_bf16(w_in + (int64_t)sen[q] * d, d, &rs);
        }

        /* cpu:571-679 repulsion over all batches once the sentence is done */
        if (sentence_position >= sentence_length) {
            rep_sentence_position = 0;
            if (sentence_length > 0) {
                int num_batch = (sentence_length + B - 1) / B;
                for (int n_b = 0; n_b < num_batch; n_b++) {
                    int min_pos = n_b * B;
                    int max_pos = min_pos + (B - 1) < sentence_length - 1 ? min_pos + (B - 1) : sentence_length - 1;
                    int input_size = max_pos - min_pos + 1;
                    if (input_size <= 0) break;
                    int rep_b;
                    x = p->strong_rng ? strong_step(x) : orc_lcg(x);
                    t->st.draws++;
                    const int end_gap = MUT(ORC_MUT_BATCH_END) ? sentence_length - min_pos : (sentence_length - 1) - min_pos;
                    if (MUT(ORC_MUT_REPB_2W)) rep_b = (int)(x % (uint64_t)(W * 2));
                    else if (min_pos <= W - 1) rep_b = (int)(x % (uint64_t)W);
                    else if (end_gap <= W - 1) rep_b = (int)(x % (uint64_t)W);
                    else rep_b = (int)(x % (uint64_t)(W * 2));
                    int K = rep_b * NEG;
                    for (int ng = 0; ng < K; ng++) {
                        x = p->strong_rng ? strong_step(x) : orc_lcg(x);
                        t->st.draws++;
                        int s = t->table[(MUT(ORC_MUT_TABLE_INDEX) ? x : x >> 16) % (uint64_t)ORC_TABLE_SIZE];
                        if (!s) s = (int)(x % (uint64_t)(t->V - 1) + 1);
                        for (;;) {
                            if (MUT(ORC_MUT_NO_REJECTION)) break;
                            int found = 0;
                            for (int q = 0; q < sentence_length; q++)
                                if (sen[q] == s) { found = 1; break; }
                            if (!found) break;
                            x = p->strong_rng ? strong_step(x) : orc_lcg(x);
                            t->st.draws++;
                            s = (int)(x % (uint64_t)(t->V - 1) + 1);
                        }
                        shared_negative[ng] = s;
                    }
                    if (t->tr && !t->tr->overflow) {
                        if (t->tr->n_batches >= t->tr->cap_batches || t->tr->n_negs + K > t->tr->cap_negs) {
                            t->tr->overflow = 1;
                        } else {
                            t->tr->rep_b[t->tr->n_batches++] = rep_b;
                            memcpy(t->tr->negs + t->tr->n_negs, shared_negative, (size_t)K * sizeof(int32_t));
                            t->tr->n_negs += K;
                        }
                    }
                    t->st.batches++;
                    t->st.negatives += (uint64_t)K;
                    t->st.rep_pairs += (uint64_t)K * (uint64_t)input_size;
                    for (int i = 0; i < input_size; i++)
                        memcpy(M_in + (int64_t)i * d, w_in + (int64_t)sen[min_pos + i] * d, (size_t)d * sizeof(float));
                    for (int i = 0; i < K; i++)
                        memcpy(M_out + (int64_t)i * d, w_out + (int64_t)shared_negative[i] * d, (size_t)d * sizeof(float));
                    orc_gemm_nt(K, input_size, d, M_out, d, M_in, d, M_grad, input_size);
                    for (int i = 0; i < K; i++) {
                        for (int j = 0; j < input_size; j++) {
                            float f = M_grad[i * input_size + j];
                            double lp = softplus((double)f);
                            t->st.loss_rep += lp;
                            if (t->cv && t->cv->n_buckets > 0) {
                                int64_t k = (int64_t)(sent_key / (uint64_t)t->cv->bucket_words);
                                if (k < t->cv->n_buckets) { t->cv->loss[k] += lp; t->cv->pairs[k] += 1; }
                            }
                            const int label = MUT(ORC_MUT_REP_LABEL) ? 1 : 0;
                            if (f > ORC_MAX_EXP) f = (label - 1) * alpha;
                            else if (f < -ORC_MAX_EXP) f = label * alpha;
                            else f = (label - et[(int)((f + ORC_MAX_EXP) * EXP_RES)]) * alpha;
                            M_grad[i * input_size + j] = f;
                        }
                    }
                    orc_gemm_nn(K, d, input_size, M_grad, input_size, M_in, d, M_out_upd, d);
                    orc_gemm_tn(input_size, d, K, M_grad, input_size, M_out, d, M_in, d);
                    for (int i = 0; i < input_size; i++) {
                        int64_t src = (int64_t)i * d, des = (int64_t)sen[min_pos + i] * d;
                        for (int j = 0; j < d; j++) w_in[des + j] += M_in[src + j];
                    }
                    for (int i = 0; i < K; i++) {
                        int64_t src = (int64_t)i * d, des = (int64_t)shared_negative[i] * d;
                        for (int j = 0; j < d; j++) w_out[des + j] += M_out_upd[src + j];
                    }
                    if (p->bf16) {
                        for (int i = 0; i < input_size; i++) round_row_bf16(w_in + (int64_t)sen[min_pos + i] * d, d, &rs);
                        for (int i = 0; i < K; i++) round_row_bf16(w_out + (int64_t)shared_negative[i] * d, d, &rs);
                    }
                    rep_sentence_position += input_size;
                }
            }
        }
        /* cpu:682-685 */
        if (rep_sentence_position >= sentence_length) {
            if (sent_trace_idx >= 0) t->tr->n_tokens += sentence_length;
            sent_trace_idx = -1;
            sentence_length = 0;
            continue;
        }
    }
    t->rng = x;
    t->st.final_rng = x;
    t->st.final_alpha = alpha;
    free(M_in);
    free(M_out);
    free(M_out_upd);
    free(M_grad);
    free(shared_negative);
}

int orc_train(const orc_params *p, const int32_t *stream, uint64_t local_words,
              const uint32_t *counts, int32_t vocab_size, uint64_t train_words,
              const int32_t *table, const float *exp_table,
              float *w_in, float *w_out, orc_stats *stats,
              orc_trace *trace, orc_curve *curve) {
    if (!p || !stream || !counts || !table || !exp_table || !w_in || !w_out) return -1;
    if (p->dim <= 0 || p->window <= 0 || p->batch_size <= 0 || p->iter <= 0 || vocab_size < 2) return -2;
    uint64_t wca = 0;
    float a = p->alpha;
    orc_thread t;
    memset(&t, 0, sizeof(t));
    t.p = p;
    t.stream = stream;
    t.local_words = local_words;
    t.cn = counts;
    t.V = vocab_size;
    t.T = train_words;
    t.table = table;
    t.et = exp_table;
    t.w_in = w_in;
    t.w_out = w_out;
    t.rng = p->seed; /* thread id 0 -> 0, cpu:401 */
    t.tr = trace;
    t.cv = curve;
    t.wc_actual = &wca;
    t.alpha_shared = &a;
    t.shared = 0;
    orc_thread_run(&t);
    if (stats) *stats = t.st;
    return 0;
}

static double orc_now(void) {
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return (double)tv.tv_sec + (double)tv.tv_usec * 1e-6;
}

double orc_train_hogwild(const orc_params *p, const int32_t *stream, uint64_t n_words,
                         const uint32_t *counts, int32_t vocab_size, uint64_t train_words,
                         const int32_t *table, const float *exp_table,
                         float *w_in, float *w_out, orc_stats *stats) {
    int nt = p->threads > 0 ? p->threads : 1;
    uint64_t per = n_words / (uint64_t)nt + (n_words % (uint64_t)nt > 0 ? 1 : 0); /* cpu:408 */
    uint64_t wca = 0;
    float a = p->alpha;
    orc_thread *ts = (orc_thread *)calloc((size_t)nt, sizeof(orc_thread));
    int32_t **slices = (int32_t **)calloc((size_t)nt, sizeof(int32_t *));
    for (int id = 0; id < nt; id++) {
        uint64_t lo = per * (uint64_t)id, hi = lo + per;
        if (lo > n_words) lo = n_words;
        if (hi > n_words) hi = n_words;
        slices[id] = (int32_t *)malloc((size_t)(hi - lo + 1) * sizeof(int32_t));
        memcpy(slices[id], stream + lo, (size_t)(hi - lo) * sizeof(int32_t));
        slices[id][hi - lo] = 0;
        orc_thread *t = &ts[id];
        t->p = p;
        t->stream = slices[id];
        t->local_words = hi - lo;
        t->cn = counts;
        t->V = vocab_size;
        t->T = train_words;
        t->table = table;
        t->et = exp_table;
        t->w_in = w_in;
        t->w_out = w_out;
        t->rng = (uint64_t)id;
        t->wc_actual = &wca;
        t->alpha_shared = &a;
        t->shared = 1;
    }
    double t0 = orc_now();
#pragma omp parallel for num_threads(nt) schedule(static, 1)
    for (int id = 0; id < nt; id++) orc_thread_run(&ts[id]);
    double t1 = orc_now();
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        for (int id = 0; id < nt; id++) {
            orc_stats *s = &ts[id].st;
            stats->raw_words += s->raw_words;
            stats->trained_tokens += s->trained_tokens;
            stats->sentences += s->sentences;
            stats->pairs += s->pairs;
            stats->batches += s->batches;
            stats->negatives += s->negatives;
            stats->rep_pairs += s->rep_pairs;
            stats->loss_attr += s->loss_attr;
            stats->loss_rep += s->loss_rep;
            stats->draws += s->draws;
        }
        stats->final_rng = ts[0].st.final_rng; /* thread 0's */
        stats->final_alpha = ts[0].st.final_alpha;
    }
    for (int id = 0; id < nt; id++) free(slices[id]);
    free(slices);
    free(ts);
    return t1 - t0;
}

/* cpu:397-416 as the reference runs it: each of p->threads threads opens the
 * text file, seeks to byte file_size * id / threads, loads up to
 * ceil(file_words / threads) in-vocabulary words (loadStream, cpu:359-370;
 * ReadWordIndex, cpu:146-196) and trains them with seed id; the returned wall
 * time covers loading and training, like the reference's Elapsed
 * (cpu:395, 702-706).  file_words slices the file (the reference uses
 * train_words for both; a sample of a larger corpus keeps that corpus's
 * train_words for subsampling and the learning rate).  *load_max gets the
 * longest per-thread load time. */
double orc_train_hogwild_file(const orc_params *p, const orc_vocab *v, const char *path, uint64_t file_words,
                              uint64_t train_words, const int32_t *table, const float *exp_table,
                              float *w_in, float *w_out, orc_stats *stats, double *load_max) {
    int nt = p->threads > 0 ? p->threads : 1;
    uint64_t per = file_words / (uint64_t)nt + (file_words % (uint64_t)nt > 0 ? 1 : 0); /* cpu:408 */
    uint64_t wca = 0;
    float a = p->alpha;
    orc_thread *ts = (orc_thread *)calloc((size_t)nt, sizeof(orc_thread));
    int32_t **streams = (int32_t **)calloc((size_t)nt, sizeof(int32_t *));
    int failed = 0;
    double lmax = 0.0;
    double t0 = orc_now();
#pragma omp parallel for num_threads(nt) schedule(static, 1) reduction(| : failed) reduction(max : lmax)
    for (int id = 0; id < nt; id++) {
        double l0 = orc_now();
        streams[id] = (int32_t *)malloc((size_t)(per + 1) * sizeof(int32_t));
        int64_t n = streams[id] ? orc_load_stream(v, path, v->file_size * (uint64_t)id / (uint64_t)nt, per, streams[id])
                                : -1;
        double l1 = orc_now();
        if (l1 - l0 > lmax) lmax = l1 - l0;
        if (n < 0) {
            failed |= 1;
            continue;
        }
        orc_thread *t = &ts[id];
        t->p = p;
        t->stream = streams[id];
        t->local_words = (uint64_t)n;
        t->cn = v->cn;
        t->V = v->size;
        t->T = train_words;
        t->table = table;
        t->et = exp_table;
        t->w_in = w_in;
        t->w_out = w_out;
        t->rng = (uint64_t)id;
        t->wc_actual = &wca;
        t->alpha_shared = &a;
        t->shared = 1;
        orc_thread_run(t);
    }
    double t1 = orc_now();
    if (load_max) *load_max = lmax;
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        for (int id = 0; id < nt; id++) {
            orc_stats *s = &ts[id].st;
            stats->raw_words += s->raw_words;
            stats->trained_tokens += s->trained_tokens;
            stats->sentences += s->sentences;
            stats->pairs += s->pairs;
            stats->batches += s->batches;
            stats->negatives += s->negatives;
            stats->rep_pairs += s->rep_pairs;
            stats->loss_attr += s->loss_attr;
            stats->loss_rep += s->loss_rep;
            stats->draws += s->draws;
        }
        stats->final_rng = ts[0].st.final_rng; /* thread 0's */
        stats->final_alpha = ts[0].st.final_alpha;
    }
    for (int id = 0; id < nt; id++) free(streams[id]);
    free(streams);
    free(ts);
    return failed ? -1.0 : t1 - t0;
}

/* cpu:721-737: header "V d\n", then per word "%s " + d * "%f " (or d raw
 * float32 with -binary 1) + "\n"; only W_in is written. */
int orc_save_model(const char *path, const orc_vocab *v, const float *w_in, int32_t dim, int32_t binary) {
    FILE *fo = fopen(path, "wb");
    if (!fo) return -1;
    fprintf(fo, "%d %d\n", v->size, dim);
    for (int32_t a = 0; a < v->size; a++) {
        fprintf(fo, "%s ", v->word[a]);
        if (binary) fwrite(w_in + (int64_t)a * dim, sizeof(float), (size_t)dim, fo);
        else
            for (int32_t b = 0; b < dim; b++) fprintf(fo, "%f ", w_in[(int64_t)a * dim + b]);
        fprintf(fo, "\n");
    }
    fclose(fo);
    return 0;
}

/* ---------------------------------------------- window loss in FP64 ---- */
/* Per window: logits F[i][0] = <W_in[ctx_i], W_out[centre]> (label 1) and
 * F[i][k] = <W_in[ctx_i], W_out[neg_k]> (label 0); loss = sum softplus(-F_i0)
 * + sum softplus(F_ik).  Same pair semantics as cpu:551-559 (label 1) and
 * cpu:633-648 (label 0) with the exact sigmoid in place of expTable. */
double orc_window_loss_f64(int64_t n_windows, int32_t nctx, int32_t nneg, int32_t dim,
                           const int32_t *centre, const int32_t *ctx, const int32_t *neg,
                           const double *w_in, const double *w_out) {
    double loss = 0.0;
    for (int64_t w = 0; w < n_windows; w++) {
        for (int32_t i = 0; i < nctx; i++) {
            const double *u = w_in + (int64_t)ctx[w * nctx + i] * dim;
            for (int32_t k = 0; k <= nneg; k++) {
                int32_t row = k == 0 ? centre[w] : neg[w * nneg + k - 1];
                const double *v = w_out + (int64_t)row * dim;
                double f = 0.0;
                for (int32_t j = 0; j < dim; j++) f += u[j] * v[j];
                loss += k == 0 ? softplus(-f) : softplus(f);
            }
        }
    }
    return loss;
}

void orc_window_grad_f64(int64_t n_windows, int32_t nctx, int32_t nneg, int32_t dim,
                         const int32_t *centre, const int32_t *ctx, const int32_t *neg,
                         const double *w_in, const double *w_out,
                         double *g_in, double *g_out) {
    for (int64_t w = 0; w < n_windows; w++) {
        for (int32_t i = 0; i < nctx; i++) {
            int32_t ru = ctx[w * nctx + i];
            const double *u = w_in + (int64_t)ru * dim;
            for (int32_t k = 0; k <= nneg; k++) {
                int32_t rv = k == 0 ? centre[w] : neg[w * nneg + k - 1];
                const double *v = w_out + (int64_t)rv * dim;
                double f = 0.0;
                for (int32_t j = 0; j < dim; j++) f += u[j] * v[j];
                double s = 1.0 / (1.0 + exp(-f));
                double gf = k == 0 ? s - 1.0 : s; /* dL/df */
                for (int32_t j = 0; j < dim; j++) {
                    g_in[(int64_t)ru * dim + j] += gf * v[j];
                    g_out[(int64_t)rv * dim + j] += gf * u[j];
                }
            }
        }
    }
}
