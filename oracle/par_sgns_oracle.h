/*
 * par_sgns_oracle.h -- CPU restatement of the PAR-Word2Vec CPU trainer
 * (PAR_Word2Vec_cpu.cpp), used ONLY as test infrastructure: the parity tests,
 * the smoke check and the CPU-baseline leg of the benchmark link it as the
 * checker.  Nothing in the product library calls into it.
 *
 * Every function cites the reference lines it restates.  The restatement keeps
 * the reference's float/double evaluation order so that, built without FMA
 * contraction (-ffp-contract=off, no -march), it is bit-reproducible.  The one
 * deliberate substitution: MKL cblas_sgemm (cpu:633, 652, 655) is replaced by a
 * fixed ascending-k triple loop (acc = 0; acc += a*b), because MKL's internal
 * summation order is unspecified and the MKL headers are not in this image.
 *
 * Parity pins (tests/test_oracle_*.py): LCG and InitNet known answers,
 * EXP_RESOLUTION, expTable against float64 math, unigram table against a
 * brute-force transcription, GEMMs against numpy float64, the single-pair
 * update in closed form, and the whole trainer against an independent
 * pure-Python float32 transcription on tiny corpora.
 */
#ifndef PAR_SGNS_ORACLE_H
#define PAR_SGNS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_EXP_TABLE_SIZE 1000
#define ORC_MAX_EXP 6
#define ORC_MAX_SENTENCE 1000
#define ORC_TABLE_SIZE 100000000
#define ORC_MAX_STRING 100

/* Hyper-parameters of one training run (reference globals, cpu:64-70). */
typedef struct orc_params {
    int32_t dim;        /* -size        (hidden_size) */
    int32_t window;     /* -window */
    int32_t negative;   /* -negative */
    int32_t batch_size; /* -batch-size */
    int32_t iter;       /* -iter */
    int32_t threads;    /* -threads (only the Hogwild variant uses >1) */
    float alpha;        /* -alpha (starting_alpha) */
    float sample;       /* -sample */
    /* Test-only options (0 reproduces the reference): */
    uint64_t seed;      /* LCG seed of the single-thread trainer (the reference's thread 0 uses 0,
                           cpu:401); other values give independent runs of the same schedule rules */
    int32_t bf16;       /* 1: emulate bf16 tables with stochastic rounding -- a row is rounded where
                           the GPU trainer writes it back (a W_out centre row after its pairs, a
                           W_in row when its word leaves the window, every row of a repulsion batch
                           after the batch), with its own xorshift32 noise stream */
    int32_t strong_rng; /* 1: draws from a well-mixed generator instead of the reference's LCG, whose
                           low bits repeat with short periods (x_n's parity alternates with n) */
    int32_t alpha_per_sentence; /* 1: the learning rate of each sentence from the words read before
                           it (the GPU trainer's rule) instead of a refresh every 10000 words
                           (cpu:431-446) */
    int32_t reserved;
} orc_params;

/* Counters and the FP64 loss a run produces. */
typedef struct orc_stats {
    uint64_t raw_words;      /* stream words consumed, incl. </s>, all epochs */
    uint64_t trained_tokens; /* tokens kept after subsampling (centre steps) */
    uint64_t sentences;      /* non-empty sentences trained */
    uint64_t pairs;          /* attraction pairs */
    uint64_t batches;        /* repulsion batches */
    uint64_t negatives;      /* shared negatives drawn (sum of K over batches) */
    uint64_t rep_pairs;      /* repulsion (negative, input) pairs = sum K*B_b */
    double loss_attr;        /* sum of -log sigmoid(f) over attraction pairs */
    double loss_rep;         /* sum of -log sigmoid(-f) over repulsion pairs */
    uint64_t draws;          /* LCG draws consumed */
    uint64_t final_rng;      /* LCG state at the end */
    float final_alpha;
} orc_stats;

/* Optional record of every random decision, in consumption order. */
typedef struct orc_trace {
    int64_t cap_tokens, cap_batches, cap_negs, cap_sent;
    int32_t *tokens;      /* kept tokens of each trained sentence, concatenated */
    uint8_t *b;           /* attraction window shrink per centre (parallel to tokens) */
    int32_t *sent_len;    /* length of each trained sentence */
    float *sent_alpha;    /* [2*sent]: alpha of centre 0, alpha of the rest */
    int32_t *rep_b;       /* per batch */
    int32_t *negs;        /* all shared negatives, concatenated */
    int64_t n_tokens, n_batches, n_negs, n_sent;
    int32_t overflow;
} orc_trace;

/* Loss curve: FP64 loss summed into buckets of `bucket_words` raw words,
 * keyed by the raw-word offset (epoch*T + word_count before loading). */
typedef struct orc_curve {
    int64_t bucket_words;
    int64_t n_buckets;
    double *loss;       /* attraction + repulsion */
    double *pairs;
    double *loss_attr;  /* attraction pairs only (may be NULL) */
    double *pairs_attr;
} orc_curve;

/* Planted misreadings of the reference (test infrastructure for the
 * infrastructure): the pin tests switch one on at a time and check that the
 * oracle's known answers then fail.  0 (none) in every parity use. */
enum {
    ORC_MUT_NONE = 0,
    ORC_MUT_ALPHA_AFTER_LOAD = 1, /* learning-rate check after the sentence load (cpu:431-446 is before it) */
    ORC_MUT_ALPHA_GE = 2,         /* ">=" in the 10000-word check (cpu:431 has ">") */
    ORC_MUT_NO_REJECTION = 3,     /* no in-sentence rejection of negatives (cpu:609-617) */
    ORC_MUT_REPB_2W = 4,          /* rep_b = x % 2W for every batch (cpu:590-601: x % W at the ends) */
    ORC_MUT_WINDOW_INCL = 5,      /* window loop to 2W+1-b inclusive (cpu:536 stops before) */
    ORC_MUT_PAIR_ORDER = 6,       /* W_in takes g * the updated W_out (cpu:562-564 saves tmp first) */
    ORC_MUT_UNIGRAM_LINEAR = 7,   /* unigram table without the 0.75 power (cpu:98-112) */
    ORC_MUT_SUBSAMPLE_FLIP = 8,   /* subsampling keeps exactly the words the reference drops (cpu:462-463) */
    ORC_MUT_TABLE_INDEX = 9,      /* table slot x % size instead of (x >> 16) % size (cpu:606) */
    ORC_MUT_WCA_LOCAL = 10,       /* epoch end adds local_words - last, not word_count - last (cpu:515-516) */
    ORC_MUT_REP_LABEL = 11,       /* repulsion gradient with label 1 instead of 0 (cpu:641-647) */
    ORC_MUT_BATCH_END = 12,       /* end-of-sentence batch rule with L - min_pos (cpu:594 has (L-1) - min_pos) */
};
void orc_set_mutation(int32_t m);
int32_t orc_mutation(void);

uint64_t orc_lcg(uint64_t x);                                   /* cpu:461 */
float orc_exp_resolution(void);                                 /* cpu:70 */
void orc_exp_table(float *table_1001);                          /* cpu:822-826 */
void orc_init_net(int64_t vocab_size, int32_t dim, float *w_in, float *w_out); /* cpu:337-357 */
int orc_unigram_table(const uint32_t *counts, int32_t vocab_size, int32_t *table); /* cpu:96-117 */
int orc_unigram_table_n(const uint32_t *counts, int32_t vocab_size, int32_t *table, int64_t table_size);

/* Text front end (cpu:120-335, 359-370). */
typedef struct orc_vocab orc_vocab;
orc_vocab *orc_vocab_learn(const char *path, int32_t min_count);  /* cpu:253-292 */
orc_vocab *orc_vocab_read(const char *vocab_path, const char *train_path, int32_t min_count); /* cpu:301-335 */
int orc_vocab_save(const orc_vocab *v, const char *path);         /* cpu:294-299 */
int32_t orc_vocab_size(const orc_vocab *v);
uint64_t orc_vocab_train_words(const orc_vocab *v);
uint64_t orc_vocab_file_size(const orc_vocab *v);
uint32_t orc_vocab_count(const orc_vocab *v, int32_t i);
const char *orc_vocab_word(const orc_vocab *v, int32_t i);
int32_t orc_vocab_search(const orc_vocab *v, const char *word);
void orc_vocab_counts(const orc_vocab *v, uint32_t *out);
void orc_vocab_free(orc_vocab *v);
/* Reads up to max_words in-vocabulary ids starting at byte `offset`,
 * appends the 0 sentinel, returns the number of words read. */
int64_t orc_load_stream(const orc_vocab *v, const char *path, uint64_t offset,
                        uint64_t max_words, int32_t *stream);

/* Single-thread trainer (cpu:372-700 with num_threads = 1, thread id 0).
 * stream has local_words entries followed by the 0 sentinel. */
int orc_train(const orc_params *p, const int32_t *stream, uint64_t local_words,
              const uint32_t *counts, int32_t vocab_size, uint64_t train_words,
              const int32_t *table, const float *exp_table,
              float *w_in, float *w_out, orc_stats *stats,
              orc_trace *trace, orc_curve *curve);

/* OpenMP Hogwild trainer used as the CPU throughput baseline: p->threads
 * threads, thread t seeded with t (cpu:401) and owning the t-th contiguous
 * ceil(T/threads) slice of the id stream (the reference slices the file by
 * bytes, cpu:405-416).  Returns the training wall time in seconds. */
double orc_train_hogwild(const orc_params *p, const int32_t *stream, uint64_t n_words,
                         const uint32_t *counts, int32_t vocab_size, uint64_t train_words,
                         const int32_t *table, const float *exp_table,
                         float *w_in, float *w_out, orc_stats *stats);

/* The Hogwild trainer as the reference's harness runs it (cpu:397-416):
 * per-thread byte-offset slices of a text file, loading inside the timed
 * region.  Returns the wall time of loading + training (-1 on error);
 * *load_max (if given) gets the longest per-thread load time. */
double orc_train_hogwild_file(const orc_params *p, const orc_vocab *v, const char *path, uint64_t file_words,
                              uint64_t train_words, const int32_t *table, const float *exp_table,
                              float *w_in, float *w_out, orc_stats *stats, double *load_max);

/* Row-major naive GEMMs with the oracle's fixed summation order. */
void orc_gemm_nt(int m, int n, int k, const float *a, int lda, const float *b, int ldb, float *c, int ldc);
void orc_gemm_nn(int m, int n, int k, const float *a, int lda, const float *b, int ldb, float *c, int ldc);
void orc_gemm_tn(int m, int n, int k, const float *a, int lda, const float *b, int ldb, float *c, int ldc);

/* Embedding writer (cpu:721-737). */
int orc_save_model(const char *path, const orc_vocab *v, const float *w_in, int32_t dim, int32_t binary);

/* SGNS window-batch loss in FP64 (used to pin the adjoint kernels):
 * windows of (centre, nctx contexts, nneg shared negatives); contexts are
 * W_in rows, centre/negatives are W_out rows, exactly as the attraction
 * (label 1) and repulsion (label 0) pairs of the trainer. */
double orc_window_loss_f64(int64_t n_windows, int32_t nctx, int32_t nneg, int32_t dim,
                           const int32_t *centre, const int32_t *ctx, const int32_t *neg,
                           const double *w_in, const double *w_out);
void orc_window_grad_f64(int64_t n_windows, int32_t nctx, int32_t nneg, int32_t dim,
                         const int32_t *centre, const int32_t *ctx, const int32_t *neg,
                         const double *w_in, const double *w_out,
                         double *g_in, double *g_out);

#ifdef __cplusplus
}
#endif
#endif
