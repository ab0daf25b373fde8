/*
 * par_sgns.h -- device C-ABI of the MI355X PAR-SGNS trainer (libpar_sgns.so).
 *
 * Replaces the reference's host->device boundary of PAR_Word2Vec_gpu_text.cu:
 * the device globals it owns (text.cu:84-86), their cudaMalloc/cudaMemcpy
 * staging (text.cu:112-113, 371-374, 841-868, 963), and the launch of
 * par_sgns_kernel (text.cu:381, 899, 912).  Instead of globals and exit(0) on
 * error (text.cu:42-48), a handle owns all device memory and one HIP stream,
 * every call returns a status code, and par_sgns_last_error() explains the
 * last failure.  No torch or HIP types appear in the signatures.
 *
 * Data layout contract (PAR_Word2Vec_cpu.cpp:80, 338-339): W_in and W_out are
 * row-major [vocab_size][dim] float32 on the host side of this interface; on
 * the device each row is padded to `row_stride` elements (zeros) and stored in
 * the configured precision.  Token id 0 is "</s>".
 *
 * One handle per host thread.  Calls on one handle are serialised on its
 * stream; every call that returns results synchronises that stream.
 */
#ifndef PAR_SGNS_H
#define PAR_SGNS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PAR_SGNS_OK = 0,
    PAR_SGNS_EINVAL = -1,   /* bad argument or unsupported shape */
    PAR_SGNS_ENOMEM = -2,   /* device allocation failed */
    PAR_SGNS_EDEVICE = -3,  /* HIP runtime error */
    PAR_SGNS_ESTATE = -4,   /* call out of order (e.g. train before upload) */
};

enum par_sgns_precision { PAR_SGNS_FP32 = 0, PAR_SGNS_BF16 = 1 };
/* How a bf16 table entry absorbs an fp32 update: stochastic rounding (the
 * default; unbiased for updates below half an ulp) or round-to-nearest
 * (ties away from zero), for both tables or round-to-nearest for one table
 * only (NEAREST_OUT: W_out rows nearest, W_in rows stochastic; NEAREST_IN the
 * reverse).  Training modes only; the deterministic mode is fp32. */
enum par_sgns_rounding {
    PAR_SGNS_ROUND_STOCHASTIC = 0,
    PAR_SGNS_ROUND_NEAREST = 1,
    PAR_SGNS_ROUND_NEAREST_OUT = 2,
    PAR_SGNS_ROUND_NEAREST_IN = 3,
};

/* How plain/persistent modes evaluate a repulsion batch (the three SGEMMs
 * of PAR_Word2Vec_cpu.cpp:633-656).  AUTO: matrix cores on bf16 tables,
 * vector ALUs on fp32 tables.  MATRIX (bf16 tables only): G = N C^T on
 * bf16 MFMA with fp32 sums, the gradient rounded to bf16 for the two update
 * products.  VECTOR: fp32 dot products, register blocks of negatives. */
enum par_sgns_repulsion { PAR_SGNS_REP_AUTO = 0, PAR_SGNS_REP_VECTOR = 1, PAR_SGNS_REP_MATRIX = 2 };

/* How the matrix-core instance writes a row back.  ATOMIC: a packed-bf16
 * atomic adds the step from the row as the wavefront read it to its rounded
 * new value, so a concurrent update is never lost.  The others store the
 * rounded value with plain stores, like the reference CPU trainer's in-place
 * Hogwild updates (cpu:561-565, 658-674): a concurrent update of the row
 * between the wavefront's read and its store is overwritten.  STORE_OUT
 * stores W_out rows (centres and shared negatives, held for one centre or one
 * batch) and keeps atomics for W_in rows (held in the window ring for 2W+1
 * centres, so concurrent updates of a frequent word's row are common there);
 * STORE_IN the reverse; STORE stores both tables; HOT stores rows with
 * id >= hot_rows (the less frequent words) of both tables. */
enum par_sgns_write_back {
    PAR_SGNS_WB_ATOMIC = 0,
    PAR_SGNS_WB_STORE_OUT = 1,
    PAR_SGNS_WB_STORE_IN = 2,
    PAR_SGNS_WB_STORE = 3,
    PAR_SGNS_WB_HOT = 4,
};

enum par_sgns_mode {
    PAR_SGNS_DETERMINISTIC = 0, /* one wavefront replays an uploaded plan exactly */
    PAR_SGNS_PLAIN = 1,         /* one workgroup (one wavefront) per sentence */
    PAR_SGNS_PERSISTENT = 2,    /* resident wavefronts pull sentences from a queue */
};

typedef struct par_sgns_handle_s *par_sgns_handle;

typedef struct par_sgns_config {
    int32_t vocab_size;   /* V, including "</s>" */
    int32_t dim;          /* -size, 1..512 */
    int32_t window;       /* -window, 1..31 (the ring of 2W+1 slots is one wavefront wide) */
    int32_t negative;     /* -negative, 0..64 */
    int32_t batch_size;   /* -batch-size (repulsion batch B), 1..64 */
    int32_t precision;    /* par_sgns_precision */
    int32_t device;       /* HIP device ordinal */
    int32_t rounding;     /* bf16 write-back: par_sgns_rounding */
} par_sgns_config;

typedef struct par_sgns_train_args {
    int32_t mode;         /* par_sgns_mode */
    int32_t iter;         /* total epochs of the schedule (alpha decay) */
    int32_t epoch;        /* current epoch, 0-based */
    int32_t compute_loss; /* accumulate the FP64 SGNS loss */
    float alpha;          /* starting learning rate alpha0 */
    int32_t waves_per_block; /* 1..8; persistent mode: 0 = the most resident waves */
    int32_t blocks;          /* persistent mode; 0 = tuned default */
    int32_t use_graph;       /* persistent mode: replay a captured hipGraph */
    int32_t schedule;        /* plain/persistent: 0 = Philox draws over the uploaded
                                corpus, 1 = the uploaded exact plan (reference draws) */
    int32_t repulsion;       /* plain/persistent: par_sgns_repulsion */
    uint64_t train_words; /* T = sum of vocabulary counts */
    uint64_t seed;        /* Philox key for windows, negatives, rounding */
    int64_t sent_begin;   /* sentence range [begin, end) of the uploaded corpus */
    int64_t sent_end;     /* (deterministic mode: of the uploaded plan) */
    int32_t sentences_per_dequeue; /* persistent mode: queue slots a wavefront takes per
                                      dequeue, 1..64; 0 = tuned default (1 untuned) */
    int32_t min_sentences_per_wave; /* Hogwild staleness bound.  Persistent mode, default
                                       grid: at least this many sentences per wavefront
                                       (fewer waves on small ranges); 0 = 24.  Plain mode,
                                       if > 0: at most range / this many sentences in flight
                                       (consecutive launches of that many sentences where
                                       fewer than that fit on the device at once); 0 = one
                                       launch over the range */
    int32_t write_back;   /* par_sgns_write_back; the vector repulsion (fp32 tables, or
                             bf16 on request) accepts ATOMIC and STORE_OUT */
    int32_t hot_rows;     /* PAR_SGNS_WB_HOT: rows with id < hot_rows keep their atomics */
} par_sgns_train_args;

typedef struct par_sgns_stats {
    double loss_attr;     /* sum of -log sigmoid(f), attraction pairs */
    double loss_rep;      /* sum of -log sigmoid(-f), repulsion pairs */
    uint64_t tokens;      /* centre tokens trained */
    uint64_t sentences;
    uint64_t pairs;       /* attraction pairs */
    uint64_t batches;     /* repulsion batches */
    uint64_t negatives;   /* sum over batches of K */
    uint64_t rep_pairs;   /* sum over batches of K * batch rows */
    double kernel_ms;     /* device time of the training launches (HIP events) */
    /* Algorithmic work of the pass (what the reference's per-pair and
     * per-batch arithmetic moves, cpu:529-674, independent of caching):
     * bytes = 4 tokens + e*dim*(4 tokens + 2 negatives), e = 4 (fp32) or
     * 2 (bf16) -- each trained token's W_in and W_out rows read and written
     * once, each shared negative's W_out row read and written once per batch;
     * flops = 6*dim*(pairs + rep_pairs). */
    double alg_bytes;
    double alg_flops;
    /* Launch shape used: workgroups, wavefronts per workgroup, queue slots per
     * dequeue (plain mode: one sentence per wavefront, dequeue 0). */
    int32_t blocks;
    int32_t waves_per_block;
    int32_t sentences_per_dequeue;
    int32_t reserved;
} par_sgns_stats;

/* Handle lifetime.  Replaces the device allocations of text.cu:348-376,
 * 841-868 and the cudaFree calls of text.cu:969-977. */
int par_sgns_create(const par_sgns_config *cfg, par_sgns_handle *out);
int par_sgns_destroy(par_sgns_handle h);
const char *par_sgns_last_error(par_sgns_handle h);
/* Device row stride (elements) of W_in/W_out. */
int32_t par_sgns_row_stride(par_sgns_handle h);

/* expTable upload (text.cu:1092-1100); n must be 1001. */
int par_sgns_set_exp_table(par_sgns_handle h, const float *table, int32_t n);

/* Negative sampler: entries[V] = slots per word in the reference's unigram
 * table (cpu:96-117).  Built into an on-device alias table; the table's
 * "</s>" slots are folded into a uniform draw over [1, V-1] as cpu:606-608
 * does.  Replaces the 400 MB d_table upload (text.cu:112-113). */
int par_sgns_set_sampler(par_sgns_handle h, const int64_t *entries, int32_t vocab_size);

/* n draws of the on-device negative sampler (alias table + Philox, before
 * in-sentence rejection), keyed like the trainer's draws for sentence ids
 * sent, sent+1, ... (8192 draws each).  For checking the distribution. */
int par_sgns_sample_negatives(par_sgns_handle h, uint64_t seed, int64_t sent, int64_t n, int32_t *out);

/* Weights (text.cu:371-374 upload, text.cu:963 download).  Host arrays are
 * [V][dim] float32; NULL skips that table.  BF16 handles round to nearest on
 * upload. */
int par_sgns_set_weights(par_sgns_handle h, const float *w_in, const float *w_out);
int par_sgns_get_weights(par_sgns_handle h, float *w_in, float *w_out);

/* Sentence CSR for the fast modes (d_corpus / d_sen_len_ptr, text.cu:841-845):
 * tokens[n_tokens], sent_off[n_sent+1] (int64), sent_raw[n_sent] = raw-word
 * offset of each sentence in the epoch (learning-rate schedule).  Sentences
 * longer than 1000 tokens are rejected. */
int par_sgns_set_corpus(par_sgns_handle h, const int32_t *tokens, int64_t n_tokens,
                        const int64_t *sent_off, const int64_t *sent_raw, int64_t n_sent);

/* Exact schedule for the deterministic mode (layout of ph_plan_copy in
 * par_host.h). */
int par_sgns_set_plan(par_sgns_handle h, const int32_t *tokens, const uint8_t *b, int64_t n_tokens,
                      const int64_t *sent_off, const float *sent_alpha, int64_t n_sent,
                      const int64_t *batch_off, const int64_t *neg_off, int64_t n_batches,
                      const int32_t *negs, int64_t n_negs);

/* One training pass over a sentence range (the epoch loop body of
 * text.cu:879-952: one par_sgns_kernel launch per epoch there). */
int par_sgns_train(par_sgns_handle h, const par_sgns_train_args *args, par_sgns_stats *stats);

/* Everything a later par_sgns_train with the same arguments needs, without
 * running it: the longest-first queue order of the sentence range and, with
 * use_graph, the captured hipGraph.  Lets a caller keep host work out of a
 * timed region. */
int par_sgns_prepare(par_sgns_handle h, const par_sgns_train_args *args);

/* One measured launch shape of the persistent kernel. */
typedef struct par_sgns_tune_point {
    int32_t waves_per_block;
    int32_t waves_per_cu;          /* resident wavefronts per CU (grid = CUs * this / block) */
    int32_t sentences_per_dequeue;
    int32_t reserved;
    double kernel_ms;              /* median over the repetitions */
} par_sgns_tune_point;

/* Launch-shape autotuner (no counterpart in the reference, whose launch shape
 * is fixed at gamma*56 blocks of `dim` threads, text.cu:848-868, 899).  Times
 * the persistent kernel on the sentence range of `args` (mode must be
 * PAR_SGNS_PERSISTENT) for every feasible combination of waves per block
 * {1, 2, 4, 8}, resident waves per CU {4, 6, 8} (bounded by the LDS a wave
 * needs) and sentences per dequeue {1, 2, 4}, `reps` runs each (median).  Both
 * tables are copied aside first and restored before every run and at the end,
 * so tuning trains nothing.  The fastest shape becomes the handle's default
 * for later calls that leave waves_per_block, blocks and
 * sentences_per_dequeue at 0.  Every measurement goes to points[0..cap) in
 * trial order, their number to *n_points (may exceed cap); the chosen shape to
 * *best.  Needs device memory for one copy of both tables. */
int par_sgns_autotune(par_sgns_handle h, const par_sgns_train_args *args, int32_t reps, par_sgns_tune_point *points,
                      int32_t cap, int32_t *n_points, par_sgns_tune_point *best);
/* Sets the handle's default persistent launch shape to a saved tuning (what
 * par_sgns_autotune would choose; no counterpart in the reference): waves per
 * block 1..8, resident waves per CU 1..8 (a multiple of the block; capped at
 * what the LDS holds), sentences per dequeue 1..64.  All three 0 restore the
 * built-in default (the block with the most resident waves, one sentence per
 * dequeue). */
int par_sgns_set_launch_shape(par_sgns_handle h, int32_t waves_per_block, int32_t waves_per_cu,
                              int32_t sentences_per_dequeue);

/* SGNS window-batch loss and its adjoint (no counterpart in the reference,
 * which fuses them into the update at cpu:551-565, 633-656).  Window w has
 * centre[w] (W_out row, label 1), ctx[w*nctx..] (W_in rows) and shared
 * negatives neg[w*nneg..] (W_out rows, label 0); the logits are
 * F = X [c, n_1..n_nneg]^T per window (nctx <= 32, nneg <= 7).  loss_out
 * receives the FP64 sum of -log sigmoid(F_i0) + sum_k -log sigmoid(-F_ik)
 * with the exact sigmoid.  The adjoint writes dense gradients dL/dW_in and
 * dL/dW_out (FP32, [V][dim], zeroed first) to the handle's gradient buffers,
 * fetched with par_sgns_get_grads.  flags: PAR_SGNS_WIN_FP64 evaluates the
 * logits and the loss in FP64 (gradient checks); PAR_SGNS_WIN_MFMA runs the
 * (nctx x dim)(dim x (1+nneg)) contraction on the matrix cores with FP32
 * accumulation: BF16 matrix cores on BF16 tables (below 4 GiB - 4 KiB each,
 * any window), FP32 matrix cores on FP32 tables (windows of at most 16 rows,
 * 1 + nctx + nneg <= 16). */
enum { PAR_SGNS_WIN_MFMA = 1, PAR_SGNS_WIN_FP64 = 2 };
int par_sgns_loss_forward(par_sgns_handle h, const int32_t *centre, const int32_t *ctx, const int32_t *neg,
                          int64_t n_windows, int32_t nctx, int32_t nneg, int32_t flags, double *loss_out);
int par_sgns_loss_adjoint(par_sgns_handle h, const int32_t *centre, const int32_t *ctx, const int32_t *neg,
                          int64_t n_windows, int32_t nctx, int32_t nneg, int32_t flags, double *loss_out);
int par_sgns_get_grads(par_sgns_handle h, float *g_in, float *g_out);

/* par_sgns_loss_forward without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  Same arguments, checks and result as
 * par_sgns_loss_forward, but the window indices are staged and the FP64 loss
 * is accumulated in a fixed 8 MiB device array that the library holds per
 * device until the process ends (looked up at par_sgns_create), instead of
 * handle buffers grown with hipMalloc.  Limit: n_windows * (1 + nctx + nneg)
 * <= 2,097,088 indices (65,536 windows of 10 contexts and 5 negatives use
 * 1,048,576); larger batches return PAR_SGNS_EINVAL.  Every handle on one
 * device shares that array and it keeps the last call's contents (the loss
 * accumulators are cleared by each call): calls that overlap on one device
 * (handles on different host threads or streams) race and must be
 * serialised by the caller. */
int par_sgns_loss_forward_noalloc(par_sgns_handle h, const int32_t *centre, const int32_t *ctx, const int32_t *neg,
                                  int64_t n_windows, int32_t nctx, int32_t nneg, int32_t flags, double *loss_out);

/* Device time (ms) and launch count of the most recent kernel of the given
 * kind: 0 = training, 1 = loss forward, 2 = adjoint. */
int par_sgns_last_kernel_time(par_sgns_handle h, int32_t kind, double *ms, int32_t *launches);

#ifdef __cplusplus
}
#endif
#endif
