/*
 * par_host.h -- host-side C-ABI of the MI355X PAR-SGNS trainer.
 *
 * The reference keeps this layer as global-state functions inside each
 * trainer executable (PAR_Word2Vec_cpu.cpp:96-370, PAR_Word2Vec_gpu_text.cu:
 * 743-828).  Here it is a C++ library (libpar_host.so) with plain pointers and
 * sizes so the CLI, Python (ctypes) and any other FFI can drive it.
 *
 * Conventions: functions return 0 (or a non-negative count) on success and a
 * negative PH_E* code on failure; ph_last_error() returns the message of the
 * calling thread's last failure.  Nothing here touches the GPU.
 *
 * Token ids follow the reference: id 0 is "</s>", ids 1.. are sorted by count
 * descending with ties in first-appearance order (cpu:204-226).
 */
#ifndef PAR_HOST_H
#define PAR_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PH_OK = 0,
    PH_EINVAL = -1,  /* bad argument */
    PH_EIO = -2,     /* file missing or unreadable */
    PH_ENOMEM = -3,  /* allocation failed */
    PH_ERANGE = -4,  /* output capacity too small */
};

const char *ph_last_error(void);

/* ------------------------------------------------------------ vocabulary */
typedef struct ph_vocab ph_vocab;

/* LearnVocabFromTrainFile + SortVocab (cpu:253-292, 204-226; tokenizer
 * ReadWord cpu:120-144; ReduceVocab cpu:229-251).  NULL on failure. */
ph_vocab *ph_vocab_learn_text(const char *train_path, int32_t min_count);
/* ReadVocab (cpu:301-335). */
ph_vocab *ph_vocab_read(const char *vocab_path, int32_t min_count);
/* SaveVocab (cpu:294-299). */
int ph_vocab_save(const ph_vocab *v, const char *path);
/* Vocabulary from a raw integer corpus: raw ids in [1, raw_vocab], 0 ends a
 * line.  The word string of raw id r is "w<r>"; the result is identical to
 * ph_vocab_learn_text on the text rendering of the same corpus. */
ph_vocab *ph_vocab_from_raw(const int32_t *raw, int64_t n, int32_t raw_vocab, int32_t min_count);
void ph_vocab_free(ph_vocab *v);
int32_t ph_vocab_size(const ph_vocab *v);
uint64_t ph_vocab_train_words(const ph_vocab *v);
int ph_vocab_counts(const ph_vocab *v, uint32_t *out);         /* out[V] */
const char *ph_vocab_word(const ph_vocab *v, int32_t id);
int32_t ph_vocab_lookup(const ph_vocab *v, const char *word);   /* -1 if absent */
/* raw id -> vocab id (-1 when pruned) for vocabularies from ph_vocab_from_raw */
int ph_vocab_raw_map(const ph_vocab *v, int32_t *out, int32_t raw_vocab);

/* ---------------------------------------------------------------- corpus */
/* loadStream (cpu:359-370): in-vocabulary ids of the whole file, OOV skipped,
 * a word that ends the file without a newline dropped as ReadWordIndex does.
 * Writes at most cap ids; returns the count (>cap means ERANGE). */
int64_t ph_load_text_stream(const ph_vocab *v, const char *train_path, int32_t *out, int64_t cap);
/* Raw integer corpus -> id stream (pruned words skipped, line ends -> 0). */
int64_t ph_map_raw_stream(const ph_vocab *v, const int32_t *raw, int64_t n, int32_t *out, int64_t cap);

/* Binary corpus: the vocabulary (words, counts, train_words) and the id
 * stream of a text corpus in one file, so a large corpus is tokenised once.
 * The reference re-reads text every run (ReadWord + SearchVocab, cpu:120-196;
 * LearnVocabFromTrainFile cpu:253-292; loadStream cpu:359-370); these three
 * calls have no counterpart there.  Layout, little-endian: 64-byte header
 * ("PSGNSBC1", header size 64, int32 V, uint64 train_words, int64 n_ids,
 * uint64 words_bytes, uint64 ids_offset, 16 reserved bytes), uint32 counts[V],
 * the V words NUL-terminated, zero padding to ids_offset (a multiple of 8),
 * int32 ids[n_ids].  Ids keep the vocabulary's order, so a corpus saved from
 * ph_vocab_learn_text + ph_load_text_stream loads back identical. */
int ph_corpus_save(const char *path, const ph_vocab *v, const int32_t *stream, int64_t n);
/* The vocabulary of a binary corpus (no re-sort or pruning); *n_ids (may be
 * NULL) receives the stream length.  NULL on a missing or malformed file. */
ph_vocab *ph_corpus_vocab(const char *path, int64_t *n_ids);
/* The id stream of a binary corpus: with out NULL returns its length; else
 * writes it (cap >= length, ERANGE otherwise) and checks every id < V. */
int64_t ph_corpus_ids(const char *path, int32_t *out, int64_t cap);

/* ------------------------------------------------------------ subsampling */
/* One-pass subsampling into a sentence CSR for the GPU trainers (the role of
 * PAR_Word2Vec_gpu_text.cu:785-828; keep rule of cpu:458-463).  A sentence
 * ends at a 0 in the stream (not stored) or after max_len kept tokens; a final
 * line not terminated by 0 is dropped (cpu:472).  The keep draw for raw
 * position i is a counter-based hash of (seed, i), so the result does not
 * depend on the thread count.  Empty sentences are not emitted.
 *   tokens[cap_tokens], sent_off[cap_sent+1], sent_raw[cap_sent]: raw-stream
 *   offset of each sentence's first word (drives the learning-rate schedule).
 * Returns the number of sentences; *n_tokens receives the kept-token count. */
int64_t ph_subsample(const int32_t *stream, int64_t n, const uint32_t *counts, int32_t vocab_size,
                     uint64_t train_words, float sample, uint64_t seed, int32_t max_len,
                     int32_t *tokens, int64_t cap_tokens, int64_t *sent_off, int64_t *sent_raw,
                     int64_t cap_sent, int64_t *n_tokens);

/* ---------------------------------------------------------- unigram table */
/* Entries per word of the reference's 1e8-slot unigram^0.75 table
 * (cpu:96-117), computed without materialising the table. */
int ph_unigram_entries(const uint32_t *counts, int32_t vocab_size, int64_t table_size, int64_t *entries);

/* ------------------------------------------------- deterministic schedule */
/* Replays the reference's single-thread random schedule (thread 0 seed,
 * cpu:401) over `iter` epochs: subsampling draws, the window shrink b of every
 * centre, the repulsion window rep_b of every batch and every shared negative
 * including in-sentence rejection (cpu:430-627), and the learning rate each
 * centre sees (cpu:431-446).  The plan is independent of the weights, so the
 * device executes it afterwards. */
typedef struct ph_plan ph_plan;
typedef struct ph_plan_info {
    int64_t n_sent, n_tokens, n_batches, n_negs;
    uint64_t draws, final_rng, raw_words;
    float final_alpha;
} ph_plan_info;
ph_plan *ph_plan_build(const int32_t *stream, int64_t n_words, const uint32_t *counts,
                       int32_t vocab_size, uint64_t train_words, int32_t window, int32_t negative,
                       int32_t batch_size, int32_t iter, float alpha, float sample);
int ph_plan_info_get(const ph_plan *p, ph_plan_info *info);
/* Copies the plan arrays out (any pointer may be NULL):
 *  tokens[n_tokens], b[n_tokens], sent_off[n_sent+1], sent_alpha[2*n_sent],
 *  batch_off[n_sent+1] (first batch of each sentence), neg_off[n_batches+1],
 *  negs[n_negs]. */
int ph_plan_copy(const ph_plan *p, int32_t *tokens, uint8_t *b, int64_t *sent_off, float *sent_alpha,
                 int64_t *batch_off, int64_t *neg_off, int32_t *negs);
void ph_plan_free(ph_plan *p);

/* ------------------------------------------------------------- reference */
/* expTable (cpu:822-826), 1001 floats. */
void ph_exp_table(float *out1001);
/* InitNet (cpu:337-357): W_in from the LCG seeded with 1, W_out = 0. */
void ph_init_net(int64_t vocab_size, int32_t dim, float *w_in, float *w_out);

/* --------------------------------------------------------- synthetic data */
/* Zipf(s) corpus with planted clusters.  Writes n_tokens raw ids in
 * [1, raw_vocab] plus line terminators (0) into out (cap >= n_tokens + lines).
 * Line lengths: fixed `line_len` when lognorm_sigma <= 0, else
 * round(exp(N(mu, sigma))) clipped to [min_len, max_len].  Clusters: n_clusters
 * groups of cluster_size distinct words drawn from the rank band
 * [band_lo, band_hi); for every occurrence of a cluster word, each other
 * position within +-span on its line is replaced, with probability p_replace,
 * by a uniform member of that cluster (triggers are read from the unplanted
 * draw).  cluster_words[n_clusters*cluster_size] receives the raw ids.
 * Returns the stream length. */
int64_t ph_synth_zipf(int64_t n_tokens, int32_t raw_vocab, double zipf_s, int32_t line_len,
                      double lognorm_mu, double lognorm_sigma, int32_t min_len, int32_t max_len,
                      int32_t n_clusters, int32_t cluster_size, int32_t band_lo, int32_t band_hi,
                      int32_t span, double p_replace, uint64_t seed,
                      int32_t *out, int64_t cap, int32_t *cluster_words);
/* DeepWalk corpus on a degree-corrected stochastic block model: n_nodes in
 * n_comm equal communities, target degrees power-law(exponent) rescaled to
 * mean_degree, a fraction p_intra of edge ends inside the community; walks
 * uniform, walks_per_node walks of walk_len nodes, one line each.  Raw id of
 * node u is u+1; community of node u is u % n_comm (written to comm[n_nodes]).
 * Returns the stream length. */
int64_t ph_synth_sbm_walks(int32_t n_nodes, int32_t n_comm, double exponent, double mean_degree,
                           double p_intra, int32_t walks_per_node, int32_t walk_len, uint64_t seed,
                           int32_t *out, int64_t cap, int32_t *comm);
/* The graph behind ph_synth_sbm_walks with the same arguments and seed:
 * degree[n_nodes] receives each node's degree, *intra_edges the number of
 * edges inside a community.  Returns the number of edges. */
int64_t ph_synth_sbm_graph(int32_t n_nodes, int32_t n_comm, double exponent, double mean_degree,
                           double p_intra, uint64_t seed, int64_t *degree, int64_t *intra_edges);
/* Writes the raw corpus as text: "w<id>" tokens, one line per sentence.
 * Graph corpora use the same rendering. */
int ph_write_raw_text(const char *path, const int32_t *raw, int64_t n);

/* ---------------------------------------------------------------- output */
/* saveModel (cpu:721-737): "V d\n", then per word "<word> " followed by d
 * "%f " values (or d raw float32 when binary) and "\n". */
int ph_save_embeddings(const char *path, const ph_vocab *v, const float *w_in, int32_t dim, int32_t binary);

#ifdef __cplusplus
}
#endif
#endif
