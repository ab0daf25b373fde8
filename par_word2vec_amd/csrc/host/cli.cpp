// cli.cpp -- par_word2vec: drop-in command line for the PAR-Word2Vec trainer.
//
// Flags and defaults follow PAR_Word2Vec_cpu.cpp:64-70, 787-818 (ArgPos
// semantics: last value wins only where the reference reads it; unknown
// flags are ignored); the GPU trainer's -gamma / -max-num-sen
// (PAR_Word2Vec_gpu_text.cu:1083-1086) and -threads / -disk are accepted and
// ignored.  Output: saveModel's format (cpu:721-737) and an "Elapsed <s>"
// timing file like cpu:702-706 / text.cu:957-959.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "par_host.h"
#include "par_sgns.h"

namespace {

int arg_pos(const char *name, int argc, char **argv) {  // cpu:713-719
    for (int a = 1; a < argc; a++)
        if (!std::strcmp(name, argv[a])) return a;
    return -1;
}

struct Opts {
    std::string train, output, save_vocab, read_vocab, mode = "persistent", precision = "fp32", time_file;
    std::string train_bin, save_bin;  // binary corpus (vocabulary + id stream): read instead of -train / write
    int size = 100, window = 5, negative = 5, iter = 5, min_count = 5, batch = 24, binary = 0, debug = 2;
    int device = 0;
    float alpha = 0.025f, sample = 1e-3f;
    unsigned long long seed = 1;
    long long chunk_words = 10000000;
    int autotune = 0;  // persistent mode: tune the launch shape on the first chunk, this many runs per shape
    int min_sent_wave = 0;  // staleness bound: persistent grid floor (0 = 16); plain mode: sentences in flight <= n / this (0 = off)
    std::string write_back = "atomic";  // bf16 matrix-core write-back policy
    std::string rounding = "stochastic";  // bf16 write-back rounding (per table: par_sgns_rounding)
    int hot_rows = 0;                   // write_back hot: rows with id < this keep atomics
};

void usage() {
    std::printf(
        "par_word2vec -- PAR-Word2Vec attraction-repulsion SGNS on MI355X\n\n"
        "  -train <file> -output <file> -size <int> (100) -window <int> (5) -sample <float> (1e-3)\n"
        "  -negative <int> (5) -iter <int> (5) -min-count <int> (5) -alpha <float> (0.025)\n"
        "  -batch-size <int> (24) -binary <int> (0) -debug <int> (2) -save-vocab <file> -read-vocab <file>\n"
        "  -mode det|plain|persistent (persistent) -precision fp32|bf16 (fp32) -seed <int> -device <int>\n"
        "  -chunk-words <int> (10000000) -time-file <file> (PAR_Word2Vec_gpu_time)\n"
        "  -autotune <runs> (0): persistent mode, tune the launch shape on the first chunk first\n"
        "  -min-sentences-per-wave <int>: staleness bound; persistent mode: sentences per wave (24);\n"
        "    plain mode: at most n / this sentences in flight (0: off)\n"
        "  -write-back atomic|store_out|store_in|store|hot (atomic) -hot-rows <int>: bf16 row write-back\n"
        "  -rounding stochastic|nearest|nearest_out|nearest_in (stochastic): bf16 write-back rounding\n"
        "  -train-bin <file>: binary corpus (vocabulary + ids) in place of -train / -read-vocab\n"
        "  -save-bin <file>: write the binary corpus of -train (then train only if -output is given)\n"
        "  accepted and ignored: -threads -disk -gamma -max-num-sen\n");
}

int fail(const char *what, const char *msg) {
    std::fprintf(stderr, "par_word2vec: %s: %s\n", what, msg ? msg : "");
    return 1;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc == 1) {
        usage();
        return 0;
    }
    Opts o;
    int i;
    auto str = [&](const char *f, std::string &dst) { if ((i = arg_pos(f, argc, argv)) > 0 && i + 1 < argc) dst = argv[i + 1]; };
    auto num = [&](const char *f, int &dst) { if ((i = arg_pos(f, argc, argv)) > 0 && i + 1 < argc) dst = std::atoi(argv[i + 1]); };
    auto flt = [&](const char *f, float &dst) { if ((i = arg_pos(f, argc, argv)) > 0 && i + 1 < argc) dst = (float)std::atof(argv[i + 1]); };
    num("-size", o.size);
    str("-train", o.train);
    str("-save-vocab", o.save_vocab);
    str("-read-vocab", o.read_vocab);
    num("-debug", o.debug);
    num("-binary", o.binary);
    flt("-alpha", o.alpha);
    str("-output", o.output);
    num("-window", o.window);
    flt("-sample", o.sample);
    num("-negative", o.negative);
    num("-iter", o.iter);
    num("-min-count", o.min_count);
    num("-batch-size", o.batch);
    str("-mode", o.mode);
    str("-precision", o.precision);
    num("-device", o.device);
    num("-autotune", o.autotune);
    num("-min-sentences-per-wave", o.min_sent_wave);
    str("-write-back", o.write_back);
    str("-rounding", o.rounding);
    num("-hot-rows", o.hot_rows);
    str("-time-file", o.time_file);
    str("-train-bin", o.train_bin);
    str("-save-bin", o.save_bin);
    if ((i = arg_pos("-seed", argc, argv)) > 0 && i + 1 < argc) o.seed = std::strtoull(argv[i + 1], nullptr, 10);
    if ((i = arg_pos("-chunk-words", argc, argv)) > 0 && i + 1 < argc) o.chunk_words = std::atoll(argv[i + 1]);
    if (o.time_file.empty()) o.time_file = "PAR_Word2Vec_gpu_time";
    const int mode = o.mode == "det" ? PAR_SGNS_DETERMINISTIC : o.mode == "plain" ? PAR_SGNS_PLAIN : PAR_SGNS_PERSISTENT;
    if (o.mode != "det" && o.mode != "plain" && o.mode != "persistent") return fail("-mode", "det, plain or persistent");
    if (o.precision != "fp32" && o.precision != "bf16") return fail("-precision", "fp32 or bf16");
    const char *roundings[] = {"stochastic", "nearest", "nearest_out", "nearest_in"};
    int rounding = -1;
    for (int k = 0; k < 4; k++)
        if (o.rounding == roundings[k]) rounding = k;
    if (rounding < 0) return fail("-rounding", "stochastic, nearest, nearest_out or nearest_in");
    if (o.train.empty() && o.train_bin.empty()) return fail("-train", "training file (or -train-bin) required");
    if (!o.train.empty() && !o.train_bin.empty()) return fail("-train-bin", "give -train or -train-bin, not both");
    if (!o.train_bin.empty() && !o.save_bin.empty()) return fail("-save-bin", "the corpus is binary already");

    std::printf("hidden size: %d\nwindow size: %d\nnumber of negative samples: %d\nbatch size: %d\n"
                "number of iterations: %d\nstarting learning rate: %.5f\nmode: %s, precision: %s\n"
                "starting training using file: %s\n\n",
                o.size, o.window, o.negative, o.batch, o.iter, o.alpha, o.mode.c_str(), o.precision.c_str(),
                o.train_bin.empty() ? o.train.c_str() : o.train_bin.c_str());

    // vocabulary (cpu:378-384) and stream (cpu:359-370), or both from a binary corpus
    const auto l0 = std::chrono::steady_clock::now();
    int64_t n = 0;
    ph_vocab *v = !o.train_bin.empty()     ? ph_corpus_vocab(o.train_bin.c_str(), &n)
                  : o.read_vocab.empty() ? ph_vocab_learn_text(o.train.c_str(), o.min_count)
                                         : ph_vocab_read(o.read_vocab.c_str(), o.min_count);
    if (!v) return fail("vocabulary", ph_last_error());
    if (!o.save_vocab.empty() && ph_vocab_save(v, o.save_vocab.c_str()) < 0) return fail("save-vocab", ph_last_error());
    const int32_t V = ph_vocab_size(v);
    const uint64_t T = ph_vocab_train_words(v);
    if (o.debug > 0) std::printf("Vocab size: %d\nWords in train file: %llu\n", V, (unsigned long long)T);
    if (o.output.empty() && o.save_bin.empty()) return 0;  // cpu:385
    std::vector<int32_t> stream;
    if (!o.train_bin.empty()) {
        stream.resize((size_t)n);
        if (ph_corpus_ids(o.train_bin.c_str(), stream.data(), n) < 0) return fail("stream", ph_last_error());
    } else {
        n = ph_load_text_stream(v, o.train.c_str(), nullptr, 0);
        if (n < 0) return fail("stream", ph_last_error());
        stream.resize((size_t)n);
        if (ph_load_text_stream(v, o.train.c_str(), stream.data(), n) < 0) return fail("stream", ph_last_error());
    }
    if (o.debug > 0)
        std::printf("corpus loaded: %lld ids in %.2f s\n", (long long)n,
                    std::chrono::duration<double>(std::chrono::steady_clock::now() - l0).count());
    if (!o.save_bin.empty()) {
        if (ph_corpus_save(o.save_bin.c_str(), v, stream.data(), n) < 0) return fail("save-bin", ph_last_error());
        if (o.output.empty()) return 0;
    }
    std::vector<uint32_t> counts((size_t)V);
    ph_vocab_counts(v, counts.data());

    // device handle, tables, weights (cpu:387-388; text.cu:348-376, 1092-1100)
    par_sgns_config cfg{V, o.size, o.window, o.negative, o.batch,
                        o.precision == "bf16" ? PAR_SGNS_BF16 : PAR_SGNS_FP32, o.device, rounding};
    par_sgns_handle h = nullptr;
    if (par_sgns_create(&cfg, &h) != 0) return fail("device", h ? par_sgns_last_error(h) : "invalid configuration");
    std::vector<float> et(1001);
    ph_exp_table(et.data());
    std::vector<float> w_in((size_t)V * o.size), w_out((size_t)V * o.size);
    ph_init_net(V, o.size, w_in.data(), w_out.data());
    if (par_sgns_set_exp_table(h, et.data(), 1001) || par_sgns_set_weights(h, w_in.data(), w_out.data()))
        return fail("upload", par_sgns_last_error(h));
    std::vector<int64_t> entries((size_t)V);
    if (ph_unigram_entries(counts.data(), V, 100000000, entries.data()) < 0) return fail("unigram", ph_last_error());
    if (mode != PAR_SGNS_DETERMINISTIC && par_sgns_set_sampler(h, entries.data(), V))
        return fail("sampler", par_sgns_last_error(h));

    double train_s = 0.0;
    par_sgns_train_args ta{};
    ta.mode = mode;
    ta.iter = o.iter;
    ta.alpha = o.alpha;
    ta.train_words = T;
    ta.seed = o.seed;
    ta.use_graph = mode == PAR_SGNS_PERSISTENT;
    ta.min_sentences_per_wave = o.min_sent_wave;
    if (mode != PAR_SGNS_DETERMINISTIC) {
        const char *wb[] = {"atomic", "store_out", "store_in", "store", "hot"};
        ta.write_back = -1;
        for (int k = 0; k < 5; k++)
            if (o.write_back == wb[k]) ta.write_back = k;
        if (ta.write_back < 0) return fail("-write-back", "atomic, store_out, store_in, store or hot");
        ta.hot_rows = o.hot_rows;
    }
    par_sgns_stats st{};
    if (mode == PAR_SGNS_DETERMINISTIC) {
        ph_plan *pl = ph_plan_build(stream.data(), n, counts.data(), V, T, o.window, o.negative, o.batch, o.iter,
                                    o.alpha, o.sample);
        if (!pl) return fail("schedule", ph_last_error());
        ph_plan_info pi{};
        ph_plan_info_get(pl, &pi);
        std::vector<int32_t> tok((size_t)pi.n_tokens), negs((size_t)pi.n_negs + 1);
        std::vector<uint8_t> b((size_t)pi.n_tokens);
        std::vector<int64_t> so((size_t)pi.n_sent + 1), bo((size_t)pi.n_sent + 1), no((size_t)pi.n_batches + 1);
        std::vector<float> sa((size_t)pi.n_sent * 2);
        ph_plan_copy(pl, tok.data(), b.data(), so.data(), sa.data(), bo.data(), no.data(), negs.data());
        ph_plan_free(pl);
        if (par_sgns_set_plan(h, tok.data(), b.data(), pi.n_tokens, so.data(), sa.data(), pi.n_sent, bo.data(),
                              no.data(), pi.n_batches, negs.data(), pi.n_negs))
            return fail("plan", par_sgns_last_error(h));
        ta.sent_begin = 0;
        ta.sent_end = pi.n_sent;
        auto t0 = std::chrono::steady_clock::now();
        if (par_sgns_train(h, &ta, &st)) return fail("train", par_sgns_last_error(h));
        train_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    } else {
        std::vector<int32_t> tok((size_t)n);
        std::vector<int64_t> off((size_t)n / 2 + 2), raw((size_t)n / 2 + 2);
        for (int ep = 0; ep < o.iter; ep++) {
            int64_t nt = 0;
            // every kept sentence holds >= 1 token and ends at a 0 or at 1000 tokens
            int64_t S = ph_subsample(stream.data(), n, counts.data(), V, T, o.sample, o.seed * 1000003ULL + (uint64_t)ep,
                                     1000, tok.data(), (int64_t)tok.size(), off.data(), raw.data(),
                                     (int64_t)off.size() - 1, &nt);
            if (S < 0) return fail("subsample", ph_last_error());
            if (par_sgns_set_corpus(h, tok.data(), nt, off.data(), raw.data(), S))
                return fail("corpus", par_sgns_last_error(h));
            ta.epoch = ep;
            if (ep == 0 && o.autotune > 0 && mode == PAR_SGNS_PERSISTENT && S > 0) {  // untimed
                int64_t e = 1;
                while (e < S && raw[e] - raw[0] < o.chunk_words) e++;
                par_sgns_train_args tt = ta;
                tt.sent_begin = 0;
                tt.sent_end = e;
                par_sgns_tune_point best{};
                int32_t np = 0;
                if (par_sgns_autotune(h, &tt, o.autotune, nullptr, 0, &np, &best))
                    return fail("autotune", par_sgns_last_error(h));
                std::printf("launch shape (%d shapes timed): %d waves per block, %d per CU, %d sentences per dequeue, "
                            "%.3f ms on the first chunk\n", np, best.waves_per_block, best.waves_per_cu,
                            best.sentences_per_dequeue, best.kernel_ms);
            }
            auto t0 = std::chrono::steady_clock::now();
            int64_t b = 0;
            uint64_t tokens = 0;
            while (b < S) {  // chunks of ~chunk_words raw words
                int64_t e = b + 1;
                while (e < S && raw[e] - raw[b] < o.chunk_words) e++;
                ta.sent_begin = b;
                ta.sent_end = e;
                if (par_sgns_train(h, &ta, &st)) return fail("train", par_sgns_last_error(h));
                if (o.debug > 1 && b == 0)
                    std::printf("launch shape: %d blocks x %d waves, %d sentences per dequeue\n", st.blocks,
                                st.waves_per_block, st.sentences_per_dequeue);
                tokens += st.tokens;
                b = e;
            }
            double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            train_s += s;
            if (o.debug > 1)
                std::printf("epoch %d: %llu tokens, %.3f s, %.2f Mwords/s\n", ep + 1, (unsigned long long)tokens, s,
                            (double)n / s / 1e6);
        }
    }
    if (par_sgns_get_weights(h, w_in.data(), nullptr)) return fail("download", par_sgns_last_error(h));
    par_sgns_destroy(h);
    FILE *ft = std::fopen(o.time_file.c_str(), "w");
    if (ft) {
        std::fprintf(ft, "Elapsed %.2lf", train_s);
        std::fclose(ft);
    }
    if (ph_save_embeddings(o.output.c_str(), v, w_in.data(), o.size, o.binary) < 0) return fail("output", ph_last_error());
    ph_vocab_free(v);
    return 0;
}
