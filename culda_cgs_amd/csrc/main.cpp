// main.cpp -- the `culda` command line (reference: src_culda/main.cpp:40-194).
//
//   culda -k <topics> -t <iterations> -a <alpha> -b <beta> -c <chunks>
//         -i <input prefix> [-o <output prefix>] [-g <gpus>] [-s <ignored>]
//         [--seed N] [--json] [--slot-tokens N] [--fused-phi 0|1]
//         [--checkpoint PATH] [--resume PATH]
//
// Same flags and per-iteration line as the reference.  Differences: every
// flag has a default (the reference leaves omitted fields uninitialised,
// culda_argument.h:50-67), -a/-b accept a fraction such as 50/1024 (atof
// read "50/1024" as 50), -g is clamped to the one device the library drives,
// and -o writes the phi and theta save files.
#include <cctype>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include "../../include/culda.h"

namespace {

struct Argument {
    int numGPUs = 1;
    int k = 1024;
    int iteration = 100;
    int numWorkers = 1;
    int numChunks = 1;
    float alpha = 50.0f / 1024.0f;
    float beta = 0.01f;
    std::string inputFilePrefix;
    std::string outputFilePrefix;
    unsigned long long seed = 1;
    bool json = false;
    int slotTokens = 0;
    int fusedPhi = 1;
    std::string checkpoint, resume;
};

bool is_uint(const char *s) {
    if (!*s) return false;
    for (; *s; ++s)
        if (!isdigit((unsigned char)*s)) return false;
    return true;
}

float parse_real(const char *flag, const char *s) {
    char *end = nullptr;
    double v = strtod(s, &end);
    if (end == s) throw std::invalid_argument(std::string(flag) + " should be followed by a number");
    if (*end == '/') {
        const char *q = end + 1;
        double d = strtod(q, &end);
        if (end == q || d == 0.0) throw std::invalid_argument(std::string(flag) + ": bad fraction");
        v /= d;
    }
    if (*end) throw std::invalid_argument(std::string(flag) + " should be followed by a number");
    return (float)v;
}

int parse_uint(const char *flag, const char *s) {
    if (!is_uint(s)) throw std::invalid_argument(std::string(flag) + " should be followed by a positive integer");
    return atoi(s);
}

Argument ParseArgument(int argc, char **argv) {
    if (argc == 1) throw std::invalid_argument("No argument found");
    Argument a;
    for (int i = 1; i < argc; ++i) {
        std::string f = argv[i];
        auto need = [&](const char *what) -> const char * {
            if (i + 1 >= argc) throw std::invalid_argument(std::string("need to specify ") + what + " after " + f);
            return argv[++i];
        };
        if (f == "-g") a.numGPUs = parse_uint("-g", need("a value"));
        else if (f == "-k") a.k = parse_uint("-k", need("a value"));
        else if (f == "-t") a.iteration = parse_uint("-t", need("a value"));
        else if (f == "-s") a.numWorkers = parse_uint("-s", need("a value"));
        else if (f == "-a") a.alpha = parse_real("-a", need("a value"));
        else if (f == "-b") a.beta = parse_real("-b", need("a value"));
        else if (f == "-c") a.numChunks = parse_uint("-c", need("a value"));
        else if (f == "-i") a.inputFilePrefix = need("a file name prefix");
        else if (f == "-o") a.outputFilePrefix = need("a file name prefix");
        else if (f == "--seed") a.seed = strtoull(need("a value"), nullptr, 10);
        else if (f == "--json") a.json = true;
        else if (f == "--slot-tokens") a.slotTokens = parse_uint("--slot-tokens", need("a value"));
        else if (f == "--fused-phi") a.fusedPhi = parse_uint("--fused-phi", need("a value"));
        else if (f == "--checkpoint") a.checkpoint = need("a path");
        else if (f == "--resume") a.resume = need("a path");
        else throw std::invalid_argument("unknown option " + f);
    }
    if (a.inputFilePrefix.empty()) throw std::invalid_argument("an input prefix (-i) is required");
    if (a.k % 32 != 0) {
        printf("Warning: number of topics(k) has been rounded to multiples of 32.\n");
        a.k = (a.k + 31) / 32 * 32;
    }
    if (a.numChunks < 1) throw std::invalid_argument("-c must be at least 1");
    if (a.numGPUs != 1) {
        printf("Warning: number of GPUs(%d) is larger than device count(1), rounded to 1\n", a.numGPUs);
        a.numGPUs = 1;
    }
    if (!(a.alpha > 0) || !(a.beta > 0)) throw std::invalid_argument("alpha and beta must be positive");
    return a;
}

int die(const char *what, int st) {
    fprintf(stderr, "Error: %s failed (%d): %s\n", what, st, culda_last_error());
    return 2;
}

}  // namespace

int main(int argc, char **argv) {
    Argument arg;
    printf("Parsing arguments ...\n");
    try {
        arg = ParseArgument(argc, argv);
    } catch (std::invalid_argument &e) {
        printf("Error: %s\n", e.what());
        return 1;
    }
    printf("numGPUs     :%d\nk           :%d\niteration   :%d\nnumWorkers  :%d\nnumChunks   :%d\n", arg.numGPUs,
           arg.k, arg.iteration, arg.numWorkers, arg.numChunks);
    printf("alpha       :%.6f\nbeta        :%.6f\nprefix      :%s\noutfile     :%s\n\n", arg.alpha, arg.beta,
           arg.inputFilePrefix.c_str(), arg.outputFilePrefix.c_str());

    printf("Loading document ...\n");
    culda_corpus *corpus = nullptr;
    int st = culda_corpus_load(arg.inputFilePrefix.c_str(), arg.numChunks, &corpus);
    if (st) return die("loading the corpus", st);
    culda_corpus_info info;
    culda_corpus_info_get(corpus, &info);
    printf("numDocs  :  %d\nnumWords :  %d\nnumTokens:  %lld\n\n", info.num_docs, info.num_words,
           (long long)info.num_tokens);

    culda_params p{};
    p.k = arg.k;
    p.alpha = arg.alpha;
    p.beta = arg.beta;
    p.seed = arg.seed;
    p.device = 0;
    p.slot_tokens = arg.slotTokens;
    p.fused_phi = arg.fusedPhi;
    culda_ctx *ctx = nullptr;
    if ((st = culda_create(&p, &ctx))) return die("creating the device context", st);
    if ((st = culda_upload(ctx, corpus))) return die("uploading the corpus", st);
    int first = 1;
    if (!arg.resume.empty()) {
        if ((st = culda_load_checkpoint(ctx, arg.resume.c_str(), &first))) return die("loading the checkpoint", st);
        printf("Resumed from %s at iteration %d\n", arg.resume.c_str(), first);
    } else {
        printf("Initialize the topic for tokens ...\n\n");
        if ((st = culda_init_z(ctx))) return die("initialising z", st);
    }
    if ((st = culda_rebuild_all(ctx))) return die("building the initial model", st);

    double elapsed = 0;
    for (int ite = first; ite < first + arg.iteration; ++ite) {
        double ll = 0, stamp = 0;
        if ((st = culda_iterate(ctx, ite, &ll, &stamp))) return die("iteration", st);
        elapsed += stamp;
        if (arg.json)
            printf("{\"iteration\": %d, \"elapsed_s\": %.6f, \"iter_s\": %.6f, \"loglik_per_token\": %.10f, "
                   "\"mtokens_per_s\": %.3f}\n",
                   ite, elapsed, stamp, ll, info.num_tokens / stamp / 1e6);
        else
            printf("Iteration %3d: %6.2f sec, %3.2f sec, logLikelyhood = %.8f, %5.3f M\n", ite, elapsed, stamp, ll,
                   info.num_tokens / stamp / 1e6);
        fflush(stdout);
    }
    if (!arg.checkpoint.empty() && (st = culda_save_checkpoint(ctx, arg.checkpoint.c_str(), first + arg.iteration)))
        return die("saving the checkpoint", st);
    if (!arg.outputFilePrefix.empty()) {
        if ((st = culda_save_phi(ctx, arg.outputFilePrefix.c_str()))) return die("saving phi", st);
        if ((st = culda_save_theta(ctx, arg.outputFilePrefix.c_str()))) return die("saving theta", st);
    }
    culda_destroy(ctx);
    culda_corpus_free(corpus);
    return 0;
}
