// mfdsgdpp / mfdsgd -- command lines of the GPU DSGD++ and DSGD paths,
// option-compatible with the reference tools (tools/mfdsgdpp.cc:379-441,
// positional arguments :406-412, exit code 0/1 :565; tools/mfdsgd.cc:50-233,
// which adds --map-reduce).  The program name selects the algorithm: invoked
// as "mfdsgd" it runs DSGD (d x d blocks).  Extra options: --precision
// (FP64_STRICT|FP64|FP32|BF16), --device, --kernel.  Everything runs through
// the C ABI in include/dsgdpp.h.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <map>
#include <string>
#include <vector>

#include "../../include/dsgdpp.h"

namespace {

int indexOf(const std::string& s, std::initializer_list<const char*> names) {
    int k = 0;
    for (const char* n : names) {
        if (s == n) return k;
        ++k;
    }
    return -1;
}

void usage(bool dsgd) {
    fprintf(stderr,
            "Error: Options input-file, input-row-file, input-col-file, update, regularize, loss, rank, "
            "and decay have to be present\n"
            "%s [options]\n\n"
            "Options:\n"
            "  --help                produce help message\n"
            "  --input-file          filename of data matrix\n"
            "  --input-test-file     filename of test matrix\n"
            "  --input-row-file      filename of initial row factors\n"
            "  --input-col-file      filename of initial column factors\n"
            "  --output-row-file     filename of final row factors\n"
            "  --output-col-file     filename of final column factors\n"
            "  --trace               filename of trace [trace.R]\n"
            "  --trace-var           variable name for trace [traceVar]\n"
            "  --epochs              number of epochs to run [10]\n"
            "  --tasks-per-rank      number of concurrent tasks per rank [1]\n"
            "  --sgd-order           order of SGD steps [WOR] (SEQ, WR, WOR)\n"
            "  --stratum-order       order of strata [COWOR] (SEQ, RSEQ, WR, WOR, COWOR)\n"
            "  --seed                seed for random number generator (system time if not set)\n"
            "  --rank                rank of factorization\n"
            "  --update              SGD update function (Nzsl, Nzsl_L2(l), Nzsl_Nzl2(l))\n"
            "  --regularize          SGD regularization function (None)\n"
            "  --loss                loss function (Nzsl, Nzsl_L2(l), Nzsl_Nzl2(l))\n"
            "  --abs                 absolute values after every step (%s)\n"
            "  --truncate            truncation, e.g. --truncate \"(-1000,1000)\"\n"
            "  --decay               decay function (Const(e), BoldDriver(e), BoldDriver(e,dec,inc))\n"
            "  --balance             type of balancing (None, L2, NZL2)\n"
            "  --precision           FP64_STRICT (default), FP64, FP32 or BF16\n"
            "  --device              HIP device ordinal [0]\n"
            "  --kernel              epoch kernel: 0 auto, 1 one launch per stratum, 2 dataflow,\n"
            "                        3 grouped, 4 grouped with H streamed from HBM, 5 level-parallel\n"
            "  --block-cache         directory of binary block caches: a matrix file's blocked form is\n"
            "                        kept there and used while the file and the blocking are unchanged\n"
            "%s",
            dsgd ? "mfdsgd" : "mfdsgdpp",
            dsgd ? "not supported by the GPU path" : "rejected, as in the reference",
            dsgd ? "  --map-reduce          whether to use the (slower) MapReduce implementation [false]\n"
                   "                        (on one GPU both give the same updates)\n"
                 : "");
}

bool endsWith(const std::string& s, const std::string& t) {
    return s.size() >= t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0;
}

}  // namespace

int main(int argc, char** argv) {
    const std::string self = argc > 0 ? argv[0] : "";
    const bool dsgd = endsWith(self, "/mfdsgd") || self == "mfdsgd";
    std::map<std::string, std::string> opt;
    std::vector<std::string> pos;
    bool help = false, absFlag = false;
    static const char* known[] = {"input-file", "input-test-file", "input-row-file", "input-col-file",
                                  "output-row-file", "output-col-file", "trace", "trace-var", "epochs",
                                  "tasks-per-rank", "sgd-order", "stratum-order", "seed", "rank",
                                  "update", "regularize", "loss", "truncate", "decay", "balance",
                                  "precision", "device", "kernel", "block-cache"};
    for (int a = 1; a < argc; ++a) {
        std::string s = argv[a];
        if (s.rfind("--", 0) != 0) {
            pos.push_back(s);
            continue;
        }
        std::string k = s.substr(2), v;
        if (k == "help") { help = true; continue; }
        if (k == "abs") { absFlag = true; continue; }
        const size_t eq = k.find('=');
        if (eq != std::string::npos) {
            v = k.substr(eq + 1);
            k = k.substr(0, eq);
        } else if (a + 1 < argc) {
            v = argv[++a];
        } else {
            fprintf(stderr, "Error: the option '--%s' requires an argument\n", k.c_str());
            return 1;
        }
        bool ok = dsgd && k == "map-reduce";
        for (const char* n : known) ok = ok || k == n;
        if (!ok) {
            fprintf(stderr, "Error: unrecognised option '--%s'\n", k.c_str());
            return 1;
        }
        opt[k] = v;
    }
    static const char* posNames[] = {"input-file", "input-test-file", "input-row-file",
                                     "input-col-file", "output-row-file", "output-col-file"};
    for (size_t k = 0; k < pos.size() && k < 6; ++k) opt[posNames[k]] = pos[k];
    for (const char* r : {"input-file", "input-row-file", "input-col-file", "update", "regularize", "loss",
                          "rank", "decay"})
        if (!opt.count(r)) help = true;
    if (help) {
        usage(dsgd);
        return 1;
    }
    if (opt.count("map-reduce") && opt["map-reduce"] != "true" && opt["map-reduce"] != "false" &&
        opt["map-reduce"] != "1" && opt["map-reduce"] != "0") {
        fprintf(stderr, "Error: the argument ('%s') for option '--map-reduce' is invalid\n", opt["map-reduce"].c_str());
        return 1;
    }
    auto get = [&](const char* k, const char* def) { return opt.count(k) ? opt[k] : std::string(def); };
    const int sgd = indexOf(get("sgd-order", "WOR"), {"SEQ", "WR", "WOR"});
    const int str = indexOf(get("stratum-order", "COWOR"), {"SEQ", "RSEQ", "WR", "WOR", "COWOR"});
    const int bal = indexOf(get("balance", "None"), {"None", "L2", "NZL2"});
    const int prec = indexOf(get("precision", "FP64_STRICT"), {"FP64_STRICT", "FP64", "FP32", "BF16"});
    if (sgd < 0) { fprintf(stderr, "Invalid arguments for sgdOrder; expected \"SEQ\", \"WR\" or \"WOR\"\n"); return 1; }
    if (str < 0) { fprintf(stderr, "Invalid arguments for stratumOrder; expected \"SEQ\", \"RSEQ\", \"WR\", \"WOR\" or \"COWOR\"\n"); return 1; }
    if (bal < 0) { fprintf(stderr, "Invalid arguments for balance; expected \"None\", \"L2\" or \"Nzl2\"\n"); return 1; }
    if (prec < 0) { fprintf(stderr, "Invalid arguments for precision\n"); return 1; }
    // mfdsgdpp's runArgs accepts no --abs combination (mfdsgdpp.cc:165, :353);
    // mfdsgd's generated combinations do, but not on this GPU path
    if (absFlag) {
        fprintf(stderr, "Invalid combination of update, regularize, and loss arguments\n");
        return 1;
    }
    const uint64_t seed = opt.count("seed") ? strtoull(opt["seed"].c_str(), nullptr, 10) : (uint64_t)time(nullptr);
    std::string trunc = get("truncate", "");
    dsgdpp_cfg cfg{};
    cfg.abi_version = DSGDPP_ABI_VERSION;
    const std::string upd = opt["update"], reg = opt["regularize"], los = opt["loss"];
    cfg.update = upd.c_str();
    cfg.regularize = reg.c_str();
    cfg.loss = los.c_str();
    cfg.truncate = trunc.c_str();
    cfg.rank = atoll(opt["rank"].c_str());
    cfg.world_size = 1;
    cfg.tasks_per_rank = atoi(get("tasks-per-rank", "1").c_str());
    cfg.sgd_order = sgd;
    cfg.stratum_order = str;
    cfg.seed = seed;
    cfg.precision = prec;
    cfg.device = atoi(get("device", "0").c_str());
    cfg.kernel = atoi(get("kernel", "0").c_str());
    cfg.algorithm = dsgd ? DSGDPP_ALGO_DSGD : DSGDPP_ALGO_DSGDPP;
    printf(dsgd ? "DSGD options\n" : "DSGD++ options\n");
    printf("    Seed: %llu\n    Epochs: %s\n    Factorization rank: %s\n"
           "    Update function: %s\n    Regularize function: %s\n    Loss function: %s\n    Decay: %s\n"
           "    Tasks per rank: %d\n    Precision: %s\n",
           (unsigned long long)seed, get("epochs", "10").c_str(), opt["rank"].c_str(), upd.c_str(), reg.c_str(),
           los.c_str(), opt["decay"].c_str(), cfg.tasks_per_rank, get("precision", "FP64_STRICT").c_str());
    dsgdpp_ctx* ctx = nullptr;
    if (dsgdpp_create(&cfg, &ctx) != DSGDPP_OK) {
        fprintf(stderr, "%s\n", dsgdpp_create_error());
        return 1;
    }
    // with --block-cache DIR, DIR/<file name>.<d>.blk holds the blocked form of each matrix file
    const std::string cacheDir = get("block-cache", "");
    auto load = [&](int which, const std::string& file) {
        if (cacheDir.empty()) return dsgdpp_load_mmc(ctx, which, file.c_str());
        const size_t slash = file.find_last_of('/');
        const std::string cache = cacheDir + "/" + (slash == std::string::npos ? file : file.substr(slash + 1)) +
                                  "." + std::to_string(cfg.tasks_per_rank) + (dsgd ? ".dsgd" : "") + ".blk";
        int32_t hit = 0;
        const int r = dsgdpp_load_mmc_cached(ctx, which, file.c_str(), cache.c_str(), &hit);
        if (r == 0) printf("%s %s\n", hit ? "Read block cache" : "Wrote block cache", cache.c_str());
        return r;
    };
    int rc = load(DSGDPP_TRAIN, opt["input-file"]);
    const std::string testFile = get("input-test-file", "");
    if (rc == 0 && !testFile.empty()) rc = load(DSGDPP_TEST, testFile);
    if (rc == 0) rc = dsgdpp_load_factors_mma(ctx, opt["input-row-file"].c_str(), opt["input-col-file"].c_str());
    const std::string trace = get("trace", ""), traceVar = get("trace-var", "trace"), decay = opt["decay"];
    if (rc == 0) {
        dsgdpp_run_cfg run{};
        run.epochs = atoll(get("epochs", "10").c_str());
        run.decay = decay.c_str();
        run.balance = bal;
        run.verbose = 1;
        run.trace_path = trace.empty() ? nullptr : trace.c_str();
        run.trace_var = traceVar.c_str();
        run.input_file = opt["input-file"].c_str();
        rc = dsgdpp_run(ctx, &run);
        if (rc == 0 && !trace.empty()) printf("Writing trace to %s\n", trace.c_str());
    }
    if (rc == 0) {
        const std::string ow = get("output-row-file", ""), oh = get("output-col-file", "");
        rc = dsgdpp_write_factors_mma(ctx, ow.c_str(), oh.c_str());
    }
    if (rc != 0) fprintf(stderr, "%s\n", dsgdpp_last_error(ctx));
    dsgdpp_destroy(ctx);
    return rc == 0 ? 0 : 1;
}
