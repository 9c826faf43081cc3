/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 * mfdsgdpp-compatible command line over the CPU restatement (oracle.h): the
 * timed CPU baseline.  Options, defaults and exit codes follow
 * tools/mfdsgdpp.cc:379-441,565; --threads is extra (worker threads; results
 * do not depend on it).
 */
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <map>
#include <string>
#include <vector>

#include "oracle.h"

static int orderOf(const std::string& s, const char* const* names, int n) {
    for (int k = 0; k < n; ++k)
        if (s == names[k]) return k;
    return -1;
}

int main(int argc, char** argv) {
    std::map<std::string, std::string> opt;
    std::vector<std::string> pos;
    static const char* known[] = {"input-file", "input-test-file", "input-row-file", "input-col-file",
                                  "output-row-file", "output-col-file", "trace", "trace-var", "epochs",
                                  "tasks-per-rank", "sgd-order", "stratum-order", "seed", "rank",
                                  "update", "regularize", "loss", "truncate", "decay", "balance",
                                  "threads", "algorithm", "map-reduce"};
    bool help = false, absFlag = false;
    for (int a = 1; a < argc; ++a) {
        std::string s = argv[a];
        if (s.rfind("--", 0) == 0) {
            std::string k = s.substr(2), v;
            size_t eq = k.find('=');
            if (k == "help") { help = true; continue; }
            if (k == "abs") { absFlag = true; continue; }
            if (eq != std::string::npos) { v = k.substr(eq + 1); k = k.substr(0, eq); }
            else if (a + 1 < argc) v = argv[++a];
            else { fprintf(stderr, "Error: missing value for --%s\n", k.c_str()); return 1; }
            bool ok = false;
            for (auto* n : known) ok |= k == n;
            if (!ok) { fprintf(stderr, "Error: unrecognised option --%s\n", k.c_str()); return 1; }
            opt[k] = v;
        } else {
            pos.push_back(s);
        }
    }
    static const char* posNames[] = {"input-file", "input-test-file", "input-row-file",
                                     "input-col-file", "output-row-file", "output-col-file"};
    for (size_t k = 0; k < pos.size() && k < 6; ++k) opt[posNames[k]] = pos[k];
    const char* req[] = {"input-file", "input-row-file", "input-col-file", "update", "regularize",
                         "loss", "rank", "decay"};
    bool missing = false;
    for (auto* r : req) missing |= !opt.count(r);
    if (help || missing) {
        fprintf(stderr, "Error: Options input-file, input-row-file, input-col-file, update, regularize, "
                        "loss, rank, and decay have to be present\n");
        return 1;
    }
    uint64_t seed = opt.count("seed") ? strtoull(opt["seed"].c_str(), nullptr, 10) : (uint64_t)time(nullptr);
    long long epochs = opt.count("epochs") ? atoll(opt["epochs"].c_str()) : 10;
    int tasks = opt.count("tasks-per-rank") ? atoi(opt["tasks-per-rank"].c_str()) : 1;
    int threads = opt.count("threads") ? atoi(opt["threads"].c_str()) : tasks;
    static const char* sgdNames[] = {"SEQ", "WR", "WOR"};
    static const char* strNames[] = {"SEQ", "RSEQ", "WR", "WOR", "COWOR"};
    int sgd = orderOf(opt.count("sgd-order") ? opt["sgd-order"] : "WOR", sgdNames, 3);
    int str = orderOf(opt.count("stratum-order") ? opt["stratum-order"] : "COWOR", strNames, 5);
    if (sgd < 0) { fprintf(stderr, "Invalid arguments for sgdOrder; expected \"SEQ\", \"WR\" or \"WOR\"\n"); return 1; }
    if (str < 0) { fprintf(stderr, "Invalid arguments for stratumOrder\n"); return 1; }
    std::string bal = opt.count("balance") ? opt["balance"] : "None";
    int balance = bal == "None" ? 0 : bal == "L2" ? 1 : bal == "NZL2" ? 2 : -1;
    if (balance < 0) { fprintf(stderr, "Invalid arguments for balance; expected \"None\", \"L2\" or \"Nzl2\"\n"); return 1; }
    if (absFlag) {
        fprintf(stderr, "Invalid combination of update, regularize, and loss arguments\n");
        return 1;
    }
    char err[512] = {0};
    orc_ctx* c = orc_create(opt["update"].c_str(), opt["regularize"].c_str(), opt["loss"].c_str(),
                            opt.count("truncate") ? opt["truncate"].c_str() : nullptr,
                            atoll(opt["rank"].c_str()), 1, tasks, sgd, str, seed, threads, err,
                            sizeof err);
    if (!c) { fprintf(stderr, "%s\n", err); return 1; }
    // --algorithm DSGD: the mfdsgd tool (tools/mfdsgd.cc: d x d blocks)
    if (opt.count("algorithm") && opt["algorithm"] != "DSGDPP" &&
        (opt["algorithm"] != "DSGD" || orc_set_algorithm(c, 1) != 0)) {
        fprintf(stderr, "Invalid arguments for algorithm; expected \"DSGDPP\" or \"DSGD\"\n");
        orc_destroy(c);
        return 1;
    }
    int rc = orc_load_mmc(c, 0, opt["input-file"].c_str());
    if (rc == 0 && opt.count("input-test-file") && !opt["input-test-file"].empty())
        rc = orc_load_mmc(c, 1, opt["input-test-file"].c_str());
    if (rc == 0) rc = orc_load_factors_mma(c, opt["input-row-file"].c_str(), opt["input-col-file"].c_str());
    if (rc == 0)
        rc = orc_run(c, epochs, opt["decay"].c_str(), balance,
                     opt.count("trace") ? opt["trace"].c_str() : nullptr,
                     opt.count("trace-var") ? opt["trace-var"].c_str() : "trace",
                     opt["input-file"].c_str(), nullptr, nullptr, nullptr, nullptr, 1);
    if (rc == 0)
        rc = orc_write_factors_mma(c, opt.count("output-row-file") ? opt["output-row-file"].c_str() : nullptr,
                                   opt.count("output-col-file") ? opt["output-col-file"].c_str() : nullptr);
    if (rc != 0) fprintf(stderr, "%s\n", orc_last_error(c));
    orc_destroy(c);
    return rc == 0 ? 0 : 1;
}
