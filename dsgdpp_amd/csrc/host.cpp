// Host side of dsgdpp_amd: the C ABI (include/dsgdpp.h), functor-string
// parsing, the d x 2d block layout, the per-epoch stratum schedule and task
// streams, the decay/loss/trace run loop, Matrix Market I/O and the binary
// block cache.  Device work is in kernels.hip, kernels_grouped.hip, ingest.hip
// and mlp.hip.
#include <hip/hip_runtime.h>
#include <sys/stat.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <fstream>
#include <map>
#include <mutex>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/dsgdpp.h"
#include "internal.h"
#include "philox.h"

using namespace dsgdpp;

namespace {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPC(call)                                                                      \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            throw Error(DSGDPP_E_DEVICE, std::string("HIP error: ") + hipGetErrorString(e_) + \
                                             " at " #call);                             \
    } while (0)

int hostThreads() {
    unsigned h = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(h, 16u));
}

int pow2_floor(int v) {
    int p = 1;
    while (2 * p <= v) p *= 2;
    return p;
}

// set in the workers of parallelForEach: work they call runs on that thread
thread_local bool tl_inWorker = false;

template <typename F>
void parallelFor(int64_t n, F f) {
    int th = (int)std::min<int64_t>(hostThreads(), std::max<int64_t>(n, 1));
    if (th <= 1 || tl_inWorker) { f(0, n); return; }
    std::vector<std::thread> ts;
    for (int k = 0; k < th; ++k) ts.emplace_back(f, n * k / th, n * (k + 1) / th);
    for (auto& t : ts) t.join();
}

// f(i) for i in [0, n) on the host threads, items taken in order from a shared
// counter; the first exception a worker caught is rethrown after all joined
template <typename F>
void parallelForEach(int64_t n, F f) {
    const int th = (int)std::min<int64_t>(hostThreads(), std::max<int64_t>(n, 1));
    if (th <= 1 || tl_inWorker) {
        for (int64_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<int64_t> next{0};
    std::exception_ptr err;
    std::mutex mu;
    std::vector<std::thread> ts;
    for (int k = 0; k < th; ++k)
        ts.emplace_back([&] {
            tl_inWorker = true;
            for (int64_t i; (i = next.fetch_add(1)) < n;) {
                try {
                    f(i);
                } catch (...) {
                    std::lock_guard<std::mutex> g(mu);
                    if (!err) err = std::current_exception();
                }
            }
        });
    for (auto& t : ts) t.join();
    if (err) std::rethrow_exception(err);
}

// ---------------------------------------------------------------------------
// functor strings: "Name(a,b,..)" or "Name" (tools/parse.h:39-61)
// ---------------------------------------------------------------------------
bool splitFunctor(const std::string& s, std::string& name, std::vector<double>& args) {
    args.clear();
    const size_t open = s.find('(');
    if (s.empty() || s.back() != ')' || open == std::string::npos) {
        name = s;
        return true;
    }
    name = s.substr(0, open);
    const std::string inner = s.substr(open + 1, s.size() - open - 2);
    std::stringstream ss(inner);
    std::string tok;
    while (std::getline(ss, tok, ',')) {
        char* end = nullptr;
        const double v = strtod(tok.c_str(), &end);
        if (end == tok.c_str() || *end != '\0') return false;
        args.push_back(v);
    }
    if (inner.empty() || inner.back() == ',') return false;   // lexical_cast of an empty token fails
    return true;
}

// computeDefaultBlockOffsets (mf/matrix/distributed_matrix_impl.cc:24-40):
// the first size%blocks blocks hold one extra row; out[blocks] = size.
std::vector<int64_t> defaultOffsets(int64_t size, int64_t blocks) {
    std::vector<int64_t> off(blocks + 1, 0);
    const int64_t base = size / blocks, extra = size % blocks;
    for (int64_t b = 1; b <= blocks; ++b) off[b] = off[b - 1] + base + (b - 1 < extra ? 1 : 0);
    return off;
}

// ---------------------------------------------------------------------------
// master stream (stands in for rg::Random32, see philox.h)
// ---------------------------------------------------------------------------
struct Stream {
    uint32_t k0 = 0, k1 = 0;
    uint64_t next = 0;
    uint64_t u64() { return stream_draw(k0, k1, next++); }
    uint64_t below(uint64_t b) { return bounded(u64(), b); }
    // rg::shuffle stand-in: Fisher-Yates from the back (n-1 draws)
    void shuffle(uint32_t* a, uint32_t n) {
        for (uint32_t i = n; i > 1; --i) std::swap(a[i - 1], a[below(i)]);
    }
};

// d x d DSGD schedule, row = sub-epoch, column = task (mf/sgd/dsgd_impl.h:246-325)
std::vector<uint32_t> dsgdSchedule(uint32_t w, uint32_t t, int order, Stream& rng) {
    const uint32_t d = w * t;
    std::vector<uint32_t> s((size_t)d * d);
    for (uint32_t e = 0; e < d; ++e)
        for (uint32_t id = 0; id < d; ++id) s[(size_t)e * d + id] = (e + id) % d;
    auto identity = [](uint32_t k) {
        std::vector<uint32_t> v(k);
        std::iota(v.begin(), v.end(), 0u);
        return v;
    };
    if (order == DSGDPP_STRATUM_SEQ) return s;
    if (order == DSGDPP_STRATUM_RSEQ) {
        std::vector<uint32_t> relabel = identity(d);
        rng.shuffle(relabel.data(), d);
        for (auto& v : s) v = relabel[v];
    } else if (order == DSGDPP_STRATUM_WR) {
        for (uint32_t e = 0; e < d; ++e) rng.shuffle(&s[(size_t)e * d], d);
    } else if (order == DSGDPP_STRATUM_WOR) {
        // mf::shuffle (mf/matrix/op/shuffle.h:37-49): rows by P_r, then columns by P_c
        std::vector<uint32_t> rowp = identity(d), colp = identity(d);
        rng.shuffle(rowp.data(), d);
        rng.shuffle(colp.data(), d);
        std::vector<uint32_t> out(s.size());
        for (uint32_t e = 0; e < d; ++e)
            for (uint32_t id = 0; id < d; ++id) out[(size_t)e * d + colp[id]] = s[(size_t)rowp[e] * d + id];
        s.swap(out);
    } else if (order == DSGDPP_STRATUM_COWOR) {
        std::vector<uint32_t> relabel = identity(d);
        rng.shuffle(relabel.data(), d);
        const std::vector<uint32_t> groups = dsgdSchedule(w, 1, DSGDPP_STRATUM_WOR, rng);
        for (uint32_t gi = 0; gi < w; ++gi)
            for (uint32_t gj = 0; gj < w; ++gj) {
                const uint32_t goff = groups[(size_t)gi * w + gj] * t;
                const std::vector<uint32_t> local = dsgdSchedule(t, 1, DSGDPP_STRATUM_WOR, rng);
                for (uint32_t ti = 0; ti < t; ++ti)
                    for (uint32_t tj = 0; tj < t; ++tj)
                        s[(size_t)(gi * t + ti) * d + gj * t + tj] = relabel[goff + local[(size_t)ti * t + tj]];
            }
    } else {
        throw Error(DSGDPP_E_ARG, "Invalid stratum order");
    }
    return s;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct HostBlocks {
    std::vector<int64_t> start;   // d*2d+1
    int64_t nnz = 0;
    bool loaded = false;
};

// Everything one epoch needs besides eps and the factors.
struct EpochPlan {
    bool drawn = false, built = false, grouped = false, dataflow = false, level = false;
    uint64_t dataVersion = 0;
    std::vector<uint32_t> schedule;   // 2d x d
    std::vector<uint64_t> seeds;      // d
    std::vector<int64_t> voff;        // 2d x d
    std::vector<int64_t> goff;        // 2d x d: first group of each visit (grouped epoch)
    std::vector<int64_t> loff;        // 2d x d: first level start of each visit (level epoch)
    std::vector<int32_t> need;        // 2d x d
    std::vector<PermJob> jobs;
    std::vector<int32_t> chains;
    int64_t maxBlock = 0;             // largest block visited
    // grouped prep: chains in size classes (largest block of the chain up to
    // each bound of prep_classes()), one launch per class; class k holds
    // chains [clsStart[k], clsStart[k + 1]) with blocks of at most clsCap[k]
    std::vector<int32_t> clsStart;
    std::vector<int64_t> clsCap;
    uint32_t* dSched = nullptr;
    int64_t* dVoff = nullptr;
    int64_t* dGoff = nullptr;
    int32_t* dNeed = nullptr;
    PermJob* dJobs = nullptr;
    int32_t* dChains = nullptr;
    uint32_t* perm = nullptr;
    Pt32* grec = nullptr;
    int32_t* ngroups = nullptr;
    LevelBuffers LB;                  // level epoch: records by level, level starts, prep scratch
    int64_t* dLoff = nullptr;
    size_t jobsCap = 0, chainsCap = 0;
    int64_t permCap = 0, groupCap = 0, levelCap = 0;
    hipEvent_t evStart = nullptr, evEnd = nullptr;
};

}  // namespace

struct dsgdpp_ctx {
    std::string err;
    std::string updateStr, regularizeStr, lossStr, truncateStr;
    UpdParams upd;
    double lossLambda = 0;
    int64_t r = 0;
    uint32_t w = 1, t = 1, d = 1;
    int sgdOrder = DSGDPP_SGD_WOR, stratumOrder = DSGDPP_STRATUM_COWOR;
    uint64_t seed = 0;
    int precision = DSGDPP_FP64_STRICT, device = 0, kernel = 0;
    // DSGD runs on the d x 2d machinery: its column block k is block 2k, every
    // odd column block is empty (zero columns), and its sub-epoch s is
    // sub-epoch 2s; the odd sub-epochs visit the empty blocks, which every
    // kernel skips and which draw nothing.
    int algorithm = DSGDPP_ALGO_DSGDPP;
    bool dsgd() const { return algorithm == DSGDPP_ALGO_DSGD; }
    Stream master;

    int64_t m = -1, n = -1;
    std::vector<int64_t> off1, off2;
    HostBlocks blocks[2];
    bool haveFactors = false;

    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    DevProblem P;
    void* dPts[2] = {nullptr, nullptr};
    int64_t* dBlk[2] = {nullptr, nullptr};
    void* dRid = nullptr;   // training points' row ranks in their blocks (grouped prep)
    double* dPartials = nullptr;
    double* hPinned = nullptr;
    int32_t* dReady = nullptr;
    // dataflow launches are ordinary launches: their waits give up after a
    // time limit and set *dStatus, which the epoch turns into an error
    int32_t* dStatus = nullptr;
    int32_t* hStatus = nullptr;   // pinned
    int wallKhz = 0;              // wall-clock ticks per millisecond
    bool dataflowOk = false;
    bool groupedOk = false;
    std::string groupedWhy;
    // level epoch (kernel 5): chosen at load for blocks with few points per
    // row and column, or forced; never inside a batch (batches run the
    // grouped kernel)
    bool levelOk = false;
    std::string levelWhy;
    LevelIndex LX;
    int levelK = 1;                // workgroups per task
    int32_t* dLevelBar = nullptr;  // d level barrier counters
    int batches = 0;
    GroupedBuffers G;
    hipStream_t prepStream = nullptr;
    EpochPlan plan[2];
    int cur = 0, lastPlan = -1;
    bool lookahead = true;
    bool kprof = false;
    uint64_t dataVersion = 0;
    int64_t maxBlock = 0;   // largest training block
    // two-slot groups (FP32 tile kernel, rank <= 64): per block, set at load
    int pairsMode = -1;                 // DSGDPP_PAIRS: 0 never, 1 every eligible block, else by the chain rule
    std::vector<uint8_t> pairBlk;       // d x 2d, empty when no block qualifies
    uint8_t* dPairs = nullptr;
    int64_t pairBlocks = 0;

    size_t factorBytes() const {
        return precision <= DSGDPP_FP64 ? 8 : precision == DSGDPP_FP32 ? 4 : 2;
    }
    bool fp64() const { return precision <= DSGDPP_FP64; }
};

namespace {

thread_local std::string g_createError;

// Upper bounds of the grouped prep's size classes (points in a chain's largest
// block); DSGDPP_PREP_CLASSES="b1,b2,..." (ascending) replaces them.
const std::vector<int64_t>& prep_classes() {
    static const std::vector<int64_t> bounds = [] {
        std::vector<int64_t> b;
        if (const char* e = getenv("DSGDPP_PREP_CLASSES")) {
            const char* q = e;
            while (*q) {
                char* end = nullptr;
                const long long v = strtoll(q, &end, 10);
                if (end == q) break;
                if (v > 0 && (b.empty() || v > b.back())) b.push_back(v);
                q = *end == ',' ? end + 1 : end;
            }
            return b;
        }
        return std::vector<int64_t>(std::begin(PREP_CLASS_BOUNDS), std::end(PREP_CLASS_BOUNDS));
    }();
    return bounds;
}

template <typename F>
int guarded(dsgdpp_ctx* c, F f) {
    if (!c) return DSGDPP_E_ARG;
    try {
        HIPC(hipSetDevice(c->device));
        f();
        return DSGDPP_OK;
    } catch (const Error& e) {
        c->err = e.what();
        return e.code;
    } catch (const std::invalid_argument& e) {
        c->err = e.what();
        return DSGDPP_E_ARG;
    } catch (const std::bad_alloc&) {
        c->err = "out of host memory";
        return DSGDPP_E_DEVICE;
    } catch (const std::exception& e) {
        c->err = e.what();
        return DSGDPP_E_DEVICE;
    }
}

// Time limit of the dataflow waits, in wall-clock ticks: DSGDPP_WAIT_LIMIT_MS
// (default 30,000 ms; a legitimate wait lasts at most one epoch), at least
// one tick.
uint32_t dataflowWaitTicks(int wallKhz) {
    double ms = 30000.0;
    if (const char* e = getenv("DSGDPP_WAIT_LIMIT_MS")) ms = strtod(e, nullptr);
    const double ticks = ms * (double)wallKhz;
    return (uint32_t)std::min(4.0e9, std::max(1.0, ticks));
}

const char* kGaveUp =
    "a dataflow wait gave up (a workgroup of the launch was not resident, e.g. another kernel that waits on its "
    "own workgroups held the device)";

void dfree(void*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}
template <typename T>
void dalloc(T*& p, size_t bytes) {
    void* q = nullptr;
    if (p) (void)hipFree(p);
    p = nullptr;
    HIPC(hipMalloc(&q, std::max<size_t>(bytes, 16)));
    p = (T*)q;
}

void configure(dsgdpp_ctx* c, const dsgdpp_cfg* cfg) {
    if (!cfg) throw Error(DSGDPP_E_ARG, "null configuration");
    if (cfg->abi_version != DSGDPP_ABI_VERSION) throw Error(DSGDPP_E_ARG, "ABI version mismatch");
    c->updateStr = cfg->update ? cfg->update : "";
    c->regularizeStr = cfg->regularize ? cfg->regularize : "";
    c->lossStr = cfg->loss ? cfg->loss : "";
    c->truncateStr = cfg->truncate ? cfg->truncate : "";
    std::string un, rn, ln;
    std::vector<double> ua, ra, la;
    if (!splitFunctor(c->updateStr, un, ua)) throw Error(DSGDPP_E_ARG, "Invalid argument in " + c->updateStr);
    if (!splitFunctor(c->regularizeStr, rn, ra)) throw Error(DSGDPP_E_ARG, "Invalid argument in " + c->regularizeStr);
    if (!splitFunctor(c->lossStr, ln, la)) throw Error(DSGDPP_E_ARG, "Invalid argument in " + c->lossStr);
    // accepted combinations: tools/mfdsgdpp.cc:162-358
    const char* combos = "Invalid combination of update, regularize, and loss arguments";
    if (rn != "None") throw Error(DSGDPP_E_ARG, combos);
    if (!ra.empty()) throw Error(DSGDPP_E_ARG, "Invalid number of arguments in " + c->regularizeStr);
    if (un == "Nzsl" && ln == "Nzsl") {
        if (!ua.empty()) throw Error(DSGDPP_E_ARG, "Invalid number of arguments in " + c->updateStr);
        if (!la.empty()) throw Error(DSGDPP_E_ARG, "Invalid number of arguments in " + c->lossStr);
        c->upd.kind = UPD_NZSL;
    } else if ((un == "Nzsl_L2" && ln == "Nzsl_L2") || (un == "Nzsl_Nzl2" && ln == "Nzsl_Nzl2")) {
        if (ua.size() != 1) throw Error(DSGDPP_E_ARG, "Invalid number of arguments in " + c->updateStr);
        if (la.size() != 1) throw Error(DSGDPP_E_ARG, "Invalid number of arguments in " + c->lossStr);
        c->upd.kind = un == "Nzsl_L2" ? UPD_NZSL_L2 : UPD_NZSL_NZL2;
        c->upd.lambda = ua[0];
        c->lossLambda = la[0];
    } else {
        throw Error(DSGDPP_E_ARG, combos);
    }
    c->upd.truncate = false;
    if (!c->truncateStr.empty()) {   // parseTruncate (parse.h:115-139)
        std::string tn;
        std::vector<double> ta;
        if (!splitFunctor(c->truncateStr, tn, ta) || !tn.empty() || ta.size() != 2)
            throw Error(DSGDPP_E_ARG, "Error in argument truncate with options " + c->truncateStr);
        c->upd.truncate = true;
        c->upd.tmin = ta[0];
        c->upd.tmax = ta[1];
    }
    if (cfg->rank <= 0) throw Error(DSGDPP_E_ARG, "rank must be positive");
    if (cfg->rank > 256) throw Error(DSGDPP_E_UNSUPPORTED, "rank above 256 is not supported by the GPU path");
    if (cfg->world_size <= 0 || cfg->tasks_per_rank <= 0) throw Error(DSGDPP_E_ARG, "invalid blocking");
    if (cfg->sgd_order < 0 || cfg->sgd_order > 2) throw Error(DSGDPP_E_ARG, "Invalid sgd order");
    if (cfg->stratum_order < 0 || cfg->stratum_order > 4) throw Error(DSGDPP_E_ARG, "Invalid stratum order");
    if (cfg->precision < 0 || cfg->precision > 3) throw Error(DSGDPP_E_ARG, "Invalid precision");
    c->r = cfg->rank;
    c->w = (uint32_t)cfg->world_size;
    c->t = (uint32_t)cfg->tasks_per_rank;
    c->d = c->w * c->t;
    c->sgdOrder = cfg->sgd_order;
    c->stratumOrder = cfg->stratum_order;
    c->seed = cfg->seed;
    c->precision = cfg->precision;
    c->device = cfg->device;
    c->kernel = cfg->kernel;
    if (c->kernel < 0 || c->kernel > 5) throw Error(DSGDPP_E_ARG, "kernel must be 0..5");
    if (cfg->algorithm != DSGDPP_ALGO_DSGDPP && cfg->algorithm != DSGDPP_ALGO_DSGD)
        throw Error(DSGDPP_E_ARG, "algorithm must be DSGDPP_ALGO_DSGDPP or DSGDPP_ALGO_DSGD");
    c->algorithm = cfg->algorithm;
    master_key(c->seed, &c->master.k0, &c->master.k1);
    c->master.next = 0;
}

// ---------------------------------------------------------------------------
// blocking: ReadMatrixBlocksMm (read_impl.h:365-486) puts each entry into the
// block of its (row, column) with block-local indices and sorts every block
// row-major (read_impl.h:163; ublas sort sums duplicates).  Here the device
// does it (ingest.hip): blocks sorted stably, duplicates summed in input
// order, the result one array of 16-byte point records, block after block
// (d x 2d, row-major), with global indices.
// ---------------------------------------------------------------------------

// Blocks whose visits run two-slot groups (FP32 on the tile kernel, rank <= 64).
// A visit's groups number at least its busiest column's point count c, and at
// least n/4 (four slots) or n/2 (two slots) for n points; a two-slot group
// costs about 45 instructions against 54, so two slots win when
// 45 max(c, n/2) < 54 max(c, n/4), i.e. c > 5n/12.  The prep and the epoch
// kernel read the same flags, and a four-slot loop would also run a two-slot
// visit correctly (its slots 2 and 3 hold dummy records).
void choosePairs(dsgdpp_ctx* c) {
    c->pairBlk.clear();
    c->pairBlocks = 0;
    c->G.pairs = nullptr;
    if (c->precision != DSGDPP_FP32 || !c->groupedOk || c->G.wide || c->P.ld > 64 || c->pairsMode == 0 ||
        c->G.max_cols > 16384)
        return;
    const std::vector<int64_t>& start = c->blocks[0].start;
    const size_t nb = start.size() - 1;
    std::vector<int32_t> top;
    if (c->pairsMode != 1)
        device_block_top(false, c->dPts[0], c->dBlk[0], c->off2.data(), (int64_t)c->off2.size() - 1, (int64_t)nb,
                         c->G.max_cols, top, c->stream);
    std::vector<uint8_t> flag(nb, 0);
    for (size_t b = 0; b < nb; ++b) {
        const int64_t n = start[b + 1] - start[b];
        if (n > 0) flag[b] = c->pairsMode == 1 || 12 * (int64_t)top[b] > 5 * n ? 1 : 0;
        c->pairBlocks += flag[b];
    }
    if (c->pairBlocks == 0) return;
    c->pairBlk.swap(flag);
    dalloc(c->dPairs, c->pairBlk.size());
    HIPC(hipMemcpy(c->dPairs, c->pairBlk.data(), c->pairBlk.size(), hipMemcpyHostToDevice));
    c->G.pairs = c->dPairs;
}

// Level epoch (kernel 5): a block runs a level at a time, one lane per point
// (kernels_level.hip), which pays when its rows and columns hold few points
// each.  Auto (kernel 0) takes it when the mean non-empty block holds at least
// LEVEL_MIN_MEAN points and LEVEL_DEPTH_RATIO times the longest row or column
// run of any block -- a lower bound of that block's level count.  A level
// costs a few microseconds whatever its size (a barrier of the task's
// workgroups and a round trip of the factor rows), a four-slot group of the
// grouped kernel about 0.12 us, so a block of n points and D levels is faster
// level by level when n / D is some hundreds or more.
constexpr double LEVEL_MIN_MEAN = 2048;
constexpr double LEVEL_DEPTH_RATIO = 256;

void freeLevelIndex(dsgdpp_ctx* c) {
    void* q;
    q = c->LX.corder; dfree(q);
    q = c->LX.cinfo; dfree(q);
    q = c->LX.rinfo; dfree(q);
    c->LX = LevelIndex{};
}

void chooseLevel(dsgdpp_ctx* c) {
    freeLevelIndex(c);
    c->levelOk = false;
    c->levelWhy.clear();
    if (c->kernel != 0 && c->kernel != 5) return;
    auto refuse = [&](const std::string& why) {
        c->levelWhy = why;
        if (c->kernel == 5) throw Error(DSGDPP_E_UNSUPPORTED, "level kernel: " + why);
    };
    const std::vector<int64_t>& start = c->blocks[0].start;
    const int64_t nb = (int64_t)start.size() - 1;
    int64_t nonempty = 0;
    for (int64_t b = 0; b < nb; ++b) nonempty += start[b + 1] > start[b] ? 1 : 0;
    const double mean = nonempty ? (double)c->blocks[0].nnz / (double)nonempty : 0.;
    std::string why;
    if (!level_supported(c->P, c->sgdOrder, &why)) return refuse(why);
    int K = level_tasks_width(c->P, c->upd, c->device);
    if (K < 1) return refuse("its d workgroups do not all fit on the device at once");
    if (const char* e = getenv("DSGDPP_LEVEL_K")) K = std::max(1, std::min(K, atoi(e)));
    if (c->kernel == 0 && mean < LEVEL_MIN_MEAN) return refuse("blocks too small");
    std::vector<int32_t> col0(c->off2.begin(), c->off2.end());
    int32_t* d0 = nullptr;
    dalloc(d0, col0.size() * sizeof(int32_t));
    std::vector<int32_t> runmax;
    try {
        HIPC(hipMemcpy(d0, col0.data(), col0.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        build_level_index(c->P, d0, nb, (int)(2 * c->d), c->LX, runmax, c->stream);
    } catch (...) {
        void* q = d0;
        dfree(q);
        freeLevelIndex(c);
        throw;
    }
    { void* q = d0; dfree(q); }
    int32_t deep = 0;
    for (int32_t v : runmax) deep = std::max(deep, v);
    if (c->kernel == 0 && (double)deep * LEVEL_DEPTH_RATIO > mean) {
        freeLevelIndex(c);
        return refuse("a row or column holds too many of a block's points");
    }
    c->levelK = K;
    c->levelOk = true;
}

// The derived state of a loaded matrix whose point records and block starts
// are on the device: for the training matrix the counts, 1/count arrays and
// the FP32 records' row factors (nnz12, mf/matrix/op/nnz.h:34-58), block
// extents, kernel eligibility and the two-slot flags.
void finishBlocks(dsgdpp_ctx* c, int which) {
    HostBlocks& hb = c->blocks[which];
    hb.loaded = true;
    if (which != DSGDPP_TRAIN) {
        c->P.test_pts = c->dPts[1];
        c->P.test_blk = c->dBlk[1];
        c->P.test_nnz = hb.nnz;
        return;
    }
    DerivedOut dv;
    try {
        device_derive(c->P, true, c->dPts[0], hb.nnz, dv, c->stream);
    } catch (...) {
        for (void* q : {dv.inv1, dv.inv2, (void*)dv.cnt1, (void*)dv.cnt2}) if (q) (void)hipFree(q);
        throw;
    }
    for (const void* q : {c->P.inv1, c->P.inv2, (const void*)c->P.cnt1, (const void*)c->P.cnt2})
        if (q) (void)hipFree(const_cast<void*>(q));
    c->P.inv1 = dv.inv1;
    c->P.inv2 = dv.inv2;
    c->P.cnt1 = dv.cnt1;
    c->P.cnt2 = dv.cnt2;
    c->P.pts = c->dPts[0];
    c->P.blk = c->dBlk[0];
    c->P.nnz = hb.nnz;
    // grouped epoch: block extents and eligibility
    int64_t maxBlock = 0;
    for (size_t b = 0; b + 1 < hb.start.size(); ++b) maxBlock = std::max(maxBlock, hb.start[b + 1] - hb.start[b]);
    int maxRows = 0, maxCols = 0;
    for (size_t b = 0; b + 1 < c->off1.size(); ++b) maxRows = std::max<int>(maxRows, (int)(c->off1[b + 1] - c->off1[b]));
    for (size_t b = 0; b + 1 < c->off2.size(); ++b) maxCols = std::max<int>(maxCols, (int)(c->off2[b + 1] - c->off2[b]));
    c->G.max_rows = maxRows;
    c->G.max_cols = maxCols;
    c->maxBlock = maxBlock;
    bool wide = false;
    c->groupedOk = grouped_supported(c->P, c->upd, maxBlock, maxRows, maxCols, c->device, &c->groupedWhy, &wide,
                                     c->kernel == 4);
    c->G.wide = wide;
    dfree(c->dRid);   // row ranks of the previous training matrix, if any
    c->G.rid = nullptr;
    c->G.prep_rows = 0;
    if (c->kernel == 4 && !c->groupedOk) throw Error(DSGDPP_E_UNSUPPORTED, "wide grouped kernel: " + c->groupedWhy);
    choosePairs(c);
    if (c->groupedOk) {
        // the prep's row tags are kept by row rank: one per row with points
        dalloc(c->dRid, (size_t)std::max<int64_t>(hb.nnz, 1) * sizeof(uint16_t));
        c->G.rid = (const uint16_t*)c->dRid;
        c->G.prep_rows = launch_row_ranks(c->dPts[0], c->fp64(), c->dBlk[0], (int64_t)hb.start.size() - 1,
                                          (uint16_t*)c->dRid, c->stream);
        std::vector<int32_t> o2(c->off2.begin(), c->off2.end());
        int32_t* d2p = const_cast<int32_t*>(c->G.off2);
        dalloc(d2p, o2.size() * sizeof(int32_t));
        HIPC(hipMemcpy(d2p, o2.data(), o2.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        c->G.off2 = d2p;
    }
    if (c->kernel == 3 && !c->groupedOk) throw Error(DSGDPP_E_UNSUPPORTED, "grouped kernel: " + c->groupedWhy);
    chooseLevel(c);
}

void uploadBlocks(dsgdpp_ctx* c, int which, int64_t nnz, const int32_t* I, const int32_t* J, const double* X) {
    IngestOut io;
    try {
        device_ingest(IngestGeom{c->m, c->n, c->d, c->dsgd()}, c->fp64(), which == DSGDPP_TRAIN, nnz, I, J, X, io,
                      c->stream);
    } catch (...) {
        if (io.pts) (void)hipFree(io.pts);
        if (io.start_dev) (void)hipFree(io.start_dev);
        throw;
    }
    dfree(c->dPts[which]);
    c->dPts[which] = io.pts;
    void* q = c->dBlk[which];
    dfree(q);
    c->dBlk[which] = io.start_dev;
    HostBlocks& hb = c->blocks[which];
    hb.start.swap(io.start);
    hb.nnz = io.nnz;
    finishBlocks(c, which);
}

// The checks and layout set-up every load of a matrix shares: dimensions
// agree with the loaded data, the block grid fits, the block offsets and, on
// the first load, the factor layout.
void beginLoad(dsgdpp_ctx* c, int which, int64_t m, int64_t n) {
    if (which != DSGDPP_TRAIN && which != DSGDPP_TEST) throw Error(DSGDPP_E_ARG, "which must be TRAIN or TEST");
    if (m <= 0 || n <= 0) throw Error(DSGDPP_E_ARG, "invalid matrix dimensions");
    if (m >= INT32_MAX || n >= INT32_MAX) throw Error(DSGDPP_E_UNSUPPORTED, "dimensions must fit in int32");
    if (c->m >= 0 && (c->m != m || c->n != n)) throw Error(DSGDPP_E_ARG, "matrix dimensions differ from the loaded data");
    if (m < (int64_t)c->d || n < (int64_t)(c->dsgd() ? c->d : 2 * c->d))
        throw Error(DSGDPP_E_ARG, c->dsgd() ? "matrix is smaller than the d x d block grid"
                                            : "matrix is smaller than the d x 2d block grid");
    const bool first = c->m < 0;
    HIPC(hipStreamSynchronize(c->prepStream));   // a look-ahead plan may still read the old blocks
    ++c->dataVersion;
    c->m = m;
    c->n = n;
    c->off1 = defaultOffsets(m, c->d);
    if (c->dsgd()) {   // DSGD's d column blocks at the even positions, empty odd blocks between
        const std::vector<int64_t> o = defaultOffsets(n, c->d);
        c->off2.assign(2 * (size_t)c->d + 1, 0);
        for (uint32_t k = 0; k < c->d; ++k) {
            c->off2[2 * k] = o[k];
            c->off2[2 * k + 1] = o[k + 1];
        }
        c->off2[2 * c->d] = n;
    } else {
        c->off2 = defaultOffsets(n, 2 * c->d);
    }
    if (first) {
        c->P.m = m;
        c->P.n = n;
        c->P.r = c->r;
        c->P.d = (int)c->d;
        c->P.precision = c->precision;
        // rows padded to 16 bytes; FP64 rows to 4 elements and rows above rank
        // 64 to 8 (the grouped kernel's lanes hold 4 or 8 elements of a row,
        // never part of the next)
        const int64_t vec = c->r > 64 ? std::max<int64_t>(8, 16 / (int64_t)c->factorBytes())
                                      : std::max<int64_t>(c->factorBytes() == 8 ? 4 : 1, 16 / (int64_t)c->factorBytes());
        c->P.ld = (c->r + vec - 1) / vec * vec;
        c->dataflowOk = dataflow_fits(c->P, c->device);
    }
}

void loadCoo(dsgdpp_ctx* c, int which, int64_t m, int64_t n, int64_t nnz, const int32_t* I,
             const int32_t* J, const double* X) {
    if (nnz < 0) throw Error(DSGDPP_E_ARG, "invalid matrix dimensions");
    if (nnz > 0 && (!I || !J || !X)) throw Error(DSGDPP_E_ARG, "null index or value array");
    for (int64_t p = 0; p < nnz; ++p)
        if (I[p] < 0 || I[p] >= m || J[p] < 0 || J[p] >= n)
            throw Error(DSGDPP_E_ARG, "index out of range at entry " + std::to_string(p));
    if (nnz >= (int64_t)UINT32_MAX) throw Error(DSGDPP_E_UNSUPPORTED, "too many nonzeros");
    beginLoad(c, which, m, n);
    uploadBlocks(c, which, nnz, I, J, X);
}

// factors: host doubles <-> device storage (W row-major, H column-major r x n)
uint16_t toBf16(float f) {
    uint32_t b;
    memcpy(&b, &f, 4);
    if ((b & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((b >> 16) | 0x40u);
    b += 0x7fffu + ((b >> 16) & 1u);
    return (uint16_t)(b >> 16);
}
float fromBf16(uint16_t u) {
    uint32_t b = (uint32_t)u << 16;
    float f;
    memcpy(&f, &b, 4);
    return f;
}

void putFactor(dsgdpp_ctx* c, void*& dst, const double* src, int64_t rows) {
    const int64_t r = c->r, ld = c->P.ld;
    const size_t es = c->factorBytes();
    // FACTOR_TAIL_ROWS spare rows after the last one (zero rows, then the dummy row of the grouped kernel)
    std::vector<uint8_t> buf((size_t)(rows + FACTOR_TAIL_ROWS) * ld * es, 0);
    for (int64_t i = 0; i < rows; ++i)
        for (int64_t z = 0; z < r; ++z) {
            const double v = src[i * r + z];
            uint8_t* p = buf.data() + ((size_t)i * ld + z) * es;
            if (es == 8) memcpy(p, &v, 8);
            else if (es == 4) { const float f = (float)v; memcpy(p, &f, 4); }
            else { const uint16_t h = toBf16((float)v); memcpy(p, &h, 2); }
        }
    if (!dst) dalloc(dst, buf.size());
    HIPC(hipMemcpy(dst, buf.data(), buf.size(), hipMemcpyHostToDevice));
}

void getFactor(dsgdpp_ctx* c, const void* src, double* dst, int64_t rows) {
    const int64_t r = c->r, ld = c->P.ld;
    const size_t es = c->factorBytes();
    std::vector<uint8_t> buf((size_t)rows * ld * es);   // the spare rows stay on the device
    HIPC(hipMemcpy(buf.data(), src, buf.size(), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < rows; ++i)
        for (int64_t z = 0; z < r; ++z) {
            const uint8_t* p = buf.data() + ((size_t)i * ld + z) * es;
            double v;
            if (es == 8) memcpy(&v, p, 8);
            else if (es == 4) { float f; memcpy(&f, p, 4); v = f; }
            else { uint16_t h; memcpy(&h, p, 2); v = fromBf16(h); }
            dst[i * r + z] = v;
        }
}

void requireReady(dsgdpp_ctx* c) {
    if (!c->blocks[0].loaded) throw Error(DSGDPP_E_STATE, "load the training matrix first");
    if (!c->haveFactors) throw Error(DSGDPP_E_STATE, "set the factors first");
}

// ---------------------------------------------------------------------------
// one epoch: DsgdPpRunner::epoch (mf/sgd/dsgdpp_impl.h:282-349)
//
// An epoch's plan -- schedule, task seeds, visit orders and groups -- depends
// only on the master stream and the data, not on eps or the factors.  So while
// epoch e runs on the main stream, the plan of epoch e+1 is drawn (in the same
// master-stream order as without look-ahead) and its visit-order kernel runs
// on a second stream; epoch e+1 then only waits for that kernel's event.
// ---------------------------------------------------------------------------
void drawPlan(dsgdpp_ctx* c, EpochPlan& P) {
    const uint32_t d = c->d, d2 = 2 * d;
    Stream& rng = c->master;
    if (c->dsgd()) {
        // DsgdRunner::epoch (dsgd_impl.h:329-350): one d x d schedule, then the
        // task seeds; DSGD's sub-epoch s and column block k are sub-epoch 2s
        // and block 2k here, and sub-epoch 2s + 1 visits the empty block 2k + 1
        const std::vector<uint32_t> s0 = dsgdSchedule(c->w, c->t, c->stratumOrder, rng);
        P.schedule.assign((size_t)d2 * d, 0);
        for (uint32_t i = 0; i < d; ++i)
            for (uint32_t id = 0; id < d; ++id) {
                P.schedule[(size_t)(2 * i) * d + id] = 2 * s0[(size_t)i * d + id];
                P.schedule[(size_t)(2 * i + 1) * d + id] = 2 * s0[(size_t)i * d + id] + 1;
            }
        P.seeds.resize(d);
        for (uint32_t id = 0; id < d; ++id) P.seeds[id] = rng.u64();
        P.drawn = true;
        P.built = false;
        return;
    }
    // column-block halves for even / odd sub-epochs (:289-323)
    std::vector<uint32_t> even(d), odd(d);
    if (c->stratumOrder == DSGDPP_STRATUM_SEQ) {
        for (uint32_t b = 0; b < d; ++b) { even[b] = 2 * b; odd[b] = 2 * b + 1; }
    } else {
        // rg::sample(d, 2d) stand-in: sorted, by selection sampling
        std::vector<uint8_t> chosen(d2, 0);
        uint32_t need = d;
        for (uint32_t b = 0; b < d2 && need > 0; ++b)
            if (rng.below(d2 - b) < need) { chosen[b] = 1; --need; }
        uint32_t e = 0, o = 0;
        for (uint32_t b = 0; b < d2; ++b) (chosen[b] ? even[e++] : odd[o++]) = b;
    }
    // two d x d schedules interleaved into 2d x d (:327-336)
    const std::vector<uint32_t> s0 = dsgdSchedule(c->w, c->t, c->stratumOrder, rng);
    const std::vector<uint32_t> s1 = dsgdSchedule(c->w, c->t, c->stratumOrder, rng);
    P.schedule.assign((size_t)d2 * d, 0);
    for (uint32_t i = 0; i < d; ++i)
        for (uint32_t id = 0; id < d; ++id) {
            P.schedule[(size_t)(2 * i) * d + id] = even[s0[(size_t)i * d + id]];
            P.schedule[(size_t)(2 * i + 1) * d + id] = odd[s1[(size_t)i * d + id]];
        }
    // task seeds (mpi2::seed, :344): one master draw per task, in task order
    P.seeds.resize(d);
    for (uint32_t id = 0; id < d; ++id) P.seeds[id] = rng.u64();
    P.drawn = true;
    P.built = false;
}

// Visit-order jobs, dataflow dependencies and the visit-order kernel of a
// drawn plan, all enqueued on `s`.  With launchPrep false the caller launches
// the grouped prep (one launch for many plans) and records P.evEnd.
// ahead: a look-ahead plan, whose prep runs beside the current epoch
void buildPlan(dsgdpp_ctx* c, EpochPlan& P, hipStream_t s, bool launchPrep = true, bool ahead = false) {
    const uint32_t d = c->d, d2 = 2 * d;
    const std::vector<int64_t>& bs = c->blocks[0].start;
    auto visitBlock = [&](uint32_t id, uint32_t sub) { return (size_t)id * d2 + P.schedule[(size_t)sub * d + id]; };
    // every (task, sub-epoch) visit gets its own run of the visit-order buffers:
    // a WR stratum order may send a task to the same block twice in one epoch
    P.voff.assign((size_t)d2 * d, 0);
    P.goff.assign((size_t)d2 * d, 0);
    P.loff.assign((size_t)d2 * d, 0);
    int64_t total = 0, gtotal = 0, ordinal = 0;
    for (uint32_t id = 0; id < d; ++id)
        for (uint32_t sub = 0; sub < d2; ++sub) {
            const size_t b = visitBlock(id, sub);
            const int64_t nb = bs[b + 1] - bs[b];
            P.voff[(size_t)sub * d + id] = total;
            P.goff[(size_t)sub * d + id] = gtotal;
            P.loff[(size_t)sub * d + id] = total + ordinal++;   // a visit's level starts: at most nb + 1
            total += nb;
            gtotal += (2 * nb + GROUP_PAD - 1) / GROUP_PAD * GROUP_PAD;   // groups < 2 points, padded
        }
    P.level = c->levelOk && c->batches == 0;
    P.grouped = !P.level && c->groupedOk && (c->kernel == 0 || c->kernel == 3 || c->kernel == 4);
    P.dataflow = P.level || P.grouped || (c->kernel != 1 && c->dataflowOk);
    std::vector<PermJob> jobs;
    std::vector<int32_t> chains;
    int64_t maxBlock = 0;
    if (c->sgdOrder != DSGDPP_SGD_SEQ || P.grouped || P.level) {
        // each task's stream is consumed block after block in sub-epoch order
        // (one SgdRunner per task per epoch, :112)
        std::vector<std::vector<PermJob>> perTask(d);
        parallelFor(d, [&](int64_t lo, int64_t hi) {
            for (int64_t id = lo; id < hi; ++id) {
                uint32_t k0, k1;
                task_key(P.seeds[id], &k0, &k1);
                uint64_t draw = 0;
                int64_t lastN = 0, lastBase = -1;
                for (uint32_t sub = 0; sub < d2; ++sub) {
                    if (c->dsgd() && (sub & 1)) continue;   // an empty column block no DSGD task visits
                    const size_t b = visitBlock((uint32_t)id, sub);
                    const int64_t nb = bs[b + 1] - bs[b];
                    const int64_t base = P.voff[(size_t)sub * d + id];
                    const uint32_t b2 = P.schedule[(size_t)sub * d + id];
                    const int32_t pairs = P.grouped && !c->pairBlk.empty() && c->pairBlk[b] ? 1 : 0;
                    PermJob j{base, nb, -1, draw, k0, k1, bs[b], (int32_t)c->off1[id], (int32_t)c->off2[b2],
                              (int32_t)(sub * d + id), pairs, P.goff[(size_t)sub * d + id]};
                    if (c->sgdOrder == DSGDPP_SGD_WR) {
                        if (nb > 0) perTask[id].push_back(j);
                        draw += (uint64_t)nb + 2;
                    } else if (c->sgdOrder == DSGDPP_SGD_SEQ) {
                        if (nb > 0) perTask[id].push_back(j);
                    } else {
                        if (nb == 0) continue;
                        j.carry_base = (nb == lastN) ? lastBase : -1;
                        perTask[id].push_back(j);
                        draw += (uint64_t)nb;
                        lastN = nb;
                        lastBase = base;
                    }
                }
            }
        });
        for (uint32_t id = 0; id < d; ++id)
            for (const PermJob& j : perTask[id]) {
                // a chain is a run of visits whose permutations carry over (WOR only)
                if (!(c->sgdOrder == DSGDPP_SGD_WOR && j.carry_base >= 0)) chains.push_back((int32_t)jobs.size());
                jobs.push_back(j);
                maxBlock = std::max(maxBlock, j.n);
            }
        chains.push_back((int32_t)jobs.size());
        P.clsStart.assign(1, 0);
        P.clsCap.clear();
        if (P.grouped) {
            // chains of small blocks first, by size class: a class's launches
            // need LDS for its largest block only, so more of them share a CU
            const std::vector<int64_t>& bounds = prep_classes();
            const int ncls = (int)bounds.size() + 1;
            auto cls_of = [&](int64_t mx) {
                int k = 0;
                while (k < ncls - 1 && mx > bounds[k]) ++k;
                return k;
            };
            std::vector<int64_t> mxs(chains.size() - 1);
            std::vector<int32_t> count(ncls, 0);
            for (size_t k = 0; k + 1 < chains.size(); ++k) {
                int64_t mx = 0;
                for (int32_t q = chains[k]; q < chains[k + 1]; ++q) mx = std::max(mx, jobs[q].n);
                mxs[k] = mx;
                ++count[cls_of(mx)];
            }
            std::vector<PermJob> j2;
            std::vector<int32_t> c2;
            j2.reserve(jobs.size());
            c2.reserve(chains.size());
            for (int cl = 0; cl < ncls; ++cl) {
                if (count[cl] == 0) continue;
                int64_t cap = 0;
                for (size_t k = 0; k + 1 < chains.size(); ++k) {
                    if (cls_of(mxs[k]) != cl) continue;
                    c2.push_back((int32_t)j2.size());
                    for (int32_t q = chains[k]; q < chains[k + 1]; ++q) j2.push_back(jobs[q]);
                    cap = std::max(cap, mxs[k]);
                }
                P.clsStart.push_back((int32_t)(c2.size()));
                P.clsCap.push_back(cap);
            }
            c2.push_back((int32_t)j2.size());
            jobs.swap(j2);
            chains.swap(c2);
        }
        if (total > P.permCap) {
            P.permCap = total;
            dalloc(P.perm, (size_t)total * sizeof(uint32_t));
        }
        if (P.grouped && gtotal > P.groupCap) {
            P.groupCap = gtotal;
            dalloc(P.grec, (size_t)gtotal * GROUP_SLOTS * sizeof(Pt32));
        }
        if (P.grouped && !P.ngroups) dalloc(P.ngroups, (size_t)d2 * d * sizeof(int32_t));
        if (P.level && total > P.levelCap) {
            P.levelCap = total;
            dalloc(P.LB.lrec, (size_t)total * sizeof(uint4));
            dalloc(P.LB.lstart, (size_t)(total + (int64_t)d2 * d) * sizeof(int32_t));
            dalloc(P.LB.pos, (size_t)total * sizeof(uint32_t));
            dalloc(P.LB.prv, (size_t)total * 2 * sizeof(int32_t));
            dalloc(P.LB.lev, (size_t)total * sizeof(int32_t));
            for (uint32_t*& f : P.LB.fy) dalloc(f, (size_t)(total + (int64_t)d2 * d) * sizeof(uint32_t));
        }
        if (P.level && !P.LB.nlev) dalloc(P.LB.nlev, (size_t)d2 * d * sizeof(int32_t));
        if (P.level && !P.dLoff) dalloc(P.dLoff, (size_t)d2 * d * sizeof(int64_t));
        P.LB.loff = P.dLoff;
    }
    if (jobs.size() > P.jobsCap) {
        P.jobsCap = jobs.size() * 2;
        dalloc(P.dJobs, P.jobsCap * sizeof(PermJob));
    }
    if (chains.size() > P.chainsCap) {
        P.chainsCap = chains.size() * 2;
        dalloc(P.dChains, P.chainsCap * sizeof(int32_t));
    }
    if (!P.dSched) dalloc(P.dSched, (size_t)d2 * d * sizeof(uint32_t));
    if (!P.dVoff) dalloc(P.dVoff, (size_t)d2 * d * sizeof(int64_t));
    if (P.grouped && !P.dGoff) dalloc(P.dGoff, (size_t)d2 * d * sizeof(int64_t));
    // dataflow dependencies: a visit to column block b2 at sub-epoch s waits for
    // the non-empty visits to b2 in sub-epochs < s
    P.need.assign((size_t)d2 * d, 0);
    if (P.dataflow) {
        std::vector<int32_t> cnt(d2, 0);
        for (uint32_t sub = 0; sub < d2; ++sub) {
            for (uint32_t id = 0; id < d; ++id) P.need[(size_t)sub * d + id] = cnt[P.schedule[(size_t)sub * d + id]];
            for (uint32_t id = 0; id < d; ++id) {
                const size_t b = visitBlock(id, sub);
                if (bs[b + 1] > bs[b]) cnt[P.schedule[(size_t)sub * d + id]]++;
            }
        }
        if (!P.dNeed) dalloc(P.dNeed, P.need.size() * sizeof(int32_t));
    }
    // host staging stays alive until the copies complete (the plan owns it)
    P.jobs.swap(jobs);
    P.chains.swap(chains);
    HIPC(hipMemcpyAsync(P.dSched, P.schedule.data(), P.schedule.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(P.dVoff, P.voff.data(), P.voff.size() * sizeof(int64_t), hipMemcpyHostToDevice, s));
    if (P.grouped)
        HIPC(hipMemcpyAsync(P.dGoff, P.goff.data(), P.goff.size() * sizeof(int64_t), hipMemcpyHostToDevice, s));
    if (P.level)
        HIPC(hipMemcpyAsync(P.dLoff, P.loff.data(), P.loff.size() * sizeof(int64_t), hipMemcpyHostToDevice, s));
    if (P.dataflow)
        HIPC(hipMemcpyAsync(P.dNeed, P.need.data(), P.need.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
    if (!P.jobs.empty())
        HIPC(hipMemcpyAsync(P.dJobs, P.jobs.data(), P.jobs.size() * sizeof(PermJob), hipMemcpyHostToDevice, s));
    if (!P.chains.empty())
        HIPC(hipMemcpyAsync(P.dChains, P.chains.data(), P.chains.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
    DevProblem Dp = c->P;
    Dp.perm = P.perm;
    GroupedBuffers G = c->G;
    G.grec = P.grec;
    G.ngroups = P.ngroups;
    G.pprof = nullptr;
    if (c->kprof && P.grouped) {   // prep clock sums, one set per plan slot
        if (!c->G.pprof) {
            unsigned long long* q = nullptr;
            dalloc(q, 8 * sizeof(unsigned long long));
            c->G.pprof = q;
        }
        G.pprof = c->G.pprof + (&P == &c->plan[0] ? 0 : 4);
        HIPC(hipMemsetAsync(G.pprof, 0, 4 * sizeof(unsigned long long), s));
    }
    const int nchains = (int)P.chains.size() - 1;
    P.maxBlock = maxBlock;
    HIPC(hipEventRecord(P.evStart, s));
    if (!launchPrep && P.grouped) {
        P.built = true;
        P.dataVersion = c->dataVersion;
        return;
    }
    if (P.level) {
        // the visit orders, then each visit's levels
        if (c->sgdOrder == DSGDPP_SGD_WOR) launch_permute_par(P.dJobs, P.dChains, nchains, P.dLoff, P.perm, P.LB, s);
        launch_level_prep(Dp, c->LX, P.LB, P.dJobs, (int)P.jobs.size(),
                          c->sgdOrder == DSGDPP_SGD_WOR ? P.perm : nullptr, s);
    } else if (P.grouped) {
        // blocks whose prep state does not fit in LDS keep it in a scratch buffer
        size_t spill = 0;
        for (size_t k = 0; k < P.clsCap.size(); ++k)
            spill = std::max(spill, prep_spill_bytes(P.clsCap[k], G.prep_rows, G.max_cols,
                                                     P.clsStart[k + 1] - P.clsStart[k]));
        if (spill > c->G.spill_bytes) {
            HIPC(hipStreamSynchronize(s));   // a prep of the other plan may still use the old one
            void* q = c->G.spill;
            dfree(q);
            c->G.spill = nullptr;
            dalloc(c->G.spill, spill);
            c->G.spill_bytes = spill;
        }
        G.spill = c->G.spill;
        G.spill_bytes = c->G.spill_bytes;
        for (size_t k = 0; k < P.clsCap.size(); ++k)
            launch_prep_grouped(Dp, G, P.dJobs, P.dChains + P.clsStart[k], P.clsStart[k + 1] - P.clsStart[k],
                                c->sgdOrder, P.clsCap[k], s, ahead);
    } else if (c->sgdOrder == DSGDPP_SGD_WOR)
        launch_permute_wor(P.dJobs, P.dChains, nchains, P.perm, maxBlock, s);
    else if (c->sgdOrder == DSGDPP_SGD_WR)
        launch_visit_wr(P.dJobs, (int)P.jobs.size(), P.perm, s);
    HIPC(hipEventRecord(P.evEnd, s));
    P.built = true;
    P.dataVersion = c->dataVersion;
}

void epoch(dsgdpp_ctx* c, double eps, dsgdpp_epoch_stats* st) {
    requireReady(c);
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t d = c->d, d2 = 2 * d;
    EpochPlan& P = c->plan[c->cur];
    if (!P.drawn) drawPlan(c, P);
    if (!P.built || P.dataVersion != c->dataVersion) {
        HIPC(hipStreamSynchronize(c->prepStream));
        buildPlan(c, P, c->prepStream);
        if (c->kprof) {
            const auto b1 = std::chrono::steady_clock::now();
            HIPC(hipEventSynchronize(P.evEnd));
            const auto b2 = std::chrono::steady_clock::now();
            fprintf(stderr, "[kprof] plan built in the epoch: draw+host %.2f ms, prep kernels waited %.2f ms\n",
                    std::chrono::duration<double, std::milli>(b1 - t0).count(),
                    std::chrono::duration<double, std::milli>(b2 - b1).count());
        }
    }
    HIPC(hipStreamWaitEvent(c->stream, P.evEnd, 0));
    if (P.dataflow && !c->dReady) dalloc(c->dReady, (size_t)d2 * sizeof(int32_t));
    HIPC(hipEventRecord(c->ev[1], c->stream));
    EpochLaunch L;
    L.sched = P.dSched;
    L.voff = P.dVoff;
    L.goff = P.dGoff;
    L.need = P.dNeed;
    L.ready = c->dReady;
    L.use_perm = c->sgdOrder == DSGDPP_SGD_SEQ ? 0 : 1;
    L.dataflow = P.dataflow ? 1 : 0;
    L.eps = eps;
    const uint32_t waitTicks = dataflowWaitTicks(c->wallKhz);
    if (P.dataflow) {
        L.status = c->dStatus;
        L.wait_ticks = waitTicks;
        *c->hStatus = 0;
    }
    DevProblem Dp = c->P;
    Dp.perm = P.perm;
    if (P.level) {
        if (!c->dLevelBar) dalloc(c->dLevelBar, (size_t)4 * d * sizeof(int32_t));
        LevelBuffers LB = P.LB;
        LB.bar = c->dLevelBar;
        LB.K = c->levelK;
        if (c->kprof) {
            if (!c->G.prof) {
                unsigned long long* q = nullptr;
                dalloc(q, (size_t)d * 5 * sizeof(unsigned long long));
                c->G.prof = q;
            }
            LB.prof = c->G.prof;
        }
        launch_epoch_levels(Dp, c->upd, LB, L, c->stream);
    } else if (P.grouped) {
        GroupedBuffers G = c->G;
        G.grec = P.grec;
        G.ngroups = P.ngroups;
        if (c->kprof && !c->G.prof) {
            unsigned long long* q = nullptr;
            dalloc(q, (size_t)d * 5 * sizeof(unsigned long long));
            c->G.prof = q;
        }
        G.prof = c->G.prof;
        launch_epoch_grouped(Dp, c->upd, G, L, c->stream);
    } else {
        launch_epoch(Dp, c->upd, L, c->stream);
    }
    if (P.dataflow)
        HIPC(hipMemcpyAsync(c->hStatus, c->dStatus, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipEventRecord(c->ev[2], c->stream));
    const int used = c->cur;
    c->lastPlan = used;
    P.drawn = false;   // consumed: the next use of this slot draws a new plan
    c->cur ^= 1;
    // look ahead: the next epoch's plan, drawn now in master-stream order
    if (c->lookahead) {
        EpochPlan& N = c->plan[c->cur];
        drawPlan(c, N);
        buildPlan(c, N, c->prepStream, true, true);
    }
    HIPC(hipEventSynchronize(c->ev[2]));
    const auto t1 = std::chrono::steady_clock::now();
    if (P.dataflow && *c->hStatus != 0)
        throw Error(DSGDPP_E_DEVICE, std::string("epoch failed: ") + kGaveUp + " after " +
                                         std::to_string(waitTicks / std::max(1, c->wallKhz)) +
                                         " ms; the factors are undefined until set again");
    if (c->kprof && P.grouped && c->G.pprof) {
        unsigned long long h[4];
        HIPC(hipMemcpy(h, c->G.pprof + (used == 0 ? 0 : 4), sizeof(h), hipMemcpyDeviceToHost));
        const double v = (double)std::max<unsigned long long>(1, h[3]);
        fprintf(stderr, "[kprof] prep clocks/visit: visit order %.3g scheduling %.3g fill %.3g (%llu visits)\n",
                h[0] / v, h[1] / v, h[2] / v, h[3]);
    }
    if (c->kprof && P.level && c->G.prof) {
        // wall-clock sums of the level epoch kernel (first workgroup of each task)
        std::vector<unsigned long long> h((size_t)d * 5);
        HIPC(hipMemcpy(h.data(), c->G.prof, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        double w = 0, l = 0, b = 0, n = 0, loc = 0;
        for (size_t t = 0; t < d; ++t) {
            w += h[5 * t]; l += h[5 * t + 1]; b += h[5 * t + 2]; n += h[5 * t + 3]; loc += h[5 * t + 4];
        }
        const double us = 1e3 / std::max(1, c->wallKhz);
        fprintf(stderr, "[kprof] level epoch, mean per task: dataflow wait %.1f us, levels %.1f us, barriers %.1f us, "
                        "%.0f levels (K = %d workgroups per task, %.0f of %u tasks on one XCD)\n", w / d * us, l / d * us,
                b / d * us, n / d, c->levelK, loc, d);
    }
    if (c->kprof && P.grouped && c->G.prof) {
        // clock counters of the grouped epoch kernel: the slowest task and the mean
        std::vector<unsigned long long> h((size_t)d * 5);
        HIPC(hipMemcpy(h.data(), c->G.prof, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        double mw = 0, ms = 0, ml = 0, mt = 0, mg = 0;
        size_t worst = 0;
        for (size_t t = 0; t < d; ++t) {
            mw += h[5 * t]; ms += h[5 * t + 1]; ml += h[5 * t + 2]; mt += h[5 * t + 3]; mg += h[5 * t + 4];
            if (h[5 * t + 2] > h[5 * worst + 2]) worst = t;
        }
        fprintf(stderr, "[kprof] mean clocks/task: wait %.3g (mover's dataflow wait %.3g) groups %.3g total %.3g; groups/task %.0f; "
                        "busiest task %zu: loop %.3g clocks, %llu groups (%.1f clocks/group)\n",
                mw / d, ms / d, ml / d, mt / d, mg / d, worst, (double)h[5 * worst + 2], h[5 * worst + 4],
                (double)h[5 * worst + 2] / std::max<unsigned long long>(1, h[5 * worst + 4]));
    }
    if (st) {
        float a = 0, b = 0;
        HIPC(hipEventSynchronize(P.evEnd));
        HIPC(hipEventElapsedTime(&a, P.evStart, P.evEnd));
        HIPC(hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
        st->seconds = std::chrono::duration<double>(t1 - t0).count();
        st->perm_seconds = a / 1e3;
        st->sgd_seconds = b / 1e3;
        st->device_seconds = b / 1e3;
        st->updates = c->blocks[0].nnz;
        st->launches = P.dataflow ? 1 : d2;
        st->kernel = P.level ? 5 : P.grouped ? (c->G.wide ? 4 : 3) : P.dataflow ? 2 : 1;
    }
}

double trainLoss(dsgdpp_ctx* c) {
    requireReady(c);
    double l = device_nzsl(c->P, 0, c->dPartials, c->hPinned, c->stream);
    if (c->upd.kind == UPD_NZSL_L2) {
        const double a = device_sumsq(c->P, 0, nullptr, 0, c->m, c->dPartials, c->hPinned, c->stream);
        const double b = device_sumsq(c->P, 1, nullptr, 0, c->n, c->dPartials, c->hPinned, c->stream);
        l = l + c->lossLambda * (a + b);   // SumLoss + L2Loss (loss.h:99, l2.h:165-167)
    } else if (c->upd.kind == UPD_NZSL_NZL2) {
        const double a = device_sumsq(c->P, 0, c->P.cnt1, 0, c->m, c->dPartials, c->hPinned, c->stream);
        const double b = device_sumsq(c->P, 1, c->P.cnt2, 0, c->n, c->dPartials, c->hPinned, c->stream);
        l = l + c->lossLambda * (a + b);
    }
    return l;
}

double testLoss(dsgdpp_ctx* c) {
    if (!c->blocks[1].loaded) throw Error(DSGDPP_E_STATE, "no test matrix loaded");
    if (!c->haveFactors) throw Error(DSGDPP_E_STATE, "set the factors first");
    return device_nzsl(c->P, 1, c->dPartials, c->hPinned, c->stream);
}

// balanceSimple (mf/matrix/op/balance_impl.cc:102-127)
void balance(dsgdpp_ctx* c, int type) {
    if (type == DSGDPP_BALANCE_NONE) return;
    if (type != DSGDPP_BALANCE_L2 && type != DSGDPP_BALANCE_NZL2) throw Error(DSGDPP_E_ARG, "invalid balance type");
    requireReady(c);
    const double* w1 = type == DSGDPP_BALANCE_NZL2 ? c->P.cnt1 : nullptr;
    const double* w2 = type == DSGDPP_BALANCE_NZL2 ? c->P.cnt2 : nullptr;
    const double regW = device_sumsq(c->P, 0, w1, 0, c->m, c->dPartials, c->hPinned, c->stream);
    const double regH = device_sumsq(c->P, 1, w2, 0, c->n, c->dPartials, c->hPinned, c->stream);
    double wf = std::sqrt(std::sqrt(regH / regW)), hf;
    if (std::isnan(wf)) wf = hf = 0;
    else hf = 1. / wf;
    device_scale(c->P, 0, wf, c->stream);
    device_scale(c->P, 1, hf, c->stream);
    HIPC(hipStreamSynchronize(c->stream));
}

// ---------------------------------------------------------------------------
// Matrix Market (read_impl.h:46-256, write_impl.h:94-118)
// ---------------------------------------------------------------------------
std::string readAll(const std::string& fn) {
    FILE* f = fopen(fn.c_str(), "rb");
    if (!f) throw Error(DSGDPP_E_IO, "Cannot open file " + fn);
    std::string s;
    if (fseeko(f, 0, SEEK_END) == 0) {   // a regular file: one allocation
        const off_t n = ftello(f);
        if (n > 0) s.reserve((size_t)n);
        fseeko(f, 0, SEEK_SET);
    }
    std::vector<char> buf(1 << 24);
    size_t got;
    while ((got = fread(buf.data(), 1, buf.size(), f)) > 0) s.append(buf.data(), got);
    const bool bad = ferror(f) != 0;
    fclose(f);
    if (bad) throw Error(DSGDPP_E_IO, "Cannot read file " + fn);
    return s;
}

struct LineReader {
    const std::string& s;
    size_t pos = 0;
    int64_t lineNo = 0;
    explicit LineReader(const std::string& t) : s(t) {}
    bool next(const char*& b, const char*& e) {
        if (pos >= s.size()) return false;
        size_t nl = s.find('\n', pos);
        if (nl == std::string::npos) nl = s.size();
        b = s.data() + pos;
        e = s.data() + nl;
        pos = nl + 1;
        ++lineNo;
        return true;
    }
};

bool onlySpace(const char* b, const char* e) {
    for (; b < e; ++b)
        if (!isspace((unsigned char)*b)) return false;
    return true;
}

// The data lines of a Matrix Market body, parsed on all host threads: the
// `count` lines from byte `start` (file line lineNo0 + 1 on) go to
// parse(p, line) (a null-terminated copy), which returns 0, or an error
// message prefix ("Parse error", "Index out of range"); lines past them must
// be blank.  The error reported is the sequential reader's: the first one in
// file order, then a short file.
template <typename F>
void parseBody(const std::string& text, size_t start, int64_t lineNo0, int64_t count, const std::string& fn,
               F parse) {
    const size_t end = text.size();
    const int th = hostThreads();
    std::vector<size_t> cut(th + 1, end);
    cut[0] = std::min(start, end);
    for (int k = 1; k < th; ++k) {   // chunks begin at line starts
        size_t c = std::max(cut[k - 1], cut[0] + (end - cut[0]) / th * k);
        if (c > cut[0] && c < end) {
            const void* nl = memchr(text.data() + c - 1, '\n', end - (c - 1));
            c = nl ? (size_t)((const char*)nl - text.data()) + 1 : end;
        }
        cut[k] = c;
    }
    auto lines_in = [&](size_t a, size_t b) {   // LineReader's lines in [a, b)
        int64_t n = 0;
        for (const char* q = text.data() + a; (q = (const char*)memchr(q, '\n', text.data() + b - q)); ++q) ++n;
        return n + ((b > a && text[b - 1] != '\n') ? 1 : 0);
    };
    std::vector<int64_t> first(th + 1, 0);
    {
        std::vector<std::thread> ts;
        for (int k = 0; k < th; ++k) ts.emplace_back([&, k] { first[k + 1] = lines_in(cut[k], cut[k + 1]); });
        for (auto& t : ts) t.join();
    }
    for (int k = 0; k < th; ++k) first[k + 1] += first[k];
    std::vector<int64_t> errLine(th, INT64_MAX);
    std::vector<std::string> errMsg(th);
    {
        std::vector<std::thread> ts;
        for (int k = 0; k < th; ++k)
            ts.emplace_back([&, k] {
                std::string line;
                int64_t L = first[k];
                for (size_t pos = cut[k]; pos < cut[k + 1]; ++L) {
                    const char* nl = (const char*)memchr(text.data() + pos, '\n', cut[k + 1] - pos);
                    const size_t e = nl ? (size_t)(nl - text.data()) : cut[k + 1];
                    const char* msg = nullptr;
                    if (L < count) {
                        line.assign(text.data() + pos, e - pos);
                        msg = parse(L, line);
                    } else if (!onlySpace(text.data() + pos, text.data() + e)) {
                        msg = "Unexpected input";
                    }
                    if (msg) {
                        errLine[k] = lineNo0 + L + 1;
                        errMsg[k] = msg;
                        return;
                    }
                    pos = e + 1;
                }
            });
        for (auto& t : ts) t.join();
    }
    int best = -1;
    for (int k = 0; k < th; ++k)
        if (errLine[k] != INT64_MAX && (best < 0 || errLine[k] < errLine[best])) best = k;
    if (best >= 0)
        throw Error(DSGDPP_E_IO, errMsg[best] + " at line " + std::to_string(errLine[best]) + " of " + fn);
    if (first[th] < count) throw Error(DSGDPP_E_IO, "Unexpected EOF in file " + fn);
}

void readHeader(LineReader& lr, const std::string& fn, const char* banner, std::string& dims) {
    const char *b, *e;
    if (!lr.next(b, e)) throw Error(DSGDPP_E_IO, "Unexpected EOF in file " + fn);
    std::string first(b, e);
    while (!first.empty() && isspace((unsigned char)first.back())) first.pop_back();
    if (first != banner) throw Error(DSGDPP_E_IO, "Wrong matrix-market banner in file " + fn + ": " + std::string(b, e));
    while (true) {
        if (!lr.next(b, e)) throw Error(DSGDPP_E_IO, "Unexpected EOF in file " + fn);
        if (b < e && *b == '%') continue;
        dims.assign(b, e);
        return;
    }
}

void loadMmc(dsgdpp_ctx* c, int which, const std::string& fn) {
    const std::string text = readAll(fn);
    LineReader lr(text);
    std::string dims;
    readHeader(lr, fn, "%%MatrixMarket matrix coordinate real general", dims);
    long long m, n, nnz;
    char junk[2];
    if (sscanf(dims.c_str(), "%lld %lld %lld %1s", &m, &n, &nnz, junk) != 3 || m <= 0 || n <= 0 || nnz < 0)
        throw Error(DSGDPP_E_IO, "Invalid matrix dimensions in file " + fn + ": " + dims);
    std::vector<int32_t> I(nnz), J(nnz);
    std::vector<double> X(nnz);
    parseBody(text, lr.pos, lr.lineNo, nnz, fn, [&](int64_t p, const std::string& line) -> const char* {
        const char* s = line.c_str();
        char* q = nullptr;
        const unsigned long i = strtoul(s, &q, 10);
        if (q == s) return "Parse error";
        s = q;
        const unsigned long j = strtoul(s, &q, 10);
        if (q == s) return "Parse error";
        s = q;
        const double x = strtod(s, &q);
        if (q == s || !onlySpace(q, line.c_str() + line.size())) return "Parse error";
        if (i < 1 || j < 1 || (long long)i > m || (long long)j > n) return "Index out of range";
        I[p] = (int32_t)(i - 1);
        J[p] = (int32_t)(j - 1);
        X[p] = x;
        return nullptr;
    });
    loadCoo(c, which, m, n, nnz, I.data(), J.data(), X.data());
}

// column-major values of a "matrix array" file
void readMma(const std::string& fn, int64_t& rows, int64_t& cols, std::vector<double>& cm) {
    const std::string text = readAll(fn);
    LineReader lr(text);
    std::string dims;
    readHeader(lr, fn, "%%MatrixMarket matrix array real general", dims);
    long long a, b2;
    char junk[2];
    if (sscanf(dims.c_str(), "%lld %lld %1s", &a, &b2, junk) != 2 || a <= 0 || b2 <= 0)
        throw Error(DSGDPP_E_IO, "Invalid matrix dimensions in file " + fn + ": " + dims);
    rows = a;
    cols = b2;
    cm.resize((size_t)a * b2);
    parseBody(text, lr.pos, lr.lineNo, (int64_t)cm.size(), fn, [&](int64_t p, const std::string& line) -> const char* {
        char* q = nullptr;
        const double x = strtod(line.c_str(), &q);
        if (q == line.c_str() || !onlySpace(q, line.c_str() + line.size())) return "Parse error";
        cm[p] = x;
        return nullptr;
    });
}

void writeMma(const std::string& fn, int64_t rows, int64_t cols, const std::vector<double>& cm) {
    std::ofstream out(fn);
    if (!out.is_open()) throw Error(DSGDPP_E_IO, "Cannot open file " + fn);
    out.precision(16);   // digits10 + 1 (write_impl.h:99)
    out << "%%MatrixMarket matrix array real general\n"
        << "% First line: ROWS COLUMNS\n"
        << "% Subsequent lines: entries in column-major order\n"
        << rows << " " << cols << "\n";
    for (double v : cm) out << v << "\n";
    if (!out) throw Error(DSGDPP_E_IO, "Cannot write file " + fn);
}

// ---------------------------------------------------------------------------
// Binary block cache: a loaded matrix in the device layout -- blocked, sorted,
// duplicates summed -- with the blocking it was made for, so a later load of
// the same data skips parsing, sorting and blocking.  Little-endian:
//   BlockCacheHeader, int64 start[d * 2d + 1], 16-byte point records[nnz].
// ---------------------------------------------------------------------------
struct BlockCacheHeader {
    char magic[8];            // "DSGDBLK1"
    uint32_t version;         // 1
    uint32_t record_bytes;    // 16
    uint32_t fp64;            // records carry x as double (FP64 modes) or float
    uint32_t algorithm;       // DSGDPP_ALGO_*
    int64_t m, n;
    uint32_t d, reserved;
    int64_t nnz;              // point records
    int64_t src_size;         // the Matrix Market file it was read from (0: none)
    int64_t src_mtime_ns;
};
static_assert(sizeof(BlockCacheHeader) == 72, "cache header layout");

bool fileStamp(const std::string& fn, int64_t& size, int64_t& mtime_ns) {
    struct stat st;
    if (stat(fn.c_str(), &st) != 0) return false;
    size = (int64_t)st.st_size;
    mtime_ns = (int64_t)st.st_mtim.tv_sec * 1000000000LL + (int64_t)st.st_mtim.tv_nsec;
    return true;
}

void saveBlocks(dsgdpp_ctx* c, int which, const std::string& fn, int64_t src_size, int64_t src_mtime_ns) {
    if (which != DSGDPP_TRAIN && which != DSGDPP_TEST) throw Error(DSGDPP_E_ARG, "which must be TRAIN or TEST");
    const HostBlocks& hb = c->blocks[which];
    if (!hb.loaded) throw Error(DSGDPP_E_STATE, "matrix not loaded");
    BlockCacheHeader h{};
    memcpy(h.magic, "DSGDBLK1", 8);
    h.version = 1;
    h.record_bytes = 16;
    h.fp64 = c->fp64() ? 1 : 0;
    h.algorithm = (uint32_t)c->algorithm;
    h.m = c->m;
    h.n = c->n;
    h.d = c->d;
    h.nnz = hb.nnz;
    h.src_size = src_size;
    h.src_mtime_ns = src_mtime_ns;
    std::vector<char> recs((size_t)hb.nnz * 16);
    HIPC(hipStreamSynchronize(c->stream));
    if (hb.nnz > 0) HIPC(hipMemcpy(recs.data(), c->dPts[which], recs.size(), hipMemcpyDeviceToHost));
    const std::string tmp = fn + ".tmp";
    {
        std::ofstream out(tmp, std::ios::binary);
        if (!out.is_open()) throw Error(DSGDPP_E_IO, "Cannot open file " + tmp);
        out.write((const char*)&h, sizeof h);
        out.write((const char*)hb.start.data(), (std::streamsize)(hb.start.size() * sizeof(int64_t)));
        out.write(recs.data(), (std::streamsize)recs.size());
        if (!out) throw Error(DSGDPP_E_IO, "Cannot write file " + tmp);
    }
    if (rename(tmp.c_str(), fn.c_str()) != 0) throw Error(DSGDPP_E_IO, "Cannot write file " + fn);
}

bool readCacheHeader(const std::string& fn, BlockCacheHeader& h) {
    std::ifstream in(fn, std::ios::binary);
    if (!in.is_open()) return false;
    in.read((char*)&h, sizeof h);
    return (bool)in && memcmp(h.magic, "DSGDBLK1", 8) == 0 && h.version == 1 && h.record_bytes == 16;
}

template <typename Pt>
void checkCachedRecords(const dsgdpp_ctx* c, const std::vector<int64_t>& start, const std::vector<char>& raw,
                        const std::string& fn) {
    // every record inside its block and the block sorted row-major without
    // repeats, as the blocking leaves it; anything else is not a cache file
    const Pt* pts = (const Pt*)raw.data();
    const uint32_t d2 = 2 * c->d;
    for (size_t b = 0; b + 1 < start.size(); ++b) {
        const int64_t r0 = c->off1[b / d2], r1 = c->off1[b / d2 + 1], q0 = c->off2[b % d2], q1 = c->off2[b % d2 + 1];
        for (int64_t p = start[b]; p < start[b + 1]; ++p) {
            const Pt& e = pts[p];
            const bool in = e.i >= r0 && e.i < r1 && e.j >= q0 && e.j < q1;
            const bool order = p == start[b] || e.i > pts[p - 1].i || (e.i == pts[p - 1].i && e.j > pts[p - 1].j);
            if (!in || !order) throw Error(DSGDPP_E_IO, "Corrupt block cache " + fn);
        }
    }
}

void loadBlocks(dsgdpp_ctx* c, int which, const std::string& fn) {
    BlockCacheHeader h;
    if (!readCacheHeader(fn, h)) throw Error(DSGDPP_E_IO, "Not a block cache file: " + fn);
    if (h.fp64 != (c->fp64() ? 1u : 0u) || h.algorithm != (uint32_t)c->algorithm || h.d != c->d)
        throw Error(DSGDPP_E_ARG, "block cache " + fn + " was written for another blocking, algorithm or precision");
    if (h.nnz < 0 || h.nnz >= (int64_t)UINT32_MAX) throw Error(DSGDPP_E_IO, "Corrupt block cache " + fn);
    std::ifstream in(fn, std::ios::binary);
    in.seekg(sizeof h);
    const size_t nb = (size_t)c->d * 2 * c->d;
    std::vector<int64_t> start(nb + 1);
    in.read((char*)start.data(), (std::streamsize)(start.size() * sizeof(int64_t)));
    std::vector<char> raw((size_t)h.nnz * 16);
    in.read(raw.data(), (std::streamsize)raw.size());
    if (!in) throw Error(DSGDPP_E_IO, "Unexpected EOF in block cache " + fn);
    if (start[0] != 0 || start[nb] != h.nnz) throw Error(DSGDPP_E_IO, "Corrupt block cache " + fn);
    for (size_t b = 0; b < nb; ++b)
        if (start[b + 1] < start[b]) throw Error(DSGDPP_E_IO, "Corrupt block cache " + fn);
    beginLoad(c, which, h.m, h.n);
    if (c->fp64()) checkCachedRecords<Pt64>(c, start, raw, fn);
    else checkCachedRecords<Pt32>(c, start, raw, fn);
    void* pts = nullptr;
    dalloc(pts, raw.size());
    dfree(c->dPts[which]);
    c->dPts[which] = pts;
    if (!raw.empty()) HIPC(hipMemcpy(pts, raw.data(), raw.size(), hipMemcpyHostToDevice));
    dalloc(c->dBlk[which], start.size() * sizeof(int64_t));
    HIPC(hipMemcpy(c->dBlk[which], start.data(), start.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    HostBlocks& hb = c->blocks[which];
    hb.start.swap(start);
    hb.nnz = h.nnz;
    finishBlocks(c, which);
}

// load_mmc through the cache: the cache when it was written from this very
// file (size and modification time) with this blocking, else the file, then
// the cache (re)written
bool loadMmcCached(dsgdpp_ctx* c, int which, const std::string& fn, const std::string& cache) {
    int64_t size = 0, mtime = 0;
    if (!fileStamp(fn, size, mtime)) throw Error(DSGDPP_E_IO, "Cannot open file " + fn);
    BlockCacheHeader h;
    if (readCacheHeader(cache, h) && h.src_size == size && h.src_mtime_ns == mtime &&
        h.fp64 == (c->fp64() ? 1u : 0u) && h.algorithm == (uint32_t)c->algorithm && h.d == c->d) {
        loadBlocks(c, which, cache);
        return true;
    }
    loadMmc(c, which, fn);
    saveBlocks(c, which, cache, size, mtime);
    return false;
}

// ---------------------------------------------------------------------------
// R trace (mf/trace.h:46-262)
// ---------------------------------------------------------------------------
struct TraceEntry {
    int64_t it;
    double loss, test, eps, tEps, tEpoch, tLoss, tTest;   // times in ns
};

std::string formatTrace(const std::vector<TraceEntry>& es, const std::string& var,
                        const std::map<std::string, double>& nums,
                        const std::map<std::string, std::string>& strs) {
    std::stringstream o;
    o << var << " <- list(trace=list(";
    double elapsed = 0;
    for (size_t k = 0; k < es.size(); ++k) {
        const TraceEntry& e = es[k];
        o << (k ? ", " : "") << std::endl;
        o << "list(it=" << e.it << ", epoch=" << e.it << ", loss=" << e.loss << ", loss.test=";
        if (std::isnan(e.test)) o << "NA, ";
        else o << e.test << ", ";
        elapsed += (e.tEps + e.tEpoch + e.tLoss + e.tTest) / 1E9;
        o << "time=list(elapsed=" << elapsed << ", eps=" << e.tEps / 1E9 << ", epoch=" << e.tEpoch / 1E9
          << ", loss=" << e.tLoss / 1E9 << ", loss.test=" << e.tTest / 1E9 << "),  eps=";
        if (std::isnan(e.eps)) o << "NA";
        else o << e.eps;
        o << ")";
    }
    o << std::endl << ")";
    if (!nums.empty() || !strs.empty()) {
        o << ", info=list(";
        bool any = false;
        for (const auto& kv : nums) { o << (any ? ", " : "") << kv.first << " = " << kv.second; any = true; }
        for (const auto& kv : strs) { o << (any ? ", " : "") << kv.first << " = \"" << kv.second << "\""; any = true; }
        o << ")";
    }
    o << ");";
    return o.str();
}

void run(dsgdpp_ctx* c, const dsgdpp_run_cfg* rc) {
    if (!rc) throw Error(DSGDPP_E_ARG, "null run configuration");
    requireReady(c);
    std::string dn;
    std::vector<double> da;
    if (!splitFunctor(rc->decay ? rc->decay : "", dn, da)) throw Error(DSGDPP_E_ARG, "invalid decay");
    bool bold = false;
    double eps0 = 0, dec = 0.5, inc = 1.05;
    if (dn == "BoldDriver" && (da.size() == 1 || da.size() == 3)) {   // parse.h:171-198
        bold = true;
        eps0 = da[0];
        if (da.size() == 3) { dec = da[1]; inc = da[2]; }
    } else if (dn == "Const" && da.size() == 1) {
        eps0 = da[0];
    } else {
        throw Error(DSGDPP_E_ARG, "Error in argument 'decay': Unknown or invalid decay '" + std::string(rc->decay ? rc->decay : "") + "'");
    }
    if (rc->epochs < 0) throw Error(DSGDPP_E_ARG, "negative epoch count");
    using clk = std::chrono::steady_clock;
    auto ns = [](clk::time_point a, clk::time_point b) {
        return (double)std::chrono::duration_cast<std::chrono::nanoseconds>(b - a).count();
    };
    const bool verbose = rc->verbose != 0;
    const bool haveTest = c->blocks[1].loaded;
    std::vector<TraceEntry> es;
    // defaultRunner (mf/sgd/sgd_impl.h:34-145)
    auto a = clk::now();
    double cur = trainLoss(c);
    auto b = clk::now();
    double tl = ns(a, b);
    if (verbose) printf("Loss: %g (%.6fs)\n", cur, tl / 1e9);
    double test = NAN, tt = 0;
    if (haveTest) {
        a = clk::now(); test = testLoss(c); b = clk::now(); tt = ns(a, b);
        if (verbose) printf("Test loss: %g (%.6fs)\n", test, tt / 1e9);
    }
    es.push_back({0, cur, test, NAN, 0, 0, tl, tt});
    double prev = 0, eps = eps0;
    for (int64_t ep = 0; ep < rc->epochs; ++ep) {
        a = clk::now();
        if (ep > 0 && bold) eps = (prev <= cur) ? eps * dec : eps * inc;   // decay_bolddriver.h:59-67
        b = clk::now();
        const double te = ns(a, b);
        if (verbose) printf("Step size: %g\nStarting epoch %lld\n", eps, (long long)ep + 1);
        a = clk::now();
        epoch(c, eps, nullptr);
        b = clk::now();
        const double tep = ns(a, b);
        if (verbose) printf("Finished epoch %lld (%.6fs)\n", (long long)ep + 1, tep / 1e9);
        balance(c, rc->balance);
        prev = cur;
        a = clk::now(); cur = trainLoss(c); b = clk::now(); tl = ns(a, b);
        if (verbose) printf("Loss: %g (%.6fs)\n", cur, tl / 1e9);
        test = NAN;
        tt = 0;
        if (haveTest) {
            a = clk::now(); test = testLoss(c); b = clk::now(); tt = ns(a, b);
            if (verbose) printf("Test loss: %g (%.6fs)\n", test, tt / 1e9);
        }
        es.push_back({ep + 1, cur, test, eps, te, tep, tl, tt});
        if (verbose) fflush(stdout);
    }
    for (size_t k = 0; k < es.size(); ++k) {
        if (rc->loss) rc->loss[k] = es[k].loss;
        if (rc->test_loss) rc->test_loss[k] = es[k].test;
        if (rc->eps) rc->eps[k] = es[k].eps;
        if (rc->time_epoch) rc->time_epoch[k] = es[k].tEpoch / 1e9;
    }
    if (rc->trace_path && *rc->trace_path) {
        // Args::createTraceFields (tools/detail/mfdsgd-args.h:53-63)
        const std::map<std::string, double> nums{{"rank", (double)c->r}, {"nodes", (double)c->w}, {"threads", (double)c->t}};
        const std::map<std::string, std::string> strs{{"update", c->updateStr},
                                                      {"regularize", c->regularizeStr},
                                                      {"input_file", rc->input_file ? rc->input_file : ""},
                                                      {"sample_matrix", ""},
                                                      {"barrier", "false"}};
        std::ofstream out(rc->trace_path);
        if (!out.is_open()) throw Error(DSGDPP_E_IO, std::string("Cannot open file ") + rc->trace_path);
        out << formatTrace(es, rc->trace_var && *rc->trace_var ? rc->trace_var : "trace", nums, strs);
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* dsgdpp_create_error(void) { return g_createError.c_str(); }

int dsgdpp_create(const dsgdpp_cfg* cfg, dsgdpp_ctx** out) {
    if (!out) return DSGDPP_E_ARG;
    *out = nullptr;
    dsgdpp_ctx* c = new dsgdpp_ctx();
    try {
        configure(c, cfg);
        int count = 0;
        HIPC(hipGetDeviceCount(&count));
        if (c->device < 0 || c->device >= count) throw Error(DSGDPP_E_DEVICE, "no such HIP device");
        HIPC(hipSetDevice(c->device));
        HIPC(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIPC(hipStreamCreateWithFlags(&c->prepStream, hipStreamNonBlocking));
        for (auto& e : c->ev) HIPC(hipEventCreate(&e));
        for (EpochPlan& P : c->plan) {
            HIPC(hipEventCreate(&P.evStart));
            HIPC(hipEventCreate(&P.evEnd));
        }
        const char* la = getenv("DSGDPP_LOOKAHEAD");
        c->lookahead = !(la && la[0] == '0');
        const char* kp = getenv("DSGDPP_KPROF");
        c->kprof = kp && kp[0] == '1';
        const char* pm = getenv("DSGDPP_PAIRS");
        c->pairsMode = pm && (pm[0] == '0' || pm[0] == '1') ? pm[0] - '0' : -1;
        dalloc(c->dPartials, (size_t)partials_capacity() * sizeof(double));
        HIPC(hipHostMalloc((void**)&c->hPinned, sizeof(double) * 8, hipHostMallocDefault));
        HIPC(hipHostMalloc((void**)&c->hStatus, sizeof(int32_t), hipHostMallocDefault));
        *c->hStatus = 0;
        dalloc(c->dStatus, sizeof(int32_t));
        HIPC(hipDeviceGetAttribute(&c->wallKhz, hipDeviceAttributeWallClockRate, c->device));
        if (c->wallKhz <= 0) c->wallKhz = 100000;
        noalloc_scratch(c->device);   // the device's address of the dsgdpp_loss_noalloc scratch
    } catch (const std::exception& e) {
        g_createError = e.what();
        const int code = dynamic_cast<const Error*>(&e) ? static_cast<const Error&>(e).code : DSGDPP_E_DEVICE;
        dsgdpp_destroy(c);
        return code;
    }
    *out = c;
    return DSGDPP_OK;
}

void dsgdpp_destroy(dsgdpp_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->prepStream) (void)hipStreamSynchronize(c->prepStream);
    for (int k = 0; k < 2; ++k) { dfree(c->dPts[k]); void* p = c->dBlk[k]; dfree(p); }
    dfree(c->dRid);
    { void* p = c->G.spill; dfree(p); c->G.spill = nullptr; }
    for (EpochPlan& P : c->plan) {
        void* q;
        q = P.dSched; dfree(q);
        q = P.dVoff; dfree(q);
        q = P.dGoff; dfree(q);
        q = P.dNeed; dfree(q);
        q = P.dJobs; dfree(q);
        q = P.dChains; dfree(q);
        q = P.perm; dfree(q);
        q = P.grec; dfree(q);
        q = P.ngroups; dfree(q);
        q = P.LB.lrec; dfree(q);
        q = P.LB.lstart; dfree(q);
        q = P.LB.pos; dfree(q);
        q = P.LB.prv; dfree(q);
        q = P.LB.lev; dfree(q);
        q = P.LB.nlev; dfree(q);
        for (uint32_t* f : P.LB.fy) { q = f; dfree(q); }
        q = P.dLoff; dfree(q);
        if (P.evStart) (void)hipEventDestroy(P.evStart);
        if (P.evEnd) (void)hipEventDestroy(P.evEnd);
    }
    freeLevelIndex(c);
    { void* q = c->dLevelBar; dfree(q); }
    dfree(c->P.W);
    dfree(c->P.H);
    void* v;
    v = const_cast<void*>(c->P.inv1); dfree(v);
    v = const_cast<void*>(c->P.inv2); dfree(v);
    v = const_cast<double*>(c->P.cnt1); dfree(v);
    v = const_cast<double*>(c->P.cnt2); dfree(v);
    v = c->dPartials; dfree(v);
    v = const_cast<int32_t*>(c->G.off2); dfree(v);
    v = c->G.prof; dfree(v);
    v = c->G.pprof; dfree(v);
    v = c->dReady; dfree(v);
    v = c->dStatus; dfree(v);
    { void* q = c->dPairs; dfree(q); c->dPairs = nullptr; }
    if (c->hPinned) (void)hipHostFree(c->hPinned);
    if (c->hStatus) (void)hipHostFree(c->hStatus);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->prepStream) (void)hipStreamDestroy(c->prepStream);
    delete c;
}

const char* dsgdpp_last_error(const dsgdpp_ctx* c) { return c ? c->err.c_str() : "null context"; }

int dsgdpp_load_coo(dsgdpp_ctx* c, int32_t which, int64_t m, int64_t n, int64_t nnz, const int32_t* i,
                    const int32_t* j, const double* x) {
    return guarded(c, [&] { loadCoo(c, which, m, n, nnz, i, j, x); });
}

int dsgdpp_load_mmc(dsgdpp_ctx* c, int32_t which, const char* path) {
    return guarded(c, [&] {
        if (!path) throw Error(DSGDPP_E_ARG, "null path");
        loadMmc(c, which, path);
    });
}

int dsgdpp_save_blocks(dsgdpp_ctx* c, int32_t which, const char* path) {
    return guarded(c, [&] {
        if (!path || !*path) throw Error(DSGDPP_E_ARG, "empty path");
        saveBlocks(c, which, path, 0, 0);
    });
}

int dsgdpp_load_blocks(dsgdpp_ctx* c, int32_t which, const char* path) {
    return guarded(c, [&] {
        if (!path || !*path) throw Error(DSGDPP_E_ARG, "empty path");
        loadBlocks(c, which, path);
    });
}

int dsgdpp_load_mmc_cached(dsgdpp_ctx* c, int32_t which, const char* path, const char* cache_path,
                           int32_t* from_cache) {
    return guarded(c, [&] {
        if (!path || !cache_path || !*cache_path) throw Error(DSGDPP_E_ARG, "empty path");
        const bool hit = loadMmcCached(c, which, path, cache_path);
        if (from_cache) *from_cache = hit ? 1 : 0;
    });
}

int dsgdpp_set_factors(dsgdpp_ctx* c, const double* W, const double* H) {
    return guarded(c, [&] {
        if (c->m < 0) throw Error(DSGDPP_E_STATE, "load the training matrix first");
        if (!W || !H) throw Error(DSGDPP_E_ARG, "null factor array");
        putFactor(c, c->P.W, W, c->m);
        putFactor(c, c->P.H, H, c->n);
        c->haveFactors = true;
    });
}

int dsgdpp_get_factors(dsgdpp_ctx* c, double* W, double* H) {
    return guarded(c, [&] {
        if (!c->haveFactors) throw Error(DSGDPP_E_STATE, "no factors");
        HIPC(hipStreamSynchronize(c->stream));
        if (W) getFactor(c, c->P.W, W, c->m);
        if (H) getFactor(c, c->P.H, H, c->n);
    });
}

int dsgdpp_load_factors_mma(dsgdpp_ctx* c, const char* wpath, const char* hpath) {
    return guarded(c, [&] {
        if (c->m < 0) throw Error(DSGDPP_E_STATE, "load the training matrix first");
        int64_t a, b;
        std::vector<double> cm;
        readMma(wpath ? wpath : "", a, b, cm);
        if (a != c->m || b != c->r) throw Error(DSGDPP_E_ARG, "row factor matrix is " + std::to_string(a) + " x " + std::to_string(b));
        std::vector<double> W((size_t)a * b);
        for (int64_t z = 0; z < b; ++z)
            for (int64_t i = 0; i < a; ++i) W[i * b + z] = cm[z * a + i];
        readMma(hpath ? hpath : "", a, b, cm);
        if (a != c->r || b != c->n) throw Error(DSGDPP_E_ARG, "column factor matrix is " + std::to_string(a) + " x " + std::to_string(b));
        putFactor(c, c->P.W, W.data(), c->m);
        putFactor(c, c->P.H, cm.data(), c->n);
        c->haveFactors = true;
    });
}

int dsgdpp_write_factors_mma(dsgdpp_ctx* c, const char* wpath, const char* hpath) {
    return guarded(c, [&] {
        if (!c->haveFactors) throw Error(DSGDPP_E_STATE, "no factors");
        HIPC(hipStreamSynchronize(c->stream));
        const int64_t r = c->r;
        if (wpath && *wpath) {
            std::vector<double> W((size_t)c->m * r), cm((size_t)c->m * r);
            getFactor(c, c->P.W, W.data(), c->m);
            for (int64_t z = 0; z < r; ++z)
                for (int64_t i = 0; i < c->m; ++i) cm[z * c->m + i] = W[i * r + z];
            writeMma(wpath, c->m, r, cm);
        }
        if (hpath && *hpath) {
            std::vector<double> H((size_t)c->n * r);
            getFactor(c, c->P.H, H.data(), c->n);
            writeMma(hpath, r, c->n, H);
        }
    });
}

int dsgdpp_grouped_ok(dsgdpp_ctx* c, int32_t* ok, char* why, int64_t cap) {
    return guarded(c, [&] {
        if (!ok) throw Error(DSGDPP_E_ARG, "null output");
        if (!c->blocks[0].loaded) throw Error(DSGDPP_E_STATE, "load the training matrix first");
        *ok = c->groupedOk ? 1 : 0;
        if (why && cap > 0) {
            const std::string w = !c->groupedOk ? c->groupedWhy
                                  : c->pairBlocks ? "two-slot groups in " + std::to_string(c->pairBlocks) + " of " +
                                                        std::to_string((size_t)c->d * 2 * c->d) + " blocks"
                                                  : std::string();
            const size_t k = std::min<size_t>(w.size(), (size_t)cap - 1);
            memcpy(why, w.data(), k);
            why[k] = 0;
        }
    });
}

int dsgdpp_epoch(dsgdpp_ctx* c, double eps, dsgdpp_epoch_stats* stats) {
    return guarded(c, [&] { epoch(c, eps, stats); });
}

int dsgdpp_loss(dsgdpp_ctx* c, int32_t which, double* out) {
    return guarded(c, [&] {
        if (!out) throw Error(DSGDPP_E_ARG, "null output");
        if (which == DSGDPP_TRAIN) *out = trainLoss(c);
        else if (which == DSGDPP_TEST) *out = testLoss(c);
        else throw Error(DSGDPP_E_ARG, "which must be TRAIN or TEST");
    });
}

int dsgdpp_loss_noalloc(dsgdpp_ctx* c, int32_t which, double* d_out, void* stream) {
    return guarded(c, [&] {
        if (which != DSGDPP_TRAIN && which != DSGDPP_TEST) throw Error(DSGDPP_E_ARG, "which must be TRAIN or TEST");
        if (!d_out) throw Error(DSGDPP_E_ARG, "null output pointer");
        requireReady(c);
        if (which == DSGDPP_TEST && !c->blocks[1].loaded) throw Error(DSGDPP_E_STATE, "no test matrix loaded");
        double* sp = noalloc_scratch_lookup(c->device);
        if (!sp) throw Error(DSGDPP_E_STATE, "the no-allocation scratch of this device is not set up");
        const int64_t need = device_loss_noalloc(c->P, which, c->upd.kind, c->lossLambda, sp, d_out,
                                                 (hipStream_t)stream);
        if (need > 0)
            throw Error(DSGDPP_E_UNSUPPORTED, "the loss reduction needs " + std::to_string(need) +
                                                  " partial sums; the no-allocation scratch holds " +
                                                  std::to_string(noalloc_partials()));
    });
}

int dsgdpp_balance(dsgdpp_ctx* c, int32_t type) {
    return guarded(c, [&] { balance(c, type); });
}

int dsgdpp_run(dsgdpp_ctx* c, const dsgdpp_run_cfg* rc) {
    return guarded(c, [&] { run(c, rc); });
}

int dsgdpp_get_schedule(dsgdpp_ctx* c, uint32_t* out) {
    return guarded(c, [&] {
        if (c->lastPlan < 0) throw Error(DSGDPP_E_STATE, "no epoch has run");
        if (!out) throw Error(DSGDPP_E_ARG, "null output");
        const auto& sch = c->plan[c->lastPlan].schedule;
        if (c->dsgd()) {   // d x d in DSGD's own numbering
            for (size_t i = 0; i < c->d; ++i)
                for (size_t id = 0; id < c->d; ++id) out[i * c->d + id] = sch[(2 * i) * c->d + id] / 2;
        } else {
            std::copy(sch.begin(), sch.end(), out);
        }
    });
}

int dsgdpp_get_task_seeds(dsgdpp_ctx* c, uint64_t* out) {
    return guarded(c, [&] {
        if (c->lastPlan < 0) throw Error(DSGDPP_E_STATE, "no epoch has run");
        if (!out) throw Error(DSGDPP_E_ARG, "null output");
        const auto& sd = c->plan[c->lastPlan].seeds;
        std::copy(sd.begin(), sd.end(), out);
    });
}

int dsgdpp_get_permutation(dsgdpp_ctx* c, int32_t task, int32_t subepoch, uint32_t* out, int64_t cap,
                           int64_t* n) {
    return guarded(c, [&] {
        if (c->lastPlan < 0) throw Error(DSGDPP_E_STATE, "no epoch has run");
        if (task < 0 || task >= (int)c->d || subepoch < 0 || subepoch >= (int)((c->dsgd() ? 1 : 2) * c->d))
            throw Error(DSGDPP_E_ARG, "task or sub-epoch out of range");
        if (c->dsgd()) subepoch *= 2;
        const EpochPlan& Pl = c->plan[c->lastPlan];
        const size_t b = (size_t)task * 2 * c->d + Pl.schedule[(size_t)subepoch * c->d + task];
        const int64_t nb = c->blocks[0].start[b + 1] - c->blocks[0].start[b];
        const int64_t base = Pl.voff[(size_t)subepoch * c->d + task];
        if (n) *n = nb;
        if (!out) return;
        if (cap < nb) throw Error(DSGDPP_E_ARG, "output buffer too small");
        if (c->sgdOrder == DSGDPP_SGD_SEQ) {
            for (int64_t k = 0; k < nb; ++k) out[k] = (uint32_t)k;
        } else if (nb > 0) {
            HIPC(hipStreamSynchronize(c->stream));
            HIPC(hipMemcpy(out, Pl.perm + base, (size_t)nb * sizeof(uint32_t), hipMemcpyDeviceToHost));
        }
    });
}

int dsgdpp_block_nnz(dsgdpp_ctx* c, int32_t which, int64_t* out) {
    return guarded(c, [&] {
        if (which != 0 && which != 1) throw Error(DSGDPP_E_ARG, "which must be TRAIN or TEST");
        if (!c->blocks[which].loaded) throw Error(DSGDPP_E_STATE, "matrix not loaded");
        const auto& s = c->blocks[which].start;
        if (c->dsgd()) {   // d x d: DSGD's block k is column block 2k
            for (size_t b1 = 0; b1 < c->d; ++b1)
                for (size_t k = 0; k < c->d; ++k) {
                    const size_t b = b1 * 2 * c->d + 2 * k;
                    out[b1 * c->d + k] = s[b + 1] - s[b];
                }
        } else {
            for (size_t b = 0; b + 1 < s.size(); ++b) out[b] = s[b + 1] - s[b];
        }
    });
}

int dsgdpp_get_block(dsgdpp_ctx* c, int32_t which, int32_t b1, int32_t b2, int32_t* i, int32_t* j, double* x,
                     int64_t cap, int64_t* n) {
    return guarded(c, [&] {
        if (which != 0 && which != 1) throw Error(DSGDPP_E_ARG, "which must be TRAIN or TEST");
        if (!c->blocks[which].loaded) throw Error(DSGDPP_E_STATE, "matrix not loaded");
        const int nb2 = c->dsgd() ? (int)c->d : 2 * (int)c->d;
        if (b1 < 0 || b1 >= (int)c->d || b2 < 0 || b2 >= nb2) throw Error(DSGDPP_E_ARG, "block out of range");
        const size_t b = (size_t)b1 * 2 * c->d + (size_t)(c->dsgd() ? 2 * b2 : b2);
        const auto& st = c->blocks[which].start;
        const int64_t k = st[b + 1] - st[b];
        if (n) *n = k;
        if (!i && !j && !x) return;
        if (cap < k) throw Error(DSGDPP_E_ARG, "output buffer too small");
        std::vector<unsigned char> raw((size_t)k * 16);
        HIPC(hipStreamSynchronize(c->stream));
        if (k > 0) HIPC(hipMemcpy(raw.data(), (const char*)c->dPts[which] + st[b] * 16, raw.size(), hipMemcpyDeviceToHost));
        for (int64_t p = 0; p < k; ++p) {
            const unsigned char* e = raw.data() + p * 16;
            int32_t ii, jj;
            memcpy(&ii, e, 4);
            memcpy(&jj, e + 4, 4);
            double v;
            if (c->fp64()) memcpy(&v, e + 8, 8);
            else { float f; memcpy(&f, e + 8, 4); v = f; }
            if (i) i[p] = ii;
            if (j) j[p] = jj;
            if (x) x[p] = v;
        }
    });
}

int dsgdpp_dims(dsgdpp_ctx* c, int64_t* out) {
    return guarded(c, [&] {
        out[0] = c->m;
        out[1] = c->n;
        out[2] = c->r;
        out[3] = c->d;
        out[4] = c->blocks[0].nnz;
        out[5] = c->blocks[1].nnz;
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched epochs: several contexts (independent factorizations of any shape,
// rank <= 128, own lambda and step size) advance one epoch each in a single
// persistent launch.  Each context keeps its data, factors, random streams
// and look-ahead plans; the batch gathers their kernel arguments into one
// device table and replays a captured graph of the launch every epoch.
// ---------------------------------------------------------------------------
struct dsgdpp_batch {
    std::vector<dsgdpp_ctx*> ctx;
    int device = 0, precision = 0, kind = 0;
    bool trunc = false;
    size_t argBytes = 0, lds = 0;
    int nblocks = 0;
    hipStream_t stream = nullptr;
    hipStream_t prepStream = nullptr;   // the look-ahead preps of all contexts, one launch per index width
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<PrepDesc> desc;
    std::vector<int2> cmap;
    PrepDesc* dDesc = nullptr;
    int2* dCmap = nullptr;
    size_t descCap = 0, cmapCap = 0;
    uint8_t* hArgs = nullptr;    // pinned host staging of the argument table
    void* dArgs = nullptr;
    int2* dTmap = nullptr;        // (problem, task) items, workgroup after workgroup
    int32_t* dWstart = nullptr;   // nblocks + 1: each workgroup's first item
    int tasks = 0;                // items in all
    std::vector<int32_t*> ready;
    std::vector<int> dsz;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool useGraph = true;
    // The launch is an ordinary one (a replayed graph node), so residency of
    // all its workgroups rests on the occupancy check at create time; its
    // dataflow waits give up after waitMs and set *dStatus instead of hanging.
    int32_t* dStatus = nullptr;
    int32_t* hStatus = nullptr;   // pinned
    int wallKhz = 0;              // wall-clock ticks per millisecond
    bool counted = false;         // its contexts' batch counts include it
    double* dLoss = nullptr;      // batch loss: the partial sums, then one total per context
    double* hLoss = nullptr;      // pinned: the totals
    std::string err;
};

namespace {

thread_local std::string g_batchCreateError;

void batchDestroy(dsgdpp_batch* b) {
    if (!b) return;
    if (b->counted)
        for (dsgdpp_ctx* c : b->ctx) --c->batches;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->hArgs) (void)hipHostFree(b->hArgs);
    if (b->hStatus) (void)hipHostFree(b->hStatus);
    if (b->hLoss) (void)hipHostFree(b->hLoss);
    void* q = b->dStatus;
    dfree(q);
    q = b->dLoss;
    dfree(q);
    q = b->dArgs;
    dfree(q);
    q = b->dTmap;
    dfree(q);
    q = b->dWstart;
    dfree(q);
    q = b->dDesc;
    dfree(q);
    q = b->dCmap;
    dfree(q);
    if (b->prepStream) (void)hipStreamSynchronize(b->prepStream);
    if (b->prepStream) (void)hipStreamDestroy(b->prepStream);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}


int deviceCUs(int device) {
    int cus = 0;
    HIPC(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    return std::max(1, cus);
}

// The batch's (problem, task) items on at most `slots` workgroups (one per CU
// at most: two tasks on a CU run at about half speed each).  When every task
// has a workgroup of its own, that is all.  Otherwise, longest-processing-time
// first: problems in decreasing order of work per task (points per task,
// rank classes above 64 weighted 1.7), each problem's d tasks onto the d
// least loaded workgroups, which then hold their items in that one order of
// the problems and at most one task of each problem (the batch kernel's
// progress argument).  Returns the number of workgroups.
int packBatch(const std::vector<dsgdpp_ctx*>& ctx, int slots, std::vector<int2>& items, std::vector<int32_t>& wstart) {
    int tasks = 0;
    for (dsgdpp_ctx* c : ctx) tasks += (int)c->d;
    items.clear();
    wstart.clear();
    if (tasks <= slots) {
        for (size_t k = 0; k < ctx.size(); ++k)
            for (int t = 0; t < (int)ctx[k]->d; ++t) items.push_back(make_int2((int)k, t));
        for (int w = 0; w <= tasks; ++w) wstart.push_back(w);
        return tasks;
    }
    std::vector<double> work(ctx.size());
    for (size_t k = 0; k < ctx.size(); ++k) {
        if ((int)ctx[k]->d > slots)
            throw Error(DSGDPP_E_UNSUPPORTED, "context " + std::to_string(k) + " has more tasks than the device has CUs");
        work[k] = (double)ctx[k]->blocks[0].nnz / ctx[k]->d * (ctx[k]->P.ld > 64 ? 1.7 : 1.0);
    }
    std::vector<int> order(ctx.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return work[a] > work[c]; });
    std::vector<double> load(slots, 0.0);
    std::vector<std::vector<int2>> lists(slots);
    std::vector<int> ws(slots);
    std::iota(ws.begin(), ws.end(), 0);
    for (int k : order) {
        std::stable_sort(ws.begin(), ws.end(), [&](int a, int c) { return load[a] < load[c]; });
        for (int t = 0; t < (int)ctx[k]->d; ++t) {
            lists[ws[t]].push_back(make_int2(k, t));
            load[ws[t]] += work[k];
        }
    }
    for (int w = 0; w < slots; ++w) {
        wstart.push_back((int32_t)items.size());
        items.insert(items.end(), lists[w].begin(), lists[w].end());
    }
    wstart.push_back((int32_t)items.size());
    return slots;
}

void batchEpoch(dsgdpp_batch* b, const double* eps, dsgdpp_epoch_stats* st) {
    const auto t0 = std::chrono::steady_clock::now();
    const size_t K = b->ctx.size();
    const uint32_t waitTicks = dataflowWaitTicks(b->wallKhz);
    int64_t updates = 0;
    for (size_t k = 0; k < K; ++k) {
        dsgdpp_ctx* c = b->ctx[k];
        requireReady(c);
        EpochPlan& P = c->plan[c->cur];
        if (!P.drawn) drawPlan(c, P);
        if (!P.built || P.dataVersion != c->dataVersion || P.level) {   // a level plan predates the batch
            HIPC(hipStreamSynchronize(c->prepStream));
            buildPlan(c, P, c->prepStream);
        }
        if (!P.grouped) throw Error(DSGDPP_E_UNSUPPORTED, "batched epochs need the grouped kernel");
        HIPC(hipStreamWaitEvent(b->stream, P.evEnd, 0));
        HIPC(hipEventRecord(c->ev[3], c->stream));   // the context's own earlier work comes first
        HIPC(hipStreamWaitEvent(b->stream, c->ev[3], 0));
        EpochLaunch L;
        L.sched = P.dSched;
        L.voff = P.dVoff;
        L.goff = P.dGoff;
        L.need = P.dNeed;
        L.ready = c->dReady;
        L.dataflow = 1;
        L.eps = eps[k];
        DevProblem Dp = c->P;
        Dp.perm = P.perm;
        GroupedBuffers G = c->G;
        G.grec = P.grec;
        G.ngroups = P.ngroups;
        fill_grouped_args(Dp, c->upd, G, L, b->dStatus, waitTicks, b->hArgs + k * b->argBytes);
        updates += c->blocks[0].nnz;
    }
    const auto t_prep = std::chrono::steady_clock::now();
    HIPC(hipMemcpyAsync(b->dArgs, b->hArgs, K * b->argBytes, hipMemcpyHostToDevice, b->stream));
    HIPC(hipEventRecord(b->ev0, b->stream));
    if (b->useGraph && !b->exec) {
        // one graph: the counters' reset and the launch; the table it reads is refreshed above
        HIPC(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        launch_epoch_batch(b->precision, b->kind, b->trunc, b->dArgs, b->dTmap, b->dWstart, b->nblocks, b->lds,
                           b->ready.data(), b->dsz.data(), (int)K, b->dStatus, b->stream);
        HIPC(hipStreamEndCapture(b->stream, &b->graph));
        HIPC(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    if (b->useGraph) HIPC(hipGraphLaunch(b->exec, b->stream));
    else launch_epoch_batch(b->precision, b->kind, b->trunc, b->dArgs, b->dTmap, b->dWstart, b->nblocks, b->lds,
                            b->ready.data(), b->dsz.data(), (int)K, b->dStatus, b->stream);
    HIPC(hipMemcpyAsync(b->hStatus, b->dStatus, sizeof(int32_t), hipMemcpyDeviceToHost, b->stream));
    HIPC(hipEventRecord(b->ev1, b->stream));
    const auto t_launch = std::chrono::steady_clock::now();
    // every context's next plan is prepared while the batch runs: the host
    // parts on the host threads, the grouped preps of all of them in one
    // launch per class of prep LDS
    HIPC(hipStreamWaitEvent(b->prepStream, b->ev0, 0));
    b->desc.clear();
    b->cmap.clear();
    std::vector<EpochPlan*> ahead;
    // chains by (index width, workgroups per CU that the problem's prep LDS
    // allows, rounded down to a power of two): one launch per class, so a
    // problem with tall blocks does not set the LDS, and with it the
    // occupancy, of every other problem's prep; the rounding merges classes
    // (fewer launch tails: batch64 179.5 -> 178.4 ms per epoch)
    struct PrepClass {
        std::vector<int2> chains;
        size_t lds = 0, order_lds = 0;
    };
    std::map<std::pair<int, int>, PrepClass> classes;
    for (size_t k = 0; k < K; ++k) {
        dsgdpp_ctx* c = b->ctx[k];
        c->lastPlan = c->cur;
        c->plan[c->cur].drawn = false;
        c->cur ^= 1;
    }
    // the contexts' plans are independent: drawn and built on the host threads
    // (one context per thread at a time; their copies go to the prep stream)
    parallelForEach((int64_t)K, [&](int64_t k) {
        dsgdpp_ctx* c = b->ctx[k];
        if (!c->lookahead) return;
        HIPC(hipSetDevice(b->device));
        drawPlan(c, c->plan[c->cur]);
        buildPlan(c, c->plan[c->cur], b->prepStream, false);
    });
    for (size_t k = 0; k < K; ++k) {
        dsgdpp_ctx* c = b->ctx[k];
        if (!c->lookahead) continue;
        EpochPlan& N = c->plan[c->cur];
        DevProblem Dp = c->P;
        Dp.perm = N.perm;
        GroupedBuffers G = c->G;
        G.grec = N.grec;
        G.ngroups = N.ngroups;
        // one descriptor per size class of the problem's chains
        for (size_t q = 0; q < N.clsCap.size(); ++q) {
            const PrepDesc D = prep_desc(Dp, G, N.dJobs, N.dChains, c->sgdOrder, N.clsCap[q]);
            const int idx = (int)b->desc.size();
            b->desc.push_back(D);
            const size_t l = prep_desc_lds(D);
            PrepClass& pc =
                classes[{prep_narrow(N.clsCap[q]) ? 1 : 0, pow2_floor((int)std::max<size_t>(1, PREP_LDS_PER_CU / l))}];
            for (int ch = N.clsStart[q]; ch < N.clsStart[q + 1]; ++ch) pc.chains.push_back(make_int2(idx, ch));
            pc.lds = std::max(pc.lds, l);
            pc.order_lds = std::max(pc.order_lds, prep_desc_order_lds(D));
        }
        ahead.push_back(&N);
    }
    if (!ahead.empty()) {
        // the classes with the least LDS per workgroup first: the launches of
        // one stream start in order, and those can start beside the batch's
        // own workgroups (a class that needs more LDS than a CU has left would
        // hold back every launch after it until the batch ends)
        std::vector<std::pair<std::pair<int, int>, PrepClass*>> order;
        for (auto& kv : classes) order.push_back({kv.first, &kv.second});
        std::sort(order.begin(), order.end(), [](const auto& x, const auto& y) { return x.first.second > y.first.second; });
        std::vector<int> start;
        for (auto& o : order) {
            start.push_back((int)b->cmap.size());
            b->cmap.insert(b->cmap.end(), o.second->chains.begin(), o.second->chains.end());
        }
        if (b->desc.size() > b->descCap) {
            b->descCap = b->desc.size();
            dalloc(b->dDesc, b->descCap * sizeof(PrepDesc));
        }
        if (b->cmap.size() > b->cmapCap) {
            b->cmapCap = b->cmap.size();
            dalloc(b->dCmap, b->cmapCap * sizeof(int2));
        }
        HIPC(hipMemcpyAsync(b->dDesc, b->desc.data(), b->desc.size() * sizeof(PrepDesc), hipMemcpyHostToDevice,
                            b->prepStream));
        if (!b->cmap.empty())
            HIPC(hipMemcpyAsync(b->dCmap, b->cmap.data(), b->cmap.size() * sizeof(int2), hipMemcpyHostToDevice,
                                b->prepStream));
        for (size_t q = 0; q < order.size(); ++q)
            launch_prep_grouped_multi(b->dDesc, b->dCmap + start[q], (int)order[q].second->chains.size(),
                                      order[q].first.first == 1, order[q].second->lds, b->prepStream,
                                      order[q].second->order_lds);
        for (EpochPlan* N : ahead) HIPC(hipEventRecord(N->evEnd, b->prepStream));
    }
    const auto t_ahead = std::chrono::steady_clock::now();
    HIPC(hipEventSynchronize(b->ev1));
    // the contexts' own streams see the batch's writes before their next work
    for (dsgdpp_ctx* c : b->ctx) HIPC(hipStreamWaitEvent(c->stream, b->ev1, 0));
    if (*b->hStatus != 0)
        throw Error(DSGDPP_E_DEVICE, std::string("batched epoch failed: ") + kGaveUp + " after " +
                                         std::to_string(waitTicks / std::max(1, b->wallKhz)) +
                                         " ms; the factors of every context in this batch are undefined until set "
                                         "again");
    const auto t1 = std::chrono::steady_clock::now();
    if (b->ctx[0]->kprof) {
        auto ms = [](auto a, auto z) { return std::chrono::duration<double, std::milli>(z - a).count(); };
        fprintf(stderr, "[kprof] batch host ms: arguments %.2f launch %.2f look-ahead plans %.2f wait %.2f\n",
                ms(t0, t_prep), ms(t_prep, t_launch), ms(t_launch, t_ahead), ms(t_ahead, t1));
    }
    if (st) {
        float ms = 0;
        HIPC(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        memset(st, 0, sizeof(*st));
        st->seconds = std::chrono::duration<double>(t1 - t0).count();
        st->sgd_seconds = ms / 1e3;
        st->device_seconds = ms / 1e3;
        st->updates = updates;
        st->launches = 1;
        st->kernel = 3;
    }
}

}  // namespace

extern "C" {

int dsgdpp_batch_create(dsgdpp_ctx* const* ctxs, int32_t count, int32_t use_graph, dsgdpp_batch** out) {
    if (!out || !ctxs || count <= 0) {
        g_batchCreateError = "invalid arguments";
        return DSGDPP_E_ARG;
    }
    *out = nullptr;
    dsgdpp_batch* b = new dsgdpp_batch();
    try {
        for (int k = 0; k < count; ++k) {
            dsgdpp_ctx* c = ctxs[k];
            if (!c) throw Error(DSGDPP_E_ARG, "null context");
            for (int q = 0; q < k; ++q)
                if (ctxs[q] == c) throw Error(DSGDPP_E_ARG, "a context appears twice");
            requireReady(c);
            if (c->precision != PREC_FP32 && c->precision != PREC_BF16)
                throw Error(DSGDPP_E_UNSUPPORTED, "context " + std::to_string(k) +
                                                      ": batched epochs run FP32/BF16 factors (grouped kernel)");
            if (!c->groupedOk || !(c->kernel == 0 || c->kernel == 3 || c->kernel == 4))
                throw Error(DSGDPP_E_UNSUPPORTED, "context " + std::to_string(k) + " cannot use the grouped kernel: " +
                                                      (c->groupedOk ? std::string("kernel forced") : c->groupedWhy));
            if (prep_spill_bytes(c->maxBlock, c->G.prep_rows, c->G.max_cols, 1) > 0)
                throw Error(DSGDPP_E_UNSUPPORTED, "context " + std::to_string(k) +
                                                      ": its blocks' prep state does not fit in LDS (batched preps "
                                                      "keep it there; use more tasks)");
            if (k == 0) {
                b->device = c->device;
                b->precision = c->precision;
                b->kind = c->upd.kind;
                b->trunc = c->upd.truncate;
            } else if (c->device != b->device || c->precision != b->precision || c->upd.kind != b->kind ||
                       c->upd.truncate != b->trunc) {
                throw Error(DSGDPP_E_UNSUPPORTED, "batched contexts need the same device, precision, update and truncation");
            }
            b->ctx.push_back(c);
        }
        for (dsgdpp_ctx* c : b->ctx) ++c->batches;   // their plans now take the grouped kernel
        b->counted = true;
        HIPC(hipSetDevice(b->device));
        b->useGraph = use_graph != 0;
        b->argBytes = grouped_args_bytes(b->precision);   // the kernel indexes the table by sizeof
        for (int k = 0; k < count; ++k) {
            dsgdpp_ctx* c = b->ctx[k];
            b->tasks += (int)c->d;
            b->lds = std::max(b->lds, batch_epoch_lds(c->G.max_cols, c->P.ld, c->G.wide));
            if (!c->dReady) dalloc(c->dReady, (size_t)2 * c->d * sizeof(int32_t));
            b->ready.push_back(c->dReady);
            b->dsz.push_back((int)c->d);
            HIPC(hipStreamSynchronize(c->stream));
        }
        const int cap = batch_capacity(b->precision, b->kind, b->trunc, b->lds, b->device);
        std::vector<int2> items;
        std::vector<int32_t> wstart;
        int slots = std::min(cap, deviceCUs(b->device));
        if (const char* e = getenv("DSGDPP_BATCH_SLOTS")) slots = std::max(1, std::min(slots, atoi(e)));
        b->nblocks = packBatch(b->ctx, slots, items, wstart);
        HIPC(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        HIPC(hipStreamCreateWithFlags(&b->prepStream, hipStreamNonBlocking));
        HIPC(hipEventCreate(&b->ev0));
        HIPC(hipEventCreate(&b->ev1));
        HIPC(hipHostMalloc((void**)&b->hArgs, (size_t)count * b->argBytes, hipHostMallocDefault));
        HIPC(hipHostMalloc((void**)&b->hStatus, sizeof(int32_t), hipHostMallocDefault));
        *b->hStatus = 0;
        dalloc(b->dStatus, sizeof(int32_t));
        HIPC(hipDeviceGetAttribute(&b->wallKhz, hipDeviceAttributeWallClockRate, b->device));
        if (b->wallKhz <= 0) b->wallKhz = 100000;
        dalloc(b->dArgs, (size_t)count * b->argBytes);
        dalloc(b->dTmap, items.size() * sizeof(int2));
        HIPC(hipMemcpy(b->dTmap, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice));
        dalloc(b->dWstart, wstart.size() * sizeof(int32_t));
        HIPC(hipMemcpy(b->dWstart, wstart.data(), wstart.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    } catch (const std::exception& e) {
        g_batchCreateError = e.what();
        const Error* ee = dynamic_cast<const Error*>(&e);
        batchDestroy(b);
        return ee ? ee->code : DSGDPP_E_DEVICE;
    }
    *out = b;
    return DSGDPP_OK;
}

const char* dsgdpp_batch_create_error(void) { return g_batchCreateError.c_str(); }
const char* dsgdpp_batch_last_error(dsgdpp_batch* b) { return b ? b->err.c_str() : ""; }

void dsgdpp_batch_destroy(dsgdpp_batch* b) { batchDestroy(b); }

// Every context's loss (the value dsgdpp_loss gives), queued on the batch's
// stream after the work of the contexts' own streams, with one wait.  The
// partial sums sit in the batch's own buffer, one context after another.
int dsgdpp_batch_loss(dsgdpp_batch* b, int32_t which, double* out) {
    if (!b || !out) return DSGDPP_E_ARG;
    try {
        if (which != DSGDPP_TRAIN && which != DSGDPP_TEST) throw Error(DSGDPP_E_ARG, "which must be TRAIN or TEST");
        HIPC(hipSetDevice(b->device));
        const size_t K = b->ctx.size();
        const int parts = noalloc_partials();
        if (!b->dLoss) {
            dalloc(b->dLoss, ((size_t)parts + 4 + K) * sizeof(double));
            HIPC(hipHostMalloc((void**)&b->hLoss, K * sizeof(double), hipHostMallocDefault));
        }
        double* const totals = b->dLoss + parts + 4;
        for (size_t k = 0; k < K; ++k) {
            dsgdpp_ctx* c = b->ctx[k];
            requireReady(c);
            if (which == DSGDPP_TEST && !c->blocks[1].loaded) throw Error(DSGDPP_E_STATE, "no test matrix loaded");
            HIPC(hipEventRecord(c->ev[3], c->stream));   // the context's own earlier work comes first
            HIPC(hipStreamWaitEvent(b->stream, c->ev[3], 0));
            const int64_t need = device_loss_noalloc(c->P, which, c->upd.kind, c->lossLambda, b->dLoss, totals + k,
                                                     b->stream);
            if (need > 0) throw Error(DSGDPP_E_UNSUPPORTED, "the loss reduction needs more partial sums");
        }
        HIPC(hipMemcpyAsync(b->hLoss, totals, K * sizeof(double), hipMemcpyDeviceToHost, b->stream));
        HIPC(hipStreamSynchronize(b->stream));
        memcpy(out, b->hLoss, K * sizeof(double));
        return DSGDPP_OK;
    } catch (const Error& e) {
        b->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        b->err = e.what();
        return DSGDPP_E_DEVICE;
    }
}

int dsgdpp_batch_epoch(dsgdpp_batch* b, const double* eps, dsgdpp_epoch_stats* stats) {
    if (!b || !eps) return DSGDPP_E_ARG;
    try {
        HIPC(hipSetDevice(b->device));
        batchEpoch(b, eps, stats);
        return DSGDPP_OK;
    } catch (const Error& e) {
        b->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        b->err = e.what();
        return DSGDPP_E_DEVICE;
    }
}

}  // extern "C"
