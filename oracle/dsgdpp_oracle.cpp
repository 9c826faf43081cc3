/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Plain C++ CPU restatement of the DSGD++ stratified-SGD path of uma-pi1/DSGDpp.
 * Each function cites the reference file:line it follows.  Arithmetic is
 * compiled with -ffp-contract=off so that every product and sum is rounded
 * exactly as written (the reference's own NativeRelease build may contract,
 * CMakeLists.txt:6; its rounding is therefore flag dependent).
 *
 * Parity pins (tests/test_oracle.py):
 *   Philox4x32-10        -- Random123 known-answer vectors
 *   update functors      -- hand-computed closed forms (r=2/3), long double brute force
 *   losses               -- closed forms / brute force on tiny inputs
 *   block offsets        -- closed form, exhaustive on tiny sizes
 *   schedules            -- invariants + SEQ closed form (dsgdpp_impl.h:292-299)
 *   permute              -- validity, carry-over rule, chi^2 uniformity
 *   BoldDriver           -- closed form eps_k = eps0 * inc^#improve * dec^#worse
 *   trace                -- golden strings written from the format in mf/trace.h
 */
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <mutex>
#include <numeric>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

namespace orc {

// ----------------------------------------------------------------------------
// Random numbers.  Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11).  Draw k of a stream with key (k0,k1)
// is the 64-bit half (k&1) of Philox(counter = {k>>1 (64 bit), 0, 0}, key).
// nextInt(b) = floor(u64 * b / 2^64) (bias <= b/2^64).
// ----------------------------------------------------------------------------
static inline void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)M0 * c[0];
        uint64_t p1 = (uint64_t)M1 * c[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
        k0 += W0; k1 += W1;
    }
}

struct Rng {
    uint32_t k0 = 0, k1 = 0;
    uint64_t n = 0;          // draws consumed
    uint64_t cache = 0;      // odd half of the last block
    bool cached = false;
    Rng() {}
    Rng(uint32_t a, uint32_t b) : k0(a), k1(b) {}
    static Rng master(uint64_t seed) {   // stands in for rg::Random32(seed), mfdsgdpp.cc:542
        return Rng((uint32_t)seed, (uint32_t)(seed >> 32) ^ 0xD5A61266u);
    }
    static Rng task(uint64_t s) { return Rng((uint32_t)s, (uint32_t)(s >> 32)); }
    uint64_t next64() {
        uint64_t k = n++;
        if ((k & 1) && cached) { cached = false; return cache; }
        uint32_t c[4] = {(uint32_t)(k >> 1), (uint32_t)(k >> 33), 0u, 0u};
        philox(c, k0, k1);
        uint64_t lo = (uint64_t)c[0] | ((uint64_t)c[1] << 32);
        uint64_t hi = (uint64_t)c[2] | ((uint64_t)c[3] << 32);
        if (k & 1) return hi;
        cache = hi; cached = true;
        return lo;
    }
    uint64_t nextInt(uint64_t bound) {
        return (uint64_t)(((unsigned __int128)next64() * bound) >> 64);
    }
    double nextDouble() { return (double)(next64() >> 11) * 0x1.0p-53; }
};

// rg::shuffle stand-in: Fisher-Yates from the back, n-1 draws.
template <typename T>
static void shuffle(T* a, size_t n, Rng& r) {
    for (size_t i = n; i-- > 1;) {
        size_t j = (size_t)r.nextInt(i + 1);
        std::swap(a[i], a[j]);
    }
}

// rg::sample stand-in: sorted k-subset of [0,N) (selection sampling, Knuth Alg. S).
static std::vector<uint32_t> sample(Rng& r, uint32_t k, uint32_t N) {
    std::vector<uint32_t> out;
    out.reserve(k);
    uint32_t need = k;
    for (uint32_t i = 0; i < N && need > 0; ++i)
        if (r.nextInt(N - i) < need) { out.push_back(i); --need; }
    return out;
}

// ----------------------------------------------------------------------------
// Schedules.  Row = sub-epoch, column = task id, entry = column block.
// ----------------------------------------------------------------------------
struct Mat {
    uint32_t rows = 0, cols = 0;
    std::vector<uint32_t> a;
    Mat() {}
    Mat(uint32_t r, uint32_t c) : rows(r), cols(c), a((size_t)r * c) {}
    uint32_t& operator()(uint32_t i, uint32_t j) { return a[(size_t)i * cols + j]; }
};

// mf::shuffle (mf/matrix/op/shuffle.h:23-49): row permutation P_r then column
// permutation P_c, each from generateRandomPermutationMatrix (one shuffle of
// the identity).  prod(r, m): new row i = old row v[i]; prod(m, c): new column
// u[j] = old column j.
static void shuffleMatrix(Mat& m, Rng& r) {
    std::vector<uint32_t> v(m.rows), u(m.cols);
    std::iota(v.begin(), v.end(), 0u);
    shuffle(v.data(), v.size(), r);
    std::iota(u.begin(), u.end(), 0u);
    shuffle(u.data(), u.size(), r);
    Mat t(m.rows, m.cols);
    for (uint32_t i = 0; i < m.rows; ++i)
        for (uint32_t j = 0; j < m.cols; ++j) t(i, j) = m(v[i], j);
    for (uint32_t i = 0; i < m.rows; ++i)
        for (uint32_t j = 0; j < m.cols; ++j) m(i, u[j]) = t(i, j);
}

// computeDsgdSchedule (mf/sgd/dsgd_impl.h:246-325)
static Mat dsgdSchedule(uint32_t w, uint32_t t, int order, Rng& r) {
    uint32_t d = w * t;
    Mat s(d, d);
    for (uint32_t e = 0; e < d; ++e)
        for (uint32_t id = 0; id < d; ++id) s(e, id) = (e + id) % d;   // :253
    switch (order) {
    case 0: break;                                                    // SEQ :257
    case 1: {                                                         // RSEQ :260-273
        std::vector<uint32_t> b(d);
        std::iota(b.begin(), b.end(), 0u);
        shuffle(b.data(), d, r);
        for (auto& x : s.a) x = b[x];
        break;
    }
    case 2:                                                           // WR :275-283
        for (uint32_t e = 0; e < d; ++e) shuffle(&s.a[(size_t)e * d], d, r);
        break;
    case 3: shuffleMatrix(s, r); break;                               // WOR :284-287
    case 4: {                                                         // COWOR :288-318
        std::vector<uint32_t> b(d);
        std::iota(b.begin(), b.end(), 0u);
        shuffle(b.data(), d, r);
        Mat g = dsgdSchedule(w, 1, 3, r);
        for (uint32_t bi = 0; bi < w; ++bi)
            for (uint32_t bj = 0; bj < w; ++bj) {
                uint32_t go = g(bi, bj) * t;
                Mat l = dsgdSchedule(t, 1, 3, r);
                for (uint32_t ti = 0; ti < t; ++ti)
                    for (uint32_t tj = 0; tj < t; ++tj)
                        s(bi * t + ti, bj * t + tj) = b[go + l(ti, tj)];
            }
        break;
    }
    default: throw std::runtime_error("Invalid stratum order");
    }
    return s;
}

// DsgdPpRunner::epoch schedule part (mf/sgd/dsgdpp_impl.h:289-336) + the task
// seeds that mpi2::seed sends (:344): one next64() per task in task order.
static Mat dsgdppSchedule(uint32_t w, uint32_t t, int order, Rng& r,
                          std::vector<uint64_t>& seeds) {
    uint32_t d = w * t;
    std::vector<uint32_t> b0(d), b1(d);
    if (order == 0) {
        for (uint32_t b = 0; b < d; ++b) { b0[b] = 2 * b; b1[b] = 2 * b + 1; }   // :292-299
    } else if (order >= 1 && order <= 4) {
        b0 = sample(r, d, 2 * d);                                               // :307
        uint32_t i0 = 0, i1 = 0;
        for (uint32_t b = 0; i1 < d && b < 2 * d; ++b) {                         // :308-317
            if (i0 < d && b0[i0] == b) { ++i0; continue; }
            b1[i1++] = b;
        }
    } else {
        throw std::runtime_error("Invalid stratum order");
    }
    Mat s0 = dsgdSchedule(w, t, order, r);                                      // :327
    Mat s1 = dsgdSchedule(w, t, order, r);                                      // :328
    Mat s(2 * d, d);
    for (uint32_t i = 0; i < d; ++i)
        for (uint32_t id = 0; id < d; ++id) {
            s(2 * i, id) = b0[s0(i, id)];                                        // :333
            s(2 * i + 1, id) = b1[s1(i, id)];                                    // :334
        }
    seeds.resize(d);
    for (uint32_t id = 0; id < d; ++id) seeds[id] = r.next64();
    return s;
}

// SgdRunner::permute (mf/sgd/sgd_impl.h:255-278): reset to the identity only
// when the size differs; first draw nextInt(n), then nextInt(n-s) for s=1..k-1
// (the last, nextInt(1), is drawn but unused).
static void permute(Rng& r, std::vector<uint32_t>& p, uint64_t n, uint64_t k) {
    if (p.size() != n) {
        p.resize(n);
        for (uint64_t i = 0; i < n; ++i) p[i] = (uint32_t)i;
    }
    uint64_t cur = 0, next = r.nextInt(n);
    for (uint64_t s = 1; s < k; ++s) {
        cur = next;
        next = s + r.nextInt(n - s);
        std::swap(p[s - 1], p[cur]);
    }
}

// computeDefaultBlockOffsets (mf/matrix/distributed_matrix_impl.cc:24-40);
// off has blocks+1 entries, off[blocks] = size.
static std::vector<int64_t> blockOffsets(int64_t size, int64_t blocks) {
    std::vector<int64_t> off(blocks + 1);
    int64_t minSize = size / blocks, rem = size % blocks;
    for (int64_t i = 0; i < blocks; ++i) {
        if (i == 0) off[i] = 0;
        else {
            off[i] = minSize + off[i - 1];
            if (rem > 0) { ++off[i]; --rem; }
        }
    }
    off[blocks] = size;
    return off;
}

// ----------------------------------------------------------------------------
// Functor strings (tools/parse.h:39-61): "Name(a,b,...)" or "Name".
// ----------------------------------------------------------------------------
static bool parseArg(const std::string& s, std::string& name, std::vector<double>& args) {
    args.clear();
    if (s.size() >= 2 && s.back() == ')') {
        size_t p = s.find('(');
        if (p != std::string::npos) {
            name = s.substr(0, p);
            std::string inner = s.substr(p + 1, s.size() - p - 2);
            size_t st = 0;
            while (true) {
                size_t e = inner.find(',', st);
                std::string tok = inner.substr(st, e == std::string::npos ? std::string::npos : e - st);
                char* endp = nullptr;
                const char* c = tok.c_str();
                double v = strtod(c, &endp);
                if (endp == c || *endp != 0) return false;   // lexical_cast rejects spaces
                args.push_back(v);
                if (e == std::string::npos) break;
                st = e + 1;
            }
            return true;
        }
    }
    name = s;
    return true;
}

// ----------------------------------------------------------------------------
// Problem data.
// ----------------------------------------------------------------------------
struct Block {
    std::vector<uint32_t> i, j;   // block-local, sorted row-major (read_impl.h:163)
    std::vector<double> x;
};

struct Barrier {
    std::mutex m;
    std::condition_variable cv;
    int n, waiting = 0;
    uint64_t gen = 0;
    explicit Barrier(int n) : n(n) {}
    void wait() {
        std::unique_lock<std::mutex> l(m);
        uint64_t g = gen;
        if (++waiting == n) { waiting = 0; ++gen; cv.notify_all(); }
        else cv.wait(l, [&] { return gen != g; });
    }
};

enum UpdKind { U_NZSL = 0, U_NZSL_L2 = 1, U_NZSL_NZL2 = 2 };

struct Ctx {
    std::string err;
    // configuration
    int updKind = U_NZSL_L2;
    double updLambda = 0, lossLambda = 0;
    bool truncate = false;
    double tmin = NAN, tmax = NAN;
    std::string updateStr, regularizeStr, lossStr;
    int64_t r = 0;
    uint32_t w = 1, t = 1, d = 1;
    // algorithm: 0 DSGD++ (d x 2d blocks, 2d sub-epochs), 1 DSGD (d x d, d sub-epochs)
    int algo = 0;
    uint32_t nb2 = 2;   // column blocks
    int sgdOrder = 2, stratumOrder = 4;
    uint64_t seed = 0;
    int threads = 1;
    Rng master;
    // data
    int64_t m = -1, n = -1;
    std::vector<int64_t> off1, off2;
    std::vector<Block> train, test;
    bool haveTrain = false, haveTest = false;
    int64_t nnzTrain = 0, nnzTest = 0;
    std::vector<int64_t> nnz1, nnz2;
    int64_t nnz12max = 0;
    std::vector<double> W, H;
    bool haveFactors = false;
    // last epoch
    Mat schedule;
    std::vector<uint64_t> seeds;
    bool record = false;
    std::vector<std::vector<uint32_t>> recPerm;   // [task*2d + subepoch]
    int arith = 0;   // 0 FP64 reference; 1 FP32 / 2 BF16 / 3 FP64 grouped-GPU emulation
    int mutant = 0;  // planted mistake of an emulation (UpdateArgs::mutant), pin tests only
};

static void setErr(Ctx* c, const std::string& e) { c->err = e; }

static int blockIndex(const Ctx* c, int64_t b1, int64_t b2) { return (int)(b1 * c->nb2 + b2); }

// ReadMatrixBlocksMm (read_impl.h:365-486) + block sort (read_impl.h:163):
// entries go to the block of (i,j) with block-local indices; each block is then
// sorted row-major.  ublas coordinate_matrix::sort sums duplicates; here the
// sort is stable and duplicates are summed in input order.
static void blockCoo(Ctx* c, int64_t m, int64_t n, int64_t nnz, const int32_t* I,
                     const int32_t* J, const double* X, std::vector<Block>& out) {
    uint32_t d = c->d, d2 = c->nb2;
    std::vector<uint32_t> rb(m), cb(n);
    for (uint32_t b = 0; b < d; ++b)
        for (int64_t i = c->off1[b]; i < c->off1[b + 1]; ++i) rb[i] = b;
    for (uint32_t b = 0; b < d2; ++b)
        for (int64_t j = c->off2[b]; j < c->off2[b + 1]; ++j) cb[j] = b;
    std::vector<int64_t> cnt((size_t)d * d2 + 1, 0);
    for (int64_t p = 0; p < nnz; ++p) cnt[(size_t)rb[I[p]] * d2 + cb[J[p]] + 1]++;
    for (size_t b = 0; b < (size_t)d * d2; ++b) cnt[b + 1] += cnt[b];
    std::vector<int64_t> order(nnz);
    {
        std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
        for (int64_t p = 0; p < nnz; ++p) order[pos[(size_t)rb[I[p]] * d2 + cb[J[p]]]++] = p;
    }
    out.assign((size_t)d * d2, Block());
    auto work = [&](size_t lo, size_t hi) {
        for (size_t b = lo; b < hi; ++b) {
            int64_t s = cnt[b], e = cnt[b + 1];
            std::vector<int64_t> idx(order.begin() + s, order.begin() + e);
            std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t q) {
                if (I[a] != I[q]) return I[a] < I[q];
                return J[a] < J[q];
            });
            Block& B = out[b];
            uint32_t b1 = (uint32_t)(b / d2), b2 = (uint32_t)(b % d2);
            for (size_t k = 0; k < idx.size(); ++k) {
                int64_t p = idx[k];
                uint32_t li = (uint32_t)(I[p] - c->off1[b1]), lj = (uint32_t)(J[p] - c->off2[b2]);
                if (!B.i.empty() && B.i.back() == li && B.j.back() == lj) {
                    B.x.back() += X[p];
                } else {
                    B.i.push_back(li); B.j.push_back(lj); B.x.push_back(X[p]);
                }
            }
        }
    };
    size_t nb = (size_t)d * d2;
    int th = std::max(1, std::min<int>(c->threads, (int)nb));
    std::vector<std::thread> ts;
    for (int k = 0; k < th; ++k) ts.emplace_back(work, nb * k / th, nb * (k + 1) / th);
    for (auto& x : ts) x.join();
}

// nnz12 (mf/matrix/op/nnz.h:34-58), summed over blocks
static void computeNnz12(Ctx* c) {
    c->nnz1.assign(c->m, 0);
    c->nnz2.assign(c->n, 0);
    c->nnzTrain = 0;
    for (uint32_t b1 = 0; b1 < c->d; ++b1)
        for (uint32_t b2 = 0; b2 < c->nb2; ++b2) {
            const Block& B = c->train[blockIndex(c, b1, b2)];
            for (size_t p = 0; p < B.i.size(); ++p) {
                c->nnz1[c->off1[b1] + B.i[p]]++;
                c->nnz2[c->off2[b2] + B.j[p]]++;
            }
            c->nnzTrain += (int64_t)B.i.size();
        }
    c->nnz12max = 0;
    for (auto v : c->nnz1) c->nnz12max = std::max(c->nnz12max, v);
    for (auto v : c->nnz2) c->nnz12max = std::max(c->nnz12max, v);
}

// ----------------------------------------------------------------------------
// Update functors on global row i / column j.
// ----------------------------------------------------------------------------
struct UpdateArgs {
    int kind; double lambda; int64_t r;
    bool truncate; double tmin, tmax;
    // Planted mistakes for the tests that show the emulation's pins catch
    // them (0 = none): 1 drops the regularization term, 2 updates h with the
    // new w, 3 swaps the row and column counts.  Emulations only.
    int mutant = 0;
};

// UpdateNzsl (update-nzsl.h:34-49), UpdateNzslL2 (update-nzsl-l2.h:34-57),
// UpdateNzslNzl2 (update-nzsl-nzl2.h:34-53), UpdateTruncate (update-truncate.h:46-66).
static inline void updatePoint(const UpdateArgs& a, double* w, double* h, double x, double eps,
                               double n1, double n2) {
    const int64_t r = a.r;
    double wh = 0;
    for (int64_t z = 0; z < r; ++z) wh += w[z] * h[z];
    if (a.kind == U_NZSL) {
        double f = eps * -2. * (x - wh);
        for (int64_t z = 0; z < r; ++z) {
            double t = w[z];
            w[z] -= f * h[z];
            h[z] -= f * t;
        }
    } else if (a.kind == U_NZSL_L2) {
        double f1 = eps * -2. * (x - wh);
        double f2 = eps * 2. * a.lambda;
        double f3 = 1. / n1;
        double f4 = 1. / n2;
        for (int64_t z = 0; z < r; ++z) {
            double t = w[z];
            w[z] -= f1 * h[z] + f2 * w[z] * f3;
            h[z] -= f1 * t + f2 * h[z] * f4;
        }
    } else {
        double f1 = eps * -2. * (x - wh);
        double f2 = eps * 2. * a.lambda;
        for (int64_t z = 0; z < r; ++z) {
            double t = w[z];
            w[z] -= f1 * h[z] + f2 * w[z];
            h[z] -= f1 * t + f2 * h[z];
        }
    }
    if (a.truncate) {
        for (int64_t z = 0; z < r; ++z) {
            double& ww = w[z];
            if (ww <= a.tmin) ww = a.tmin + eps;
            else if (ww >= a.tmax) ww = a.tmax - eps;
            double& hh = h[z];
            if (hh <= a.tmin) hh = a.tmin + eps;
            else if (hh >= a.tmax) hh = a.tmax - eps;
        }
    }
}

// Float emulation of the GPU's grouped FP32/BF16 update (not part of the
// reference): lane l of a 16-lane row holds elements [4l, 4l+4); its partial
// dot is a 4-term fmaf chain; the row sum is ((q0+q1)+(q2+q3)) with
// q_k = (d_4k + d_4k+1) + (d_4k+2 + d_4k+3); the update is
// w' = fmaf(-f1, h, fmaf(-a, w, w)), h' = fmaf(-f1, w, fmaf(-b, h, h)); with
// BF16 storage every result is rounded to bf16 (nearest even).
static inline float bf16Round(float f) {
    uint32_t b;
    memcpy(&b, &f, 4);
    if ((b & 0x7fffffffu) > 0x7f800000u) b = (b | 0x00400000u) & 0xffff0000u;
    else { b += 0x7fffu + ((b >> 16) & 1u); b &= 0xffff0000u; }
    float o;
    memcpy(&o, &b, 4);
    return o;
}

static inline void updatePointF(const UpdateArgs& a, int mode, double* w, double* h, double xd, double epsd,
                                double n1, double n2) {
    // 16 lanes per point; lane l holds elements [E l, E l + E), E = 4 up to
    // rank 64 and 8 up to rank 128, and sums its products in element order
    const int64_t r = a.r;
    const int E = r > 64 ? 8 : 4;
    float wf[128], hf[128];
    for (int64_t z = 0; z < 16 * E; ++z) {
        wf[z] = z < r ? (float)w[z] : 0.f;
        hf[z] = z < r ? (float)h[z] : 0.f;
    }
    float dl[16];
    for (int l = 0; l < 16; ++l) {
        float p = wf[E * l] * hf[E * l];
        for (int e = 1; e < E; ++e) p = fmaf(wf[E * l + e], hf[E * l + e], p);
        dl[l] = p;
    }
    float q[4];
    for (int k = 0; k < 4; ++k) q[k] = (dl[4 * k] + dl[4 * k + 1]) + (dl[4 * k + 2] + dl[4 * k + 3]);
    const float wh = (q[0] + q[1]) + (q[2] + q[3]);
    const float eps = (float)epsd, x = (float)xd, lam = (float)a.lambda;
    const float f1 = (eps * -2.f) * (x - wh);
    const float f2 = (eps * 2.f) * lam;
    float aa = 0.f, bb = 0.f;
    if (a.mutant == 3) std::swap(n1, n2);
    if (a.kind == U_NZSL_L2) { aa = f2 * (float)(1. / n1); bb = f2 * (float)(1. / n2); }
    else if (a.kind == U_NZSL_NZL2) { aa = f2; bb = f2; }
    if (a.mutant == 1) aa = bb = 0.f;
    const float tmin = (float)a.tmin, tmax = (float)a.tmax;
    for (int64_t z = 0; z < r; ++z) {
        float wn = fmaf(-f1, hf[z], fmaf(-aa, wf[z], wf[z]));
        float hn = fmaf(-f1, a.mutant == 2 ? wn : wf[z], fmaf(-bb, hf[z], hf[z]));
        if (a.truncate) {
            wn = (wn <= tmin) ? tmin + eps : ((wn >= tmax) ? tmax - eps : wn);
            hn = (hn <= tmin) ? tmin + eps : ((hn >= tmax) ? tmax - eps : hn);
        }
        if (mode == 2) { wn = bf16Round(wn); hn = bf16Round(hn); }
        w[z] = wn;
        h[z] = hn;
    }
}

// Double emulation of the GPU's grouped FP64 update (not part of the
// reference; rank <= 64): the lane layout and summation order of the float
// emulation above, in double with fused multiply-adds.
static inline void updatePointD(const UpdateArgs& a, double* w, double* h, double x, double eps, double n1,
                                double n2) {
    const int64_t r = a.r;
    const int L = r > 64 ? 8 : 4;   // elements per lane: 4 up to rank 64, 8 up to 128
    double wd[128], hd[128];
    for (int64_t z = 0; z < 128; ++z) {
        wd[z] = z < r ? w[z] : 0.;
        hd[z] = z < r ? h[z] : 0.;
    }
    double dl[16];
    for (int l = 0; l < 16; ++l) {
        double p = wd[L * l] * hd[L * l];
        for (int e = 1; e < L; ++e) p = std::fma(wd[L * l + e], hd[L * l + e], p);
        dl[l] = p;
    }
    double q[4];
    for (int k = 0; k < 4; ++k) q[k] = (dl[4 * k] + dl[4 * k + 1]) + (dl[4 * k + 2] + dl[4 * k + 3]);
    const double wh = (q[0] + q[1]) + (q[2] + q[3]);
    const double f1 = (eps * -2.) * (x - wh);
    const double f2 = (eps * 2.) * a.lambda;
    double aa = 0., bb = 0.;
    if (a.mutant == 3) std::swap(n1, n2);
    if (a.kind == U_NZSL_L2) { aa = f2 * (1. / n1); bb = f2 * (1. / n2); }
    else if (a.kind == U_NZSL_NZL2) { aa = f2; bb = f2; }
    if (a.mutant == 1) aa = bb = 0.;
    for (int64_t z = 0; z < r; ++z) {
        double wn = std::fma(-f1, hd[z], std::fma(-aa, wd[z], wd[z]));
        double hn = std::fma(-f1, a.mutant == 2 ? wn : wd[z], std::fma(-bb, hd[z], hd[z]));
        if (a.truncate) {
            wn = (wn <= a.tmin) ? a.tmin + eps : ((wn >= a.tmax) ? a.tmax - eps : wn);
            hn = (hn <= a.tmin) ? a.tmin + eps : ((hn >= a.tmax) ? a.tmax - eps : hn);
        }
        w[z] = wn;
        h[z] = hn;
    }
}

// One task's SgdRunner (sgd.h:111; one per task per epoch, dsgdpp_impl.h:112).
struct TaskRunner {
    Rng rng;
    std::vector<uint32_t> perm;
};

// SgdRunner::epoch -> update -> updateSequential/updateWr/updateWor
// (sgd_impl.h:182-351) on block (b1,b2).
static void runBlock(Ctx* c, TaskRunner& tr, uint32_t b1, uint32_t b2, double eps,
                     std::vector<uint32_t>* rec) {
    const Block& B = c->train[blockIndex(c, b1, b2)];
    const uint64_t n = B.i.size();
    const int64_t r = c->r;
    const int64_t o1 = c->off1[b1], o2 = c->off2[b2];
    UpdateArgs ua{c->updKind, c->updLambda, r, c->truncate, c->tmin, c->tmax, c->mutant};
    double* W = c->W.data();
    double* H = c->H.data();
    const int mode = c->arith;
    auto step = [&](uint64_t p) {
        int64_t gi = o1 + B.i[p], gj = o2 + B.j[p];
        if (mode == 0)
            updatePoint(ua, W + gi * r, H + gj * r, B.x[p], eps, (double)c->nnz1[gi], (double)c->nnz2[gj]);
        else if (mode == 3)
            updatePointD(ua, W + gi * r, H + gj * r, B.x[p], eps, (double)c->nnz1[gi], (double)c->nnz2[gj]);
        else
            updatePointF(ua, mode, W + gi * r, H + gj * r, B.x[p], eps, (double)c->nnz1[gi], (double)c->nnz2[gj]);
    };
    if (c->sgdOrder == 0) {                                   // SEQ :198-217
        for (uint64_t p = 0; p < n; ++p) step(p);
        if (rec) { rec->resize(n); std::iota(rec->begin(), rec->end(), 0u); }
    } else if (c->sgdOrder == 1) {                            // WR :220-251
        uint64_t next = tr.rng.nextInt(n), next2 = tr.rng.nextInt(n);
        if (rec) rec->clear();
        for (uint64_t s = 0; s < n; ++s) {
            uint64_t cur = next;
            next = next2;
            next2 = tr.rng.nextInt(n);
            if (rec) rec->push_back((uint32_t)cur);
            step(cur);
        }
    } else {                                                  // WOR :282-351
        if (rec) rec->clear();
        if (n == 0) return;
        permute(tr.rng, tr.perm, n, n);
        const uint32_t* pp = tr.perm.data();
        const uint32_t* bi = B.i.data();
        const uint32_t* bj = B.j.data();
        for (uint64_t s = 0; s < n; ++s) {
            if (s + 1 < n) {                                  // prefetch as :319-327
                uint32_t q = pp[s + 1];
                __builtin_prefetch(W + (o1 + bi[q]) * r, 1, 2);
                __builtin_prefetch(H + (o2 + bj[q]) * r, 1, 2);
            }
            step(pp[s]);
        }
        if (rec) rec->assign(tr.perm.begin(), tr.perm.end());
    }
}

// DsgdPpRunner::epoch (dsgdpp_impl.h:282-349) + DsgdPpTask::run (:78-276) on
// one node: d tasks, 2d sub-epochs, barrier after each sub-epoch (:259-261).
// DSGD: DsgdRunner::epoch (dsgd_impl.h:329-350) + DsgdTask::run (:81-234):
// the d x d schedule of computeDsgdSchedule, then one seed per task in task
// order (mpi2::seed, :344), d sub-epochs, task id on block (id, schedule(s, id))
// (:111-112), one SgdRunner per task per epoch (:106), barrier after each
// sub-epoch (:221-225).
static double epoch(Ctx* c, double eps) {
    auto t0 = std::chrono::steady_clock::now();
    const uint32_t d = c->d;
    if (c->algo == 1) {
        c->schedule = dsgdSchedule(c->w, c->t, c->stratumOrder, c->master);
        c->seeds.resize(d);
        for (uint32_t id = 0; id < d; ++id) c->seeds[id] = c->master.next64();
    } else {
        c->schedule = dsgdppSchedule(c->w, c->t, c->stratumOrder, c->master, c->seeds);
    }
    const uint32_t subs = c->schedule.rows;
    std::vector<TaskRunner> runners(d);
    for (uint32_t id = 0; id < d; ++id) runners[id].rng = Rng::task(c->seeds[id]);
    if (c->record) c->recPerm.assign((size_t)d * subs, std::vector<uint32_t>());
    int th = std::max(1, std::min<int>(c->threads, (int)d));
    auto subepochs = [&](int k, Barrier* bar) {
        for (uint32_t s = 0; s < subs; ++s) {
            for (uint32_t id = k; id < d; id += th) {
                std::vector<uint32_t>* rec = c->record ? &c->recPerm[(size_t)id * subs + s] : nullptr;
                runBlock(c, runners[id], id, c->schedule(s, id), eps, rec);
            }
            if (bar) bar->wait();
        }
    };
    if (th == 1) {
        subepochs(0, nullptr);
    } else {
        Barrier bar(th);
        std::vector<std::thread> ts;
        for (int k = 0; k < th; ++k) ts.emplace_back(subepochs, k, &bar);
        for (auto& x : ts) x.join();
    }
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

// ----------------------------------------------------------------------------
// Losses.
// ----------------------------------------------------------------------------
// nzsl (nzsl.h:30-53) per block, block partials summed row-major (sum.h:28-30)
static double nzsl(const Ctx* c, const std::vector<Block>& blocks) {
    const uint32_t d = c->d, d2 = c->nb2;
    const int64_t r = c->r;
    std::vector<double> part((size_t)d * d2, 0.0);
    auto work = [&](size_t lo, size_t hi) {
        for (size_t b = lo; b < hi; ++b) {
            const Block& B = blocks[b];
            int64_t o1 = c->off1[b / d2], o2 = c->off2[b % d2];
            double res = 0;
            for (size_t p = 0; p < B.i.size(); ++p) {
                const double* w = &c->W[(o1 + B.i[p]) * r];
                const double* h = &c->H[(o2 + B.j[p]) * r];
                double ip = 0;
                for (int64_t z = 0; z < r; ++z) ip += w[z] * h[z];
                double diff = B.x[p] - ip;
                res += diff * diff;
            }
            part[b] = res;
        }
    };
    size_t nb = part.size();
    int th = std::max(1, std::min<int>(c->threads, (int)nb));
    std::vector<std::thread> ts;
    for (int k = 0; k < th; ++k) ts.emplace_back(work, nb * k / th, nb * (k + 1) / th);
    for (auto& x : ts) x.join();
    double s = 0;
    for (double v : part) s += v;
    return s;
}

// l2 of a dense matrix, per block std::accumulate in storage order
// (l2.h:40-44, distributed :141-146), blocks summed in order (sum.h:28-30)
static double l2Blocks(const std::vector<double>& A, const std::vector<int64_t>& off, int64_t r) {
    double total = 0;
    for (size_t b = 0; b + 1 < off.size(); ++b) {
        double acc = 0;
        for (int64_t q = off[b] * r; q < off[b + 1] * r; ++q) acc = acc + A[q] * A[q];
        total += acc;
    }
    return total;
}

// nzl2 (nzl2.h:29-76), accumulator initialised to 0 (the reference leaves it
// uninitialised, nzl2.h:32,67)
static double nzl2Blocks(const std::vector<double>& A, const std::vector<int64_t>& off, int64_t r,
                         const std::vector<int64_t>& cnt) {
    double total = 0;
    for (size_t b = 0; b + 1 < off.size(); ++b) {
        double res = 0;
        for (int64_t i = off[b]; i < off[b + 1]; ++i) {
            double v = 0;
            for (int64_t z = 0; z < r; ++z) { double vv = A[i * r + z]; v += vv * vv; }
            res += (double)cnt[i] * v;
        }
        total += res;
    }
    return total;
}

// SumLoss (loss.h:97-100) of NzslLoss and L2Loss / Nzl2Loss
static double trainLoss(const Ctx* c) {
    double l = nzsl(c, c->train);
    if (c->updKind == U_NZSL_L2)
        l = l + c->lossLambda * (l2Blocks(c->W, c->off1, c->r) + l2Blocks(c->H, c->off2, c->r));
    else if (c->updKind == U_NZSL_NZL2)
        l = l + c->lossLambda * (nzl2Blocks(c->W, c->off1, c->r, c->nnz1) +
                                 nzl2Blocks(c->H, c->off2, c->r, c->nnz2));
    return l;
}

// balanceSimple (balance_impl.cc:102-127)
static void balance(Ctx* c, int type) {
    if (type == 0) return;
    double regW, regH;
    if (type == 1) {
        regW = l2Blocks(c->W, c->off1, c->r);
        regH = l2Blocks(c->H, c->off2, c->r);
    } else {
        regW = nzl2Blocks(c->W, c->off1, c->r, c->nnz1);
        regH = nzl2Blocks(c->H, c->off2, c->r, c->nnz2);
    }
    double wf = std::sqrt(std::sqrt(regH / regW)), hf;
    if (std::isnan(wf)) { wf = 0; hf = 0; }
    else hf = 1. / wf;
    for (auto& v : c->W) v *= wf;
    for (auto& v : c->H) v *= hf;
}

// ----------------------------------------------------------------------------
// Matrix Market I/O (read_impl.h:46-116, 195-256; write_impl.h:94-118).
// ----------------------------------------------------------------------------
static std::string rtrim(const std::string& s) {
    size_t e = s.find_last_not_of(" \t\r\n");
    return e == std::string::npos ? std::string() : s.substr(0, e + 1);
}
static bool blankTail(const char* p) {
    while (*p) { if (!isspace((unsigned char)*p)) return false; ++p; }
    return true;
}

static bool readMmc(const std::string& fn, int64_t& m, int64_t& n, std::vector<int32_t>& I,
                    std::vector<int32_t>& J, std::vector<double>& X, std::string& err) {
    std::ifstream in(fn);
    if (!in.is_open()) { err = "Cannot open file " + fn; return false; }
    std::string line;
    if (!std::getline(in, line)) { err = "Unexpected EOF in file " + fn; return false; }
    if (rtrim(line) != "%%MatrixMarket matrix coordinate real general") {
        err = "Wrong matrix-market banner in file " + fn + ": " + line; return false;
    }
    bool got = false;
    while ((got = (bool)std::getline(in, line)) && !line.empty() && line[0] == '%') {}
    if (!got || (!line.empty() && line[0] == '%')) { err = "Unexpected EOF in file " + fn; return false; }
    long long a, b, nz;
    char junk[2] = {0, 0};
    int k = sscanf(line.c_str(), "%lld %lld %lld %1s", &a, &b, &nz, junk);
    if (k != 3) { err = "Invalid matrix dimensions in file " + fn + ": " + line; return false; }
    m = a; n = b;
    I.resize(nz); J.resize(nz); X.resize(nz);
    for (long long p = 0; p < nz; ++p) {
        if (!std::getline(in, line)) { err = "Unexpected EOF in file " + fn; return false; }
        const char* s = line.c_str();
        char* e = nullptr;
        unsigned long i = strtoul(s, &e, 10);
        if (e == s) { err = "Parse error in " + fn + ": " + line; return false; }
        s = e;
        unsigned long j = strtoul(s, &e, 10);
        if (e == s) { err = "Parse error in " + fn + ": " + line; return false; }
        s = e;
        double x = strtod(s, &e);
        if (e == s || !blankTail(e)) { err = "Parse error in " + fn + ": " + line; return false; }
        I[p] = (int32_t)(i - 1); J[p] = (int32_t)(j - 1); X[p] = x;
    }
    while (std::getline(in, line))
        if (!blankTail(line.c_str())) { err = "Unexpected input in " + fn + ": " + line; return false; }
    return true;
}

static bool readMma(const std::string& fn, int64_t& rows, int64_t& cols, std::vector<double>& cm,
                    std::string& err) {
    std::ifstream in(fn);
    if (!in.is_open()) { err = "Cannot open file " + fn; return false; }
    std::string line;
    if (!std::getline(in, line)) { err = "Unexpected EOF in file " + fn; return false; }
    if (rtrim(line) != "%%MatrixMarket matrix array real general") {
        err = "Wrong matrix-market banner in file " + fn + ": " + line; return false;
    }
    bool got = false;
    while ((got = (bool)std::getline(in, line)) && !line.empty() && line[0] == '%') {}
    if (!got) { err = "Unexpected EOF in file " + fn; return false; }
    long long a, b;
    char junk[2] = {0, 0};
    if (sscanf(line.c_str(), "%lld %lld %1s", &a, &b, junk) != 2) {
        err = "Invalid matrix dimensions in file " + fn + ": " + line; return false;
    }
    rows = a; cols = b;
    cm.resize((size_t)a * b);
    for (size_t p = 0; p < cm.size(); ++p) {
        if (!std::getline(in, line)) { err = "Unexpected EOF in file " + fn; return false; }
        char* e = nullptr;
        double x = strtod(line.c_str(), &e);
        if (e == line.c_str() || !blankTail(e)) { err = "Parse error in " + fn + ": " + line; return false; }
        cm[p] = x;
    }
    return true;
}

static bool writeMma(const std::string& fn, int64_t rows, int64_t cols,
                     const std::function<double(int64_t, int64_t)>& at, std::string& err) {
    std::ofstream out(fn);
    if (!out.is_open()) { err = "Cannot open file " + fn; return false; }
    out.precision(16);
    out << "%%MatrixMarket matrix array real general\n";
    out << "% First line: ROWS COLUMNS\n";
    out << "% Subsequent lines: entries in column-major order\n";
    out << rows << " " << cols << "\n";
    for (int64_t j = 0; j < cols; ++j)
        for (int64_t i = 0; i < rows; ++i) out << at(i, j) << "\n";
    return true;
}

// ----------------------------------------------------------------------------
// Trace (mf/trace.h:46-248).
// ----------------------------------------------------------------------------
struct TraceRow {
    int64_t it; double loss, test, eps, tEps, tEpoch, tLoss, tTest;
};

static std::string traceString(const std::vector<TraceRow>& rows, const std::string& var,
                               const std::map<std::string, double>& dbl,
                               const std::map<std::string, std::string>& str) {
    std::stringstream out;
    out << var << " <- list(" << "trace=list(";
    std::string sep;
    double elapsed = 0;
    for (const auto& e : rows) {
        out << sep << std::endl;
        out << "list(" << "it=" << e.it << ", " << "epoch=" << e.it << ", " << "loss=" << e.loss << ", "
            << "loss.test=";
        if (std::isnan(e.test)) out << "NA" << ", ";
        else out << e.test << ", ";
        elapsed += (e.tEps + e.tEpoch + e.tLoss + e.tTest) / 1E9;
        out << "time=list(elapsed=" << elapsed << ", eps=" << e.tEps / 1E9 << ", epoch=" << e.tEpoch / 1E9
            << ", loss=" << e.tLoss / 1E9 << ", loss.test=" << e.tTest / 1E9 << "), ";
        out << " eps=";
        if (std::isnan(e.eps)) out << "NA";
        else out << e.eps;
        out << ")";
        sep = ", ";
    }
    out << std::endl << ")";
    if (!dbl.empty() || !str.empty()) {
        out << ", info=list(";
        std::string s2;
        for (const auto& kv : dbl) { out << s2 << kv.first << " = " << kv.second; s2 = ", "; }
        s2 = dbl.empty() ? "" : ", ";
        for (const auto& kv : str) { out << s2 << kv.first << " = \"" << kv.second << "\""; s2 = ", "; }
        out << ")";
    }
    out << ");";
    return out.str();
}

static bool configure(Ctx* c, const char* update, const char* regularize, const char* loss,
                      const char* truncate) {
    std::string un, rn, ln;
    std::vector<double> ua, ra, la;
    c->updateStr = update ? update : "";
    c->regularizeStr = regularize ? regularize : "";
    c->lossStr = loss ? loss : "";
    if (!parseArg(c->updateStr, un, ua) || !parseArg(c->regularizeStr, rn, ra) ||
        !parseArg(c->lossStr, ln, la)) {
        setErr(c, "Invalid argument in functor string"); return false;
    }
    // runArgs accepted combinations (tools/mfdsgdpp.cc:162-358)
    if (rn != "None" || !ra.empty()) { setErr(c, "Invalid combination of update, regularize, and loss arguments"); return false; }
    if (un == "Nzsl" && ln == "Nzsl") {
        if (!ua.empty() || !la.empty()) { setErr(c, "Invalid number of arguments"); return false; }
        c->updKind = U_NZSL;
    } else if (un == "Nzsl_L2" && ln == "Nzsl_L2") {
        if (ua.size() != 1 || la.size() != 1) { setErr(c, "Invalid number of arguments"); return false; }
        c->updKind = U_NZSL_L2; c->updLambda = ua[0]; c->lossLambda = la[0];
    } else if (un == "Nzsl_Nzl2" && ln == "Nzsl_Nzl2") {
        if (ua.size() != 1 || la.size() != 1) { setErr(c, "Invalid number of arguments"); return false; }
        c->updKind = U_NZSL_NZL2; c->updLambda = ua[0]; c->lossLambda = la[0];
    } else {
        setErr(c, "Invalid combination of update, regularize, and loss arguments"); return false;
    }
    c->truncate = false;
    if (truncate && *truncate) {   // parseTruncate (parse.h:115-139): "(min,max)"
        std::string tn; std::vector<double> ta;
        if (!parseArg(truncate, tn, ta) || !tn.empty() || ta.size() != 2) {
            setErr(c, "Error in argument truncate"); return false;
        }
        c->truncate = true; c->tmin = ta[0]; c->tmax = ta[1];
    }
    return true;
}

}  // namespace orc

using namespace orc;

struct orc_ctx : public Ctx {};

extern "C" {

orc_ctx* orc_create(const char* update, const char* regularize, const char* loss,
                    const char* truncate, int64_t rank, int world, int tasks_per_rank,
                    int sgd_order, int stratum_order, uint64_t seed, int threads, char* err,
                    int errlen) {
    orc_ctx* c = new orc_ctx();
    auto fail = [&](const std::string& e) -> orc_ctx* {
        if (err && errlen > 0) snprintf(err, errlen, "%s", e.c_str());
        delete c;
        return nullptr;
    };
    if (!configure(c, update, regularize, loss, truncate)) return fail(c->err);
    if (rank <= 0) return fail("rank must be positive");
    if (world <= 0 || tasks_per_rank <= 0) return fail("invalid blocking");
    if (sgd_order < 0 || sgd_order > 2) return fail("Invalid sgd order");
    if (stratum_order < 0 || stratum_order > 4) return fail("Invalid stratum order");
    c->r = rank; c->w = world; c->t = tasks_per_rank; c->d = world * tasks_per_rank;
    c->nb2 = 2 * c->d;
    c->sgdOrder = sgd_order; c->stratumOrder = stratum_order; c->seed = seed;
    c->threads = threads > 0 ? threads : 1;
    c->master = Rng::master(seed);
    return c;
}

void orc_destroy(orc_ctx* c) { delete c; }
const char* orc_last_error(orc_ctx* c) { return c ? c->err.c_str() : "null context"; }

int orc_load_coo(orc_ctx* c, int which, int64_t m, int64_t n, int64_t nnz, const int32_t* i,
                 const int32_t* j, const double* x) {
    if (which != 0 && which != 1) { setErr(c, "which must be 0 or 1"); return -1; }
    if (c->m >= 0 && (c->m != m || c->n != n)) { setErr(c, "dimension mismatch"); return -1; }
    if (m < (int64_t)c->d || n < (int64_t)c->nb2) { setErr(c, "matrix smaller than the block grid"); return -1; }
    for (int64_t p = 0; p < nnz; ++p)
        if (i[p] < 0 || i[p] >= m || j[p] < 0 || j[p] >= n) { setErr(c, "index out of range"); return -1; }
    c->m = m; c->n = n;
    c->off1 = blockOffsets(m, c->d);
    c->off2 = blockOffsets(n, c->nb2);
    if (which == 0) {
        blockCoo(c, m, n, nnz, i, j, x, c->train);
        c->haveTrain = true;
        computeNnz12(c);
    } else {
        blockCoo(c, m, n, nnz, i, j, x, c->test);
        c->haveTest = true;
        c->nnzTest = 0;
        for (auto& B : c->test) c->nnzTest += (int64_t)B.i.size();
    }
    return 0;
}

int orc_load_mmc(orc_ctx* c, int which, const char* path) {
    int64_t m, n;
    std::vector<int32_t> I, J;
    std::vector<double> X;
    if (!readMmc(path, m, n, I, J, X, c->err)) return -1;
    return orc_load_coo(c, which, m, n, (int64_t)X.size(), I.data(), J.data(), X.data());
}

int orc_set_factors(orc_ctx* c, const double* W, const double* H) {
    if (c->m < 0) { setErr(c, "load the data matrix first"); return -1; }
    c->W.assign(W, W + c->m * c->r);
    c->H.assign(H, H + c->n * c->r);
    c->haveFactors = true;
    return 0;
}

int orc_load_factors_mma(orc_ctx* c, const char* wpath, const char* hpath) {
    int64_t a, b;
    std::vector<double> cm;
    if (!readMma(wpath, a, b, cm, c->err)) return -1;
    if (a != c->m || b != c->r) { setErr(c, "row factor dimensions do not match"); return -1; }
    std::vector<double> W((size_t)a * b);
    for (int64_t z = 0; z < b; ++z)
        for (int64_t i = 0; i < a; ++i) W[i * b + z] = cm[z * a + i];
    if (!readMma(hpath, a, b, cm, c->err)) return -1;
    if (a != c->r || b != c->n) { setErr(c, "column factor dimensions do not match"); return -1; }
    return orc_set_factors(c, W.data(), cm.data());
}

int orc_get_factors(orc_ctx* c, double* W, double* H) {
    if (!c->haveFactors) { setErr(c, "no factors"); return -1; }
    if (W) std::copy(c->W.begin(), c->W.end(), W);
    if (H) std::copy(c->H.begin(), c->H.end(), H);
    return 0;
}

int orc_write_factors_mma(orc_ctx* c, const char* wpath, const char* hpath) {
    int64_t r = c->r;
    if (wpath && *wpath &&
        !writeMma(wpath, c->m, r, [&](int64_t i, int64_t z) { return c->W[i * r + z]; }, c->err))
        return -1;
    if (hpath && *hpath &&
        !writeMma(hpath, r, c->n, [&](int64_t z, int64_t j) { return c->H[j * r + z]; }, c->err))
        return -1;
    return 0;
}

int orc_epoch(orc_ctx* c, double eps, double* seconds) {
    if (!c->haveTrain || !c->haveFactors) { setErr(c, "data and factors required"); return -1; }
    double s = epoch(c, eps);
    if (seconds) *seconds = s;
    return 0;
}

int orc_loss(orc_ctx* c, int which, double* out) {
    if (!c->haveFactors) { setErr(c, "no factors"); return -1; }
    if (which == 0) {
        if (!c->haveTrain) { setErr(c, "no training data"); return -1; }
        *out = trainLoss(c);
    } else {
        if (!c->haveTest) { setErr(c, "no test data"); return -1; }
        *out = nzsl(c, c->test);
    }
    return 0;
}

int orc_balance(orc_ctx* c, int type) {
    balance(c, type);
    return 0;
}

int orc_get_schedule(orc_ctx* c, uint32_t* out) {
    if (c->schedule.a.empty()) { setErr(c, "no epoch run yet"); return -1; }
    std::copy(c->schedule.a.begin(), c->schedule.a.end(), out);
    return 0;
}

int orc_set_record(orc_ctx* c, int on) { c->record = on != 0; return 0; }

int orc_set_algorithm(orc_ctx* c, int algo) {
    if (algo != 0 && algo != 1) { setErr(c, "algorithm must be 0 (DSGD++) or 1 (DSGD)"); return -1; }
    if (c->m >= 0) { setErr(c, "set the algorithm before loading data"); return -1; }
    c->algo = algo;
    c->nb2 = algo == 1 ? c->d : 2 * c->d;
    return 0;
}

int orc_set_arith(orc_ctx* c, int mode) {
    const int mutant = mode >> 4;
    mode &= 15;
    if (mode < 0 || mode > 3 || mutant < 0 || mutant > 3 || (mutant && mode == 0)) {
        setErr(c, "invalid arithmetic mode"); return -1;
    }
    c->mutant = mutant;
    if ((mode == 1 || mode == 2) && c->r > 128) { setErr(c, "float emulation needs rank <= 128"); return -1; }
    if (mode == 3 && c->r > 128) { setErr(c, "FP64 grouped emulation needs rank <= 128"); return -1; }
    c->arith = mode;
    // values as the GPU stores them: x as float, factors rounded to the storage type
    if (mode == 1 || mode == 2) {
        for (auto* blocks : {&c->train, &c->test})
            for (auto& B : *blocks)
                for (auto& v : B.x) v = (float)v;
        for (auto* A : {&c->W, &c->H})
            for (auto& v : *A) v = mode == 2 ? bf16Round((float)v) : (float)v;
    }
    return 0;
}

int orc_get_permutation(orc_ctx* c, int task, int subepoch, uint32_t* out, int64_t cap, int64_t* n) {
    if (!c->record || c->recPerm.empty()) { setErr(c, "recording off"); return -1; }
    if (task < 0 || task >= (int)c->d || subepoch < 0 || subepoch >= (int)c->schedule.rows) {
        setErr(c, "task/subepoch out of range"); return -1;
    }
    const auto& p = c->recPerm[(size_t)task * c->schedule.rows + subepoch];
    *n = (int64_t)p.size();
    if (out) {
        if (cap < (int64_t)p.size()) { setErr(c, "buffer too small"); return -1; }
        std::copy(p.begin(), p.end(), out);
    }
    return 0;
}

int orc_get_task_seeds(orc_ctx* c, uint64_t* out) {
    if (c->seeds.empty()) { setErr(c, "no epoch run yet"); return -1; }
    std::copy(c->seeds.begin(), c->seeds.end(), out);
    return 0;
}

int orc_block_nnz(orc_ctx* c, int which, int64_t* out) {
    const auto& v = which == 0 ? c->train : c->test;
    if (v.empty()) { setErr(c, "no data"); return -1; }
    for (size_t b = 0; b < v.size(); ++b) out[b] = (int64_t)v[b].i.size();
    return 0;
}

int orc_get_block(orc_ctx* c, int which, int b1, int b2, int32_t* i, int32_t* j, double* x,
                  int64_t cap, int64_t* n) {
    const auto& v = which == 0 ? c->train : c->test;
    if (v.empty()) { setErr(c, "no data"); return -1; }
    const Block& B = v[blockIndex(c, b1, b2)];
    *n = (int64_t)B.i.size();
    if (cap < *n) return *n == 0 ? 0 : -1;
    for (int64_t p = 0; p < *n; ++p) {
        i[p] = (int32_t)(B.i[p] + c->off1[b1]);
        j[p] = (int32_t)(B.j[p] + c->off2[b2]);
        x[p] = B.x[p];
    }
    return 0;
}

int orc_dims(orc_ctx* c, int64_t* out) {
    out[0] = c->m; out[1] = c->n; out[2] = c->r; out[3] = c->d;
    out[4] = c->nnzTrain; out[5] = c->nnzTest;
    return 0;
}

int orc_run(orc_ctx* c, int64_t epochs, const char* decay, int bal, const char* trace_path,
            const char* trace_var, const char* input_file, double* lossOut, double* testOut,
            double* epsOut, double* timeEpochOut, int verbose) {
    // parseDecay (parse.h:142-235): Const(e), BoldDriver(e), BoldDriver(e,dec,inc)
    std::string dn;
    std::vector<double> da;
    if (!parseArg(decay ? decay : "", dn, da)) { setErr(c, "invalid decay"); return -1; }
    bool bold;
    double eps0, dec = 0.5, inc = 1.05;
    if (dn == "BoldDriver" && (da.size() == 1 || da.size() == 3)) {
        bold = true; eps0 = da[0];
        if (da.size() == 3) { dec = da[1]; inc = da[2]; }
    } else if (dn == "Const" && da.size() == 1) {
        bold = false; eps0 = da[0];   // the reference never stores Const's epsilon (parse.h:148-153)
    } else {
        setErr(c, "Unknown or invalid decay"); return -1;
    }
    if (!c->haveTrain || !c->haveFactors) { setErr(c, "data and factors required"); return -1; }
    using clk = std::chrono::steady_clock;
    auto ns = [](clk::time_point a, clk::time_point b) {
        return (double)std::chrono::duration_cast<std::chrono::nanoseconds>(b - a).count();
    };
    std::vector<TraceRow> rows;
    // defaultRunner (sgd_impl.h:34-145)
    auto t0 = clk::now();
    double cur = trainLoss(c);
    auto t1 = clk::now();
    double tl = ns(t0, t1);
    if (verbose) printf("Loss: %g (%.6fs)\n", cur, tl / 1e9);
    double test = NAN, tt = 0;
    if (c->haveTest) {
        t0 = clk::now(); test = nzsl(c, c->test); t1 = clk::now(); tt = ns(t0, t1);
        if (verbose) printf("Test loss: %g (%.6fs)\n", test, tt / 1e9);
    }
    rows.push_back({0, cur, test, NAN, 0, 0, tl, tt});
    double prev = 0, eps = eps0;
    for (int64_t e = 0; e < epochs; ++e) {
        t0 = clk::now();
        if (e == 0) eps = eps0;                                   // BoldDriver :59-67
        else if (bold) eps = (prev <= cur) ? eps * dec : eps * inc;
        t1 = clk::now();
        double te = ns(t0, t1);
        if (verbose) printf("Step size: %g\nStarting epoch %lld\n", eps, (long long)e + 1);
        t0 = clk::now(); epoch(c, eps); t1 = clk::now();
        double tep = ns(t0, t1);
        if (verbose) printf("Finished epoch %lld (%.6fs)\n", (long long)e + 1, tep / 1e9);
        balance(c, bal);
        prev = cur;
        t0 = clk::now(); cur = trainLoss(c); t1 = clk::now(); tl = ns(t0, t1);
        if (verbose) printf("Loss: %g (%.6fs)\n", cur, tl / 1e9);
        test = NAN; tt = 0;
        if (c->haveTest) {
            t0 = clk::now(); test = nzsl(c, c->test); t1 = clk::now(); tt = ns(t0, t1);
            if (verbose) printf("Test loss: %g (%.6fs)\n", test, tt / 1e9);
        }
        rows.push_back({e + 1, cur, test, eps, te, tep, tl, tt});
    }
    for (size_t k = 0; k < rows.size(); ++k) {
        if (lossOut) lossOut[k] = rows[k].loss;
        if (testOut) testOut[k] = rows[k].test;
        if (epsOut) epsOut[k] = rows[k].eps;
        if (timeEpochOut) timeEpochOut[k] = rows[k].tEpoch / 1e9;
    }
    if (trace_path && *trace_path) {
        // Args::createTraceFields (tools/detail/mfdsgd-args.h:53-63)
        std::map<std::string, double> dbl{{"rank", (double)c->r}, {"nodes", (double)c->w},
                                          {"threads", (double)c->t}};
        std::map<std::string, std::string> str{{"update", c->updateStr},
                                               {"regularize", c->regularizeStr},
                                               {"input_file", input_file ? input_file : ""},
                                               {"sample_matrix", ""},
                                               {"barrier", "false"}};
        std::string s = traceString(rows, trace_var && *trace_var ? trace_var : "trace", dbl, str);
        std::ofstream out(trace_path);
        out << s;
    }
    return 0;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    philox(c, key[0], key[1]);
    for (int k = 0; k < 4; ++k) out[k] = c[k];
}

void orc_rng_draws(uint64_t seed, int kind, uint64_t bound, int64_t count, uint64_t* out) {
    Rng r = kind == 0 ? Rng::master(seed) : Rng::task(seed);
    for (int64_t k = 0; k < count; ++k) out[k] = bound ? r.nextInt(bound) : r.next64();
}

int orc_dsgd_schedule(int world, int tasks, int order, uint64_t seed, uint32_t* out) {
    Rng r = Rng::master(seed);
    try {
        Mat s = dsgdSchedule(world, tasks, order, r);
        std::copy(s.a.begin(), s.a.end(), out);
    } catch (...) { return -1; }
    return 0;
}

int orc_dsgdpp_schedule(int world, int tasks, int order, uint64_t seed, uint32_t* out,
                        uint64_t* task_seeds) {
    Rng r = Rng::master(seed);
    std::vector<uint64_t> seeds;
    try {
        Mat s = dsgdppSchedule(world, tasks, order, r, seeds);
        std::copy(s.a.begin(), s.a.end(), out);
        if (task_seeds) std::copy(seeds.begin(), seeds.end(), task_seeds);
    } catch (...) { return -1; }
    return 0;
}

void orc_block_offsets(int64_t size, int64_t blocks, int64_t* out) {
    auto o = blockOffsets(size, blocks);
    std::copy(o.begin(), o.end(), out);
}

void orc_update_point(int kind, double lambda, int64_t r, double* w, double* h, double x,
                      double eps, double nnz1, double nnz2, double tmin, double tmax,
                      int truncate) {
    UpdateArgs a{kind, lambda, r, truncate != 0, tmin, tmax};
    updatePoint(a, w, h, x, eps, nnz1, nnz2);
}

int orc_update_point_arith(int arith, int mutant, int kind, double lambda, int64_t r, double* w, double* h,
                           double x, double eps, double nnz1, double nnz2, double tmin, double tmax,
                           int truncate) {
    if (arith < 0 || arith > 3 || (arith > 0 && r > 128) || mutant < 0 || mutant > 3 || (mutant && !arith))
        return -1;
    UpdateArgs a{kind, lambda, r, truncate != 0, tmin, tmax, mutant};
    if (arith == 0) updatePoint(a, w, h, x, eps, nnz1, nnz2);
    else if (arith == 3) updatePointD(a, w, h, x, eps, nnz1, nnz2);
    else updatePointF(a, arith, w, h, x, eps, nnz1, nnz2);
    return 0;
}

void orc_permute(uint64_t key64, uint64_t first_draw, int64_t n, uint32_t* perm, int carry) {
    Rng r = Rng::task(key64);
    for (uint64_t k = 0; k < first_draw; ++k) r.next64();
    std::vector<uint32_t> p;
    if (carry) p.assign(perm, perm + n);
    permute(r, p, (uint64_t)n, (uint64_t)n);
    std::copy(p.begin(), p.end(), perm);
}

int orc_format_trace(int n, const double* loss, const double* test_loss, int has_test,
                     const double* eps, const double* t_eps, const double* t_epoch,
                     const double* t_loss, const double* t_test, const char* var,
                     const char* info_kv, char* out, int64_t cap) {
    std::vector<TraceRow> rows;
    for (int k = 0; k < n; ++k)
        rows.push_back({k, loss[k], has_test ? test_loss[k] : NAN, k == 0 ? NAN : eps[k],
                        t_eps[k], t_epoch[k], t_loss[k], has_test ? t_test[k] : 0.0});
    // info_kv: "d:name=value;s:name=value;..."
    std::map<std::string, double> dbl;
    std::map<std::string, std::string> str;
    std::string kv = info_kv ? info_kv : "";
    size_t st = 0;
    while (st < kv.size()) {
        size_t e = kv.find(';', st);
        std::string item = kv.substr(st, e == std::string::npos ? std::string::npos : e - st);
        size_t eq = item.find('=');
        if (item.size() > 2 && eq != std::string::npos) {
            std::string name = item.substr(2, eq - 2), val = item.substr(eq + 1);
            if (item[0] == 'd') dbl[name] = atof(val.c_str());
            else str[name] = val;
        }
        if (e == std::string::npos) break;
        st = e + 1;
    }
    std::string s = traceString(rows, var ? var : "trace", dbl, str);
    if ((int64_t)s.size() + 1 > cap) return (int)s.size() + 1;
    memcpy(out, s.c_str(), s.size() + 1);
    return 0;
}

}  // extern "C"
