// executor.cpp -- the batched / served executor behind include/ptk_batch.h (libptk_batch.so).
//
// Host side of the batching and serving layer: an HBM arena with first-fit sub-allocation
// (admission by datablock footprint), device descriptor arrays per instance kind, work lists
// rebuilt on the host whenever the set of resident instances changes, and one iteration ("tick")
// executed either
//   * batched: one captured hipGraph of the persistent multi-instance kernels of batch.hsaco,
//     whose work-list lengths live in device memory so the same graph serves every mix, or
//   * plain: the baseline, one kernel launch per task per instance (the six Yee tasks of a step,
//     the seven groupby tasks, the two EM tasks) in submission order on one stream.
// Every HIP call goes through the C ABI of libptk_hip.so.

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "ptk_batch.h"
#include "ptk_hip.h"

namespace {

thread_local std::string g_err;

int Fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(call)                                                                              \
    do {                                                                                          \
        int _s = (call);                                                                          \
        if (_s != PTK_HIP_OK)                                                                     \
            return Fail(PTK_BATCH_EHIP, std::string(#call) + ": " + ptk_hip_last_error_string()); \
    } while (0)

constexpr unsigned long long kAlign = 256;
constexpr int kFdtdCap = 16384, kGbCap = 8192, kGmmCap = 8192;
// device work-list lengths: FDTD tiles, FDTD instances, groupby chunk units, groupby buckets,
// GMM accumulate units, GMM finalize units, three-step items, three-step blocks, groupby tiles,
// groupby half buckets, groupby quarter buckets
constexpr int kCounts = 11;

unsigned long long Up(unsigned long long x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Host images of the device int4 / int2 work-list entries.
struct alignas(16) int4 {
    int x, y, z, w;
};
struct alignas(8) int2 {
    int x, y;
};

// Kernel arguments packed as the kernels' formal parameters.
struct Args {
    std::vector<unsigned char> b;
    template <class T>
    Args &add(const T &v) {
        const size_t a = alignof(T) < 8 ? alignof(T) : 8;
        while (b.size() % a) b.push_back(0);
        const unsigned char *p = (const unsigned char *)&v;
        b.insert(b.end(), p, p + sizeof(T));
        return *this;
    }
    Args &ptr(const void *p) { return add<const void *>(p); }
};

// First-fit sub-allocation of the arena, coalescing on free.
class Arena {
public:
    void Init(unsigned long long size) {
        size_ = size;
        free_.clear();
        free_[0] = size;
        used_ = peak_ = 0;
    }
    bool Alloc(unsigned long long bytes, unsigned long long *off) {
        bytes = Up(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const unsigned long long rest = it->second - bytes, at = it->first + bytes;
            free_.erase(it);
            if (rest) free_[at] = rest;
            used_ += bytes;
            peak_ = std::max(peak_, used_);
            return true;
        }
        return false;
    }
    void Free(unsigned long long off, unsigned long long bytes) {
        bytes = Up(bytes);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    unsigned long long size() const { return size_; }
    unsigned long long used() const { return used_; }
    unsigned long long peak() const { return peak_; }

private:
    unsigned long long size_ = 0, used_ = 0, peak_ = 0;
    std::map<unsigned long long, unsigned long long> free_;
};

struct Inst {
    int id = -1, kind = 0, state = PTK_STATE_QUEUED, keep = 0, slot = -1;
    // FDTD: a=nx, b=ny, c=nz; groupby: rows=a, numKeys=b; GMM: n=a, d=b, k=c
    long long a = 0;
    int b = 0, c = 0, steps = 1, stepsDone = 0;
    double zipf = 0;
    unsigned long long seed = 0, off = 0, bytes = 0;
    double tSubmit = -1, tAdmit = -1, tDone = -1;
    unsigned long long csum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    PTK_FDTD_INST fd{};
    PTK_GROUPBY_INST gb{};
    PTK_GMM_INST gm{};
    int *gmmFirst = nullptr;
};

// A device work list with its pinned host staging copy.
template <class T>
struct List {
    T *dev = nullptr, *host = nullptr;
    size_t cap = 0, n = 0;
};

}  // namespace

struct ptk_batch {
    int mode = PTK_BATCH_BATCHED;
    std::chrono::steady_clock::time_point t0;
    ptk_hip_stream stream = nullptr;
    ptk_hip_event evStart = nullptr, evStop = nullptr;
    ptk_hip_module mBatch = nullptr, mFdtd = nullptr, mGb = nullptr, mGmm = nullptr, mUtil = nullptr;
    ptk_hip_function fH = nullptr, fE = nullptr, fStep = nullptr, fStep3 = nullptr, fAdv = nullptr, fGbHist = nullptr,
                     fGbBScan = nullptr, fGbPart = nullptr, fGbKCount = nullptr, fGbKScan = nullptr, fGbRes = nullptr,
                     fGbTot = nullptr, fGmmAcc = nullptr, fGmmFin = nullptr, fInitFdtd = nullptr,
                     fInitGb = nullptr, fInitGmm = nullptr, fInitGmmModel = nullptr, fSumWords = nullptr,
                     fSumPairs = nullptr, fSumGroups = nullptr, fSumJobs = nullptr, fFill = nullptr;
    ptk_hip_function fSix[6] = {}, fPGb[7] = {},  // the groupby graph's seven kernels, in order
                     fPGmmAcc = nullptr, fPGmmFin = nullptr;
    void *arenaBase = nullptr;
    Arena arena;
    PTK_FDTD_INST *dFdtd = nullptr;
    PTK_GROUPBY_INST *dGb = nullptr;
    PTK_GMM_INST *dGmm = nullptr;
    std::vector<int> freeFdtd, freeGb, freeGmm;
    int *dCounts = nullptr, *hCounts = nullptr;
    unsigned long long *dSums = nullptr, *hSums = nullptr;  // checksum staging, 8 per retiring instance
    size_t sumCap = 0;
    List<int4> tiles, tiles3;  // FDTD tiles of the current form; three-step items (when cur3)
    List<int> t3start;         // three-step items of virtual block b: [t3start[b], t3start[b+1])
    List<int> fdtdActive;
    List<int2> gbUnits, gbBuckets, gbTiles, gbHalves, gbQuarters, gmmAcc, gmmFin;
    List<PTK_CSUM_JOB> csumJobs;
    double *dCdf = nullptr;
    int cdfKeys = -1;
    double cdfS = 0;
    long long cdfA = 0;
    // captured iteration per mix of kinds (bit 0 FDTD, 1 groupby, 2 GMM; bit 3 three FDTD steps)
    ptk_hip_graph graphs[16] = {};
    bool dirty = true;
    bool useGraphs = true;  // batched iterations as captured graphs (else direct, grid-fitted launches)
    int fdtdCols = 0, fdtdDepth = 8, gbRows = 262144, fdtdGrid = 4096;
    int cus = 256;                   // compute units of the device
    int fdtdSteps = 3, grid3 = 256;  // FDTD steps per launch (1 or 3); three-step virtual blocks (one per CU)
    int curCols = 512, curDepth = 16;  // the shape the current work lists were built with
    bool curStep = false;              // the current FDTD tiles are whole-step (overlapped) tiles
    bool cur3 = false;                 // tiles3 holds the three-step tiles of the resident set
    std::vector<Inst> insts;
    std::vector<int> queue, active;
    unsigned long long nTicks = 0, nLaunches = 0, nGraphs = 0;
    double execMs = 0;  // device time of the iterations' work (admission and retirement excluded)

    double Now() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    int Launch(ptk_hip_function f, unsigned gx, unsigned gy, unsigned gz, unsigned bx, Args &a) {
        ++nLaunches;
        HIPCHK(ptk_hip_launch(f, gx, gy, gz, bx, 1, 1, 0, stream, a.b.data(), a.b.size()));
        return PTK_BATCH_OK;
    }
};

namespace {

void DropGraphs(ptk_batch *B) {
    for (ptk_hip_graph &g : B->graphs)
        if (g) {
            ptk_hip_graph_destroy(g);
            g = nullptr;
        }
}

template <class T>
int Reserve(ptk_batch *B, List<T> &l, size_t need) {
    if (need <= l.cap) return PTK_BATCH_OK;
    size_t cap = std::max<size_t>(need, std::max<size_t>(1024, l.cap * 2));
    if (l.dev) {
        HIPCHK(ptk_hip_stream_synchronize(B->stream));
        ptk_hip_free(l.dev);
        ptk_hip_host_free(l.host);
        l.dev = nullptr;
        l.host = nullptr;
    }
    HIPCHK(ptk_hip_malloc((void **)&l.dev, cap * sizeof(T)));
    HIPCHK(ptk_hip_host_malloc((void **)&l.host, cap * sizeof(T)));
    l.cap = cap;
    DropGraphs(B);  // the captured iterations hold the old list address
    return PTK_BATCH_OK;
}

unsigned long long FdtdBytes(int nx, int ny, int nz) {
    unsigned long long s = 0;
    for (int f = 0; f < 6; ++f) s += Up(4ull * ptk_fdtd_field_elems(f, nx, ny, nz));
    return 2 * s;
}

unsigned long long GbBytes(long long rows, int numKeys) {
    const unsigned long long K = (unsigned long long)numKeys + 1;
    return 4 * Up(4ull * rows) + Up(8 * (3 * K + 1)) + Up(8 * K) + Up(8 * (3 + 3 * K)) +
           Up((unsigned long long)ptk_gb_scratch_bytes(rows, numKeys));
}

GMM_DIMS GmmDims(int n, int d, int k, int steps) {
    GMM_DIMS g{};
    g.N = n;
    g.D = d;
    g.K = k;
    const long long tiles = ((long long)n + 127) / 128;
    g.blocks = (int)std::max(1LL, (tiles + 15) / 16);  // 16 tiles (2048 points) per unit
    g.maxIter = steps;
    return g;
}

unsigned long long GmmBytes(int n, int d, int k, int steps) {
    GMM_DIMS g = GmmDims(n, d, k, steps);
    return Up(4ull * n * d) + Up(8ull * ptk_gmm_model_elems(&g)) + Up(8ull * g.blocks * ptk_gmm_partial_stride(&g)) +
           Up(4ull * k);
}

int Load(ptk_hip_module m, const char *name, ptk_hip_function *f) {
    HIPCHK(ptk_hip_module_get_function(f, m, name));
    return PTK_BATCH_OK;
}

int Submit(ptk_batch *B, Inst &in, int *id) {
    if (in.bytes > B->arena.size()) return Fail(PTK_BATCH_ENOMEM, "instance larger than the arena");
    in.id = (int)B->insts.size();
    in.tSubmit = B->Now();
    B->insts.push_back(in);
    B->queue.push_back(in.id);
    if (id) *id = in.id;
    return PTK_BATCH_OK;
}

int PopSlot(std::vector<int> &fr) {
    if (fr.empty()) return -1;
    int s = fr.back();
    fr.pop_back();
    return s;
}

// The Zipf inverse-CDF table of the groupby generator (ptk_groupby.cpp MakeInputs).
int EnsureCdf(ptk_batch *B, int numKeys, double s, long long *A) {
    const long long K = (long long)numKeys + 1;
    if (B->cdfKeys == numKeys && B->cdfS == s) {
        *A = B->cdfA;
        return PTK_BATCH_OK;
    }
    std::vector<double> cdf((size_t)K);
    double acc = 0;
    for (long long r = 0; r < K; ++r) {
        acc += std::pow((double)(r + 1), -s);
        cdf[(size_t)r] = acc;
    }
    for (double &c : cdf) c /= acc;
    long long a = 40503;
    auto gcd = [](long long x, long long y) { while (y) { long long t = x % y; x = y; y = t; } return x; };
    while (gcd(a, K) != 1) ++a;
    HIPCHK(ptk_hip_stream_synchronize(B->stream));
    if (B->dCdf) ptk_hip_free(B->dCdf);
    HIPCHK(ptk_hip_malloc((void **)&B->dCdf, sizeof(double) * (size_t)K));
    HIPCHK(ptk_hip_memcpy_htod(B->dCdf, cdf.data(), sizeof(double) * (size_t)K));
    B->cdfKeys = numKeys;
    B->cdfS = s;
    B->cdfA = a;
    *A = a;
    return PTK_BATCH_OK;
}

// Places an instance's datablocks in the arena and generates its inputs on the device.
int Admit(ptk_batch *B, Inst &in) {
    char *base = (char *)B->arenaBase + in.off;
    unsigned long long o = 0;
    auto carve = [&](unsigned long long bytes) {
        char *p = base + o;
        o += Up(bytes);
        return p;
    };
    if (in.kind == PTK_INST_FDTD) {
        in.slot = PopSlot(B->freeFdtd);
        if (in.slot < 0) return Fail(PTK_BATCH_ENOMEM, "too many resident FDTD instances");
        PTK_FDTD_INST &d = in.fd;
        for (int s = 0; s < 2; ++s)
            for (int f = 0; f < 6; ++f) d.f[s][f] = (float *)carve(4ull * ptk_fdtd_field_elems(f, (int)in.a, in.b, in.c));
        d.p = ptk_fdtd_make_params((int)in.a, in.b, in.c, 0.5f);
        d.steps = in.steps;
        d.done = 0;
        HIPCHK(ptk_hip_memcpy_htod_async(B->dFdtd + in.slot, &d, sizeof(d), B->stream));
        Args a;
        a.ptr(B->dFdtd).add<int>(in.slot).add<unsigned long long>(in.seed);
        if (int r = B->Launch(B->fInitFdtd, 512, 1, 1, 256, a)) return r;
    } else if (in.kind == PTK_INST_GROUPBY) {
        in.slot = PopSlot(B->freeGb);
        if (in.slot < 0) return Fail(PTK_BATCH_ENOMEM, "too many resident groupby instances");
        PTK_GROUPBY_INST &d = in.gb;
        const unsigned long long K = (unsigned long long)in.b + 1;
        d.n = in.a;
        d.numKeys = in.b;
        d.chunkRows = B->gbRows;
        d.keys = (const int *)carve(4ull * in.a);
        d.vals = (const int *)carve(4ull * in.a);
        d.outKeys = (int *)carve(4ull * in.a);
        d.outVals = (int *)carve(4ull * in.a);
        d.table = (long long *)carve(8 * (3 * K + 1));
        d.offsets = (long long *)carve(8 * K);
        d.groups = (long long *)carve(8 * (3 + 3 * K));
        d.scratch = carve((unsigned long long)ptk_gb_scratch_bytes(in.a, in.b));
        HIPCHK(ptk_hip_memcpy_htod_async(B->dGb + in.slot, &d, sizeof(d), B->stream));
        long long A = 0;
        const double *cdf = nullptr;
        if (in.zipf > 0) {
            if (int r = EnsureCdf(B, in.b, in.zipf, &A)) return r;
            cdf = B->dCdf;
        }
        Args a;
        a.ptr(d.keys).ptr(d.vals).add<long long>(d.n).ptr(cdf).add<int>(d.numKeys).add<long long>(A).add<unsigned long long>(in.seed);
        if (int r = B->Launch(B->fInitGb, 1024, 1, 1, 256, a)) return r;
    } else {
        in.slot = PopSlot(B->freeGmm);
        if (in.slot < 0) return Fail(PTK_BATCH_ENOMEM, "too many resident GMM instances");
        PTK_GMM_INST &d = in.gm;
        d.d = GmmDims((int)in.a, in.b, in.c, in.steps);
        d.X = (const float *)carve(4ull * in.a * in.b);
        d.model = (double *)carve(8ull * ptk_gmm_model_elems(&d.d));
        d.partials = (double *)carve(8ull * d.d.blocks * ptk_gmm_partial_stride(&d.d));
        in.gmmFirst = (int *)carve(4ull * in.c);
        HIPCHK(ptk_hip_memcpy_htod_async(B->dGmm + in.slot, &d, sizeof(d), B->stream));
        Args f;
        f.ptr(in.gmmFirst).add<unsigned long long>((unsigned long long)in.c).add<unsigned>(0x7fffffffu);
        if (int r = B->Launch(B->fFill, 1, 1, 1, 256, f)) return r;
        Args a;
        a.ptr(d.X).add<int>(d.d.N).add<int>(d.d.D).add<int>(d.d.K).add<unsigned long long>(in.seed).ptr(in.gmmFirst);
        if (int r = B->Launch(B->fInitGmm, 1024, 1, 1, 256, a)) return r;
        Args m;
        m.ptr(d.model).add(d.d).ptr(d.X).ptr(in.gmmFirst);
        if (int r = B->Launch(B->fInitGmmModel, 1, 1, 1, 256, m)) return r;
    }
    in.state = PTK_STATE_ACTIVE;
    in.tAdmit = B->Now();
    in.stepsDone = 0;
    return PTK_BATCH_OK;
}

// Work lists of the resident instances for the batched kernels.
// Whole-step tiles own 62 x 6 columns (blocks of 64 x 8 threads overlapping by one cell);
// three-step tiles own 58 x 10 (blocks of 64 x 16 threads overlapping by three cells).
constexpr int kStepTileI = 62, kStepTileJ = 6, kStep3TileI = 58, kStep3TileJ = 10;

// TRUE when every resident FDTD instance can take the whole-step kernel (32-bit buffer offsets).
bool StepFits(const ptk_batch *B) {
    for (int id : B->active) {
        const Inst &in = B->insts[(size_t)id];
        if (in.kind == PTK_INST_FDTD && (in.a + 1) * (long long)(in.b + 1) * (in.c + 1) >= (1ll << 29)) return false;
    }
    return true;
}

// FDTD tiles of the resident instances: half-step column tiles (cols, depth), or whole-step
// tiles of `depth` planes when cols == 0.
size_t FdtdTiles(const ptk_batch *B, int cols, int depth) {
    size_t nt = 0;
    for (int id : B->active) {
        const Inst &in = B->insts[(size_t)id];
        if (in.kind != PTK_INST_FDTD) continue;
        const long long slabs = (in.c + 1 + depth - 1) / depth;
        if (cols == 0) {
            nt += (size_t)(((in.a + 1 + kStepTileI - 1) / kStepTileI) * ((in.b + 1 + kStepTileJ - 1) / kStepTileJ) * slabs);
        } else {
            const long long ncol = (in.a + 1) * (long long)(in.b + 1);
            nt += (size_t)(((ncol + cols - 1) / cols) * slabs);
        }
    }
    return nt;
}

// The three-step persistent schedule over G virtual blocks (one per CU by default): the FDTD
// tiles of the resident instances in order, q = N / G whole-depth tiles per block taken
// round-robin (block b: tiles b, b + G, ...; at any time the blocks stream neighbouring tiles at
// about the same planes, so their overlap is shared in L2), then the planes of the remaining
// tiles cut into G equal runs in (tile, plane) order, one per block.  Items are grouped by
// block; at most N + G of them.
void BuildSchedule3(ptk_batch *B, const std::vector<int4> &t3, int G) {
    const size_t N = t3.size(), q = N / (size_t)G;
    long long W = 0;
    for (size_t i = q * (size_t)G; i < N; ++i) W += t3[i].w;
    size_t idx = q * (size_t)G;  // first remaining tile whose planes reach past `pos`
    long long pos = 0;          // planes of the remaining tiles before tile idx
    B->tiles3.n = 0;
    for (int b = 0; b < G; ++b) {
        B->t3start.host[b] = (int)B->tiles3.n;
        for (size_t r = 0; r < q; ++r) {
            const int4 &t = t3[r * (size_t)G + (size_t)b];
            B->tiles3.host[B->tiles3.n++] = int4{t.x, t.y, t.z, 0 | (t.w << 16)};
        }
        const long long lo = W * b / G, hi = W * (b + 1) / G;
        while (idx < N && pos + t3[idx].w <= lo) pos += t3[idx++].w;
        size_t j = idx;
        long long pj = pos, w = lo;
        while (w < hi) {
            const int4 &t = t3[j];
            const int k0 = (int)(w - pj), k1 = (int)std::min<long long>(t.w, hi - pj);
            B->tiles3.host[B->tiles3.n++] = int4{t.x, t.y, t.z, k0 | (k1 << 16)};
            w = pj + k1;
            if (k1 == t.w) pj += t3[j++].w;
        }
    }
    B->t3start.host[G] = (int)B->tiles3.n;
}

int BuildLists(ptk_batch *B) {
    // The tuned tile shape, made smaller (depth first, then columns) while the resident FDTD
    // work would give fewer than four tiles per CU: a few small instances must still spread over
    // the whole device.  Tiling never changes results.
    int cols = B->fdtdCols, depth = B->fdtdDepth;
    if (cols == 0 && !StepFits(B)) cols = 512, depth = 16;  // an instance too large for the whole-step form
    while (FdtdTiles(B, cols, depth) < 1024 && (depth > 2 || cols > 64)) {
        if (depth > 2) depth = (depth + 1) / 2;
        else if (cols == 0) break;
        else cols /= 2;
    }
    if (B->curStep != (cols == 0)) DropGraphs(B);  // the captured iterations hold the FDTD kernels
    B->curCols = cols;
    B->curDepth = depth;
    B->curStep = cols == 0;
    // The three-step schedule beside them when the form is the whole step and asks for three
    // steps per launch: tiles {slot, x, y, depth} of every resident FDTD instance.
    const bool three = cols == 0 && B->fdtdSteps == 3;
    std::vector<int4> t3;
    if (three)
        for (int id : B->active) {
            const Inst &in = B->insts[(size_t)id];
            if (in.kind != PTK_INST_FDTD) continue;
            const int ti3 = (int)((in.a + kStep3TileI) / kStep3TileI), tj3 = (in.b + kStep3TileJ) / kStep3TileJ;
            for (int y = 0; y < tj3; ++y)
                for (int x = 0; x < ti3; ++x) t3.push_back(int4{in.slot, x, y, in.c + 1});
        }
    const size_t nt3 = three ? t3.size() + (size_t)B->grid3 : 0;
    B->cur3 = three;
    size_t nt = FdtdTiles(B, cols, depth), nf = 0, ngu = 0, ngi = 0, ngm = 0, nga = 0, ngf = 0;
    for (int id : B->active) {
        const Inst &in = B->insts[(size_t)id];
        if (in.kind == PTK_INST_FDTD) {
            ++nf;
        } else if (in.kind == PTK_INST_GROUPBY) {
            ngu += (size_t)((in.a + in.gb.chunkRows - 1) / in.gb.chunkRows);
            ngi += (size_t)ptk_gb_buckets(in.b);
            ngm += (size_t)ptk_gb_tiles_max(in.a, in.b);
        } else {
            nga += (size_t)in.gm.d.blocks;
            ngf += (size_t)in.gm.d.K + 1;
        }
    }
    if (int r = Reserve(B, B->tiles, nt)) return r;
    if (int r = Reserve(B, B->tiles3, nt3)) return r;
    if (int r = Reserve(B, B->t3start, three ? (size_t)B->grid3 + 1 : 0)) return r;
    if (int r = Reserve(B, B->fdtdActive, nf)) return r;
    if (int r = Reserve(B, B->gbUnits, ngu)) return r;
    if (int r = Reserve(B, B->gbBuckets, ngi)) return r;
    if (int r = Reserve(B, B->gbTiles, ngm)) return r;
    if (int r = Reserve(B, B->gbHalves, 2 * ngi)) return r;
    if (int r = Reserve(B, B->gbQuarters, 4 * ngi)) return r;
    if (int r = Reserve(B, B->gmmAcc, nga)) return r;
    if (int r = Reserve(B, B->gmmFin, ngf)) return r;
    B->tiles.n = B->tiles3.n = B->t3start.n = B->fdtdActive.n = B->gbUnits.n = B->gbBuckets.n = B->gbTiles.n = 0;
    B->gbHalves.n = B->gbQuarters.n = B->gmmAcc.n = B->gmmFin.n = 0;
    for (int id : B->active) {
        const Inst &in = B->insts[(size_t)id];
        if (in.kind == PTK_INST_FDTD) {
            const int ncol = (int)((in.a + 1) * (long long)(in.b + 1)), kdepth = in.c + 1;
            const int ti = (int)((in.a + kStepTileI) / kStepTileI), tj = (in.b + kStepTileJ) / kStepTileJ;
            // ceil(kdepth / depth) slabs of equal depth (within one plane): a short last slab
            // would leave the device's last round of tiles uneven
            const int slabs = (kdepth + depth - 1) / depth;
            for (int sl = 0; sl < slabs; ++sl) {
                const int k0 = (int)((long long)sl * kdepth / slabs), k1 = (int)((long long)(sl + 1) * kdepth / slabs);
                if (cols == 0) {
                    for (int y = 0; y < tj; ++y)
                        for (int x = 0; x < ti; ++x) B->tiles.host[B->tiles.n++] = int4{in.slot, x, y, k0 | (k1 << 16)};
                } else {
                    for (int c0 = 0; c0 < ncol; c0 += cols)
                        B->tiles.host[B->tiles.n++] = int4{in.slot, c0, std::min(cols, ncol - c0), k0 | (k1 << 16)};
                }
            }
            B->fdtdActive.host[B->fdtdActive.n++] = in.slot;
        } else if (in.kind == PTK_INST_GROUPBY) {
            const long long chunks = (in.a + in.gb.chunkRows - 1) / in.gb.chunkRows;
            for (long long c = 0; c < chunks; ++c) B->gbUnits.host[B->gbUnits.n++] = int2{in.slot, (int)c};
            const long long nb = ptk_gb_buckets(in.b);
            for (long long q = 0; q < nb; ++q) B->gbBuckets.host[B->gbBuckets.n++] = int2{in.slot, (int)q};
            for (long long q = 0; q < 2 * nb; ++q) B->gbHalves.host[B->gbHalves.n++] = int2{in.slot, (int)q};
            for (long long q = 0; q < 4 * nb; ++q) B->gbQuarters.host[B->gbQuarters.n++] = int2{in.slot, (int)q};
            // every tile the instance may have; those past its tiles (known on the device) do nothing
            for (long long q = 0; q < ptk_gb_tiles_max(in.a, in.b); ++q) B->gbTiles.host[B->gbTiles.n++] = int2{in.slot, (int)q};
        } else {
            for (int b = 0; b < in.gm.d.blocks; ++b) B->gmmAcc.host[B->gmmAcc.n++] = int2{in.slot, b};
            for (int k = 0; k <= in.gm.d.K; ++k) B->gmmFin.host[B->gmmFin.n++] = int2{in.slot, k};
        }
    }
    B->hCounts[0] = (int)B->tiles.n;
    B->hCounts[1] = (int)B->fdtdActive.n;
    B->hCounts[2] = (int)B->gbUnits.n;
    B->hCounts[3] = (int)B->gbBuckets.n;
    B->hCounts[4] = (int)B->gmmAcc.n;
    B->hCounts[5] = (int)B->gmmFin.n;
    if (three) {
        BuildSchedule3(B, t3, B->grid3);
        B->t3start.n = (size_t)B->grid3 + 1;
    }
    B->hCounts[6] = (int)B->tiles3.n;
    B->hCounts[7] = three ? B->grid3 : 0;
    B->hCounts[8] = (int)B->gbTiles.n;
    B->hCounts[9] = (int)B->gbHalves.n;
    B->hCounts[10] = (int)B->gbQuarters.n;
    auto up = [&](auto &l) -> int {
        if (l.n) HIPCHK(ptk_hip_memcpy_htod_async(l.dev, l.host, l.n * sizeof(*l.host), B->stream));
        return PTK_BATCH_OK;
    };
    if (int r = up(B->tiles)) return r;
    if (int r = up(B->tiles3)) return r;
    if (int r = up(B->t3start)) return r;
    if (int r = up(B->fdtdActive)) return r;
    if (int r = up(B->gbUnits)) return r;
    if (int r = up(B->gbBuckets)) return r;
    if (int r = up(B->gbTiles)) return r;
    if (int r = up(B->gbHalves)) return r;
    if (int r = up(B->gbQuarters)) return r;
    if (int r = up(B->gmmAcc)) return r;
    if (int r = up(B->gmmFin)) return r;
    HIPCHK(ptk_hip_memcpy_htod_async(B->dCounts, B->hCounts, kCounts * sizeof(int), B->stream));
    return PTK_BATCH_OK;
}

// The batched iteration's launches (captured once into a graph).
// The batched iteration's launches.  Captured into a graph, the grids are the configured
// persistent sizes (the work-list lengths are read on the device); launched directly (fit), each
// grid is cut to the current list length, which the host knows.
int IssueBatched(ptk_batch *B, bool fdtd, bool gb, bool gmm, bool advance, bool fit = false, int steps = 1) {
    auto grid = [&](int configured, int count) {
        return (unsigned)(fit ? std::max(1, std::min(configured, count)) : configured);
    };
    const int *hc = B->hCounts;
    if (fdtd && steps == 3) {  // three whole steps per launch (tiles3)
        Args t;
        t.ptr(B->tiles3.dev).ptr(B->dCounts + 6).ptr(B->t3start.dev).ptr(B->dFdtd);
        ++B->nLaunches;  // one block per virtual block of the schedule
        HIPCHK(ptk_hip_launch(B->fStep3, (unsigned)B->grid3, 1, 1, 64, 16, 1, 0, B->stream, t.b.data(), t.b.size()));
    } else if (fdtd && B->curStep) {
        Args t;
        t.ptr(B->tiles.dev).ptr(B->dCounts + 0).ptr(B->dFdtd);
        ++B->nLaunches;
        HIPCHK(ptk_hip_launch(B->fStep, grid(B->fdtdGrid, hc[0]), 1, 1, 64, 8, 1, 0, B->stream, t.b.data(), t.b.size()));
    } else if (fdtd) {
        Args h, e;
        h.ptr(B->tiles.dev).ptr(B->dCounts + 0).ptr(B->dFdtd);
        e.ptr(B->tiles.dev).ptr(B->dCounts + 0).ptr(B->dFdtd);
        if (int r = B->Launch(B->fH, grid(B->fdtdGrid, hc[0]), 1, 1, 256, h)) return r;
        if (int r = B->Launch(B->fE, grid(B->fdtdGrid, hc[0]), 1, 1, 256, e)) return r;
    }
    if (advance) {
        Args a;
        a.ptr(B->fdtdActive.dev).ptr(B->dCounts + 1).ptr(B->dFdtd);
        if (int r = B->Launch(B->fAdv, grid(16, (hc[1] + 255) / 256), 1, 1, 256, a)) return r;
    }
    if (gmm) {
        Args a, f;
        a.ptr(B->gmmAcc.dev).ptr(B->dCounts + 4).ptr(B->dGmm);
        f.ptr(B->gmmFin.dev).ptr(B->dCounts + 5).ptr(B->dGmm);
        // the matrix-core accumulate holds two blocks per CU: a 512-block grid is resident at once
        if (int r = B->Launch(B->fGmmAcc, grid(512, hc[4]), 1, 1, 256, a)) return r;
        if (int r = B->Launch(B->fGmmFin, grid(512, hc[5]), 1, 1, 256, f)) return r;
    }
    if (gb) {  // the groupby graph's seven passes, each over its work list
        Args u, b, p, k, h, s, q;
        u.ptr(B->gbUnits.dev).ptr(B->dCounts + 2).ptr(B->dGb);
        b.ptr(B->gbBuckets.dev).ptr(B->dCounts + 3).ptr(B->dGb);
        p.ptr(B->gbUnits.dev).ptr(B->dCounts + 2).ptr(B->dGb);
        k.ptr(B->gbTiles.dev).ptr(B->dCounts + 8).ptr(B->dGb);
        h.ptr(B->gbHalves.dev).ptr(B->dCounts + 9).ptr(B->dGb);
        s.ptr(B->gbTiles.dev).ptr(B->dCounts + 8).ptr(B->dGb);
        q.ptr(B->gbQuarters.dev).ptr(B->dCounts + 10).ptr(B->dGb);
        if (int r = B->Launch(B->fGbHist, grid(2048, hc[2]), 1, 1, PTK_GB_HIST_THREADS, u)) return r;
        if (int r = B->Launch(B->fGbBScan, grid(2048, hc[3]), 1, 1, PTK_GB_SCAN_THREADS, b)) return r;
        if (int r = B->Launch(B->fGbPart, grid(2 * B->cus, hc[2]), 1, 1, PTK_GB_PART_THREADS, p)) return r;
        if (int r = B->Launch(B->fGbKCount, grid(8192, hc[8]), 1, 1, PTK_GB_COUNT_THREADS, k)) return r;
        if (int r = B->Launch(B->fGbKScan, grid(4096, hc[9]), 1, 1, PTK_GB_KEYSCAN_THREADS, h)) return r;
        if (int r = B->Launch(B->fGbRes, grid(8192, hc[8]), 1, 1, PTK_GB_TILE_THREADS, s)) return r;
        if (int r = B->Launch(B->fGbTot, grid(8192, hc[10]), 1, 1, PTK_GB_TOTALS_THREADS, q)) return r;
    }
    return PTK_BATCH_OK;
}

// Plain per-task dispatch of one iteration of one instance.
int IssuePlain(ptk_batch *B, const Inst &in) {
    if (in.kind == PTK_INST_FDTD) {
        const PTK_FDTD_INST &d = in.fd;
        const int cur = in.stepsDone & 1, nxt = cur ^ 1;
        // (own, in1, in2) of Hx, Hy, Hz, Ex, Ey, Ez (graphfdtd.cpp:448-560); E reads the new H
        static const int own[6] = {0, 1, 2, 3, 4, 5}, i1[6] = {4, 3, 3, 1, 0, 0}, i2[6] = {5, 5, 4, 2, 2, 1};
        for (int t = 0; t < 6; ++t) {
            const int src = t < 3 ? cur : nxt;
            int d1, d2, d3;
            ptk_fdtd_extent(t, d.p.Nx, d.p.Ny, d.p.Nz, &d1, &d2, &d3);
            Args a;
            a.ptr(d.f[cur][own[t]]).ptr(d.f[src][i1[t]]).ptr(d.f[src][i2[t]]).add(d.p).ptr(d.f[nxt][own[t]]);
            const unsigned gx = (unsigned)(((long long)d1 * d2 + 255) / 256), slabs = (unsigned)((d3 + 7) / 8);
            if (int r = B->Launch(B->fSix[t], gx, slabs, 1, 256, a)) return r;
        }
    } else if (in.kind == PTK_INST_GROUPBY) {
        const PTK_GROUPBY_INST &d = in.gb;
        const unsigned chunks = (unsigned)std::max(1LL, ptk_gb_chunks(d.n)), nb = (unsigned)ptk_gb_buckets(d.numKeys);
        const unsigned tiles = (unsigned)ptk_gb_tiles_max(d.n, d.numKeys);
        const int n = (int)d.n, nk = d.numKeys;
        Args h, b, p, k, s, r6, t;
        h.ptr(d.keys).add<int>(n).add<int>(nk).ptr(d.scratch);
        b.ptr(d.scratch).add<int>(n).add<int>(nk);
        p.ptr(d.keys).ptr(d.vals).add<int>(n).add<int>(nk).ptr(d.scratch);
        k.ptr(d.scratch).add<int>(n).add<int>(nk);
        s.ptr(d.scratch).add<int>(n).add<int>(nk).ptr(d.offsets);
        r6.ptr(d.scratch).add<int>(n).add<int>(nk).ptr(d.offsets).ptr(d.outKeys).ptr(d.outVals);
        t.ptr(d.scratch).add<int>(n).add<int>(nk).ptr(d.table).ptr(d.groups);
        if (int r = B->Launch(B->fPGb[0], chunks, 1, 1, PTK_GB_HIST_THREADS, h)) return r;
        if (int r = B->Launch(B->fPGb[1], nb, 1, 1, PTK_GB_SCAN_THREADS, b)) return r;
        if (int r = B->Launch(B->fPGb[2], std::min(chunks, (unsigned)(2 * B->cus)), 1, 1, PTK_GB_PART_THREADS, p)) return r;
        if (int r = B->Launch(B->fPGb[3], tiles, 1, 1, PTK_GB_COUNT_THREADS, k)) return r;
        if (int r = B->Launch(B->fPGb[4], 2 * nb, 1, 1, PTK_GB_KEYSCAN_THREADS, s)) return r;
        if (int r = B->Launch(B->fPGb[5], tiles, 1, 1, PTK_GB_TILE_THREADS, r6)) return r;
        if (int r = B->Launch(B->fPGb[6], 4 * nb, 1, 1, PTK_GB_TOTALS_THREADS, t)) return r;
    } else {
        const PTK_GMM_INST &d = in.gm;
        Args a, f;
        a.ptr(d.X).add(d.d).ptr(d.model).ptr(d.partials);
        f.ptr(d.model).ptr(d.partials).add(d.d);
        if (int r = B->Launch(B->fPGmmAcc, (unsigned)d.d.blocks, 1, 1, 256, a)) return r;
        if (int r = B->Launch(B->fPGmmFin, (unsigned)d.d.K + 1, 1, 1, 256, f)) return r;
    }
    return PTK_BATCH_OK;
}

// Device checksums of the completed instances, then their memory goes back to the arena.
int Retire(ptk_batch *B, const std::vector<int> &done) {
    if (done.empty()) return PTK_BATCH_OK;
    if (done.size() * 8 > B->sumCap) {
        HIPCHK(ptk_hip_stream_synchronize(B->stream));
        if (B->dSums) ptk_hip_free(B->dSums);
        if (B->hSums) ptk_hip_host_free(B->hSums);
        B->sumCap = std::max<size_t>(done.size() * 8, 4096);
        HIPCHK(ptk_hip_malloc((void **)&B->dSums, B->sumCap * 8));
        HIPCHK(ptk_hip_host_malloc((void **)&B->hSums, B->sumCap * 8));
    }
    std::vector<unsigned long long> zero(done.size() * 8, 0);
    HIPCHK(ptk_hip_memcpy_htod_async(B->dSums, zero.data(), zero.size() * 8, B->stream));
    // one job per result array, all summed by one launch per 65535 jobs
    std::vector<PTK_CSUM_JOB> jobs;
    for (size_t i = 0; i < done.size(); ++i) {
        const Inst &in = B->insts[(size_t)done[i]];
        unsigned long long *s = B->dSums + 8 * i;
        if (in.kind == PTK_INST_FDTD) {
            const int cur = in.stepsDone & 1;
            for (int f = 0; f < 6; ++f)
                jobs.push_back({in.fd.f[cur][f], nullptr, (long long)ptk_fdtd_field_elems(f, (int)in.a, in.b, in.c), s + f, 0, 0});
        } else if (in.kind == PTK_INST_GROUPBY) {
            jobs.push_back({in.gb.groups, nullptr, 0, s + 0, 2, 0});
            jobs.push_back({in.gb.outKeys, nullptr, in.a, s + 1, 0, 0});
            jobs.push_back({in.gb.outVals, nullptr, in.a, s + 2, 0, 0});
            jobs.push_back({in.gb.keys, in.gb.vals, in.a, s + 3, 1, 0});
            jobs.push_back({in.gb.outKeys, in.gb.outVals, in.a, s + 4, 1, 0});
        } else {
            jobs.push_back({in.gm.model, nullptr, 2 * ptk_gmm_model_elems(&in.gm.d), s + 0, 0, 0});
        }
    }
    if (int r = Reserve(B, B->csumJobs, jobs.size())) return r;
    memcpy(B->csumJobs.host, jobs.data(), jobs.size() * sizeof(PTK_CSUM_JOB));
    HIPCHK(ptk_hip_memcpy_htod_async(B->csumJobs.dev, B->csumJobs.host, jobs.size() * sizeof(PTK_CSUM_JOB), B->stream));
    for (size_t j0 = 0; j0 < jobs.size(); j0 += 65535) {
        Args a;
        a.ptr(B->csumJobs.dev + j0);
        const unsigned nj = (unsigned)std::min<size_t>(65535, jobs.size() - j0);
        if (int r = B->Launch(B->fSumJobs, 64, nj, 1, 256, a)) return r;
    }
    HIPCHK(ptk_hip_memcpy_dtoh_async(B->hSums, B->dSums, done.size() * 64, B->stream));
    HIPCHK(ptk_hip_stream_synchronize(B->stream));
    for (size_t i = 0; i < done.size(); ++i) {
        Inst &in = B->insts[(size_t)done[i]];
        memcpy(in.csum, B->hSums + 8 * i, 64);
        in.state = PTK_STATE_DONE;
        in.tDone = B->Now();
        if (in.kind == PTK_INST_FDTD) B->freeFdtd.push_back(in.slot);
        else if (in.kind == PTK_INST_GROUPBY) B->freeGb.push_back(in.slot);
        else B->freeGmm.push_back(in.slot);
        if (!in.keep) {
            B->arena.Free(in.off, in.bytes);
            in.state = PTK_STATE_RELEASED;
        }
    }
    return PTK_BATCH_OK;
}

}  // namespace

extern "C" {

const char *ptk_batch_last_error(void) { return g_err.c_str(); }

unsigned long long ptk_batch_footprint(int kind, long long a, int b_, int c, int steps) {
    if (kind == PTK_INST_FDTD) return FdtdBytes((int)a, b_, c);
    if (kind == PTK_INST_GROUPBY) return GbBytes(a, b_);
    return GmmBytes((int)a, b_, c, steps);
}

int ptk_batch_create(const char *kernel_dir, unsigned long long arena_bytes, int mode, ptk_batch **out) {
    if (!kernel_dir || !out || arena_bytes == 0 || (mode != PTK_BATCH_PLAIN && mode != PTK_BATCH_BATCHED))
        return Fail(PTK_BATCH_EINVAL, "bad argument");
    *out = nullptr;
    ptk_batch *B = new ptk_batch();
    B->mode = mode;
    B->t0 = std::chrono::steady_clock::now();
    auto fail = [&](int r) {
        ptk_batch_destroy(B);
        return r;
    };
    const std::string dir(kernel_dir);
    int r = PTK_BATCH_OK;
    if (ptk_hip_stream_create(&B->stream) || ptk_hip_event_create(&B->evStart, 1) || ptk_hip_event_create(&B->evStop, 1))
        return fail(Fail(PTK_BATCH_EHIP, std::string("stream/event: ") + ptk_hip_last_error_string()));
    {
        int dev = 0;
        ptk_hip_device_props_t props;
        if (ptk_hip_get_device(&dev) == PTK_HIP_OK && ptk_hip_get_device_props(dev, &props) == PTK_HIP_OK &&
            props.compute_units > 0)
            B->cus = props.compute_units;
        B->grid3 = B->cus;  // the three-step schedule: one virtual block per CU
    }
    const std::pair<ptk_hip_module *, const char *> mods[] = {
        {&B->mBatch, "batch.hsaco"}, {&B->mFdtd, "fdtd.hsaco"}, {&B->mGb, "groupby.hsaco"},
        {&B->mGmm, "gmm.hsaco"}, {&B->mUtil, "ptk_util.hsaco"}};
    for (auto &m : mods)
        if (ptk_hip_module_load(m.first, (dir + "/" + m.second).c_str()))
            return fail(Fail(PTK_BATCH_EHIP, std::string("load ") + m.second + ": " + ptk_hip_last_error_string()));
    const std::pair<ptk_hip_function *, const char *> bf[] = {
        {&B->fH, "BatchFdtdH"}, {&B->fE, "BatchFdtdE"}, {&B->fStep, "BatchFdtdStep"},
        {&B->fStep3, "BatchFdtdSteps3"}, {&B->fAdv, "BatchFdtdAdvance"},
        {&B->fGbHist, "BatchGroupbyHistogram"}, {&B->fGbBScan, "BatchGroupbyBucketScan"},
        {&B->fGbPart, "BatchGroupbyPartition"}, {&B->fGbKCount, "BatchGroupbyKeyCount"},
        {&B->fGbKScan, "BatchGroupbyKeyScan"}, {&B->fGbRes, "BatchGroupbyReshuffle"}, {&B->fGbTot, "BatchGroupbyTotals"},
        {&B->fGmmAcc, "BatchGmmAccumulateMfma"}, {&B->fGmmFin, "BatchGmmFinalize"},
        {&B->fInitFdtd, "BatchInitFdtd"}, {&B->fInitGb, "BatchInitGroupby"}, {&B->fInitGmm, "BatchInitGmm"},
        {&B->fInitGmmModel, "BatchInitGmmModel"}, {&B->fSumWords, "BatchChecksumWords"},
        {&B->fSumPairs, "BatchChecksumPairs"}, {&B->fSumGroups, "BatchChecksumGroups"},
        {&B->fSumJobs, "BatchChecksumJobs"}};
    for (auto &f : bf)
        if ((r = Load(B->mBatch, f.second, f.first))) return fail(r);
    static const char *six[6] = {"HxComputation", "HyComputation", "HzComputation",
                                 "ExComputation", "EyComputation", "EzComputation"};
    for (int t = 0; t < 6; ++t)
        if ((r = Load(B->mFdtd, six[t], &B->fSix[t]))) return fail(r);
    static const char *gbk[7] = {"GroupbyHistogram", "GroupbyBucketScan", "GroupbyPartition", "GroupbyKeyCount",
                                 "GroupbyKeyScan", "GroupbyReshuffle", "GroupbyTotals"};
    for (int i = 0; i < 7; ++i)
        if ((r = Load(B->mGb, gbk[i], &B->fPGb[i]))) return fail(r);
    if ((r = Load(B->mGmm, "GmmAccumulateMfma", &B->fPGmmAcc)) || (r = Load(B->mGmm, "GmmFinalize", &B->fPGmmFin)) ||
        (r = Load(B->mUtil, "PtkFillWords", &B->fFill)))
        return fail(r);
    if (ptk_hip_malloc(&B->arenaBase, arena_bytes))
        return fail(Fail(PTK_BATCH_ENOMEM, std::string("arena: ") + ptk_hip_last_error_string()));
    B->arena.Init(arena_bytes);
    if (ptk_hip_malloc((void **)&B->dFdtd, sizeof(PTK_FDTD_INST) * kFdtdCap) ||
        ptk_hip_malloc((void **)&B->dGb, sizeof(PTK_GROUPBY_INST) * kGbCap) ||
        ptk_hip_malloc((void **)&B->dGmm, sizeof(PTK_GMM_INST) * kGmmCap) ||
        ptk_hip_malloc((void **)&B->dCounts, kCounts * sizeof(int)) || ptk_hip_host_malloc((void **)&B->hCounts, kCounts * sizeof(int)))
        return fail(Fail(PTK_BATCH_EHIP, std::string("tables: ") + ptk_hip_last_error_string()));
    for (int i = kFdtdCap - 1; i >= 0; --i) B->freeFdtd.push_back(i);
    for (int i = kGbCap - 1; i >= 0; --i) B->freeGb.push_back(i);
    for (int i = kGmmCap - 1; i >= 0; --i) B->freeGmm.push_back(i);
    memset(B->hCounts, 0, kCounts * sizeof(int));
    *out = B;
    return PTK_BATCH_OK;
}

int ptk_batch_destroy(ptk_batch *B) {
    if (!B) return PTK_BATCH_OK;
    if (B->stream) ptk_hip_stream_synchronize(B->stream);
    DropGraphs(B);
    auto freeList = [](auto &l) {
        if (l.dev) ptk_hip_free(l.dev);
        if (l.host) ptk_hip_host_free(l.host);
    };
    freeList(B->tiles);
    freeList(B->tiles3);
    freeList(B->t3start);
    freeList(B->fdtdActive);
    freeList(B->gbBuckets);
    freeList(B->gbTiles);
    freeList(B->gbHalves);
    freeList(B->gbQuarters);
    freeList(B->gbUnits);
    freeList(B->gmmAcc);
    freeList(B->gmmFin);
    freeList(B->csumJobs);
    if (B->dSums) ptk_hip_free(B->dSums);
    if (B->hSums) ptk_hip_host_free(B->hSums);
    if (B->dCdf) ptk_hip_free(B->dCdf);
    if (B->dCounts) ptk_hip_free(B->dCounts);
    if (B->hCounts) ptk_hip_host_free(B->hCounts);
    if (B->dFdtd) ptk_hip_free(B->dFdtd);
    if (B->dGb) ptk_hip_free(B->dGb);
    if (B->dGmm) ptk_hip_free(B->dGmm);
    if (B->arenaBase) ptk_hip_free(B->arenaBase);
    for (ptk_hip_module m : {B->mBatch, B->mFdtd, B->mGb, B->mGmm, B->mUtil})
        if (m) ptk_hip_module_unload(m);
    if (B->evStart) ptk_hip_event_destroy(B->evStart);
    if (B->evStop) ptk_hip_event_destroy(B->evStop);
    if (B->stream) ptk_hip_stream_destroy(B->stream);
    delete B;
    return PTK_BATCH_OK;
}

int ptk_batch_submit_fdtd(ptk_batch *B, int nx, int ny, int nz, int steps, unsigned long long seed, int keep, int *id) {
    if (!B || nx < 1 || ny < 1 || nz < 1 || nz >= 65535 || steps < 1) return Fail(PTK_BATCH_EINVAL, "bad FDTD shape");
    for (int f = 0; f < 6; ++f)
        if (ptk_fdtd_field_elems(f, nx, ny, nz) >= (1ull << 31)) return Fail(PTK_BATCH_EINVAL, "FDTD field too large");
    Inst in;
    in.kind = PTK_INST_FDTD;
    in.a = nx;
    in.b = ny;
    in.c = nz;
    in.steps = steps;
    in.seed = seed;
    in.keep = keep;
    in.bytes = FdtdBytes(nx, ny, nz);
    return Submit(B, in, id);
}

int ptk_batch_submit_groupby(ptk_batch *B, long long rows, int num_keys, double zipf_s, unsigned long long seed,
                             int keep, int *id) {
    if (!B || rows < 1 || rows > 0x7fffffffLL || num_keys < 0 || (long long)num_keys + 1 > PTK_GB_MAX_KEYS)
        return Fail(PTK_BATCH_EINVAL, "bad groupby shape");
    Inst in;
    in.kind = PTK_INST_GROUPBY;
    in.a = rows;
    in.b = num_keys;
    in.zipf = zipf_s;
    in.seed = seed;
    in.keep = keep;
    in.bytes = GbBytes(rows, num_keys);
    return Submit(B, in, id);
}

int ptk_batch_submit_gmm(ptk_batch *B, int n, int d, int k, int steps, unsigned long long seed, int keep, int *id) {
    if (!B || n < 1 || d < 1 || d > PTK_GMM_MAX_D || k < 1 || k > PTK_GMM_MAX_K || k > n || steps < 1 ||
        (long long)n * d > 0x3fffffffLL)
        return Fail(PTK_BATCH_EINVAL, "bad GMM shape");
    Inst in;
    in.kind = PTK_INST_GMM;
    in.a = n;
    in.b = d;
    in.c = k;
    in.steps = steps;
    in.seed = seed;
    in.keep = keep;
    in.bytes = GmmBytes(n, d, k, steps);
    return Submit(B, in, id);
}

int ptk_batch_run(ptk_batch *B, int max_admit, int max_iters, int *niters, int *nactive, int *ncompleted) {
    if (!B || max_iters < 1) return Fail(PTK_BATCH_EINVAL, "bad argument");
    // admission: in submission order while the footprint fits
    int admitted = 0;
    size_t qi = 0;
    while (qi < B->queue.size() && (max_admit < 0 || admitted < max_admit)) {
        Inst &in = B->insts[(size_t)B->queue[qi]];
        if (!B->arena.Alloc(in.bytes, &in.off)) break;
        if (int r = Admit(B, in)) {
            B->arena.Free(in.off, in.bytes);
            return r;
        }
        B->active.push_back(in.id);
        ++admitted;
        ++qi;
    }
    B->queue.erase(B->queue.begin(), B->queue.begin() + (long)qi);
    if (admitted) B->dirty = true;
    if (!B->active.empty() && B->mode == PTK_BATCH_BATCHED && B->dirty) {
        if (int r = BuildLists(B)) return r;
        B->dirty = false;
    }
    // FDTD steps per batched launch this round: three while every resident FDTD instance has at
    // least three left (three is odd, so the current field set stays stepsDone & 1), else one
    int T = B->mode == PTK_BATCH_BATCHED && B->cur3 ? 3 : 1;
    for (int id : B->active) {
        const Inst &in = B->insts[(size_t)id];
        if (in.kind == PTK_INST_FDTD && in.steps - in.stepsDone < 3) T = 1;
    }
    // iterations every resident instance still needs before the first completes
    int iters = 0, mask = 0, need = max_iters;
    for (int id : B->active) {
        const Inst &in = B->insts[(size_t)id];
        mask |= in.kind == PTK_INST_FDTD ? 1 : in.kind == PTK_INST_GROUPBY ? 2 : 4;
        need = std::min(need, in.kind == PTK_INST_GROUPBY ? 1
                              : in.kind == PTK_INST_FDTD  ? (in.steps - in.stepsDone) / T
                                                          : in.steps - in.stepsDone);
    }
    if (!B->active.empty()) {
        iters = std::max(1, need);
        // the events bracket only the iterations: admission work queued before them is excluded
        HIPCHK(ptk_hip_event_record(B->evStart, B->stream));
        if (B->mode == PTK_BATCH_BATCHED && !B->useGraphs) {
            for (int it = 0; it < iters; ++it)
                if (int r = IssueBatched(B, mask & 1, mask & 2, mask & 4, mask & 1, true, T)) return r;
        } else if (B->mode == PTK_BATCH_BATCHED) {
            const int key = mask | (T == 3 ? 8 : 0);
            if (!B->graphs[key]) {  // one captured iteration per mix of instance kinds and FDTD form
                HIPCHK(ptk_hip_stream_begin_capture(B->stream));
                int r = IssueBatched(B, mask & 1, mask & 2, mask & 4, mask & 1, false, T);
                ptk_hip_graph g = nullptr;
                int e = ptk_hip_stream_end_capture(B->stream, &g);
                if (r) return r;
                if (e) return Fail(PTK_BATCH_EHIP, std::string("capture: ") + ptk_hip_last_error_string());
                B->graphs[key] = g;
            }
            for (int it = 0; it < iters; ++it) {
                HIPCHK(ptk_hip_graph_launch(B->graphs[key], B->stream));
                ++B->nGraphs;
            }
        } else {
            for (int it = 0; it < iters; ++it) {
                for (int id : B->active)
                    if (int r = IssuePlain(B, B->insts[(size_t)id])) return r;
                for (int id : B->active) ++B->insts[(size_t)id].stepsDone;  // parity of the next round
            }
            for (int id : B->active) B->insts[(size_t)id].stepsDone -= iters;
        }
        HIPCHK(ptk_hip_event_record(B->evStop, B->stream));
        HIPCHK(ptk_hip_event_synchronize(B->evStop));
        float ms = 0;
        HIPCHK(ptk_hip_event_elapsed_ms(&ms, B->evStart, B->evStop));
        B->execMs += ms;
        B->nTicks += (unsigned long long)iters;
    }
    std::vector<int> done, keepActive;
    for (int id : B->active) {
        Inst &in = B->insts[(size_t)id];
        if (in.kind == PTK_INST_GROUPBY) {
            in.stepsDone = 1;
            done.push_back(id);
        } else if ((in.stepsDone += in.kind == PTK_INST_FDTD ? iters * T : iters) >= in.steps) {
            done.push_back(id);
        } else {
            keepActive.push_back(id);
        }
    }
    if (!done.empty()) {
        B->active.swap(keepActive);
        B->dirty = true;
        if (int r = Retire(B, done)) return r;
    }
    if (niters) *niters = iters;
    if (nactive) *nactive = (int)B->active.size();
    if (ncompleted) *ncompleted = (int)done.size();
    return PTK_BATCH_OK;
}

int ptk_batch_tick(ptk_batch *B, int max_admit, int *nactive, int *ncompleted) {
    return ptk_batch_run(B, max_admit, 1, nullptr, nactive, ncompleted);
}

int ptk_batch_drain(ptk_batch *B) {
    if (!B) return Fail(PTK_BATCH_EINVAL, "null executor");
    while (!B->queue.empty() || !B->active.empty()) {
        const size_t q = B->queue.size(), a = B->active.size();
        if (int r = ptk_batch_run(B, -1, 1 << 30, nullptr, nullptr, nullptr)) return r;
        if (a == 0 && B->active.empty() && B->queue.size() == q)
            return Fail(PTK_BATCH_ENOMEM, "a queued instance does not fit the free arena");
    }
    return PTK_BATCH_OK;
}

int ptk_batch_status(ptk_batch *B, int id, int *state, double *ts, double *ta, double *td) {
    if (!B || id < 0 || id >= (int)B->insts.size()) return Fail(PTK_BATCH_EINVAL, "bad instance id");
    const Inst &in = B->insts[(size_t)id];
    if (state) *state = in.state;
    if (ts) *ts = in.tSubmit;
    if (ta) *ta = in.tAdmit;
    if (td) *td = in.tDone;
    return PTK_BATCH_OK;
}

int ptk_batch_checksums(ptk_batch *B, int id, unsigned long long out[8]) {
    if (!B || id < 0 || id >= (int)B->insts.size() || !out) return Fail(PTK_BATCH_EINVAL, "bad instance id");
    const Inst &in = B->insts[(size_t)id];
    if (in.state < PTK_STATE_DONE) return Fail(PTK_BATCH_ESTATE, "instance not complete");
    memcpy(out, in.csum, 64);
    return PTK_BATCH_OK;
}

int ptk_batch_read(ptk_batch *B, int id, int which, void *host, unsigned long long *bytes) {
    if (!B || id < 0 || id >= (int)B->insts.size()) return Fail(PTK_BATCH_EINVAL, "bad instance id");
    const Inst &in = B->insts[(size_t)id];
    if (in.state != PTK_STATE_DONE) return Fail(PTK_BATCH_ESTATE, "instance not complete or released");
    const void *src = nullptr;
    unsigned long long n = 0;
    if (in.kind == PTK_INST_FDTD && which >= 0 && which < 6) {
        src = in.fd.f[in.stepsDone & 1][which];
        n = 4 * ptk_fdtd_field_elems(which, (int)in.a, in.b, in.c);
    } else if (in.kind == PTK_INST_GROUPBY && which >= 0 && which < 6) {
        const unsigned long long K = (unsigned long long)in.b + 1;
        const void *p[6] = {in.gb.keys, in.gb.vals, in.gb.groups, in.gb.offsets, in.gb.outKeys, in.gb.outVals};
        const unsigned long long s[6] = {4ull * in.a, 4ull * in.a, 8 * (3 + 3 * K), 8 * K, 4ull * in.a, 4ull * in.a};
        src = p[which];
        n = s[which];
    } else if (in.kind == PTK_INST_GMM && (which == 0 || which == 1)) {
        src = which == 0 ? (const void *)in.gm.X : (const void *)in.gm.model;
        n = which == 0 ? 4ull * in.a * in.b : 8ull * ptk_gmm_model_elems(&in.gm.d);
    } else {
        return Fail(PTK_BATCH_EINVAL, "bad result index");
    }
    if (bytes) *bytes = n;
    if (host) {
        HIPCHK(ptk_hip_stream_synchronize(B->stream));
        HIPCHK(ptk_hip_memcpy_dtoh(host, src, n));
    }
    return PTK_BATCH_OK;
}

int ptk_batch_release(ptk_batch *B, int id) {
    if (!B || id < 0 || id >= (int)B->insts.size()) return Fail(PTK_BATCH_EINVAL, "bad instance id");
    Inst &in = B->insts[(size_t)id];
    if (in.state != PTK_STATE_DONE) return Fail(PTK_BATCH_ESTATE, "instance not complete or already released");
    HIPCHK(ptk_hip_stream_synchronize(B->stream));
    B->arena.Free(in.off, in.bytes);
    in.state = PTK_STATE_RELEASED;
    return PTK_BATCH_OK;
}

int ptk_batch_set_tiles(ptk_batch *B, int cols, int depth, int gbRows) {
    if (!B || cols < 0 || depth < 1 || depth > 65535 || gbRows < PTK_GB_CHUNK || gbRows % PTK_GB_CHUNK)
        return Fail(PTK_BATCH_EINVAL, "bad tile shape");
    B->fdtdCols = cols;
    B->fdtdDepth = depth;
    B->gbRows = gbRows;
    B->dirty = true;
    return PTK_BATCH_OK;
}

int ptk_batch_set_fdtd_steps(ptk_batch *B, int steps) {
    if (!B || (steps != 1 && steps != 3)) return Fail(PTK_BATCH_EINVAL, "FDTD steps per launch must be 1 or 3");
    B->fdtdSteps = steps;
    B->dirty = true;
    return PTK_BATCH_OK;
}

int ptk_batch_get_fdtd_steps(ptk_batch *B, int *steps, int *blocks) {
    if (!B) return Fail(PTK_BATCH_EINVAL, "null executor");
    if (steps) *steps = B->fdtdSteps;
    if (blocks) *blocks = B->grid3;
    return PTK_BATCH_OK;
}

int ptk_batch_set_launch(ptk_batch *B, int graphs) {
    if (!B) return Fail(PTK_BATCH_EINVAL, "null executor");
    B->useGraphs = graphs != 0;
    return PTK_BATCH_OK;
}

int ptk_batch_get_tiles(ptk_batch *B, int *cols, int *depth, int *gbRows) {
    if (!B) return Fail(PTK_BATCH_EINVAL, "null executor");
    if (cols) *cols = B->fdtdCols;
    if (depth) *depth = B->fdtdDepth;
    if (gbRows) *gbRows = B->gbRows;
    return PTK_BATCH_OK;
}

// Times one FDTD iteration (the half-step pair, the whole step or three whole steps) over the
// resident instances for each candidate tile shape and persistent grid size (an iteration only
// rewrites the next field set from the current one, so repeating it changes nothing) and keeps
// the fastest per time step.  Groupby chunk rows are fixed at
// admission and are not tuned here.
int ptk_batch_get_tuning(ptk_batch *B, int out[5]) {
    if (!B || !out) return Fail(PTK_BATCH_EINVAL, "null argument");
    out[0] = B->fdtdCols;
    out[1] = B->fdtdDepth;
    out[2] = B->fdtdGrid;
    out[3] = B->fdtdSteps;
    out[4] = B->grid3;
    return PTK_BATCH_OK;
}

int ptk_batch_set_tuning(ptk_batch *B, const int in[5]) {
    if (!B || !in) return Fail(PTK_BATCH_EINVAL, "null argument");
    if (in[0] < 0 || in[1] < 1 || in[1] > 65535 || in[2] < 1 || in[2] > (1 << 20) || (in[3] != 1 && in[3] != 3) ||
        in[4] < 1 || in[4] > (1 << 16))
        return Fail(PTK_BATCH_EINVAL, "tuning out of range");
    B->fdtdCols = in[0];
    B->fdtdDepth = in[1];
    B->fdtdGrid = in[2];
    B->fdtdSteps = in[3];
    B->grid3 = in[4];
    DropGraphs(B);  // the captured iterations hold the grid sizes
    B->dirty = true;
    return PTK_BATCH_OK;
}

int ptk_batch_autotune(ptk_batch *B, double *best_ms) {
    if (!B) return Fail(PTK_BATCH_EINVAL, "null executor");
    bool anyFdtd = false;
    for (int id : B->active) anyFdtd |= B->insts[(size_t)id].kind == PTK_INST_FDTD;
    if (!anyFdtd || B->mode != PTK_BATCH_BATCHED) {
        if (best_ms) *best_ms = 0;
        return PTK_BATCH_OK;
    }
    // {columns, planes, steps per launch} of half-step tiles; {0, planes, 1}: whole-step
    // overlapped tiles; {0, planes, 3}: three steps per launch on overlapped tiles (compared per
    // step; the one-step whole-step tiles it falls back to near the end keep their shape)
    static const int cand[][3] = {{0, 4, 1}, {0, 8, 1}, {0, 16, 1}, {256, 8, 1}, {512, 16, 1}, {1024, 32, 1},
                                  {0, 0, 3}};
    static const int grids[] = {2048, 4096, 8192};
    const int grids3[] = {B->cus, 2 * B->cus};  // virtual blocks of the three-step schedule
    double best = 1e30;
    int bc = B->fdtdCols, bd = B->fdtdDepth, bg = B->fdtdGrid, bs = B->fdtdSteps, bg3 = B->grid3;
    const int oneCols = bs == 3 ? 0 : bc, oneDepth = bd;  // the one-step shape kept beside three steps
    for (auto &c : cand) {
        const int T = c[2];
        B->fdtdSteps = T;
        B->fdtdCols = T == 3 ? 0 : c[0];
        B->fdtdDepth = T == 3 ? (oneCols == 0 ? oneDepth : 8) : c[1];
        for (int g : T == 3 ? std::vector<int>(grids3, grids3 + 2) : std::vector<int>(grids, grids + 3)) {
            (T == 3 ? B->grid3 : B->fdtdGrid) = g;
            if (int r = BuildLists(B)) return r;  // the three-step schedule depends on the grid
            if (T == 3 && !B->cur3) break;        // an instance too large for the whole-step forms
            if (int r = IssueBatched(B, true, false, false, false, false, T)) return r;  // warm
            HIPCHK(ptk_hip_event_record(B->evStart, B->stream));
            for (int rep = 0; rep < 2; ++rep)
                if (int r = IssueBatched(B, true, false, false, false, false, T)) return r;
            HIPCHK(ptk_hip_event_record(B->evStop, B->stream));
            HIPCHK(ptk_hip_event_synchronize(B->evStop));
            float ms = 0;
            HIPCHK(ptk_hip_event_elapsed_ms(&ms, B->evStart, B->evStop));
            const double perStep = ms / 2 / T;
            if (perStep < best) {
                best = perStep;
                bs = T;
                if (T == 3) {
                    bc = 0;
                    bd = B->fdtdDepth;
                    bg3 = g;
                } else {
                    bc = c[0];
                    bd = c[1];
                    bg = g;
                }
            }
        }
    }
    B->fdtdSteps = bs;
    B->grid3 = bg3;
    B->fdtdCols = bc;
    B->fdtdDepth = bd;
    B->fdtdGrid = bg;
    DropGraphs(B);  // the captured iterations hold the grid size
    if (int r = BuildLists(B)) return r;
    B->dirty = false;
    if (best_ms) *best_ms = best;
    return PTK_BATCH_OK;
}

int ptk_batch_stats(ptk_batch *B, unsigned long long out[7]) {
    if (!B || !out) return Fail(PTK_BATCH_EINVAL, "bad argument");
    out[6] = (unsigned long long)(B->execMs * 1000.0);
    out[0] = B->nTicks;
    out[1] = B->nLaunches;
    out[2] = B->nGraphs;
    out[3] = B->arena.used();
    out[4] = B->arena.peak();
    out[5] = B->arena.size();
    return PTK_BATCH_OK;
}

}  // extern "C"
