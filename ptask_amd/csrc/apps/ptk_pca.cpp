// ptk_pca.cpp -- Gram-Schmidt PCA as an unrolled PTask graph (the role of
// ptaskUnitTest/graphcupca.cpp), against ptaskapi.h only.
//
// The reference unrolls one task per BLAS call (graphcupca.cpp:317-1540).  This graph unrolls the
// fused task kernels of pca.hsaco: for each component k
//   PcaStart [-> PcaFinish of k-1] -> J x (PcaReduce -> PcaOrthonormalize [-> PcaGram unless last])
// between PcaLoad and PcaDeflate + PcaFinish of the last component.  PcaStart of k forms component
// k-1's scores t = R p in the read of R that deflates it, so no separate projection pass reads R.
// The residual R, the scores T and the state S (loads, singular
// values, partial sums) are datablocks chained through in/out port pairs, so every task runs on
// the device-resident copy; each task has its own sticky PCA_DIMS carrying its component index.
// X (column-major M x N binary32) is N(0,1) or U[0,1) from the shared generator, or --data.
// --capture records the first run's launches and replays each later run as one hipGraph launch
// (Graph::EnableLaunchCapture).  Prints one JSON object; --dump writes X, then T (M x K), P (N x K), R (M x N) column-major
// binary32 and L (K binary64).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ptaskapi.h"
#include "ptk_pca.h"
#include "ptk_prng.h"

using namespace PTask;

namespace {

struct Options {
    std::string kernelFile, data, dump, dist = "normal";
    int m = 4096, n = 512, k = 4, j = 10;
    int center = 1, blocks = 0, gr = 0;
    int repeat = 1, warmup = 0;
    unsigned long long seed = 2011;
    bool profile = false;
    // HIP stream policy: one shared stream (the graph is one chain of dependent tasks, which then
    // needs no cross-stream event waits; 13% faster per run at 65536 x 2048, K=16, J=10), or one
    // per task
    std::string streams = "shared";
    bool capture = false;  // record the first run's launches and replay later runs as one hipGraph
    int threads = 0;               // graph runner threads (0: the runtime default)
};

uint64_t Fnv1a(const void *p, size_t n) {
    const unsigned char *b = (const unsigned char *)p;
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < n; ++i) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}

enum { BR = 0, BT = 1, BS = 2 };

struct PcaGraph {
    Graph *g = nullptr;
    bool hip = false;
    std::string kernels;
    UINT uid = 0;
    DatablockTemplate *tmpl[3] = {nullptr, nullptr, nullptr};
    DatablockTemplate *tDims = nullptr;
    OutputPort *cur[3] = {nullptr, nullptr, nullptr};
    std::vector<std::pair<GraphInputChannel *, PCA_DIMS>> dimsIn;
    int steps = 0;

    bool Step(const char *kernel, std::initializer_list<int> blocks, const PCA_DIMS &d, PTASKDIM3 grid,
              PTASKDIM3 block) {
        CompiledKernel *ck = Runtime::GetCompiledKernel((char *)kernels.c_str(), (char *)kernel);
        if (!ck) {
            fprintf(stderr, "kernel %s unavailable: %s\n", kernel, Runtime::GetLastErrorText());
            return false;
        }
        std::vector<Port *> ins, outs;
        std::vector<int> bl(blocks);
        static const char *names[3] = {"R", "T", "S"};
        for (size_t i = 0; i < bl.size(); ++i) {
            ins.push_back(Runtime::CreatePort(INPUT_PORT, tmpl[bl[i]], uid++, (char *)names[bl[i]], (UINT)i, (UINT)i));
            outs.push_back(Runtime::CreatePort(OUTPUT_PORT, tmpl[bl[i]], uid++, (char *)names[bl[i]], (UINT)i));
        }
        Port *dp = Runtime::CreatePort(STICKY_PORT, tDims, uid++, (char *)"dims", (UINT)bl.size());
        ins.push_back(dp);
        const std::string name = std::string(kernel) + "_" + std::to_string(steps++);
        Task *t = g->AddTask(ck, (UINT)ins.size(), ins.data(), (UINT)outs.size(), outs.data(), name.c_str());
        if (!t) return false;
        if (hip) t->SetBlockAndGridSize(grid, block);
        else t->SetComputeGeometry(1, 1, 1);
        for (size_t i = 0; i < bl.size(); ++i) {
            g->AddInternalChannel(cur[bl[i]], ins[i], (name + "." + names[bl[i]]).c_str());
            cur[bl[i]] = static_cast<OutputPort *>(outs[i]);
        }
        dimsIn.push_back(std::make_pair(g->AddInputChannel(dp, (name + ".dims").c_str()), d));
        return true;
    }
};

}  // namespace

int main(int argc, char **argv) {
    Options o;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> const char * {
            if (i + 1 >= argc) { fprintf(stderr, "missing value for %s\n", a.c_str()); exit(2); }
            return argv[++i];
        };
        if (a == "--kernels") o.kernelFile = next();
        else if (a == "--m") o.m = atoi(next());
        else if (a == "--n") o.n = atoi(next());
        else if (a == "--k") o.k = atoi(next());
        else if (a == "--j") o.j = atoi(next());
        else if (a == "--no-center") o.center = 0;
        else if (a == "--blocks") o.blocks = atoi(next());
        else if (a == "--gr") o.gr = atoi(next());
        else if (a == "--repeat") o.repeat = atoi(next());
        else if (a == "--warmup") o.warmup = atoi(next());
        else if (a == "--seed") o.seed = strtoull(next(), nullptr, 0);
        else if (a == "--dist") o.dist = next();
        else if (a == "--data") o.data = next();
        else if (a == "--dump") o.dump = next();
        else if (a == "--profile") o.profile = true;
        else if (a == "--streams") o.streams = next();
        else if (a == "--threads") o.threads = atoi(next());
        else if (a == "--capture") o.capture = true;
        else { fprintf(stderr, "unknown option %s\n", a.c_str()); return 2; }
    }
    PCA_DIMS d{};
    d.M = o.m;
    d.N = o.n;
    d.K = o.k;
    d.J = o.j;
    d.ld = ptk_pca_ld(o.n);
    d.Mp = ptk_pca_mp(std::max(o.m, 1));
    d.center = o.center;
    const long long rBytes = 4ll * d.Mp * d.ld;
    if (o.kernelFile.empty() || o.m < 1 || o.n < 1 || o.n > PTK_PCA_MAX_N || o.k < 1 || o.k > PTK_PCA_MAX_K ||
        o.k > o.n || o.j < 1 || o.repeat < 1 || rBytes > 0xF0000000LL || (o.dist != "normal" && o.dist != "uniform") ||
        (o.streams != "shared" && o.streams != "task") || o.threads < 0) {
        fprintf(stderr, "usage: ptk_pca --kernels <pca.hsaco|hosttasks.so> [--m M] [--n N<=4096] [--k K<=64] "
                        "[--j J] [--no-center] [--blocks G] [--gr GR] [--repeat R] [--warmup W] [--seed S] "
                        "[--dist normal|uniform] [--data x.f32] [--dump file] [--profile] [--streams shared|task] "
                        "[--threads T] [--capture]\n");
        return 2;
    }
    const int rb = (d.ld / 4 + 255) / 256 <= 2 ? 8 : 4;  // rows per slab of the row passes
    const int slabs = d.Mp / rb;
    // Row-pass blocks: one and a half per compute unit of the device (set once the runtime
    // knows it; 384 on the host).  Measured at 65536 x 2048, K=16, J=10, captured runs on 256
    // CUs: 256 blocks 27.6 ms, 320-384 26.3-26.7, 448 27.0, 512 28.0, 1024 29.5 per run.
    auto rowBlocks = [&](int cus) { return o.blocks > 0 ? o.blocks : std::min(slabs, cus > 0 ? 3 * cus / 2 : 384); };
    d.G = rowBlocks(0);
    d.GR = o.gr > 0 ? std::min(o.gr, d.G) : std::min(d.G, 8);  // 8: measured best at 65536 x 2048

    std::vector<float> X((size_t)o.m * o.n);
    if (!o.data.empty()) {
        FILE *fp = fopen(o.data.c_str(), "rb");
        if (!fp || fread(X.data(), sizeof(float), X.size(), fp) != X.size()) {
            fprintf(stderr, "cannot read %d x %d binary32 values from %s\n", o.m, o.n, o.data.c_str());
            return 2;
        }
        fclose(fp);
    } else {
        for (size_t i = 0; i < X.size(); ++i) {
            if (o.dist == "uniform") {
                X[i] = (float)ptk_rand_unit_f64(o.seed, 20, i);
            } else {
                const double u1 = 1.0 - ptk_rand_unit_f64(o.seed, 21, i), u2 = ptk_rand_unit_f64(o.seed, 22, i);
                X[i] = (float)(std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
            }
        }
    }

    Runtime::SetExitOnRuntimeFailure(FALSE);
    Runtime::SetTaskStreamPolicy(o.streams == "shared" ? PTK_STREAMS_SHARED : PTK_STREAMS_PER_TASK);
    if (o.threads) Runtime::SetTaskThreadPoolSize((UINT)o.threads);
    if (o.profile) Runtime::SetProfilePSDispatch(TRUE);
    if (PTFAILED(Runtime::Initialize())) return 1;
    PcaGraph G;
    G.kernels = o.kernelFile;
    G.hip = ptaskutils::SelectAcceleratorClass(o.kernelFile.c_str()) == ACCELERATOR_CLASS_HIP;
    if (G.hip) {
        Accelerator *acc = Runtime::GetAccelerator(ACCELERATOR_CLASS_HIP, 0);
        d.G = rowBlocks(acc ? (int)acc->GetCoreCount() : 0);
        d.GR = o.gr > 0 ? std::min(o.gr, d.G) : std::min(d.G, 8);
    }
    G.g = new Graph("pca");
    DatablockTemplate *tX = Runtime::GetDatablockTemplate((char *)"pcaX", sizeof(float), (UINT)o.m, (UINT)o.n, 1);
    G.tmpl[BR] = Runtime::GetDatablockTemplate((char *)"pcaR", sizeof(float) * d.ld, (UINT)d.Mp, 1, 1);
    G.tmpl[BT] = Runtime::GetDatablockTemplate((char *)"pcaT", sizeof(float), (UINT)d.Mp, (UINT)d.K, 1);
    G.tmpl[BS] = Runtime::GetDatablockTemplate((char *)"pcaS", 4, (UINT)(ptk_pca_state_bytes(&d) / 4), 1, 1);
    G.tDims = Runtime::GetDatablockTemplate((char *)"pcaDims", sizeof(PCA_DIMS), PTPARM_BYVALSTRUCT);

    // PcaLoad: X -> R, T, S
    CompiledKernel *kLoad = Runtime::GetCompiledKernel((char *)o.kernelFile.c_str(), (char *)"PcaLoad");
    if (!kLoad) {
        fprintf(stderr, "kernel PcaLoad unavailable: %s\n", Runtime::GetLastErrorText());
        return 1;
    }
    Port *lIn[2] = {Runtime::CreatePort(INPUT_PORT, tX, G.uid++, (char *)"X", 0),
                    Runtime::CreatePort(STICKY_PORT, G.tDims, G.uid++, (char *)"dims", 1)};
    Port *lOut[3] = {Runtime::CreatePort(OUTPUT_PORT, G.tmpl[BR], G.uid++, (char *)"R", 2),
                     Runtime::CreatePort(OUTPUT_PORT, G.tmpl[BT], G.uid++, (char *)"T", 3),
                     Runtime::CreatePort(OUTPUT_PORT, G.tmpl[BS], G.uid++, (char *)"S", 4)};
    Task *tl = G.g->AddTask(kLoad, 2, lIn, 3, lOut, "PcaLoad");
    if (G.hip) tl->SetBlockAndGridSize(PTASKDIM3(d.Mp / 64, 1, 1), PTASKDIM3(256, 1, 1));
    else tl->SetComputeGeometry(1, 1, 1);
    GraphInputChannel *xIn = G.g->AddInputChannel(lIn[0], "X");
    G.dimsIn.push_back(std::make_pair(G.g->AddInputChannel(lIn[1], "PcaLoad.dims"), d));
    for (int b = 0; b < 3; ++b) G.cur[b] = static_cast<OutputPort *>(lOut[b]);

    const PTASKDIM3 b256(256, 1, 1), b1024(1024, 1, 1);
    const PTASKDIM3 gRows(d.G, 1, 1), gRed((d.ld + 63) / 64, d.GR, 1), gOne(1, 1, 1);
    bool built = true;
    for (int k = 0; k < o.k && built; ++k) {
        PCA_DIMS dk = d, dp = d;
        dk.k = k;
        dp.k = k - 1;
        built = G.Step("PcaStart", {BR, BT, BS}, dk, gRows, b256);
        if (built && k > 0) built = G.Step("PcaFinish", {BS}, dp, gOne, b256);
        for (int j = 1; j <= o.j && built; ++j) {
            built = G.Step("PcaReduce", {BS}, dk, gRed, b256) && G.Step("PcaOrthonormalize", {BS}, dk, gOne, b1024);
            if (built && j < o.j) built = G.Step("PcaGram", {BR, BS}, dk, gRows, b256);
        }
    }
    PCA_DIMS dl = d;
    dl.k = o.k - 1;
    built = built && G.Step("PcaDeflate", {BR, BT, BS}, dl, gRows, b256) &&
            G.Step("PcaFinish", {BS}, dl, gOne, b256);
    if (!built) {
        Runtime::Terminate();
        return 1;
    }
    GraphOutputChannel *outR = G.g->AddOutputChannel(G.cur[BR], "R.out");
    GraphOutputChannel *outT = G.g->AddOutputChannel(G.cur[BT], "T.out");
    GraphOutputChannel *outS = G.g->AddOutputChannel(G.cur[BS], "S.out");
    if (o.capture && !G.g->EnableLaunchCapture(TRUE)) {
        fprintf(stderr, "launch capture: %s\n", Runtime::GetLastErrorText());
        Runtime::Terminate();
        return 1;
    }
    G.g->Run();
    for (auto &c : G.dimsIn) {
        Datablock *bd = Runtime::AllocateDatablock(G.tDims, &c.second, sizeof(PCA_DIMS), nullptr);
        c.first->Push(bd);
        bd->Release();
    }
    Datablock *bX = Runtime::AllocateDatablock(tX, X.data(), (UINT)(X.size() * sizeof(float)), nullptr);
    Accelerator *hipAcc = G.hip ? Runtime::GetAccelerator(ACCELERATOR_CLASS_HIP, 0) : nullptr;
    CHighResolutionTimer timer(gran_msec);
    double timedMs = 0;
    std::vector<unsigned char> res[3];
    std::vector<std::pair<double, UINT>> base;
    unsigned long long h2d = 0, d2h = 0;
    bool ok = true;
    const int runs = o.warmup + o.repeat;
    for (int r = 0; r < runs && ok; ++r) {
        if (r == o.warmup) {
            G.g->WaitQuiescent(600000);
            if (hipAcc) hipAcc->Synchronize();
            Runtime::ResetTransferStatistics();
            base.clear();
            for (UINT t = 0; t < G.g->GetTaskCount(); ++t)
                base.push_back(std::make_pair(G.g->GetTaskByIndex(t)->GetTotalDispatchDeviceMs(),
                                              G.g->GetTaskByIndex(t)->GetDispatchCount()));
            timer.reset();
        }
        xIn->Push(bX);
        const bool last = r == runs - 1;
        GraphOutputChannel *outs[3] = {outR, outT, outS};
        Datablock *got[3] = {nullptr, nullptr, nullptr};
        for (int b = 0; b < 3; ++b) {
            got[b] = outs[b]->Pull(600000);
            if (!got[b]) {
                fprintf(stderr, "no output %d: %s\n", b, Runtime::GetLastErrorText());
                ok = false;
            }
        }
        if (ok && last) {
            // the timed region ends when the last run's scores and state are on the host
            for (int b = BT; b <= BS; ++b) {
                got[b]->Lock();
                const unsigned char *p = (const unsigned char *)got[b]->GetDataPointer(FALSE);
                res[b].assign(p, p + got[b]->GetDataBufferLogicalSizeBytes());
                got[b]->Unlock();
            }
            timedMs = timer.elapsed(false);
            Runtime::GetTransferStatistics(&h2d, nullptr, &d2h, nullptr);
            if (!o.dump.empty()) {
                got[BR]->Lock();
                const unsigned char *p = (const unsigned char *)got[BR]->GetDataPointer(FALSE);
                res[BR].assign(p, p + got[BR]->GetDataBufferLogicalSizeBytes());
                got[BR]->Unlock();
            }
        }
        for (int b = 0; b < 3; ++b)
            if (got[b]) got[b]->Release();
    }
    std::string taskJson;
    double gramMs = 0;
    unsigned gramN = 0;
    if (o.profile) {
        // per-kernel totals over the timed runs
        std::vector<std::string> names;
        std::vector<double> ms;
        std::vector<unsigned> cnt;
        for (UINT t = 0; t < G.g->GetTaskCount(); ++t) {
            Task *task = G.g->GetTaskByIndex(t);
            std::string nm = task->GetTaskName();
            nm = nm.substr(0, nm.find('_'));
            double v = task->GetTotalDispatchDeviceMs();
            unsigned c = task->GetDispatchCount();
            if (t < base.size()) {
                v -= base[t].first;
                c -= base[t].second;
            }
            size_t i = std::find(names.begin(), names.end(), nm) - names.begin();
            if (i == names.size()) {
                names.push_back(nm);
                ms.push_back(0);
                cnt.push_back(0);
            }
            ms[i] += v;
            cnt[i] += c;
        }
        for (size_t i = 0; i < names.size(); ++i) {
            char buf[256];
            snprintf(buf, sizeof(buf), "%s{\"task\": \"%s\", \"dispatches\": %u, \"device_ms\": %.4f}",
                     taskJson.empty() ? "" : ", ", names[i].c_str(), cnt[i], ms[i]);
            taskJson += buf;
            if (names[i] == "PcaGram") {
                gramMs = ms[i];
                gramN = cnt[i];
            }
        }
    }
    bX->Release();
    const unsigned tasks = G.g->GetTaskCount();
    const bool captureArmed = G.g->IsLaunchCaptureArmed();
    const unsigned captureReplays = G.g->GetLaunchCaptureReplayCount();
    G.g->Stop();
    G.g->Teardown();
    Graph::DestroyGraph(G.g);
    std::vector<double> L((size_t)o.k, 0.0);
    if (ok && res[BS].size() >= (size_t)ptk_pca_state_bytes(&d))
        memcpy(L.data(), res[BS].data() + ptk_pca_off_L(&d), sizeof(double) * (size_t)o.k);
    if (ok && !o.dump.empty()) {
        FILE *fp = fopen(o.dump.c_str(), "wb");
        if (fp) {
            fwrite(X.data(), sizeof(float), X.size(), fp);
            const float *T = (const float *)res[BT].data();
            for (int k = 0; k < o.k; ++k) fwrite(T + (size_t)k * d.Mp, sizeof(float), (size_t)o.m, fp);
            const float *P = (const float *)(res[BS].data() + ptk_pca_off_P(&d));
            for (int k = 0; k < o.k; ++k) fwrite(P + (size_t)k * d.ld, sizeof(float), (size_t)o.n, fp);
            const float *R = (const float *)res[BR].data();
            std::vector<float> col((size_t)o.m);
            for (int c = 0; c < o.n; ++c) {
                for (int m = 0; m < o.m; ++m) col[(size_t)m] = R[(size_t)m * d.ld + c];
                fwrite(col.data(), sizeof(float), col.size(), fp);
            }
            fwrite(L.data(), sizeof(double), L.size(), fp);
            fclose(fp);
        }
    }
    // algorithmic HBM bytes of one run: PcaLoad reads X (twice when centring) and writes R; per
    // component PcaStart reads R (and writes it back when deflating) and J-1 PcaGram read it;
    // PcaDeflate reads and writes it once.
    const double xB = 4.0 * o.m * o.n, rB = (double)rBytes;
    const double runBytes = xB * (o.center ? 2 : 1) + rB + (double)o.k * o.j * rB + (o.k - 1) * rB + 2 * rB;
    const double msRun = o.repeat ? timedMs / o.repeat : 0;
    std::string lj;
    for (int k = 0; k < o.k; ++k) {
        char buf[40];
        snprintf(buf, sizeof(buf), "%s%.17g", k ? ", " : "", L[(size_t)k]);
        lj += buf;
    }
    printf("{\"ok\": %s, \"backend\": \"%s\", \"m\": %d, \"n\": %d, \"k\": %d, \"j\": %d, \"center\": %d, "
           "\"G\": %d, \"GR\": %d, \"tasks\": %u, \"runs\": %d, \"timed_runs\": %d, \"timed_ms\": %.4f, "
           "\"ms_per_run\": %.4f, \"runs_per_s\": %.3f, \"run_bytes\": %.0f, \"gram_ms\": %.4f, \"gram_dispatches\": %u, "
           "\"gram_bytes\": %lld, \"h2d_count_timed\": %llu, \"d2h_count_timed\": %llu, \"kernels\": [%s], "
           "\"L\": [%s], \"T_fnv1a\": \"%016llx\", \"S_fnv1a\": \"%016llx\", "
           "\"capture\": {\"requested\": %s, \"armed\": %s, \"replays\": %u}}\n",
           ok ? "true" : "false", G.hip ? "hip" : "host", o.m, o.n, o.k, o.j, o.center, d.G, d.GR, tasks, runs,
           o.repeat, timedMs, msRun, msRun > 0 ? 1e3 / msRun : 0.0, runBytes, gramMs, gramN, rBytes, h2d, d2h,
           taskJson.c_str(), lj.c_str(), (unsigned long long)Fnv1a(res[BT].data(), res[BT].size()),
           (unsigned long long)Fnv1a(res[BS].data(), res[BS].size()), o.capture ? "true" : "false",
           captureArmed ? "true" : "false", captureReplays);
    Runtime::Terminate();
    return ok ? 0 : 1;
}
