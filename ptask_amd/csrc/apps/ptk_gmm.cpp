// ptk_gmm.cpp -- EM for a diagonal Gaussian mixture as a PTask graph, against ptaskapi.h only.
//
// The reference sketches this graph (ptaskUnitTest/graphemgmm.cpp:124-211: an E task and an M
// task per iteration, the first fed from a graph input, the last feeding a graph output) but
// never builds it; its EM runs on the host (graphemgmm.cpp:92-122, NR3 Gaumixmod).  Two tasks
// carry one EM step here:
//   GmmAccumulate (X sticky, dims sticky, model)      -> partials  (E-step + M-step sums)
//   GmmFinalize   (model in/out, partials, dims)      -> model     (M-step update)
// --mode looped:   both tasks carry a MF_GENERAL_ITERATOR meta port; the model is loop-carried
//                  (switch channel for the first iteration, back-channels closed on
//                  ENDITERATION, output channel opened on ENDITERATION).
// --mode unrolled: 2 x steps tasks, as graphemgmm.cpp:195-208 lays them out.
//
// The model block (binary64) holds the step count, mu, var, frac and the log-likelihood of every
// E-step; the initial model follows Gaumixmod's constructor (frac 1/K, var 1e-10) with means
// taken from K of the points.  Points come from a seeded mixture (means U[-10,10], standard
// deviations U[0.5,2] per dimension, uniform component choice) or from --data (raw binary32,
// N x D).  Prints one JSON object; --dump writes X, the initial means and the final model.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ptaskapi.h"
#include "ptk_gmm.h"
#include "ptk_prng.h"

using namespace PTask;

namespace {

struct Options {
    std::string kernelFile, data, dump, mode = "looped", init;
    long long n = 100000;
    int d = 8, k = 8, steps = 10;
    int blocks = 0;  // GmmAccumulate blocks; 0 = min(tiles, 768): three 128-point tiles per CU
    int repeat = 1, warmup = 0;
    unsigned long long seed = 0x6a5d39;
    bool profile = false;
    // HIP stream policy: one shared stream (each EM step is a chain of two dependent tasks, which
    // then needs no cross-stream event waits), or one per task
    std::string streams = "shared";
    std::string accumulate;  // the accumulate kernel: valu (GmmAccumulate) or mfma (the HIP default)
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

// Seeded mixture: component c of point p is uniform on [0, K); coordinates are mean + sd * z with
// z from Box-Muller over two independent uniforms of the shared generator.  first[c] receives the
// index of the first point drawn from component c (-1 if none).
void MakePoints(const Options &o, std::vector<float> &X, std::vector<long long> &first) {
    const int D = o.d, K = o.k;
    std::vector<double> mean((size_t)K * D), sd((size_t)K * D);
    for (size_t i = 0; i < mean.size(); ++i) {
        mean[i] = -10.0 + 20.0 * ptk_rand_unit_f64(o.seed, 10, i);
        sd[i] = 0.5 + 1.5 * ptk_rand_unit_f64(o.seed, 11, i);
    }
    X.resize((size_t)o.n * D);
    first.assign((size_t)K, -1);
    for (long long p = 0; p < o.n; ++p) {
        int c = (int)(ptk_rand_unit_f64(o.seed, 12, (uint64_t)p) * K);
        if (c >= K) c = K - 1;
        if (first[(size_t)c] < 0) first[(size_t)c] = p;
        for (int m = 0; m < D; ++m) {
            const uint64_t i = (uint64_t)p * D + m;
            const double u1 = 1.0 - ptk_rand_unit_f64(o.seed, 13, i);  // (0, 1]
            const double u2 = ptk_rand_unit_f64(o.seed, 14, i);
            const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
            X[(size_t)i] = (float)(mean[(size_t)c * D + m] + sd[(size_t)c * D + m] * z);
        }
    }
}

// JSON number, or null when not finite
std::string Num(double v) {
    char buf[40];
    if (!std::isfinite(v)) return "null";
    snprintf(buf, sizeof(buf), "%.9g", v);
    return buf;
}

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
        else if (a == "--n") o.n = atoll(next());
        else if (a == "--d") o.d = atoi(next());
        else if (a == "--k") o.k = atoi(next());
        else if (a == "--steps") o.steps = atoi(next());
        else if (a == "--mode") o.mode = next();
        else if (a == "--init") o.init = next();
        else if (a == "--blocks") o.blocks = atoi(next());
        else if (a == "--repeat") o.repeat = atoi(next());
        else if (a == "--warmup") o.warmup = atoi(next());
        else if (a == "--seed") o.seed = strtoull(next(), nullptr, 0);
        else if (a == "--data") o.data = next();
        else if (a == "--dump") o.dump = next();
        else if (a == "--profile") o.profile = true;
        else if (a == "--streams") o.streams = next();
        else if (a == "--accumulate") o.accumulate = next();
        else { fprintf(stderr, "unknown option %s\n", a.c_str()); return 2; }
    }
    const long long tiles = (o.n + 127) / 128;
    if (o.blocks <= 0) o.blocks = (int)std::max(1LL, std::min(tiles, 768LL));
    if (o.kernelFile.empty() || o.n < 1 || o.n > 0x7fffffffLL || o.d < 1 || o.d > PTK_GMM_MAX_D ||
        o.k < 1 || o.k > PTK_GMM_MAX_K || o.k > o.n || o.steps < 1 || o.repeat < 1 ||
        (o.mode != "looped" && o.mode != "unrolled") || (long long)o.n * o.d > 0x3fffffffLL ||
        (!o.init.empty() && o.init != "strided" && o.init != "components") ||
        (o.init == "components" && !o.data.empty()) || (o.streams != "shared" && o.streams != "task") ||
        (!o.accumulate.empty() && o.accumulate != "valu" && o.accumulate != "mfma")) {
        fprintf(stderr, "usage: ptk_gmm --kernels <gmm.hsaco|hosttasks.so> [--n N] [--d D<=32] "
                        "[--k K<=64] [--steps S] [--mode looped|unrolled] [--init strided|components] [--blocks B] [--repeat R] "
                        "[--warmup W] [--seed S] [--data x.f32] [--dump file] [--profile] [--streams shared|task] "
                        "[--accumulate valu|mfma]\n");
        return 2;
    }
    std::vector<float> X;
    std::vector<long long> first;
    if (!o.data.empty()) {
        X.resize((size_t)o.n * o.d);
        FILE *fp = fopen(o.data.c_str(), "rb");
        if (!fp || fread(X.data(), sizeof(float), X.size(), fp) != X.size()) {
            fprintf(stderr, "cannot read %lld x %d binary32 points from %s\n", o.n, o.d, o.data.c_str());
            return 2;
        }
        fclose(fp);
    } else {
        MakePoints(o, X, first);
    }
    GMM_DIMS dims{};
    dims.N = (int)o.n;
    dims.D = o.d;
    dims.K = o.k;
    dims.blocks = o.blocks;
    dims.maxIter = o.steps;
    const long long modelElems = ptk_gmm_model_elems(&dims);
    const long long partElems = ptk_gmm_partial_stride(&dims) * dims.blocks;
    // Initial model: Gaumixmod's constructor (gaumixmod.h:14-27).  Means are K of the points:
    // "strided" takes points c*N/K; "components" (the default for generated points) takes the
    // first point drawn from each generating component, so that no two initial means share a
    // component -- with shared ones, unregularised EM lets the less central mean shrink onto a
    // single point (variance 0, then 0/0), in binary64 as in binary32.
    if (o.init.empty()) o.init = o.data.empty() ? "components" : "strided";
    std::vector<double> model0((size_t)modelElems, 0.0);
    model0[1] = (double)o.n;
    model0[2] = o.d;
    model0[3] = o.k;
    for (int c = 0; c < o.k; ++c) {
        long long p = (long long)c * o.n / o.k;
        if (o.init == "components" && first[(size_t)c] >= 0) p = first[(size_t)c];
        for (int m = 0; m < o.d; ++m) {
            model0[(size_t)(ptk_gmm_mu(&dims) + (long long)c * o.d + m)] = X[(size_t)(p * o.d + m)];
            model0[(size_t)(ptk_gmm_var(&dims) + (long long)c * o.d + m)] = 1.0e-10;
        }
        model0[(size_t)(ptk_gmm_frac(&dims) + c)] = 1.0 / o.k;
    }
    ptk_gmm_fill_table(&dims, model0.data());

    Runtime::SetExitOnRuntimeFailure(FALSE);
    Runtime::SetTaskStreamPolicy(o.streams == "shared" ? PTK_STREAMS_SHARED : PTK_STREAMS_PER_TASK);
    if (o.profile) Runtime::SetProfilePSDispatch(TRUE);
    if (PTFAILED(Runtime::Initialize())) return 1;
    const bool hip = ptaskutils::SelectAcceleratorClass(o.kernelFile.c_str()) == ACCELERATOR_CLASS_HIP;
    DatablockTemplate *tX = Runtime::GetDatablockTemplate((char *)"gmmPoints", sizeof(float) * o.d, (UINT)o.n, 1, 1);
    DatablockTemplate *tModel = Runtime::GetDatablockTemplate((char *)"gmmModel", sizeof(double), (UINT)modelElems, 1, 1);
    DatablockTemplate *tPart = Runtime::GetDatablockTemplate((char *)"gmmPartials", sizeof(double), (UINT)partElems, 1, 1);
    DatablockTemplate *tDims = Runtime::GetDatablockTemplate((char *)"gmmDims", sizeof(GMM_DIMS), PTPARM_BYVALSTRUCT);
    DatablockTemplate *tIter = Runtime::GetDatablockTemplate((char *)"gmmIters", sizeof(int), PTPARM_INT);
    // the accumulate task: GmmAccumulate on the VALU, or GmmAccumulateMfma with its products on
    // the matrix cores (same arguments and partial rows; HIP code objects only)
    if (o.accumulate.empty())
        o.accumulate = ptaskutils::SelectAcceleratorClass(o.kernelFile.c_str()) == ACCELERATOR_CLASS_HIP ? "mfma" : "valu";
    const char *accName = o.accumulate == "mfma" ? "GmmAccumulateMfma" : "GmmAccumulate";
    CompiledKernel *kAcc = Runtime::GetCompiledKernel((char *)o.kernelFile.c_str(), (char *)accName);
    CompiledKernel *kFin = Runtime::GetCompiledKernel((char *)o.kernelFile.c_str(), (char *)"GmmFinalize");
    if (!kAcc || !kFin) {
        fprintf(stderr, "GMM kernels unavailable: %s\n", Runtime::GetLastErrorText());
        Runtime::Terminate();
        return 1;
    }
    Graph *g = new Graph("emgmm");
    UINT uid = 0;
    std::vector<GraphInputChannel *> xIn, dimsIn, modelIn, iterIn;
    GraphOutputChannel *modelOut = nullptr;
    const bool looped = o.mode == "looped";
    const int stages = looped ? 1 : o.steps;
    const int finGrid = o.k + 1;
    OutputPort *prevModel = nullptr;
    for (int s = 0; s < stages; ++s) {
        const std::string u = looped ? "" : "_" + std::to_string(s);
        Port *aIn[4] = {Runtime::CreatePort(STICKY_PORT, tX, uid++, (char *)"X", 0),
                        Runtime::CreatePort(STICKY_PORT, tDims, uid++, (char *)"dims", 1),
                        Runtime::CreatePort(INPUT_PORT, tModel, uid++, (char *)"model", 2), nullptr};
        Port *aOut[1] = {Runtime::CreatePort(OUTPUT_PORT, tPart, uid++, (char *)"partials", 3)};
        Port *fIn[4] = {Runtime::CreatePort(INPUT_PORT, tModel, uid++, (char *)"model", 0, 0),
                        Runtime::CreatePort(INPUT_PORT, tPart, uid++, (char *)"partials", 1),
                        Runtime::CreatePort(STICKY_PORT, tDims, uid++, (char *)"dims", 2), nullptr};
        Port *fOut[1] = {Runtime::CreatePort(OUTPUT_PORT, tModel, uid++, (char *)"model(out)", 0)};
        // The next finalize writes the model in place while the next accumulate reads it, but the
        // finalize also consumes that accumulate's partials, so it cannot start before the read
        // ends: the two share the block and the runtime makes no copy of it.
        fOut[0]->SetSuppressClones(TRUE);
        UINT na = 3, nf = 3;
        if (looped) {
            aIn[3] = Runtime::CreatePort(META_PORT, tIter, uid++, (char *)"iterations");
            aIn[3]->SetMetaFunction(MF_GENERAL_ITERATOR);
            fIn[3] = Runtime::CreatePort(META_PORT, tIter, uid++, (char *)"iterations");
            fIn[3]->SetMetaFunction(MF_GENERAL_ITERATOR);
            na = nf = 4;
        }
        Task *ta = g->AddTask(kAcc, na, aIn, 1, aOut, ("GmmAccumulate" + u).c_str());
        Task *tf = g->AddTask(kFin, nf, fIn, 1, fOut, ("GmmFinalize" + u).c_str());
        if (hip) {
            ta->SetBlockAndGridSize(PTASKDIM3(o.blocks, 1, 1), PTASKDIM3(256, 1, 1));
            tf->SetBlockAndGridSize(PTASKDIM3(finGrid, 1, 1), PTASKDIM3(256, 1, 1));
        } else {
            ta->SetComputeGeometry(1, 1, 1);
            tf->SetComputeGeometry(1, 1, 1);
        }
        xIn.push_back(g->AddInputChannel(aIn[0], ("X" + u).c_str()));
        dimsIn.push_back(g->AddInputChannel(aIn[1], ("dims->acc" + u).c_str()));
        dimsIn.push_back(g->AddInputChannel(fIn[2], ("dims->fin" + u).c_str()));
        g->AddInternalChannel(aOut[0], fIn[1], ("partials" + u).c_str());
        if (looped) {
            modelIn.push_back(g->AddInputChannel(aIn[2], "model->acc.init", TRUE));
            modelIn.push_back(g->AddInputChannel(fIn[0], "model->fin.init", TRUE));
            InternalChannel *ba = g->AddInternalChannel(fOut[0], aIn[2], "model->acc.back");
            ba->SetPredicationType(CE_SRC, CGATEFN_CLOSE_ON_ENDITERATION);
            InternalChannel *bf = g->AddInternalChannel(fOut[0], fIn[0], "model->fin.back");
            bf->SetPredicationType(CE_SRC, CGATEFN_CLOSE_ON_ENDITERATION);
            iterIn.push_back(g->AddInputChannel(aIn[3], "iters->acc"));
            iterIn.push_back(g->AddInputChannel(fIn[3], "iters->fin"));
            g->BindControlPropagationPort(aIn[3], aOut[0]);
            g->BindControlPropagationPort(fIn[3], fOut[0]);
            modelOut = g->AddOutputChannel(fOut[0], "model.out");
            modelOut->SetPredicationType(CE_SRC, CGATEFN_OPEN_ON_ENDITERATION);
        } else {
            if (s == 0) {
                modelIn.push_back(g->AddInputChannel(aIn[2], "model->acc_0"));
                modelIn.push_back(g->AddInputChannel(fIn[0], "model->fin_0"));
            } else {
                g->AddInternalChannel(prevModel, aIn[2], ("model->acc" + u).c_str());
                g->AddInternalChannel(prevModel, fIn[0], ("model->fin" + u).c_str());
            }
            prevModel = static_cast<OutputPort *>(fOut[0]);
            if (s == stages - 1) modelOut = g->AddOutputChannel(fOut[0], "model.out");
        }
    }
    g->Run();
    Datablock *bDims = Runtime::AllocateDatablock(tDims, &dims, sizeof(dims), nullptr);
    for (auto *c : dimsIn) c->Push(bDims);
    bDims->Release();
    Datablock *bX = Runtime::AllocateDatablock(tX, X.data(), (UINT)(X.size() * sizeof(float)), nullptr);
    for (auto *c : xIn) c->Push(bX);
    bX->Release();
    Accelerator *hipAcc = hip ? Runtime::GetAccelerator(ACCELERATOR_CLASS_HIP, 0) : nullptr;
    CHighResolutionTimer timer(gran_msec);
    double timedMs = 0;
    std::vector<double> result;
    bool ok = true;
    const int runs = o.warmup + o.repeat;
    unsigned long long h2d = 0, d2h = 0;
    std::vector<std::pair<double, UINT>> base;
    for (int r = 0; r < runs && ok; ++r) {
        if (r == o.warmup) {
            g->WaitQuiescent(600000);
            if (hipAcc) hipAcc->Synchronize();
            Runtime::ResetTransferStatistics();
            base.clear();
            for (UINT t = 0; t < g->GetTaskCount(); ++t)
                base.push_back(std::make_pair(g->GetTaskByIndex(t)->GetTotalDispatchDeviceMs(),
                                              g->GetTaskByIndex(t)->GetDispatchCount()));
            timer.reset();
        }
        Datablock *bModel = Runtime::AllocateDatablock(tModel, model0.data(), (UINT)(model0.size() * sizeof(double)), nullptr);
        for (auto *c : modelIn) c->Push(bModel);
        bModel->Release();
        if (looped) {
            Datablock *it = Runtime::AllocateDatablock(tIter, &o.steps, sizeof(int), nullptr);
            for (auto *c : iterIn) c->Push(it);
            it->Release();
        }
        Datablock *res = modelOut->Pull(600000);
        if (!res) {
            fprintf(stderr, "no model output: %s\n", Runtime::GetLastErrorText());
            ok = false;
            break;
        }
        if (r == runs - 1) {
            res->Lock();
            const double *p = (const double *)res->GetDataPointer(FALSE);
            result.assign(p, p + modelElems);
            res->Unlock();
            timedMs = timer.elapsed(false);
            Runtime::GetTransferStatistics(&h2d, nullptr, &d2h, nullptr);
        }
        res->Release();
    }
    std::string taskJson;
    if (o.profile) {
        for (UINT t = 0; t < g->GetTaskCount(); ++t) {
            Task *task = g->GetTaskByIndex(t);
            double ms = task->GetTotalDispatchDeviceMs();
            UINT nd = task->GetDispatchCount();
            if (t < base.size()) {
                ms -= base[t].first;
                nd -= base[t].second;
            }
            char buf[256];
            snprintf(buf, sizeof(buf), "%s{\"task\": \"%s\", \"dispatches\": %u, \"device_ms\": %.4f}",
                     taskJson.empty() ? "" : ", ", task->GetTaskName(), nd, ms);
            taskJson += buf;
        }
    }
    g->Stop();
    g->Teardown();
    Graph::DestroyGraph(g);
    if (ok && !o.dump.empty()) {
        FILE *fp = fopen(o.dump.c_str(), "wb");
        if (fp) {
            fwrite(X.data(), sizeof(float), X.size(), fp);
            fwrite(model0.data() + ptk_gmm_mu(&dims), sizeof(double), (size_t)o.k * o.d, fp);
            fwrite(result.data(), sizeof(double), result.size(), fp);
            fclose(fp);
        }
    }
    // log-likelihood history and the extremes of the final model, for diagnostics
    std::string hist;
    double minVar = 0, minFrac = 0;
    int badVar = 0;
    if (!result.empty()) {
        for (int s = 0; s < o.steps; ++s) {
            char buf[40];
            snprintf(buf, sizeof(buf), "%s%.17g", s ? ", " : "", result[(size_t)(ptk_gmm_ll(&dims) + s)]);
            hist += std::isfinite(result[(size_t)(ptk_gmm_ll(&dims) + s)]) ? buf : (s ? ", null" : "null");
        }
        minVar = result[(size_t)ptk_gmm_var(&dims)];
        for (long long i = 0; i < (long long)o.k * o.d; ++i) {
            const double v = result[(size_t)(ptk_gmm_var(&dims) + i)];
            if (!(v > 0) || !std::isfinite(v)) ++badVar;
            else minVar = std::min(minVar, v);
        }
        minFrac = result[(size_t)ptk_gmm_frac(&dims)];
        for (int c = 0; c < o.k; ++c) minFrac = std::min(minFrac, result[(size_t)(ptk_gmm_frac(&dims) + c)]);
    }
    const double stepsDone = result.empty() ? 0 : result[0];
    const double ll = result.empty() ? 0 : result[(size_t)(ptk_gmm_ll(&dims) + o.steps - 1)];
    const double msRun = o.repeat ? timedMs / o.repeat : 0.0;
    char llText[40] = "null";
    if (std::isfinite(ll)) snprintf(llText, sizeof(llText), "%.17g", ll);
    printf("{\"ok\": %s, \"backend\": \"%s\", \"accumulate\": \"%s\", \"mode\": \"%s\", \"init\": \"%s\", \"n\": %lld, \"d\": %d, \"k\": %d, "
           "\"steps\": %d, \"blocks\": %d, \"runs\": %d, \"timed_runs\": %d, \"timed_ms\": %.4f, "
           "\"ms_per_run\": %.4f, \"point_steps_per_s\": %.1f, \"steps_done\": %.0f, \"loglike\": %s, "
           "\"h2d_count_timed\": %llu, \"d2h_count_timed\": %llu, \"tasks\": [%s], \"model_fnv1a\": \"%016llx\", "
           "\"min_var\": %s, \"nonpositive_var\": %d, \"min_frac\": %s, \"loglike_history\": [%s]}\n",
           ok ? "true" : "false", hip ? "hip" : "host", o.accumulate.c_str(), o.mode.c_str(), o.init.c_str(), o.n, o.d, o.k,
           o.steps, o.blocks,
           runs, o.repeat, timedMs, msRun, msRun > 0 ? (double)o.n * o.steps / (msRun / 1e3) : 0.0,
           stepsDone, llText, h2d, d2h, taskJson.c_str(),
           (unsigned long long)Fnv1a(result.data(), result.size() * sizeof(double)), Num(minVar).c_str(), badVar,
           Num(minFrac).c_str(),
           hist.c_str());
    Runtime::Terminate();
    return ok && stepsDone == o.steps ? 0 : 1;
}
