// ptk_fdtd.cpp -- the FDTD Yee graph as a PTask application (the role of
// ptaskUnitTest/graphfdtd.cpp), written against include/ptask/ptaskapi.h only.
//
// Two graph shapes:
//   looped   -- six tasks; each has a MF_GENERAL_ITERATOR meta port whose iteration count
//               arrives once per run, back-channels predicated CGATEFN_CLOSE_ON_ENDITERATION
//               and output channels CGATEFN_OPEN_ON_ENDITERATION (the idiom of
//               ptaskUnitTest/iteration.cpp:887-1060).
//   unrolled -- six tasks per time step, step i feeding step i+1 (graphfdtd.cpp:735-739).
// The same graph runs HIP task kernels (a gfx950 code object) or host-task functions (a
// shared library); the kernel file decides, as Runtime::GetCompiledKernel does.
//
// Prints one JSON object: timings, dispatch counts, host<->device traffic and a checksum per
// field (FNV-1a over the bytes, plus the float64 sum), and optionally dumps the final fields.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ptaskapi.h"
#include "ptk_fdtd.h"
#include "ptk_prng.h"

using namespace PTask;

namespace {

const char *kFieldNames[6] = {"Hx", "Hy", "Hz", "Ex", "Ey", "Ez"};
const char *kKernelNames[6] = {"HxComputation", "HyComputation", "HzComputation",
                               "ExComputation", "EyComputation", "EzComputation"};
// (own, in1, in2) per task, in step order (graphfdtd.cpp:448-560)
const int kTaskInputs[6][3] = {{0, 4, 5}, {1, 3, 5}, {2, 3, 4}, {3, 1, 2}, {4, 0, 2}, {5, 0, 1}};

struct Options {
    int nx = 32, ny = 32, nz = 32, steps = 10, blocks = 1, repeat = 1;
    std::string mode = "looped";
    std::string kernelFile;
    std::string geometry = "default";
    std::string dump;
    unsigned long long seed = 0x5EED;
    float courant = 0.5f;
    int threads = 0;
    bool singleThreaded = false;
    bool eager = false;
    bool profile = false;
};

bool Parse(int argc, char **argv, Options &o) {
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&](const char *what) -> const char * {
            if (i + 1 >= argc) {
                fprintf(stderr, "missing value for %s\n", what);
                exit(2);
            }
            return argv[++i];
        };
        if (a == "--nx") o.nx = atoi(next("--nx"));
        else if (a == "--ny") o.ny = atoi(next("--ny"));
        else if (a == "--nz") o.nz = atoi(next("--nz"));
        else if (a == "--n") o.nx = o.ny = o.nz = atoi(next("--n"));
        else if (a == "--steps") o.steps = atoi(next("--steps"));
        else if (a == "--blocks") o.blocks = atoi(next("--blocks"));
        else if (a == "--repeat") o.repeat = atoi(next("--repeat"));
        else if (a == "--mode") o.mode = next("--mode");
        else if (a == "--kernels") o.kernelFile = next("--kernels");
        else if (a == "--geometry") o.geometry = next("--geometry");
        else if (a == "--dump") o.dump = next("--dump");
        else if (a == "--seed") o.seed = strtoull(next("--seed"), nullptr, 0);
        else if (a == "--courant") o.courant = (float)atof(next("--courant"));
        else if (a == "--threads") o.threads = atoi(next("--threads"));
        else if (a == "--single-threaded") o.singleThreaded = true;
        else if (a == "--eager") o.eager = true;
        else if (a == "--profile") o.profile = true;
        else {
            fprintf(stderr, "unknown option %s\n", a.c_str());
            return false;
        }
    }
    return !o.kernelFile.empty() && o.nx > 0 && o.ny > 0 && o.nz > 0 && o.steps > 0;
}

uint64_t Fnv1a(const void *p, size_t n) {
    const unsigned char *b = (const unsigned char *)p;
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < n; ++i) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}

struct FdtdGraph {
    Options o;
    bool hip = false;
    Graph *g = nullptr;
    DatablockTemplate *field[6] = {};
    DatablockTemplate *params = nullptr;
    DatablockTemplate *iters = nullptr;
    std::vector<GraphInputChannel *> fieldInputs[6];  // every channel that takes an initial field
    std::vector<GraphInputChannel *> paramInputs;
    std::vector<GraphInputChannel *> iterInputs;
    GraphOutputChannel *outputs[6] = {};
    UINT uid = 0;

    void Geometry(Task *t, int f) {
        int d1, d2, d3;
        ptk_fdtd_extent(f, o.nx, o.ny, o.nz, &d1, &d2, &d3);
        if (!hip) {
            t->SetComputeGeometry(1, 1, 1);  // a host task computes the whole array per call
            return;
        }
        if (o.geometry == "reference") {
            // one thread per cell, as graphfdtd.cpp:439-446 sets for its CUDA kernels
            t->SetBlockAndGridSize(PTASKDIM3((d1 + 31) / 32, d2 * d3, 1), PTASKDIM3(32, 1, 1));
        } else if (o.geometry == "compute") {
            t->SetComputeGeometry(d1, d2, d3);
        } else {
            // column walk: one lane per (i,j) column, k split in slabs of ~8 planes
            int gx = (d1 * d2 + 255) / 256, slabs = (d3 + 7) / 8;
            t->SetBlockAndGridSize(PTASKDIM3(gx, slabs, 1), PTASKDIM3(256, 1, 1));
        }
    }

    // One Yee update task with ports (own, in1, in2, Config) -> out (+ a meta port if looped).
    Task *AddUpdate(int t, bool looped, Port **ins, Port **outs, int step) {
        CompiledKernel *k = Runtime::GetCompiledKernel((char *)o.kernelFile.c_str(), (char *)kKernelNames[t]);
        if (!k) return nullptr;
        const int own = kTaskInputs[t][0];
        for (int a = 0; a < 3; ++a) {
            int f = kTaskInputs[t][a];
            ins[a] = Runtime::CreatePort(INPUT_PORT, field[f], uid++, (char *)kFieldNames[f], a);
        }
        ins[3] = Runtime::CreatePort(STICKY_PORT, params, uid++, (char *)"Config", 3);
        UINT nIn = 4;
        if (looped) ins[nIn++] = Runtime::CreatePort(META_PORT, iters, uid++, (char *)"iterations");
        outs[0] = Runtime::CreatePort(OUTPUT_PORT, field[own], uid++, (char *)kFieldNames[own], 4);
        if (looped) outs[0]->SetMetaFunction(MF_NONE), ins[4]->SetMetaFunction(MF_GENERAL_ITERATOR);
        std::string name = std::string(kFieldNames[own]) + "Comp" + (step >= 0 ? "_" + std::to_string(step) : "");
        Task *task = g->AddTask(k, nIn, ins, 1, outs, name.c_str());
        if (task) Geometry(task, own);
        return task;
    }

    bool Build() {
        hip = ptaskutils::SelectAcceleratorClass(o.kernelFile.c_str()) == ACCELERATOR_CLASS_HIP;
        for (int f = 0; f < 6; ++f) {
            int d1, d2, d3;
            ptk_fdtd_extent(f, o.nx, o.ny, o.nz, &d1, &d2, &d3);
            std::string nm = std::string(kFieldNames[f]) + "Size";
            field[f] = Runtime::GetDatablockTemplate((char *)nm.c_str(), sizeof(float), d1, d2, d3);
        }
        params = Runtime::GetDatablockTemplate((char *)"fdtdparms", sizeof(FDTD_PARAMS), PTPARM_BYVALSTRUCT);
        iters = Runtime::GetDatablockTemplate((char *)"fdtditers", sizeof(int), PTPARM_INT);
        g = new Graph("fdtd");
        return o.mode == "looped" ? BuildLooped() : BuildUnrolled();
    }

    bool BuildLooped() {
        Port *ins[6][5], *outs[6][1];
        for (int t = 0; t < 6; ++t)
            if (!AddUpdate(t, true, ins[t], outs[t], -1)) return false;
        for (int t = 0; t < 6; ++t) {
            const int own = kTaskInputs[t][0];
            // first iteration: every field input from the application (switch channels win)
            for (int a = 0; a < 3; ++a) {
                int f = kTaskInputs[t][a];
                std::string nm = std::string(kFieldNames[t]) + "Comp." + kFieldNames[f] + ".in";
                if (t < 3 || a == 0) fieldInputs[f].push_back(g->AddInputChannel(ins[t][a], nm.c_str(), TRUE));
            }
            paramInputs.push_back(g->AddInputChannel(ins[t][3], (std::string(kFieldNames[own]) + ".cfg").c_str()));
            iterInputs.push_back(g->AddInputChannel(ins[t][4], (std::string(kFieldNames[own]) + ".iters").c_str()));
            g->BindControlPropagationPort(ins[t][4], outs[t][0]);
            // loop-carried: own field back to itself
            InternalChannel *back = g->AddInternalChannel(outs[t][0], ins[t][0], (std::string(kFieldNames[own]) + ".back").c_str());
            back->SetPredicationType(CE_SRC, CGATEFN_CLOSE_ON_ENDITERATION);
            outputs[own] = g->AddOutputChannel(outs[t][0], (std::string(kFieldNames[own]) + ".out").c_str());
            outputs[own]->SetPredicationType(CE_SRC, CGATEFN_OPEN_ON_ENDITERATION);
            if (o.eager) outputs[own]->SetViewMaterializationPolicy(VIEWMATERIALIZATIONPOLICY_EAGER);
        }
        // H -> E within a step (always open); E -> H into the next step (closed at the end)
        for (int t = 3; t < 6; ++t)
            for (int a = 1; a < 3; ++a) {
                int f = kTaskInputs[t][a];
                g->AddInternalChannel(outs[f][0], ins[t][a], (std::string(kFieldNames[f]) + "->" + kKernelNames[t]).c_str());
            }
        for (int t = 0; t < 3; ++t)
            for (int a = 1; a < 3; ++a) {
                int f = kTaskInputs[t][a];
                InternalChannel *c = g->AddInternalChannel(outs[f][0], ins[t][a], (std::string(kFieldNames[f]) + "->" + kKernelNames[t]).c_str());
                c->SetPredicationType(CE_SRC, CGATEFN_CLOSE_ON_ENDITERATION);
            }
        return true;
    }

    bool BuildUnrolled() {
        Port *prevOut[6] = {};
        for (int s = 0; s < o.steps; ++s) {
            Port *ins[6][5], *outs[6][1];
            for (int t = 0; t < 6; ++t)
                if (!AddUpdate(t, false, ins[t], outs[t], s)) return false;
            std::string u = "_" + std::to_string(s);
            for (int t = 0; t < 6; ++t) {
                paramInputs.push_back(g->AddInputChannel(ins[t][3], (std::string(kKernelNames[t]) + ".cfg" + u).c_str()));
                for (int a = 0; a < 3; ++a) {
                    int f = kTaskInputs[t][a];
                    bool fromThisStep = t >= 3 && a > 0;  // E tasks read the new H
                    std::string nm = std::string(kKernelNames[t]) + "." + kFieldNames[f] + u;
                    if (fromThisStep) g->AddInternalChannel(outs[f][0], ins[t][a], nm.c_str());
                    else if (s == 0) fieldInputs[f].push_back(g->AddInputChannel(ins[t][a], nm.c_str()));
                    else g->AddInternalChannel(prevOut[f], ins[t][a], nm.c_str());
                }
            }
            for (int t = 0; t < 6; ++t) prevOut[kTaskInputs[t][0]] = outs[t][0];
        }
        for (int f = 0; f < 6; ++f) {
            outputs[f] = g->AddOutputChannel(prevOut[f], (std::string(kFieldNames[f]) + ".out").c_str());
            if (o.eager) outputs[f]->SetViewMaterializationPolicy(VIEWMATERIALIZATIONPOLICY_EAGER);
        }
        for (GraphInputChannel *c : paramInputs) c->SetNoDraw();
        return true;
    }
};

}  // namespace

int main(int argc, char **argv) {
    Options o;
    if (!Parse(argc, argv, o)) {
        fprintf(stderr, "usage: ptk_fdtd --kernels <fdtd.hsaco|hosttasks.so> [--n N | --nx --ny --nz] "
                        "[--steps S] [--mode looped|unrolled] [--blocks B] [--repeat R] [--seed S] "
                        "[--geometry default|reference|compute] [--dump file] [--eager] [--profile]\n");
        return 2;
    }
    Runtime::SetExitOnRuntimeFailure(FALSE);
    if (o.threads) Runtime::SetTaskThreadPoolSize((UINT)o.threads);
    if (o.profile) Runtime::SetProfilePSDispatch(TRUE);
    if (o.mode == "unrolled") Runtime::SetDefaultChannelCapacity(std::max(4, o.blocks));
    if (PTFAILED(Runtime::Initialize())) {
        fprintf(stderr, "Runtime::Initialize failed\n");
        return 1;
    }
    FdtdGraph G;
    G.o = o;
    if (!G.Build()) {
        fprintf(stderr, "graph construction failed: %s\n", Runtime::GetLastErrorText());
        return 1;
    }
    // seeded initial fields (element n of field f = ptk_rand_sym_f32(seed, f, n))
    std::vector<std::vector<float>> init(6);
    for (int f = 0; f < 6; ++f) {
        init[f].resize((size_t)ptk_fdtd_field_elems(f, o.nx, o.ny, o.nz));
        for (size_t n = 0; n < init[f].size(); ++n) init[f][n] = ptk_rand_sym_f32(o.seed, (uint64_t)f, n);
    }
    FDTD_PARAMS prm = ptk_fdtd_make_params(o.nx, o.ny, o.nz, o.courant);
    G.g->Run(o.singleThreaded ? TRUE : FALSE);

    Datablock *cfg = Runtime::AllocateDatablock(G.params, &prm, sizeof(prm), nullptr);
    for (GraphInputChannel *c : G.paramInputs) c->Push(cfg);
    cfg->Release();

    CHighResolutionTimer timer(gran_msec);
    std::vector<double> runMs;
    unsigned long long h2d0 = 0, d2h0 = 0, h2db = 0, d2hb = 0;
    std::vector<float> final[6];
    const int runs = o.mode == "looped" ? o.repeat : 1;
    const int blocks = o.mode == "looped" ? 1 : o.blocks;
    unsigned long long dispatch0 = 0;
    bool ok = true;
    for (int r = 0; r < runs && ok; ++r) {
        Runtime::ResetTransferStatistics();
        dispatch0 = G.g->GetDispatchCount();
        timer.reset();
        for (int b = 0; b < blocks; ++b) {
            for (int f = 0; f < 6; ++f) {
                Datablock *db = Runtime::AllocateDatablock(G.field[f], init[f].data(),
                                                           (UINT)(init[f].size() * sizeof(float)), nullptr);
                for (GraphInputChannel *c : G.fieldInputs[f]) c->Push(db);
                db->Release();
            }
            if (o.mode == "looped") {
                Datablock *it = Runtime::AllocateDatablock(G.iters, &o.steps, sizeof(int), nullptr);
                for (GraphInputChannel *c : G.iterInputs) c->Push(it);
                it->Release();
            }
        }
        for (int b = 0; b < blocks && ok; ++b) {
            for (int f = 0; f < 6; ++f) {
                Datablock *res = G.outputs[f]->Pull(600000);
                if (!res) {
                    fprintf(stderr, "no output on %s: %s\n", kFieldNames[f], Runtime::GetLastErrorText());
                    ok = false;
                    break;
                }
                res->Lock();
                const float *p = (const float *)res->GetDataPointer(FALSE);
                final[f].assign(p, p + init[f].size());
                res->Unlock();
                res->Release();
            }
        }
        runMs.push_back(timer.elapsed(false));
        Runtime::GetTransferStatistics(&h2d0, &h2db, &d2h0, &d2hb);
    }
    unsigned long long dispatches = G.g->GetDispatchCount() - dispatch0;
    double devMs = 0;
    if (o.profile)
        for (UINT t = 0; t < G.g->GetTaskCount(); ++t) devMs += G.g->GetTaskByIndex(t)->GetTotalDispatchDeviceMs();
    G.g->Stop();
    G.g->Teardown();
    Graph::DestroyGraph(G.g);

    if (ok && !o.dump.empty()) {
        FILE *fp = fopen(o.dump.c_str(), "wb");
        for (int f = 0; fp && f < 6; ++f) fwrite(final[f].data(), sizeof(float), final[f].size(), fp);
        if (fp) fclose(fp);
    }
    double best = runMs.empty() ? 0 : runMs[0];
    for (double m : runMs) best = std::min(best, m);
    const double cells = (double)o.nx * o.ny * o.nz;
    printf("{\"ok\": %s, \"mode\": \"%s\", \"backend\": \"%s\", \"grid\": [%d, %d, %d], \"steps\": %d, "
           "\"blocks\": %d, \"runs\": %d, \"best_ms\": %.4f, \"last_ms\": %.4f, "
           "\"mcups\": %.3f, \"dispatches_last_run\": %llu, \"h2d_count\": %llu, \"h2d_bytes\": %llu, "
           "\"d2h_count\": %llu, \"d2h_bytes\": %llu, \"device_ms_profiled\": %.4f, \"checksums\": [",
           ok ? "true" : "false", o.mode.c_str(), G.hip ? "hip" : "host", o.nx, o.ny, o.nz, o.steps,
           blocks, (int)runMs.size(), best, runMs.empty() ? 0.0 : runMs.back(),
           best > 0 ? cells * o.steps * blocks / (best / 1e3) / 1e6 : 0.0, dispatches, h2d0, h2db,
           d2h0, d2hb, devMs);
    for (int f = 0; f < 6; ++f) {
        double s = 0;
        for (float v : final[f]) s += v;
        printf("%s{\"field\": \"%s\", \"fnv1a\": \"%016llx\", \"sum\": %.17g}", f ? ", " : "", kFieldNames[f],
               (unsigned long long)Fnv1a(final[f].data(), final[f].size() * sizeof(float)), s);
    }
    printf("]}\n");
    fflush(stdout);
    Runtime::Terminate();
    return ok ? 0 : 1;
}
