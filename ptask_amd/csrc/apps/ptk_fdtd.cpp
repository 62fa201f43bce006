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

// A graph form: the tasks of one time step, in order, with the fields each reads and writes.
// Formal parameters: the input fields (0..n-1), the FDTD_PARAMS block (n), the outputs (n+1..).
struct TaskSpec {
    const char *kernel;
    std::vector<int> in;
    std::vector<int> out;
};

std::vector<TaskSpec> FormTasks(const std::string &form) {
    if (form == "halves")  // all H updates, then all E updates (72 B per cell-step)
        return {{"FdtdHFields", {0, 1, 2, 3, 4, 5}, {0, 1, 2}},
                {"FdtdEFields", {3, 4, 5, 0, 1, 2}, {3, 4, 5}}};
    if (form == "fused")  // one whole step (48 B per cell-step)
        return {{"FdtdStepFused", {0, 1, 2, 3, 4, 5}, {0, 1, 2, 3, 4, 5}}};
    if (form == "overlap")  // one whole step on overlapped tiles (48 B per cell-step + overlap)
        return {{"FdtdStepOverlap", {0, 1, 2, 3, 4, 5}, {0, 1, 2, 3, 4, 5}}};
    if (form == "overlap2")  // two whole steps per task (24 B per cell-step + overlap)
        return {{"FdtdTwoStepOverlap", {0, 1, 2, 3, 4, 5}, {0, 1, 2, 3, 4, 5}}};
    if (form == "overlap3")  // three whole steps per task (16 B per cell-step + overlap)
        return {{"FdtdThreeStepOverlap", {0, 1, 2, 3, 4, 5}, {0, 1, 2, 3, 4, 5}}};
    if (form == "overlap3p")  // the same on a persistent grid of one block per CU
        return {{"FdtdThreeStepPersistent", {0, 1, 2, 3, 4, 5}, {0, 1, 2, 3, 4, 5}}};
    // the reference's six tasks (graphfdtd.cpp:448-560): (own, in1, in2) -> own
    return {{"HxComputation", {0, 4, 5}, {0}}, {"HyComputation", {1, 3, 5}, {1}},
            {"HzComputation", {2, 3, 4}, {2}}, {"ExComputation", {3, 1, 2}, {3}},
            {"EyComputation", {4, 0, 2}, {4}}, {"EzComputation", {5, 0, 1}, {5}}};
}

// Time steps one task of the form advances the fields by.
int StepsPerTask(const std::string &form) {
    return form == "overlap2" ? 2 : form == "overlap3" || form == "overlap3p" ? 3 : 1;
}

struct Options {
    int nx = 32, ny = 32, nz = 32, steps = 10, blocks = 1, repeat = 1;
    std::string mode = "looped";
    std::string form = "six";
    std::string kernelFile;
    std::string geometry = "default";
    int tileI = 0, tileJ = 0, slab = 0;  // tile shape and k-slab depth (0: the form's default)
    std::string dump;
    unsigned long long seed = 0x5EED;
    float courant = 0.5f;
    int threads = 0;
    int hostConcurrency = 0;  // host tasks running at once (0: the runtime default)
    bool singleThreaded = false;
    bool eager = false;
    bool profile = false;
    bool chain = false;  // run r+1 starts from run r's device-resident outputs
    int warmup = 1;      // chained runs before the timed ones (the first one uploads)
    std::string streams = "task";  // HIP stream policy: one per task, or one shared
    std::string serialize;    // write the graph as XML (and stop, unless it was read from XML)
    std::string deserialize;  // read the graph from XML instead of building it
    std::string pool;         // input blocks from a runtime-wide ("global") or graph ("graph") pool
    int poolBlocks = 6;       // blocks per field pool
    int injectHost = 0, injectDevice = 0;  // allocation failures forced before every run after the first
    bool capture = false;     // record the first run's launches and replay them as one hipGraph
    int seedStep = 0;         // run r starts from seed + r * seedStep (the dump is the last run's)
    int graphs = 1;           // independent graphs run at once (graph g: seed + g, priority g)
    std::string priorities;   // comma-separated task priority per graph (default: all equal)
    std::string sched = "priority";  // dispatch admission across graphs: priority | fifo
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
        else if (a == "--form") o.form = next("--form");
        else if (a == "--kernels") o.kernelFile = next("--kernels");
        else if (a == "--geometry") o.geometry = next("--geometry");
        else if (a == "--tile") {
            if (sscanf(next("--tile"), "%dx%dx%d", &o.tileI, &o.tileJ, &o.slab) != 3) return false;
        } else if (a == "--dump") o.dump = next("--dump");
        else if (a == "--seed") o.seed = strtoull(next("--seed"), nullptr, 0);
        else if (a == "--courant") o.courant = (float)atof(next("--courant"));
        else if (a == "--threads") o.threads = atoi(next("--threads"));
        else if (a == "--host-concurrency") o.hostConcurrency = atoi(next("--host-concurrency"));
        else if (a == "--single-threaded") o.singleThreaded = true;
        else if (a == "--eager") o.eager = true;
        else if (a == "--profile") o.profile = true;
        else if (a == "--chain") o.chain = true;
        else if (a == "--warmup") o.warmup = atoi(next("--warmup"));
        else if (a == "--streams") o.streams = next("--streams");
        else if (a == "--serialize") o.serialize = next("--serialize");
        else if (a == "--deserialize") o.deserialize = next("--deserialize");
        else if (a == "--pool") o.pool = next("--pool");
        else if (a == "--pool-blocks") o.poolBlocks = atoi(next("--pool-blocks"));
        else if (a == "--inject-host-failures") o.injectHost = atoi(next("--inject-host-failures"));
        else if (a == "--inject-device-failures") o.injectDevice = atoi(next("--inject-device-failures"));
        else if (a == "--graphs") o.graphs = atoi(next("--graphs"));
        else if (a == "--capture") o.capture = true;
        else if (a == "--seed-step") o.seedStep = atoi(next("--seed-step"));
        else if (a == "--priorities") o.priorities = next("--priorities");
        else if (a == "--sched") o.sched = next("--sched");
        else {
            fprintf(stderr, "unknown option %s\n", a.c_str());
            return false;
        }
    }
    if (o.tileI && (o.tileI < 1 || o.tileJ < 2 || o.slab < 1 || o.tileI * o.tileJ > 1024)) return false;
    if (o.form == "overlap" && o.tileI && (o.tileI != 64 || o.tileJ < 3 || o.tileJ > 16)) return false;
    if (o.form == "overlap2" || o.form == "overlap3" || o.form == "overlap3p") {
        const int T = StepsPerTask(o.form);
        if (o.steps < T) {  // the remainder of steps % T runs in one task after the T-step tasks
            fprintf(stderr, "--form %s advances %d steps per task: --steps must be at least %d\n",
                    o.form.c_str(), T, T);
            return false;
        }
        if (o.tileI && (o.tileI != 64 || o.tileJ < 2 * T + 1 || o.tileJ > 16)) return false;
    }
    // the overlapped-tile kernel addresses each field with 32-bit buffer offsets
    if (o.form.compare(0, 7, "overlap") == 0 && (long long)(o.nx + 1) * (o.ny + 1) * (o.nz + 1) >= (1ll << 29)) {
        fprintf(stderr, "the overlap forms need (nx+1)(ny+1)(nz+1) < 2^29\n");
        return false;
    }
    return (!o.kernelFile.empty() || !o.deserialize.empty()) && o.nx > 0 && o.ny > 0 && o.nz > 0 && o.steps > 0 &&
           o.graphs >= 1 && o.graphs <= 64 && (o.pool.empty() || o.pool == "global" || o.pool == "graph") &&
           o.poolBlocks >= 1;
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
    std::vector<TaskSpec> spec;
    Graph *g = nullptr;
    DatablockTemplate *field[6] = {};
    DatablockTemplate *params = nullptr;
    DatablockTemplate *iters = nullptr;
    std::vector<GraphInputChannel *> fieldInputs[6];  // every channel that takes an initial field
    std::vector<GraphInputChannel *> paramInputs;
    std::vector<GraphInputChannel *> iterInputs;
    GraphOutputChannel *outputs[6] = {};
    UINT uid = 0;

    // Compute units of the first HIP accelerator (256 on an MI355X if the runtime cannot say).
    static int DeviceCores() {
        Accelerator *a = Runtime::GetAccelerator(ACCELERATOR_CLASS_HIP, 0);
        if (!a) return 256;
        ACCELERATOR_DESCRIPTOR d;
        a->GetDescriptor(&d);
        return d.nCoreCount ? (int)d.nCoreCount : 256;
    }

    void Geometry(Task *t, const TaskSpec &ts) {
        if (!hip) {
            t->SetComputeGeometry(1, 1, 1);  // a host task computes its whole arrays per call
            return;
        }
        // --tile applies to the form's own task; a remainder task keeps its kernel's defaults
        const bool tiled = o.tileI && !strcmp(ts.kernel, spec[0].kernel);
        if (!strcmp(ts.kernel, "FdtdThreeStepPersistent")) {
            // 64 x R thread blocks on a persistent grid of G blocks (--tile 64xRxG; default one
            // per compute unit of the device, whose LDS holds one such block)
            const int R = tiled ? o.tileJ : 16, G = tiled ? o.slab : DeviceCores();
            t->SetBlockAndGridSize(PTASKDIM3(G, 1, 1), PTASKDIM3(64, R, 1));
            return;
        }
        if (!strcmp(ts.kernel, "FdtdTwoStepOverlap") || !strcmp(ts.kernel, "FdtdThreeStepOverlap")) {
            // 64 x R thread blocks owning (64-2T) x (R-2T) columns each, k slabs of `slab` planes
            // (deeper slabs for three steps, whose pipeline runs 4 planes past each slab end)
            const int T = !strcmp(ts.kernel, "FdtdThreeStepOverlap") ? 3 : 2, ow = 64 - 2 * T;
            const int R = tiled ? o.tileJ : 16, slab = tiled ? o.slab : T == 3 ? 64 : 16, oh = R - 2 * T;
            t->SetBlockAndGridSize(PTASKDIM3((o.nx + ow) / ow, (o.ny + oh) / oh, (o.nz + slab) / slab),
                                   PTASKDIM3(64, R, 1));
            return;
        }
        if (!strncmp(ts.kernel, "FdtdStepOverlap", 15)) {
            // 64 x R thread blocks owning 62 x (R-2) columns each, k slabs of `slab` planes
            const int R = tiled ? o.tileJ : 8, slab = tiled ? o.slab : 8;
            t->SetBlockAndGridSize(PTASKDIM3((o.nx + 62) / 62, (o.ny + R - 2) / (R - 2), (o.nz + slab) / slab),
                                   PTASKDIM3(64, R, 1));
            return;
        }
        if (ts.out.size() == 6) {  // whole-step kernel: 64x8 tiles, slabs of 16 planes
            t->SetBlockAndGridSize(PTASKDIM3((o.nx + 64) / 64, (o.ny + 8) / 8, (o.nz + 16) / 16),
                                   PTASKDIM3(64, 8, 1));
            return;
        }
        int d1, d2, d3;
        if (ts.out.size() == 3) {  // half steps walk the union domain
            d1 = o.nx + 1, d2 = o.ny + 1, d3 = o.nz + 1;
        } else {
            ptk_fdtd_extent(ts.out[0], o.nx, o.ny, o.nz, &d1, &d2, &d3);
        }
        if (o.geometry == "reference") {
            // one thread per cell, as graphfdtd.cpp:439-446 sets for its CUDA kernels
            t->SetBlockAndGridSize(PTASKDIM3((d1 + 31) / 32, d2 * d3, 1), PTASKDIM3(32, 1, 1));
        } else if (o.geometry == "compute") {
            t->SetComputeGeometry(d1, d2, d3);
        } else if (ts.out.size() == 3 && o.tileI) {
            // half steps with an explicit tileI x tileJ block and k slabs of ~slab planes
            t->SetBlockAndGridSize(PTASKDIM3((d1 + o.tileI - 1) / o.tileI, (d2 + o.tileJ - 1) / o.tileJ,
                                             (d3 + o.slab - 1) / o.slab),
                                   PTASKDIM3(o.tileI, o.tileJ, 1));
        } else {
            // column walk: one lane per (i,j) column, k split in slabs of ~8 planes
            const int bx = ts.out.size() == 3 ? 512 : 256;
            int gx = (d1 * d2 + bx - 1) / bx, slabs = (d3 + 7) / 8;
            t->SetBlockAndGridSize(PTASKDIM3(gx, slabs, 1), PTASKDIM3(bx, 1, 1));
        }
    }

    // One task of the form: input ports (fields), the sticky Config port, an iteration meta port
    // when looped, output ports.
    Task *AddTask(const TaskSpec &ts, bool looped, std::vector<Port *> &ins, std::vector<Port *> &outs,
                  Port *&meta, int step) {
        CompiledKernel *k = Runtime::GetCompiledKernel((char *)o.kernelFile.c_str(), (char *)ts.kernel);
        if (!k) return nullptr;
        const UINT n = (UINT)ts.in.size();
        ins.clear();
        outs.clear();
        for (UINT a = 0; a < n; ++a)
            ins.push_back(Runtime::CreatePort(INPUT_PORT, field[ts.in[a]], uid++, (char *)kFieldNames[ts.in[a]], a));
        ins.push_back(Runtime::CreatePort(STICKY_PORT, params, uid++, (char *)"Config", n));
        meta = nullptr;
        if (looped) {
            meta = Runtime::CreatePort(META_PORT, iters, uid++, (char *)"iterations");
            meta->SetMetaFunction(MF_GENERAL_ITERATOR);
            ins.push_back(meta);
        }
        // A field block goes to tasks that only read it and to this application, which only reads
        // what it pulls: no consumer writes it in place, so the consumers share it uncopied.
        for (size_t b = 0; b < ts.out.size(); ++b) {
            outs.push_back(Runtime::CreatePort(OUTPUT_PORT, field[ts.out[b]], uid++,
                                               (char *)kFieldNames[ts.out[b]], n + 1 + (UINT)b));
            outs.back()->SetSuppressClones(TRUE);
        }
        std::string name = std::string(ts.kernel) + (step >= 0 ? "_" + std::to_string(step) : step == -2 ? "_tail" : "");
        Task *task = g->AddTask(k, (UINT)ins.size(), ins.data(), (UINT)outs.size(), outs.data(), name.c_str());
        if (task) Geometry(task, ts);
        return task;
    }

    bool Build() {
        hip = ptaskutils::SelectAcceleratorClass(o.kernelFile.c_str()) == ACCELERATOR_CLASS_HIP;
        spec = FormTasks(o.form);
        if (o.form == "overlap" && o.tileJ <= 8) spec[0].kernel = "FdtdStepOverlap8";  // half the LDS
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

    // The graph from an XML file written by Graph::Serialize.  Its boundary is found by type and
    // template: graph inputs carrying a field, the Config parameters or the iteration count, and
    // one graph output per field.
    bool Load() {
        g = new Graph("fdtd");
        if (!g->Deserialize(o.deserialize.c_str())) return false;
        hip = g->GetTaskCount() && g->GetTaskByIndex(0)->GetAcceleratorClass() == ACCELERATOR_CLASS_HIP;
        for (int f = 0; f < 6; ++f) field[f] = Runtime::GetDatablockTemplate((char *)(std::string(kFieldNames[f]) + "Size").c_str());
        params = Runtime::GetDatablockTemplate((char *)"fdtdparms");
        iters = Runtime::GetDatablockTemplate((char *)"fdtditers");
        for (UINT i = 0; i < g->GetChannelCount(); ++i) {
            Channel *c = g->GetChannelByIndex(i);
            DatablockTemplate *t = c->GetTemplate();
            if (c->GetType() == CT_GRAPH_INPUT) {
                if (t == params) paramInputs.push_back((GraphInputChannel *)c);
                else if (t == iters) iterInputs.push_back((GraphInputChannel *)c);
                for (int f = 0; f < 6; ++f)
                    if (t == field[f]) fieldInputs[f].push_back((GraphInputChannel *)c);
            } else if (c->GetType() == CT_GRAPH_OUTPUT) {
                for (int f = 0; f < 6; ++f)
                    if (t == field[f]) {
                        if (outputs[f]) return Runtime::HandleError("two graph outputs carry %s", kFieldNames[f]);
                        outputs[f] = (GraphOutputChannel *)c;
                        // as when built here: the consumers of a field only read it
                        if (Port *src = c->GetBoundPort(CE_SRC)) src->SetSuppressClones(TRUE);
                    }
            }
        }
        for (int f = 0; f < 6; ++f) {
            if (!field[f] || fieldInputs[f].empty() || !outputs[f])
                return Runtime::HandleError("%s has no graph input or output for %s", o.deserialize.c_str(), kFieldNames[f]);
            // the field templates fix the grid the file was written for
            int d1, d2, d3;
            ptk_fdtd_extent(f, o.nx, o.ny, o.nz, &d1, &d2, &d3);
            if (field[f]->GetXElementCount() != (UINT)d1 || field[f]->GetYElementCount() != (UINT)d2 ||
                field[f]->GetZElementCount() != (UINT)d3)
                return Runtime::HandleError("%s was written for another grid than %dx%dx%d", o.deserialize.c_str(), o.nx, o.ny, o.nz);
        }
        if (!params || paramInputs.empty() || (o.mode == "looped" && iterInputs.empty()))
            return Runtime::HandleError("%s lacks the Config or iteration inputs", o.deserialize.c_str());
        return true;
    }

    // producer[f] = (task index, output slot) of field f within one step
    void Producers(int prodTask[6], int prodSlot[6]) {
        for (size_t t = 0; t < spec.size(); ++t)
            for (size_t b = 0; b < spec[t].out.size(); ++b) {
                prodTask[spec[t].out[b]] = (int)t;
                prodSlot[spec[t].out[b]] = (int)b;
            }
    }

    // The remainder of the steps when the form's task advances several (steps % T): one task
    // after the loop, or after the unrolled tasks, advancing the last one or two steps.
    bool TailSpec(TaskSpec &ts) const {
        const int r = o.steps % StepsPerTask(o.form);
        if (!r) return false;
        ts = {r == 2 ? "FdtdTwoStepOverlap" : "FdtdStepOverlap", {0, 1, 2, 3, 4, 5}, {0, 1, 2, 3, 4, 5}};
        return true;
    }

    // A remainder task fed by `src[f]`, the producers of the fields after the last full task; its
    // outputs are the graph's.  `endOfLoop`: the feeding channels open on the loop's last
    // iteration only.
    bool AddTail(const TaskSpec &ts, Port *const src[6], bool endOfLoop) {
        std::vector<Port *> ins, outs;
        Port *meta = nullptr;
        if (!AddTask(ts, false, ins, outs, meta, -2)) return false;
        for (size_t a = 0; a < ts.in.size(); ++a) {
            const int f = ts.in[a];
            InternalChannel *c = g->AddInternalChannel(src[f], ins[a], (std::string(kFieldNames[f]) + "->tail").c_str());
            if (endOfLoop) c->SetPredicationType(CE_SRC, CGATEFN_OPEN_ON_ENDITERATION);
        }
        paramInputs.push_back(g->AddInputChannel(ins[ts.in.size()], "tail.cfg"));
        for (size_t b = 0; b < ts.out.size(); ++b) {
            const int f = ts.out[b];
            outputs[f] = g->AddOutputChannel(outs[b], (std::string(kFieldNames[f]) + ".out").c_str());
            if (o.eager) outputs[f]->SetViewMaterializationPolicy(VIEWMATERIALIZATIONPOLICY_EAGER);
        }
        return true;
    }

    // Looped: every task carries a MF_GENERAL_ITERATOR meta port.  A field produced earlier in
    // the step arrives on an always-open channel; a field produced by the same or a later task
    // is loop-carried: the application's switch channel feeds the first iteration, a
    // back-channel closed on ENDITERATION feeds the rest (iteration.cpp:887-1060).
    bool BuildLooped() {
        const size_t T = spec.size();
        TaskSpec tailSpec;
        const bool tail = TailSpec(tailSpec);
        std::vector<std::vector<Port *>> ins(T), outs(T);
        std::vector<Port *> metas(T);
        for (size_t t = 0; t < T; ++t)
            if (!AddTask(spec[t], true, ins[t], outs[t], metas[t], -1)) return false;
        int prodTask[6], prodSlot[6];
        Producers(prodTask, prodSlot);
        for (size_t t = 0; t < T; ++t) {
            std::string tn = spec[t].kernel;
            for (size_t a = 0; a < spec[t].in.size(); ++a) {
                const int f = spec[t].in[a];
                const int pt = prodTask[f];
                Port *src = outs[pt][prodSlot[f]];
                std::string nm = std::string(kFieldNames[f]) + "->" + tn + "." + std::to_string(a);
                if (pt < (int)t) {
                    g->AddInternalChannel(src, ins[t][a], nm.c_str());
                } else {
                    fieldInputs[f].push_back(g->AddInputChannel(ins[t][a], (nm + ".init").c_str(), TRUE));
                    InternalChannel *c = g->AddInternalChannel(src, ins[t][a], (nm + ".back").c_str());
                    c->SetPredicationType(CE_SRC, CGATEFN_CLOSE_ON_ENDITERATION);
                }
            }
            paramInputs.push_back(g->AddInputChannel(ins[t][spec[t].in.size()], (tn + ".cfg").c_str()));
            iterInputs.push_back(g->AddInputChannel(metas[t], (tn + ".iters").c_str()));
            for (size_t b = 0; b < spec[t].out.size(); ++b) {
                g->BindControlPropagationPort(metas[t], outs[t][b]);
                if (tail) continue;  // the remainder task takes the loop's last outputs
                const int f = spec[t].out[b];
                outputs[f] = g->AddOutputChannel(outs[t][b], (std::string(kFieldNames[f]) + ".out").c_str());
                outputs[f]->SetPredicationType(CE_SRC, CGATEFN_OPEN_ON_ENDITERATION);
                if (o.eager) outputs[f]->SetViewMaterializationPolicy(VIEWMATERIALIZATIONPOLICY_EAGER);
            }
        }
        if (tail) {
            Port *src[6];
            for (int f = 0; f < 6; ++f) src[f] = outs[prodTask[f]][prodSlot[f]];
            return AddTail(tailSpec, src, true);
        }
        return true;
    }

    // Unrolled: the tasks of every step are stamped out (graphfdtd.cpp:735-739); a field comes
    // from this step's producer if it ran earlier in the step, else from the previous step.
    bool BuildUnrolled() {
        const size_t T = spec.size();
        int prodTask[6], prodSlot[6];
        Producers(prodTask, prodSlot);
        std::vector<std::vector<Port *>> prevOuts;
        for (int s = 0; s < o.steps / StepsPerTask(o.form); ++s) {
            std::vector<std::vector<Port *>> ins(T), outs(T);
            Port *meta = nullptr;
            for (size_t t = 0; t < T; ++t)
                if (!AddTask(spec[t], false, ins[t], outs[t], meta, s)) return false;
            const std::string u = "_" + std::to_string(s);
            for (size_t t = 0; t < T; ++t) {
                std::string tn = spec[t].kernel;
                paramInputs.push_back(g->AddInputChannel(ins[t][spec[t].in.size()], (tn + ".cfg" + u).c_str()));
                for (size_t a = 0; a < spec[t].in.size(); ++a) {
                    const int f = spec[t].in[a];
                    const int pt = prodTask[f];
                    std::string nm = tn + "." + std::to_string(a) + "." + kFieldNames[f] + u;
                    if (pt < (int)t) g->AddInternalChannel(outs[pt][prodSlot[f]], ins[t][a], nm.c_str());
                    else if (s == 0) fieldInputs[f].push_back(g->AddInputChannel(ins[t][a], nm.c_str()));
                    else g->AddInternalChannel(prevOuts[pt][prodSlot[f]], ins[t][a], nm.c_str());
                }
            }
            prevOuts = outs;
        }
        TaskSpec tailSpec;
        if (TailSpec(tailSpec)) {
            Port *src[6];
            for (int f = 0; f < 6; ++f) src[f] = prevOuts[prodTask[f]][prodSlot[f]];
            if (!AddTail(tailSpec, src, false)) return false;
        } else {
            for (int f = 0; f < 6; ++f) {
                outputs[f] = g->AddOutputChannel(prevOuts[prodTask[f]][prodSlot[f]], (std::string(kFieldNames[f]) + ".out").c_str());
                if (o.eager) outputs[f]->SetViewMaterializationPolicy(VIEWMATERIALIZATIONPOLICY_EAGER);
            }
        }
        for (GraphInputChannel *c : paramInputs) c->SetNoDraw();
        return true;
    }
};

// --graphs G: G independent graphs of the same form live at once (the concurrent-graph scenario
// of rttests.cpp:936-1333 and the per-instance priorities of graphbench.cpp:752-845).  Graph g
// starts from seed + g and its tasks run at priority priorities[g]; every graph's inputs are
// pushed before any output is pulled, so their dispatches contend for the device together.
// Dumps go to <dump>.<g>.
int RunConcurrent(const Options &o) {
    std::vector<int> prio(o.graphs, 0);
    if (!o.priorities.empty()) {
        size_t i = 0;
        for (int g = 0; g < o.graphs; ++g) {
            size_t j = o.priorities.find(',', i);
            prio[g] = atoi(o.priorities.substr(i, j == std::string::npos ? std::string::npos : j - i).c_str());
            if (j == std::string::npos) {
                for (int h = g + 1; h < o.graphs; ++h) prio[h] = prio[g];
                break;
            }
            i = j + 1;
        }
    }
    std::vector<FdtdGraph> G(o.graphs);
    for (int g = 0; g < o.graphs; ++g) {
        G[g].o = o;
        G[g].o.seed = o.seed + (unsigned long long)g;
        if (!G[g].Build()) {
            fprintf(stderr, "graph %d construction failed: %s\n", g, Runtime::GetLastErrorText());
            return 1;
        }
        for (UINT t = 0; t < G[g].g->GetTaskCount(); ++t) G[g].g->GetTaskByIndex(t)->SetPriority(prio[g]);
    }
    FDTD_PARAMS prm = ptk_fdtd_make_params(o.nx, o.ny, o.nz, o.courant);
    for (auto &fg : G) {
        fg.g->Run(FALSE);
        Datablock *cfg = Runtime::AllocateDatablock(fg.params, &prm, sizeof(prm), nullptr);
        for (GraphInputChannel *c : fg.paramInputs) c->Push(cfg);
        cfg->Release();
    }
    std::vector<std::vector<std::vector<float>>> fin(o.graphs, std::vector<std::vector<float>>(6));
    std::vector<unsigned long long> disp0(o.graphs);
    CHighResolutionTimer timer(gran_msec);
    for (int g = 0; g < o.graphs; ++g) {
        FdtdGraph &fg = G[g];
        disp0[g] = fg.g->GetDispatchCount();
        for (int f = 0; f < 6; ++f) {
            std::vector<float> init((size_t)ptk_fdtd_field_elems(f, o.nx, o.ny, o.nz));
            for (size_t n = 0; n < init.size(); ++n) init[n] = ptk_rand_sym_f32(fg.o.seed, (uint64_t)f, n);
            fin[g][f].resize(init.size());
            Datablock *db = Runtime::AllocateDatablock(fg.field[f], init.data(), (UINT)(init.size() * sizeof(float)), nullptr);
            for (GraphInputChannel *c : fg.fieldInputs[f]) c->Push(db);
            db->Release();
        }
        if (o.mode == "looped") {
            int iters = o.steps / StepsPerTask(o.form);
            Datablock *it = Runtime::AllocateDatablock(fg.iters, &iters, sizeof(int), nullptr);
            for (GraphInputChannel *c : fg.iterInputs) c->Push(it);
            it->Release();
        }
    }
    bool ok = true;
    for (int g = 0; g < o.graphs && ok; ++g)
        for (int f = 0; f < 6; ++f) {
            Datablock *res = G[g].outputs[f]->Pull(600000);
            if (!res) {
                fprintf(stderr, "graph %d: no output on %s: %s\n", g, kFieldNames[f], Runtime::GetLastErrorText());
                ok = false;
                break;
            }
            res->Lock();
            const float *p = (const float *)res->GetDataPointer(FALSE);
            fin[g][f].assign(p, p + fin[g][f].size());
            res->Unlock();
            res->Release();
        }
    const double ms = timer.elapsed(false);
    std::string js;
    for (int g = 0; g < o.graphs; ++g) {
        FdtdGraph &fg = G[g];
        char buf[160];
        snprintf(buf, sizeof(buf), "%s{\"graph\": %d, \"seed\": %llu, \"priority\": %d, \"dispatches\": %llu, \"fnv1a\": [",
                 g ? ", " : "", g, fg.o.seed, prio[g], (unsigned long long)(fg.g->GetDispatchCount() - disp0[g]));
        js += buf;
        for (int f = 0; f < 6; ++f) {
            snprintf(buf, sizeof(buf), "%s\"%016llx\"", f ? ", " : "",
                     (unsigned long long)Fnv1a(fin[g][f].data(), fin[g][f].size() * sizeof(float)));
            js += buf;
        }
        js += "]}";
        fg.g->Stop();
        fg.g->Teardown();
        Graph::DestroyGraph(fg.g);
        if (ok && !o.dump.empty()) {
            FILE *fp = fopen((o.dump + "." + std::to_string(g)).c_str(), "wb");
            for (int f = 0; fp && f < 6; ++f) fwrite(fin[g][f].data(), sizeof(float), fin[g][f].size(), fp);
            if (fp) fclose(fp);
        }
    }
    printf("{\"ok\": %s, \"mode\": \"%s\", \"form\": \"%s\", \"backend\": \"%s\", \"grid\": [%d, %d, %d], "
           "\"steps\": %d, \"graphs\": %d, \"sched\": \"%s\", \"ms\": %.4f, \"results\": [%s]}\n",
           ok ? "true" : "false", o.mode.c_str(), o.form.c_str(), G[0].hip ? "hip" : "host", o.nx, o.ny, o.nz,
           o.steps, o.graphs, o.sched.c_str(), ms, js.c_str());
    fflush(stdout);
    Runtime::Terminate();
    return ok ? 0 : 1;
}

}  // namespace

int main(int argc, char **argv) {
    Options o;
    if (!Parse(argc, argv, o)) {
        fprintf(stderr, "usage: ptk_fdtd --kernels <fdtd.hsaco|hosttasks.so> [--n N | --nx --ny --nz] "
                        "[--steps S] [--mode looped|unrolled] [--blocks B] [--repeat R] [--seed S] "
                        "[--form six|halves|fused|overlap|overlap2|overlap3|overlap3p] [--geometry default|reference|compute] [--tile IxJxK] [--dump file] "
                        "[--eager] [--profile] [--chain] [--serialize graph.xml] [--deserialize graph.xml] "
                        "[--graphs G [--priorities p0,p1,...] [--sched priority|fifo]] [--capture]\n");
        return 2;
    }
    Runtime::SetExitOnRuntimeFailure(FALSE);
    if (o.threads) Runtime::SetTaskThreadPoolSize((UINT)o.threads);
    if (o.hostConcurrency > 0) Runtime::SetMaximumHostConcurrency(o.hostConcurrency);
    if (o.profile) Runtime::SetProfilePSDispatch(TRUE);
    Runtime::SetTaskStreamPolicy(o.streams == "shared" ? PTK_STREAMS_SHARED : PTK_STREAMS_PER_TASK);
    if (o.mode == "unrolled") Runtime::SetDefaultChannelCapacity(std::max(4, o.blocks));
    if (o.graphs > 1) {
        if (!o.deserialize.empty() || !o.serialize.empty() || o.chain || o.blocks != 1 ||
            (o.sched != "priority" && o.sched != "fifo")) {
            fprintf(stderr, "--graphs runs built graphs once each: no --chain, --blocks, --serialize or --deserialize\n");
            return 2;
        }
        Runtime::SetSchedulingMode(o.sched == "fifo" ? SCHEDMODE_FIFO : SCHEDMODE_PRIORITY);
    }
    if (PTFAILED(Runtime::Initialize())) {
        fprintf(stderr, "Runtime::Initialize failed\n");
        return 1;
    }
    if (o.graphs > 1) return RunConcurrent(o);
    FdtdGraph G;
    G.o = o;
    if (!(o.deserialize.empty() ? G.Build() : G.Load())) {
        fprintf(stderr, "graph construction failed: %s\n", Runtime::GetLastErrorText());
        return 1;
    }
    if (o.pool == "global") {
        for (int f = 0; f < 6; ++f) Runtime::RequireBlockPool(G.field[f], o.poolBlocks);
    } else if (o.pool == "graph") {
        for (int f = 0; f < 6; ++f)
            if (!G.g->RequireBlockPool(G.field[f], o.poolBlocks)) {
                fprintf(stderr, "graph block pool: %s\n", Runtime::GetLastErrorText());
                return 1;
            }
    }
    Runtime::ResetMemoryStatistics();
    if (o.capture && !G.g->EnableLaunchCapture(TRUE)) {
        fprintf(stderr, "launch capture: %s\n", Runtime::GetLastErrorText());
        return 1;
    }
    if (!o.serialize.empty()) {
        // write the graph (built here or read from XML) and stop
        BOOL ok = G.g->Serialize(o.serialize.c_str());
        printf("{\"ok\": %s, \"serialized\": \"%s\", \"tasks\": %u, \"channels\": %u}\n", ok ? "true" : "false",
               o.serialize.c_str(), G.g->GetTaskCount(), G.g->GetChannelCount());
        G.g->Teardown();
        Graph::DestroyGraph(G.g);
        Runtime::Terminate();
        return ok ? 0 : 1;
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
    // chained: warmup runs, then the timed runs, then one run whose outputs are downloaded
    const int runs = o.chain ? o.warmup + o.repeat + 1 : o.repeat;
    std::vector<std::pair<double, UINT>> profileBase;
    const int blocks = o.mode == "looped" ? 1 : o.blocks;
    unsigned long long dispatch0 = 0;
    bool ok = true;
    std::vector<Datablock *> carried;  // chained runs: last run's outputs, still on the device
    Accelerator *hipAcc = G.hip ? Runtime::GetAccelerator(ACCELERATOR_CLASS_HIP, 0) : nullptr;
    double chainedMs = 0;
    CHighResolutionTimer chainTimer(gran_msec);
    for (int r = 0; r < runs && ok; ++r) {
        const bool last = r == runs - 1;
        if (o.chain && r == o.warmup) {
            // timed region of chained runs: from a drained device to the last run's completion
            G.g->WaitQuiescent(600000);
            if (hipAcc) hipAcc->Synchronize();
            for (UINT t = 0; t < G.g->GetTaskCount(); ++t) G.g->GetTaskByIndex(t)->GetTotalDispatchDeviceMs();
            profileBase.clear();
            for (UINT t = 0; t < G.g->GetTaskCount(); ++t)
                profileBase.push_back(std::make_pair(G.g->GetTaskByIndex(t)->GetTotalDispatchDeviceMs(),
                                                     G.g->GetTaskByIndex(t)->GetDispatchCount()));
            chainTimer.reset();
        }
        Runtime::ResetTransferStatistics();
        if (r > 0) {
            Runtime::SetAllocationFailureInjection(HOST_MEMORY_SPACE_ID, o.injectHost);
            if (hipAcc) Runtime::SetAllocationFailureInjection(hipAcc->GetMemorySpaceId(), o.injectDevice);
        }
        if (o.seedStep && r > 0 && carried.empty())  // this run's initial fields (off the clock)
            for (int f = 0; f < 6; ++f)
                for (size_t n = 0; n < init[f].size(); ++n)
                    init[f][n] = ptk_rand_sym_f32(o.seed + (unsigned long long)(r * o.seedStep), (uint64_t)f, n);
        dispatch0 = G.g->GetDispatchCount();
        timer.reset();
        for (int b = 0; b < blocks; ++b) {
            for (int f = 0; f < 6; ++f) {
                Datablock *db = nullptr;
                if (!carried.empty()) {
                    db = carried[f];
                } else if (o.pool == "graph") {
                    db = G.g->RequestPooledBlock(G.field[f], (UINT)(init[f].size() * sizeof(float)), 0, 0);
                    void *dst = db ? db->GetDataPointer(TRUE) : nullptr;
                    if (dst) memcpy(dst, init[f].data(), init[f].size() * sizeof(float));
                    else if (db) db->Release(), db = nullptr;
                } else {
                    db = Runtime::AllocateDatablock(G.field[f], init[f].data(),
                                                    (UINT)(init[f].size() * sizeof(float)), nullptr);
                }
                if (!db) {
                    fprintf(stderr, "run %d: no block for %s: %s\n", r, kFieldNames[f], Runtime::GetLastErrorText());
                    ok = false;
                    break;
                }
                for (GraphInputChannel *c : G.fieldInputs[f]) c->Push(db);
                db->Release();
            }
            carried.clear();
            if (!ok) break;
            if (o.mode == "looped") {
                int iters = o.steps / StepsPerTask(o.form);  // graph iterations
                Datablock *it = Runtime::AllocateDatablock(G.iters, &iters, sizeof(int), nullptr);
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
                if (o.chain && !last) {
                    res->ClearAllControlSignals();
                    carried.push_back(res);  // keep our reference; never touched on the host
                    continue;
                }
                res->Lock();
                const float *p = (const float *)res->GetDataPointer(FALSE);
                final[f].assign(p, p + init[f].size());
                res->Unlock();
                res->Release();
            }
        }
        if (o.chain && r == runs - 2) {
            // every timed run has been issued and pulled; wait for the device to finish them
            G.g->WaitQuiescent(600000);
            if (hipAcc) hipAcc->Synchronize();
            chainedMs = chainTimer.elapsed(false);
        }
        runMs.push_back(timer.elapsed(false));
        Runtime::GetTransferStatistics(&h2d0, &h2db, &d2h0, &d2hb);
        if (o.chain && !last && r > 0)
            fprintf(stderr, "run %d: %.3f ms, h2d %llu, d2h %llu\n", r, runMs.back(), h2d0, d2h0);
    }
    unsigned long long dispatches = G.g->GetDispatchCount() - dispatch0;
    // per-task device time of the timed runs (profiling events resolved now, off the clock)
    std::string taskJson;
    double devMs = 0;
    if (o.profile) {
        for (UINT t = 0; t < G.g->GetTaskCount(); ++t) {
            Task *task = G.g->GetTaskByIndex(t);
            double ms = task->GetTotalDispatchDeviceMs();
            double n = task->GetDispatchCount();
            if (t < profileBase.size()) {
                ms -= profileBase[t].first;
                n -= profileBase[t].second;
            }
            if (o.chain) n -= blocks;  // the untimed download run
            devMs += ms;
            char buf[256];
            snprintf(buf, sizeof(buf), "%s{\"task\": \"%s\", \"dispatches\": %.0f, \"device_ms\": %.4f}",
                     taskJson.empty() ? "" : ", ", task->GetTaskName(), n, ms);
            taskJson += buf;
        }
    }
    const bool captureArmed = G.g->IsLaunchCaptureArmed();
    const unsigned captureReplays = G.g->GetLaunchCaptureReplayCount();
    G.g->Stop();
    G.g->Teardown();
    Graph::DestroyGraph(G.g);

    if (ok && !o.dump.empty()) {
        FILE *fp = fopen(o.dump.c_str(), "wb");
        for (int f = 0; fp && f < 6; ++f) fwrite(final[f].data(), sizeof(float), final[f].size(), fp);
        if (fp) fclose(fp);
    }
    const double cells = (double)o.nx * o.ny * o.nz;
    double best = runMs.empty() ? 0 : runMs[0];
    for (double m : runMs) best = std::min(best, m);
    // chained: the timed region covers o.repeat runs from a drained device to completion
    double timedMs = o.chain ? chainedMs : best;
    int timedRuns = o.chain ? o.repeat : 1;
    printf("{\"ok\": %s, \"mode\": \"%s\", \"form\": \"%s\", \"backend\": \"%s\", \"grid\": [%d, %d, %d], \"steps\": %d, "
           "\"blocks\": %d, \"runs\": %d, \"chained\": %s, \"timed_runs\": %d, \"timed_ms\": %.4f, "
           "\"ms_per_run\": %.4f, \"best_ms\": %.4f, "
           "\"mcups\": %.3f, \"dispatches_last_run\": %llu, \"h2d_count\": %llu, \"h2d_bytes\": %llu, "
           "\"d2h_count\": %llu, \"d2h_bytes\": %llu, \"device_ms_profiled\": %.4f, \"tasks\": [%s], \"checksums\": [",
           ok ? "true" : "false", o.mode.c_str(), o.form.c_str(), G.hip ? "hip" : "host", o.nx, o.ny, o.nz, o.steps,
           blocks, (int)runMs.size(), o.chain ? "true" : "false", timedRuns, timedMs,
           timedRuns ? timedMs / timedRuns : 0.0, best,
           timedMs > 0 ? cells * o.steps * blocks * timedRuns / (timedMs / 1e3) / 1e6 : 0.0, dispatches, h2d0, h2db,
           d2h0, d2hb, devMs, taskJson.c_str());
    for (int f = 0; f < 6; ++f) {
        double s = 0;
        for (float v : final[f]) s += v;
        printf("%s{\"field\": \"%s\", \"fnv1a\": \"%016llx\", \"sum\": %.17g}", f ? ", " : "", kFieldNames[f],
               (unsigned long long)Fnv1a(final[f].data(), final[f].size() * sizeof(float)), s);
    }
    Runtime::PTK_MEMORY_STATISTICS ms = {};
    Runtime::GetMemoryStatistics(&ms);
    printf("], \"capture\": {\"requested\": %s, \"armed\": %s, \"replays\": %u}",
           o.capture ? "true" : "false", captureArmed ? "true" : "false", captureReplays);
    printf(", \"memory\": {\"allocation_failures\": %llu, \"recovered\": %llu, \"unrecovered\": %llu, "
           "\"bytes_reclaimed\": %llu, \"pool_hits\": %llu, \"pool_misses\": %llu}",
           ms.ullAllocationFailures, ms.ullAllocationsRecovered, ms.ullAllocationsUnrecovered, ms.ullBytesReclaimed,
           ms.ullPooledBlockHits, ms.ullPooledBlockMisses);
    printf(", \"run_ms\": [");
    for (size_t r = 0; r < runMs.size(); ++r) printf("%s%.4f", r ? ", " : "", runMs[r]);
    printf("]}\n");
    fflush(stdout);
    Runtime::Terminate();
    return ok ? 0 : 1;
}
