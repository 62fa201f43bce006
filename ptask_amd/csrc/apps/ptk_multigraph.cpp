// ptk_multigraph.cpp -- many live graphs on one runtime: dispatch admission by priority or
// arrival, global quiescence, graph pause/resume, and concurrent looped graphs (reference
// scenarios: rttests.cpp:936-1333 run_concurrent_graph_test / run_ultra_concurrent_graph_test,
// graphbench.cpp:752-845 priorities per graph instance and scheduling modes).
//
// The host-task library named by --kernels provides
//   RecordTask(tag int* in, int* out)     appends the tag to a process-wide log, holds the
//                                         dispatch for about 2 ms, copies the tag out
//   RecordLog(int* out, int max) -> n     the log so far (called here through dlsym)
//   ScaleVector(float* in, float s, int n, float* out)   out = s * in
//   ScaleInPlace(float* A, float s, int n)               A = s * A (an in/out port)
//   AddVectors(float* A, float* B, int n, float* out)    out = A + B
//
// Scenarios (--scenario):
//   contend     G graphs of K independent RecordTask tasks each, graph g's tasks at priority
//               --priorities[g].  The runtime is quiesced and the tasks get their inputs one at
//               a time, task-major (g0t0, g1t0, ..., g0t1, ...), each once the previous one waits
//               for admission, so the arrival order is known; then the quiescence ends.  With
//               --host-slots 1 the log is the admission order: by priority, then arrival, in
//               SCHEDMODE_PRIORITY; by arrival in SCHEDMODE_FIFO.
//   pause       one RecordTask graph stopped after Run: inputs pushed while it is stopped are
//               not dispatched until Run resumes it.
//   stall       a ScaleVector graph missing its scalar: the graph monitor watchdog (--watchdog-ms)
//               reports the stall, then the missing input arrives and the task runs.
//   concurrent  G graphs, each one ScaleVector task whose MF_SIMPLE_ITERATOR meta port repeats
//               its dispatch --iterations times, all running at once; every output is checked.
//   fanout      one output port feeding both an in/out consumer and a reader (reference
//               OutputPort.cpp:370-440): P = s1*A goes to ScaleInPlace (W = s2*P, in place) and
//               to AddVectors, whose second input is W, so the reader runs after the writer and
//               must still see P: out = P + W.  Repeated --iterations times on one graph.
// Prints one JSON object.
#include <dlfcn.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "ptaskapi.h"

using namespace PTask;

namespace {

struct Options {
    std::string kernels, scenario = "contend", mode = "priority", priorities;
    int graphs = 2, tasks = 4, hostSlots = 1, iterations = 5, elems = 1024, watchdogMs = 0;
};

std::vector<int> ParseList(const std::string &s) {
    std::vector<int> v;
    size_t i = 0;
    while (i < s.size()) {
        size_t j = s.find(',', i);
        if (j == std::string::npos) j = s.size();
        v.push_back(atoi(s.substr(i, j - i).c_str()));
        i = j + 1;
    }
    return v;
}

struct RecordGraph {
    Graph *g = nullptr;
    std::vector<GraphInputChannel *> in;
    std::vector<GraphOutputChannel *> out;
};

// One graph of `tasks` independent RecordTask tasks at priority `prio`.
bool BuildRecordGraph(const Options &o, int gi, int prio, DatablockTemplate *tInt, UINT &uid, RecordGraph &R) {
    R.g = new Graph(("record" + std::to_string(gi)).c_str());
    for (int t = 0; t < o.tasks; ++t) {
        CompiledKernel *k = Runtime::GetCompiledKernel((char *)o.kernels.c_str(), (char *)"RecordTask");
        if (!k) return false;
        Port *pin = Runtime::CreatePort(INPUT_PORT, tInt, uid++, (char *)"tag", 0);
        Port *pout = Runtime::CreatePort(OUTPUT_PORT, tInt, uid++, (char *)"tagOut", 1);
        const std::string nm = "g" + std::to_string(gi) + "t" + std::to_string(t);
        Task *task = R.g->AddTask(k, 1, &pin, 1, &pout, nm.c_str());
        if (!task) return false;
        task->SetComputeGeometry(1, 1, 1);
        task->SetPriority(prio);
        R.in.push_back(R.g->AddInputChannel(pin, (nm + ".in").c_str()));
        R.out.push_back(R.g->AddOutputChannel(pout, (nm + ".out").c_str()));
    }
    return true;
}

int Tag(const RecordGraph &, int gi, int t) { return gi * 1000 + t; }

bool PushTag(GraphInputChannel *c, DatablockTemplate *tInt, int tag) {
    Datablock *b = Runtime::AllocateDatablock(tInt, &tag, sizeof(tag), nullptr);
    if (!b) return false;
    c->Push(b);
    b->Release();
    return true;
}

bool PullTag(GraphOutputChannel *c, int want) {
    Datablock *b = c->Pull(60000);
    if (!b) return false;
    b->Lock();
    int v = 0;
    memcpy(&v, b->GetDataPointer(FALSE), sizeof(v));
    b->Unlock();
    b->Release();
    return v == want;
}

using LogFn = int (*)(int *, int);

std::vector<int> ReadLog(const Options &o) {
    std::vector<int> v(4096);
    void *h = dlopen(o.kernels.c_str(), RTLD_NOW | RTLD_NOLOAD);  // the runtime's own handle
    if (!h) h = dlopen(o.kernels.c_str(), RTLD_NOW);
    LogFn f = h ? (LogFn)dlsym(h, "RecordLog") : nullptr;
    const int n = f ? f(v.data(), (int)v.size()) : 0;
    v.resize((size_t)n);
    return v;
}

std::string JsonList(const std::vector<int> &v) {
    std::string s = "[";
    for (size_t i = 0; i < v.size(); ++i) s += (i ? ", " : "") + std::to_string(v[i]);
    return s + "]";
}

int Contend(const Options &o) {
    std::vector<int> prio = ParseList(o.priorities);
    prio.resize((size_t)o.graphs, 0);
    DatablockTemplate *tInt = Runtime::GetDatablockTemplate((char *)"tag", sizeof(int), 1, 1, 1);
    UINT uid = 0;
    std::vector<RecordGraph> G((size_t)o.graphs);
    for (int gi = 0; gi < o.graphs; ++gi)
        if (!BuildRecordGraph(o, gi, prio[(size_t)gi], tInt, uid, G[(size_t)gi])) {
            fprintf(stderr, "graph construction failed: %s\n", Runtime::GetLastErrorText());
            return 1;
        }
    Runtime::GlobalQuiesce();
    for (auto &R : G) R.g->Run();
    std::vector<int> arrival;
    for (int t = 0; t < o.tasks; ++t)
        for (int gi = 0; gi < o.graphs; ++gi) {
            PushTag(G[(size_t)gi].in[(size_t)t], tInt, Tag(G[(size_t)gi], gi, t));
            arrival.push_back(Tag(G[(size_t)gi], gi, t));
            // this dispatch waits for admission before the next input is pushed
            auto t0 = std::chrono::steady_clock::now();
            while (Runtime::GetWaitingDispatchCount() < (int)arrival.size()) {
                if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(30)) {
                    fprintf(stderr, "dispatch %zu did not reach admission\n", arrival.size());
                    Runtime::EndGlobalQuiescence();
                    return 1;
                }
                std::this_thread::sleep_for(std::chrono::microseconds(200));
            }
        }
    const int waited = Runtime::GetWaitingDispatchCount();
    Runtime::EndGlobalQuiescence();
    bool ok = true;
    for (int gi = 0; gi < o.graphs; ++gi)
        for (int t = 0; t < o.tasks; ++t) ok = PullTag(G[(size_t)gi].out[(size_t)t], Tag(G[(size_t)gi], gi, t)) && ok;
    const std::vector<int> log = ReadLog(o);
    for (auto &R : G) {
        R.g->Stop();
        R.g->Teardown();
        Graph::DestroyGraph(R.g);
    }
    printf("{\"scenario\": \"contend\", \"ok\": %s, \"mode\": \"%s\", \"graphs\": %d, \"tasks\": %d, "
           "\"host_slots\": %d, \"waiting_at_release\": %d, \"priorities\": %s, \"arrival\": %s, "
           "\"order\": %s}\n",
           ok ? "true" : "false", o.mode.c_str(), o.graphs, o.tasks, o.hostSlots, waited,
           JsonList(prio).c_str(), JsonList(arrival).c_str(), JsonList(log).c_str());
    return ok ? 0 : 1;
}

int Pause(const Options &o) {
    DatablockTemplate *tInt = Runtime::GetDatablockTemplate((char *)"tag", sizeof(int), 1, 1, 1);
    UINT uid = 0;
    RecordGraph R;
    if (!BuildRecordGraph(o, 0, 0, tInt, uid, R)) return 1;
    R.g->Run();
    R.g->Stop();  // quiesced: runnable, no dispatch until Run
    for (int t = 0; t < o.tasks; ++t) PushTag(R.in[(size_t)t], tInt, Tag(R, 0, t));
    std::this_thread::sleep_for(std::chrono::milliseconds(100));
    const size_t whileStopped = ReadLog(o).size();
    const unsigned long long dispatchesStopped = R.g->GetDispatchCount();
    R.g->Run();
    bool ok = true;
    for (int t = 0; t < o.tasks; ++t) ok = PullTag(R.out[(size_t)t], Tag(R, 0, t)) && ok;
    const unsigned long long dispatchesAfter = R.g->GetDispatchCount();
    R.g->Stop();
    R.g->Teardown();
    Graph::DestroyGraph(R.g);
    printf("{\"scenario\": \"pause\", \"ok\": %s, \"tasks\": %d, \"logged_while_stopped\": %zu, "
           "\"dispatches_while_stopped\": %llu, \"dispatches_after_resume\": %llu}\n",
           ok ? "true" : "false", o.tasks, whileStopped, dispatchesStopped, dispatchesAfter);
    return ok ? 0 : 1;
}

int Concurrent(const Options &o) {
    DatablockTemplate *tVec = Runtime::GetDatablockTemplate((char *)"vec", sizeof(float), o.elems, 1, 1);
    DatablockTemplate *tScale = Runtime::GetDatablockTemplate((char *)"vscale", sizeof(float), PTPARM_FLOAT);
    DatablockTemplate *tN = Runtime::GetDatablockTemplate((char *)"vecdims", sizeof(int), PTPARM_INT);
    DatablockTemplate *tIter = Runtime::GetDatablockTemplate((char *)"dbiter", sizeof(int), PTPARM_INT);
    UINT uid = 0;
    struct G1 {
        Graph *g;
        GraphInputChannel *a, *s, *n, *it;
        GraphOutputChannel *out;
    };
    std::vector<G1> G((size_t)o.graphs);
    for (int gi = 0; gi < o.graphs; ++gi) {
        G1 &x = G[(size_t)gi];
        x.g = new Graph(("scale" + std::to_string(gi)).c_str());
        CompiledKernel *k = Runtime::GetCompiledKernel((char *)o.kernels.c_str(), (char *)"ScaleVector");
        if (!k) return 1;
        Port *in[4] = {Runtime::CreatePort(INPUT_PORT, tVec, uid++, (char *)"A", 0),
                       Runtime::CreatePort(STICKY_PORT, tScale, uid++, (char *)"scalar", 1),
                       Runtime::CreatePort(STICKY_PORT, tN, uid++, (char *)"N", 2),
                       Runtime::CreatePort(META_PORT, tIter, uid++, (char *)"iter_meta")};
        in[3]->SetMetaFunction(MF_SIMPLE_ITERATOR);
        Port *out = Runtime::CreatePort(OUTPUT_PORT, tVec, uid++, (char *)"Aout", 3);
        Task *t = x.g->AddTask(k, 4, in, 1, &out, ("scale" + std::to_string(gi)).c_str());
        if (!t) return 1;
        t->SetComputeGeometry(1, 1, 1);
        x.a = x.g->AddInputChannel(in[0], "A");
        x.s = x.g->AddInputChannel(in[1], "scalar");
        x.n = x.g->AddInputChannel(in[2], "N");
        x.it = x.g->AddInputChannel(in[3], "iterations");
        x.out = x.g->AddOutputChannel(out, "Aout");
        x.g->Run();
    }
    // every graph gets its inputs before any result is pulled: all run at once
    std::vector<std::vector<float>> A((size_t)o.graphs);
    for (int gi = 0; gi < o.graphs; ++gi) {
        G1 &x = G[(size_t)gi];
        A[(size_t)gi].resize((size_t)o.elems);
        for (int e = 0; e < o.elems; ++e) A[(size_t)gi][(size_t)e] = (float)(gi + 1) + 0.25f * (float)e;
        const float s = 1.0f + 0.125f * (float)gi;
        const int n = o.elems, iters = o.iterations;
        Datablock *bs = Runtime::AllocateDatablock(tScale, (void *)&s, sizeof(s), nullptr);
        Datablock *bn = Runtime::AllocateDatablock(tN, (void *)&n, sizeof(n), nullptr);
        Datablock *bi = Runtime::AllocateDatablock(tIter, (void *)&iters, sizeof(iters), nullptr);
        Datablock *ba = Runtime::AllocateDatablock(tVec, A[(size_t)gi].data(), (UINT)(n * sizeof(float)), nullptr);
        x.s->Push(bs);
        x.n->Push(bn);
        x.it->Push(bi);
        x.a->Push(ba);
        bs->Release();
        bn->Release();
        bi->Release();
        ba->Release();
    }
    int bad = 0;
    unsigned long long dispatches = 0;
    for (int gi = 0; gi < o.graphs; ++gi) {
        G1 &x = G[(size_t)gi];
        Datablock *r = x.out->Pull(60000);
        if (!r) {
            ++bad;
            continue;
        }
        r->Lock();
        const float *p = (const float *)r->GetDataPointer(FALSE);
        const float s = 1.0f + 0.125f * (float)gi;
        for (int e = 0; e < o.elems; ++e)
            if (p[e] != s * A[(size_t)gi][(size_t)e]) {
                ++bad;
                break;
            }
        r->Unlock();
        r->Release();
        dispatches += x.g->GetDispatchCount();
    }
    for (auto &x : G) {
        x.g->Stop();
        x.g->Teardown();
        Graph::DestroyGraph(x.g);
    }
    printf("{\"scenario\": \"concurrent\", \"ok\": %s, \"graphs\": %d, \"iterations\": %d, \"bad_graphs\": %d, "
           "\"dispatches\": %llu}\n",
           bad ? "false" : "true", o.graphs, o.iterations, bad, dispatches);
    return bad ? 1 : 0;
}

// stall: one ScaleVector graph whose vector arrives while its scalar does not, so the task can
// never dispatch: with the graph monitor watchdog on (threshold --watchdog-ms) the stall is
// reported.  Then the missing inputs arrive, the task runs, and an idle graph with nothing pending
// draws no further report.
int Stall(const Options &o) {
    DatablockTemplate *tVec = Runtime::GetDatablockTemplate((char *)"vec", sizeof(float), (UINT)o.elems, 1, 1);
    DatablockTemplate *tScale = Runtime::GetDatablockTemplate((char *)"scale", sizeof(float), PTPARM_FLOAT);
    DatablockTemplate *tN = Runtime::GetDatablockTemplate((char *)"n", sizeof(int), PTPARM_INT);
    CompiledKernel *k = Runtime::GetCompiledKernel((char *)o.kernels.c_str(), (char *)"ScaleVector");
    if (!k) return 1;
    UINT uid = 0;
    Graph *g = new Graph("stall");
    Port *in[3] = {Runtime::CreatePort(INPUT_PORT, tVec, uid++, (char *)"A", 0),
                   Runtime::CreatePort(STICKY_PORT, tScale, uid++, (char *)"scalar", 1),
                   Runtime::CreatePort(STICKY_PORT, tN, uid++, (char *)"N", 2)};
    Port *out = Runtime::CreatePort(OUTPUT_PORT, tVec, uid++, (char *)"Aout", 3);
    Task *t = g->AddTask(k, 3, in, 1, &out, "scale");
    if (!t) return 1;
    t->SetComputeGeometry(1, 1, 1);
    GraphInputChannel *ca = g->AddInputChannel(in[0], "A"), *cs = g->AddInputChannel(in[1], "scalar");
    GraphInputChannel *cn = g->AddInputChannel(in[2], "N");
    GraphOutputChannel *co = g->AddOutputChannel(out, "Aout");
    g->Run();
    std::vector<float> A((size_t)o.elems);
    for (int e = 0; e < o.elems; ++e) A[(size_t)e] = 0.5f * (float)e;
    const float s = 3.0f;
    const int n = o.elems;
    Datablock *ba = Runtime::AllocateDatablock(tVec, A.data(), (UINT)(n * sizeof(float)), nullptr);
    ca->Push(ba);
    ba->Release();
    std::this_thread::sleep_for(std::chrono::milliseconds(4 * o.watchdogMs));
    const UINT during = g->GetWatchdogReportCount();
    Datablock *bs = Runtime::AllocateDatablock(tScale, (void *)&s, sizeof(s), nullptr);
    Datablock *bn = Runtime::AllocateDatablock(tN, (void *)&n, sizeof(n), nullptr);
    cs->Push(bs);
    cn->Push(bn);
    bs->Release();
    bn->Release();
    bool ok = true;
    Datablock *r = co->Pull(60000);
    if (!r) {
        ok = false;
    } else {
        r->Lock();
        const float *p = (const float *)r->GetDataPointer(FALSE);
        for (int e = 0; e < n; ++e) ok = ok && p[e] == s * A[(size_t)e];
        r->Unlock();
        r->Release();
    }
    std::this_thread::sleep_for(std::chrono::milliseconds(4 * o.watchdogMs));
    const UINT after = g->GetWatchdogReportCount();
    g->Stop();
    g->Teardown();
    Graph::DestroyGraph(g);
    printf("{\"scenario\": \"stall\", \"ok\": %s, \"reports_during_stall\": %u, \"reports_after\": %u, "
           "\"runtime_reports\": %llu}\n",
           ok ? "true" : "false", during, after, Runtime::GetWatchdogReportCount());
    return ok ? 0 : 1;
}

// fanout: P = ScaleVector(A, s1) fans out to W = ScaleInPlace(P, s2) (in/out: port 0 in and out)
// and to R = AddVectors(P, W).  R cannot dispatch before W has produced W, so a writer that
// scaled the shared block in place would hand R s2*P for P; the runtime must give the writer its
// own copy instead.  out = s1*A + s2*s1*A, exact in binary32 for the values used.
int Fanout(const Options &o) {
    DatablockTemplate *tVec = Runtime::GetDatablockTemplate((char *)"vec", sizeof(float), (UINT)o.elems, 1, 1);
    DatablockTemplate *tScale = Runtime::GetDatablockTemplate((char *)"scale", sizeof(float), PTPARM_FLOAT);
    DatablockTemplate *tN = Runtime::GetDatablockTemplate((char *)"n", sizeof(int), PTPARM_INT);
    CompiledKernel *kP = Runtime::GetCompiledKernel((char *)o.kernels.c_str(), (char *)"ScaleVector");
    CompiledKernel *kW = Runtime::GetCompiledKernel((char *)o.kernels.c_str(), (char *)"ScaleInPlace");
    CompiledKernel *kR = Runtime::GetCompiledKernel((char *)o.kernels.c_str(), (char *)"AddVectors");
    if (!kP || !kW || !kR) return 1;
    UINT uid = 0;
    Graph *g = new Graph("fanout");
    Port *pin[3] = {Runtime::CreatePort(INPUT_PORT, tVec, uid++, (char *)"A", 0),
                    Runtime::CreatePort(STICKY_PORT, tScale, uid++, (char *)"s1", 1),
                    Runtime::CreatePort(STICKY_PORT, tN, uid++, (char *)"N", 2)};
    Port *pout = Runtime::CreatePort(OUTPUT_PORT, tVec, uid++, (char *)"P", 3);
    Port *win[3] = {Runtime::CreatePort(INPUT_PORT, tVec, uid++, (char *)"W", 0, 0),
                    Runtime::CreatePort(STICKY_PORT, tScale, uid++, (char *)"s2", 1),
                    Runtime::CreatePort(STICKY_PORT, tN, uid++, (char *)"N", 2)};
    Port *wout = Runtime::CreatePort(OUTPUT_PORT, tVec, uid++, (char *)"W(out)", 0);
    Port *rin[3] = {Runtime::CreatePort(INPUT_PORT, tVec, uid++, (char *)"X", 0),
                    Runtime::CreatePort(INPUT_PORT, tVec, uid++, (char *)"Y", 1),
                    Runtime::CreatePort(STICKY_PORT, tN, uid++, (char *)"N", 2)};
    Port *rout = Runtime::CreatePort(OUTPUT_PORT, tVec, uid++, (char *)"out", 3);
    Task *tP = g->AddTask(kP, 3, pin, 1, &pout, "P");
    Task *tW = g->AddTask(kW, 3, win, 1, &wout, "W");
    Task *tR = g->AddTask(kR, 3, rin, 1, &rout, "R");
    if (!tP || !tW || !tR) return 1;
    for (Task *t : {tP, tW, tR}) t->SetComputeGeometry(1, 1, 1);
    GraphInputChannel *ca = g->AddInputChannel(pin[0], "A"), *cs1 = g->AddInputChannel(pin[1], "s1");
    GraphInputChannel *cs2 = g->AddInputChannel(win[1], "s2");
    GraphInputChannel *cn[3] = {g->AddInputChannel(pin[2], "Np"), g->AddInputChannel(win[2], "Nw"),
                                g->AddInputChannel(rin[2], "Nr")};
    g->AddInternalChannel(pout, win[0], "P->W");
    g->AddInternalChannel(pout, rin[0], "P->R");
    g->AddInternalChannel(wout, rin[1], "W->R");
    GraphOutputChannel *co = g->AddOutputChannel(rout, "out");
    g->Run();
    const float s1 = 1.5f, s2 = 4.0f;
    const int n = o.elems;
    Datablock *b1 = Runtime::AllocateDatablock(tScale, (void *)&s1, sizeof(s1), nullptr);
    Datablock *b2 = Runtime::AllocateDatablock(tScale, (void *)&s2, sizeof(s2), nullptr);
    cs1->Push(b1);
    cs2->Push(b2);
    b1->Release();
    b2->Release();
    for (GraphInputChannel *c : cn) {
        Datablock *bn = Runtime::AllocateDatablock(tN, (void *)&n, sizeof(n), nullptr);
        c->Push(bn);
        bn->Release();
    }
    bool ok = true;
    int bad = 0;
    for (int it = 0; it < o.iterations; ++it) {
        std::vector<float> A((size_t)n);
        for (int e = 0; e < n; ++e) A[(size_t)e] = (float)(it + 1) + 0.25f * (float)(e % 1024);
        Datablock *ba = Runtime::AllocateDatablock(tVec, A.data(), (UINT)(n * sizeof(float)), nullptr);
        ca->Push(ba);
        ba->Release();
        Datablock *r = co->Pull(60000);
        if (!r) {
            ok = false;
            break;
        }
        r->Lock();
        const float *p = (const float *)r->GetDataPointer(FALSE);
        for (int e = 0; e < n; ++e) {
            const float P = s1 * A[(size_t)e];
            if (p[e] != P + s2 * P) ++bad;
        }
        r->Unlock();
        r->Release();
    }
    g->Stop();
    g->Teardown();
    Graph::DestroyGraph(g);
    ok = ok && bad == 0;
    printf("{\"scenario\": \"fanout\", \"ok\": %s, \"iterations\": %d, \"elems\": %d, \"mismatches\": %d}\n",
           ok ? "true" : "false", o.iterations, n, bad);
    return ok ? 0 : 1;
}

}  // namespace

int main(int argc, char **argv) {
    Options o;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> const char * {
            if (i + 1 >= argc) {
                fprintf(stderr, "missing value for %s\n", a.c_str());
                exit(2);
            }
            return argv[++i];
        };
        if (a == "--kernels") o.kernels = next();
        else if (a == "--scenario") o.scenario = next();
        else if (a == "--mode") o.mode = next();
        else if (a == "--priorities") o.priorities = next();
        else if (a == "--graphs") o.graphs = atoi(next());
        else if (a == "--tasks") o.tasks = atoi(next());
        else if (a == "--host-slots") o.hostSlots = atoi(next());
        else if (a == "--iterations") o.iterations = atoi(next());
        else if (a == "--elems") o.elems = atoi(next());
        else if (a == "--watchdog-ms") o.watchdogMs = atoi(next());
        else {
            fprintf(stderr, "unknown option %s\n", a.c_str());
            return 2;
        }
    }
    if (o.kernels.empty() || o.graphs < 1 || o.tasks < 1 || o.hostSlots < 1 || o.iterations < 1 || o.elems < 1 ||
        (o.mode != "priority" && o.mode != "fifo")) {
        fprintf(stderr, "usage: ptk_multigraph --kernels <hosttasks.so> [--scenario contend|pause|concurrent|stall|fanout] "
                        "[--mode priority|fifo] [--graphs G] [--tasks K] [--priorities p0,p1,...] "
                        "[--host-slots S] [--iterations I] [--elems N] [--watchdog-ms T]\n");
        return 2;
    }
    Runtime::SetExitOnRuntimeFailure(FALSE);
    Runtime::SetUseHIP(FALSE);  // host tasks only
    Runtime::SetMaximumHostConcurrency(o.hostSlots);
    Runtime::SetSchedulingMode(o.mode == "priority" ? SCHEDMODE_PRIORITY : SCHEDMODE_FIFO);
    if (o.watchdogMs > 0) {
        Runtime::SetUseGraphMonitorWatchdog(TRUE);
        Runtime::SetDispatchWatchdogThreshold((DWORD)o.watchdogMs);
    }
    // enough runner threads per graph for every task of a graph to wait for admission at once
    Runtime::SetTaskThreadPoolSize((UINT)std::max(o.tasks, 1));
    if (PTFAILED(Runtime::Initialize())) {
        fprintf(stderr, "Runtime::Initialize failed\n");
        return 1;
    }
    int r = 2;
    if (o.scenario == "contend") r = Contend(o);
    else if (o.scenario == "pause") r = Pause(o);
    else if (o.scenario == "concurrent") r = Concurrent(o);
    else if (o.scenario == "stall") r = Stall(o);
    else if (o.scenario == "fanout") r = Fanout(o);
    else fprintf(stderr, "unknown scenario %s\n", o.scenario.c_str());
    Runtime::Terminate();
    return r;
}
