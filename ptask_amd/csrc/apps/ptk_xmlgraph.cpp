// ptk_xmlgraph.cpp -- a host graph that uses every structural feature Graph::Serialize writes,
// for XML round trips (reference: graph serialization modes 1 and 2 of the unit tests, e.g.
// graphcuadd.cpp:70-80,148-155 and controlpropagation.cpp:184-186,328-335: build or read the
// graph, then look its channels up by name and run it).
//
// The host-task library named by --kernels provides ScaleVector(float* in, float s, int n,
// float* out), out = s * in.  The graph:
//   multi-channel "A"      feeds scale1.A and scale3.A
//   scale1                 s and n from initializer channels (templates with initial values 3
//                          and the vector length), A's control codes propagate to its output
//   scale2 (priority 5)    A from scale1 on an internal channel, s from graph input "s2", n from
//                          an initializer, output on "out2"
//   scale3                 s from "s3" (a switch channel), n from an initializer, output on
//                          "out3" (capacity 7, eager views); scale3.A controls the output
//                          port's gate (initially closed: a controller block that is not all
//                          zero opens it for that dispatch) and propagates control codes to
//                          channel "out3"
//   probe                  never fed: carries a simple-iterator meta port scoped to its input,
//                          a descriptor binding, a permanent-block port, a non-marshallable port,
//                          a predicated input channel, a template with padded pitch and a
//                          metadata channel, an explicit grid, and user-allocated outputs
// out2 = 1.5 * A, out3 = s3 * A.
//
// Usage: ptk_xmlgraph (--kernels lib.so | --read graph.xml) [--write graph.xml] [--run]
//        [--n N] [--s3 S] [--unwritable estimator|predicate|uid|channel]
// --unwritable adds what has no text form to the built graph (a geometry estimator function, a
// user-defined predicate function, a second port with an existing UniqueID, a second channel
// with an existing name), for the tests of Serialize's refusals.
// --write without --run writes the graph (built or read) and stops.  --run pushes a seeded A,
// pulls both outputs, checks them against the closed forms and prints one JSON object with
// their checksums.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ptaskapi.h"
#include "ptk_prng.h"

using namespace PTask;

namespace {

struct Options {
    std::string kernels, read, write, unwritable;
    bool run = false;
    int n = 64;
    float s3 = 2.0f;
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

void __stdcall Estimate(UINT, Datablock **, PTASKDIM3 *block, PTASKDIM3 *grid) {
    *block = PTASKDIM3(1, 1, 1);
    *grid = PTASKDIM3(1, 1, 1);
}

BOOL __stdcall Predicate(Channel *, Datablock *, PREDICATE_FAILURE_ACTION) { return TRUE; }

Graph *Build(const Options &o) {
    DatablockTemplate *tVec = Runtime::GetDatablockTemplate((char *)"xgVec", sizeof(float), (UINT)o.n, 1, 1);
    DatablockTemplate *tS = Runtime::GetDatablockTemplate((char *)"xgScale", sizeof(float), PTPARM_FLOAT);
    DatablockTemplate *tS1 = Runtime::GetDatablockTemplate((char *)"xgScale3", sizeof(float), PTPARM_FLOAT);
    DatablockTemplate *tN = Runtime::GetDatablockTemplate((char *)"xgN", sizeof(int), PTPARM_INT);
    DatablockTemplate *tIter = Runtime::GetDatablockTemplate((char *)"xgIter", sizeof(int), PTPARM_INT);
    float three = 3.0f;
    tS1->SetInitialValue(&three, sizeof(three), 1);
    tN->SetInitialValue((void *)&o.n, sizeof(int), 1);
    // a record layout with a padded pitch and a metadata channel
    BUFFERDIMENSIONS dims[NUM_DATABLOCK_CHANNELS];
    dims[DBDATA_IDX].Initialize((UINT)o.n, 1, 1, sizeof(float), (UINT)(o.n * sizeof(float) + 64));
    dims[DBMETADATA_IDX].Initialize(16, 1, 1, sizeof(int), 16 * sizeof(int));
    for (UINT c = DBMETADATA_IDX + 1; c < NUM_DATABLOCK_CHANNELS; ++c) dims[c].Initialize(0, 0, 0, 0, 0);
    DatablockTemplate *tRec = Runtime::GetDatablockTemplate((char *)"xgRecords", dims, NUM_DATABLOCK_CHANNELS, true, false);

    CompiledKernel *k = Runtime::GetCompiledKernel((char *)o.kernels.c_str(), (char *)"ScaleVector");
    if (!k) return nullptr;
    Graph *g = new Graph("xmlfeatures");
    UINT uid = 100;
    auto task = [&](const char *name, DatablockTemplate *sT, DatablockTemplate *outT, Port **in, Port **out) {
        in[0] = Runtime::CreatePort(INPUT_PORT, tVec, uid++, (char *)"A", 0);
        in[1] = Runtime::CreatePort(STICKY_PORT, sT, uid++, (char *)"s", 1);
        in[2] = Runtime::CreatePort(STICKY_PORT, tN, uid++, (char *)"n", 2);
        out[0] = Runtime::CreatePort(OUTPUT_PORT, outT, uid++, (char *)"out", 3);
        Task *t = g->AddTask(k, 3, in, 1, out, name);
        t->SetComputeGeometry(1, 1, 1);
        return t;
    };
    Port *i1[3], *o1[1], *i2[3], *o2[1], *i3[3], *o3[1];
    task("scale1", tS1, tVec, i1, o1);
    Task *t2 = task("scale2", tS, tVec, i2, o2);
    t2->SetPriority(5);
    task("scale3", tS, tVec, i3, o3);

    // probe: A, s, n as the others, plus a simple-iterator meta port; output in the record layout
    Port *ip[4], *op[1];
    ip[0] = Runtime::CreatePort(INPUT_PORT, tVec, uid++, (char *)"A", 0);
    ip[1] = Runtime::CreatePort(STICKY_PORT, tS, uid++, (char *)"s", 1);
    ip[2] = Runtime::CreatePort(INPUT_PORT, tN, uid++, (char *)"n", 2);
    ip[3] = Runtime::CreatePort(META_PORT, tIter, uid++, (char *)"repeat");
    ip[3]->SetMetaFunction(MF_SIMPLE_ITERATOR);
    op[0] = Runtime::CreatePort(OUTPUT_PORT, tRec, uid++, (char *)"out", 3);
    ip[1]->SetPermanentBlock(TRUE);
    ip[0]->SetMarshallable(FALSE);
    Task *tp = g->AddTask(k, 4, ip, 1, op, "probe");
    tp->SetBlockAndGridSize(PTASKDIM3(4, 2, 1), PTASKDIM3(64, 1, 1));
    tp->SetUserCodeAllocatesMemory(TRUE);

    Port *aDst[2] = {i1[0], i3[0]};
    g->AddMultiChannel(aDst, 2, "A");
    g->AddInitializerChannel(i1[1], "s1.init");
    g->AddInitializerChannel(i1[2], "n1.init");
    g->AddInternalChannel(o1[0], i2[0], "scale1->scale2");
    g->AddInputChannel(i2[1], "s2");
    g->AddInitializerChannel(i2[2], "n2.init");
    GraphOutputChannel *out2 = g->AddOutputChannel(o2[0], "out2");
    g->AddInputChannel(i3[1], "s3", TRUE);
    InitializerChannel *n3 = g->AddInitializerChannel(i3[2], "n3.init");
    n3->SetInitialPropagatedControlSignal(DBCTLC_NONE);
    GraphOutputChannel *out3 = g->AddOutputChannel(o3[0], "out3");
    out3->SetCapacity(7);
    out3->SetViewMaterializationPolicy(VIEWMATERIALIZATIONPOLICY_EAGER);
    (void)out2;
    g->BindControlPropagationPort(i1[0], o1[0]);
    g->BindControlPort(i3[0], o3[0], FALSE);
    g->BindControlPropagationChannel(i3[0], out3);

    GraphInputChannel *pa = g->AddInputChannel(ip[0], "probe.A");
    pa->SetPredicationType(CE_DST, CGATEFN_CLOSE_ON_EOF, PFA_FAIL_OPERATION);
    g->AddInputChannel(ip[1], "probe.s");
    g->AddInputChannel(ip[3], "probe.repeat");
    InitializerChannel *pn = g->AddInitializerChannel(ip[2], "probe.n");
    pn->SetInitialPropagatedControlSignal(DBCTLC_BOF);
    g->AddOutputChannel(op[0], "probe.out")->SetPredicationType(CE_SRC, CGATEFN_OPEN_ON_ENDITERATION);
    g->BindIterationScope(ip[3], ip[0]);
    g->BindDescriptorPort(ip[0], ip[2], DF_SIZE);

    if (o.unwritable == "estimator") {
        tp->SetGeometryEstimator(Estimate);
    } else if (o.unwritable == "predicate") {
        pa->SetPredicator(CE_DST, Predicate);
    } else if (o.unwritable == "uid" || o.unwritable == "channel") {
        // a second probe-like task; its input port reuses scale1.A's UniqueID, or its input
        // channel reuses the name "s2"
        Port *in[3] = {Runtime::CreatePort(INPUT_PORT, tVec, o.unwritable == "uid" ? i1[0]->GetUID() : uid++, (char *)"A", 0),
                       Runtime::CreatePort(STICKY_PORT, tS, uid++, (char *)"s", 1),
                       Runtime::CreatePort(STICKY_PORT, tN, uid++, (char *)"n", 2)};
        Port *out[1] = {Runtime::CreatePort(OUTPUT_PORT, tVec, uid++, (char *)"out", 3)};
        g->AddTask(k, 3, in, 1, out, "extra");
        g->AddInputChannel(in[0], o.unwritable == "channel" ? "s2" : "extra.A");
    }
    return g;
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
        if (a == "--kernels") o.kernels = next();
        else if (a == "--read") o.read = next();
        else if (a == "--write") o.write = next();
        else if (a == "--run") o.run = true;
        else if (a == "--n") o.n = atoi(next());
        else if (a == "--s3") o.s3 = (float)atof(next());
        else if (a == "--unwritable") o.unwritable = next();
        else { fprintf(stderr, "unknown option %s\n", a.c_str()); return 2; }
    }
    if ((o.kernels.empty() == o.read.empty()) || o.n < 1) {
        fprintf(stderr, "usage: ptk_xmlgraph (--kernels lib.so | --read graph.xml) [--write graph.xml] [--run] [--n N] [--s3 S]\n");
        return 2;
    }
    Runtime::SetExitOnRuntimeFailure(FALSE);
    Runtime::SetUseHIP(FALSE);
    if (PTFAILED(Runtime::Initialize())) return 1;
    Graph *g = nullptr;
    if (o.read.empty()) {
        g = Build(o);
    } else {
        g = new Graph();
        if (!g->Deserialize(o.read.c_str())) {
            Graph::DestroyGraph(g);
            g = nullptr;
        }
    }
    if (!g) {
        printf("{\"ok\": false, \"error\": \"%s\"}\n", Runtime::GetLastErrorText());
        Runtime::Terminate();
        return 1;
    }
    bool ok = true;
    if (!o.write.empty()) ok = g->Serialize(o.write.c_str()) == TRUE;
    std::string json = "{\"graph\": \"" + std::string(g->GetName()) + "\", \"tasks\": " + std::to_string(g->GetTaskCount()) +
                       ", \"channels\": " + std::to_string(g->GetChannelCount());
    if (ok && o.run) {
        // channels looked up by name, as an application of a deserialized graph does
        Channel *A = g->GetChannel("A"), *s2 = g->GetChannel("s2"), *s3 = g->GetChannel("s3");
        Channel *out2 = g->GetChannel("out2"), *out3 = g->GetChannel("out3");
        DatablockTemplate *tVec = Runtime::GetDatablockTemplate((char *)"xgVec");
        DatablockTemplate *tS = Runtime::GetDatablockTemplate((char *)"xgScale");
        if (!A || !s2 || !s3 || !out2 || !out3 || !tVec || !tS) {
            fprintf(stderr, "the graph lacks a channel or template this program feeds\n");
            ok = false;
        } else {
            g->Run();
            std::vector<float> a((size_t)o.n);
            for (int i = 0; i < o.n; ++i) a[(size_t)i] = ptk_rand_sym_f32(0x3A1, 0, (uint64_t)i);
            float half = 0.5f;
            Datablock *bA = Runtime::AllocateDatablock(tVec, a.data(), (UINT)(a.size() * sizeof(float)), nullptr);
            Datablock *b2 = Runtime::AllocateDatablock(tS, &half, sizeof(half), nullptr);
            Datablock *b3 = Runtime::AllocateDatablock(tS, &o.s3, sizeof(o.s3), nullptr);
            s2->Push(b2);
            s3->Push(b3);
            A->Push(bA);
            bA->Release();
            b2->Release();
            b3->Release();
            Channel *outs[2] = {out2, out3};
            const float scale[2] = {3.0f * 0.5f, o.s3};
            for (int q = 0; q < 2 && ok; ++q) {
                Datablock *r = outs[q]->Pull(60000);
                if (!r) {
                    fprintf(stderr, "no block on %s\n", outs[q]->GetName());
                    ok = false;
                    break;
                }
                r->Lock();
                const float *p = (const float *)r->GetDataPointer(FALSE);
                int bad = 0;
                for (int i = 0; i < o.n; ++i) {
                    // out2 = 0.5 * (3 * A); out3 = s3 * A, each product rounded as the task does
                    const float want = q == 0 ? 0.5f * (3.0f * a[(size_t)i]) : o.s3 * a[(size_t)i];
                    bad += p[i] != want;
                }
                char buf[160];
                snprintf(buf, sizeof(buf), ", \"%s\": {\"scale\": %g, \"bad\": %d, \"fnv1a\": \"%016llx\"}", outs[q]->GetName(),
                         scale[q], bad, (unsigned long long)Fnv1a(p, (size_t)o.n * sizeof(float)));
                json += buf;
                ok = ok && bad == 0;
                r->Unlock();
                r->Release();
            }
            g->Stop();
        }
    }
    json += std::string(", \"ok\": ") + (ok ? "true" : "false") + "}";
    printf("%s\n", json.c_str());
    g->Teardown();
    Graph::DestroyGraph(g);
    Runtime::Terminate();
    return ok ? 0 : 1;
}
