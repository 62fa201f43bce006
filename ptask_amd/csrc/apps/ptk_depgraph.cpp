// ptk_depgraph.cpp -- host tasks with a dependent HIP accelerator (Task::BindDependentAcceleratorClass,
// Port::BindDependentAccelerator; hosttask.h), as the reference's cuBLAS tests build them.
//
// --scenario sgemm: the graph of ptaskUnitTest/ptaskcublas.cpp run_cublas_task (:363-490): one
//   host task SGemm (EX form, platform objects requested) with inputs A, B, C(in/out) bound to
//   the dependent accelerator and a sticky SGEMMPARMS port; C = A B (column-major, m x k times
//   k x n) through rocBLAS on the task's dependent stream.  --repeat R pushes R independent
//   problems through the running graph.
// --scenario scale: a legacy-form dependent host task ScaleDeviceVector (A in/out on the device,
//   a sticky scalar, n) iterated by an MF_SIMPLE_ITERATOR meta port, its output consumed by a
//   second, plain host task (ScaleVector, host pointers), so the block moves device -> host.
//
// --kernels names the dependent host-task library (libptk_hosttasks.so), --host-kernels the plain
// host-task library for the scale scenario.  Prints one JSON object with the largest relative
// error against a binary64 host computation.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ptaskapi.h"

using namespace PTask;

namespace {

typedef struct _sgemm_parms_t {
    int ax, ay, bx, by, cx, cy;
} SGEMMPARMS;

float Rnd(unsigned long long &s) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return (float)((double)(s >> 40) / (double)(1ull << 24)) - 0.5f;
}

int Sgemm(const std::string &lib, int m, int k, int n, int repeat) {
    const UINT stride = sizeof(float);
    Graph *g = new Graph();
    DatablockTemplate *tA = Runtime::GetDatablockTemplate((char *)"dbA_float", stride, (UINT)m, (UINT)k, 1);
    DatablockTemplate *tB = Runtime::GetDatablockTemplate((char *)"dbB_float", stride, (UINT)k, (UINT)n, 1);
    DatablockTemplate *tC = Runtime::GetDatablockTemplate((char *)"dbC_float", stride, (UINT)m, (UINT)n, 1);
    DatablockTemplate *tP = Runtime::GetDatablockTemplate((char *)"sgemmparms", sizeof(SGEMMPARMS), 1, 1, 1);
    CompiledKernel *kern = Runtime::GetCompiledKernel((char *)lib.c_str(), (char *)"SGemm");
    if (!kern) {
        fprintf(stderr, "GetCompiledKernel SGemm failed: %s\n", Runtime::GetLastErrorText());
        return 1;
    }
    Port *in[4], *out[1];
    UINT uid = 0;
    in[0] = Runtime::CreatePort(INPUT_PORT, tA, uid++, (char *)"A", 0);
    in[1] = Runtime::CreatePort(INPUT_PORT, tB, uid++, (char *)"B", 1);
    in[2] = Runtime::CreatePort(INPUT_PORT, tC, uid++, (char *)"C(in)", 2, 0);
    in[3] = Runtime::CreatePort(STICKY_PORT, tP, uid++, (char *)"parms", 3);
    out[0] = Runtime::CreatePort(OUTPUT_PORT, tC, uid++, (char *)"C(out)", 2);
    for (int i = 0; i < 3; ++i) in[i]->BindDependentAccelerator(ACCELERATOR_CLASS_CUDA, 0);
    out[0]->BindDependentAccelerator(ACCELERATOR_CLASS_CUDA, 0);
    Task *t = g->AddTask(kern, 4, in, 1, out, (char *)"AxBTask");
    if (!t || !t->BindDependentAcceleratorClass(ACCELERATOR_CLASS_CUDA, 1, TRUE)) {
        fprintf(stderr, "dependent binding failed: %s\n", Runtime::GetLastErrorText());
        return 1;
    }
    GraphInputChannel *cA = g->AddInputChannel(in[0], (char *)"AInputChannel");
    GraphInputChannel *cB = g->AddInputChannel(in[1], (char *)"BInputChannel");
    GraphInputChannel *cC = g->AddInputChannel(in[2], (char *)"CInputChannel");
    GraphInputChannel *cP = g->AddInputChannel(in[3], (char *)"ABConstChannel");
    GraphOutputChannel *cOut = g->AddOutputChannel(out[0], (char *)"outputChannel");
    g->Run();
    SGEMMPARMS prm = {m, k, k, n, m, n};
    Datablock *bP = Runtime::AllocateDatablock(tP, &prm, sizeof(prm), cP);
    cP->Push(bP);
    bP->Release();
    double maxRel = 0;
    unsigned long long seed = 12345;
    for (int r = 0; r < repeat; ++r) {
        std::vector<float> A((size_t)m * k), B((size_t)k * n), C((size_t)m * n, 0.f);
        for (float &v : A) v = Rnd(seed);
        for (float &v : B) v = Rnd(seed);
        Datablock *bA = Runtime::AllocateDatablock(tA, A.data(), stride * (UINT)A.size(), cA);
        Datablock *bB = Runtime::AllocateDatablock(tB, B.data(), stride * (UINT)B.size(), cB);
        Datablock *bC = Runtime::AllocateDatablock(tC, C.data(), stride * (UINT)C.size(), cC);
        cA->Push(bA);
        cB->Push(bB);
        cC->Push(bC);
        bA->Release();
        bB->Release();
        bC->Release();
        Datablock *res = cOut->Pull();
        if (!res) return 1;
        res->Lock();
        const float *p = (const float *)res->GetDataPointer(FALSE);
        double scale = 0, err = 0;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i) {
                double s = 0;
                for (int q = 0; q < k; ++q) s += (double)A[(size_t)q * m + i] * (double)B[(size_t)j * k + q];
                scale = std::max(scale, std::fabs(s));
                err = std::max(err, std::fabs(s - (double)p[(size_t)j * m + i]));
            }
        res->Unlock();
        res->Release();
        maxRel = std::max(maxRel, err / std::max(scale, 1e-30));
    }
    g->Stop();
    g->Teardown();
    Graph::DestroyGraph(g);
    printf("{\"ok\": %s, \"scenario\": \"sgemm\", \"m\": %d, \"k\": %d, \"n\": %d, \"repeat\": %d, \"max_rel_err\": %.3g}\n",
           maxRel <= 1e-5 ? "true" : "false", m, k, n, repeat, maxRel);
    return maxRel <= 1e-5 ? 0 : 1;
}

int Scale(const std::string &lib, const std::string &hostLib, int n, int iterations, float s) {
    const UINT stride = sizeof(float);
    Graph *g = new Graph();
    DatablockTemplate *tV = Runtime::GetDatablockTemplate((char *)"dbV_float", stride, (UINT)n, 1, 1);
    DatablockTemplate *tS = Runtime::GetDatablockTemplate((char *)"vscale", stride, PTPARM_FLOAT);
    DatablockTemplate *tN = Runtime::GetDatablockTemplate((char *)"vlen", sizeof(int), PTPARM_INT);
    DatablockTemplate *tI = Runtime::GetDatablockTemplate((char *)"viter", sizeof(int), PTPARM_INT);
    CompiledKernel *kDev = Runtime::GetCompiledKernel((char *)lib.c_str(), (char *)"ScaleDeviceVector");
    CompiledKernel *kHost = Runtime::GetCompiledKernel((char *)hostLib.c_str(), (char *)"ScaleVector");
    if (!kDev || !kHost) {
        fprintf(stderr, "GetCompiledKernel failed: %s\n", Runtime::GetLastErrorText());
        return 1;
    }
    UINT uid = 0;
    Port *in1[4] = {Runtime::CreatePort(INPUT_PORT, tV, uid++, (char *)"A", 0, 0),
                    Runtime::CreatePort(STICKY_PORT, tS, uid++, (char *)"s", 1),
                    Runtime::CreatePort(STICKY_PORT, tN, uid++, (char *)"n", 2),
                    Runtime::CreatePort(META_PORT, tI, uid++, (char *)"iterations", 3)};
    Port *out1[1] = {Runtime::CreatePort(OUTPUT_PORT, tV, uid++, (char *)"A(out)", 0)};
    in1[0]->BindDependentAccelerator(ACCELERATOR_CLASS_HIP, 0);
    out1[0]->BindDependentAccelerator(ACCELERATOR_CLASS_HIP, 0);
    in1[3]->SetMetaFunction(MF_SIMPLE_ITERATOR);
    Task *t1 = g->AddTask(kDev, 4, in1, 1, out1, (char *)"DeviceScale");
    if (!t1 || !t1->BindDependentAcceleratorClass(ACCELERATOR_CLASS_HIP, 1, FALSE)) {
        fprintf(stderr, "dependent binding failed: %s\n", Runtime::GetLastErrorText());
        return 1;
    }
    // ScaleVector(float* A, float s, int n, float* out): out = s * A, on the host
    Port *in2[3] = {Runtime::CreatePort(INPUT_PORT, tV, uid++, (char *)"A2", 0),
                    Runtime::CreatePort(STICKY_PORT, tS, uid++, (char *)"s2", 1),
                    Runtime::CreatePort(STICKY_PORT, tN, uid++, (char *)"n2", 2)};
    Port *out2[1] = {Runtime::CreatePort(OUTPUT_PORT, tV, uid++, (char *)"out", 3)};
    Task *t2 = g->AddTask(kHost, 3, in2, 1, out2, (char *)"HostScale");
    if (!t2) return 1;
    GraphInputChannel *cA = g->AddInputChannel(in1[0], (char *)"A");
    GraphInputChannel *cS = g->AddInputChannel(in1[1], (char *)"s");
    GraphInputChannel *cN = g->AddInputChannel(in1[2], (char *)"n");
    GraphInputChannel *cI = g->AddInputChannel(in1[3], (char *)"iterations");
    GraphInputChannel *cS2 = g->AddInputChannel(in2[1], (char *)"s2");
    GraphInputChannel *cN2 = g->AddInputChannel(in2[2], (char *)"n2");
    g->AddInternalChannel(out1[0], in2[0], (char *)"device->host");
    GraphOutputChannel *cOut = g->AddOutputChannel(out2[0], (char *)"out");
    g->Run();
    const float one = 1.0f;
    auto push = [&](GraphInputChannel *c, DatablockTemplate *tt, const void *v, UINT cb) {
        Datablock *b = Runtime::AllocateDatablock(tt, (void *)v, cb, c);
        c->Push(b);
        b->Release();
    };
    push(cS, tS, &s, sizeof(s));
    push(cN, tN, &n, sizeof(n));
    push(cS2, tS, &one, sizeof(one));
    push(cN2, tN, &n, sizeof(n));
    std::vector<float> A((size_t)n);
    unsigned long long seed = 99;
    for (float &v : A) v = Rnd(seed);
    push(cA, tV, A.data(), stride * (UINT)n);
    push(cI, tI, &iterations, sizeof(iterations));
    Datablock *res = cOut->Pull();
    if (!res) return 1;
    res->Lock();
    const float *p = (const float *)res->GetDataPointer(FALSE);
    double maxRel = 0;
    for (int i = 0; i < n; ++i) {
        float w = A[(size_t)i];
        for (int it = 0; it < iterations; ++it) w *= s;
        maxRel = std::max(maxRel, std::fabs((double)p[i] - (double)w) / std::max(1e-30, std::fabs((double)w)));
    }
    res->Unlock();
    res->Release();
    g->Stop();
    g->Teardown();
    Graph::DestroyGraph(g);
    printf("{\"ok\": %s, \"scenario\": \"scale\", \"n\": %d, \"iterations\": %d, \"max_rel_err\": %.3g}\n",
           maxRel == 0 ? "true" : "false", n, iterations, maxRel);
    return maxRel == 0 ? 0 : 1;
}

}  // namespace

int main(int argc, char **argv) {
    std::string lib, hostLib, scenario = "sgemm";
    int m = 64, k = 48, n = 32, repeat = 2, elems = 4096, iterations = 3;
    float s = 1.5f;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> const char * {
            if (i + 1 >= argc) { fprintf(stderr, "missing value for %s\n", a.c_str()); exit(2); }
            return argv[++i];
        };
        if (a == "--kernels") lib = next();
        else if (a == "--host-kernels") hostLib = next();
        else if (a == "--scenario") scenario = next();
        else if (a == "--m") m = atoi(next());
        else if (a == "--k") k = atoi(next());
        else if (a == "--n") n = atoi(next());
        else if (a == "--repeat") repeat = atoi(next());
        else if (a == "--elems") elems = atoi(next());
        else if (a == "--iterations") iterations = atoi(next());
        else if (a == "--scalar") s = (float)atof(next());
        else {
            fprintf(stderr, "usage: ptk_depgraph --kernels LIB [--scenario sgemm|scale] [--host-kernels LIB] ...\n");
            return 2;
        }
    }
    if (lib.empty() || (scenario == "scale" && hostLib.empty()) || m < 1 || k < 1 || n < 1 || elems < 1) {
        fprintf(stderr, "usage: ptk_depgraph --kernels LIB [--scenario sgemm|scale] [--host-kernels LIB] ...\n");
        return 2;
    }
    Runtime::SetExitOnRuntimeFailure(FALSE);
    if (PTFAILED(Runtime::Initialize())) return 1;
    const int rc = scenario == "sgemm" ? Sgemm(lib, m, k, n, repeat) : Scale(lib, hostLib, elems, iterations, s);
    Runtime::Terminate();
    return rc;
}
