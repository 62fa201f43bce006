// ptk_iteration.cpp -- the simple-iteration graph of ptaskUnitTest/iteration.cpp
// (run_simple_iteration_task, :67-180) built with the Runtime:: helper calls the reference test
// uses: one task scales a vector in place, its MF_SIMPLE_ITERATOR meta port repeats the dispatch
// --iterations times, the element count reaches the kernel through a port derived from the
// vector (BindDerivedPort, DF_SIZE), and the graph is run, fed and drained through
// Runtime::RunGraph / Push / Pull.
//
// --kernels names a host-task library with ScaleInPlace(float* A, float s, int n) or a gfx950
// code object with a kernel of that signature.  Prints one JSON object with the result vector's
// checksum and the largest relative difference from A * s^iterations computed on the host.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ptaskapi.h"

using namespace PTask;

int main(int argc, char **argv) {
    std::string kernels, op = "ScaleInPlace";
    int n = 1000, iterations = 4;
    float scalar = 2.0f;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> const char * {
            if (i + 1 >= argc) { fprintf(stderr, "missing value for %s\n", a.c_str()); exit(2); }
            return argv[++i];
        };
        if (a == "--kernels") kernels = next();
        else if (a == "--op") op = next();
        else if (a == "--n") n = atoi(next());
        else if (a == "--iterations") iterations = atoi(next());
        else if (a == "--scalar") scalar = (float)atof(next());
        else {
            fprintf(stderr, "usage: ptk_iteration --kernels FILE [--op NAME] [--n N] [--iterations I] [--scalar S]\n");
            return 2;
        }
    }
    if (kernels.empty() || n < 1 || iterations < 1) {
        fprintf(stderr, "usage: ptk_iteration --kernels FILE [--op NAME] [--n N] [--iterations I] [--scalar S]\n");
        return 2;
    }
    Runtime::SetExitOnRuntimeFailure(FALSE);
    if (PTFAILED(Runtime::Initialize())) return 1;
    std::vector<float> A((size_t)n), want((size_t)n);
    for (int i = 0; i < n; ++i) A[(size_t)i] = (float)((i * 37) % 101) / 101.0f - 0.5f;
    for (int i = 0; i < n; ++i) {
        float v = A[(size_t)i];
        for (int it = 0; it < iterations; ++it) v = scalar * v;
        want[(size_t)i] = v;
    }
    const UINT stride = sizeof(float);
    Graph *pGraph = new Graph();
    DatablockTemplate *pDataTemplate = Runtime::GetDatablockTemplate((char *)"dbV1_float", stride, (UINT)n, 1, 1);
    DatablockTemplate *pScaleTemplate = Runtime::GetDatablockTemplate((char *)"vscale", stride, PTPARM_FLOAT);
    DatablockTemplate *pParmTemplate = Runtime::GetDatablockTemplate((char *)"vecdims", sizeof(int), PTPARM_INT);
    DatablockTemplate *pIterTemplate = Runtime::GetDatablockTemplate((char *)"dbiter", sizeof(int), PTPARM_INT);
    CompiledKernel *pKernel = Runtime::GetCompiledKernel((char *)kernels.c_str(), (char *)op.c_str());
    if (!pKernel) {
        fprintf(stderr, "GetCompiledKernel %s:%s failed: %s\n", kernels.c_str(), op.c_str(), Runtime::GetLastErrorText());
        return 1;
    }
    Port *in[4], *out[1];
    UINT uid = 0;
    in[0] = Runtime::CreatePort(INPUT_PORT, pDataTemplate, uid++, (char *)"A_top", 0, 0);
    in[1] = Runtime::CreatePort(STICKY_PORT, pScaleTemplate, uid++, (char *)"scalar_top", 1);
    in[2] = Runtime::CreatePort(STICKY_PORT, pParmTemplate, uid++, (char *)"N_top", 2);
    in[3] = Runtime::CreatePort(META_PORT, pIterTemplate, uid++, (char *)"iter_meta", 3);
    out[0] = Runtime::CreatePort(OUTPUT_PORT, pDataTemplate, uid++, (char *)"A(out)_top", 0);
    Runtime::SetPortMetaFunction(in[3], MF_SIMPLE_ITERATOR);
    Task *pTop = Runtime::AddTask(pGraph, pKernel, 4, in, 1, out, (char *)"top");
    if (!pTop) {
        fprintf(stderr, "AddTask failed: %s\n", Runtime::GetLastErrorText());
        return 1;
    }
    pTop->SetComputeGeometry(1, 1, 1);
    pTop->SetBlockAndGridSize(PTASKDIM3((n + 255) / 256, 1, 1), PTASKDIM3(256, 1, 1));
    GraphInputChannel *pAIn = pGraph->AddInputChannel(in[0], (char *)"AInputChannel");
    GraphInputChannel *pScaleIn = pGraph->AddInputChannel(in[1], (char *)"ScalarChannel");
    GraphInputChannel *pIterIn = pGraph->AddInputChannel(in[3], (char *)"IterationChannel");
    GraphOutputChannel *pAOut = pGraph->AddOutputChannel(out[0], (char *)"outputChannel_left_A");
    Runtime::BindDerivedPort(pGraph, in[0], in[2]);
    const PTRESULT sem = Runtime::CheckGraphSemantics(pGraph, FALSE, FALSE);
    Runtime::RunGraph(pGraph);
    Datablock *pScale = Runtime::AllocateDatablock(pScaleTemplate, &scalar, sizeof(scalar), pScaleIn);
    pScaleIn->Push(pScale);
    Datablock *pA = Runtime::AllocateDatablock(pDataTemplate, A.data(), stride * (UINT)n, pAIn);
    Datablock *pIters = Runtime::AllocateDatablock(pIterTemplate, &iterations, sizeof(iterations), pIterIn);
    pA->Lock();
    pA->SetRecordCount((UINT)n);
    pA->Unlock();
    Runtime::Push(pAIn, pA);
    Runtime::Push(pIterIn, pIters);
    pA->Release();
    pIters->Release();
    pScale->Release();
    Datablock *pOut = Runtime::Pull(pAOut);
    double maxRel = 0;
    unsigned long long h = 1469598103934665603ull;
    bool ok = pOut != nullptr;
    if (pOut) {
        pOut->Lock();
        const float *p = (const float *)pOut->GetDataPointer(FALSE);
        for (int i = 0; i < n; ++i) {
            const double d = std::fabs((double)p[i] - (double)want[(size_t)i]);
            const double r = d / std::max(1e-30, std::fabs((double)want[(size_t)i]));
            maxRel = std::max(maxRel, r);
            unsigned u;
            memcpy(&u, &p[i], 4);
            h = (h ^ u) * 1099511628211ull;
        }
        pOut->Unlock();
        pOut->Release();
    }
    const UINT phys = Runtime::GetPhysicalAcceleratorCount();
    const BOOL inv = Runtime::CheckInvariantCondition(ok, (char *)"the output was pulled");
    pGraph->Stop();
    pGraph->Teardown();
    Graph::DestroyGraph(pGraph);
    Runtime::Terminate();
    printf("{\"ok\": %s, \"n\": %d, \"iterations\": %d, \"max_rel_diff\": %.3g, \"fnv1a\": \"%016llx\", "
           "\"semantics\": %ld, \"physical_accelerators\": %u, \"invariant\": %s}\n",
           ok && maxRel == 0 ? "true" : "false", n, iterations, maxRel, h, (long)sem, phys, inv ? "true" : "false");
    return ok && maxRel == 0 ? 0 : 1;
}
