// ptk_pfxsum.cpp -- the reference's prefix-sum graph (ptaskUnitTest/cpfxsum.cpp:174-299,
// run_cpfxsum_task) built against ptaskapi.h with the same call sequence: one task bound to
// `pfxsum` with ports pin (formal 0), a sticky int N (formal 2) and pout (formal 1), 256-thread
// blocks on a grid of 16 blocks, the input pushed with PT_ACCESS_HOST_WRITE |
// PT_ACCESS_ACCELERATOR_READ, and the output pulled and checked against the exclusive prefix sum
// computed on the host (host_pfxsum, pfxsum.cpp:56-69).
//
// --kernels names scan.hsaco (the HIP kernel) or a host-task library exporting pfxsum.  The input
// is --data FILE (int32, as the fixtures of cpfxsum.cpp / pfxsum.cpp are stored) or --n N seeded
// values.  Prints one JSON object; --dump writes the last run's output (int32).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ptaskapi.h"

using namespace PTask;

int main(int argc, char **argv) {
    std::string kernels, data, dump;
    int n = 24, iterations = 1;
    unsigned long long seed = 1;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> const char * {
            if (i + 1 >= argc) {
                fprintf(stderr, "missing value for %s\n", a.c_str());
                exit(2);
            }
            return argv[++i];
        };
        if (a == "--kernels") kernels = next();
        else if (a == "--data") data = next();
        else if (a == "--n") n = atoi(next());
        else if (a == "--iterations") iterations = atoi(next());
        else if (a == "--seed") seed = strtoull(next(), nullptr, 0);
        else if (a == "--dump") dump = next();
        else {
            kernels.clear();
            break;
        }
    }
    std::vector<int> in;
    if (!data.empty()) {
        FILE *f = fopen(data.c_str(), "rb");
        if (!f) {
            fprintf(stderr, "cannot open %s\n", data.c_str());
            return 2;
        }
        int v;
        while (fread(&v, sizeof(int), 1, f) == 1) in.push_back(v);
        fclose(f);
        n = (int)in.size();
    } else if (n >= 1) {
        unsigned long long s = seed * 0x9E3779B97F4A7C15ull + 1;
        in.resize(n);
        for (int &v : in) {
            s = s * 6364136223846793005ull + 1442695040888963407ull;
            v = (int)(s >> 33) % 2001 - 1000;
        }
    }
    if (kernels.empty() || n < 1 || iterations < 1) {
        fprintf(stderr, "usage: ptk_pfxsum --kernels FILE [--data int32-file | --n N [--seed S]] [--iterations I] [--dump file]\n");
        return 2;
    }
    Runtime::SetExitOnRuntimeFailure(FALSE);
    if (PTFAILED(Runtime::Initialize())) return 1;
    const bool hip = ptaskutils::SelectAcceleratorClass(kernels.c_str()) == ACCELERATOR_CLASS_HIP;
    Graph *g = new Graph();
    DatablockTemplate *tData = Runtime::GetDatablockTemplate((char *)"dbpfxdata", sizeof(int), (UINT)n, 1, 1);
    DatablockTemplate *tParm = Runtime::GetDatablockTemplate((char *)"parms", sizeof(int), PTPARM_INT);
    CompiledKernel *k = Runtime::GetCompiledKernel((char *)kernels.c_str(), (char *)"pfxsum");
    if (!k) {
        fprintf(stderr, "GetCompiledKernel pfxsum: %s\n", Runtime::GetLastErrorText());
        return 1;
    }
    UINT uid = 0;
    Port *ins[2] = {Runtime::CreatePort(INPUT_PORT, tData, uid++, (char *)"pin", 0),
                    Runtime::CreatePort(STICKY_PORT, tParm, uid++, (char *)"N", 2)};
    Port *outs[1] = {Runtime::CreatePort(OUTPUT_PORT, tData, uid++, (char *)"pout", 1)};
    Task *t = g->AddTask(k, 2, ins, 1, outs, "PfxsumTask");
    if (!t) return 1;
    if (hip) t->SetBlockAndGridSize(PTASKDIM3(16, 1, 1), PTASKDIM3(256, 1, 1));  // cpfxsum.cpp:219
    else t->SetComputeGeometry(1, 1, 1);
    GraphInputChannel *pIn = g->AddInputChannel(ins[0], "DataInputChannel");
    GraphInputChannel *pParms = g->AddInputChannel(ins[1], "ParmsInputChannel");
    GraphOutputChannel *pOut = g->AddOutputChannel(outs[0], "DataOutputChannel");
    g->Run();
    std::vector<int> want(n), got(n, 0);
    unsigned run = 0;  // host_pfxsum in two's-complement arithmetic
    for (int i = 0; i < n; ++i) {
        want[i] = (int)run;
        run += (unsigned)in[i];
    }
    bool ok = true;
    int mismatches = 0;
    for (int it = 0; it < iterations && ok; ++it) {
        BUFFERACCESSFLAGS raw = PT_ACCESS_HOST_WRITE | PT_ACCESS_ACCELERATOR_READ;
        Datablock *db = Runtime::AllocateDatablock(tData, in.data(), (UINT)(sizeof(int) * n), pIn, raw);
        Datablock *pdb = Runtime::AllocateDatablock(tParm, &n, sizeof(n), pParms);
        pIn->Push(db);
        pParms->Push(pdb);
        db->Release();
        pdb->Release();
        Datablock *o = pOut->Pull(600000);
        if (!o) {
            fprintf(stderr, "no output: %s\n", Runtime::GetLastErrorText());
            ok = false;
            break;
        }
        o->Lock();
        memcpy(got.data(), o->GetDataPointer(FALSE), sizeof(int) * (size_t)n);
        o->Unlock();
        o->Release();
        for (int i = 0; i < n; ++i) mismatches += got[i] != want[i];
    }
    g->Stop();
    g->Teardown();
    Graph::DestroyGraph(g);
    Runtime::Terminate();
    if (!dump.empty()) {
        FILE *f = fopen(dump.c_str(), "wb");
        if (!f) return 1;
        fwrite(got.data(), sizeof(int), got.size(), f);
        fclose(f);
    }
    const bool succeeded = ok && mismatches == 0;
    printf("{\"ok\": %s, \"backend\": \"%s\", \"n\": %d, \"iterations\": %d, \"mismatches\": %d, \"succeeded\": %s}\n",
           ok ? "true" : "false", hip ? "hip" : "host", n, iterations, mismatches, succeeded ? "true" : "false");
    return succeeded ? 0 : 1;
}
