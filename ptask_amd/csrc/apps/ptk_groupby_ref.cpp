// ptk_groupby_ref.cpp -- the reference's groupby graph (ptaskUnitTest/graphgroupby.cpp:56-387 graph
// construction, :421-479 runGraph, :394-419 testResults, :541-573 run_graph_groupby_task) built
// against ptaskapi.h with the same task, port and channel sequence: countUniqueKeysIntInt ->
// allocateCountArray -> keyCountIntInt -> singleBlockPrefixSumIntAdd -> reshuffleIntInt, then
// destroyHashTableInt and deallocateCountArray, the same templates (a 256-pointer hash table, a
// 10-int key count, one pointer for the count array), formal indices, in/out routing,
// non-marshallable hash-table ports, priorities 10 on the clean-up tasks, mandatory affinity to
// the preferred accelerator and 512-thread geometry.
//
// --kernels names groupby_ref.hsaco (the HIP kernels) or the oracle's host library (their
// host-task forms).  Keys: --dist reference draws (int)(rand()/(float)RAND_MAX * numKeys) from
// the MSVC rand() sequence seeded 0x4fd478 as the reference does (graphgroupby.cpp:428,555);
// --dist wide draws rows from `keys` values spread over the whole int32 range (negative ones
// included) with a 64-bit LCG, --dist distinct gives every row its own key.  Prints one JSON object with the reference check's
// verdict; --dump writes the keys and the last run's output (int32 each).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <string>
#include <vector>

#include "ptaskapi.h"

using namespace PTask;

namespace {

const int kHashTableSize = 256, kBlockSize = 512;

struct Params {
    Graph *g = nullptr;
    int inputLength = 0, numKeys = 0, numIterations = 1;
    std::string kernels;
    bool hip = false;
    DatablockTemplate *hashTableTemplate, *inputTemplate, *keycountTemplate, *ptrTemplate, *intTemplate;
    UINT accelerator = 0;
    UINT uid = 0;
    std::vector<GraphInputChannel *> inputChannels, inputLengthChannels, hashtableChannels;
    GraphInputChannel *keyCountInChannel = nullptr;
    GraphOutputChannel *outputChannel = nullptr;
    bool ok = true;
};

void Geometry(Params &p, Task *t, int n) {
    if (!p.hip) {
        t->SetComputeGeometry(1, 1, 1);  // a host task handles the whole operation per call
        return;
    }
    t->SetBlockAndGridSize(PTASKDIM3((n + kBlockSize - 1) / kBlockSize, 1, 1), PTASKDIM3(kBlockSize, 1, 1));
}

void SetupTask(Params &p, Task *t, int geometryN, int priority = 0) {
    if (!t) {
        p.ok = false;
        return;
    }
    if (priority) t->SetPriority(priority);
    if (PTFAILED(Runtime::SetTaskAffinity(t, p.accelerator, AFFINITYTYPE_MANDATORY))) p.ok = false;
    if (geometryN > 0) {
        Geometry(p, t, geometryN);
    } else if (p.hip) {  // the clean-up tasks: one block
        t->SetBlockAndGridSize(PTASKDIM3(1, 1, 1), PTASKDIM3(geometryN == 0 ? 1 : kHashTableSize, 1, 1));
    } else {
        t->SetComputeGeometry(1, 1, 1);
    }
}

CompiledKernel *Kernel(Params &p, const char *name) {
    CompiledKernel *k = Runtime::GetCompiledKernel((char *)p.kernels.c_str(), (char *)name);
    if (!k) {
        fprintf(stderr, "GetCompiledKernel %s: %s\n", name, Runtime::GetLastErrorText());
        p.ok = false;
    }
    return k;
}

// graphgroupby.cpp:76-104
void CountDeallocator(Port *globalCountPort, Params &p) {
    Port *in[1] = {Runtime::CreatePort(INPUT_PORT, p.ptrTemplate, p.uid++, (char *)"globalCount(in)", 0)};
    Task *t = p.g->AddTask(Kernel(p, "deallocateCountArray"), 1, in, 0, nullptr, "deallocateCountArray");
    SetupTask(p, t, 0, 10);
    p.g->AddInternalChannel(globalCountPort, in[0], "globalCount_deallocateGlobalCount");
}

// graphgroupby.cpp:106-144
void CleanUpNodes(Port *hashTablePort, Port *globalCountPort, Params &p) {
    Port *in[1] = {Runtime::CreatePort(INPUT_PORT, p.hashTableTemplate, p.uid++, (char *)"hashTable(in)", 0)};
    in[0]->SetMarshallable(FALSE);
    Task *t = p.g->AddTask(Kernel(p, "destroyHashTableInt"), 1, in, 0, nullptr, "destroyHashTableInt");
    SetupTask(p, t, -1, 10);
    p.g->AddInternalChannel(hashTablePort, in[0], "hashTable_hashTableDestroy");
    CountDeallocator(globalCountPort, p);
}

// graphgroupby.cpp:147-201
void ShuffleNode(Port *globalCountPort, Port *hashTablePort, Params &p) {
    Port *in[4], *out[3];
    in[0] = Runtime::CreatePort(INPUT_PORT, p.inputTemplate, p.uid++, (char *)"in", 0);
    in[1] = Runtime::CreatePort(STICKY_PORT, p.intTemplate, p.uid++, (char *)"n", 1);
    in[2] = Runtime::CreatePort(INPUT_PORT, p.hashTableTemplate, p.uid++, (char *)"hashTable(in)", 3, 1);
    in[3] = Runtime::CreatePort(INPUT_PORT, p.ptrTemplate, p.uid++, (char *)"globalCount(in)", 4, 2);
    out[0] = Runtime::CreatePort(OUTPUT_PORT, p.inputTemplate, p.uid++, (char *)"out", 2);
    out[1] = Runtime::CreatePort(OUTPUT_PORT, p.hashTableTemplate, p.uid++, (char *)"hashTable(out)", 3);
    out[2] = Runtime::CreatePort(OUTPUT_PORT, p.ptrTemplate, p.uid++, (char *)"globalCount(out)", 4);
    out[1]->SetMarshallable(FALSE);
    in[2]->SetMarshallable(FALSE);
    Task *t = p.g->AddTask(Kernel(p, "reshuffleIntInt"), 4, in, 3, out, "reshuffleIntInt");
    SetupTask(p, t, p.inputLength);
    p.g->AddInternalChannel(globalCountPort, in[3], "globalCount_shuffle");
    p.inputChannels.push_back(p.g->AddInputChannel(in[0], "inChannel_shuffle"));
    p.inputLengthChannels.push_back(p.g->AddInputChannel(in[1], "nChannel_shuffle"));
    p.g->AddInternalChannel(hashTablePort, in[2], "hashTableChannel_shuffle");
    p.outputChannel = p.g->AddOutputChannel(out[0], "OutChannel_shuffle");
    CleanUpNodes(out[1], out[2], p);
}

// graphgroupby.cpp:204-242
void PrefixSum(Port *globalCountPort, Port *keyCountPort, Port *hashTablePort, Params &p) {
    Port *in[2], *out[1];
    in[0] = Runtime::CreatePort(INPUT_PORT, p.ptrTemplate, p.uid++, (char *)"globalCount(in)", 0, 0);
    in[1] = Runtime::CreatePort(INPUT_PORT, p.keycountTemplate, p.uid++, (char *)"keyCount(in)", 1);
    out[0] = Runtime::CreatePort(OUTPUT_PORT, p.ptrTemplate, p.uid++, (char *)"globalCount(out)", 0);
    Task *t = p.g->AddTask(Kernel(p, "singleBlockPrefixSumIntAdd"), 2, in, 1, out, "singleBlockPrefixSumIntAdd");
    SetupTask(p, t, 512);
    p.g->AddInternalChannel(keyCountPort, in[1], "keyCountChannel_prefixSum");
    p.g->AddInternalChannel(globalCountPort, in[0], "globalCountChanel_prefixSum");
    ShuffleNode(out[0], hashTablePort, p);
}

// graphgroupby.cpp:245-295
void KeyCountKernel(Port *countArray, Port *keyCountPort, Port *hashTablePort, Params &p) {
    Port *in[4], *out[1];
    in[0] = Runtime::CreatePort(INPUT_PORT, p.inputTemplate, p.uid++, (char *)"in", 0);
    in[1] = Runtime::CreatePort(STICKY_PORT, p.intTemplate, p.uid++, (char *)"n", 1);
    in[2] = Runtime::CreatePort(INPUT_PORT, p.hashTableTemplate, p.uid++, (char *)"hashTable", 2);
    in[3] = Runtime::CreatePort(INPUT_PORT, p.ptrTemplate, p.uid++, (char *)"globalCount(in)", 3, 0);
    in[2]->SetMarshallable(FALSE);
    out[0] = Runtime::CreatePort(OUTPUT_PORT, p.ptrTemplate, p.uid++, (char *)"globalCount(out)", 3);
    Task *t = p.g->AddTask(Kernel(p, "keyCountIntInt"), 4, in, 1, out, "keyCountIntInt");
    SetupTask(p, t, p.inputLength);
    p.g->AddInternalChannel(countArray, in[3], "globalCountChannel");
    p.inputChannels.push_back(p.g->AddInputChannel(in[0], "inChannel_keyCount"));
    p.inputLengthChannels.push_back(p.g->AddInputChannel(in[1], "nChannel_keyCount"));
    p.g->AddInternalChannel(hashTablePort, in[2], "hashTableChannel_keyCount");
    PrefixSum(out[0], keyCountPort, hashTablePort, p);
}

// graphgroupby.cpp:297-330
void AllocKernel(Port *keyCountPort, Port *hashTablePort, Params &p) {
    Port *in[1], *out[1];
    in[0] = Runtime::CreatePort(INPUT_PORT, p.keycountTemplate, p.uid++, (char *)"keyCount(in)", 1);
    out[0] = Runtime::CreatePort(OUTPUT_PORT, p.ptrTemplate, p.uid++, (char *)"countArray(out)", 0);
    Task *t = p.g->AddTask(Kernel(p, "allocateCountArray"), 1, in, 1, out, "allocateCountArray");
    SetupTask(p, t, 512);
    p.g->AddInternalChannel(keyCountPort, in[0], "KeyCountChannel");
    KeyCountKernel(out[0], keyCountPort, hashTablePort, p);
}

// graphgroupby.cpp:332-387
void UniqueKeyCountNode(Params &p) {
    Port *in[4], *out[2];
    in[0] = Runtime::CreatePort(INPUT_PORT, p.inputTemplate, p.uid++, (char *)"in", 1);
    in[1] = Runtime::CreatePort(STICKY_PORT, p.intTemplate, p.uid++, (char *)"n", 3);
    in[2] = Runtime::CreatePort(INPUT_PORT, p.hashTableTemplate, p.uid++, (char *)"hashTable(in)", 2, 1);
    in[3] = Runtime::CreatePort(INPUT_PORT, p.keycountTemplate, p.uid++, (char *)"keyCount(in)", 0, 0);
    out[0] = Runtime::CreatePort(OUTPUT_PORT, p.keycountTemplate, p.uid++, (char *)"keyCount(out)", 0);
    out[1] = Runtime::CreatePort(OUTPUT_PORT, p.hashTableTemplate, p.uid++, (char *)"hashTable(out)", 2);
    out[1]->SetMarshallable(FALSE);
    in[2]->SetMarshallable(FALSE);
    Task *t = p.g->AddTask(Kernel(p, "countUniqueKeysIntInt"), 4, in, 2, out, "countUniqueKeysIntInt");
    SetupTask(p, t, p.inputLength);
    p.inputChannels.push_back(p.g->AddInputChannel(in[0], "countUniqueKeysIn"));
    p.inputLengthChannels.push_back(p.g->AddInputChannel(in[1], "countUniqueKeysn"));
    p.hashtableChannels.push_back(p.g->AddInputChannel(in[2], "countUniqueKeysHashTable"));
    p.keyCountInChannel = p.g->AddInputChannel(in[3], "countUniqueKeysKeyCountIn");
    AllocKernel(out[0], out[1], p);
}

// The first enabled accelerator of the kernels' class with the most cores (graphgroupby.cpp:488-539
// prefers, in order, core count, concurrent kernels, memory, clock and platform index).
UINT PreferredAccelerator(ACCELERATOR_CLASS cls) {
    ACCELERATOR_DESCRIPTOR *best = nullptr, *d = nullptr;
    std::vector<ACCELERATOR_DESCRIPTOR *> all;
    for (UINT i = 0; Runtime::EnumerateAccelerators(cls, i, &d) != PTASK_ERR_NOT_FOUND; ++i) all.push_back(d);
    for (ACCELERATOR_DESCRIPTOR *x : all) {
        if (!x->bEnabled) continue;
        if (!best || x->nCoreCount > best->nCoreCount ||
            (x->nCoreCount == best->nCoreCount && x->nMemorySize > best->nMemorySize))
            best = x;
    }
    const UINT id = best ? best->uiAcceleratorId : 0;
    for (ACCELERATOR_DESCRIPTOR *x : all) free(x);
    return id;
}

// graphgroupby.cpp:394-419: equal keys contiguous, and the same key set as the input
int Failures(const std::vector<int> &in, const std::vector<int> &out, bool *sameSet) {
    std::set<int> inSet(in.begin(), in.end()), seen;
    int failures = 0;
    for (size_t i = 0; i + 1 < out.size(); ++i) {
        if (seen.count(out[i])) ++failures;
        if (out[i] != out[i + 1]) seen.insert(out[i]);
    }
    if (!out.empty()) {
        if (seen.count(out.back())) ++failures;
        seen.insert(out.back());
    }
    *sameSet = inSet == seen;
    return failures;
}

}  // namespace

int main(int argc, char **argv) {
    Params p;
    std::string dist = "reference", dump;
    int rows = 1024, keys = 32, iters = 1;
    unsigned long long seed = 0x4fd478;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> const char * {
            if (i + 1 >= argc) {
                fprintf(stderr, "missing value for %s\n", a.c_str());
                exit(2);
            }
            return argv[++i];
        };
        if (a == "--kernels") p.kernels = next();
        else if (a == "-r" || a == "--rows") rows = atoi(next());
        else if (a == "-c" || a == "--keys") keys = atoi(next());
        else if (a == "-i" || a == "--iterations") iters = atoi(next());
        else if (a == "--dist") dist = next();
        else if (a == "--seed") seed = strtoull(next(), nullptr, 0);
        else if (a == "--dump") dump = next();
        else {
            p.kernels.clear();
            break;
        }
    }
    if (p.kernels.empty() || rows < 1 || keys < 1 || iters < 1 ||
        (dist != "reference" && dist != "wide" && dist != "distinct")) {
        fprintf(stderr, "usage: ptk_groupby_ref --kernels FILE [-r rows] [-c keys] [-i iterations] "
                        "[--dist reference|wide|distinct] [--seed S] [--dump file]\n");
        return 2;
    }
    p.inputLength = rows;
    p.numKeys = keys;
    p.numIterations = iters;
    Runtime::SetExitOnRuntimeFailure(FALSE);
    if (PTFAILED(Runtime::Initialize())) return 1;
    p.hip = ptaskutils::SelectAcceleratorClass(p.kernels.c_str()) == ACCELERATOR_CLASS_HIP;
    p.hashTableTemplate = Runtime::GetDatablockTemplate((char *)"HashTableSize", sizeof(void *), kHashTableSize, 1, 1);
    p.inputTemplate = Runtime::GetDatablockTemplate((char *)"InputSize", sizeof(int), (UINT)rows, 1, 1);
    p.keycountTemplate = Runtime::GetDatablockTemplate((char *)"KeyCountTemplate", sizeof(int), 10, 1, 1);
    p.ptrTemplate = Runtime::GetDatablockTemplate((char *)"PtrTemplate", sizeof(int *), 1, 1, 1);
    p.intTemplate = Runtime::GetDatablockTemplate((char *)"IntTemplate", sizeof(int), PTPARM_INT);
    p.accelerator = PreferredAccelerator(p.hip ? ACCELERATOR_CLASS_HIP : ACCELERATOR_CLASS_HOST);
    p.g = new Graph();
    UniqueKeyCountNode(p);
    if (!p.ok) {
        fprintf(stderr, "graph construction failed: %s\n", Runtime::GetLastErrorText());
        return 1;
    }
    p.g->Run(FALSE);

    // graphgroupby.cpp:421-479
    std::vector<int> input(rows), output(rows, 0);
    if (dist == "reference") {
        unsigned holdrand = (unsigned)seed;  // MSVC rand(): the published LCG, RAND_MAX 32767
        for (int i = 0; i < rows; ++i) {
            holdrand = holdrand * 214013u + 2531011u;
            const int r = (int)((holdrand >> 16) & 0x7fff);
            input[i] = (int)(r / (float)32767 * keys);
        }
    } else if (dist == "wide") {
        unsigned long long s = seed * 0x9E3779B97F4A7C15ull + 1;
        for (int i = 0; i < rows; ++i) {
            s = s * 6364136223846793005ull + 1442695040888963407ull;
            const unsigned j = (unsigned)((s >> 33) % (unsigned long long)keys);
            input[i] = (int)((j * 2654435761u) ^ 0x5bd1e995u);  // `keys` values spread over int32
        }
    } else {
        for (int i = 0; i < rows; ++i) input[i] = (int)((unsigned)(rows - i) * 2654435761u);
    }
    std::vector<void *> hashTableCPU(kHashTableSize, nullptr);
    int keyCountCPU[10] = {0};
    auto t0 = std::chrono::steady_clock::now();
    for (int it = 0; it < iters && p.ok; ++it) {
        Datablock *inputBlock = Runtime::AllocateDatablock(p.inputTemplate, input.data(), (UINT)(rows * sizeof(int)), p.inputChannels[0]);
        Datablock *hashTableBlock = Runtime::AllocateDatablock(p.hashTableTemplate, hashTableCPU.data(),
                                                               (UINT)(hashTableCPU.size() * sizeof(void *)), p.hashtableChannels[0]);
        Datablock *keyCountBlock = Runtime::AllocateDatablock(p.keycountTemplate, keyCountCPU, sizeof(keyCountCPU), p.keyCountInChannel);
        Datablock *inputLenBlock = Runtime::AllocateDatablock(p.intTemplate, &p.inputLength, sizeof(p.inputLength), p.inputLengthChannels[0]);
        if (!inputBlock || !hashTableBlock || !keyCountBlock || !inputLenBlock) {
            fprintf(stderr, "allocation failure\n");
            return 1;
        }
        for (GraphInputChannel *c : p.inputChannels) c->Push(inputBlock);
        for (GraphInputChannel *c : p.hashtableChannels) c->Push(hashTableBlock);
        for (GraphInputChannel *c : p.inputLengthChannels) c->Push(inputLenBlock);
        p.keyCountInChannel->Push(keyCountBlock);
        inputBlock->Release();
        hashTableBlock->Release();
        inputLenBlock->Release();
        keyCountBlock->Release();
        Datablock *res = p.outputChannel->Pull(600000);
        if (!res) {
            fprintf(stderr, "no output: %s\n", Runtime::GetLastErrorText());
            p.ok = false;
            break;
        }
        res->Lock();
        memcpy(output.data(), res->GetDataPointer(FALSE), sizeof(int) * (size_t)rows);
        res->Unlock();
        res->Release();
    }
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    p.g->Stop();
    p.g->Teardown();
    Graph::DestroyGraph(p.g);
    Runtime::Terminate();
    bool sameSet = false;
    const int failures = Failures(input, output, &sameSet);
    const bool succeeded = p.ok && failures == 0 && sameSet;
    if (!dump.empty()) {
        FILE *f = fopen(dump.c_str(), "wb");
        if (!f) return 1;
        fwrite(input.data(), sizeof(int), input.size(), f);
        fwrite(output.data(), sizeof(int), output.size(), f);
        fclose(f);
    }
    printf("{\"ok\": %s, \"backend\": \"%s\", \"rows\": %d, \"keys\": %d, \"iterations\": %d, \"dist\": \"%s\", "
           "\"failures\": %d, \"same_key_set\": %s, \"succeeded\": %s, \"ms\": %.3f}\n",
           p.ok ? "true" : "false", p.hip ? "hip" : "host", rows, keys, iters, dist.c_str(), failures,
           sameSet ? "true" : "false", succeeded ? "true" : "false", ms);
    return succeeded ? 0 : 1;
}
