// ptk_groupby.cpp -- the groupby and prefix-sum graphs as PTask applications (the roles of
// ptaskUnitTest/graphgroupby.cpp and ptaskUnitTest/cpfxsum.cpp), against ptaskapi.h only.
//
// --workload groupby: seven tasks per run (graphgroupby.cpp builds the stages count, prefix-sum,
//   reshuffle, :106-387; here as a stable two-level radix partition; the layouts are in
//   include/ptk_groupby.h):
//     GroupbyHistogram  (keys, n, numKeys)                               -> scratch
//     GroupbyBucketScan (scratch in/out, n, numKeys)                     -> scratch
//     GroupbyPartition  (keys, vals, n, numKeys, scratch in/out)         -> scratch
//     GroupbyKeyCount   (scratch in/out, n, numKeys)                     -> scratch
//     GroupbyKeyScan    (scratch in/out, n, numKeys)                     -> scratch, offsets
//     GroupbyReshuffle  (scratch in/out, n, numKeys, offsets)            -> scratch, outKeys, outVals
//     GroupbyTotals     (scratch, n, numKeys)                            -> table, groups
//   and with --coalesce the reference's CoalesceGroups layout (HostTasks.cpp:131-197):
//     GroupbyCoalesce   (outVals, outKeys, groups, offsets)              -> coalesced, groupCounts
// --workload pfxsum: one task pfxsum(pin, pout, N) wired as cpfxsum.cpp:204-226 wires it.
// --workload scan2: the two-task large scan PfxScanReduce -> PfxScanApply.
//
// Keys: Zipf(s) over K distinct keys (rank r drawn by inverse CDF, key = r*A mod K for a fixed A
// coprime to K, so frequent keys are spread over the key space), or uniform on [0, numKeys] as
// the reference draws them (graphgroupby.cpp:428).  Values: int32 from the shared generator.
// Repeated runs reuse the same input datablocks, which stay valid on the device, so only the
// first run uploads.  Prints one JSON object; --dump writes inputs and outputs.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ptaskapi.h"
#include "ptk_groupby.h"
#include "ptk_prng.h"

using namespace PTask;

namespace {

struct Options {
    std::string workload = "groupby";
    std::string kernelFile, scanKernelFile;
    long long rows = 1 << 20;
    int numKeys = 65535;  // keys lie in [0, numKeys]
    std::string dist = "zipf";
    double zipfS = 1.1;
    unsigned long long seed = 0x4fd478;
    int repeat = 1, warmup = 0;
    std::string dump;
    bool profile = false;
    int gbGrid = 2048;  // at most this many blocks for GroupbyPartition (also at most two per CU)
    bool coalesce = false;  // append GroupbyCoalesce (the reference's CoalesceGroups layout)
    bool exact = false;     // ... with its outputs allocated to their exact sizes
    std::string serialize;    // write the graph as XML (and stop)
    std::string deserialize;  // read the graph from XML instead of building it
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

void MakeInputs(const Options &o, std::vector<int32_t> &keys, std::vector<int32_t> &vals) {
    const long long n = o.rows;
    keys.resize((size_t)n);
    vals.resize((size_t)n);
    const long long K = (long long)o.numKeys + 1;
    std::vector<double> cdf;
    long long A = 40503;
    if (o.dist == "zipf") {
        cdf.resize((size_t)K);
        double acc = 0;
        for (long long r = 0; r < K; ++r) {
            acc += std::pow((double)(r + 1), -o.zipfS);
            cdf[(size_t)r] = acc;
        }
        for (double &c : cdf) c /= acc;
        auto gcd = [](long long a, long long b) { while (b) { long long t = a % b; a = b; b = t; } return a; };
        while (gcd(A, K) != 1) ++A;
    }
    for (long long i = 0; i < n; ++i) {
        const double u = ptk_rand_unit_f64(o.seed, 1, (uint64_t)i);
        long long k;
        if (o.dist == "zipf") {
            long long r = std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin();
            if (r >= K) r = K - 1;
            k = (r * A) % K;
        } else {
            k = (long long)(u * (double)o.numKeys);  // [0, numKeys) as rand()/RAND_MAX*numKeys
        }
        keys[(size_t)i] = (int32_t)k;
        vals[(size_t)i] = (int32_t)(uint32_t)ptk_rand_u64(o.seed, 2, (uint64_t)i);
    }
}

void Dump(FILE *fp, const void *p, size_t n) {
    if (fp && n) fwrite(p, 1, n, fp);
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
        if (a == "--workload") o.workload = next();
        else if (a == "--kernels") o.kernelFile = next();
        else if (a == "--rows") o.rows = atoll(next());
        else if (a == "--keys") o.numKeys = atoi(next()) - 1;
        else if (a == "--num-keys") o.numKeys = atoi(next());
        else if (a == "--dist") o.dist = next();
        else if (a == "--zipf-s") o.zipfS = atof(next());
        else if (a == "--seed") o.seed = strtoull(next(), nullptr, 0);
        else if (a == "--repeat") o.repeat = atoi(next());
        else if (a == "--warmup") o.warmup = atoi(next());
        else if (a == "--dump") o.dump = next();
        else if (a == "--profile") o.profile = true;
        else if (a == "--gb-grid") o.gbGrid = atoi(next());
        else if (a == "--coalesce") o.coalesce = true;
        else if (a == "--coalesce-exact") o.coalesce = o.exact = true;
        else if (a == "--serialize") o.serialize = next();
        else if (a == "--deserialize") o.deserialize = next();
        else { fprintf(stderr, "unknown option %s\n", a.c_str()); return 2; }
    }
    if ((o.kernelFile.empty() && o.deserialize.empty()) || o.rows < 0 || o.numKeys < 0 || o.rows > 0x7fffffffLL ||
        (o.workload == "groupby" && (long long)o.numKeys + 1 > PTK_GB_MAX_KEYS)) {
        fprintf(stderr, "usage: ptk_groupby --kernels <groupby.hsaco|scan.hsaco|hosttasks.so> "
                        "[--workload groupby|pfxsum|scan2] [--rows N] [--keys K] [--dist zipf|uniform] "
                        "[--zipf-s S] [--seed S] [--repeat R] [--warmup W] [--dump file] [--profile] [--coalesce] "
                        "[--coalesce-exact] [--serialize graph.xml] [--deserialize graph.xml]\n");
        return 2;
    }
    Runtime::SetExitOnRuntimeFailure(FALSE);
    if (o.profile) Runtime::SetProfilePSDispatch(TRUE);
    if (PTFAILED(Runtime::Initialize())) return 1;
    bool hip = ptaskutils::SelectAcceleratorClass(o.kernelFile.c_str()) == ACCELERATOR_CLASS_HIP;
    const int n = (int)o.rows;
    std::vector<int32_t> keys, vals;
    MakeInputs(o, keys, vals);
    const UINT cbRows = (UINT)(sizeof(int32_t) * (size_t)std::max(n, 1));
    DatablockTemplate *tRows = Runtime::GetDatablockTemplate((char *)"gbRows", sizeof(int32_t), std::max(n, 1), 1, 1);
    DatablockTemplate *tInt = Runtime::GetDatablockTemplate((char *)"gbInt", sizeof(int), PTPARM_INT);
    Graph *g = new Graph(o.workload.c_str());
    UINT uid = 0;
    std::vector<GraphInputChannel *> keyIn, valIn, nIn, kIn;
    std::vector<GraphOutputChannel *> outs;
    std::vector<std::string> outNames;
    auto K = [&](const char *name) {
        CompiledKernel *k = Runtime::GetCompiledKernel((char *)o.kernelFile.c_str(), (char *)name);
        if (!k) {
            fprintf(stderr, "kernel %s unavailable: %s\n", name, Runtime::GetLastErrorText());
            exit(1);
        }
        return k;
    };
    auto geometry = [&](Task *t, int gx, int bx) {
        if (hip) t->SetBlockAndGridSize(PTASKDIM3(gx, 1, 1), PTASKDIM3(bx, 1, 1));
        else t->SetComputeGeometry(1, 1, 1);
    };
    const long long tableElems = 3ll * ((long long)o.numKeys + 1) + 1;
    const long long groupElems = 3 + 3ll * ((long long)o.numKeys + 1);
    if (!o.deserialize.empty()) {
        // the graph from an XML file written by Graph::Serialize; its inputs are found by the
        // binding of the port each feeds, its outputs are the graph outputs in creation order
        if (!g->Deserialize(o.deserialize.c_str())) return 1;
        hip = g->GetTaskCount() && g->GetTaskByIndex(0)->GetAcceleratorClass() == ACCELERATOR_CLASS_HIP;
        for (UINT i = 0; i < g->GetChannelCount(); ++i) {
            Channel *c = g->GetChannelByIndex(i);
            if (c->GetType() == CT_GRAPH_OUTPUT) {
                outs.push_back((GraphOutputChannel *)c);
                outNames.push_back(c->GetName());
                continue;
            }
            if (c->GetType() != CT_GRAPH_INPUT) continue;
            const std::string b = c->GetBoundPort(CE_DST)->GetVariableBinding();
            GraphInputChannel *in = (GraphInputChannel *)c;
            if (b == "keys" || b == "pin") keyIn.push_back(in);
            else if (b == "vals") valIn.push_back(in);
            else if (b == "n" || b == "N") nIn.push_back(in);
            else if (b == "numKeys") kIn.push_back(in);
            else {
                fprintf(stderr, "%s: graph input %s feeds unknown port %s\n", o.deserialize.c_str(), c->GetName(), b.c_str());
                return 1;
            }
        }
        if (keyIn.empty() || outs.empty()) {
            fprintf(stderr, "%s: no row input or no output\n", o.deserialize.c_str());
            return 1;
        }
    } else if (o.workload == "groupby") {
        DatablockTemplate *tTable = Runtime::GetDatablockTemplate((char *)"gbTable", sizeof(int64_t), (UINT)tableElems, 1, 1);
        DatablockTemplate *tOff = Runtime::GetDatablockTemplate((char *)"gbOffsets", sizeof(int64_t), (UINT)(o.numKeys + 1), 1, 1);
        DatablockTemplate *tGroups = Runtime::GetDatablockTemplate((char *)"gbGroups", sizeof(int64_t), (UINT)groupElems, 1, 1);
        const long long scratchBytes = ptk_gb_scratch_bytes(n, o.numKeys);
        DatablockTemplate *tScratch = Runtime::GetDatablockTemplate((char *)"gbScratch", 16, (UINT)(scratchBytes / 16), 1, 1);
        const int chunks = (int)std::max(1LL, ptk_gb_chunks(n));
        const int buckets = (int)ptk_gb_buckets(o.numKeys);
        const int tiles = (int)ptk_gb_tiles_max(n, o.numKeys);
        int cus = 256;  // GroupbyPartition: two blocks per compute unit
        if (hip)
            if (Accelerator *acc = Runtime::GetAccelerator(ACCELERATOR_CLASS_HIP, 0))
                if (acc->GetCoreCount()) cus = (int)acc->GetCoreCount();
        const int partGrid = std::max(1, std::min(std::min(chunks, o.gbGrid), 2 * cus));
        auto nk = [&](UINT i) {  // the sticky n and numKeys ports at formal indices i and i + 1
            return std::array<Port *, 2>{Runtime::CreatePort(STICKY_PORT, tInt, uid++, (char *)"n", i),
                                         Runtime::CreatePort(STICKY_PORT, tInt, uid++, (char *)"numKeys", i + 1)};
        };
        // GroupbyHistogram: rows per bucket per chunk
        auto hNK = nk(1);
        Port *hIn[3] = {Runtime::CreatePort(INPUT_PORT, tRows, uid++, (char *)"keys", 0), hNK[0], hNK[1]};
        Port *hOut[1] = {Runtime::CreatePort(OUTPUT_PORT, tScratch, uid++, (char *)"scratch", 3)};
        Task *th = g->AddTask(K("GroupbyHistogram"), 3, hIn, 1, hOut, "GroupbyHistogram");
        geometry(th, chunks, PTK_GB_HIST_THREADS);
        // GroupbyBucketScan: chunk prefixes per bucket (scratch in/out)
        auto bNK = nk(1);
        Port *bIn[3] = {Runtime::CreatePort(INPUT_PORT, tScratch, uid++, (char *)"scratch", 0, 0), bNK[0], bNK[1]};
        Port *bOut[1] = {Runtime::CreatePort(OUTPUT_PORT, tScratch, uid++, (char *)"scratch(out)", 0)};
        Task *tb = g->AddTask(K("GroupbyBucketScan"), 3, bIn, 1, bOut, "GroupbyBucketScan");
        geometry(tb, buckets, PTK_GB_SCAN_THREADS);
        // GroupbyPartition: rows by bucket into the scratch (scratch in/out)
        auto pNK = nk(2);
        Port *pIn[5] = {Runtime::CreatePort(INPUT_PORT, tRows, uid++, (char *)"keys", 0),
                        Runtime::CreatePort(INPUT_PORT, tRows, uid++, (char *)"vals", 1), pNK[0], pNK[1],
                        Runtime::CreatePort(INPUT_PORT, tScratch, uid++, (char *)"scratch", 4, 0)};
        Port *pOut[1] = {Runtime::CreatePort(OUTPUT_PORT, tScratch, uid++, (char *)"scratch(out)", 4)};
        Task *tp = g->AddTask(K("GroupbyPartition"), 5, pIn, 1, pOut, "GroupbyPartition");
        geometry(tp, partGrid, PTK_GB_PART_THREADS);
        // GroupbyKeyCount: rows per key per tile (scratch in/out)
        auto cNK = nk(1);
        Port *cIn[3] = {Runtime::CreatePort(INPUT_PORT, tScratch, uid++, (char *)"scratch", 0, 0), cNK[0], cNK[1]};
        Port *cOut[1] = {Runtime::CreatePort(OUTPUT_PORT, tScratch, uid++, (char *)"scratch(out)", 0)};
        Task *tc = g->AddTask(K("GroupbyKeyCount"), 3, cIn, 1, cOut, "GroupbyKeyCount");
        geometry(tc, tiles, PTK_GB_COUNT_THREADS);
        // GroupbyKeyScan: tile prefixes per key, offsets (scratch in/out)
        auto sNK = nk(1);
        Port *sIn[3] = {Runtime::CreatePort(INPUT_PORT, tScratch, uid++, (char *)"scratch", 0, 0), sNK[0], sNK[1]};
        Port *sOut[2] = {Runtime::CreatePort(OUTPUT_PORT, tScratch, uid++, (char *)"scratch(out)", 0),
                         Runtime::CreatePort(OUTPUT_PORT, tOff, uid++, (char *)"offsets", 3)};
        Task *ts = g->AddTask(K("GroupbyKeyScan"), 3, sIn, 2, sOut, "GroupbyKeyScan");
        geometry(ts, 2 * buckets, PTK_GB_KEYSCAN_THREADS);
        // GroupbyReshuffle: rows by key into outKeys / outVals, per-tile sums (scratch in/out)
        auto rNK = nk(1);
        Port *rIn[4] = {Runtime::CreatePort(INPUT_PORT, tScratch, uid++, (char *)"scratch", 0, 0), rNK[0], rNK[1],
                        Runtime::CreatePort(INPUT_PORT, tOff, uid++, (char *)"offsets", 3)};
        Port *rOut[3] = {Runtime::CreatePort(OUTPUT_PORT, tScratch, uid++, (char *)"scratch(out)", 0),
                         Runtime::CreatePort(OUTPUT_PORT, tRows, uid++, (char *)"outKeys", 4),
                         Runtime::CreatePort(OUTPUT_PORT, tRows, uid++, (char *)"outVals", 5)};
        Task *tr = g->AddTask(K("GroupbyReshuffle"), 4, rIn, 3, rOut, "GroupbyReshuffle");
        geometry(tr, tiles, PTK_GB_TILE_THREADS);
        // GroupbyTotals: the table and the group records
        auto tNK = nk(1);
        Port *tIn[3] = {Runtime::CreatePort(INPUT_PORT, tScratch, uid++, (char *)"scratch", 0), tNK[0], tNK[1]};
        Port *tOut[2] = {Runtime::CreatePort(OUTPUT_PORT, tTable, uid++, (char *)"table", 3),
                         Runtime::CreatePort(OUTPUT_PORT, tGroups, uid++, (char *)"groups", 4)};
        Task *tt = g->AddTask(K("GroupbyTotals"), 3, tIn, 2, tOut, "GroupbyTotals");
        geometry(tt, 4 * buckets, PTK_GB_TOTALS_THREADS);
        keyIn = {g->AddInputChannel(hIn[0], "keys->histogram"), g->AddInputChannel(pIn[0], "keys->partition")};
        valIn = {g->AddInputChannel(pIn[1], "vals->partition")};
        nIn.clear();
        kIn.clear();
        Port *nkPorts[7][2] = {{hNK[0], hNK[1]}, {bNK[0], bNK[1]}, {pNK[0], pNK[1]}, {cNK[0], cNK[1]},
                               {sNK[0], sNK[1]}, {rNK[0], rNK[1]}, {tNK[0], tNK[1]}};
        static const char *stage[7] = {"histogram", "bucketscan", "partition", "keycount", "keyscan", "reshuffle", "totals"};
        for (int i = 0; i < 7; ++i) {
            nIn.push_back(g->AddInputChannel(nkPorts[i][0], (std::string("n->") + stage[i]).c_str()));
            kIn.push_back(g->AddInputChannel(nkPorts[i][1], (std::string("nk->") + stage[i]).c_str()));
        }
        g->AddInternalChannel(hOut[0], bIn[0], "histogram->bucketscan");
        g->AddInternalChannel(bOut[0], pIn[4], "bucketscan->partition");
        g->AddInternalChannel(pOut[0], cIn[0], "partition->keycount");
        g->AddInternalChannel(cOut[0], sIn[0], "keycount->keyscan");
        g->AddInternalChannel(sOut[0], rIn[0], "keyscan->reshuffle(scratch)");
        g->AddInternalChannel(sOut[1], rIn[3], "keyscan->reshuffle(offsets)");
        g->AddInternalChannel(rOut[0], tIn[0], "reshuffle->totals");
        Port *oOut[2] = {sOut[1], tOut[1]};  // offsets, groups (for the coalesced form below)
        Port *rOutKV[2] = {rOut[1], rOut[2]};
        outs = {g->AddOutputChannel(oOut[1], "groups"), g->AddOutputChannel(rOutKV[0], "outKeys"),
                g->AddOutputChannel(rOutKV[1], "outVals")};
        outNames = {"groups", "outKeys", "outVals"};
        if (o.coalesce) {
            // GroupbyCoalesce: per group its key then its values (n + numKeys + 1 words at most),
            // and the per-group counts (HostTasks.cpp:131-197)
            DatablockTemplate *tCo = Runtime::GetDatablockTemplate((char *)"gbCoalesced", sizeof(int32_t),
                                                                   (UINT)(n + o.numKeys + 1), 1, 1);
            DatablockTemplate *tMeta = Runtime::GetDatablockTemplate((char *)"gbGroupCounts", sizeof(int32_t),
                                                                     (UINT)(o.numKeys + 1), 1, 1);
            std::vector<Port *> kIn4 = {Runtime::CreatePort(INPUT_PORT, tRows, uid++, (char *)"outVals", 0),
                                        Runtime::CreatePort(INPUT_PORT, tRows, uid++, (char *)"outKeys", 1),
                                        Runtime::CreatePort(INPUT_PORT, tGroups, uid++, (char *)"groups", 2),
                                        Runtime::CreatePort(INPUT_PORT, tOff, uid++, (char *)"offsets", 3)};
            Port *kOut2[2] = {Runtime::CreatePort(OUTPUT_PORT, tCo, uid++, (char *)"coalesced", 4),
                              Runtime::CreatePort(OUTPUT_PORT, tMeta, uid++, (char *)"groupCounts", 5)};
            Port *sOut[2] = {nullptr, nullptr};
            if (o.exact) {
                // variable-size outputs (MetaPort.cpp:724, MF_ALLOCATION_SIZE): GroupbyCoalesceSizes
                // turns the groups header into the two byte counts; each arrives on a meta port
                // bound to the output it sizes (its in/out routing index), so the pulled blocks hold
                // exactly rows + G and G words, the reference's wire format
                DatablockTemplate *tBytes = Runtime::GetDatablockTemplate((char *)"gbBytes", sizeof(int32_t), 1, 1, 1);
                Port *sIn2[1] = {Runtime::CreatePort(INPUT_PORT, tGroups, uid++, (char *)"groups", 0)};
                sOut[0] = Runtime::CreatePort(OUTPUT_PORT, tBytes, uid++, (char *)"coalescedBytes", 1);
                sOut[1] = Runtime::CreatePort(OUTPUT_PORT, tBytes, uid++, (char *)"groupCountsBytes", 2);
                Task *tz = g->AddTask(K("GroupbyCoalesceSizes"), 1, sIn2, 2, sOut, "GroupbyCoalesceSizes");
                geometry(tz, 1, 64);
                g->AddInternalChannel(oOut[1], sIn2[0], "totals->sizes(groups)");
                for (int q = 0; q < 2; ++q) {
                    Port *m = Runtime::CreatePort(META_PORT, tBytes, uid++, (char *)(q ? "groupCountsSize" : "coalescedSize"),
                                                  (UINT)-1, (UINT)q);
                    m->SetMetaFunction(MF_ALLOCATION_SIZE);
                    kIn4.push_back(m);
                }
            }
            Task *tk = g->AddTask(K("GroupbyCoalesce"), (UINT)kIn4.size(), kIn4.data(), 2, kOut2, "GroupbyCoalesce");
            geometry(tk, (int)std::max(1LL, std::min(2048LL, ((long long)n + 2047) / 2048)), 512);
            if (o.exact) {
                g->AddInternalChannel(sOut[0], kIn4[4], "sizes->coalesce(coalescedBytes)");
                g->AddInternalChannel(sOut[1], kIn4[5], "sizes->coalesce(groupCountsBytes)");
            }
            g->AddInternalChannel(rOutKV[1], kIn4[0], "shuffle->coalesce(vals)");
            g->AddInternalChannel(rOutKV[0], kIn4[1], "shuffle->coalesce(keys)");
            g->AddInternalChannel(oOut[1], kIn4[2], "totals->coalesce(groups)");
            g->AddInternalChannel(oOut[0], kIn4[3], "keyscan->coalesce(offsets)");
            outs.push_back(g->AddOutputChannel(kOut2[0], "coalesced"));
            outs.push_back(g->AddOutputChannel(kOut2[1], "groupCounts"));
            outNames.push_back("coalesced");
            outNames.push_back("groupCounts");
        }
    } else if (o.workload == "pfxsum") {
        // cpfxsum.cpp:204-226: pin at 0, pout at 1, N (sticky) at 2
        Port *in[2] = {Runtime::CreatePort(INPUT_PORT, tRows, uid++, (char *)"pin", 0),
                       Runtime::CreatePort(STICKY_PORT, tInt, uid++, (char *)"N", 2)};
        Port *out[1] = {Runtime::CreatePort(OUTPUT_PORT, tRows, uid++, (char *)"pout", 1)};
        Task *t = g->AddTask(K("pfxsum"), 2, in, 1, out, "PfxsumTask");
        geometry(t, 1, 1024);
        keyIn = {g->AddInputChannel(in[0], "DataInputChannel")};
        nIn = {g->AddInputChannel(in[1], "ParmsInputChannel")};
        outs = {g->AddOutputChannel(out[0], "DataOutputChannel")};
        outNames = {"pout"};
    } else if (o.workload == "scan2") {
        const int tiles = std::max(1, (n + 4095) / 4096);
        DatablockTemplate *tTiles = Runtime::GetDatablockTemplate((char *)"scanTiles", sizeof(int32_t), (UINT)tiles, 1, 1);
        Port *rIn[2] = {Runtime::CreatePort(INPUT_PORT, tRows, uid++, (char *)"pin", 0),
                        Runtime::CreatePort(STICKY_PORT, tInt, uid++, (char *)"N", 1)};
        Port *rOut[1] = {Runtime::CreatePort(OUTPUT_PORT, tTiles, uid++, (char *)"tileSums", 2)};
        Task *tr = g->AddTask(K("PfxScanReduce"), 2, rIn, 1, rOut, "PfxScanReduce");
        geometry(tr, std::min(tiles, 2048), 1024);
        Port *aIn[3] = {Runtime::CreatePort(INPUT_PORT, tRows, uid++, (char *)"pin", 0),
                        Runtime::CreatePort(STICKY_PORT, tInt, uid++, (char *)"N", 1),
                        Runtime::CreatePort(INPUT_PORT, tTiles, uid++, (char *)"tileSums", 2)};
        Port *aOut[1] = {Runtime::CreatePort(OUTPUT_PORT, tRows, uid++, (char *)"pout", 3)};
        Task *ta = g->AddTask(K("PfxScanApply"), 3, aIn, 1, aOut, "PfxScanApply");
        geometry(ta, std::min(tiles, 2048), 1024);
        keyIn = {g->AddInputChannel(rIn[0], "pin->reduce"), g->AddInputChannel(aIn[0], "pin->apply")};
        nIn = {g->AddInputChannel(rIn[1], "N->reduce"), g->AddInputChannel(aIn[1], "N->apply")};
        g->AddInternalChannel(rOut[0], aIn[2], "tiles");
        outs = {g->AddOutputChannel(aOut[0], "pout")};
        outNames = {"pout"};
    } else {
        fprintf(stderr, "unknown workload %s\n", o.workload.c_str());
        return 2;
    }
    if (!o.serialize.empty()) {
        // write the graph (built here or read from XML) and stop
        BOOL ok = g->Serialize(o.serialize.c_str());
        printf("{\"ok\": %s, \"serialized\": \"%s\", \"tasks\": %u, \"channels\": %u}\n", ok ? "true" : "false",
               o.serialize.c_str(), g->GetTaskCount(), g->GetChannelCount());
        g->Teardown();
        Graph::DestroyGraph(g);
        Runtime::Terminate();
        return ok ? 0 : 1;
    }
    g->Run();
    Datablock *bn = Runtime::AllocateDatablock(tInt, (void *)&n, sizeof(int), nullptr);
    for (auto *c : nIn) c->Push(bn);
    bn->Release();
    Datablock *bk = Runtime::AllocateDatablock(tInt, (void *)&o.numKeys, sizeof(int), nullptr);
    for (auto *c : kIn) c->Push(bk);
    bk->Release();
    Datablock *bKeys = Runtime::AllocateDatablock(tRows, keys.data(), cbRows, nullptr);
    Datablock *bVals = o.workload == "groupby" ? Runtime::AllocateDatablock(tRows, vals.data(), cbRows, nullptr) : nullptr;
    Accelerator *hipAcc = hip ? Runtime::GetAccelerator(ACCELERATOR_CLASS_HIP, 0) : nullptr;
    CHighResolutionTimer timer(gran_msec);
    double timedMs = 0;
    std::vector<std::vector<unsigned char>> results(outs.size());
    bool ok = true;
    const int runs = o.warmup + o.repeat;
    unsigned long long h2d = 0, d2h = 0;
    for (int r = 0; r < runs && ok; ++r) {
        if (r == o.warmup) {
            g->WaitQuiescent(600000);
            if (hipAcc) hipAcc->Synchronize();
            Runtime::ResetTransferStatistics();
            timer.reset();
        }
        for (auto *c : keyIn) c->Push(bKeys);
        for (auto *c : valIn) c->Push(bVals);
        const bool last = r == runs - 1;
        std::vector<Datablock *> got;
        for (size_t i = 0; i < outs.size(); ++i) {
            Datablock *b = outs[i]->Pull(600000);
            if (!b) {
                fprintf(stderr, "no output on %s\n", outNames[i].c_str());
                ok = false;
                break;
            }
            got.push_back(b);
        }
        if (last && ok) {
            g->WaitQuiescent(600000);
            if (hipAcc) hipAcc->Synchronize();
            timedMs = timer.elapsed(false);
            Runtime::GetTransferStatistics(&h2d, nullptr, &d2h, nullptr);
            for (size_t i = 0; i < got.size(); ++i) {
                got[i]->Lock();
                const unsigned char *p = (const unsigned char *)got[i]->GetDataPointer(FALSE);
                results[i].assign(p, p + got[i]->GetDataBufferLogicalSizeBytes());
                got[i]->Unlock();
            }
        }
        for (Datablock *b : got) b->Release();
    }
    std::string taskJson;
    if (o.profile) {
        for (UINT t = 0; t < g->GetTaskCount(); ++t) {
            Task *task = g->GetTaskByIndex(t);
            char buf[256];
            snprintf(buf, sizeof(buf), "%s{\"task\": \"%s\", \"dispatches\": %u, \"device_ms\": %.4f}",
                     taskJson.empty() ? "" : ", ", task->GetTaskName(), task->GetDispatchCount(),
                     task->GetTotalDispatchDeviceMs());
            taskJson += buf;
        }
    }
    bKeys->Release();
    if (bVals) bVals->Release();
    g->Stop();
    g->Teardown();
    Graph::DestroyGraph(g);
    if (ok && !o.dump.empty()) {
        FILE *fp = fopen(o.dump.c_str(), "wb");
        Dump(fp, keys.data(), keys.size() * 4);
        Dump(fp, vals.data(), vals.size() * 4);
        for (auto &r : results) Dump(fp, r.data(), r.size());
        if (fp) fclose(fp);
    }
    long long groupsFound = -1;
    if (o.workload == "groupby" && !results.empty() && results[0].size() >= 8) memcpy(&groupsFound, results[0].data(), 8);
    printf("{\"ok\": %s, \"workload\": \"%s\", \"backend\": \"%s\", \"rows\": %d, \"num_keys\": %d, "
           "\"dist\": \"%s\", \"runs\": %d, \"timed_runs\": %d, \"timed_ms\": %.4f, \"ms_per_run\": %.4f, "
           "\"rows_per_s\": %.1f, \"groups\": %lld, \"h2d_count_timed\": %llu, \"d2h_count_timed\": %llu, "
           "\"tasks\": [%s], \"outputs\": [",
           ok ? "true" : "false", o.workload.c_str(), hip ? "hip" : "host", n, o.numKeys, o.dist.c_str(),
           runs, o.repeat, timedMs, o.repeat ? timedMs / o.repeat : 0.0,
           timedMs > 0 ? (double)n * o.repeat / (timedMs / 1e3) : 0.0, groupsFound, h2d, d2h,
           taskJson.c_str());
    for (size_t i = 0; i < results.size(); ++i)
        printf("%s{\"name\": \"%s\", \"bytes\": %zu, \"fnv1a\": \"%016llx\"}", i ? ", " : "", outNames[i].c_str(),
               results[i].size(), (unsigned long long)Fnv1a(results[i].data(), results[i].size()));
    printf("]}\n");
    Runtime::Terminate();
    return ok ? 0 : 1;
}
