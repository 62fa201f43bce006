// pools.cpp -- allocation with recovery, the registry of live block pools, and the runtime-wide
// and graph-scoped pools (reference roles: GlobalPoolManager.cpp:120-250,665 RequireBlockPool /
// RequestBlock, ScopedPoolManager for Graph::RequireBlockPool / RequestPooledBlock (graph.h:139-
// 199), and the forced collection and retry of PCUBuffer.cpp:979-1015 / ptgc.cpp:326 when an
// allocation fails).
//
// Every block pool (an output port's, a graph's, the runtime's) registers itself.  When host or
// device storage cannot be allocated, the free blocks of every registered pool are deleted and
// the allocation is tried again; if it still fails, every accelerator is drained (blocks whose
// last reference went while a kernel still read them are gone once it ends) and the pools are
// swept once more before the last try.  Only the final failure is reported.
#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <set>
#include <tuple>

#include "ptk_internal.h"

namespace PTask {

namespace {

std::recursive_mutex &RegistryLock() {
    static std::recursive_mutex m;
    return m;
}
std::set<BlockPool *> &Registry() {
    static std::set<BlockPool *> s;
    return s;
}

// A forced failure for tests (Runtime::SetAllocationFailureInjection).
bool Injected(UINT ms) {
    if (ms >= PTK_MAX_MEMORY_SPACES) return false;
    std::atomic<int> &n = RT().injectedFailures[ms];
    int v = n.load();
    while (v > 0)
        if (n.compare_exchange_weak(v, v - 1)) return true;
    return false;
}

void DrainAccelerators() {
    for (Accelerator *a : RT().hipAccelerators) a->Synchronize();
}

// Runs `attempt` up to three times: as is, after a sweep of the pools, after draining the
// accelerators and sweeping again.  attempt() returns nullptr on failure.
template <class T, class F>
T *WithRecovery(UINT ms, F attempt) {
    T *p = Injected(ms) ? nullptr : attempt();
    if (p) return p;
    RuntimeState &S = RT();
    S.allocFailures++;
    for (int round = 0; round < 2; ++round) {
        if (round == 1) DrainAccelerators();
        S.bytesReclaimed += ReclaimPooledMemory();
        p = Injected(ms) ? nullptr : attempt();
        if (p) {
            S.allocRecovered++;
            return p;
        }
    }
    S.allocUnrecovered++;
    return nullptr;
}

}  // namespace

void RegisterPool(BlockPool *pPool) {
    std::lock_guard<std::recursive_mutex> g(RegistryLock());
    Registry().insert(pPool);
}

void UnregisterPool(BlockPool *pPool) {
    std::lock_guard<std::recursive_mutex> g(RegistryLock());
    Registry().erase(pPool);
}

// The registry lock is held throughout: a pool whose last reference is dropped elsewhere waits in
// its destructor (UnregisterPool comes first there) until the sweep is over, and such a pool has
// no free blocks (each free block holds a reference).  A pool that a sweep's own Trim destroys
// unregisters itself on this thread; the snapshot is checked against the registry before use.
size_t ReclaimPooledMemory() {
    std::lock_guard<std::recursive_mutex> g(RegistryLock());
    std::vector<BlockPool *> snapshot(Registry().begin(), Registry().end());
    size_t cb = 0;
    for (BlockPool *p : snapshot)
        if (Registry().count(p)) cb += p->Trim();
    return cb;
}

void *AllocHostStorage(size_t cb) {
    const size_t rounded = ((cb ? cb : 1) + 4095) & ~size_t(4095);
    return WithRecovery<void>(HOST_MEMORY_SPACE_ID, [&]() -> void * {
        void *p = nullptr;
        return posix_memalign(&p, 4096, rounded) == 0 ? p : nullptr;
    });
}

PBuffer *AllocDeviceStorage(Accelerator *pAcc, size_t cb) {
    PBuffer *b = WithRecovery<PBuffer>(pAcc->GetMemorySpaceId(), [&]() -> PBuffer * {
        return pAcc->NewPlatformSpecificBuffer(cb);
    });
    if (!b)
        Runtime::HandleError("out of memory on %s allocating %zu bytes (after reclaiming pooled blocks): %s",
                             pAcc->GetDeviceName(), cb, ptk_hip_last_error_string());
    return b;
}

// ---- runtime-wide pools ------------------------------------------------------------------------

namespace {

struct Requirement {
    DatablockTemplate *tmpl;  // null: untyped
    int cbData, cbMeta, blocks;
};

struct GlobalPools {
    std::mutex lock;
    std::vector<Requirement> required;
    std::map<DatablockTemplate *, BlockPool *> typed;
    std::map<UINT, BlockPool *> untyped;  // by block data size
    bool allocated = false;
};

GlobalPools &GP() {
    static GlobalPools g;
    return g;
}

UINT DefaultBlocks(int n) { return n > 0 ? (UINT)n : std::max<UINT>(1, RT().blockPoolSize); }

BlockPool *MakePool(DatablockTemplate *t, const Requirement &r) {
    BlockPool *p = new BlockPool(t, UINT_MAX);
    p->AddRef();
    p->Prefill(DefaultBlocks(r.blocks), (UINT)r.cbData, (UINT)r.cbMeta, DefaultHIPAccelerator());
    return p;
}

// Caller holds GP().lock.
void Materialize(const Requirement &r) {
    GlobalPools &G = GP();
    if (r.tmpl) {
        auto it = G.typed.find(r.tmpl);
        if (it != G.typed.end()) {
            it->second->Prefill(DefaultBlocks(r.blocks), (UINT)r.cbData, (UINT)r.cbMeta, DefaultHIPAccelerator());
            return;
        }
        r.tmpl->AddRef();
        G.typed[r.tmpl] = MakePool(r.tmpl, r);
    } else {
        auto it = G.untyped.find((UINT)r.cbData);
        if (it != G.untyped.end()) {
            it->second->Prefill(DefaultBlocks(r.blocks), (UINT)r.cbData, (UINT)r.cbMeta, DefaultHIPAccelerator());
            return;
        }
        G.untyped[(UINT)r.cbData] = MakePool(nullptr, r);
    }
}

}  // namespace

void AllocateGlobalPools() {
    GlobalPools &G = GP();
    std::lock_guard<std::mutex> g(G.lock);
    for (const Requirement &r : G.required) {
        Materialize(r);
        if (r.tmpl) r.tmpl->Release();  // the requirement's own reference
    }
    G.required.clear();
    G.allocated = true;
}

void DestroyGlobalPools() {
    GlobalPools &G = GP();
    std::vector<std::pair<DatablockTemplate *, BlockPool *>> pools;
    {
        std::lock_guard<std::mutex> g(G.lock);
        for (auto &kv : G.typed) pools.push_back(kv);
        for (auto &kv : G.untyped) pools.push_back(std::make_pair((DatablockTemplate *)nullptr, kv.second));
        G.typed.clear();
        G.untyped.clear();
        G.allocated = false;
    }
    for (auto &kv : pools) {
        kv.second->Close();
        kv.second->Release();
        if (kv.first) kv.first->Release();
    }
}

Datablock *RequestGlobalPooledBlock(DatablockTemplate *pTemplate, UINT cb, UINT cbMeta) {
    GlobalPools &G = GP();
    BlockPool *pool = nullptr;
    bool typed = false;
    {
        std::lock_guard<std::mutex> g(G.lock);
        if (pTemplate) {
            auto it = G.typed.find(pTemplate);
            if (it != G.typed.end()) {
                pool = it->second;
                typed = true;
            }
        }
        if (!pool) {
            auto it = G.untyped.lower_bound(cb);
            if (it != G.untyped.end()) pool = it->second;
        }
        if (!pool || cbMeta) return nullptr;
        pool->AddRef();
    }
    RuntimeState &S = RT();
    Datablock *b = pool->Request(cb, TRUE, FALSE);
    if (b) {
        S.poolHits++;
    } else {
        S.poolMisses++;
        // a typed pool grows when it runs dry (the new block returns to it); an untyped pool
        // hands out only the blocks it holds
        if (typed) b = pool->Request(cb, FALSE, TRUE);
    }
    pool->Release();
    return b;
}

namespace Runtime {

// reference: PTaskRuntime.h:111-112 (GlobalPoolManager::RequireBlockPool); before Initialize the
// requirement is recorded and the pools are allocated by Initialize, as the reference's are.
void RequireBlockPool(int nDataSize, int nMetaSize, int nTemplateSize, int nBlocks) {
    (void)nTemplateSize;
    if (nDataSize <= 0 || nMetaSize < 0) {
        HandleError("RequireBlockPool: a pool needs a positive data size");
        return;
    }
    GlobalPools &G = GP();
    std::lock_guard<std::mutex> g(G.lock);
    Requirement r{nullptr, nDataSize, nMetaSize, nBlocks};
    if (G.allocated) Materialize(r);
    else G.required.push_back(r);
}

void RequireBlockPool(DatablockTemplate *pTemplate, int nBlocks) {
    if (!pTemplate) {
        HandleError("RequireBlockPool: no template");
        return;
    }
    GlobalPools &G = GP();
    std::lock_guard<std::mutex> g(G.lock);
    Requirement r{pTemplate, (int)pTemplate->GetDatablockByteCount(DBDATA_IDX), 0, nBlocks};
    if (G.allocated) {
        Materialize(r);
    } else {
        pTemplate->AddRef();  // held until Initialize materializes the pool
        G.required.push_back(r);
    }
}

void GetMemoryStatistics(PTK_MEMORY_STATISTICS *p) {
    if (!p) return;
    RuntimeState &S = RT();
    p->ullAllocationFailures = S.allocFailures;
    p->ullAllocationsRecovered = S.allocRecovered;
    p->ullAllocationsUnrecovered = S.allocUnrecovered;
    p->ullBytesReclaimed = S.bytesReclaimed;
    p->ullPooledBlockHits = S.poolHits;
    p->ullPooledBlockMisses = S.poolMisses;
}

void ResetMemoryStatistics() {
    RuntimeState &S = RT();
    S.allocFailures = 0;
    S.allocRecovered = 0;
    S.allocUnrecovered = 0;
    S.bytesReclaimed = 0;
    S.poolHits = 0;
    S.poolMisses = 0;
}

void SetAllocationFailureInjection(UINT uiMemorySpaceId, int nFailures) {
    if (uiMemorySpaceId < PTK_MAX_MEMORY_SPACES) RT().injectedFailures[uiMemorySpaceId] = nFailures > 0 ? nFailures : 0;
}

}  // namespace Runtime

// ---- graph-scoped pools (reference graph.h:139-199) ---------------------------------------------

BOOL Graph::RequireBlockPool(DatablockTemplate *pTemplate, int nDataSize, int nMetaSize, int nTemplateSize,
                             int nBlocks) {
    (void)nTemplateSize;
    const int cb = nDataSize > 0 ? nDataSize : pTemplate ? (int)pTemplate->GetDatablockByteCount(DBDATA_IDX) : 0;
    if (cb <= 0) return Runtime::HandleError("%s: RequireBlockPool needs a data size", GetName());
    std::lock_guard<std::recursive_mutex> g(m_lock);
    BlockPool *&pool = pTemplate ? m_mTypedPools[pTemplate] : m_mUntypedPools[(UINT)cb];
    if (!pool) {
        pool = new BlockPool(pTemplate, UINT_MAX);
        pool->AddRef();
        if (pTemplate) pTemplate->AddRef();
    }
    const UINT want = DefaultBlocks(nBlocks);
    return pool->Prefill(want, (UINT)cb, (UINT)std::max(0, nMetaSize), DefaultHIPAccelerator()) == want;
}

BOOL Graph::RequireBlockPool(int nDataSize, int nMetaSize, int nTemplateSize, int nBlocks) {
    return RequireBlockPool(nullptr, nDataSize, nMetaSize, nTemplateSize, nBlocks);
}

BOOL Graph::RequireBlockPool(DatablockTemplate *pTemplate, int nBlocks) {
    return RequireBlockPool(pTemplate, 0, 0, 0, nBlocks);
}

// A block from this graph's pool for the template, else from its smallest untyped pool that holds
// uiDataSize bytes; nullptr when no such pool has a free block (the reference's scoped pools
// likewise only hand out what they hold).
Datablock *Graph::RequestPooledBlock(DatablockTemplate *pTemplate, UINT uiDataSize, UINT uiMetaSize,
                                     UINT uiTemplateSize) {
    (void)uiTemplateSize;
    const UINT cb = uiDataSize ? uiDataSize : pTemplate ? pTemplate->GetDatablockByteCount(DBDATA_IDX) : 0;
    BlockPool *pool = nullptr;
    {
        std::lock_guard<std::recursive_mutex> g(m_lock);
        if (pTemplate) {
            auto it = m_mTypedPools.find(pTemplate);
            if (it != m_mTypedPools.end()) pool = it->second;
        }
        if (!pool) {
            auto it = m_mUntypedPools.lower_bound(cb);
            if (it != m_mUntypedPools.end()) pool = it->second;
        }
        if (!pool) return nullptr;
        pool->AddRef();
    }
    Datablock *b = uiMetaSize ? nullptr : pool->Request(cb, TRUE, FALSE);
    pool->Release();
    RuntimeState &S = RT();
    if (b) S.poolHits++;
    else S.poolMisses++;
    return b;
}

void Graph::ReleaseScopedPools() {
    std::vector<std::pair<DatablockTemplate *, BlockPool *>> pools;
    {
        std::lock_guard<std::recursive_mutex> g(m_lock);
        for (auto &kv : m_mTypedPools) pools.push_back(kv);
        for (auto &kv : m_mUntypedPools) pools.push_back(std::make_pair((DatablockTemplate *)nullptr, kv.second));
        m_mTypedPools.clear();
        m_mUntypedPools.clear();
    }
    for (auto &kv : pools) {
        kv.second->Close();
        kv.second->Release();
        if (kv.first) kv.first->Release();
    }
}

}  // namespace PTask
