// datablock.cpp -- templates, datablocks with host/device coherence, block pools and the
// per-buffer asynchronous-dependence frontier (reference roles: datablocktemplate.cpp,
// datablock.cpp UpdateView :2235-2348, PBuffer.cpp :1233-1834, BlockPool.cpp).

#include <cstdlib>
#include <cstring>

#include "ptk_internal.h"

namespace PTask {

// ---- DatablockTemplate --------------------------------------------------------------------------

DatablockTemplate::DatablockTemplate(const char *lpszName, UINT uiStride, UINT x, UINT y, UINT z,
                                     UINT uiPitch, bool bRecordStream, bool bByteAddressable)
    : m_strName(lpszName ? lpszName : ""), m_eParamType(PTPARM_BUFFER),
      m_bRecordStream(bRecordStream), m_bByteAddressable(bByteAddressable),
      m_uiInitialElements(0), m_bHasInitialValue(FALSE), m_bExplicitlyEmpty(FALSE) {
    m_dims[DBDATA_IDX].Initialize(x, y, z, uiStride, uiPitch ? uiPitch : uiStride * x);
    for (UINT c = 1; c < NUM_DATABLOCK_CHANNELS; ++c) m_dims[c].Initialize(0, 0, 0, 0, 0);
}

DatablockTemplate::DatablockTemplate(const char *lpszName, const BUFFERDIMENSIONS *pDims, UINT nDims,
                                     bool bRecordStream, bool bByteAddressable)
    : m_strName(lpszName ? lpszName : ""), m_eParamType(PTPARM_BUFFER),
      m_bRecordStream(bRecordStream), m_bByteAddressable(bByteAddressable),
      m_uiInitialElements(0), m_bHasInitialValue(FALSE), m_bExplicitlyEmpty(FALSE) {
    for (UINT c = 0; c < NUM_DATABLOCK_CHANNELS; ++c) {
        if (pDims && c < nDims) m_dims[c] = pDims[c];
        else m_dims[c].Initialize(0, 0, 0, 0, 0);
    }
}

DatablockTemplate::DatablockTemplate(const char *lpszName, UINT uiStride, PTASK_PARM_TYPE eType)
    : m_strName(lpszName ? lpszName : ""), m_eParamType(eType), m_bRecordStream(false),
      m_bByteAddressable(false), m_uiInitialElements(0), m_bHasInitialValue(FALSE),
      m_bExplicitlyEmpty(FALSE) {
    m_dims[DBDATA_IDX].Initialize(1, 1, 1, uiStride, uiStride);
    for (UINT c = 1; c < NUM_DATABLOCK_CHANNELS; ++c) m_dims[c].Initialize(0, 0, 0, 0, 0);
}

DatablockTemplate::~DatablockTemplate() {}
char *DatablockTemplate::GetTemplateName() { return const_cast<char *>(m_strName.c_str()); }
UINT DatablockTemplate::GetStride(UINT c) { return m_dims[c].cbElementStride; }
UINT DatablockTemplate::GetPitch(UINT c) { return m_dims[c].cbPitch; }
UINT DatablockTemplate::GetXElementCount(UINT c) { return m_dims[c].uiXElements; }
UINT DatablockTemplate::GetYElementCount(UINT c) { return m_dims[c].uiYElements; }
UINT DatablockTemplate::GetZElementCount(UINT c) { return m_dims[c].uiZElements; }
UINT DatablockTemplate::GetDatablockByteCount(UINT c) { return m_dims[c].AllocationSizeBytes(); }
BUFFERDIMENSIONS DatablockTemplate::GetBufferDimensions(UINT c) { return m_dims[c]; }
PTASK_PARM_TYPE DatablockTemplate::GetParameterBaseType() { return m_eParamType; }
BOOL DatablockTemplate::IsByteAddressable() { return m_bByteAddressable; }
BOOL DatablockTemplate::DescribesRecordStream() { return m_bRecordStream; }
BOOL DatablockTemplate::IsVariableDimensioned() { return m_dims[DBDATA_IDX].AllocationSizeBytes() == 0; }
BOOL DatablockTemplate::IsScalarParameter() {
    return m_eParamType == PTPARM_INT || m_eParamType == PTPARM_FLOAT ||
           m_eParamType == PTPARM_BYVALSTRUCT || m_eParamType == PTPARM_DOUBLE;
}

void DatablockTemplate::SetInitialValue(void *lpvInitData, UINT cbData, UINT nRecordCount,
                                        BOOL bExplicitlyEmpty) {
    m_vInitialValue.assign((unsigned char *)lpvInitData, (unsigned char *)lpvInitData + (lpvInitData ? cbData : 0));
    m_uiInitialElements = nRecordCount;
    m_bHasInitialValue = TRUE;
    m_bExplicitlyEmpty = bExplicitlyEmpty;
}
const void *DatablockTemplate::GetInitialValue() { return m_vInitialValue.empty() ? nullptr : m_vInitialValue.data(); }
UINT DatablockTemplate::GetInitialValueSizeBytes() { return (UINT)m_vInitialValue.size(); }
UINT DatablockTemplate::GetInitialValueElements() { return m_uiInitialElements; }
BOOL DatablockTemplate::HasInitialValue() { return m_bHasInitialValue; }
BOOL DatablockTemplate::IsInitialValueEmpty() { return m_bExplicitlyEmpty; }

// ---- SyncPoint / frontier -------------------------------------------------------------------------

BOOL SyncPoint::QueryOutstanding() {
    if (!m_pEvent) return FALSE;
    return ptk_hip_event_query((ptk_hip_event)m_pEvent) == PTK_HIP_NOT_READY;
}

void SyncPoint::Wait() {
    if (m_pEvent) CheckHip(ptk_hip_event_synchronize((ptk_hip_event)m_pEvent), "event synchronize");
}

void SyncPoint::OnLastRelease() {
    AsyncContext *ctx = m_pContext;
    if (ctx) {
        if (m_pEvent) ctx->RecycleEvent(m_pEvent);
        ctx->Release();  // each sync point holds its context alive
    }
    delete this;
}

PBuffer::~PBuffer() {
    std::lock_guard<std::mutex> g(m_frontierLock);
    if (m_pLastWrite) m_pLastWrite->Release();
    for (SyncPoint *r : m_vReads) r->Release();
    m_vReads.clear();
    m_pLastWrite = nullptr;
}

void PBuffer::WaitOutstanding(AsyncContext *pCtx, BOOL bForWrite) {
    std::lock_guard<std::mutex> g(m_frontierLock);
    if (m_pLastWrite) pCtx->InsertFence(m_pLastWrite);
    if (bForWrite)
        for (SyncPoint *r : m_vReads) pCtx->InsertFence(r);
}

void PBuffer::AddOutstanding(SyncPoint *pSP, BOOL bWrite) {
    std::lock_guard<std::mutex> g(m_frontierLock);
    pSP->AddRef();
    if (bWrite) {
        // the writer was fenced after everything in the frontier, so it subsumes it
        if (m_pLastWrite) m_pLastWrite->Release();
        for (SyncPoint *r : m_vReads) r->Release();
        m_vReads.clear();
        m_pLastWrite = pSP;
        return;
    }
    for (SyncPoint *r : m_vReads)
        if (r == pSP) {  // already recorded by another argument of the same dispatch
            pSP->Release();
            return;
        }
    if (m_vReads.size() >= 16) {
        // keep the list short: completed reads no longer order anything
        std::vector<SyncPoint *> keep;
        for (SyncPoint *r : m_vReads) {
            if (r->QueryOutstanding()) keep.push_back(r);
            else r->Release();
        }
        m_vReads.swap(keep);
    }
    m_vReads.push_back(pSP);
}

void PBuffer::WaitAllOnHost() {
    std::vector<SyncPoint *> v;
    {
        std::lock_guard<std::mutex> g(m_frontierLock);
        if (m_pLastWrite) {
            m_pLastWrite->AddRef();
            v.push_back(m_pLastWrite);
        }
        for (SyncPoint *r : m_vReads) {
            r->AddRef();
            v.push_back(r);
        }
    }
    for (SyncPoint *s : v) {
        s->Wait();
        s->Release();
    }
}

// ---- Datablock --------------------------------------------------------------------------------------

Datablock::Datablock(DatablockTemplate *pTemplate, UINT cbData, UINT cbMeta, CONTROLSIGNAL luiCode)
    : m_pTemplate(pTemplate), m_cbLogical(cbData), m_cbAllocated(cbData), m_uiRecordCount(0),
      m_luiControl(luiCode), m_eFlags(PT_ACCESS_DEFAULT), m_bSealed(FALSE),
      m_uiDBUID(RT().nextDBUID++), m_pHost(nullptr), m_pPoolOwner(nullptr) {
    if (m_pTemplate) m_pTemplate->AddRef();
    if (cbMeta) m_vMeta.assign(cbMeta, 0);
    for (UINT i = 0; i < PTK_MAX_MEMORY_SPACES; ++i) {
        m_pBuffers[i] = nullptr;
        m_eState[i] = BSTATE_NO_ENTRY;
    }
}

Datablock::~Datablock() {
    for (UINT i = 0; i < PTK_MAX_MEMORY_SPACES; ++i) {
        if (m_pBuffers[i]) {
            // do not free device storage under an in-flight kernel
            m_pBuffers[i]->WaitAllOnHost();
            delete m_pBuffers[i];
        }
    }
    free(m_pHost);
    if (m_pTemplate) m_pTemplate->Release();
}

void Datablock::OnLastRelease() {
    BlockPool *pool = m_pPoolOwner;
    if (pool) pool->Return(this);
    else delete this;
}

BOOL Datablock::EnsureHostBuffer() {
    if (m_pHost) return TRUE;
    size_t cb = m_cbAllocated ? m_cbAllocated : 1;
    // page-aligned and zeroed: a block nobody has written reads as zeros everywhere; a failed
    // allocation is retried after pooled blocks are reclaimed
    void *p = AllocHostStorage(cb);
    if (!p) {
        Runtime::HandleError("out of host memory allocating a %u-byte datablock", m_cbAllocated);
        return FALSE;
    }
    memset(p, 0, cb);
    m_pHost = (unsigned char *)p;
    if (m_eState[HOST_MEMORY_SPACE_ID] == BSTATE_NO_ENTRY) {
        BOOL anyValid = FALSE;
        for (UINT i = 1; i < PTK_MAX_MEMORY_SPACES; ++i)
            if (m_eState[i] == BSTATE_SHARED || m_eState[i] == BSTATE_EXCLUSIVE) anyValid = TRUE;
        m_eState[HOST_MEMORY_SPACE_ID] = anyValid ? BSTATE_INVALID : BSTATE_SHARED;
    }
    return TRUE;
}

void Datablock::SyncHostFromDevice(UINT ms) {
    Accelerator *acc = AcceleratorForMemorySpace(ms);
    if (!acc || !m_pBuffers[ms]) return;
    acc->MakeDeviceContextCurrent();
    if (m_cbLogical) acc->CopyDeviceToHost(m_pHost, m_pBuffers[ms], m_cbLogical);
    RT().d2hCount++;
    RT().d2hBytes += m_cbLogical;
    m_eState[HOST_MEMORY_SPACE_ID] = BSTATE_SHARED;
    if (m_eState[ms] == BSTATE_EXCLUSIVE) m_eState[ms] = BSTATE_SHARED;
}

void *Datablock::AcquireHostView(BOOL bWriteable, BOOL bPopulate) {
    Lock();
    if (!EnsureHostBuffer()) {
        Unlock();
        return nullptr;
    }
    BUFFER_COHERENCE_STATE &hs = m_eState[HOST_MEMORY_SPACE_ID];
    if (bPopulate && hs != BSTATE_SHARED && hs != BSTATE_EXCLUSIVE) {
        for (UINT ms = 1; ms < PTK_MAX_MEMORY_SPACES; ++ms) {
            if (m_eState[ms] == BSTATE_SHARED || m_eState[ms] == BSTATE_EXCLUSIVE) {
                SyncHostFromDevice(ms);
                break;
            }
        }
        if (hs != BSTATE_SHARED && hs != BSTATE_EXCLUSIVE) hs = BSTATE_SHARED;
    }
    if (bWriteable) {
        for (UINT ms = 1; ms < PTK_MAX_MEMORY_SPACES; ++ms) {
            if (m_pBuffers[ms]) {
                // a device read or a host->device copy of this block may still be in flight
                m_pBuffers[ms]->WaitAllOnHost();
                if (m_eState[ms] != BSTATE_NO_ENTRY) m_eState[ms] = BSTATE_INVALID;
            }
        }
        hs = BSTATE_EXCLUSIVE;
    }
    void *p = m_pHost;
    Unlock();
    return p;
}

void *Datablock::GetDataPointer(BOOL bWriteable, BOOL bSynchronize) {
    return AcquireHostView(bWriteable, bSynchronize);
}

void *Datablock::GetMetadataPointer(BOOL, BOOL) {
    return m_vMeta.empty() ? nullptr : m_vMeta.data();
}

void *Datablock::AcquireDeviceView(Accelerator *pAcc, AsyncContext *pCtx, BOOL bWriteable,
                                   BOOL bPopulate) {
    const UINT ms = pAcc->GetMemorySpaceId();
    if (ms == HOST_MEMORY_SPACE_ID || ms >= PTK_MAX_MEMORY_SPACES) {
        Runtime::HandleError("AcquireDeviceView: memory space %u is not a device", ms);
        return nullptr;
    }
    Lock();
    if (!m_pBuffers[ms]) {
        m_pBuffers[ms] = AllocDeviceStorage(pAcc, m_cbAllocated ? m_cbAllocated : 4);
        if (!m_pBuffers[ms]) {
            Unlock();
            return nullptr;
        }
    }
    PBuffer *dev = m_pBuffers[ms];
    BUFFER_COHERENCE_STATE &ds = m_eState[ms];
    const BUFFER_COHERENCE_STATE hs = m_eState[HOST_MEMORY_SPACE_ID];
    if (bPopulate && ds != BSTATE_SHARED && ds != BSTATE_EXCLUSIVE) {
        if ((hs == BSTATE_SHARED || hs == BSTATE_EXCLUSIVE) && m_pHost && m_cbLogical) {
            dev->WaitOutstanding(pCtx, TRUE);
            pAcc->CopyHostToDevice(dev, m_pHost, m_cbLogical, pCtx);
            SyncPoint *sp = pCtx->CreateSyncPoint();
            dev->AddOutstanding(sp, TRUE);
            sp->Release();
            RT().h2dCount++;
            RT().h2dBytes += m_cbLogical;
        }
        ds = BSTATE_SHARED;
        if (hs == BSTATE_EXCLUSIVE) m_eState[HOST_MEMORY_SPACE_ID] = BSTATE_SHARED;
    }
    if (bWriteable) {
        ds = BSTATE_EXCLUSIVE;
        if (m_eState[HOST_MEMORY_SPACE_ID] != BSTATE_NO_ENTRY)
            m_eState[HOST_MEMORY_SPACE_ID] = BSTATE_INVALID;
    } else if (ds == BSTATE_NO_ENTRY || ds == BSTATE_INVALID) {
        ds = BSTATE_SHARED;
    }
    void *p = dev->GetBuffer();
    Unlock();
    return p;
}

BOOL Datablock::ReserveDeviceBuffer(Accelerator *pAcc) {
    const UINT ms = pAcc ? pAcc->GetMemorySpaceId() : HOST_MEMORY_SPACE_ID;
    if (ms == HOST_MEMORY_SPACE_ID || ms >= PTK_MAX_MEMORY_SPACES) return FALSE;
    Lock();
    if (!m_pBuffers[ms]) m_pBuffers[ms] = AllocDeviceStorage(pAcc, m_cbAllocated ? m_cbAllocated : 4);
    const BOOL ok = m_pBuffers[ms] != nullptr;
    Unlock();
    return ok;
}

size_t Datablock::GetFootprintBytes() {
    size_t cb = m_pHost ? m_cbAllocated : 0;
    for (UINT i = 1; i < PTK_MAX_MEMORY_SPACES; ++i)
        if (m_pBuffers[i]) cb += m_pBuffers[i]->GetSize();
    return cb;
}

Datablock *Datablock::CloneForWriter() {
    Datablock *c = new Datablock(m_pTemplate, m_cbLogical, 0, m_luiControl);
    c->AddRef();  // the caller's reference
    Lock();
    c->m_uiRecordCount = m_uiRecordCount;
    c->m_bSealed = m_bSealed;
    c->m_eFlags = m_eFlags;
    c->m_vMeta = m_vMeta;
    // a valid device copy first (the copy stays on the device, where a writing task of the
    // device class reads it), else the host copy
    BOOL ok = TRUE, copied = FALSE;
    for (UINT ms = 1; ms < PTK_MAX_MEMORY_SPACES && !copied; ++ms) {
        if ((m_eState[ms] != BSTATE_SHARED && m_eState[ms] != BSTATE_EXCLUSIVE) || !m_pBuffers[ms]) continue;
        Accelerator *acc = AcceleratorForMemorySpace(ms);
        AsyncContext *ctx = acc ? acc->GetTransferContext() : nullptr;
        PBuffer *dst = acc && ctx ? AllocDeviceStorage(acc, m_cbAllocated ? m_cbAllocated : 4) : nullptr;
        copied = TRUE;
        if (!dst) {
            ok = FALSE;
            break;
        }
        acc->MakeDeviceContextCurrent();
        m_pBuffers[ms]->WaitOutstanding(ctx, FALSE);  // after the block's last write
        if (m_cbLogical) ok = acc->CopyDeviceToDevice(dst, m_pBuffers[ms], m_cbLogical, ctx);
        SyncPoint *sp = ctx->CreateSyncPoint();
        dst->AddOutstanding(sp, TRUE);
        m_pBuffers[ms]->AddOutstanding(sp, FALSE);  // a later writer of this block waits for the copy
        sp->Release();
        c->m_pBuffers[ms] = dst;
        c->m_eState[ms] = BSTATE_EXCLUSIVE;
    }
    const BUFFER_COHERENCE_STATE hs = m_eState[HOST_MEMORY_SPACE_ID];
    if (!copied && (hs == BSTATE_SHARED || hs == BSTATE_EXCLUSIVE) && m_pHost) {
        ok = c->EnsureHostBuffer();
        if (ok && m_cbLogical) memcpy(c->m_pHost, m_pHost, m_cbLogical);
        c->m_eState[HOST_MEMORY_SPACE_ID] = BSTATE_EXCLUSIVE;
    }
    Unlock();
    if (!ok) {
        c->Release();  // not yet referenced elsewhere: deleted
        return nullptr;
    }
    return c;
}

PBuffer *Datablock::GetPlatformBuffer(UINT ms) { return ms < PTK_MAX_MEMORY_SPACES ? m_pBuffers[ms] : nullptr; }
void Datablock::SetPoolOwner(BlockPool *pPool) { m_pPoolOwner = pPool; }
BlockPool *Datablock::GetPoolOwner() { return m_pPoolOwner; }

void Datablock::ResetForReuse() {
    m_luiControl = DBCTLC_NONE;
    m_uiRecordCount = 0;
    m_bSealed = FALSE;
}

void Datablock::MarkDeviceOnly(UINT ms) {
    Lock();
    if (ms < PTK_MAX_MEMORY_SPACES && m_pBuffers[ms]) {
        m_eState[ms] = BSTATE_EXCLUSIVE;
        if (m_eState[HOST_MEMORY_SPACE_ID] != BSTATE_NO_ENTRY) m_eState[HOST_MEMORY_SPACE_ID] = BSTATE_INVALID;
    }
    Unlock();
}

DatablockTemplate *Datablock::GetTemplate() { return m_pTemplate; }
UINT Datablock::GetDataBufferLogicalSizeBytes() { return m_cbLogical; }
UINT Datablock::GetDataBufferAllocatedSizeBytes() { return m_cbAllocated; }
UINT Datablock::GetMetaBufferLogicalSizeBytes() { return (UINT)m_vMeta.size(); }

void Datablock::GrowDataBuffer(UINT uiNewSize) {
    Lock();
    if (uiNewSize > m_cbAllocated) {
        // grow on the host; device copies are dropped and re-created on next use
        AcquireHostView(FALSE, TRUE);
        void *p = AllocHostStorage(uiNewSize);
        if (!p) {
            Unlock();
            Runtime::HandleError("out of host memory growing a datablock to %u bytes", uiNewSize);
            return;
        }
        memset(p, 0, uiNewSize);
        if (m_pHost) memcpy(p, m_pHost, m_cbLogical);
        free(m_pHost);
        m_pHost = (unsigned char *)p;
        for (UINT ms = 1; ms < PTK_MAX_MEMORY_SPACES; ++ms) {
            if (m_pBuffers[ms]) {
                m_pBuffers[ms]->WaitAllOnHost();
                delete m_pBuffers[ms];
                m_pBuffers[ms] = nullptr;
            }
            m_eState[ms] = BSTATE_NO_ENTRY;
        }
        m_eState[HOST_MEMORY_SPACE_ID] = BSTATE_EXCLUSIVE;
        m_cbAllocated = uiNewSize;
    }
    m_cbLogical = uiNewSize;
    Unlock();
}

void Datablock::Seal(UINT uiRecordCount, UINT uiDataSize, UINT uiMetaDataSize, UINT) {
    m_uiRecordCount = uiRecordCount;
    if (uiDataSize <= m_cbAllocated) m_cbLogical = uiDataSize;
    if (uiMetaDataSize) m_vMeta.resize(uiMetaDataSize, 0);
    m_bSealed = TRUE;
}

BOOL Datablock::IsSealed() { return m_bSealed; }
UINT Datablock::GetRecordCount() { return m_uiRecordCount; }
void Datablock::SetRecordCount(UINT n) { m_uiRecordCount = n; }
BOOL Datablock::HasValidDataBuffer() {
    for (UINT i = 0; i < PTK_MAX_MEMORY_SPACES; ++i)
        if (m_eState[i] == BSTATE_SHARED || m_eState[i] == BSTATE_EXCLUSIVE) return TRUE;
    return FALSE;
}
CONTROLSIGNAL Datablock::GetControlSignals() { return m_luiControl; }
BOOL Datablock::TestControlSignal(CONTROLSIGNAL c) { return TESTSIGNAL(m_luiControl, c); }
VOID Datablock::SetControlSignal(CONTROLSIGNAL c) { m_luiControl |= c; }
VOID Datablock::ClearControlSignal(CONTROLSIGNAL c) { m_luiControl &= ~c; }
VOID Datablock::ClearAllControlSignals() { m_luiControl = DBCTLC_NONE; }
BUFFERACCESSFLAGS Datablock::GetAccessFlags() { return m_eFlags; }
void Datablock::SetAccessFlags(BUFFERACCESSFLAGS e) { m_eFlags = e; }
UINT Datablock::GetDBUID() { return m_uiDBUID; }
BUFFER_COHERENCE_STATE Datablock::GetCoherenceState(UINT ms) {
    return ms < PTK_MAX_MEMORY_SPACES ? m_eState[ms] : BSTATE_NO_ENTRY;
}

// ---- BlockPool ------------------------------------------------------------------------------------------

BlockPool::BlockPool(DatablockTemplate *pTemplate, UINT uiMaxFree)
    : m_pTemplate(pTemplate), m_uiMaxFree(uiMaxFree), m_bClosed(FALSE) {
    if (m_pTemplate) m_pTemplate->AddRef();
    RegisterPool(this);
}

BlockPool::~BlockPool() {
    UnregisterPool(this);  // first: a reclaim walking the registry may be waiting to visit
    for (Datablock *b : m_vFree) {
        b->SetPoolOwner(nullptr);
        delete b;
    }
    if (m_pTemplate) m_pTemplate->Release();
}

Datablock *BlockPool::Request(UINT cb, BOOL bAtLeast, BOOL bCreate) {
    Datablock *reuse = nullptr;
    {
        std::lock_guard<std::mutex> g(m_lock);
        size_t best = m_vFree.size();
        for (size_t i = 0; i < m_vFree.size(); ++i) {
            const UINT have = m_vFree[i]->GetDataBufferAllocatedSizeBytes();
            if (have == cb) {
                best = i;
                break;
            }
            if (bAtLeast && have > cb &&
                (best == m_vFree.size() || have < m_vFree[best]->GetDataBufferAllocatedSizeBytes()))
                best = i;
        }
        if (best < m_vFree.size()) {
            reuse = m_vFree[best];
            m_vFree.erase(m_vFree.begin() + (long)best);
        }
    }
    if (reuse) {
        reuse->ResetForReuse();
        if (reuse->GetDataBufferLogicalSizeBytes() != cb) reuse->GrowDataBuffer(cb);  // sets the logical size
        reuse->AddRef();
        return reuse;
    }
    if (!bCreate) return nullptr;
    Datablock *b = new Datablock(m_pTemplate, cb, 0, DBCTLC_NONE);
    b->SetPoolOwner(this);
    AddRef();  // each live pooled block keeps its pool
    b->AddRef();
    return b;
}

UINT BlockPool::Prefill(UINT n, UINT cb, UINT cbMeta, Accelerator *pDevice) {
    UINT added = 0;
    for (UINT i = 0; i < n; ++i) {
        Datablock *b = new Datablock(m_pTemplate, cb, cbMeta, DBCTLC_NONE);
        // host storage now; device storage now too, so the first dispatch allocates nothing
        if (!b->AcquireHostView(FALSE, FALSE) || (pDevice && !b->ReserveDeviceBuffer(pDevice))) {
            delete b;
            break;
        }
        b->SetPoolOwner(this);
        AddRef();
        std::lock_guard<std::mutex> g(m_lock);
        m_vFree.push_back(b);
        ++added;
    }
    return added;
}

size_t BlockPool::Trim() {
    std::vector<Datablock *> v;
    {
        std::lock_guard<std::mutex> g(m_lock);
        v.swap(m_vFree);
    }
    size_t cb = 0;
    for (Datablock *b : v) {
        cb += b->GetFootprintBytes();
        b->SetPoolOwner(nullptr);
        delete b;
    }
    // the freed blocks' references on this pool go last: the caller holds one of its own
    for (size_t i = 0; i < v.size(); ++i) Release();
    return cb;
}

void BlockPool::Return(Datablock *pBlock) {
    bool keep = false;
    {
        std::lock_guard<std::mutex> g(m_lock);
        if (!m_bClosed && m_vFree.size() < m_uiMaxFree) {
            m_vFree.push_back(pBlock);
            keep = true;
        }
    }
    if (!keep) {
        pBlock->SetPoolOwner(nullptr);
        delete pBlock;
        Release();
    }
}

void BlockPool::Close() {
    std::vector<Datablock *> v;
    {
        std::lock_guard<std::mutex> g(m_lock);
        m_bClosed = TRUE;
        v.swap(m_vFree);
    }
    for (Datablock *b : v) {
        b->SetPoolOwner(nullptr);
        delete b;
        Release();
    }
}

size_t BlockPool::FreeCount() {
    std::lock_guard<std::mutex> g(m_lock);
    return m_vFree.size();
}

}  // namespace PTask
