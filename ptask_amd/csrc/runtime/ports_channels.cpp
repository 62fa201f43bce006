// ports_channels.cpp -- port and channel semantics (reference roles: port.cpp, InputPort.cpp,
// OutputPort.cpp, StickyPort.cpp, MetaPort.cpp, InitializerPort.cpp, channel.cpp and the
// channel subclasses).  Everything here runs with the owning graph's lock held, except the
// application-facing Channel::Push/Pull which take it themselves.

#include <chrono>
#include <cstring>

#include "ptk_internal.h"

namespace PTask {

// =============================================================================================
// Port
// =============================================================================================

Port::Port(PORTTYPE eType, DatablockTemplate *pTemplate, UINT uiId, const char *lpszBinding,
           UINT uiFormalIndex, UINT uiInOutIndex)
    : m_eType(eType), m_pTemplate(pTemplate), m_uiId(uiId),
      m_strBinding(lpszBinding ? lpszBinding : ""), m_uiFormalIndex(uiFormalIndex),
      m_uiInOutIndex(uiInOutIndex), m_pTask(nullptr), m_uiTaskIndex(0), m_bMarshallable(TRUE),
      m_eMetaFunction(MF_NONE), m_bPermanentBlock(FALSE), m_pDescribedPort(nullptr) {
    if (m_pTemplate) m_pTemplate->AddRef();
}

Port::~Port() {
    if (m_pTemplate) m_pTemplate->Release();
}

BOOL Port::IsFormalParameter() { return m_uiFormalIndex != (UINT)-1; }

void Port::BindDependentAccelerator(ACCELERATOR_CLASS accClass, int nIndex) {
    m_bDependentBinding = TRUE;
    m_eDependentClass = accClass == ACCELERATOR_CLASS_CUDA ? ACCELERATOR_CLASS_HIP : accClass;
    m_nDependentIndex = nIndex;
}

void Port::BindChannel(Channel *pChannel) { m_vChannels.push_back(pChannel); }
void Port::BindControlChannel(Channel *pChannel) { m_vControlChannels.push_back(pChannel); }

void Port::PropagateControlSignal(CONTROLSIGNAL luiCode) {
    if (!HASSIGNAL(luiCode)) return;
    for (Port *p : m_vControlPropagationPorts)
        if (p->GetPortType() == OUTPUT_PORT) static_cast<OutputPort *>(p)->AddPropagatedControlSignal(luiCode);
}

// =============================================================================================
// InputPort
// =============================================================================================

InputPort::InputPort(DatablockTemplate *pTemplate, UINT uiId, const char *lpszBinding,
                     UINT uiFormalIndex, UINT uiInOutIndex, PORTTYPE eType)
    : Port(eType, pTemplate, uiId, lpszBinding, uiFormalIndex, uiInOutIndex), m_pInOutConsumer(nullptr) {}

InputPort::~InputPort() {}

// The first channel (switch channels first) whose head block passes its destination predicate;
// blocks that fail are released on the spot (PFA_RELEASE_BLOCK).
Channel *InputPort::NextReadyChannel() {
    for (int pass = 0; pass < 2; ++pass) {
        std::vector<Channel *> &v = pass == 0 ? m_vControlChannels : m_vChannels;
        for (Channel *c : v) {
            if (c->GetType() == CT_INITIALIZER) {
                if (c->IsReady(CE_DST)) return c;
                continue;
            }
            for (;;) {
                Datablock *b = c->PeekLocked();
                if (!b) break;
                if (c->PassesPredicate(CE_DST, b)) return c;
                Datablock *drop = c->DequeueLocked();
                drop->Release();
            }
        }
    }
    return nullptr;
}

BOOL InputPort::IsReady() { return NextReadyChannel() != nullptr; }

Datablock *InputPort::Peek() {
    Channel *c = NextReadyChannel();
    if (!c) return nullptr;
    return c->GetType() == CT_INITIALIZER ? nullptr : c->PeekLocked();
}

Datablock *InputPort::Pull() {
    Channel *c = NextReadyChannel();
    if (!c) return nullptr;
    if (c->GetType() == CT_INITIALIZER) return static_cast<InitializerChannel *>(c)->Produce();
    return c->DequeueLocked();
}

void InputPort::Drain() {
    for (int pass = 0; pass < 2; ++pass) {
        std::vector<Channel *> &v = pass == 0 ? m_vControlChannels : m_vChannels;
        for (Channel *c : v) c->Drain();
    }
}

// =============================================================================================
// StickyPort
// =============================================================================================

StickyPort::StickyPort(DatablockTemplate *pTemplate, UINT uiId, const char *lpszBinding, UINT uiFormalIndex)
    : InputPort(pTemplate, uiId, lpszBinding, uiFormalIndex, (UINT)-1, STICKY_PORT), m_pSticky(nullptr) {}

StickyPort::~StickyPort() { PTSRELEASE(m_pSticky); }

BOOL StickyPort::IsReady() { return m_pSticky != nullptr || InputPort::IsReady(); }

Datablock *StickyPort::Peek() {
    Datablock *b = InputPort::Peek();
    return b ? b : m_pSticky;
}

Datablock *StickyPort::Pull() {
    Datablock *b = InputPort::Pull();
    if (b) {
        PTSRELEASE(m_pSticky);
        m_pSticky = b;  // the port keeps the channel's reference
    }
    if (m_pSticky) m_pSticky->AddRef();
    return m_pSticky;
}

void StickyPort::Drain() {
    InputPort::Drain();
    PTSRELEASE(m_pSticky);
}

// =============================================================================================
// MetaPort
// =============================================================================================

// uiAllocationOutput: for an allocation-size meta port, the index of the output whose blocks it
// sizes (the reference's CreatePort in/out index, MetaPort.cpp:724)
MetaPort::MetaPort(DatablockTemplate *pTemplate, UINT uiId, const char *lpszBinding, UINT uiFormalIndex,
                   UINT uiAllocationOutput)
    : InputPort(pTemplate, uiId, lpszBinding, uiFormalIndex, uiAllocationOutput, META_PORT),
      m_pIterationBlock(nullptr), m_uiIterMax(0), m_uiIterCount(0), m_pAllocationPort(nullptr) {}

MetaPort::~MetaPort() { PTSRELEASE(m_pIterationBlock); }

BOOL MetaPort::IsReady() {
    if (m_eMetaFunction == MF_GENERAL_ITERATOR && m_pIterationBlock) return TRUE;
    return InputPort::IsReady();
}

static UINT ReadUint(Datablock *b) {
    UINT v = 0;
    if (!b) return 0;
    const void *p = b->AcquireHostView(FALSE, TRUE);
    if (p && b->GetDataBufferLogicalSizeBytes() >= sizeof(UINT)) memcpy(&v, p, sizeof(UINT));
    return v;
}

// reference: MetaPort::ConfigureGeneralIteration / ConfigureSimpleIteration /
// allocation-size handling (MetaPort.cpp:724-1010)
void MetaPort::Configure(UINT *puiDispatchIterations, UINT *puiAllocationBytes, BOOL *pbHasAllocation) {
    switch (m_eMetaFunction) {
        case MF_GENERAL_ITERATOR: {
            if (!m_pIterationBlock) {
                m_pIterationBlock = InputPort::Pull();
                if (!m_pIterationBlock) return;
                m_uiIterMax = ReadUint(m_pIterationBlock);
                m_uiIterCount = 0;
                if (m_uiIterMax == 0) {
                    Runtime::HandleError("%s: a general iterator needs at least one iteration "
                                         "(loop bounds are read at dispatch time)",
                                         GetVariableBinding());
                    m_uiIterMax = 1;
                }
                CONTROLSIGNAL c = m_pIterationBlock->GetControlSignals();
                PropagateControlSignal(HASSIGNAL(c) ? c : DBCTLC_BEGINITERATION);
            }
            ++m_uiIterCount;
            if (m_uiIterCount >= m_uiIterMax) {
                PropagateControlSignal(DBCTLC_ENDITERATION);
                PTSRELEASE(m_pIterationBlock);
                m_uiIterMax = m_uiIterCount = 0;
            }
            *puiDispatchIterations = 1;
            break;
        }
        case MF_SIMPLE_ITERATOR: {
            Datablock *b = InputPort::Pull();
            if (b) {
                *puiDispatchIterations = ReadUint(b);
                PropagateControlSignal(b->GetControlSignals());
                b->Release();
            }
            break;
        }
        case MF_ALLOCATION_SIZE:
        case MF_PADDED_ALLOCATION_SIZE: {
            Datablock *b = InputPort::Pull();
            if (b) {
                *puiAllocationBytes = ReadUint(b);
                *pbHasAllocation = TRUE;
                PropagateControlSignal(b->GetControlSignals());
                b->Release();
            }
            break;
        }
        default: {
            Datablock *b = InputPort::Pull();
            if (b) {
                PropagateControlSignal(b->GetControlSignals());
                b->Release();
            }
            break;
        }
    }
}

void MetaPort::Finalize() {}

void MetaPort::Drain() {
    InputPort::Drain();
    PTSRELEASE(m_pIterationBlock);
    m_uiIterMax = m_uiIterCount = 0;
}

// =============================================================================================
// InitializerPort
// =============================================================================================

InitializerPort::InitializerPort(DatablockTemplate *pTemplate, UINT uiId, const char *lpszBinding,
                                 UINT uiFormalIndex, UINT uiInOutIndex)
    : InputPort(pTemplate, uiId, lpszBinding, uiFormalIndex, uiInOutIndex, INITIALIZER_PORT) {}

BOOL InitializerPort::IsReady() { return TRUE; }

Datablock *InitializerPort::Pull() {
    DatablockTemplate *t = m_pTemplate;
    UINT cb = t ? t->GetDatablockByteCount() : 0;
    Datablock *b = new Datablock(t, cb, 0, DBCTLC_NONE);
    b->AddRef();
    if (t && t->HasInitialValue() && t->GetInitialValue()) {
        unsigned char *p = (unsigned char *)b->AcquireHostView(TRUE, FALSE);
        UINT n = t->GetInitialValueSizeBytes();
        // a short initial value is replicated across the block (one record repeated)
        for (UINT off = 0; n && off < cb; off += n) memcpy(p + off, t->GetInitialValue(), std::min(n, cb - off));
    }
    return b;
}

// =============================================================================================
// OutputPort
// =============================================================================================

OutputPort::OutputPort(DatablockTemplate *pTemplate, UINT uiId, const char *lpszBinding,
                       UINT uiFormalIndex, UINT uiInOutIndex)
    : Port(OUTPUT_PORT, pTemplate, uiId, lpszBinding, uiFormalIndex, uiInOutIndex),
      m_pInOutProducer(nullptr), m_luiPropagated(DBCTLC_NONE), m_pControlPort(nullptr),
      m_bGateOpen(TRUE), m_bInitiallyOpen(TRUE), m_pPool(nullptr) {}

OutputPort::~OutputPort() { ReleasePool(); }

BOOL OutputPort::IsFormalParameter() {
    return m_uiFormalIndex != (UINT)-1 && m_pInOutProducer == nullptr;
}

BOOL OutputPort::HasCapacity() {
    for (Channel *c : m_vChannels)
        if (!c->HasCapacity()) return FALSE;
    return TRUE;
}

Datablock *OutputPort::GetDestinationBlock(UINT cbBytes) {
    UINT cb = cbBytes ? cbBytes : (m_pTemplate ? m_pTemplate->GetDatablockByteCount() : 0);
    if (cb == 0) {
        Runtime::HandleError("%s: output of a variable-size template needs an allocation meta port",
                             GetVariableBinding());
        return nullptr;
    }
    if (!m_pPool) {
        m_pPool = new BlockPool(m_pTemplate, RT().blockPoolSize);
        m_pPool->AddRef();
    }
    Datablock *b = m_pPool->Request(cb);
    if (m_pTemplate && m_pTemplate->GetStride() && !m_pTemplate->IsScalarParameter())
        b->SetRecordCount(cb / m_pTemplate->GetStride());
    return b;
}

// reference: OutputPort::Push (OutputPort.cpp:340-440), GetDownstreamWriterPortCount and
// GetDownstreamReadonlyPortCount (OutputPort.cpp:1641-1750).  Of the channels a block passes, a
// graph output channel (the application may write what it pulls) and a channel into an in/out
// port are writers, the rest readers.  Readers share the block.  When a reader and a writer, or
// two writers, share it, every writer gets its own copy -- but the first writer keeps the block
// when there is no reader -- so no consumer sees another's writes.  SetSuppressClones(TRUE) on
// the port pushes the one block to all.
void OutputPort::Push(Datablock *pBlock) {
    if (!pBlock) return;
    if (IsGateOpen()) {
        auto writer = [](Channel *c) {
            if (c->GetType() == CT_GRAPH_OUTPUT) return true;
            Port *d = c->GetBoundPort(CE_DST);
            return d && d->GetPortType() == INPUT_PORT && static_cast<InputPort *>(d)->GetInOutConsumer() != nullptr;
        };
        UINT nReaders = 0, nWriters = 0;
        for (Channel *c : m_vChannels) {
            if (!c->PassesPredicate(CE_SRC, pBlock)) continue;
            if (writer(c)) ++nWriters;
            else ++nReaders;
        }
        const bool clones = !m_bSuppressClones && ((nReaders && nWriters) || nWriters > 1);
        UINT writersSeen = 0;
        for (Channel *c : m_vChannels) {
            if (!c->PassesPredicate(CE_SRC, pBlock)) continue;
            if (clones && writer(c) && (nReaders > 0 || writersSeen++ > 0)) {
                Graph *g = m_pTask ? m_pTask->GetGraph() : nullptr;
                if (g && g->GetReplay()) g->GetReplay()->OnBlockCopied();
                Datablock *copy = pBlock->CloneForWriter();
                if (!copy) {
                    Runtime::HandleError("%s: no memory for the copy a writing consumer needs", GetVariableBinding());
                    continue;
                }
                c->Enqueue(copy);
                copy->Release();
                continue;
            }
            c->Enqueue(pBlock);
        }
    }
    if (m_pControlPort) m_bGateOpen = m_bInitiallyOpen;
    pBlock->Release();
}

void OutputPort::ReleasePool() {
    if (m_pPool) {
        m_pPool->Close();
        m_pPool->Release();
        m_pPool = nullptr;
    }
}

// =============================================================================================
// Channel
// =============================================================================================

Channel::Channel(Graph *pGraph, DatablockTemplate *pTemplate, const char *lpszName, CHANNELTYPE eType)
    : m_pGraph(pGraph), m_pTemplate(pTemplate), m_strName(lpszName ? lpszName : ""), m_eType(eType),
      m_uiCapacity((UINT)Runtime::GetDefaultChannelCapacity()), m_bDraw(true), m_dwTimeout(INFINITE),
      m_pSrcPort(nullptr), m_pDstPort(nullptr), m_eViewPolicy(Runtime::GetDefaultOutputViewMaterializationPolicy()) {
    if (m_pTemplate) m_pTemplate->AddRef();
    for (int i = 0; i < 2; ++i) {
        m_ePredicate[i] = CGATEFN_NONE;
        m_lpfnPredicate[i] = nullptr;
        m_ePFA[i] = PFA_RELEASE_BLOCK;
    }
}

Channel::~Channel() {
    for (Datablock *b : m_q) b->Release();
    m_q.clear();
    if (m_pTemplate) m_pTemplate->Release();
}

CONTROLSIGNAL Channel::GetControlSignalsOfInterest(CHANNELPREDICATE e) {
    switch (e) {
        case CGATEFN_CLOSE_ON_EOF:
        case CGATEFN_OPEN_ON_EOF: return DBCTLC_EOF;
        case CGATEFN_CLOSE_ON_BOF:
        case CGATEFN_OPEN_ON_BOF: return DBCTLC_BOF;
        case CGATEFN_OPEN_ON_BEGINITERATION:
        case CGATEFN_CLOSE_ON_BEGINITERATION: return DBCTLC_BEGINITERATION;
        case CGATEFN_OPEN_ON_ENDITERATION:
        case CGATEFN_CLOSE_ON_ENDITERATION: return DBCTLC_ENDITERATION;
        default: return DBCTLC_NONE;
    }
}

// reference: Channel::PassesPredicate, channel.cpp:1619 -- the canonical gate functions
BOOL Channel::PassesPredicate(CHANNELENDPOINTTYPE e, Datablock *pBlock) {
    CHANNELPREDICATE p = m_ePredicate[e];
    if (p == CGATEFN_NONE) return TRUE;
    CONTROLSIGNAL c = pBlock ? pBlock->GetControlSignals() : DBCTLC_NONE;
    switch (p) {
        case CGATEFN_CLOSE_ON_EOF: return !TESTSIGNAL(c, DBCTLC_EOF);
        case CGATEFN_OPEN_ON_EOF: return TESTSIGNAL(c, DBCTLC_EOF);
        case CGATEFN_OPEN_ON_BEGINITERATION: return TESTSIGNAL(c, DBCTLC_BEGINITERATION);
        case CGATEFN_CLOSE_ON_BEGINITERATION: return !TESTSIGNAL(c, DBCTLC_BEGINITERATION);
        case CGATEFN_OPEN_ON_ENDITERATION: return TESTSIGNAL(c, DBCTLC_ENDITERATION);
        case CGATEFN_CLOSE_ON_ENDITERATION: return !TESTSIGNAL(c, DBCTLC_ENDITERATION);
        case CGATEFN_DEVNULL: return FALSE;
        case CGATEFN_CLOSE_ON_BOF: return !TESTSIGNAL(c, DBCTLC_BOF);
        case CGATEFN_OPEN_ON_BOF: return TESTSIGNAL(c, DBCTLC_BOF);
        case CGATEFN_USER_DEFINED:
            return m_lpfnPredicate[e] ? m_lpfnPredicate[e](this, pBlock, m_ePFA[e]) : TRUE;
        default: return TRUE;
    }
}

void Channel::SetPredicator(CHANNELENDPOINTTYPE e, LPFNCHANNELPREDICATE lpfn) {
    m_lpfnPredicate[e] = lpfn;
    m_ePredicate[e] = lpfn ? CGATEFN_USER_DEFINED : CGATEFN_NONE;
}

void Channel::SetPredicationType(CHANNELENDPOINTTYPE e, CHANNELPREDICATE ePredicate) {
    m_ePredicate[e] = ePredicate;
}

void Channel::SetPredicationType(CHANNELENDPOINTTYPE e, CHANNELPREDICATE ePredicate,
                                 PREDICATE_FAILURE_ACTION pfa) {
    m_ePredicate[e] = ePredicate;
    m_ePFA[e] = pfa;
}

void Channel::BindPort(Port *pPort, CHANNELENDPOINTTYPE type) {
    if (type == CE_SRC) m_pSrcPort = pPort;
    else m_pDstPort = pPort;
}

Port *Channel::UnbindPort(CHANNELENDPOINTTYPE type) {
    Port *p = GetBoundPort(type);
    BindPort(nullptr, type);
    return p;
}

Port *Channel::GetBoundPort(CHANNELENDPOINTTYPE type) { return type == CE_SRC ? m_pSrcPort : m_pDstPort; }

BOOL Channel::HasCapacity() { return m_q.size() < m_uiCapacity; }

void Channel::PushDescriptors(Datablock *pBlock) {
    (void)pBlock;
}

void Channel::Enqueue(Datablock *pBlock) {
    if (m_eType == CT_GRAPH_OUTPUT && m_pGraph->GetReplay()) m_pGraph->GetReplay()->OnOutputEnqueue(this, pBlock);
    pBlock->AddRef();
    m_q.push_back(pBlock);
    if (m_pDstPort) m_pGraph->SignalPortTask(m_pDstPort);
    if (m_eType == CT_GRAPH_OUTPUT) m_pGraph->NotifyDataArrived();
}

Datablock *Channel::PeekLocked() { return m_q.empty() ? nullptr : m_q.front(); }

Datablock *Channel::DequeueLocked() {
    if (m_q.empty()) return nullptr;
    Datablock *b = m_q.front();
    m_q.pop_front();
    if (m_pSrcPort) m_pGraph->SignalPortTask(m_pSrcPort);
    m_pGraph->NotifyRoom();
    return b;
}

BOOL Channel::IsReady(CHANNELENDPOINTTYPE type) {
    std::lock_guard<std::recursive_mutex> g(m_pGraph->GraphLock());
    return type == CE_DST ? !m_q.empty() : HasCapacity();
}

BOOL GraphOutputChannel::IsReady(CHANNELENDPOINTTYPE type) {
    std::lock_guard<std::recursive_mutex> g(m_pGraph->GraphLock());
    return type == CE_DST ? !m_q.empty() : TRUE;
}

BOOL Channel::PushToQueue(Datablock *pBlock, DWORD dwTimeout) {
    std::unique_lock<std::recursive_mutex> lk(m_pGraph->GraphLock());
    if (m_eType == CT_GRAPH_INPUT || m_eType == CT_MULTI) {
        auto room = [&] { return HasCapacity() || !m_pGraph->IsAlive(); };
        if (dwTimeout == INFINITE) m_pGraph->RoomAvailable().wait(lk, room);
        else if (!m_pGraph->RoomAvailable().wait_for(lk, std::chrono::milliseconds(dwTimeout), room))
            return FALSE;
        if (!m_pGraph->IsAlive()) return FALSE;
        if (m_eType == CT_GRAPH_INPUT && m_pGraph->GetReplay()) m_pGraph->GetReplay()->OnInputPush(this, pBlock);
    }
    Enqueue(pBlock);
    return TRUE;
}

// reference: Channel::Push, channel.cpp:1125 -- the channel takes its own reference
BOOL Channel::Push(Datablock *pBlock, DWORD dwTimeout) {
    if (!pBlock) return FALSE;
    return PushToQueue(pBlock, dwTimeout == 0xFFFFFFFF ? m_dwTimeout : dwTimeout);
}

// reference: Channel::Pull, channel.cpp:523 -- the caller owns the returned reference
Datablock *Channel::Pull(DWORD dwTimeout) {
    std::unique_lock<std::recursive_mutex> lk(m_pGraph->GraphLock());
    if (dwTimeout == 0xFFFFFFFF) dwTimeout = m_dwTimeout;
    if (m_eType == CT_GRAPH_OUTPUT) {
        auto has = [&] { return !m_q.empty() || !m_pGraph->IsAlive(); };
        if (dwTimeout == INFINITE) m_pGraph->DataArrived().wait(lk, has);
        else m_pGraph->DataArrived().wait_for(lk, std::chrono::milliseconds(dwTimeout), has);
    }
    return DequeueLocked();
}

Datablock *Channel::Peek() {
    std::lock_guard<std::recursive_mutex> g(m_pGraph->GraphLock());
    return PeekLocked();
}

void Channel::SetCapacity(UINT n) { m_uiCapacity = n ? n : 1; }
UINT Channel::GetCapacity() { return m_uiCapacity; }

size_t Channel::GetQueueDepth() {
    std::lock_guard<std::recursive_mutex> g(m_pGraph->GraphLock());
    return m_q.size();
}

void Channel::Drain() {
    std::lock_guard<std::recursive_mutex> g(m_pGraph->GraphLock());
    for (Datablock *b : m_q) b->Release();
    m_q.clear();
    m_pGraph->NotifyRoom();
}

// ---- InitializerChannel -------------------------------------------------------------------------

Datablock *InitializerChannel::Produce() {
    DatablockTemplate *t = m_pTemplate;
    UINT cb = t ? t->GetDatablockByteCount() : 0;
    Datablock *b = new Datablock(t, cb, 0, m_luiControl);
    b->AddRef();
    if (t && t->HasInitialValue() && t->GetInitialValue()) {
        unsigned char *p = (unsigned char *)b->AcquireHostView(TRUE, FALSE);
        UINT n = t->GetInitialValueSizeBytes();
        for (UINT off = 0; n && off < cb; off += n) memcpy(p + off, t->GetInitialValue(), std::min(n, cb - off));
    }
    return b;
}

BOOL InitializerChannel::IsReady(CHANNELENDPOINTTYPE type) {
    if (type == CE_SRC) return TRUE;
    if (m_ePredicate[CE_DST] == CGATEFN_NONE) return TRUE;
    // judge the predicate on the control code a produced block would carry
    Datablock probe(nullptr, 0, 0, m_luiControl);
    probe.AddRef();
    BOOL ok = PassesPredicate(CE_DST, &probe);
    return ok;
}

Datablock *InitializerChannel::Pull(DWORD) { return Produce(); }

// ---- MultiChannel --------------------------------------------------------------------------------

BOOL MultiChannel::Push(Datablock *pBlock, DWORD dwTimeout) {
    if (!pBlock) return FALSE;
    std::unique_lock<std::recursive_mutex> lk(m_pGraph->GraphLock());
    auto room = [&] {
        if (!m_pGraph->IsAlive()) return true;
        for (Channel *c : m_vMembers)
            if (!c->HasCapacity()) return false;
        return true;
    };
    if (dwTimeout == 0xFFFFFFFF || dwTimeout == INFINITE) m_pGraph->RoomAvailable().wait(lk, room);
    else if (!m_pGraph->RoomAvailable().wait_for(lk, std::chrono::milliseconds(dwTimeout), room)) return FALSE;
    for (Channel *c : m_vMembers) c->Enqueue(pBlock);
    return TRUE;
}

BOOL MultiChannel::IsReady(CHANNELENDPOINTTYPE type) {
    std::lock_guard<std::recursive_mutex> g(m_pGraph->GraphLock());
    for (Channel *c : m_vMembers) {
        if (type == CE_SRC && !c->HasCapacity()) return FALSE;
        if (type == CE_DST && c->QueueDepthLocked() == 0) return FALSE;
    }
    return TRUE;
}

}  // namespace PTask
