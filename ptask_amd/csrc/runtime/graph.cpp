// graph.cpp -- tasks and graphs: parameter layout, readiness, the three-phase dispatch and the
// runner threads (reference roles: task.cpp Schedule/Execute/Dispatch :1407-3906,
// cutask.cpp ComputeParameterOffsets :435-528, graph.cpp AddTask/Run/runner threads
// :829-4070, Scheduler.cpp Schedule :4551-4673).

#include <algorithm>
#include <chrono>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>

#include "ptk_internal.h"

namespace PTask {

// =============================================================================================
// Task
// =============================================================================================

Task::Task(Graph *pGraph, CompiledKernel *pKernel, ACCELERATOR_CLASS eClass, const char *lpszName)
    : m_bQueued(FALSE), m_pGraph(pGraph), m_pKernel(pKernel), m_eClass(eClass),
      m_strName(lpszName ? lpszName : ""), m_uiArgBytes(0), m_bLayoutDone(FALSE), m_tgx(1), m_tgy(1),
      m_tgz(1), m_bExplicitGeometry(FALSE), m_lpfnEstimator(nullptr), m_nPriority(0),
      m_pMandatoryAccelerator(nullptr), m_bUserCodeAllocates(FALSE), m_bDispatching(FALSE),
      m_uiDispatchCount(0), m_dLastDeviceMs(0), m_dTotalDeviceMs(0) {}

Task::~Task() {
    for (Port *p : m_vInputs) p->Release();
    for (Port *p : m_vOutputs) p->Release();
}

void Task::SetComputeGeometry(int tgx, int tgy, int tgz) {
    m_tgx = std::max(1, tgx);
    m_tgy = std::max(1, tgy);
    m_tgz = std::max(1, tgz);
}

void Task::SetBlockAndGridSize(PTASKDIM3 grid, PTASKDIM3 block) {
    m_grid = grid;
    m_block = block;
    m_bExplicitGeometry = TRUE;
}

void Task::SetPriority(int nPTaskPrio, int) { m_nPriority = nPTaskPrio; }

BOOL Task::SetAffinity(Accelerator *pAccelerator, AFFINITYTYPE eType) {
    if (!pAccelerator) return FALSE;
    if (eType == AFFINITYTYPE_MANDATORY || eType == AFFINITYTYPE_STRONG) {
        if (pAccelerator->GetClass() != m_eClass) return FALSE;
        m_pMandatoryAccelerator = pAccelerator;
    }
    return TRUE;
}

// reference: Task::BindDependentAcceleratorClass (task.cpp): one instance per class, recorded
// for the dispatch; the dependent accelerator is the HIP device (CUDA requests map to it).
BOOL Task::BindDependentAcceleratorClass(ACCELERATOR_CLASS cls, int nRequiredInstances,
                                         BOOL bRequestDependentPSObjects) {
    if (cls == ACCELERATOR_CLASS_CUDA) cls = ACCELERATOR_CLASS_HIP;
    if (m_eClass != ACCELERATOR_CLASS_HOST)
        return Runtime::HandleError("%s: dependent accelerators are for host tasks", GetTaskName());
    if (cls != ACCELERATOR_CLASS_HIP)
        return Runtime::HandleError("%s: dependent accelerator class %d has no backend here", GetTaskName(), (int)cls);
    if (nRequiredInstances != 1 ||
        std::find(m_vDependentClasses.begin(), m_vDependentClasses.end(), cls) != m_vDependentClasses.end())
        return Runtime::HandleError("%s: malformed dependent binding: %d instance(s), at most one per class",
                                   GetTaskName(), nRequiredInstances);
    m_vDependentClasses.push_back(cls);
    m_bRequestDependentPSObjects = bRequestDependentPSObjects;
    return TRUE;
}

void Task::AttachPorts(UINT nIn, Port **ppIn, UINT nOut, Port **ppOut) {
    for (UINT i = 0; i < nIn; ++i) {
        ppIn[i]->AddRef();
        ppIn[i]->BindTask(this, i);
        m_vInputs.push_back(ppIn[i]);
    }
    for (UINT o = 0; o < nOut; ++o) {
        if (ppOut[o]->GetPortType() != OUTPUT_PORT)
            Runtime::HandleError("%s: output %u is not an OUTPUT_PORT", GetTaskName(), o);
        ppOut[o]->AddRef();
        ppOut[o]->BindTask(this, o);
        m_vOutputs.push_back(static_cast<OutputPort *>(ppOut[o]));
    }
    // in/out pairs: an input's routing index names the output that re-emits its block
    for (Port *p : m_vInputs) {
        if (!p->IsInOutParameter()) continue;
        UINT k = p->GetInOutRoutingIndex();
        if (p->GetPortType() == META_PORT) {
            if (k < m_vOutputs.size()) static_cast<MetaPort *>(p)->SetAllocationPort(m_vOutputs[k]);
            continue;
        }
        if (k >= m_vOutputs.size()) {
            Runtime::HandleError("%s: in/out port %s routes to missing output %u", GetTaskName(),
                                 p->GetVariableBinding(), k);
            continue;
        }
        static_cast<InputPort *>(p)->SetInOutConsumer(m_vOutputs[k]);
        m_vOutputs[k]->SetInOutProducer(static_cast<InputPort *>(p));
    }
}

// reference: CUTask::ComputeParameterOffsets (cutask.cpp:435-528) -- formal parameters in
// index order; INT and FLOAT at 4-byte alignment, by-value structs at 4-byte alignment with the
// template's byte count, buffers as 8-byte device pointers.
void Task::ComputeParameterLayout() {
    if (m_bLayoutDone) return;
    std::multimap<UINT, Binding> byIndex;
    for (size_t i = 0; i < m_vInputs.size(); ++i) {
        Port *p = m_vInputs[i];
        if (!p->IsFormalParameter()) continue;
        Binding b{};
        b.pPort = p;
        b.bInput = TRUE;
        b.uiSlot = (UINT)i;
        b.bWritable = p->GetPortType() != META_PORT && static_cast<InputPort *>(p)->GetInOutConsumer() != nullptr;
        byIndex.insert(std::make_pair(p->GetFormalParameterIndex(), b));
    }
    for (size_t o = 0; o < m_vOutputs.size(); ++o) {
        OutputPort *p = m_vOutputs[o];
        if (!p->IsFormalParameter()) continue;
        Binding b{};
        b.pPort = p;
        b.bInput = FALSE;
        b.uiSlot = (UINT)o;
        b.bWritable = TRUE;
        byIndex.insert(std::make_pair(p->GetFormalParameterIndex(), b));
    }
    UINT off = 0, expect = 0;
    for (auto &kv : byIndex) {
        Binding b = kv.second;
        if (kv.first != expect)
            Runtime::Warning((std::string(GetTaskName()) + ": formal parameter indices are not dense").c_str());
        expect = kv.first + 1;
        DatablockTemplate *t = b.pPort->GetTemplate();
        b.eType = t ? t->GetParameterBaseType() : PTPARM_BUFFER;
        auto align = [&](UINT a) { off = (off + a - 1) / a * a; };
        switch (b.eType) {
            case PTPARM_INT:
            case PTPARM_FLOAT:
                align(4);
                b.uiOffset = off;
                b.uiBytes = 4;
                break;
            case PTPARM_BYVALSTRUCT:
                align(4);
                b.uiOffset = off;
                b.uiBytes = t->GetDatablockByteCount();
                break;
            case PTPARM_DOUBLE:
                align(8);
                b.uiOffset = off;
                b.uiBytes = 8;
                break;
            default:
                align(8);
                b.uiOffset = off;
                b.uiBytes = 8;
                b.eType = PTPARM_BUFFER;
                break;
        }
        off += b.uiBytes;
        m_vBindings.push_back(b);
    }
    m_uiArgBytes = off;
    m_bLayoutDone = TRUE;
}

void Task::OnGraphComplete() { ComputeParameterLayout(); }

// Launch geometry: explicit values win, then a geometry estimator, then the compute geometry
// covered by 256-wide blocks.
void Task::GeometryFor(std::vector<Datablock *> &vIn, PTASKDIM3 *pGrid, PTASKDIM3 *pBlock) {
    if (m_bExplicitGeometry) {
        *pGrid = m_grid;
        *pBlock = m_block;
    } else if (m_lpfnEstimator) {
        m_lpfnEstimator((UINT)vIn.size(), vIn.data(), pBlock, pGrid);
    } else {
        int bx = 256;
        *pBlock = PTASKDIM3(bx, 1, 1);
        *pGrid = PTASKDIM3((m_tgx + bx - 1) / bx, m_tgy, m_tgz);
    }
    if (pGrid->x < 1) pGrid->x = 1;
    if (pGrid->y < 1) pGrid->y = 1;
    if (pGrid->z < 1) pGrid->z = 1;
    if (pBlock->x < 1) pBlock->x = 1;
    if (pBlock->y < 1) pBlock->y = 1;
    if (pBlock->z < 1) pBlock->z = 1;
}

// reference: Task::IsReadyForDispatch, task.cpp:1113
BOOL Task::IsReadyForDispatch() {
    if (m_bDispatching) return FALSE;
    for (Port *p : m_vInputs) {
        if (p->GetDescribedPort()) continue;  // fed by the port it describes
        switch (p->GetPortType()) {
            case META_PORT:
                if (!static_cast<MetaPort *>(p)->IsReady()) return FALSE;
                break;
            case INITIALIZER_PORT: break;
            default:
                if (!static_cast<InputPort *>(p)->IsReady()) return FALSE;
                break;
        }
    }
    for (OutputPort *o : m_vOutputs)
        if (!o->HasCapacity()) return FALSE;
    return TRUE;
}

static Datablock *MakeDescriptorBlock(Datablock *pDescribed, DESCRIPTORFUNC eFunc, DatablockTemplate *pTemplate) {
    RuntimeState &S = RT();
    if (eFunc == DF_SIZE) {
        // reference: Datablock::CreateSizeDescriptorBlock -- the record count, or bytes/stride
        UINT n = pDescribed ? pDescribed->GetRecordCount() : 0;
        if (pDescribed && pDescribed->GetTemplate() && pDescribed->GetTemplate()->GetStride()) {
            UINT m = pDescribed->GetDataBufferLogicalSizeBytes() / pDescribed->GetTemplate()->GetStride();
            if (m > n) n = m;
        }
        Datablock *b = new Datablock(pTemplate ? pTemplate : S.sizeDescriptorTemplate, sizeof(int), 0, DBCTLC_NONE);
        b->AddRef();
        memcpy(b->AcquireHostView(TRUE, FALSE), &n, sizeof(int));
        return b;
    }
    CONTROLSIGNAL c = pDescribed ? pDescribed->GetControlSignals() : DBCTLC_NONE;
    CONTROLSIGNAL v = 0;
    if (eFunc == DF_EOF) v = TESTSIGNAL(c, DBCTLC_EOF);
    else if (eFunc == DF_BOF) v = TESTSIGNAL(c, DBCTLC_BOF);
    else if (eFunc == DF_CONTROL_CODE) v = c;
    else if (eFunc == DF_BLOCK_UID) v = pDescribed ? pDescribed->GetDBUID() : 0;
    UINT cb = pTemplate && pTemplate->GetDatablockByteCount() ? pTemplate->GetDatablockByteCount() : sizeof(CONTROLSIGNAL);
    Datablock *b = new Datablock(pTemplate ? pTemplate : S.controlTemplate, cb, 0, DBCTLC_NONE);
    b->AddRef();
    memcpy(b->AcquireHostView(TRUE, FALSE), &v, std::min<size_t>(cb, sizeof(v)));
    return b;
}

// reference: Task::Dispatch, task.cpp:3822-3906
BOOL Task::AttemptDispatch(std::unique_lock<std::recursive_mutex> *pLock) {
    GraphReplay *replay = m_pGraph->GetReplay();
    if (replay && replay->IsArmed()) return replay->TryReplay();  // no task runs on its own then
    if (!IsReadyForDispatch()) return FALSE;
    ComputeParameterLayout();
    m_bDispatching = TRUE;
    const size_t nIn = m_vInputs.size(), nOut = m_vOutputs.size();
    std::vector<Datablock *> vIn(nIn, nullptr), vOut(nOut, nullptr);
    UINT uiIterations = 1;
    std::vector<UINT> allocBytes(nOut, 0);

    // phase 1a: take a block from every data input; propagate and gate on its control code
    for (size_t i = 0; i < nIn; ++i) {
        Port *p = m_vInputs[i];
        if (p->GetPortType() == META_PORT || p->GetDescribedPort()) continue;
        Datablock *b = static_cast<InputPort *>(p)->Pull();
        vIn[i] = b;
        if (!b) continue;
        CONTROLSIGNAL c = b->GetControlSignals();
        if (HASSIGNAL(c)) p->PropagateControlSignal(c);
        if (!p->GetGatedPorts().empty()) {
            BOOL eof = TESTSIGNAL(c, DBCTLC_EOF);
            if (!eof && b->GetDataBufferLogicalSizeBytes() >= sizeof(CONTROLSIGNAL)) {
                CONTROLSIGNAL v = 0;
                memcpy(&v, b->AcquireHostView(FALSE, TRUE), sizeof(v));
                eof = v != 0;
            }
            if (eof)
                for (Port *g : p->GetGatedPorts())
                    static_cast<OutputPort *>(g)->SignalGate();
        }
    }
    // phase 1b: descriptor ports derive their block from the port they describe
    for (size_t i = 0; i < nIn; ++i) {
        Port *p = m_vInputs[i];
        Port *described = p->GetDescribedPort();
        if (!described) continue;
        Datablock *src = nullptr;
        DESCRIPTORFUNC f = DF_SIZE;
        for (size_t j = 0; j < nIn; ++j)
            if (m_vInputs[j] == described) src = vIn[j];
        for (auto &d : described->GetDescriptorPorts())
            if (d.first == p) f = d.second;
        vIn[i] = MakeDescriptorBlock(src, f, p->GetTemplate());
    }
    // phase 1c: meta ports (iteration counts, allocation sizes, loop control)
    for (size_t i = 0; i < nIn; ++i) {
        Port *p = m_vInputs[i];
        if (p->GetPortType() != META_PORT) continue;
        MetaPort *m = static_cast<MetaPort *>(p);
        UINT alloc = 0;
        BOOL hasAlloc = FALSE;
        m->Configure(&uiIterations, &alloc, &hasAlloc);
        if (hasAlloc && m->GetAllocationPort())
            for (size_t o = 0; o < nOut; ++o)
                if (m_vOutputs[o] == m->GetAllocationPort()) allocBytes[o] = alloc;
    }
    // phase 1d: destination blocks -- the input block for in/out pairs, else a pooled block
    for (size_t o = 0; o < nOut; ++o) {
        OutputPort *op = m_vOutputs[o];
        InputPort *prod = op->GetInOutProducer();
        if (prod) {
            for (size_t i = 0; i < nIn; ++i)
                if (m_vInputs[i] == prod && vIn[i]) {
                    vOut[o] = vIn[i];
                    vOut[o]->AddRef();
                }
        } else {
            vOut[o] = op->GetDestinationBlock(allocBytes[o]);
        }
    }

    // phase 2: admission to the accelerator class shared by all live graphs, then the backend
    // binds and launches, without the graph lock
    pLock->unlock();
    BOOL ok = TRUE;
    if (uiIterations != 0) {
        DispatchArbiter &arb = GetAcceleratorClass() == ACCELERATOR_CLASS_HOST ? RT().hostArbiter : RT().hipArbiter;
        arb.Acquire(RT().schedulingMode == SCHEDMODE_PRIORITY ? GetPriority() : 0);
        ok = PlatformSpecificDispatch(vIn, vOut, uiIterations);
        arb.Release();
    }
    if (!ok) {
        for (size_t o = 0; o < nOut; ++o)
            if (vOut[o]) vOut[o]->SetControlSignal(DBCTLC_ERR);
    }
    for (size_t o = 0; o < nOut; ++o) {
        Datablock *b = vOut[o];
        if (!b) continue;
        if (m_vOutputs[o]->GetInOutProducer()) b->ClearAllControlSignals();
        b->SetControlSignal(m_vOutputs[o]->TakePropagatedControlSignals());
        // EAGER outputs: bring the host view up to date now, off the application's path
        for (UINT c = 0; c < m_vOutputs[o]->GetChannelCount(); ++c) {
            Channel *ch = m_vOutputs[o]->GetChannel(c);
            if (ch->GetType() == CT_GRAPH_OUTPUT &&
                ch->GetViewMaterializationPolicy() == VIEWMATERIALIZATIONPOLICY_EAGER &&
                ch->PassesPredicate(CE_SRC, b)) {
                b->AcquireHostView(FALSE, TRUE);
                break;
            }
        }
    }
    for (size_t i = 0; i < nIn; ++i)
        if (vIn[i]) vIn[i]->Release();
    pLock->lock();

    // phase 3: deliver outputs downstream (each Push consumes the dispatch's reference)
    for (size_t o = 0; o < nOut; ++o)
        if (vOut[o]) m_vOutputs[o]->Push(vOut[o]);
    m_uiDispatchCount++;
    m_bDispatching = FALSE;
    if (replay) replay->OnDispatchDone(this);
    return TRUE;
}

// =============================================================================================
// Graph
// =============================================================================================

Graph::Graph(const char *lpszName)
    : m_strName(lpszName ? lpszName : "graph"), m_bUnnamed(lpszName == nullptr), m_eState(PTGS_INITIALIZING),
      m_bStopRequested(FALSE), m_uiBusyRunners(0), m_bFinalized(FALSE) {
    AddRef();
}

Graph::~Graph() {
    if (m_eState != PTGS_TEARDOWNCOMPLETE) Teardown();
}

BOOL Graph::DestroyGraph(Graph *pGraph, BOOL) {
    if (!pGraph) return FALSE;
    if (pGraph->m_eState != PTGS_TEARDOWNCOMPLETE) pGraph->Teardown();
    pGraph->Release();
    return TRUE;
}

// reference: Graph::AddTask, graph.cpp:829-1075 (task factory switch :1048-1070)
Task *Graph::AddTask(CompiledKernel *pKernel, UINT nIn, Port **ppIn, UINT nOut, Port **ppOut,
                     const char *lpszTaskName) {
    if (!pKernel) {
        Runtime::HandleError("AddTask(%s): null kernel", lpszTaskName ? lpszTaskName : "");
        return nullptr;
    }
    Task *t = nullptr;
    switch (pKernel->GetAcceleratorClass()) {
        case ACCELERATOR_CLASS_HIP: t = new HIPTask(this, pKernel, lpszTaskName); break;
        case ACCELERATOR_CLASS_HOST: t = new HostTask(this, pKernel, lpszTaskName); break;
        default:
            Runtime::HandleError("AddTask(%s): no backend for %s", lpszTaskName,
                                 AccClassString(pKernel->GetAcceleratorClass()));
            return nullptr;
    }
    t->AddRef();
    t->AttachPorts(nIn, ppIn, nOut, ppOut);
    std::lock_guard<std::recursive_mutex> g(m_lock);
    m_vTasks.push_back(t);
    return t;
}

Task *Graph::GetTask(const char *lpszTaskName) {
    for (Task *t : m_vTasks)
        if (!strcmp(t->GetTaskName(), lpszTaskName)) return t;
    return nullptr;
}

void Graph::RegisterChannel(Channel *c) {
    std::lock_guard<std::recursive_mutex> g(m_lock);
    c->AddRef();
    m_vChannels.push_back(c);
    m_mChannels[c->GetName()] = c;
}

GraphInputChannel *Graph::AddInputChannel(Port *pPort, const char *lpszName, BOOL bSwitch, Port *, CONTROLSIGNAL) {
    GraphInputChannel *c = new GraphInputChannel(this, pPort->GetTemplate(), lpszName);
    if (bSwitch) pPort->BindControlChannel(c);
    else pPort->BindChannel(c);
    c->BindPort(pPort, CE_DST);
    RegisterChannel(c);
    return c;
}

InitializerChannel *Graph::AddInitializerChannel(Port *pPort, const char *lpszName, BOOL bSwitch, Port *, CONTROLSIGNAL) {
    InitializerChannel *c = new InitializerChannel(this, pPort->GetTemplate(), lpszName);
    if (bSwitch) pPort->BindControlChannel(c);
    else pPort->BindChannel(c);
    c->BindPort(pPort, CE_DST);
    RegisterChannel(c);
    return c;
}

GraphOutputChannel *Graph::AddOutputChannel(Port *pPort, const char *lpszName, Task *, CONTROLSIGNAL) {
    GraphOutputChannel *c = new GraphOutputChannel(this, pPort->GetTemplate(), lpszName);
    pPort->BindChannel(c);
    c->BindPort(pPort, CE_SRC);
    RegisterChannel(c);
    return c;
}

InternalChannel *Graph::AddInternalChannel(Port *pSrc, Port *pDst, const char *lpszName, BOOL bSwitch) {
    InternalChannel *c = new InternalChannel(this, pSrc->GetTemplate(), lpszName);
    pSrc->BindChannel(c);
    if (bSwitch) pDst->BindControlChannel(c);
    else pDst->BindChannel(c);
    c->BindPort(pSrc, CE_SRC);
    c->BindPort(pDst, CE_DST);
    RegisterChannel(c);
    return c;
}

MultiChannel *Graph::AddMultiChannel(Port **ppDst, int n, const char *lpszName, BOOL bSwitch, Port *, CONTROLSIGNAL) {
    MultiChannel *m = new MultiChannel(this, n > 0 ? ppDst[0]->GetTemplate() : nullptr, lpszName);
    for (int i = 0; i < n; ++i) {
        std::string nm = std::string(lpszName ? lpszName : "multi") + "#" + std::to_string(i);
        GraphInputChannel *c = new GraphInputChannel(this, ppDst[i]->GetTemplate(), nm.c_str());
        if (bSwitch) ppDst[i]->BindControlChannel(c);
        else ppDst[i]->BindChannel(c);
        c->BindPort(ppDst[i], CE_DST);
        c->SetNoDraw();
        RegisterChannel(c);
        m->CoalesceChannel(c);
    }
    RegisterChannel(m);
    return m;
}

Channel *Graph::GetChannel(const char *lpszName) {
    auto it = m_mChannels.find(lpszName ? lpszName : "");
    return it == m_mChannels.end() ? nullptr : it->second;
}

void Graph::BindDescriptorPort(Port *pDescribed, Port *pDescriber, DESCRIPTORFUNC func) {
    pDescribed->AddDescriptorPort(pDescriber, func);
    pDescriber->SetDescribedPort(pDescribed);
}

void Graph::BindControlPort(Port *pController, Port *pGated, BOOL bInitiallyOpen) {
    if (pGated->GetPortType() != OUTPUT_PORT) {
        Runtime::HandleError("BindControlPort: only output ports can be gated");
        return;
    }
    pController->AddGatedPort(pGated);
    static_cast<OutputPort *>(pGated)->SetControlPort(pController, bInitiallyOpen);
}

void Graph::BindControlPropagationPort(Port *pIn, Port *pOut) { pIn->AddControlPropagationPort(pOut); }
void Graph::BindIterationScope(Port *pMeta, Port *pScoped) { pMeta->AddIterationTarget(pScoped); }
void Graph::BindControlPropagationChannel(Port *pIn, Channel *pCh) { pIn->AddControlPropagationChannel(pCh); }

PTRESULT Graph::CheckGraphSemantics(BOOL bVerbose, BOOL bFailOnWarning) {
    PTRESULT r = PTASK_OK;
    for (Task *t : m_vTasks) {
        for (UINT i = 0; i < t->GetInputPortCount(); ++i) {
            Port *p = t->GetInputPort(i);
            if (p->GetPortType() == INITIALIZER_PORT || p->GetDescribedPort()) continue;
            if (!p->HasChannels()) {
                if (bVerbose)
                    Runtime::Warning((std::string(t->GetTaskName()) + ": input " + p->GetVariableBinding() +
                                      " has no channel").c_str());
                r = bFailOnWarning ? PTASK_ERR_GRAPH_MALFORMED : PTASK_WARNING_GRAPH_MALFORMED;
            }
        }
    }
    return r;
}

void Graph::SignalTask(Task *pTask) {
    if (!pTask || pTask->m_bQueued) return;
    pTask->m_bQueued = TRUE;
    m_qReady.push_back(pTask);
    m_cvReady.notify_one();
}

void Graph::SignalPortTask(Port *pPort) {
    if (pPort && pPort->GetTask()) SignalTask(pPort->GetTask());
}

unsigned long long Graph::GetDispatchCount() {
    std::lock_guard<std::recursive_mutex> g(m_lock);
    unsigned long long n = 0;
    for (Task *t : m_vTasks) n += t->GetDispatchCount();
    return n;
}

// reference: Graph::Run, graph.cpp:4037-4070 -- finalize, then start the runners
void Graph::Run(BOOL bSingleThreaded) {
    std::unique_lock<std::recursive_mutex> lk(m_lock);
    if (m_eState == PTGS_RUNNING) return;
    if (!m_bFinalized) {
        for (Task *t : m_vTasks) t->OnGraphComplete();
        CheckGraphSemantics(RT().loggingLevel > 0, FALSE);
        m_bFinalized = TRUE;
    }
    UINT n = 1;
    if (!bSingleThreaded) {
        n = RT().threadPoolSize;
        if (n == 0) {
            UINT hw = std::max(1u, std::thread::hardware_concurrency());
            n = std::min<UINT>(std::max<UINT>(1, (UINT)m_vTasks.size()), std::min<UINT>(hw, 8));
        }
    }
    m_bStopRequested = FALSE;
    m_eState = PTGS_RUNNING;
    for (UINT i = 0; i < n; ++i) m_vRunners.emplace_back(&Graph::RunnerProc, this, i);
    for (Task *t : m_vTasks) SignalTask(t);
    if (RT().useWatchdog && !m_tMonitor.joinable()) {
        m_bMonitorStop = FALSE;
        m_tMonitor = std::thread(&Graph::MonitorProc, this);
    }
}

BOOL Graph::EnableLaunchCapture(BOOL bEnable) {
    std::lock_guard<std::recursive_mutex> g(m_lock);
    if (m_eState == PTGS_RUNNING) return Runtime::HandleError("%s: EnableLaunchCapture while running", GetName());
    if (m_pReplay) {
        delete m_pReplay;
        m_pReplay = nullptr;
    }
    if (bEnable) m_pReplay = new GraphReplay(this);
    return TRUE;
}

BOOL Graph::IsLaunchCaptureArmed() {
    std::lock_guard<std::recursive_mutex> g(m_lock);
    return m_pReplay && m_pReplay->IsArmed();
}

UINT Graph::GetLaunchCaptureReplayCount() {
    std::lock_guard<std::recursive_mutex> g(m_lock);
    return m_pReplay ? m_pReplay->GetReplayCount() : 0;
}

// Work that should lead to a dispatch: a queued task, a running dispatch, or a block waiting in a
// channel or an input port.  Graph lock held.
BOOL Graph::HasPendingWorkLocked() {
    if (!m_qReady.empty() || m_uiBusyRunners) return TRUE;
    for (Channel *c : m_vChannels)
        if (c->GetType() != CT_GRAPH_OUTPUT && c->GetQueueDepth()) return TRUE;
    return FALSE;
}

// reference: Graph::GraphMonitorProc, graph.cpp:2477-2560 -- wake a few times per threshold;
// progress is a change in the graph's completed dispatch count; a running graph that makes none
// for the threshold while work is pending is reported once, until it progresses again.
void Graph::MonitorProc() {
    std::unique_lock<std::recursive_mutex> lk(m_lock);
    unsigned long long last = 0;
    for (Task *t : m_vTasks) last += t->GetDispatchCount();
    auto since = std::chrono::steady_clock::now();
    bool reported = false;
    while (!m_bMonitorStop) {
        const DWORD threshold = RT().watchdogThresholdMs;
        const DWORD period = std::max<DWORD>(10, std::min<DWORD>(threshold / 4, 500));
        m_cvMonitor.wait_for(lk, std::chrono::milliseconds(period));
        if (m_bMonitorStop) break;
        unsigned long long now = 0;
        for (Task *t : m_vTasks) now += t->GetDispatchCount();
        const auto tnow = std::chrono::steady_clock::now();
        if (now != last || m_eState != PTGS_RUNNING) {
            last = now;
            since = tnow;
            reported = false;
            continue;
        }
        const auto idleMs = std::chrono::duration_cast<std::chrono::milliseconds>(tnow - since).count();
        if (reported || idleMs <= (long long)threshold || !HasPendingWorkLocked()) continue;
        reported = true;
        m_uiWatchdogReports++;
        RT().watchdogReports++;
        std::cerr << "graph monitor: graph " << m_strName << " has completed no dispatch for " << idleMs
                  << " ms with work pending; hung?  Graph state:\n";
        ReportGraphState(std::cerr, nullptr, TRUE);
    }
}

void Graph::StopMonitor() {
    std::thread t;
    {
        std::lock_guard<std::recursive_mutex> g(m_lock);
        m_bMonitorStop = TRUE;
        m_cvMonitor.notify_all();
        t.swap(m_tMonitor);
    }
    if (t.joinable()) t.join();
}

void Graph::ReportGraphState(std::ostream &ios, const char *lpszDiagnosticDOTFilePath, BOOL) {
    std::lock_guard<std::recursive_mutex> g(m_lock);
    static const char *kStates[] = {"initializing", "runnable", "running", "quiescing", "tearing down",
                                    "torn down"};
    const int st = (int)m_eState;
    ios << "graph " << m_strName << ": " << (st >= 0 && st < 6 ? kStates[st] : "?") << ", " << m_vTasks.size()
        << " tasks, " << m_qReady.size() << " queued, " << m_uiBusyRunners << " dispatching\n";
    for (Task *t : m_vTasks) {
        ios << "  task " << t->GetTaskName() << ": " << t->GetDispatchCount() << " dispatches"
            << (t->IsDispatching() ? ", dispatching" : "") << "\n";
        for (UINT i = 0; i < t->GetInputPortCount(); ++i) {
            Port *p = t->GetInputPort(i);
            const BOOL ready = p->GetPortType() == INITIALIZER_PORT || static_cast<InputPort *>(p)->IsReady();
            ios << "    in  " << (p->GetVariableBinding() ? p->GetVariableBinding() : "?") << ": "
                << (ready ? "ready" : "waiting") << "\n";
        }
    }
    for (Channel *c : m_vChannels)
        if (c->GetQueueDepth())
            ios << "  channel " << (c->GetName() ? c->GetName() : "?") << ": " << c->GetQueueDepth()
                << " block(s) queued\n";
    ios.flush();
    if (lpszDiagnosticDOTFilePath) WriteDOTFile(lpszDiagnosticDOTFilePath, TRUE);
}

// Runner thread: pop a queued task, dispatch it if ready, re-queue it after a dispatch (it may
// be ready again).  The highest priority queued task goes first in priority mode.
void Graph::RunnerProc(UINT) {
    std::unique_lock<std::recursive_mutex> lk(m_lock);
    for (;;) {
        while (!m_bStopRequested && m_qReady.empty()) m_cvReady.wait(lk);
        if (m_bStopRequested) break;
        auto pick = m_qReady.begin();
        if (RT().schedulingMode == SCHEDMODE_PRIORITY)
            for (auto it = m_qReady.begin(); it != m_qReady.end(); ++it)
                if ((*it)->GetPriority() > (*pick)->GetPriority()) pick = it;
        Task *t = *pick;
        m_qReady.erase(pick);
        t->m_bQueued = FALSE;
        if (t->IsDispatching()) continue;  // its own runner re-queues it when done
        m_uiBusyRunners++;
        BOOL did = t->AttemptDispatch(&lk);
        m_uiBusyRunners--;
        if (did) SignalTask(t);
        if (m_qReady.empty() && m_uiBusyRunners == 0) m_cvIdle.notify_all();
    }
    m_cvIdle.notify_all();
}

BOOL Graph::WaitQuiescent(DWORD dwTimeout) {
    std::unique_lock<std::recursive_mutex> lk(m_lock);
    auto idle = [&] { return (m_qReady.empty() && m_uiBusyRunners == 0) || m_eState != PTGS_RUNNING; };
    if (dwTimeout == INFINITE) {
        m_cvIdle.wait(lk, idle);
        return TRUE;
    }
    return m_cvIdle.wait_for(lk, std::chrono::milliseconds(dwTimeout), idle);
}

void Graph::Stop() {
    StopMonitor();
    std::vector<std::thread> runners;
    {
        std::lock_guard<std::recursive_mutex> g(m_lock);
        if (m_eState != PTGS_RUNNING) return;
        m_bStopRequested = TRUE;
        m_eState = PTGS_QUIESCING;
        m_cvReady.notify_all();
        m_cvData.notify_all();
        m_cvRoom.notify_all();
        runners.swap(m_vRunners);
    }
    for (std::thread &t : runners) t.join();
    std::lock_guard<std::recursive_mutex> g(m_lock);
    for (Task *t : m_vTasks) t->m_bQueued = FALSE;
    m_qReady.clear();
    m_eState = PTGS_RUNNABLE;
}

void Graph::Reset() {
    Stop();
    std::lock_guard<std::recursive_mutex> g(m_lock);
    for (Channel *c : m_vChannels) c->Drain();
    for (Task *t : m_vTasks)
        for (UINT i = 0; i < t->GetInputPortCount(); ++i)
            static_cast<InputPort *>(t->GetInputPort(i))->Drain();
}

// reference: Graph::Teardown -- stop, drain every queue, retire pools and device resources
void Graph::Teardown() {
    Stop();
    std::lock_guard<std::recursive_mutex> g(m_lock);
    if (m_eState == PTGS_TEARDOWNCOMPLETE) return;
    m_eState = PTGS_TEARINGDOWN;
    m_cvData.notify_all();
    m_cvRoom.notify_all();
    for (Channel *c : m_vChannels) c->Drain();
    for (Task *t : m_vTasks) {
        for (UINT i = 0; i < t->GetInputPortCount(); ++i) {
            Port *p = t->GetInputPort(i);
            if (p->GetPortType() != INITIALIZER_PORT) static_cast<InputPort *>(p)->Drain();
        }
        t->ReleaseDeviceResources();
        for (UINT o = 0; o < t->GetOutputPortCount(); ++o) t->GetOutputPort(o)->ReleasePool();
    }
    for (Task *t : m_vTasks) t->Release();
    m_vTasks.clear();
    for (Channel *c : m_vChannels) c->Release();
    m_vChannels.clear();
    m_mChannels.clear();
    ReleaseScopedPools();
    if (m_pReplay) {
        delete m_pReplay;
        m_pReplay = nullptr;
    }
    m_eState = PTGS_TEARDOWNCOMPLETE;
}

void Graph::WriteDOTFile(const char *lpszFileName, BOOL bDrawPorts) {
    std::ofstream f(lpszFileName);
    if (!f) return;
    f << "digraph \"" << m_strName << "\" {\n  rankdir=LR;\n";
    std::map<Task *, int> id;
    int k = 0;
    for (Task *t : m_vTasks) {
        id[t] = k;
        f << "  t" << k++ << " [shape=box,label=\"" << t->GetTaskName() << "\\n"
          << AccClassString(t->GetAcceleratorClass()) << "\"];\n";
    }
    int io = 0;
    for (Channel *c : m_vChannels) {
        if (!c->ShouldDraw()) continue;
        Port *s = c->GetBoundPort(CE_SRC), *d = c->GetBoundPort(CE_DST);
        std::string src, dst;
        if (s && s->GetTask()) src = "t" + std::to_string(id[s->GetTask()]);
        else {
            src = "in" + std::to_string(io++);
            f << "  " << src << " [shape=point];\n";
        }
        if (d && d->GetTask()) dst = "t" + std::to_string(id[d->GetTask()]);
        else {
            dst = "out" + std::to_string(io++);
            f << "  " << dst << " [shape=doublecircle,label=\"\"];\n";
        }
        f << "  " << src << " -> " << dst << " [label=\"" << c->GetName();
        if (bDrawPorts && d) f << "\\n" << d->GetVariableBinding();
        f << "\"];\n";
    }
    f << "}\n";
}

}  // namespace PTask
