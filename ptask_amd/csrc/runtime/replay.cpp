// replay.cpp -- launch capture of a graph: one recorded run replayed as one hipGraph launch per
// later run (no counterpart in the reference, whose dispatch path launches every task on its own,
// task.cpp:3822-3906 and cutask.cpp:633-697).
//
// Graph::EnableLaunchCapture(TRUE) before Run arms recording.  The next run goes through the
// ordinary dispatch path; every HIP launch it makes (kernel, geometry, packed arguments and the
// datablock behind each pointer argument) is recorded in host dispatch order, which is a valid
// order of the run: a task dispatches only once its inputs exist, and later writers of a block
// were dispatched after its earlier readers.  When every task has dispatched exactly once the
// recording is sealed:
//   - blocks the application pushed into graph input channels get private device staging
//     buffers, and so do blocks delivered to graph output channels; every other block the run
//     touched stays pinned (the recording holds a reference) and is reused by every replay;
//   - the launches, their pointer arguments re-targeted to the staging buffers, are captured in
//     order on a private stream and instantiated as one hipGraph.
// From then on, when every data input channel holds a block, a replay (graph lock held):
// copies each input block into its staging buffer on the replay stream, launches the hipGraph,
// takes a fresh block from each graph output's port pool, copies the staging output into it and
// enqueues it on the output channel with a sync point, and counts one dispatch per task.  No
// task dispatches individually while the capture is armed.
//
// Supported: graphs whose tasks are all HIP tasks on one accelerator, with no meta ports and no
// predicated channels, and a run in which each task dispatches once (an unrolled graph).  A
// recording that does not fit, or new values on a sticky (parameter) channel later, turn capture
// off for good and the graph continues with ordinary dispatch.
#include <algorithm>
#include <cstring>
#include <map>
#include <set>

#include "ptk_internal.h"

namespace PTask {

GraphReplay::GraphReplay(Graph *pGraph) : m_pGraph(pGraph) {}

GraphReplay::~GraphReplay() { Release(); }

void GraphReplay::Release() {
    std::lock_guard<std::mutex> g(m_recLock);
    if (m_pAccelerator) m_pAccelerator->MakeDeviceContextCurrent();
    if (m_pContext) m_pContext->SynchronizeContext();
    if (m_hGraph) ptk_hip_graph_destroy(m_hGraph);
    m_hGraph = nullptr;
    for (auto &kv : m_mStaging)
        if (kv.second.dptr) ptk_hip_free(kv.second.dptr);
    m_mStaging.clear();
    for (Datablock *b : m_vPinned) b->Release();
    m_vPinned.clear();
    for (Launch &l : m_vLaunches)
        for (auto &pa : l.ptrs) pa.pBlock->Release();
    m_vLaunches.clear();
    if (!m_bKeysOnly) {
        for (auto &kv : m_mInputs) kv.second->Release();
        for (auto &kv : m_mOutputs) kv.second->Release();
    }
    m_mInputs.clear();
    m_mOutputs.clear();
    if (m_pContext) m_pContext->Release();
    m_pContext = nullptr;
}

void GraphReplay::Fail(const char *why) {
    if (m_eState == FAILED) return;
    m_eState = FAILED;
    Runtime::Warning((std::string("graph ") + m_pGraph->GetName() + ": launch capture off: " + why).c_str());
}

// ---- recording --------------------------------------------------------------------------------

void GraphReplay::OnInputPush(Channel *pChannel, Datablock *pBlock) {
    Port *dst = pChannel->GetBoundPort(CE_DST);
    const bool param = dst && dst->GetPortType() == STICKY_PORT;
    if (m_eState == ARMED) {
        // new parameters would change the baked by-value arguments
        if (param) Fail("a sticky port received new parameters");
        return;
    }
    if (m_eState != RECORDING || param) return;
    if (!dst || dst->GetPortType() != INPUT_PORT) return Fail("a graph input feeds a port that is not an input port");
    if (m_mInputs.count(pChannel)) return Fail("a graph input received two blocks in the recorded run");
    pBlock->AddRef();
    m_mInputs[pChannel] = pBlock;
}

void GraphReplay::OnOutputEnqueue(Channel *pChannel, Datablock *pBlock) {
    if (m_eState != RECORDING) return;
    if (m_mOutputs.count(pChannel)) return Fail("a graph output received two blocks in the recorded run");
    pBlock->AddRef();
    m_mOutputs[pChannel] = pBlock;
}

void GraphReplay::OnLaunch(Task *pTask, Accelerator *pAcc, ptk_hip_function fn, const PTASKDIM3 &grid,
                           const PTASKDIM3 &block, const unsigned char *args, UINT cbArgs,
                           const std::vector<PtrArg> &ptrs, UINT uiIterations) {
    std::lock_guard<std::mutex> g(m_recLock);
    if (m_eState != RECORDING) return;
    if (m_pAccelerator && m_pAccelerator != pAcc) return Fail("tasks run on more than one accelerator");
    m_pAccelerator = pAcc;
    Launch l;
    l.pTask = pTask;
    l.fn = fn;
    l.grid = grid;
    l.block = block;
    l.args.assign(args, args + cbArgs);
    l.ptrs = ptrs;
    for (auto &pa : l.ptrs) pa.pBlock->AddRef();
    l.uiIterations = uiIterations;
    m_vLaunches.push_back(std::move(l));
}

// Graph lock held.  Seals the recording once every task has dispatched exactly once.
void GraphReplay::OnDispatchDone(Task *pTask) {
    if (m_eState != RECORDING) return;
    if (pTask->GetAcceleratorClass() != ACCELERATOR_CLASS_HIP) return Fail("a task is not a HIP task");
    if (!m_sDispatched.insert(pTask).second) return Fail("a task dispatched twice in the recorded run");
    if (m_sDispatched.size() == m_pGraph->GetTaskCount()) Seal();
}

void GraphReplay::Seal() {
    std::lock_guard<std::mutex> g(m_recLock);
    if (m_vLaunches.empty() || !m_pAccelerator) return Fail("the recorded run launched nothing");
    for (UINT i = 0; i < m_pGraph->GetChannelCount(); ++i) {
        Channel *c = m_pGraph->GetChannelByIndex(i);
        if (c->GetType() == CT_INITIALIZER || c->GetType() == CT_MULTI) return Fail("initializer or multi channels");
        if (c->GetPredicationType(CE_SRC) != CGATEFN_NONE || c->GetPredicationType(CE_DST) != CGATEFN_NONE)
            return Fail("predicated channels");
        Port *dst = c->GetBoundPort(CE_DST);
        if (c->GetType() == CT_GRAPH_INPUT && dst && dst->GetPortType() == INPUT_PORT && !m_mInputs.count(c))
            return Fail("a data input received nothing in the recorded run");
        if (c->GetType() == CT_GRAPH_OUTPUT && !m_mOutputs.count(c)) return Fail("an output received nothing");
    }
    for (UINT t = 0; t < m_pGraph->GetTaskCount(); ++t) {
        Task *task = m_pGraph->GetTaskByIndex(t);
        for (UINT i = 0; i < task->GetInputPortCount(); ++i)
            if (task->GetInputPort(i)->GetPortType() == META_PORT) return Fail("meta ports");
    }
    m_pAccelerator->MakeDeviceContextCurrent();
    m_pAccelerator->Synchronize();  // the recorded run is complete before its blocks are reused
    m_pContext = static_cast<HIPAsyncContext *>(m_pAccelerator->CreateAsyncContext(ASYNCCTXT_TASK));
    if (!m_pContext) return Fail("no stream");
    // staging buffers: one per distinct input block and per distinct output block
    auto stage = [&](Datablock *b, bool in) -> bool {
        Staging &s = m_mStaging[b];
        if (s.dptr) {
            s.in = s.in || in;
            s.out = s.out || !in;
            return true;
        }
        s.bytes = std::max<size_t>(b->GetDataBufferAllocatedSizeBytes(), 4);
        s.logical = b->GetDataBufferLogicalSizeBytes();
        s.in = in;
        s.out = !in;
        return ptk_hip_malloc(&s.dptr, s.bytes) == PTK_HIP_OK;
    };
    for (auto &kv : m_mInputs)
        if (!stage(kv.second, true)) return Fail("staging allocation failed");
    for (auto &kv : m_mOutputs)
        if (!stage(kv.second, false)) return Fail("staging allocation failed");
    // pin every other block the run touched
    std::set<Datablock *> pinned;
    for (Launch &l : m_vLaunches)
        for (auto &pa : l.ptrs)
            if (!m_mStaging.count(pa.pBlock) && pinned.insert(pa.pBlock).second) {
                pa.pBlock->AddRef();
                m_vPinned.push_back(pa.pBlock);
            }
    if (!CaptureLaunches()) return Fail("stream capture failed");
    m_uiLaunchesPerReplay = 0;
    for (Launch &l : m_vLaunches) m_uiLaunchesPerReplay += l.uiIterations;
    // the recorded input and output blocks go back to the application and the pools; their
    // addresses stay as the keys of the staging buffers, never dereferenced again
    for (Launch &l : m_vLaunches)
        for (auto &pa : l.ptrs) pa.pBlock->Release();
    m_vLaunches.clear();
    for (auto &kv : m_mInputs) kv.second->Release();
    for (auto &kv : m_mOutputs) kv.second->Release();
    m_bKeysOnly = true;
    m_eState = ARMED;
}

// The re-targeted launches in recorded order, captured on the replay stream.  One stream: a
// capture forked over several streams (events recorded on the capturing stream, waited on by the
// others) ended the process with a segmentation fault inside the HIP runtime on the first run,
// so the captured graph is a chain and independent launches (the three H or E updates of a Yee
// step) run one after another.
bool GraphReplay::CaptureLaunches() {
    ptk_hip_stream s = m_pContext->GetStream();
    if (ptk_hip_stream_begin_capture(s) != PTK_HIP_OK) return false;
    bool ok = true;
    for (Launch &l : m_vLaunches) {
        std::vector<unsigned char> args = l.args;
        for (auto &pa : l.ptrs) {
            void *p = nullptr;
            auto it = m_mStaging.find(pa.pBlock);
            if (it != m_mStaging.end()) p = it->second.dptr;
            else p = pa.pBlock->GetPlatformBuffer(m_pAccelerator->GetMemorySpaceId())->GetBuffer();
            memcpy(args.data() + pa.uiOffset, &p, sizeof(void *));
        }
        for (UINT it = 0; it < l.uiIterations && ok; ++it)
            ok = ptk_hip_launch(l.fn, (unsigned)l.grid.x, (unsigned)l.grid.y, (unsigned)l.grid.z, (unsigned)l.block.x,
                                (unsigned)l.block.y, (unsigned)l.block.z, 0, s, args.data(), args.size()) == PTK_HIP_OK;
    }
    ptk_hip_graph graph = nullptr;
    const int rc = ptk_hip_stream_end_capture(s, &graph);
    if (!ok || rc != PTK_HIP_OK || !graph) {
        if (graph) ptk_hip_graph_destroy(graph);
        return false;
    }
    m_hGraph = graph;
    return true;
}

// ---- replay -----------------------------------------------------------------------------------

// Graph lock held.  TRUE when a replay ran.
BOOL GraphReplay::TryReplay() {
    if (m_eState != ARMED) return FALSE;
    // every data input holds a block, and inputs recorded as one block hold one block again
    std::map<Datablock *, Datablock *> slot;  // recorded block -> this run's block
    for (auto &kv : m_mInputs) {
        Datablock *b = kv.first->Peek();
        if (!b) return FALSE;
        auto it = slot.find(kv.second);
        if (it != slot.end() && it->second != b) {
            Fail("inputs recorded as one block received different blocks");
            return FALSE;
        }
        slot[kv.second] = b;
    }
    // take the blocks
    std::vector<Datablock *> taken;
    for (auto &kv : m_mInputs) taken.push_back(kv.first->Pull(0));
    m_pAccelerator->MakeDeviceContextCurrent();
    const UINT ms = m_pAccelerator->GetMemorySpaceId();
    ptk_hip_stream s = m_pContext->GetStream();
    bool ok = true;
    // copy in
    for (auto &kv : slot) {
        Staging &st = m_mStaging[kv.first];
        Datablock *b = kv.second;
        void *src = b->AcquireDeviceView(m_pAccelerator, m_pContext, FALSE, TRUE);
        PBuffer *pb = b->GetPlatformBuffer(ms);
        if (!src || !pb) {
            ok = false;
            break;
        }
        pb->WaitOutstanding(m_pContext, FALSE);
        const size_t cb = std::min<size_t>(b->GetDataBufferLogicalSizeBytes(), st.bytes);
        if (cb && ptk_hip_memcpy_dtod_async(st.dptr, src, cb, s) != PTK_HIP_OK) ok = false;
        SyncPoint *sp = m_pContext->CreateSyncPoint();
        if (sp) {
            pb->AddOutstanding(sp, FALSE);
            sp->Release();
        }
    }
    if (ok) ok = ptk_hip_graph_launch(m_hGraph, s) == PTK_HIP_OK;
    // copy out into fresh blocks of each output port's pool
    std::vector<std::pair<Channel *, Datablock *>> delivered;
    for (auto &kv : m_mOutputs) {
        if (!ok) break;
        Channel *c = kv.first;
        Staging &st = m_mStaging[kv.second];
        OutputPort *op = static_cast<OutputPort *>(c->GetBoundPort(CE_SRC));
        Datablock *nb = op ? op->GetDestinationBlock((UINT)st.logical) : nullptr;
        void *dst = nb ? nb->AcquireDeviceView(m_pAccelerator, m_pContext, TRUE, FALSE) : nullptr;
        PBuffer *pb = nb ? nb->GetPlatformBuffer(ms) : nullptr;
        if (!dst || !pb) {
            if (nb) nb->Release();
            ok = false;
            break;
        }
        pb->WaitOutstanding(m_pContext, TRUE);
        if (st.logical && ptk_hip_memcpy_dtod_async(dst, st.dptr, st.logical, s) != PTK_HIP_OK) ok = false;
        SyncPoint *sp = m_pContext->CreateSyncPoint();
        if (sp) {
            pb->AddOutstanding(sp, TRUE);
            sp->Release();
        }
        delivered.push_back(std::make_pair(c, nb));
    }
    for (Datablock *b : taken)
        if (b) b->Release();
    if (!ok) {
        for (auto &d : delivered) {
            d.second->SetControlSignal(DBCTLC_ERR);
            d.first->Enqueue(d.second);
            d.second->Release();
        }
        Runtime::HandleError("graph %s: replay failed: %s", m_pGraph->GetName(), ptk_hip_last_error_string());
        Fail("a replay failed");
        return TRUE;
    }
    for (auto &d : delivered) {
        d.first->Enqueue(d.second);  // the channel takes its own reference
        d.second->Release();
    }
    for (UINT t = 0; t < m_pGraph->GetTaskCount(); ++t) m_pGraph->GetTaskByIndex(t)->NoteReplayedDispatch();
    m_uiReplays++;
    return TRUE;
}

}  // namespace PTask
