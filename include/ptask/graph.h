/* graph.h -- a dataflow graph of tasks, ports and channels (reference: ptask/graph.h).
 *
 * Execution model: Run() starts runner threads that serve a ready queue.  A task is queued
 * whenever something that could make it ready happens (a block arrives on one of its inputs, a
 * downstream channel drains); a runner re-checks readiness under the graph lock and dispatches.
 * One lock per graph guards all channel queues and port state; backends run outside it. */
#ifndef PTK_GRAPH_H
#define PTK_GRAPH_H

#include <atomic>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <ostream>
#include <string>
#include <thread>
#include <vector>

#include "Lockable.h"
#include "ReferenceCounted.h"
#include "channel.h"
#include "primitive_types.h"

namespace PTask {

class CompiledKernel;
class Task;
class BlockPool;
class GraphReplay;
class Port;
class Datablock;

class Graph : public ReferenceCounted {
public:
    Graph(const char *lpszName = NULL);
    virtual ~Graph();
    static BOOL DestroyGraph(Graph *pGraph, BOOL bGCSweep = FALSE);

    virtual void Run(BOOL bSingleThreaded = FALSE);
    virtual void Reset();
    virtual void Stop();
    virtual void Teardown();
    virtual GRAPHSTATE GetState() { return m_eState; }
    BOOL IsRunning() { return m_eState == PTGS_RUNNING; }
    BOOL IsStopped() { return m_eState != PTGS_RUNNING; }
    BOOL IsTorndown() { return m_eState == PTGS_TEARDOWNCOMPLETE; }
    const char *GetName() { return m_strName.c_str(); }

    virtual Task *AddTask(CompiledKernel *pKernel, UINT uiInputPortCount, Port **pvInputPorts,
                          UINT uiOutputPorts, Port **pvOutputPorts, const char *lpszTaskName);
    virtual Task *GetTask(const char *lpszTaskName);
    virtual GraphInputChannel *AddInputChannel(Port *pPort, const char *lpszChannelName = NULL,
                                               BOOL bSwitchChannel = FALSE,
                                               Port *pDefaultTriggerPort = NULL,
                                               CONTROLSIGNAL luiTriggerSignal = DBCTLC_NONE);
    virtual InitializerChannel *AddInitializerChannel(Port *pPort, const char *lpszChannelName = NULL,
                                                      BOOL bSwitchChannel = FALSE,
                                                      Port *pDefaultTriggerPort = NULL,
                                                      CONTROLSIGNAL luiTriggerSignal = DBCTLC_NONE);
    virtual GraphOutputChannel *AddOutputChannel(Port *pPort, const char *lpszChannelName = NULL,
                                                 Task *pTriggerTask = NULL,
                                                 CONTROLSIGNAL luiTriggerSignal = DBCTLC_NONE);
    virtual InternalChannel *AddInternalChannel(Port *pSrc, Port *pDst,
                                                const char *lpszChannelName = NULL,
                                                BOOL bSwitchChannel = FALSE);
    virtual MultiChannel *AddMultiChannel(Port **ppDst, int nDestPorts,
                                          const char *lpszChannelName = NULL,
                                          BOOL bSwitchChannel = FALSE,
                                          Port *pDefaultTriggerPort = NULL,
                                          CONTROLSIGNAL luiTriggerSignal = DBCTLC_NONE);
    virtual Channel *GetChannel(const char *lpszChannelName);
    virtual void BindDescriptorPort(Port *pDescribedPort, Port *pDescriberPort,
                                    DESCRIPTORFUNC func = DF_SIZE);
    virtual void BindControlPort(Port *pController, Port *pGatedPort, BOOL bInitiallyOpen = FALSE);
    virtual void BindControlPropagationPort(Port *pInputPort, Port *pOutputPort);
    virtual void BindIterationScope(Port *pMetaPort, Port *pScopedPort);
    virtual void BindControlPropagationChannel(Port *pInputPort, Channel *pControlledChannel);
    /* Graphviz rendering of tasks and channels. */
    virtual void WriteDOTFile(const char *lpszFileName, BOOL bDrawPorts = FALSE);
    /* Structural checks: every formal parameter bound, every input port fed, indices dense. */
    virtual PTRESULT CheckGraphSemantics(BOOL bVerbose = TRUE, BOOL bFailOnWarning = FALSE);
    UINT GetTaskCount() { return (UINT)m_vTasks.size(); }
    Task *GetTaskByIndex(UINT i) { return m_vTasks[i]; }
    /* Every channel in creation order, the members of a multi-channel included. */
    UINT GetChannelCount() { return (UINT)m_vChannels.size(); }
    Channel *GetChannelByIndex(UINT i) { return m_vChannels[i]; }
    /* Writes the graph's structure (templates, kernels, tasks and ports, channels, bindings,
     * predication, geometry) as XML (reference: Graph::Serialize, graph.cpp:6452-6468,
     * XMLWriter::WriteGraph XMLWriter.cpp:124-245).  Fails (HandleError, FALSE) on what has no
     * text form: a geometry estimator or a user-defined predicate function, duplicate port UIDs
     * or channel names. */
    virtual BOOL Serialize(const char *lpszFileName);
    /* Builds the graph written by Serialize into this empty graph: templates are registered with
     * the runtime (an existing template of the same name must match), kernels are loaded with
     * Runtime::GetCompiledKernel, then tasks, channels and bindings are recreated (reference:
     * Graph::Deserialize, graph.cpp:6477-6490, XMLReader::ReadGraph XMLReader.cpp:383-404).
     * A kernel file recorded for CUDA (.ptx, .cu) is looked for as the HIP code object of the
     * same path and stem (.hsaco), and CUDA tasks become HIP tasks.  Returns FALSE (after
     * HandleError) on a malformed file or a kernel that cannot be loaded. */
    virtual BOOL Deserialize(const char *lpszFileName);
    /* Total dispatches since Run(), across all tasks. */
    unsigned long long GetDispatchCount();
    /* Blocks until no task is ready or dispatching and every runner is idle (or the timeout in
     * ms elapses); returns TRUE when quiescent. */
    BOOL WaitQuiescent(DWORD dwTimeout = INFINITE);

    /* ---- runtime internals ---------------------------------------------------------------- */
    std::recursive_mutex &GraphLock() { return m_lock; }
    std::condition_variable_any &DataArrived() { return m_cvData; }
    std::condition_variable_any &RoomAvailable() { return m_cvRoom; }
    /* Queue a task for a readiness check (graph lock held). */
    void SignalTask(Task *pTask);
    void SignalPortTask(Port *pPort);
    /* Wake waiters on output channels / input-channel capacity (graph lock held). */
    void NotifyDataArrived() { m_cvData.notify_all(); }
    void NotifyRoom() { m_cvRoom.notify_all(); }
    BOOL IsAlive() { return m_eState != PTGS_TEARINGDOWN && m_eState != PTGS_TEARDOWNCOMPLETE; }

    /* ---- diagnostics (reference Graph::ReportGraphState, graph.h:1476; the monitor of
     * GraphMonitorProc, graph.cpp:2477-2560) ----------------------------------------------------
     * A textual state dump: every task with its dispatch count, whether it is dispatching, and the
     * readiness of each input port; every channel holding blocks.  With a path, the DOT rendering
     * is written there too.  While Runtime::SetUseGraphMonitorWatchdog(TRUE) was in force at Run,
     * a monitor thread dumps the state to stderr when no dispatch has completed for
     * Runtime::GetDispatchWatchdogThreshold() ms although work is pending (once per stall). */
    void ReportGraphState(std::ostream &ios, const char *lpszDiagnosticDOTFilePath, BOOL bForceGraphDisplay);
    UINT GetWatchdogReportCount() { return m_uiWatchdogReports; }

    /* ---- launch capture (no reference counterpart) ---------------------------------------------
     * EnableLaunchCapture(TRUE) before Run: the next run is recorded and, if every task is a HIP
     * task that dispatches exactly once per run (no meta ports, no predicated channels), every
     * later run replays it as one hipGraph launch: inputs are copied into staging buffers, the
     * outputs are fresh pooled blocks, and each task's dispatch count still advances by one per
     * run.  New values on a sticky port, or a recording that does not fit, return the graph to
     * ordinary dispatch.  IsLaunchCaptureArmed: replays are active. */
    BOOL EnableLaunchCapture(BOOL bEnable);
    BOOL IsLaunchCaptureArmed();
    UINT GetLaunchCaptureReplayCount();
    class GraphReplay *GetReplay() { return m_pReplay; }

    /* ---- graph-scoped block pools (reference graph.h:139-199, ScopedPoolManager) ------------
     * Pre-allocated blocks (host storage, and device storage when a HIP accelerator exists) that
     * RequestPooledBlock hands out and that return to the pool on their last Release; released
     * at Teardown.  nBlocks 0: the runtime's default block pool size. */
    BOOL RequireBlockPool(DatablockTemplate *pTemplate, int nDataSize, int nMetaSize, int nTemplateSize,
                          int nBlocks = 0);
    BOOL RequireBlockPool(int nDataSize, int nMetaSize, int nTemplateSize, int nBlocks = 0);
    BOOL RequireBlockPool(DatablockTemplate *pTemplate, int nBlocks = 0);
    /* A free block of the template's pool (else of the smallest untyped pool holding uiDataSize
     * bytes), with one reference; NULL when none is free. */
    Datablock *RequestPooledBlock(DatablockTemplate *pTemplate, UINT uiDataSize, UINT uiMetaSize,
                                  UINT uiTemplateSize);

private:
    void RunnerProc(UINT uiIndex);
    void MonitorProc();
    void StopMonitor();
    BOOL HasPendingWorkLocked();
    void ReleaseScopedPools();
    std::thread m_tMonitor;
    std::condition_variable_any m_cvMonitor;
    BOOL m_bMonitorStop = FALSE;
    std::atomic<UINT> m_uiWatchdogReports{0};
    std::map<DatablockTemplate *, BlockPool *> m_mTypedPools;
    std::map<UINT, BlockPool *> m_mUntypedPools;
    class GraphReplay *m_pReplay = nullptr;
    void RegisterChannel(Channel *pChannel);
    std::string SerializeLocked();

    std::string m_strName;
    bool m_bUnnamed;  // constructed without a name: Deserialize takes the file's
    GRAPHSTATE m_eState;
    std::recursive_mutex m_lock;
    std::condition_variable_any m_cvReady;
    std::condition_variable_any m_cvData;
    std::condition_variable_any m_cvRoom;
    std::condition_variable_any m_cvIdle;
    std::deque<Task *> m_qReady;
    std::vector<Task *> m_vTasks;
    std::vector<Channel *> m_vChannels;
    std::map<std::string, Channel *> m_mChannels;
    std::vector<std::thread> m_vRunners;
    BOOL m_bStopRequested;
    UINT m_uiBusyRunners;
    BOOL m_bFinalized;
};

}  // namespace PTask

#endif
