/* task.h -- a graph node bound to one compiled kernel (reference: ptask/task.h).
 *
 * Dispatch (reference: Task::Dispatch, ptask/task.cpp:3822-3906) runs in three phases: with
 * the graph lock held the task takes one block from every input, runs its meta ports and picks
 * destination blocks; the lock is dropped while the backend binds and launches; the lock is
 * retaken to push outputs downstream and wake the tasks that may have become ready. */
#ifndef PTK_TASK_H
#define PTK_TASK_H

#include <mutex>
#include <string>
#include <vector>

#include "Lockable.h"
#include "ReferenceCounted.h"
#include "primitive_types.h"

namespace PTask {

class Accelerator;
class AsyncContext;
class CompiledKernel;
class Datablock;
class Graph;
class Port;
class OutputPort;

/* Per-dispatch geometry callback (reference: Task::SetGeometryEstimator). */
typedef void(__stdcall *LPFNGEOMETRYESTIMATOR)(UINT nArguments, Datablock **ppArguments,
                                               PTASKDIM3 *pBlockSize, PTASKDIM3 *pGridSize);

class Task : public ReferenceCounted, public Lockable {
public:
    Task(Graph *pGraph, CompiledKernel *pKernel, ACCELERATOR_CLASS eClass, const char *lpszName);
    virtual ~Task();

    virtual const char *GetTaskName() { return m_strName.c_str(); }
    /* Host tasks: the function runs tgx*tgy*tgz times per dispatch.  HIP tasks: the number of
     * work items to cover when no explicit grid is given. */
    virtual void SetComputeGeometry(int tgx = 1, int tgy = 1, int tgz = 1);
    /* Explicit launch geometry, honoured exactly (reference: cutask.cpp:633-697). */
    virtual void SetBlockAndGridSize(PTASKDIM3 grid, PTASKDIM3 block);
    virtual void SetGeometryEstimator(LPFNGEOMETRYESTIMATOR lpfn) { m_lpfnEstimator = lpfn; }
    LPFNGEOMETRYESTIMATOR GetGeometryEstimator() { return m_lpfnEstimator; }
    void GetComputeGeometry(int *tgx, int *tgy, int *tgz) { *tgx = m_tgx; *tgy = m_tgy; *tgz = m_tgz; }
    /* TRUE once SetBlockAndGridSize was called; the grid and block then come from it. */
    BOOL GetBlockAndGridSize(PTASKDIM3 *grid, PTASKDIM3 *block) {
        *grid = m_grid;
        *block = m_block;
        return m_bExplicitGeometry;
    }
    virtual void SetPriority(int nPTaskPrio, int nProxyOSPrio = 0);
    int GetPriority() { return m_nPriority; }
    BOOL SetAffinity(Accelerator *pAccelerator, AFFINITYTYPE eAffinityType);
    Accelerator *GetMandatoryAccelerator() { return m_pMandatoryAccelerator; }
    void SetUserCodeAllocatesMemory(BOOL b) { m_bUserCodeAllocates = b; }
    BOOL GetUserCodeAllocatesMemory() { return m_bUserCodeAllocates; }
    ACCELERATOR_CLASS GetAcceleratorClass() { return m_eClass; }
    CompiledKernel *GetCompiledKernel() { return m_pKernel; }
    Graph *GetGraph() { return m_pGraph; }
    UINT GetInputPortCount() { return (UINT)m_vInputs.size(); }
    UINT GetOutputPortCount() { return (UINT)m_vOutputs.size(); }
    Port *GetInputPort(UINT i) { return m_vInputs[i]; }
    OutputPort *GetOutputPort(UINT i) { return m_vOutputs[i]; }
    UINT GetDispatchCount() { return m_uiDispatchCount; }
    /* A dispatch that a replay of the graph's captured launches performed for this task. */
    void NoteReplayedDispatch() { m_uiDispatchCount++; }
    /* Dependent accelerators (reference: Task::BindDependentAcceleratorClass, task.h:768,
     * task.cpp): a host task that also needs nRequiredInstances (1) accelerators of class cls to
     * run -- its ports bound with Port::BindDependentAccelerator receive device pointers, and with
     * bRequestDependentPSObjects the task is called in the LPFNDEPHOSTTASKEX form with the
     * dependent stream (hosttask.h).  ACCELERATOR_CLASS_CUDA requests bind to HIP on this build.
     * FALSE (and an error) for a second binding of a class or a count other than 1. */
    virtual BOOL BindDependentAcceleratorClass(ACCELERATOR_CLASS cls, int nRequiredInstances = 1,
                                               BOOL bRequestDependentPSObjects = FALSE);
    virtual int GetDependentBindingClassCount() { return (int)m_vDependentClasses.size(); }
    virtual BOOL HasDependentAcceleratorBindings() { return !m_vDependentClasses.empty(); }
    ACCELERATOR_CLASS GetDependentAcceleratorClass(int i) { return m_vDependentClasses[(size_t)i]; }
    BOOL GetRequestDependentPSObjects() { return m_bRequestDependentPSObjects; }
    /* Device time of the last dispatch in ms (filled when Runtime::SetProfilePSDispatch(TRUE),
     * from events around the launches, as ptask/cutask.cpp:709-776). */
    double GetLastDispatchDeviceMs() { return m_dLastDeviceMs; }
    /* Waits for the profiled dispatches still in flight. */
    virtual double GetTotalDispatchDeviceMs() { return m_dTotalDeviceMs; }

    /* ---- runtime internals ---------------------------------------------------------------- */
    void AttachPorts(UINT nIn, Port **ppIn, UINT nOut, Port **ppOut);
    /* Graph lock held. */
    BOOL IsReadyForDispatch();
    /* Runs one dispatch if ready; called by a graph runner with the graph lock held (the lock
     * is released during the backend phase).  Returns TRUE if a dispatch happened. */
    BOOL AttemptDispatch(std::unique_lock<std::recursive_mutex> *pGraphLock);
    BOOL IsDispatching() { return m_bDispatching; }
    BOOL m_bQueued;
    void ComputeParameterLayout();
    virtual void OnGraphComplete();
    virtual void ReleaseDeviceResources() {}

protected:
    struct Binding {
        Port *pPort;          // the port that supplies the argument
        BOOL bInput;          // input side (input, sticky, initializer) or output side
        UINT uiSlot;          // index into the dispatch's input or output block array
        PTASK_PARM_TYPE eType;
        UINT uiOffset;        // byte offset in the packed argument buffer
        UINT uiBytes;
        BOOL bWritable;       // an in/out input, or an output
    };
    /* Backend phase: bind the blocks and run the kernel `uiIterations` times. */
    virtual BOOL PlatformSpecificDispatch(std::vector<Datablock *> &vIn,
                                          std::vector<Datablock *> &vOut, UINT uiIterations) = 0;
    void GeometryFor(std::vector<Datablock *> &vIn, PTASKDIM3 *pGrid, PTASKDIM3 *pBlock);

    Graph *m_pGraph;
    CompiledKernel *m_pKernel;
    ACCELERATOR_CLASS m_eClass;
    std::string m_strName;
    std::vector<Port *> m_vInputs;
    std::vector<OutputPort *> m_vOutputs;
    std::vector<Binding> m_vBindings;
    UINT m_uiArgBytes;
    BOOL m_bLayoutDone;
    int m_tgx, m_tgy, m_tgz;
    BOOL m_bExplicitGeometry;
    PTASKDIM3 m_grid, m_block;
    LPFNGEOMETRYESTIMATOR m_lpfnEstimator;
    int m_nPriority;
    Accelerator *m_pMandatoryAccelerator;
    BOOL m_bUserCodeAllocates;
    BOOL m_bDispatching;
    UINT m_uiDispatchCount;
    double m_dLastDeviceMs;
    double m_dTotalDeviceMs;
    std::vector<ACCELERATOR_CLASS> m_vDependentClasses;
    BOOL m_bRequestDependentPSObjects = FALSE;
};

}  // namespace PTask

#endif
