/* port.h -- task ports: where channels attach to a task's formal parameters
 * (reference: ptask/port.h, InputPort.h, OutputPort.h, StickyPort.h, MetaPort.h,
 * InitializerPort.h; created by Runtime::CreatePort, ptask/PTaskRuntime.h:231-237).
 *
 * Port state is only touched with the owning graph's lock held (or before the graph runs). */
#ifndef PTK_PORT_H
#define PTK_PORT_H

#include <deque>
#include <string>
#include <utility>
#include <vector>

#include "Lockable.h"
#include "ReferenceCounted.h"
#include "primitive_types.h"

namespace PTask {

class Channel;
class Datablock;
class DatablockTemplate;
class Task;
class Graph;
class OutputPort;
class BlockPool;

class Port : public ReferenceCounted, public Lockable {
public:
    Port(PORTTYPE eType, DatablockTemplate *pTemplate, UINT uiId, const char *lpszBinding,
         UINT uiFormalIndex, UINT uiInOutIndex);
    virtual ~Port();

    PORTTYPE GetPortType() { return m_eType; }
    DatablockTemplate *GetTemplate() { return m_pTemplate; }
    UINT GetUID() { return m_uiId; }
    char *GetVariableBinding() { return const_cast<char *>(m_strBinding.c_str()); }
    UINT GetFormalParameterIndex() { return m_uiFormalIndex; }
    /* TRUE when the port supplies a kernel argument: it has a formal index, is not a meta port
     * and is not the output half of an in/out pair (which shares its input's argument). */
    virtual BOOL IsFormalParameter();
    UINT GetInOutRoutingIndex() { return m_uiInOutIndex; }
    BOOL IsInOutParameter() { return m_uiInOutIndex != (UINT)-1; }
    Task *GetTask() { return m_pTask; }
    virtual BOOL IsSticky() { return FALSE; }
    void SetMarshallable(BOOL b) { m_bMarshallable = b; }
    BOOL IsMarshallable() { return m_bMarshallable; }
    void SetMetaFunction(METAFUNCTION e) { m_eMetaFunction = e; }
    METAFUNCTION GetMetaFunction() { return m_eMetaFunction; }
    /* Materialise this port's block in the memory of the task's dependent accelerator of class
     * accClass (instance nIndex) and pass a device pointer to a dependent host task (reference:
     * Port::BindDependentAccelerator, port.h:1185).  ACCELERATOR_CLASS_CUDA requests bind to the
     * HIP accelerator on this build. */
    virtual void BindDependentAccelerator(ACCELERATOR_CLASS accClass, int nIndex);
    virtual BOOL HasDependentAcceleratorBinding() { return m_bDependentBinding; }
    ACCELERATOR_CLASS GetDependentAcceleratorClass() { return m_eDependentClass; }
    int GetDependentAcceleratorIndex() { return m_nDependentIndex; }
    /* A sticky or in/out block that survives across dispatches until replaced. */
    void SetPermanentBlock(BOOL b) { m_bPermanentBlock = b; }
    BOOL IsPermanentBlock() { return m_bPermanentBlock; }
    /* An output port whose block goes to a reader and a writer (an in/out consumer or a graph
     * output channel), or to two writers, pushes copies to the writers; TRUE here pushes the one
     * block to all of them, for graphs whose structure already orders the writes after the reads
     * (reference: Port::SetSuppressClones, port.h:1274, OutputPort.cpp:1760). */
    virtual void SetSuppressClones(BOOL b) { m_bSuppressClones = b; }
    virtual BOOL GetSuppressClones() { return m_bSuppressClones; }

    virtual void BindChannel(Channel *pChannel);
    virtual void BindControlChannel(Channel *pChannel);
    UINT GetChannelCount() { return (UINT)m_vChannels.size(); }
    Channel *GetChannel(UINT i) { return m_vChannels[i]; }
    UINT GetControlChannelCount() { return (UINT)m_vControlChannels.size(); }
    Channel *GetControlChannel(UINT i) { return m_vControlChannels[i]; }
    BOOL HasChannels() { return !m_vChannels.empty() || !m_vControlChannels.empty(); }

    /* ---- runtime internals ---------------------------------------------------------------- */
    void BindTask(Task *pTask, UINT uiIndex) { m_pTask = pTask; m_uiTaskIndex = uiIndex; }
    UINT GetTaskIndex() { return m_uiTaskIndex; }
    void AddControlPropagationPort(Port *pOutput) { m_vControlPropagationPorts.push_back(pOutput); }
    std::vector<Port *> &GetControlPropagationPorts() { return m_vControlPropagationPorts; }
    void AddControlPropagationChannel(Channel *p) { m_vControlPropagationChannels.push_back(p); }
    std::vector<Channel *> &GetControlPropagationChannels() { return m_vControlPropagationChannels; }
    /* Pushes a control code to every output port / channel this port propagates to. */
    void PropagateControlSignal(CONTROLSIGNAL luiCode);
    void AddGatedPort(Port *pGated) { m_vGatedPorts.push_back(pGated); }
    std::vector<Port *> &GetGatedPorts() { return m_vGatedPorts; }
    void AddDescriptorPort(Port *pDescriber, DESCRIPTORFUNC eFunc) {
        m_vDescriptorPorts.push_back(std::make_pair(pDescriber, eFunc));
    }
    std::vector<std::pair<Port *, DESCRIPTORFUNC>> &GetDescriptorPorts() { return m_vDescriptorPorts; }
    void SetDescribedPort(Port *p) { m_pDescribedPort = p; }
    Port *GetDescribedPort() { return m_pDescribedPort; }
    void AddIterationTarget(Port *p) { m_vIterationTargets.push_back(p); }
    std::vector<Port *> &GetIterationTargets() { return m_vIterationTargets; }

protected:
    PORTTYPE m_eType;
    DatablockTemplate *m_pTemplate;
    UINT m_uiId;
    std::string m_strBinding;
    UINT m_uiFormalIndex;
    UINT m_uiInOutIndex;
    Task *m_pTask;
    UINT m_uiTaskIndex;
    BOOL m_bMarshallable;
    METAFUNCTION m_eMetaFunction;
    BOOL m_bPermanentBlock;
    std::vector<Channel *> m_vChannels;
    std::vector<Channel *> m_vControlChannels;
    std::vector<Port *> m_vControlPropagationPorts;
    std::vector<Channel *> m_vControlPropagationChannels;
    std::vector<Port *> m_vGatedPorts;
    std::vector<std::pair<Port *, DESCRIPTORFUNC>> m_vDescriptorPorts;
    Port *m_pDescribedPort;
    std::vector<Port *> m_vIterationTargets;
    BOOL m_bDependentBinding = FALSE;
    BOOL m_bSuppressClones = FALSE;
    ACCELERATOR_CLASS m_eDependentClass = ACCELERATOR_CLASS_UNKNOWN;
    int m_nDependentIndex = 0;
};

/* A port that consumes blocks.  Switch ("control") channels are preferred over normal ones;
 * a block failing a channel's destination predicate is released, not delivered. */
class InputPort : public Port {
public:
    InputPort(DatablockTemplate *pTemplate, UINT uiId, const char *lpszBinding, UINT uiFormalIndex,
              UINT uiInOutIndex, PORTTYPE eType = INPUT_PORT);
    virtual ~InputPort();
    /* TRUE when a block can be taken now. */
    virtual BOOL IsReady();
    /* Takes the next block (caller owns one reference), or NULL. */
    virtual Datablock *Pull();
    /* The next block without taking it. */
    virtual Datablock *Peek();
    void SetInOutConsumer(OutputPort *p) { m_pInOutConsumer = p; }
    OutputPort *GetInOutConsumer() { return m_pInOutConsumer; }
    virtual void Drain();

protected:
    Channel *NextReadyChannel();
    OutputPort *m_pInOutConsumer;
};

/* Keeps its last block and re-binds it on every dispatch until a new one arrives
 * (reference: ptask/StickyPort.h). */
class StickyPort : public InputPort {
public:
    StickyPort(DatablockTemplate *pTemplate, UINT uiId, const char *lpszBinding, UINT uiFormalIndex);
    virtual ~StickyPort();
    BOOL IsSticky() override { return TRUE; }
    BOOL IsReady() override;
    /* Returns the sticky block with one new reference for the caller. */
    Datablock *Pull() override;
    Datablock *Peek() override;
    void Drain() override;

protected:
    Datablock *m_pSticky;
};

/* Carries control data, never a kernel argument: iteration counts and allocation sizes
 * (reference: ptask/MetaPort.h; general iteration MetaPort.cpp:927-1010). */
class MetaPort : public InputPort {
public:
    MetaPort(DatablockTemplate *pTemplate, UINT uiId, const char *lpszBinding, UINT uiFormalIndex,
             UINT uiAllocationOutput = (UINT)-1);
    virtual ~MetaPort();
    BOOL IsFormalParameter() override { return FALSE; }
    BOOL IsReady() override;
    /* Runs the meta function before a dispatch.  Returns the dispatch iteration count (simple
     * iteration) and the allocation size (allocation meta ports) through the out-parameters. */
    void Configure(UINT *puiDispatchIterations, UINT *puiAllocationBytes, BOOL *pbHasAllocation);
    /* Called after a dispatch (general iteration bookkeeping is done in Configure). */
    void Finalize();
    void Drain() override;
    void SetAllocationPort(OutputPort *p) { m_pAllocationPort = p; }
    OutputPort *GetAllocationPort() { return m_pAllocationPort; }
    BOOL InGeneralIteration() { return m_pIterationBlock != nullptr; }

protected:
    Datablock *m_pIterationBlock;
    UINT m_uiIterMax;
    UINT m_uiIterCount;
    OutputPort *m_pAllocationPort;
};

/* Produces a block holding its template's initial value whenever one is needed
 * (reference: ptask/InitializerPort.h). */
class InitializerPort : public InputPort {
public:
    InitializerPort(DatablockTemplate *pTemplate, UINT uiId, const char *lpszBinding, UINT uiFormalIndex,
                    UINT uiInOutIndex);
    BOOL IsReady() override;
    Datablock *Pull() override;
    Datablock *Peek() override { return nullptr; }
};

/* Emits the blocks a dispatch produced to every bound channel (fan-out); recycles its blocks
 * through a pool so steady-state dispatch allocates nothing (reference: ptask/OutputPort.h,
 * GetPooledDestinationBuffer OutputPort.cpp:603). */
class OutputPort : public Port {
public:
    OutputPort(DatablockTemplate *pTemplate, UINT uiId, const char *lpszBinding, UINT uiFormalIndex,
               UINT uiInOutIndex);
    virtual ~OutputPort();
    BOOL IsFormalParameter() override;
    /* TRUE when every bound channel can accept a block. */
    BOOL HasCapacity();
    /* A destination block of `cbBytes` (template size when 0), from the pool when possible. */
    Datablock *GetDestinationBlock(UINT cbBytes);
    /* Delivers `pBlock` to the bound channels (applying source predicates) and consumes the
     * caller's reference. */
    void Push(Datablock *pBlock);
    void SetInOutProducer(InputPort *p) { m_pInOutProducer = p; }
    InputPort *GetInOutProducer() { return m_pInOutProducer; }
    void AddPropagatedControlSignal(CONTROLSIGNAL luiCode) { m_luiPropagated |= luiCode; }
    CONTROLSIGNAL TakePropagatedControlSignals() { CONTROLSIGNAL c = m_luiPropagated; m_luiPropagated = DBCTLC_NONE; return c; }
    void SetControlPort(Port *pController, BOOL bInitiallyOpen) { m_pControlPort = pController; m_bGateOpen = bInitiallyOpen; m_bInitiallyOpen = bInitiallyOpen; }
    Port *GetControlPort() { return m_pControlPort; }
    void SetGateOpen(BOOL b) { m_bGateOpen = b; }
    /* The controller saw EOF: flip the gate away from its initial state for one push
     * (reference: OutputPort::SignalGate, OutputPort.cpp:1363). */
    void SignalGate() { m_bGateOpen = !m_bInitiallyOpen; }
    BOOL IsInitiallyOpen() { return m_bInitiallyOpen; }
    BOOL IsGateOpen() { return m_pControlPort == nullptr || m_bGateOpen; }
    BlockPool *GetBlockPool() { return m_pPool; }
    void ReleasePool();

protected:
    InputPort *m_pInOutProducer;
    CONTROLSIGNAL m_luiPropagated;
    Port *m_pControlPort;
    BOOL m_bGateOpen;
    BOOL m_bInitiallyOpen;
    BlockPool *m_pPool;
};

}  // namespace PTask

#endif
