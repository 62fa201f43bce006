/* channel.h -- queues of datablocks between ports (reference: ptask/channel.h and the
 * graphInputChannel / graphOutputChannel / internalChannel / InitializerChannel / multichannel
 * headers).  Predicate names and values match channel.h:134-191 of the reference. */
#ifndef PTK_CHANNEL_H
#define PTK_CHANNEL_H

#include <deque>
#include <string>
#include <vector>

#include "Lockable.h"
#include "ReferenceCounted.h"
#include "primitive_types.h"

namespace PTask {

class Task;
class Port;
class Graph;
class Channel;
class Datablock;
class DatablockTemplate;

typedef enum endpt_type_t {
    CE_SRC = 0,
    CE_DST = 1
} CHANNELENDPOINTTYPE;

typedef enum enum_ct_t {
    CT_GRAPH_INPUT,
    CT_GRAPH_OUTPUT,
    CT_INTERNAL,
    CT_MULTI,
    CT_INITIALIZER
} CHANNELTYPE;

static const DWORD DEFAULT_CHANNEL_TIMEOUT = INFINITE;

typedef enum predfailaction_t {
    PFA_RELEASE_BLOCK = 0,
    PFA_FAIL_OPERATION = 1
} PREDICATE_FAILURE_ACTION;

typedef BOOL(__stdcall *LPFNCHANNELPREDICATE)(Channel *pChannel, Datablock *pBlock,
                                               PREDICATE_FAILURE_ACTION eFailAction);

typedef enum _gate_predicate_fns {
    CGATEFN_NONE = 0,
    CGATEFN_CLOSE_ON_EOF = 1,
    CGATEFN_OPEN_ON_EOF = 2,
    CGATEFN_OPEN_ON_BEGINITERATION = 3,
    CGATEFN_CLOSE_ON_BEGINITERATION = 4,
    CGATEFN_OPEN_ON_ENDITERATION = 5,
    CGATEFN_CLOSE_ON_ENDITERATION = 6,
    CGATEFN_DEVNULL = 7,
    CGATEFN_CLOSE_ON_BOF = 8,
    CGATEFN_OPEN_ON_BOF = 9,
    CGATEFN_USER_DEFINED = 10
} CHANNELPREDICATE;

class Channel : public ReferenceCounted, public Lockable {
public:
    Channel(Graph *pGraph, DatablockTemplate *pTemplate, const char *lpszName, CHANNELTYPE eType);
    virtual ~Channel();

    /* CE_DST: a block is waiting; CE_SRC: there is room for one more. */
    virtual BOOL IsReady(CHANNELENDPOINTTYPE type = CE_DST);
    /* Takes the oldest block; the caller owns the returned reference.  Waits up to dwTimeout
     * ms (graph output channels; internal channels never wait). */
    virtual Datablock *Pull(DWORD dwTimeout = 0xFFFFFFFF);
    virtual Datablock *Peek();
    /* Queues pBlock (the channel takes its own reference; the caller keeps theirs).  Graph input
     * channels wait up to dwTimeout ms for room. */
    virtual BOOL Push(Datablock *pBlock, DWORD dwTimeout = 0xFFFFFFFF);
    virtual void SetCapacity(UINT nCapacity);
    virtual UINT GetCapacity();
    virtual CHANNELTYPE GetType() { return m_eType; }
    DatablockTemplate *GetTemplate() { return m_pTemplate; }
    void SetNoDraw() { m_bDraw = false; }
    bool ShouldDraw() { return m_bDraw; }
    virtual char *GetName() { return const_cast<char *>(m_strName.c_str()); }
    virtual size_t GetQueueDepth();
    virtual void Drain();
    virtual void BindPort(Port *pPort, CHANNELENDPOINTTYPE type);
    virtual Port *UnbindPort(CHANNELENDPOINTTYPE type);
    virtual Port *GetBoundPort(CHANNELENDPOINTTYPE type);
    virtual void SetTimeout(DWORD dwTimeoutMilliseconds) { m_dwTimeout = dwTimeoutMilliseconds; }
    virtual LPFNCHANNELPREDICATE GetPredicator(CHANNELENDPOINTTYPE e) { return m_lpfnPredicate[e]; }
    virtual CHANNELPREDICATE GetPredicationType(CHANNELENDPOINTTYPE e) { return m_ePredicate[e]; }
    virtual void SetPredicator(CHANNELENDPOINTTYPE e, LPFNCHANNELPREDICATE lpfn);
    virtual void SetPredicationType(CHANNELENDPOINTTYPE e, CHANNELPREDICATE ePredicate);
    virtual void SetPredicationType(CHANNELENDPOINTTYPE e, CHANNELPREDICATE ePredicate,
                                    PREDICATE_FAILURE_ACTION pfa);
    virtual BOOL IsPredicated(CHANNELENDPOINTTYPE e) { return m_ePredicate[e] != CGATEFN_NONE; }
    PREDICATE_FAILURE_ACTION GetPredicateFailureAction(CHANNELENDPOINTTYPE e) { return m_ePFA[e]; }
    BOOL PassesPredicate(CHANNELENDPOINTTYPE eEndpoint, Datablock *pBlock);
    void SetViewMaterializationPolicy(VIEWMATERIALIZATIONPOLICY e) { m_eViewPolicy = e; }
    VIEWMATERIALIZATIONPOLICY GetViewMaterializationPolicy() { return m_eViewPolicy; }
    /* Control signals this channel's predicates react to (reference: channel.h
     * GetControlSignalsOfInterest). */
    static CONTROLSIGNAL GetControlSignalsOfInterest(CHANNELPREDICATE ePredicate);
    Graph *GetGraph() { return m_pGraph; }

    /* ---- runtime internals (graph lock held) ----------------------------------------------- */
    BOOL HasCapacity();
    /* Queues without waiting; applies nothing (predicates are applied by the pushing port). */
    void Enqueue(Datablock *pBlock);
    Datablock *DequeueLocked();
    Datablock *PeekLocked();
    size_t QueueDepthLocked() { return m_q.size(); }
    /* Deferred descriptor channels fed alongside this one (BindDescriptorPort). */
    void AddDescriptorChannel(Channel *pChannel, DESCRIPTORFUNC eFunc) {
        m_vDescriptorChannels.push_back(std::make_pair(pChannel, eFunc));
    }

protected:
    BOOL PushToQueue(Datablock *pBlock, DWORD dwTimeout);
    void PushDescriptors(Datablock *pBlock);

    Graph *m_pGraph;
    DatablockTemplate *m_pTemplate;
    std::string m_strName;
    CHANNELTYPE m_eType;
    std::deque<Datablock *> m_q;
    UINT m_uiCapacity;
    bool m_bDraw;
    DWORD m_dwTimeout;
    Port *m_pSrcPort;
    Port *m_pDstPort;
    CHANNELPREDICATE m_ePredicate[2];
    LPFNCHANNELPREDICATE m_lpfnPredicate[2];
    PREDICATE_FAILURE_ACTION m_ePFA[2];
    VIEWMATERIALIZATIONPOLICY m_eViewPolicy;
    std::vector<std::pair<Channel *, DESCRIPTORFUNC>> m_vDescriptorChannels;
};

class GraphInputChannel : public Channel {
public:
    GraphInputChannel(Graph *pGraph, DatablockTemplate *pTemplate, const char *lpszName)
        : Channel(pGraph, pTemplate, lpszName, CT_GRAPH_INPUT) {}
};

class GraphOutputChannel : public Channel {
public:
    GraphOutputChannel(Graph *pGraph, DatablockTemplate *pTemplate, const char *lpszName)
        : Channel(pGraph, pTemplate, lpszName, CT_GRAPH_OUTPUT) {}
    /* Output channels grow without bound: the application drains them at its own pace. */
    BOOL IsReady(CHANNELENDPOINTTYPE type = CE_DST) override;
};

class InternalChannel : public Channel {
public:
    InternalChannel(Graph *pGraph, DatablockTemplate *pTemplate, const char *lpszName)
        : Channel(pGraph, pTemplate, lpszName, CT_INTERNAL) {}
};

/* Produces template-initialised blocks on demand, subject to its destination predicate
 * (reference: ptask/InitializerChannel.h). */
class InitializerChannel : public Channel {
public:
    InitializerChannel(Graph *pGraph, DatablockTemplate *pTemplate, const char *lpszName)
        : Channel(pGraph, pTemplate, lpszName, CT_INITIALIZER), m_luiControl(DBCTLC_NONE) {}
    BOOL IsReady(CHANNELENDPOINTTYPE type = CE_DST) override;
    Datablock *Pull(DWORD dwTimeout = 0xFFFFFFFF) override;
    /* Control code stamped on produced blocks (e.g. DBCTLC_BEGINITERATION). */
    void SetInitialPropagatedControlSignal(CONTROLSIGNAL c) { m_luiControl = c; }
    CONTROLSIGNAL GetInitialPropagatedControlSignal() { return m_luiControl; }
    Datablock *Produce();

private:
    CONTROLSIGNAL m_luiControl;
};

/* One graph input that feeds several ports (reference: ptask/multichannel.h). */
class MultiChannel : public Channel {
public:
    MultiChannel(Graph *pGraph, DatablockTemplate *pTemplate, const char *lpszName)
        : Channel(pGraph, pTemplate, lpszName, CT_MULTI) {}
    void CoalesceChannel(Channel *pChannel) { m_vMembers.push_back(pChannel); }
    BOOL Push(Datablock *pBlock, DWORD dwTimeout = 0xFFFFFFFF) override;
    BOOL IsReady(CHANNELENDPOINTTYPE type = CE_DST) override;
    std::vector<Channel *> &GetMembers() { return m_vMembers; }

private:
    std::vector<Channel *> m_vMembers;
};

}  // namespace PTask

#endif
