/* hosttask.h -- host-task entry point signatures (reference: ptask/hosttask.h:22-148).
 *
 * A host task is a function in a shared library (loaded by GetCompiledKernel on a ".so" file) or
 * registered with GetHostFunctionCompiledKernel.  Plain host tasks receive their arguments as
 * host pointers (INT and FLOAT values in the pointer slot, by-value structs by pointer).  A host
 * task whose task object has a dependent accelerator class (Task::BindDependentAcceleratorClass)
 * is a dependent host task: arguments of ports bound to the dependent accelerator
 * (Port::BindDependentAccelerator) arrive as device pointers of that accelerator's memory, with
 * the device's ordinal and, in the EX form, the stream the task should enqueue its device work
 * on (a hipStream_t on this build). */
#ifndef PTK_HOSTTASK_H
#define PTK_HOSTTASK_H

#include "primitive_types.h"

namespace PTask {
class Datablock;
}

/* Plain host task: called X*Y*Z times per dispatch (Task::SetComputeGeometry). */
typedef void(__stdcall *LPFNHOSTTASK)(UINT nArguments, void **ppArguments);

/* Dependent host task, legacy form: pbIsDependentBinding[i] says whether argument i is a device
 * pointer; pvDeviceBindings[i] is then its device (an ordinal); pDeps holds the nDeps dependent
 * devices.  The task may use the devices synchronously: the runtime waits for the device after
 * the call. */
typedef void(__stdcall *LPFNDEPHOSTTASK)(UINT nArguments, void **ppArguments, BOOL *pbIsDependentBinding,
                                         void **pvDeviceBindings, UINT nDeps, void **pDeps);

/* The descriptor of the EX form (field order and sizes as the reference's). */
typedef struct _dependent_context_t {
    UINT cbDependentContext;            /* sizeof(DEPENDENTCONTEXT) */
    UINT nArguments;
    UINT nDependentAccelerators;
    UINT uiReserved0;
    void **ppArguments;
    PTask::Datablock **ppDatablocks;    /* the block behind each argument */
    BOOL *pbIsDependentBinding;         /* per argument: a device pointer */
    void **pvDeviceBindings;            /* per argument: its device (ordinal) when dependent */
    void **pDependentDevices;           /* nDependentAccelerators devices (ordinals) */
    void **pStreams;                    /* nDependentAccelerators streams (hipStream_t) when the task
                                           asked for platform objects, else NULL entries */
    char *lpszTaskName;
} DEPENDENTCONTEXT, *LPDEPENDENTCONTEXT;

/* Dependent host task, EX form: used when BindDependentAcceleratorClass was called with
 * bRequestDependentPSObjects == TRUE.  Device work enqueued on pStreams[0] is ordered before every
 * later reader of the task's outputs; the runtime does not wait for it. */
typedef void(__stdcall *LPFNDEPHOSTTASKEX)(LPDEPENDENTCONTEXT);

namespace PTask {
static const int MAXARGS = 64;
}

#endif
