/* ptaskapi.h -- the PTask public API (reference: ptask/ptaskapi.h:11-31): include this one
 * header and link libptask.so.  The MI355X backend is selected by compiling kernels into gfx950
 * code objects (".hsaco") and naming them in Runtime::GetCompiledKernel. */
#ifndef PTK_PTASKAPI_H
#define PTK_PTASKAPI_H

#include "primitive_types.h"
#include "PTaskRuntime.h"
#include "ptaskutils.h"
#include "accelerator.h"
#include "graph.h"
#include "datablock.h"
#include "datablocktemplate.h"
#include "CompiledKernel.h"
#include "hrperft.h"
#include "task.h"
#include "graphInputChannel.h"
#include "graphOutputChannel.h"
#include "internalChannel.h"
#include "InitializerChannel.h"
#include "multichannel.h"
#include "InputPort.h"
#include "OutputPort.h"
#include "StickyPort.h"
#include "MetaPort.h"
#include "shrperft.h"
#include "hosttask.h"

#endif
