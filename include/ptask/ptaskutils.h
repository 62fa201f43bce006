/* ptaskutils.h -- small helpers of the public API (reference: ptask/ptaskutils.h). */
#ifndef PTK_PTASKUTILS_H
#define PTK_PTASKUTILS_H

#include "primitive_types.h"

namespace PTask {
class ptaskutils {
public:
    /* Accelerator class implied by a kernel file name (reference: ptaskutils.cpp:105-133). */
    static ACCELERATOR_CLASS SelectAcceleratorClass(const char *lpszFileName);
    static UINT nextpow2(UINT x);
};
}  // namespace PTask

#endif
