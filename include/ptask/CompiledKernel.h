/* CompiledKernel.h -- a kernel resolved for every accelerator class able to run it
 * (reference: ptask/CompiledKernel.h; produced by Runtime::GetCompiledKernel,
 * ptask/PTaskRuntime.cpp:2904-2970). */
#ifndef PTK_COMPILEDKERNEL_H
#define PTK_COMPILEDKERNEL_H

#include <map>
#include <string>

#include "primitive_types.h"

namespace PTask {

class Accelerator;

class CompiledKernel {
public:
    CompiledKernel(const char *lpszSourceFile, const char *lpszOperation,
                   ACCELERATOR_CLASS eClass);
    ~CompiledKernel();
    const char *GetSourceFile();
    const char *GetOperation();
    ACCELERATOR_CLASS GetAcceleratorClass();
    /* Platform object for an accelerator: a ptk_hip_function for HIP, a host function pointer
     * for the host. */
    void *GetPlatformSpecificBinary(Accelerator *pAccelerator);
    void SetPlatformSpecificBinary(Accelerator *pAccelerator, void *pBinary, void *pModule);
    void *GetPlatformSpecificModule(Accelerator *pAccelerator);
    BOOL HasBinaryFor(Accelerator *pAccelerator);
    void SetGroupSize(int tgx, int tgy, int tgz) { m_tgx = tgx; m_tgy = tgy; m_tgz = tgz; }

private:
    std::string m_strFile;
    std::string m_strOp;
    ACCELERATOR_CLASS m_eClass;
    std::map<Accelerator *, std::pair<void *, void *>> m_binaries;
    int m_tgx, m_tgy, m_tgz;
};

}  // namespace PTask

#endif
