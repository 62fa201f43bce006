"""Every task kernel a kernel source defines is present in its built code object.

The applications load kernels by name from lib/kernels/<source>.hsaco (GetCompiledKernel); a
definition that drops out of a source (an include lost, a kernel renamed) would otherwise show
only when a GPU run asks for it.  This reads the sources and the built objects as text and bytes;
nothing runs on a GPU."""
import os
import re

import pytest

from ptask_amd import paths

KSRC = os.path.join(os.path.dirname(paths.KERNEL_DIR), "..", "csrc", "kernels")
DEF = re.compile(r'extern\s+"C"\s+__global__\s+(?:void\s+)?(?:__launch_bounds__\([^)]*\)\s+)?(?:void\s+)?(\w+)\s*\(')
INC = re.compile(r'#include\s+"([^"]+)"')
MACRO = re.compile(r'#define\s+(\w+)\s+extern\s+"C"\s+__global__')


def _defined(path, seen):
    """Kernel names defined in a source and in the local headers it includes."""
    if path in seen or not os.path.exists(path):
        return set()
    seen.add(path)
    text = open(path).read()
    names = {n for n in DEF.findall(text) if n != "__launch_bounds__"}
    for macro in MACRO.findall(text):
        names |= set(re.findall(r"^" + macro + r"\s+(\w+)\s*\(", text, re.M))
    for inc in INC.findall(text):
        names |= _defined(os.path.join(os.path.dirname(path), inc), seen)
    return names


SOURCES = sorted(f for f in os.listdir(KSRC) if f.endswith(".hip")) if os.path.isdir(KSRC) else []


def test_sources_found():
    assert "pca.hip" in SOURCES and "fdtd.hip" in SOURCES


@pytest.mark.parametrize("src", SOURCES)
def test_code_object_holds_every_defined_kernel(src):
    obj = paths.kernel_path(src[:-4] + ".hsaco")
    if not os.path.exists(obj):
        pytest.skip(f"{obj} not built")
    names = _defined(os.path.join(KSRC, src), set())
    assert names, f"no kernels found in {src}"
    data = open(obj, "rb").read()
    missing = sorted(n for n in names if (n + ".kd").encode() not in data)
    assert not missing, f"{obj} lacks {missing}"


def test_pca_object_holds_the_blas_tasks():
    names = _defined(os.path.join(KSRC, "pca.hip"), set())
    for n in ("PcaStart", "PcaGram", "PcaDeflate", "PcaFinish", "ptaskDcopy", "ptaskDaxpy", "ptaskDgemv",
              "ptaskDnrm2", "ptaskDscal", "ptaskDger"):
        assert n in names
