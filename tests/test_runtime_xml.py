"""XML graph serialization round trips: Graph::Serialize / Graph::Deserialize
(ptask_amd/csrc/runtime/xml.cpp; reference XMLWriter.cpp:124-620, XMLReader.cpp:383-404, and the
unit tests' serialization modes 1/2, e.g. graphcuadd.cpp:70-80,148-155).

Checked on the CPU with host tasks:
- every structural feature the writer knows survives write -> read -> write unchanged (the second
  file is the first byte for byte), and the text holds each feature where Python's own XML parser
  finds it (an independent reading of the document);
- graphs read from XML compute what the built graphs compute: the feature graph against its
  closed forms, the FDTD graphs (looped and unrolled, one-step and multi-step forms with their
  remainder task) against the oracle bit for bit, the exact-size groupby graph byte for byte;
- a file holding only the reference writer's elements (no optional additions) still loads and
  runs;
- malformed files and template mismatches fail with a message and a nonzero exit, and
  Serialize refuses function pointers and duplicate port UIDs or channel names without writing.

On the GPU: HIP FDTD graphs read from XML equal the built graphs bit for bit, and a file recorded
for CUDA (kernel file .ptx, accelerator class 2) loads onto the HIP code object of the same stem.
"""
import json
import os
import re
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import oracle
from ptask_amd import paths

XMLAPP = paths.bin_path("ptk_xmlgraph")
FDTD = paths.bin_path("ptk_fdtd")
GROUPBY = paths.bin_path("ptk_groupby")
HOSTLIB = os.path.join(oracle.ORACLE_DIR, "_build", "libptk_oracle.so")
HSACO = paths.kernel_path("fdtd.hsaco")
HSACO_FUSED = paths.kernel_path("fdtd_fused.hsaco")
SRC = os.path.join(os.path.dirname(__file__), "hosttasks", "record_tasks.c")

# elements the reference's writer emits (XMLWriter.cpp, Recorder.cpp Write methods)
REFERENCE_ELEMENTS = {
    "Graph", "DatablockTemplates", "DatablockTemplate", "Count", "Name", "Stride", "ParamType", "ElementsX",
    "ElementsY", "ElementsZ", "IsRecordStream", "IsByteAddressable", "CompiledKernels", "CompiledKernel",
    "KernelID", "SourceFile", "Operation", "Tasks", "Task", "AcceleratorClass", "InputPorts", "OutputPorts",
    "Port", "UniqueID", "OriginalIndex", "Type", "VariableBinding", "ParameterIndex", "InOutIndex", "Channels",
    "Channel", "SourcePort", "DestinationPort", "Actions", "Action", "ActionName", "ControllerPort", "GatedPort",
    "InitiallyOpen", "InputPort", "ControlledChannel", "OutputPort", "DescribedPort", "DescriberPort",
    "DescriptorFunction", "MetaPort", "ScopedPort", "TaskName", "BlockXDim", "BlockYDim", "BlockZDim", "GridXDim",
    "GridYDim", "GridZDim", "PreferredXDim", "PreferredYDim", "PreferredZDim", "ChannelName", "EndpointType",
    "CanonicalPredicate",
}


@pytest.fixture(scope="module")
def hostlib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("hosttasks") / "record_tasks.so")
    subprocess.run(["gcc", "-O2", "-shared", "-fPIC", "-o", out, SRC, "-lpthread"], check=True)
    return out


def run(app, *args, ok=True):
    cmd = [app, *map(str, args)]
    print("run:", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    if ok:
        assert p.returncode == 0, p.stderr[-3000:] + p.stdout[-2000:]
        return json.loads(p.stdout.strip().splitlines()[-1])
    return p


def read(path):
    with open(path) as f:
        return f.read()


# ---- the feature graph --------------------------------------------------------------------------

def test_feature_graph_write_read_write_is_a_fixpoint_and_runs_the_same(hostlib, tmp_path):
    a, b = str(tmp_path / "a.xml"), str(tmp_path / "b.xml")
    built = run(XMLAPP, "--kernels", hostlib, "--run", "--write", a)
    assert built["ok"] and built["tasks"] == 4 and built["channels"] == 17
    again = run(XMLAPP, "--read", a, "--write", b)
    assert again["graph"] == "xmlfeatures" and again["channels"] == 17
    assert read(a) == read(b)
    loaded = run(XMLAPP, "--read", a, "--run")
    assert loaded["ok"] and loaded["out2"]["bad"] == 0 and loaded["out3"]["bad"] == 0
    assert loaded["out2"] == built["out2"] and loaded["out3"] == built["out3"]


def test_feature_graph_document_holds_every_feature(hostlib, tmp_path):
    a = str(tmp_path / "a.xml")
    run(XMLAPP, "--kernels", hostlib, "--write", a)
    root = ET.parse(a).getroot()
    assert root.tag == "Graph" and root.findtext("Name") == "xmlfeatures"
    tmpl = {t.findtext("Name"): t for t in root.find("DatablockTemplates").findall("DatablockTemplate")}
    assert int(root.find("DatablockTemplates").findtext("Count")) == len(tmpl) == 6
    # initial values: 3.0f and the vector length 64, little-endian bytes
    assert tmpl["xgScale3"].find("InitialValue").findtext("Bytes") == np.float32(3.0).tobytes().hex()
    assert tmpl["xgN"].find("InitialValue").findtext("Bytes") == np.int32(64).tobytes().hex()
    assert tmpl["xgScale"].findtext("ParamType") == "1" and tmpl["xgVec"].findtext("ParamType") == "5"
    rec = tmpl["xgRecords"]
    assert rec.findtext("Pitch") == str(64 * 4 + 64) and rec.findtext("IsRecordStream") == "true"
    meta = rec.find("BufferDimensions")
    assert meta.findtext("Channel") == "1" and meta.findtext("ElementsX") == "16" and meta.findtext("Pitch") == "64"
    kernels = root.find("CompiledKernels").findall("CompiledKernel")
    assert len(kernels) == 1 and kernels[0].findtext("Operation") == "ScaleVector"
    tasks = {t.findtext("Name"): t for t in root.find("Tasks").findall("Task")}
    assert set(tasks) == {"scale1", "scale2", "scale3", "probe"}
    assert tasks["scale2"].findtext("Priority") == "5" and tasks["probe"].findtext("UserCodeAllocatesMemory") == "true"
    assert all(t.findtext("AcceleratorClass") == "4" for t in tasks.values())  # ACCELERATOR_CLASS_HOST
    probe = {p.findtext("VariableBinding"): p for p in tasks["probe"].find("InputPorts").findall("Port")}
    assert probe["repeat"].findtext("Type") == "3" and probe["repeat"].findtext("MetaFunction") is not None
    assert probe["s"].findtext("PermanentBlock") == "true" and probe["A"].findtext("Marshallable") == "false"
    uids = [int(p.findtext("UniqueID")) for p in root.iter("Port")]
    assert len(uids) == len(set(uids)) == 4 * 4 + 1
    chans = {c.findtext("Name"): c for c in root.find("Channels").findall("Channel")}
    assert int(root.find("Channels").findtext("Count")) == len(chans) == 15  # the 2 members sit in "A"
    members = chans["A"].find("Members").findall("Member")
    assert [m.findtext("Name") for m in members] == ["A#0", "A#1"]
    assert chans["s3"].findtext("Switch") == "true" and chans["s2"].findtext("Switch") is None
    assert chans["out3"].findtext("Capacity") == "7" and chans["out3"].findtext("ViewMaterialization") == "1"
    assert chans["probe.n"].findtext("InitialControlSignal") is not None
    assert chans["scale1->scale2"].findtext("Type") == "2"
    actions = root.find("Actions").findall("Action")
    assert int(root.find("Actions").findtext("Count")) == len(actions)
    kinds = [x.findtext("ActionName") for x in actions]
    for k in ["BindControlPort", "BindControlPropagationPort", "BindControlPropagationChannel", "BindDescriptorPort",
              "BindIterationScope", "SetComputeGeometry", "SetBlockAndGridSize", "SetPredicationType"]:
        assert k in kinds, k
    gate = next(x for x in actions if x.findtext("ActionName") == "BindControlPort")
    assert gate.findtext("InitiallyOpen") == "false"
    grid = next(x for x in actions if x.findtext("ActionName") == "SetBlockAndGridSize")
    assert grid.findtext("TaskName") == "probe" and grid.findtext("GridYDim") == "2" and grid.findtext("BlockXDim") == "64"
    preds = {(x.findtext("ChannelName"), x.findtext("EndpointType")): x for x in actions
             if x.findtext("ActionName") == "SetPredicationType"}
    assert set(preds) == {("probe.A", "1"), ("probe.out", "0")}
    assert preds[("probe.A", "1")].findtext("FailureAction") == "1"


def test_reader_ignores_layout_comments_and_entities(hostlib, tmp_path):
    """Reformatted text (one line, comments, character references) reads to the same graph."""
    a, b, c = str(tmp_path / "a.xml"), str(tmp_path / "b.xml"), str(tmp_path / "c.xml")
    run(XMLAPP, "--kernels", hostlib, "--write", a)
    text = read(a)
    squashed = re.sub(r">\s+<", "><", text)
    squashed = squashed.replace("<Name>scale2</Name>", "<!-- renamed? no --><Name>&#115;cale2</Name>")
    squashed = squashed.replace("<Count>", "<Count >", 1)
    with open(b, "w") as f:
        f.write(squashed)
    run(XMLAPP, "--read", b, "--write", c)
    assert read(c) == text


# ---- applications' graphs -----------------------------------------------------------------------

def direct(nx, ny, nz, steps, seed=0x5EED):
    F = oracle.init_fields(nx, ny, nz, seed)
    return oracle.fdtd_run(F, oracle.make_params(nx, ny, nz, 0.5), steps)


def load_dump(path, nx, ny, nz):
    raw = np.fromfile(path, dtype=np.float32)
    out, off = [], 0
    for f in range(6):
        n = int(np.prod(oracle.extent(f, nx, ny, nz)))
        out.append(raw[off:off + n])
        off += n
    assert off == raw.size
    return out


@pytest.mark.parametrize("mode,form,steps", [("looped", "six", 5), ("unrolled", "halves", 3),
                                              ("looped", "overlap3", 7), ("unrolled", "overlap2", 5)])
def test_fdtd_graph_from_xml_equals_the_oracle(tmp_path, mode, form, steps):
    nx, ny, nz = 8, 7, 6
    a, b, dump = str(tmp_path / "a.xml"), str(tmp_path / "b.xml"), str(tmp_path / "d.bin")
    geo = ["--nx", nx, "--ny", ny, "--nz", nz, "--steps", steps, "--mode", mode, "--form", form]
    w = run(FDTD, "--kernels", HOSTLIB, *geo, "--serialize", a)
    assert w["ok"]
    run(FDTD, "--deserialize", a, *geo, "--serialize", b)
    assert read(a) == read(b)
    r = run(FDTD, "--deserialize", a, *geo, "--dump", dump, "--blocks", 2 if mode == "unrolled" else 1)
    assert r["ok"] and r["backend"] == "host"
    want = direct(nx, ny, nz, steps)
    for f, got in enumerate(load_dump(dump, nx, ny, nz)):
        assert np.array_equal(got, want[f].ravel()), f


def test_groupby_exact_size_graph_from_xml_matches_the_built_graph(tmp_path):
    a, d1, d2 = str(tmp_path / "a.xml"), str(tmp_path / "1.bin"), str(tmp_path / "2.bin")
    opts = ["--rows", 3000, "--keys", 200]
    run(GROUPBY, "--kernels", HOSTLIB, *opts, "--coalesce-exact", "--serialize", a)
    built = run(GROUPBY, "--kernels", HOSTLIB, *opts, "--coalesce-exact", "--dump", d1)
    loaded = run(GROUPBY, "--deserialize", a, *opts, "--dump", d2)
    assert [o["name"] for o in loaded["outputs"]] == ["groups", "outKeys", "outVals", "coalesced", "groupCounts"]
    assert loaded["outputs"] == built["outputs"]
    with open(d1, "rb") as f1, open(d2, "rb") as f2:
        assert f1.read() == f2.read()
    root = ET.parse(a).getroot()
    metas = [p for p in root.iter("Port") if p.findtext("MetaFunction") is not None]
    assert sorted(p.findtext("InOutIndex") for p in metas) == ["0", "1"]  # each sizes its output


def strip_to_reference_schema(text, cuda=False):
    """Only the reference writer's elements; optionally recorded for CUDA (class 2, .ptx)."""
    root = ET.fromstring(text)

    def prune(e):
        for k in list(e):
            if k.tag not in REFERENCE_ELEMENTS or (e.tag == "Graph" and k.tag == "Name"):
                e.remove(k)
            else:
                prune(k)
    prune(root)
    if cuda:
        for k in root.iter("SourceFile"):
            k.text = re.sub(r"\.hsaco$", ".ptx", k.text)
        for t in root.iter("AcceleratorClass"):
            t.text = "2"
    return ET.tostring(root, encoding="unicode")


def test_reference_schema_file_loads_and_runs(tmp_path):
    nx, ny, nz, steps = 6, 5, 4, 3
    a, ref, dump = str(tmp_path / "a.xml"), str(tmp_path / "ref.xml"), str(tmp_path / "d.bin")
    geo = ["--nx", nx, "--ny", ny, "--nz", nz, "--steps", steps, "--mode", "unrolled", "--form", "six"]
    run(FDTD, "--kernels", HOSTLIB, *geo, "--serialize", a)
    with open(ref, "w") as f:
        f.write(strip_to_reference_schema(read(a)))
    assert "<Capacity>" not in read(ref) and "<Switch>" not in read(ref)
    r = run(FDTD, "--deserialize", ref, *geo, "--dump", dump)
    assert r["ok"]
    want = direct(nx, ny, nz, steps)
    for f, got in enumerate(load_dump(dump, nx, ny, nz)):
        assert np.array_equal(got, want[f].ravel()), f


# ---- errors -------------------------------------------------------------------------------------

@pytest.mark.parametrize("edit,message", [
    (lambda t: t[: len(t) // 2], "XML line"),
    (lambda t: t.replace("<ActionName>BindIterationScope</ActionName>", "<ActionName>Frobnicate</ActionName>"),
     "unknown action"),
    (lambda t: t.replace("<Count>4</Count>", "<Count>5</Count>", 1), "has Count"),
    (lambda t: t.replace("<Operation>ScaleVector</Operation>", "<Operation>NoSuchKernel</Operation>"),
     "could not be loaded"),
    (lambda t: t.replace("</ScopedPort>", "9</ScopedPort>"), "unknown port"),
    (lambda t: t.replace("</Graph>", "<Z>" * 300 + "</Z>" * 300 + "</Graph>"), "nested deeper"),
    (lambda t: t.replace("<Operation>ScaleVector", "<Operation>Scale&#;Vector"), "bad character reference"),
    (lambda t: t.replace("<Operation>ScaleVector", "<Operation>Scale&#0;Vector"), "bad character reference"),
    (lambda t: t.replace("<Operation>ScaleVector", "<Operation>Scale&#x110000;Vector"),
     "bad character reference"),
    (lambda t: t.replace("<Operation>ScaleVector", "<Operation>Scale&#12a;Vector"), "bad character reference"),
])
def test_malformed_files_fail_loudly(hostlib, tmp_path, edit, message):
    a, b = str(tmp_path / "a.xml"), str(tmp_path / "b.xml")
    run(XMLAPP, "--kernels", hostlib, "--write", a)
    with open(b, "w") as f:
        f.write(edit(read(a)))
    p = run(XMLAPP, "--read", b, ok=False)
    assert p.returncode != 0 and message in p.stderr, p.stderr


def test_template_mismatch_is_an_error(tmp_path):
    a = str(tmp_path / "a.xml")
    run(GROUPBY, "--kernels", HOSTLIB, "--rows", 3000, "--keys", 200, "--serialize", a)
    p = run(GROUPBY, "--deserialize", a, "--rows", 4000, "--keys", 200, ok=False)
    assert p.returncode != 0 and "differs from the runtime's template" in p.stderr, p.stderr


# ---- GPU ----------------------------------------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("mode,form,steps", [("looped", "overlap3p", 7), ("unrolled", "overlap2", 5),
                                              ("looped", "fused", 4)])
def test_gpu_fdtd_graph_from_xml_equals_the_built_graph(tmp_path, mode, form, steps):
    nx, ny, nz = 40, 21, 13
    a, d1, d2 = str(tmp_path / "a.xml"), str(tmp_path / "1.bin"), str(tmp_path / "2.bin")
    geo = ["--nx", nx, "--ny", ny, "--nz", nz, "--steps", steps, "--mode", mode, "--form", form]
    run(FDTD, "--kernels", HSACO_FUSED, *geo, "--serialize", a)
    built = run(FDTD, "--kernels", HSACO_FUSED, *geo, "--dump", d1)
    loaded = run(FDTD, "--deserialize", a, *geo, "--dump", d2)
    assert built["ok"] and loaded["ok"] and loaded["backend"] == "hip"
    assert loaded["dispatches_last_run"] == built["dispatches_last_run"]
    with open(d1, "rb") as f1, open(d2, "rb") as f2:
        assert f1.read() == f2.read()


@pytest.mark.gpu
def test_gpu_cuda_recorded_file_loads_onto_hip(tmp_path):
    """The reference writer's elements only, kernel file .ptx and class CUDA: the reader loads the
    .hsaco of the same stem and makes HIP tasks; the six-kernel graph equals the oracle."""
    nx, ny, nz, steps = 12, 9, 7, 3
    a, ref, dump = str(tmp_path / "a.xml"), str(tmp_path / "ref.xml"), str(tmp_path / "d.bin")
    geo = ["--nx", nx, "--ny", ny, "--nz", nz, "--steps", steps, "--mode", "unrolled", "--form", "six"]
    run(FDTD, "--kernels", HSACO, *geo, "--serialize", a)
    with open(ref, "w") as f:
        f.write(strip_to_reference_schema(read(a), cuda=True))
    assert ".ptx</SourceFile>" in read(ref) and "<AcceleratorClass>2<" in read(ref)
    r = run(FDTD, "--deserialize", ref, *geo, "--dump", dump)
    assert r["ok"] and r["backend"] == "hip"
    want = direct(nx, ny, nz, steps)
    for f, got in enumerate(load_dump(dump, nx, ny, nz)):
        assert np.array_equal(got, want[f].ravel()), f


@pytest.mark.parametrize("what,message", [
    ("estimator", "geometry estimator function"),
    ("predicate", "user-defined predicate function"),
    ("uid", "share UniqueID"),
    ("channel", "two channels are named 's2'"),
])
def test_serialize_refuses_what_has_no_text_form(hostlib, tmp_path, what, message):
    """Function pointers and ambiguous names: Serialize fails with a message and writes nothing."""
    a = tmp_path / "a.xml"
    p = run(XMLAPP, "--kernels", hostlib, "--unwritable", what, "--write", str(a), ok=False)
    assert p.returncode != 0 and message in p.stderr, p.stderr
    assert not a.exists() and not (tmp_path / "a.xml.tmp").exists()
