"""Extracts the prefix-sum input fixtures the reference's tests hold as data:
g_data[128] (ptaskUnitTest/pfxsum.cpp:22-40) and g_cdata[1024] (ptaskUnitTest/cpfxsum.cpp:25-154),
and writes them with their expected exclusive scans (host_pfxsum, pfxsum.cpp:56-69) to
pfxsum_fixtures.json.  Reads the reference source text only; run from the repository root."""
import json
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/ptaskUnitTest"


def array(path, name):
    text = open(f"{REF}/{path}").read()
    m = re.search(name + r"\[\]\s*=\s*\{(.*?)\};", text, re.S)
    return [int(x) for x in re.findall(r"-?\d+", m.group(1))]


def exclusive(a):
    out, acc = [], 0
    for v in a:
        out.append(acc)
        acc = (acc + v + 2**31) % 2**32 - 2**31
    return out


fx = {}
for key, (path, name, n) in {"g_data": ("pfxsum.cpp", "g_data", 128),
                             "g_cdata": ("cpfxsum.cpp", "g_cdata", None)}.items():
    a = array(path, name)
    if n is not None:
        assert len(a) == n, (key, len(a))
    fx[key] = {"source": f"ptaskUnitTest/{path}", "input": a, "exclusive_scan": exclusive(a)}
fx["g_cdata"]["used_prefix"] = 24  # cpfxsum.cpp:185 sets g_nCArraySize = 24
json.dump(fx, open("tests/golden/pfxsum_fixtures.json", "w"), indent=0)
print({k: len(v["input"]) for k, v in fx.items()})
