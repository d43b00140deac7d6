"""Register budget of the sweep kernels, read from the compiler's resource
report (hipcc -Rpass-analysis=kernel-resource-usage, device code only).

The sweeps are bound by VALU issue and latency; two allocation accidents cost
a third or more of a kernel's speed without changing a single result: a
private array the compiler cannot keep in registers (ScratchSize > 0, e.g. a
selection between two structs made through a pointer), and a body one
register over the 256 that two waves per SIMD allow (one wave per SIMD).
This pins both for the throughput kernels: no scratch anywhere, and at least
two waves per SIMD for the gamma = 5/3 Sod sweeps of both precisions and for
the batch kernels of the gamma = 5/3 and mixed forms.  Runs on the CPU
(cross-compilation only)."""
import os
import re
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

from hydropad_amd import build

FIELDS = {"VGPRs": "vgpr", "AGPRs": "agpr", "ScratchSize [bytes/lane]": "scratch", "Occupancy [waves/SIMD]": "waves"}


def resources(src):
    cmd = [build.hipcc()] + build.flags("fast") + ["-Rpass-analysis=kernel-resource-usage", "--cuda-device-only",
                                                  "-c", os.path.join(build.CSRC, src), "-o", os.devnull]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out, cur = {}, None
    for line in r.stderr.splitlines():
        m = re.search(r"remark: Function Name: (\S+)", line)
        if m:
            cur = out.setdefault(m.group(1), {})
            continue
        m = re.search(r"remark:\s+([A-Za-z/\[\] ]+): (\d+)", line)
        if cur is not None and m and m.group(1).strip() in FIELDS:
            cur[FIELDS[m.group(1).strip()]] = int(m.group(2))
    return out


@pytest.fixture(scope="module")
def report():
    with ThreadPoolExecutor(2) as ex:
        fast, batch = ex.map(resources, ["hp_kernels_fast.hip", "hp_batch_kernels.hip"])
    return fast, batch


def test_no_kernel_uses_scratch(report):
    for rep in report:
        assert rep, "no resource remarks parsed"
        for name, r in rep.items():
            assert r.get("scratch", 0) == 0, (name, r)


def test_throughput_sweeps_keep_two_waves(report):
    fast, batch = report
    # FastMath53 sweeps without gravity, expansion terms or fused expansion
    sod = [n for n in fast if "FastMath53" in n and re.search(r"E[df](Li[23]E)?Lb0ELb0ELb0E", n)]
    assert len(sod) == 6, sod
    held = [n for n in batch if "batch_" in n and ("FastMath53" in n or re.search(r"FastMathEf", n))]
    assert len(held) == 18, held
    for n in sod + held:
        assert batch.get(n, fast.get(n))["waves"] >= 2, (n, batch.get(n, fast.get(n)))
