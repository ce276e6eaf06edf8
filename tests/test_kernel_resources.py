"""Build-time check of the gfx950 kernels: none spills to per-lane scratch.
A kernel that indexes its argument struct with a run-time value, or keeps
more live values than the register file holds, gets private scratch memory,
which costs every lane a round trip to memory per access (a batched launch
of such a kernel once took 2 ms per frame instead of 0.3 ms). Compiles each
source for the device only, with the compiler's resource-usage remarks."""
import re
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "frostbitten-hair-webgpu_amd"))


@pytest.mark.skipif(shutil.which("/opt/rocm/bin/hipcc") is None, reason="hipcc not installed")
def test_no_kernel_uses_private_scratch(tmp_path):
    import build as B

    def remarks(src):
        cmd = [B.HIPCC, *B.COMMON_FLAGS, "--offload-arch=gfx950", "--cuda-device-only", "-c", str(B.CSRC / src),
               "-o", str(tmp_path / (src + ".o")), "-Rpass-analysis=kernel-resource-usage"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]
        return src, r.stderr

    with ThreadPoolExecutor(max_workers=4) as ex:
        outs = list(ex.map(remarks, B.SOURCES))
    spills, kernels = [], 0
    for src, text in outs:
        name = None
        for line in text.splitlines():
            m = re.search(r"Function Name: (\S+)", line)
            if m:
                name = m.group(1)
                kernels += 1
            m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
            if m and int(m.group(1)) > 0:
                spills.append(f"{src}: {name}: {m.group(1)} B/lane")
    assert kernels > 50, "no resource-usage remarks found"
    assert not spills, "kernels with private scratch:\n" + "\n".join(spills)
