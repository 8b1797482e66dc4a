"""Properties of the build (CPU): the fused step kernel's single-network
variants fit the 128 VGPRs a 1024-thread workgroup allows without spilling
to scratch memory. Every variant runs a persistent loop of cycles, so a
spill costs scratch traffic on every cycle; the kernel sits close to the
limit, and small source changes can push the register allocator over it."""
import os
import re

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPORT = os.path.join(HERE, "..", "nemosim_amd", "_build", "step_kernel.resources.txt")


def kernel_resources(text):
    """{kernel name: (VGPRs, scratch bytes per lane)} from the compiler's
    kernel-resource-usage remarks"""
    out = {}
    name = None
    for line in text.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            continue
        m = re.search(r"VGPRs: (\d+)", line)
        if m and name:
            out.setdefault(name, [None, None])[0] = int(m.group(1))
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and name:
            out.setdefault(name, [None, None])[1] = int(m.group(1))
    return out


def test_single_network_step_kernels_do_not_spill():
    if not os.path.exists(REPORT):
        pytest.skip("no build report (build with __graft_entry__.build())")
    res = kernel_resources(open(REPORT).read())
    one = {k: v for k, v in res.items() if "nemo_step_kernel_one" in k}
    assert len(one) == 4, sorted(res)
    for name, (vgprs, scratch) in one.items():
        assert vgprs is not None and vgprs <= 128, (name, vgprs)
        assert scratch == 0, "%s spills %d bytes per lane" % (name, scratch)
