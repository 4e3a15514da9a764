"""The hot kernels keep everything in registers and LDS: no scratch memory and
no VGPR spills (a value demoted to scratch costs a 2.5x slowdown of the batched
kernel).  Reads the compiler's resource report that the build writes next to
the library."""
import os
import re

import pytest

from conftest import ROOT

REPORT = os.path.join(ROOT, "swm_amd", "libswmgpu.resources.txt")
HOT = [
    "_ZN3swm11step_kernelIdLi1ELi1ELi32ELi8ELb0ELb1EEEvNS_3DynIT_EE",   # FP64 default
    "_ZN3swm11step_kernelIdLi1ELi1ELi32ELi8ELb0ELb0EEEvNS_3DynIT_EE",
    "_ZN3swm11step_kernelIdLi1ELi1ELi64ELi4ELb0ELb0EEEvNS_3DynIT_EE",
    "_ZN3swm11step_kernelIfLi1ELi1ELi32ELi8ELb0ELb1EEEvNS_3DynIT_EE",
    "_ZN3swm11step_kernelIfLi1ELi1ELi32ELi8ELb0ELb0EEEvNS_3DynIT_EE",
    "_ZN3swm11step_kernelIdLi3ELi0ELi32ELi8ELb0ELb1EEEvNS_3DynIT_EE",   # linear variant
    "_ZN3swm11step_kernelIdLi2ELi0ELi32ELi8ELb0ELb1EEEvNS_3DynIT_EE",   # mean state
    "_ZN3swm12batch_kernelIdLi1ELi1ELi32ELi8EEEvPKNS_6MemberIT_EEPKNS_10MemberStepEPKii",
    "_ZN3swm12batch_kernelIfLi1ELi1ELi32ELi8EEEvPKNS_6MemberIT_EEPKNS_10MemberStepEPKii",
    "_ZN3swm12batch_kernelIdLi1ELi0ELi32ELi8EEEvPKNS_6MemberIT_EEPKNS_10MemberStepEPKii",
    "_ZN3swm12batch_kernelIfLi1ELi0ELi32ELi8EEEvPKNS_6MemberIT_EEPKNS_10MemberStepEPKii",
    "_ZN3swm18batch_queue_kernelIdLi1ELi1ELi32ELi8EEEvPKNS_6MemberIT_EEPNS_10MemberStepEPKiiiPi",
    "_ZN3swm18batch_queue_kernelIfLi1ELi1ELi32ELi8EEEvPKNS_6MemberIT_EEPNS_10MemberStepEPKiiiPi",
    "_ZN3swm18batch_queue_kernelIdLi1ELi1ELi64ELi4EEEvPKNS_6MemberIT_EEPNS_10MemberStepEPKiiiPi",
    "_ZN3swm18batch_queue_kernelIfLi1ELi1ELi64ELi4EEEvPKNS_6MemberIT_EEPNS_10MemberStepEPKiiiPi",
]


def report():
    if not os.path.exists(REPORT):
        pytest.skip("no resource report: run __graft_entry__.build()")
    out = {}
    for block in open(REPORT).read().split("remark: Function Name: ")[1:]:
        name = block.split()[0]
        get = lambda key: int(re.search(key + r": (\d+)", block).group(1))
        out[name] = {"scratch": get(r"ScratchSize \[bytes/lane\]"), "spill": get(r"VGPRs Spill"),
                     "vgpr": get(r"VGPRs"), "lds": get(r"LDS Size \[bytes/block\]")}
    return out


@pytest.mark.parametrize("kernel", HOT)
def test_hot_kernel_has_no_scratch(kernel):
    r = report()
    assert kernel in r, "kernel not in the build"
    assert r[kernel]["scratch"] == 0 and r[kernel]["spill"] == 0, r[kernel]


def test_fp64_default_tile_fits_five_blocks_per_cu():
    """32x8 FP64: LDS under 32 KiB and at most 96 VGPRs, i.e. five 256-thread
    blocks per CU (the occupancy the step time was measured at)."""
    r = report()["_ZN3swm11step_kernelIdLi1ELi1ELi32ELi8ELb0ELb1EEEvNS_3DynIT_EE"]
    assert r["lds"] <= 32 * 1024 and r["vgpr"] <= 96, r
