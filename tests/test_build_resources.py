"""Register budgets of the hot kernels, from the compiler's resource report
that the build writes next to libmdn.so (no GPU needed).  A change that
makes the compiler spill the batched kernel's registers costs it about a
third of its speed with no other symptom, so the budget is checked here."""
import json
import os

import pytest

from md_nucleation_amd import build as B


def _resources():
    if not os.path.exists(B.RESOURCES):
        pytest.fail("%s is missing: build the library with md_nucleation_amd.build" % B.RESOURCES)
    return json.load(open(B.RESOURCES))


def _kernel(res, prefix):
    hits = [v for k, v in res.items() if k.startswith(prefix)]
    assert len(hits) == 1, (prefix, sorted(res))
    return hits[0]


def test_batched_kernel_register_budget():
    """the variant the benchmark runs (no many-body ion terms) spills
    nothing; the one with Tersoff / SW / BMH at most 32 VGPRs"""
    res = _resources()
    lean = _kernel(res, "_ZN3mdn9k_batchedIfLb0E")
    mb = _kernel(res, "_ZN3mdn9k_batchedIfLb1E")
    print("k_batched<float, false>:", lean, "k_batched<float, true>:", mb)
    assert lean["vgprs"] <= 256 and lean["vgpr_spill"] == 0
    assert mb["vgprs"] <= 256 and mb["vgpr_spill"] <= 32


def test_all_pairs_vv_kernels_do_not_spill():
    res = _resources()
    n3 = _kernel(res, "_ZN3mdn7k_vv_n3IfLb0E")
    tiled = _kernel(res, "_ZN3mdn10k_vv_tiledIfLb0E")
    print("k_vv_n3<float>:", n3, "k_vv_tiled<float>:", tiled)
    assert n3["vgpr_spill"] == 0 and tiled["vgpr_spill"] == 0
    # five waves per SIMD for the Newton-3 kernel (at most 96 VGPRs), which
    # puts every workgroup of the 12,000-atom cluster on the chip at once
    assert n3["vgprs"] <= 96


def test_resource_report_parser():
    text = """x.hip:1:1: remark: Function Name: _ZN3mdn5k_fooEv [-Rpass-analysis=kernel-resource-usage]
x.hip:1:1: remark:     VGPRs: 40 [-Rpass-analysis=kernel-resource-usage]
x.hip:1:1: remark:     ScratchSize [bytes/lane]: 16 [-Rpass-analysis=kernel-resource-usage]
x.hip:1:1: remark:     SGPRs Spill: 2 [-Rpass-analysis=kernel-resource-usage]
x.hip:1:1: remark:     VGPRs Spill: 3 [-Rpass-analysis=kernel-resource-usage]
"""
    assert B.kernel_resources(text) == {"_ZN3mdn5k_fooEv": {"vgprs": 40, "scratch": 16, "sgpr_spill": 2,
                                                             "vgpr_spill": 3}}
