"""Residency budget of the persistent kernels (host-side check, no GPU).

The persistent CG and Newton-GMRES kernels are launched with one workgroup
per part of the mesh and exchange partial sums between workgroups of the same
launch, so every workgroup of the grid must be resident at once.  The grid is
sized from the number of workgroups a CU holds.  That number is limited by the
kernel's VGPRs (the compiler's occupancy figure) and by its LDS, which the
hardware hands out in whole allocation granules: a workgroup of 54,592 B
rounds to 55,296 B at a 2 KiB granule, three of them exceed the 160 KiB of a
gfx950 CU, and a grid sized for three per CU leaves workgroups that never
start while the others wait for them.

This test compiles the two kernel sources for gfx950 with the compiler's
resource report and checks, for every instantiation, that the occupancy the
compiler reports (waves per SIMD; one wave per SIMD per 256-thread workgroup)
times the granule-rounded LDS fits the CU, and that the grid helper rounds the
same way (kLdsGranule in nsim_dev_common.h).
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "nmag-src_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"
LDS_PER_CU = 160 * 1024
GRANULE = 2048


def resource_report(src, tmp_path):
    out = subprocess.run(
        [HIPCC, "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Rpass-analysis=kernel-resource-usage", "-c", "-o", str(tmp_path / "k.o"),
         os.path.join(CSRC, src)],
        capture_output=True, text=True, check=True)
    kernels = {}
    name = None
    for line in out.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            kernels[name] = {}
            continue
        m = re.search(r"(VGPRs|Occupancy \[waves/SIMD\]|LDS Size \[bytes/block\]): (\d+)", line)
        if m and name:
            kernels[name][m.group(1).split()[0]] = int(m.group(2))
    return kernels


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
@pytest.mark.parametrize("src,kernel", [("nsim_gmres.hip", "k_gmres"), ("nsim_cg.hip", "k_cg")])
def test_persistent_kernels_fit_their_occupancy_with_granule_rounded_lds(src, kernel,
                                                                         tmp_path):
    rep = {k: v for k, v in resource_report(src, tmp_path).items() if kernel in k}
    assert rep, "no %s instantiation in the report" % kernel
    for name, r in rep.items():
        lds = r["LDS"]
        occ = r["Occupancy"]
        rounded = -(-lds // GRANULE) * GRANULE
        assert occ >= 1, (name, r)
        # the runtime's residency figure is at most the compiler's occupancy;
        # the grid helper caps it by the rounded LDS, which must not cut it
        assert occ * rounded <= LDS_PER_CU, (name, r, rounded)


def test_granule_constants_match_the_device_header():
    txt = open(os.path.join(CSRC, "nsim_dev_common.h")).read()
    assert re.search(r"kLdsGranule\s*=\s*%d;" % GRANULE, txt)
    assert re.search(r"kLdsPerCu\s*=\s*160\s*\*\s*1024;", txt)


def test_krylov_bases_above_the_kernel_capacity_fall_back_to_host_gmres():
    """kGmresMaxl bounds the LDS arrays; larger krylov_max runs the host-driven
    GMRES (gmres_grid refuses, the integrator's device grid stays 0)."""
    h = open(os.path.join(CSRC, "nsim_gmres.h")).read()
    cap = int(re.search(r"constexpr int kGmresMaxl = (\d+);", h).group(1))
    assert 300 <= cap <= 320  # the reference's default maxl (300) runs on the device
    g = open(os.path.join(CSRC, "nsim_gmres.hip")).read()
    assert "maxl > kGmresMaxl) return false" in g
    assert "grid > w.resident" in g
