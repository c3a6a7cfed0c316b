"""Register, scratch and LDS use of the software-pipelined tile kernels, read from the gfx950 code
object of this build (the kernel metadata note of build/tmm_gemm_kernels.o).

The pipelined kernels only pay off at the occupancy their launchers assume, and they were shaped
so that nothing spills: the FP64 and complex FP32 ones at two workgroups (of four waves) per CU,
the FP32 and complex FP64 ones at up to three, the FP64 128 x 256 one at one with its
accumulators in AGPRs. A compiler or source change that breaks this shows here on the CPU.
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OBJ = os.path.join(ROOT, "tiled-mm_amd", "build", "tmm_gemm_kernels.o")
LLVM = "/opt/rocm/llvm/bin"
TOOLS = [os.path.join(LLVM, t) for t in ("llvm-objcopy", "clang-offload-bundler", "llvm-readelf")]

pytestmark = pytest.mark.skipif(not os.path.exists(OBJ) or not all(map(os.path.exists, TOOLS)),
                                reason="kernel object not built or ROCm LLVM tools missing")

REGS_PER_LANE = 512      # unified VGPR + AGPR file per SIMD lane (CDNA3/4)
LDS_PER_CU = 160 * 1024  # gfx950


def _kernels(tmp_path):
    fat, co = tmp_path / "k.fatbin", tmp_path / "k.co"
    subprocess.run([TOOLS[0], f"--dump-section=.hip_fatbin={fat}", OBJ, str(tmp_path / "host.o")], check=True)
    subprocess.run([TOOLS[1], "--unbundle", "--type=o", f"--input={fat}",
                    "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--output={co}"], check=True)
    notes = subprocess.run([TOOLS[2], "--notes", str(co)], check=True, capture_output=True, text=True).stdout
    out = {}
    for block in notes.split("\n  - "):
        fields = dict(re.findall(r"\.(\w+):\s+(\S+)", block))
        if "name" in fields and "vgpr_count" in fields:
            out[fields["name"]] = {k: int(fields[k]) for k in ("vgpr_count", "agpr_count",
                                                              "private_segment_fixed_size",
                                                              "group_segment_fixed_size",
                                                              "max_flat_workgroup_size")}
    return out


def _workgroups_per_cu(k):
    """Workgroups of 256 threads (one wave per SIMD each) resident per CU: the register file
    (vgpr_count is the unified total, allocated in granules of 8) and the LDS bound it."""
    regs = -(-k["vgpr_count"] // 8) * 8
    return min(REGS_PER_LANE // regs, LDS_PER_CU // max(k["group_segment_fixed_size"], 1))


# mangled-name fragment -> (instances, workgroups per CU the launcher relies on)
EXPECTED = {
    "gemm_d_pipe_kernel": (4, 2), "gemm_d_pipe_group_kernel": (4, 2),
    "gemm_c_pipe_kernel": (9, 2), "gemm_c_pipe_group_kernel": (9, 2),
    "gemm_s_pipe_kernel": (4, 3), "gemm_s_pipe_group_kernel": (4, 3),
    "gemm_z_pipe_kernel": (9, 3), "gemm_z_pipe_group_kernel": (9, 3),
    "gemm_d_big_kernel": (4, 1), "gemm_d_big_group_kernel": (4, 1),
}


def test_pipelined_kernels_fit_without_scratch(tmp_path):
    ks = _kernels(tmp_path)
    assert len(ks) > 100
    for frag, (count, wgs) in EXPECTED.items():
        mine = {n: k for n, k in ks.items() if re.search(rf"\d{frag}I", n)}
        assert len(mine) == count, (frag, sorted(mine))
        for name, k in mine.items():
            assert k["private_segment_fixed_size"] == 0, f"{name} spills to scratch"
            assert k["max_flat_workgroup_size"] == 256, name
            assert _workgroups_per_cu(k) >= wgs, (name, k)


def test_wide_fp64_kernel_keeps_accumulators_in_agprs(tmp_path):
    """The 128 x 256 FP64 tile: 32 accumulator blocks of 4 doubles per wave (256 registers) in
    AGPRs, one wave per SIMD, 3 stages of 24 KB of LDS."""
    ks = _kernels(tmp_path)
    big = {n: k for n, k in ks.items() if "gemm_d_big" in n}
    assert big
    for name, k in big.items():
        assert k["agpr_count"] == 256, (name, k)
        assert k["vgpr_count"] <= REGS_PER_LANE, (name, k)
        assert k["group_segment_fixed_size"] == 3 * 24 * 1024, (name, k)
