"""Register, scratch and LDS budgets of the built gfx950 kernels, read on the CPU from the code objects inside
libnpde.so (the AMDGPU metadata notes of each kernel): the occupancy the kernels are designed for, kept from
drifting when the code changes.

Wavefronts per SIMD follow from the registers a wavefront holds: the 512 registers of a SIMD lane, in
granules of 8, shared by the resident wavefronts (at most 8).  A kernel that spills to scratch memory
would trade its latency hiding for memory round trips, so none may."""
import os
import re
import struct
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "neural-pde-surrogate_amd" / "lib" / "libnpde.so"
LLVM = Path(os.environ.get("ROCM_PATH", "/opt/rocm")) / "llvm" / "bin"
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"
CODE_OBJECTS = []  # the extracted gfx950 code objects (filled by the kernels fixture)


def waves_per_simd(vgprs: int) -> int:
    return min(8, 512 // (-(-vgprs // 8) * 8))


@pytest.fixture(scope="module")
def kernels(tmp_path_factory):
    """{mangled kernel name: (registers incl. accumulation registers, scratch bytes, LDS bytes)} of every gfx950
    code object bundled in libnpde.so."""
    readelf, objcopy = LLVM / "llvm-readelf", LLVM / "llvm-objcopy"
    if not LIB.exists() or not readelf.exists() or not objcopy.exists():
        pytest.skip("needs the built library and the ROCm LLVM tools")
    d = tmp_path_factory.mktemp("co")
    fat = d / "fatbin"
    subprocess.run([str(objcopy), f"--dump-section=.hip_fatbin={fat}", str(LIB), str(d / "lib.so")], check=True)
    data = fat.read_bytes()
    out = {}
    for start in (m.start() for m in re.finditer(re.escape(MAGIC), data)):
        # bundle header: magic, entry count, then per entry offset, size, triple length, triple
        (count,) = struct.unpack_from("<Q", data, start + len(MAGIC))
        pos = start + len(MAGIC) + 8
        for _ in range(count):
            off, size, tlen = struct.unpack_from("<QQQ", data, pos)
            triple = data[pos + 24:pos + 24 + tlen].decode()
            pos += 24 + tlen
            if "gfx950" not in triple or size == 0:
                continue
            co = d / f"k{len(CODE_OBJECTS)}.co"
            co.write_bytes(data[start + off:start + off + size])
            CODE_OBJECTS.append(co)
            notes = subprocess.run([str(readelf), "--notes", str(co)], capture_output=True, text=True,
                                   check=True).stdout
            for blk in notes.split("\n  - .")[1:]:
                name = re.search(r"\.name:\s+(\S+)", blk)
                vg = re.search(r"\.vgpr_count:\s+(\d+)", blk)
                if not name or not vg:
                    continue
                scratch = int(re.search(r"\.private_segment_fixed_size:\s+(\d+)", blk).group(1))
                lds = int(re.search(r"\.group_segment_fixed_size:\s+(\d+)", blk).group(1))
                out[name.group(1)] = (int(vg.group(1)), scratch, lds)
    if not out:
        pytest.skip("no gfx950 code objects found")
    return out


def stage(kernels, ns, stage_, corr, mode, scheme=1, lim=3):
    """k_stage<scheme, limiter, stage, corr, layout> (layout 0 strips, 1 row groups, 2 seams)."""
    key = f"_ZN4npde{len(ns)}{ns}7k_stageILi{scheme}ELi{lim}ELi{stage_}ELb{int(corr)}ELi{mode}E"
    hits = [v for k, v in kernels.items() if k.startswith(key)]
    assert len(hits) == 1, key
    return hits[0]


def test_no_kernel_spills_to_scratch(kernels):
    assert len(kernels) > 200
    spill = {k: v[1] for k, v in kernels.items() if v[1] != 0}
    assert not spill, spill


def test_stage_kernels_hold_their_designed_occupancy(kernels):
    """The benchmark's exact seam kernels run 3 wavefronts per SIMD (146 and 167 registers: stage 1 and
    stage 2); the 60-column-strip stage 1 keeps 4 (126) and its stage 2 3; the flux-corrected stages at
    least 2."""
    assert waves_per_simd(stage(kernels, "exact", 1, False, 2)[0]) == 3
    assert waves_per_simd(stage(kernels, "exact", 2, False, 2)[0]) == 3
    assert waves_per_simd(stage(kernels, "exact", 1, False, 0)[0]) == 4
    assert waves_per_simd(stage(kernels, "exact", 2, False, 0)[0]) == 3
    for mode in (0, 1, 2):
        for s in (1, 2):
            assert waves_per_simd(stage(kernels, "exact", s, True, mode)[0]) >= 2, (mode, s)
    # the fast-arithmetic build is no larger where the benchmark runs it
    assert stage(kernels, "fast", 2, False, 2)[0] <= stage(kernels, "exact", 2, False, 2)[0]


def test_network_and_layer_kernels_budgets(kernels):
    """The fused network kernel: two wavefronts per SIMD by registers, and two workgroups per CU (160 KiB) by
    its LDS; the layer kernels' LDS leaves room for at least two workgroups per CU."""
    net = [v for k, v in kernels.items() if k.startswith("_ZN4npde10k_flux_net")]
    assert len(net) == 1
    vg, _, lds = net[0]
    assert waves_per_simd(vg) == 2 and 2 * lds <= 160 * 1024
    conv = {k: v for k, v in kernels.items() if k.startswith("_ZN7npde_nn")}
    assert conv and all(2 * v[2] <= 160 * 1024 for v in conv.values())


def test_no_packed_fp32_arithmetic(kernels):
    """Every unit is compiled with the packed-FP32 target feature off (the Makefile's -packed-fp32-ops): no
    kernel holds a packed FP32 add, multiply or FMA; FP32 work issues as single operations."""
    objdump = LLVM / "llvm-objdump"
    if not objdump.exists():
        pytest.skip("no llvm-objdump")
    assert CODE_OBJECTS
    seen = set()
    for co in CODE_OBJECTS:  # one per translation unit (the C-ABI's holds only the static staging array)
        dis = subprocess.run([str(objdump), "-d", "--mcpu=gfx950", str(co)], capture_output=True, text=True,
                             check=True).stdout
        seen |= {op for op in ("v_add_f64", "v_mfma_f32_16x16x32_bf16", "v_fma_f64") if op in dis}
        found = sorted(set(re.findall(r"\bv_pk_(?:add|mul|fma)_f32\b", dis)))
        assert not found, (co.name, found)
    assert seen == {"v_add_f64", "v_mfma_f32_16x16x32_bf16", "v_fma_f64"}  # the kernels were disassembled
