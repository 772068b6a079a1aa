"""CPU check of the library's gfx950 machine code: no basic block reads the scalar
condition code (SCC) before writing it.

hipcc for gfx950 (ROCm 7.2) once compiled a select between two constants on a flag
loaded from LDS into an s_cselect that read SCC as left by the predecessor blocks'
loop tests, which meant different things on different paths (the relative splitting
tolerance of the small-matrix route came out negative whenever n % 4 == 0).  A block
whose first SCC reader comes before any SCC writer is where such a value enters; the
library's kernels have none, and this test keeps it so.  Builds the device code to
assembly with the library's flags (about 20 s).
"""
import os
import re
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "mr3smp_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
         "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]

READS = re.compile(r"^\s+(s_cselect_\w+|s_cbranch_scc[01]|s_addc_u32|s_subb_u32|s_cmovk_i32|"
                   r"s_cmov_b\d+)\b")
WRITES = re.compile(r"^\s+(s_cmp\w*|s_bitcmp\w*|s_add_[ui]32|s_sub_[ui]32|s_addc_u32|s_subb_u32|"
                    r"s_and_b\d+|s_or_b\d+|s_xor_b\d+|s_andn[12]_b\d+|s_orn[12]_b\d+|s_nand_b\d+|"
                    r"s_nor_b\d+|s_xnor_b\d+|s_not_b\d+|s_lshl\w*|s_lshr_b\d+|s_ashr_i\d+|s_bfe_\w+|"
                    r"s_abs_i32|s_absdiff_i32|s_min_[iu]32|s_max_[iu]32|s_bcnt\w+|s_quadmask\w*|"
                    r"s_wqm\w*|s_\w+_saveexec_b\d+)\b")


def scc_live_in_blocks(asm):
    """Names of the blocks whose first SCC read precedes every SCC write in them."""
    out, name, state = [], None, None
    for ln in asm.split("\n"):
        m = re.match(r"^(\.LBB\d+_\d+):|^; %bb\.(\d+):|^(\w+):\s*(;|$)", ln)
        if m:
            name, state = m.group(1) or m.group(2) or m.group(3), "open"
            continue
        if state != "open" or not ln.startswith("\t") or ln.strip().startswith(";"):
            continue
        if READS.match(ln):
            out.append(f"{name}: {ln.strip()}")
            state = "done"
        elif WRITES.match(ln):
            state = "done"
    return out


def test_scanner_finds_a_live_in_read():
    """The scanner flags a block that selects on SCC it did not set, and passes one
    that sets it first."""
    bad = ".LBB0_1:\n\ts_mov_b32 s2, 0xbcb00000\n\ts_cselect_b32 s5, s2, 0x3cb00000\n"
    good = ".LBB0_2:\n\ts_cmp_eq_u32 s1, 0\n\ts_cselect_b32 s5, s2, 0x3cb00000\n"
    assert scc_live_in_blocks(bad) == [".LBB0_1: s_cselect_b32 s5, s2, 0x3cb00000"]
    assert scc_live_in_blocks(good) == []


@pytest.mark.parametrize("src", ["mr3_driver.hip", "mr3_dense.hip"])
def test_no_block_reads_scc_before_writing_it(src, tmp_path):
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = str(tmp_path / (src + ".s"))
    subprocess.run([HIPCC, *FLAGS, "--cuda-device-only", "-S", "-o", out,
                    os.path.join(CSRC, src)], check=True, capture_output=True)
    bad = scc_live_in_blocks(open(out).read())
    assert not bad, bad[:5]
