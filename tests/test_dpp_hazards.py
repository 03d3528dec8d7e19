"""The Cholesky panel's hand-written DPP multiply-adds and the wait states
they need.

k_chol_panel issues v_fmac_f64 with a DPP row_newbcast operand as inline asm
(tis_bdhi.hip, fmac_row_bcast).  The compiler does not model an inline-asm
instruction's DPP source, so the hardware rules for DPP reads are checked on
the generated gfx950 code instead: a VALU write of a VGPR needs two wait
states before a DPP read of that VGPR, and a VALU write of EXEC five
(AMD CDNA3 ISA, "Manually inserted wait states").  CPU only: hipcc
cross-compiles the device code to assembly.
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tis_amd", "csrc", "tis_bdhi.hip")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _regs(tok):
    m = re.match(r"v\[(\d+):(\d+)\]", tok)
    if m:
        return set(range(int(m.group(1)), int(m.group(2)) + 1))
    m = re.match(r"v(\d+)$", tok)
    return {int(m.group(1))} if m else set()


def dpp_hazards(asm_lines):
    """(number of DPP instructions, list of violations) for one function's
    instruction lines."""
    ins = [l.strip() for l in asm_lines
           if l.strip() and not l.strip().startswith((";", ".")) and not l.strip().endswith(":")]
    bad, n = [], 0
    for i, line in enumerate(ins):
        op = line.split()[0]
        if "_dpp" not in op:
            continue
        n += 1
        src0 = _regs(line.split()[2].strip(","))
        ws, j = 0, i - 1
        while j >= 0 and ws < 5:
            prev = ins[j]
            pop = prev.split()[0]
            if pop == "s_nop":
                ws += int(prev.split()[1], 0) + 1
                j -= 1
                continue
            if pop.startswith("v_cmpx") and ws < 5:
                bad.append((prev, line))
                break
            if ws < 2 and pop.startswith("v_") and not pop.startswith(("v_readlane", "v_readfirstlane", "v_cmp")):
                dst = prev.split()[1].strip(",") if len(prev.split()) > 1 else ""
                if _regs(dst) & src0:
                    bad.append((prev, line))
                    break
            ws += 1
            j -= 1
    return n, bad


def test_hazard_scan_finds_a_planted_violation():
    ok = ["v_mov_b32_e32 v1, v2", "s_nop 1",
          "v_fmac_f64_dpp v[4:5], v[0:1], v[2:3] row_newbcast:3 row_mask:0xf bank_mask:0xf"]
    assert dpp_hazards(ok) == (1, [])
    planted = ["v_add_f64 v[0:1], v[6:7], v[8:9]",
               "v_fmac_f64_dpp v[4:5], v[0:1], v[2:3] row_newbcast:3 row_mask:0xf bank_mask:0xf"]
    n, bad = dpp_hazards(planted)
    assert n == 1 and len(bad) == 1


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_cholesky_panel_dpp_wait_states(tmp_path):
    out = tmp_path / "bdhi.s"
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "--cuda-device-only", "-S", SRC, "-o", str(out)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-2000:]
    lines = out.read_text().splitlines()
    start = next(i for i, l in enumerate(lines) if l.startswith("_ZN3tis12k_chol_panel"))
    end = next(i for i in range(start, len(lines)) if lines[i].startswith(".Lfunc_end"))
    n, bad = dpp_hazards(lines[start:end])
    # the solve's 2016 multiply-adds and the diagonal block's 1954 (465 + 1024
    # + 465 over its two waves)
    assert n == 2016 + 1954, n
    assert not bad, bad[:3]
