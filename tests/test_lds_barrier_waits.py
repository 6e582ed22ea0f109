"""Every workgroup barrier in the built kernels is reached with its LDS writes completed.

Static check of the gfx950 disassembly of the objects the build wrote (see lds_barrier_check);
no GPU is needed. The synthetic listings pin the checker itself: a write that reaches a barrier
over a loop's back edge without a wait is found, and a wait on every path clears it.
"""
import glob
import os

import pytest

from tests import lds_barrier_check as lbc

BUILD = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "soap_turbo_amd", "_build")

# listings in llvm-objdump's layout: a loop whose body ends with an LDS write and whose header
# is a barrier; the second adds the wait before the back edge
_LOOP = """
0000000000001000 <k>:
	s_waitcnt lgkmcnt(0)                                       // 000000001000:
	s_barrier                                                  // 000000001004:
	ds_read_b32 v1, v3                                         // 000000001008:
	s_waitcnt lgkmcnt(0)                                       // 00000000100C:
	v_cmp_gt_i32_e32 vcc, 0, v1                                // 000000001010:
	s_cbranch_vccnz 6                                          // 000000001014: <k+0x30>
	ds_write_b32 v3, v2                                        // 000000001018:
	{wait}                                                     // 00000000101C:
	s_branch 65528                                             // 000000001020: <k+0x4>
	s_nop 0                                                    // 000000001024:
	s_nop 0                                                    // 000000001028:
	s_nop 0                                                    // 00000000102C:
	s_endpgm                                                   // 000000001030:
"""


def _check_listing(text):
    (name, base, insns), = lbc.functions(text)
    return lbc.pending_barriers(insns, base)


def test_checker_finds_a_write_reaching_a_barrier_over_a_back_edge():
    assert _check_listing(_LOOP.format(wait="s_nop 0")) == [0x1004]


def test_checker_accepts_a_wait_before_the_back_edge():
    assert _check_listing(_LOOP.format(wait="s_waitcnt lgkmcnt(0)")) == []


def test_cross_lane_ds_ops_are_not_writes():
    text = _LOOP.format(wait="s_nop 0").replace("ds_write_b32 v3, v2", "ds_bpermute_b32 v3, v2, v1")
    assert _check_listing(text) == []


def test_built_kernels_wait_for_lds_writes_before_barriers():
    objs = sorted(glob.glob(os.path.join(BUILD, "*.hip.o")))
    if not objs:
        pytest.skip("no built HIP objects (run the build first)")
    if not os.path.exists(os.path.join(lbc.LLVM, "llvm-objdump")):
        pytest.skip("ROCm LLVM tools not found")
    found = lbc.check_objects(objs)
    assert found == {}, f"barriers reachable with LDS writes in flight: {found}"
