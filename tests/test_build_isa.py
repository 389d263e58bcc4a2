"""CPU test of the build's device-code guard (f2m_amd/build.py lost_spills):
a spill placed in a join block ahead of its exec restore runs with lanes
masked off and loses their values; the scan must flag exactly that."""
from f2m_amd.build import lost_spills

BAD = """
_Zkernel:
	s_and_saveexec_b64 s[0:1], vcc
	s_cbranch_execz .LBB0_3
.LBB0_2:
	v_add_u32_e32 v4, 0x400, v4
	s_andn2_b64 exec, exec, s[4:5]
	s_cbranch_execnz .LBB0_2
.LBB0_3:
	s_mov_b64 s[70:71], 0x1000
	scratch_store_dwordx2 off, v[24:25], off offset:1088 ; 8-byte Folded Spill
	s_or_b64 exec, exec, s[0:1]
	s_barrier
"""

GOOD = BAD.replace("""	scratch_store_dwordx2 off, v[24:25], off offset:1088 ; 8-byte Folded Spill
	s_or_b64 exec, exec, s[0:1]""", """	s_or_b64 exec, exec, s[0:1]
	scratch_store_dwordx2 off, v[24:25], off offset:1088 ; 8-byte Folded Spill""")


def test_lost_spill_is_flagged():
    hits = lost_spills(BAD)
    assert len(hits) == 1 and hits[0][0] == "_Zkernel" and "offset:1088" in hits[0][2]
    assert lost_spills(GOOD) == []
    # a private-array store (no spill comment) before the restore is program
    # semantics, not a spill
    assert lost_spills(BAD.replace(" ; 8-byte Folded Spill", "")) == []


WWM = BAD.replace("""	scratch_store_dwordx2 off, v[24:25], off offset:1088 ; 8-byte Folded Spill
""", """	v_writelane_b32 v123, s16, 7
	s_or_saveexec_b64 s[100:101], -1
	scratch_store_dword off, v123, s33 offset:696 ; 4-byte Folded Spill
	s_mov_b64 exec, s[100:101]
""")


def test_whole_wave_spill_is_not_flagged():
    # the save of the VGPR that holds spilled SGPRs runs with every lane on
    assert WWM != BAD and lost_spills(WWM) == []
    # the same store outside the whole-wave bracket is flagged
    outside = WWM.replace("\ts_or_saveexec_b64 s[100:101], -1\n", "").replace("\ts_mov_b64 exec, s[100:101]\n", "")
    assert len(lost_spills(outside)) == 1
    # a spill after the bracket closes, still before the restore, is flagged
    after = WWM.replace("\ts_mov_b64 exec, s[100:101]\n",
                        "\ts_mov_b64 exec, s[100:101]\n\tscratch_store_dword off, v7, s33 offset:4 ; 4-byte Folded Spill\n")
    assert len(lost_spills(after)) == 1


def test_built_device_code_has_no_lost_spills():
    import glob
    import os
    import pytest
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    files = glob.glob(os.path.join(root, "f2m_amd", "_build", "*.gfx950.s"))
    if not files:
        pytest.skip("library not built here")
    for f in files:
        with open(f) as fh:
            assert lost_spills(fh.read()) == [], f
