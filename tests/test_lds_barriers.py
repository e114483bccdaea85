"""No workgroup barrier of the library's gfx950 code is reached with an LDS
operation of the wave still in flight (CPU: hipcc cross-compiles to device
assembly; see vira_amd/lds_check.py for why it matters).

The checker itself is tested on a hand-written fragment that has such a path
(a store to LDS on a loop's back edge, the shape of the binning kernel's
per-chunk count hand-over before it waited) and on the same fragment with the
wait in place.
"""
import shutil

import pytest

from vira_amd import lds_check as L

_LOOP = """
	s_mov_b32 s4, 0
.LBB0_1:
	s_barrier
	ds_read_b32 v1, v0
	s_waitcnt lgkmcnt(0)
	v_add_u32_e32 v1, 1, v1
	ds_write_b32 v0, v1
{wait}	s_add_i32 s4, s4, 1
	s_cmp_lt_u32 s4, 8
	s_cbranch_scc1 .LBB0_1
; %bb.2:
	s_endpgm
""".splitlines()


def test_checker_flags_store_on_back_edge():
    bad = L.barrier_hazards([ln.replace("{wait}", "") for ln in _LOOP])
    assert len(bad) == 1 and bad[0][1].startswith("ds_write_b32")


def test_checker_accepts_waited_store():
    lines = []
    for ln in _LOOP:
        if "{wait}" in ln:
            lines.append("\ts_waitcnt lgkmcnt(0)")
            ln = ln.replace("{wait}", "")
        lines.append(ln)
    assert L.barrier_hazards(lines) == []


@pytest.mark.skipif(shutil.which(L.B.HIPCC) is None and not L.B.Path(L.B.HIPCC).exists(), reason="no hipcc")
def test_library_barriers():
    bad = L.check()
    assert bad == [], "\n".join(bad)
