"""Checks the gfx950 code of every kernel for workgroup barriers reached with
an LDS operation of the wave still in flight.

The compiler waits for a wave's LDS operations before ``s_barrier`` only where
the barrier also orders other memory; a store to LDS still in flight at the
barrier was seen to land after another wave's load issued past it.  Kernels
that hand LDS data from one wave to the others across a barrier therefore use
``block_sync()`` (device_math.hpp), which waits first.  This module compiles
each source to device assembly with the library's flags and walks the control
flow backwards from every ``s_barrier``: a path that meets an LDS operation
(``ds_*``, the register-only permutes excepted) before an ``s_waitcnt`` with
``lgkmcnt(0)`` or an earlier barrier is reported.

    python -m vira_amd.lds_check
"""
from __future__ import annotations

import collections
import concurrent.futures as cf
import re
import subprocess
import tempfile
from pathlib import Path

from . import build as B

_LABEL = (re.compile(r"^(\.LBB\w+):"), re.compile(r"^; (%bb\.\d+):"))
_BRANCH = re.compile(r"s_c?branch\w*\s+(\.LBB\w+)")


def _kernels(asm: str):
    name, body = None, None
    for line in asm.splitlines():
        m = re.match(r"^(_Z\w+):", line)
        if m:
            name, body = m.group(1), []
            continue
        if body is None:
            continue
        if line.strip().startswith("s_endpgm") or line.startswith(".Lfunc_end"):
            yield name, body
            name, body = None, None
            continue
        body.append(line.rstrip())


def _is_lds(ins: str) -> bool:
    return ins.startswith("ds_") and not ins.startswith(("ds_bpermute", "ds_permute", "ds_swizzle"))


def _waits_lds(ins: str) -> bool:
    return ins.startswith("s_waitcnt") and ("lgkmcnt(0)" in ins or re.fullmatch(r"s_waitcnt\s+0(x0)?", ins) is not None)


def barrier_hazards(lines: list[str]) -> list[tuple[str, str]]:
    """(block of the barrier, LDS instruction pending on some path) per hazard."""
    order, blocks = ["entry"], {"entry": []}
    for ln in lines:
        m = _LABEL[0].match(ln) or _LABEL[1].match(ln)
        if m:
            order.append(m.group(1))
            blocks[m.group(1)] = []
            continue
        ins = ln.split(";")[0].strip()
        if ins and not ins.startswith("."):
            blocks[order[-1]].append(ins)
    preds = collections.defaultdict(set)
    for i, b in enumerate(order):
        ins = blocks[b]
        for x in ins:
            m = _BRANCH.match(x)
            if m:
                preds[m.group(1)].add(b)
        if not (ins and ins[-1].startswith("s_branch")) and i + 1 < len(order):
            preds[order[i + 1]].add(b)
    found = []
    for b in order:
        for j, x in enumerate(blocks[b]):
            if x != "s_barrier":
                continue
            stack, seen, bad = [(b, j - 1)], set(), None
            while stack and bad is None:
                blk, k = stack.pop()
                while k >= 0:
                    y = blocks[blk][k]
                    if _waits_lds(y) or y == "s_barrier":
                        break
                    if _is_lds(y):
                        bad = y
                        break
                    k -= 1
                else:
                    for p in preds[blk]:
                        if p not in seen:
                            seen.add(p)
                            stack.append((p, len(blocks[p]) - 1))
            if bad is not None:
                found.append((b, bad))
    return found


def _device_asm(src: Path, out: Path) -> str:
    cmd = [B.HIPCC, *B.FLAGS, "--cuda-device-only", "-S", str(src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc -S failed for {src.name}:\n{r.stderr}")
    return out.read_text()


def check(sources: list[str] | None = None) -> list[str]:
    """Hazards of the given sources (default: all of the library's), one line each."""
    srcs = [B.CSRC / s for s in (sources or B.SOURCES)]
    report = []
    with tempfile.TemporaryDirectory() as tmp, cf.ThreadPoolExecutor(max_workers=8) as ex:
        futs = {ex.submit(_device_asm, s, Path(tmp) / (s.stem + ".s")): s for s in srcs}
        for fut, src in futs.items():
            for kname, body in _kernels(fut.result()):
                for blk, ins in barrier_hazards(body):
                    report.append(f"{src.name}: {kname}: s_barrier in {blk} with '{ins}' in flight")
    return report


if __name__ == "__main__":
    bad = check()
    print("\n".join(bad) if bad else "no barrier reached with LDS operations in flight")
    raise SystemExit(1 if bad else 0)
