"""The batched job schedule (csrc/batch_args.hpp job_args, which drives both
the batch kernel's tick-argument table and a lone job's per-shot launches),
checked on the host by following the data through it rather than by
restating its formulas:

* every step reads, as p^n and p^{n-1}, the buffers that the two previous
  steps of its phase wrote, reads the zero field exactly where the phase has
  no such step, and writes either a buffer it does not read or, in place,
  its own p^{n-1} (the two field buffers alternate; only a held p^{n-1}
  sends the step to the other one); a dry step (the autotuner's timing
  launches on live jobs) updates only a scratch field, in place, never the
  live state;
* paired imaging: every snapshot slot is imaged exactly once, in descending
  order; a paired step images the held slot first, and the held buffer then
  still holds the field of that slot's backward step (no step overwrote it);
* FP32: every snapshot slot is written once in the forward phase by the
  step that produces that field and read once in the backward phase at the
  matching time;
* a step may run a lean build (no snapshot epilogue; the batch engine picks
  it per tick, the plain engine per launch) exactly when it has no snapshot
  slot to write or image: FP32 steps that are not stored, 16-bit backward
  steps that do not image, never a 16-bit forward step;
* 16-bit: the scale of slot k is reduced from the per-wave maxima that step
  it_k wrote (same parity) and the codes are written from the field of step
  it_k, two steps later, with the slice maxima of the same parity; the last
  two stored steps are left to the one bookkeeping tick (or, for the last
  field, imaged straight from the buffer that holds it); a parity buffer is
  never written while a pending reduction still has to read it.
"""
import json
import os
import subprocess

import pytest

from conftest import PKG, REPO

CASES = [(nt, st, fmt) for fmt in (0, 1) for nt in range(1, 10) for st in range(1, 6)] + \
        [(40, 4, 1), (41, 3, 0), (37, 7, 1)]


@pytest.fixture(scope="module")
def schedules(tmp_path_factory):
    exe = tmp_path_factory.mktemp("sched") / "schedule_check"
    subprocess.run(["g++", "-O1", "-std=c++17", "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    "-I", os.path.join(PKG, "csrc"), "-I", os.path.join(REPO, "include"),
                    os.path.join(REPO, "tests", "native", "schedule_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)] + [f"{nt}:{st}:{fmt}" for nt, st, fmt in CASES],
                         capture_output=True, text=True, check=True).stdout
    by_case = {}
    for line in out.splitlines():
        rec = json.loads(line)
        by_case.setdefault(rec["case"], []).append(rec)
    return by_case


@pytest.mark.parametrize("nt,store,fmt", CASES)
def test_schedule_dataflow(schedules, nt, store, fmt):
    steps = schedules[f"{nt}:{store}:{fmt}"]
    s16 = fmt != 0
    assert [r["total"] for r in steps] == list(range(2 * nt + (1 if s16 else 0)))
    book = [r for r in steps if r["kind"] == "book"]
    assert len(book) == (1 if s16 else 0) and (not book or book[0]["total"] == nt)
    fwd = [r for r in steps if r["kind"] == "fwd"]
    bwd = [r for r in steps if r["kind"] == "bwd"]
    assert [r["it"] for r in fwd] == list(range(nt))
    assert [r["it"] for r in bwd] == list(range(nt - 1, -1, -1))
    stored = [it for it in range(nt) if it % store == 0]
    nsnap = (nt - 1) // store + 1
    assert len(stored) == nsnap

    # ---- fields: what each buffer holds, followed through both phases
    # (buffers 0-1 alternate, 3-4 are the hold buffers, 5 is the zero field)
    ZERO = 5
    content = {k: None for k in range(5)}
    for phase, recs in (("F", fwd), ("B", bwd)):
        for s, r in enumerate(recs):
            assert r["zc"] == (s == 0) and r["zp"] == (s <= 1)
            assert (r["cur"] == ZERO) == (s == 0) and (r["prev"] == ZERO) == (s <= 1)
            assert r["next"] not in (ZERO, r["cur"], r["hold"], 2)
            # a dry (timing) step updates only the scratch field 2, in place,
            # and reads the live p^n
            assert r["dry_next"] == r["dry_prev"] == 2 and r["dry_cur"] == r["cur"]
            # in place over p^{n-1} whenever p^{n-1} is a field buffer
            if s >= 2 and r["prev"] in (0, 1):
                assert r["next"] in (r["prev"], 3, 4)
            else:
                assert r["next"] != r["prev"]
            if s >= 1:
                assert content[r["cur"]] == (phase, s - 1)
            if s >= 2:
                assert content[r["prev"]] == (phase, s - 2)
            if r["hold"] >= 0:
                # the held field is that of the backward step of the held slot
                k2 = r["snap_in2"] if not s16 else r["s16in2"]
                assert phase == "B" and k2 >= 0 and content[r["hold"]] == ("B", nt - 1 - k2 * store)
            else:
                assert r["snap_in2"] == r["s16in2"] == r["s16max2"] == -1
            if phase == "B" and s == 0 and r["pending"] >= 0:
                # the last forward field, imaged straight from its buffer
                assert content[r["pending"]] == ("F", nt - 1)
                assert r["next"] != r["pending"]
            content[r["next"]] = (phase, s)
        if phase == "F":
            last_fwd = dict(content)

    for r in fwd + bwd:
        if s16 and r["kind"] == "fwd":
            # a 16-bit snapshot's maxima (its stored step) and codes (two
            # steps later) need the full build; the reduction of the maxima
            # (the step between) runs lean, as every other step does
            assert r["lean"] == (r["wmax"] == -1 and r["qout"] == -1)
            assert r["lean"] == all(r["it"] - d not in stored for d in (0, 2))
        elif s16:
            assert r["lean"] == (r["s16in"] == -1 and r["pending"] == -1)
        else:
            assert r["lean"] == (r["snap_out"] == -1 and r["snap_in"] == -1)
        if r["kind"] == "fwd" and not s16:
            assert r["lean"] == (r["it"] not in stored)

    # backward imaging, in the order the kernels apply it: the held slot of a
    # paired step first, then the step's own
    order = []
    for r in bwd:
        own, held = (r["s16in"], r["s16in2"]) if s16 else (r["snap_in"], r["snap_in2"])
        if s16 and r["pending"] >= 0:
            own = r["it"] // store
        if held >= 0:
            assert own == held - 1
            order.append(held)
        if own >= 0:
            assert r["it"] == own * store
            order.append(own)
    assert order == list(range(nsnap - 1, -1, -1))
    # pairing pays: at most one stored step images alone besides a pending one
    alone = [r for r in bwd if r["it"] in stored and r["hold"] < 0 and
             (r["snap_in"] >= 0 or r["s16in"] >= 0) and r["pending"] < 0]
    assert len(alone) <= 1

    if not s16:
        assert [r["snap_out"] for r in fwd] == [it // store if it in stored else -1 for it in range(nt)]
        assert all(r["wmax"] == r["red_in"] == r["qout"] == r["s16in"] == r["pending"] == -1 for r in fwd + bwd)
        return

    # ---- 16-bit scales and codes
    assert all(r["snap_out"] == r["snap_in"] == -1 for r in fwd + bwd)
    by_it = {r["it"]: r for r in fwd}
    for k, itk in enumerate(stored):
        p = k & 1
        assert by_it[itk]["wmax"] == p and by_it[itk]["wcount"] == p
        if itk + 1 <= nt - 1:
            r = by_it[itk + 1]
            assert r["red_in"] == p and r["red_count"] == p and r["red_out"] == p and r["red_out_count"] == 2 + p
        if itk + 2 <= nt - 1:
            r = by_it[itk + 2]
            assert r["qout"] == k and r["qmax"] == k and r["qparts"] == p and r["qcount"] == 2 + p
            # the codes come from p^{n-1} of that step: the field of step it_k
            # (checked by the dataflow above: prev holds F it_k at step it_k+2)
    # no step reduces or quantises a slot it should not
    reduced = [r["it"] - 1 for r in fwd if r["red_in"] >= 0]
    assert reduced == [it for it in stored if it + 1 <= nt - 1]
    quantised = [(r["it"] - 2) // store for r in fwd if r["qout"] >= 0]
    assert quantised == [k for k, it in enumerate(stored) if it + 2 <= nt - 1]
    # double buffers: a step never writes the parity it still has to read
    for r in fwd:
        if r["wmax"] >= 0 and r["red_in"] >= 0:
            assert r["wmax"] != r["red_in"]
        if r["red_out"] >= 0 and r["qparts"] >= 0:
            assert r["red_out"] != r["qparts"]
    # the bookkeeping tick settles the rest: slot nt-2's codes from its field,
    # still in the buffer that forward step nt-2 wrote (not overwritten by
    # step nt-1), and slot nt-1's slices for backward step 0
    if nt >= 2 and (nt - 2) % store == 0:
        assert any(v == ("F", nt - 2) for v in last_fwd.values())
    # backward: the last forward field is imaged pending, at step 0; every
    # other imaged slot reads its codes and its scale (the held one too)
    for s, r in enumerate(bwd):
        it = r["it"]
        if it in stored and s == 0:
            k = it // store
            assert r["pending"] >= 0 and r["s16in"] == -1 and r["qparts"] == (k & 1) and r["qcount"] == 2 + (k & 1)
        else:
            assert r["pending"] == -1
            assert r["s16max"] == r["s16in"] and r["s16max2"] == r["s16in2"]
