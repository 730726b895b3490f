"""Binary DRAT helpers for the proof tests: parsing, the comparison the GPU proof is
held to, and a reverse-unit-propagation (RUP) check of every added line.

Format (the reference's binary mode, src/cpu/proof.cpp:92-135): a line is 'a' or
'd', each literal 2v+sign as a little-endian base-128 varint (bit 7 = more), then 0.
"""
import numpy as np


def parse(b) -> list:
    b = bytes(np.asarray(b, np.uint8))
    out, i, n = [], 0, len(b)
    while i < n:
        tag = chr(b[i])
        assert tag in "ad", f"bad line tag {b[i]} at byte {i}"
        i += 1
        lits = []
        while True:
            r, sh = 0, 0
            while b[i] & 128:
                r |= (b[i] & 127) << sh
                sh += 7
                i += 1
            r |= b[i] << sh
            i += 1
            if r == 0:
                break
            lits.append(r)
        out.append((tag, tuple(lits)))
    return out


def normalized(lines) -> list:
    """Units found by batched propagation come in a device order (the trail is
    compared as a set as well): every maximal run of unit lines is sorted. A proof
    that ends with the empty clause drops the unit run right before it -- which
    units a propagation meets before its contradiction depends on that order."""
    lines = list(lines)
    if lines and lines[-1] == ("a", ()):
        k = len(lines) - 1
        while k > 0 and lines[k - 1][0] == "a" and len(lines[k - 1][1]) == 1:
            k -= 1
        lines = lines[:k] + [("a", ())]
    out, run = [], []
    for ln in lines:
        if ln[0] == "a" and len(ln[1]) == 1:
            run.append(ln)
            continue
        out.extend(sorted(run))
        run = []
        out.append(ln)
    out.extend(sorted(run))
    return out


def fast_normalized(b) -> list:
    """normalized() for large proofs, on line bytes instead of parsed literals: a line
    ends at its zero byte (every varint byte of a literal >= 2 is nonzero); a unit line
    is a tag, one varint and the zero. Runs of unit lines are sorted as byte strings
    (the comparison is of sets, so any fixed order serves)."""
    a = np.asarray(b, np.uint8)
    if not len(a):
        return []
    ends = np.flatnonzero(a == 0)
    starts = np.concatenate([[0], ends[:-1] + 1])
    # varint ends inside a line: bytes below 128 other than the tag and the zero
    term = (a < 128).astype(np.int64)
    cs = np.concatenate([[0], np.cumsum(term)])
    nlits = cs[ends] - cs[starts + 1]          # terminal bytes in (tag, zero)
    raw = a.tobytes()
    lines = [raw[s:e + 1] for s, e in zip(starts.tolist(), ends.tolist())]
    unit = (a[starts] == ord("a")) & (nlits == 1)
    empty = (a[starts] == ord("a")) & (ends == starts + 1)
    n = len(lines)
    if empty[-1]:
        k = n - 1
        while k > 0 and unit[k - 1]:
            k -= 1
        lines = lines[:k] + [lines[-1]]
        unit = np.concatenate([unit[:k], [False]])
    out, run = [], []
    for ln, u in zip(lines, unit.tolist()):
        if u:
            run.append(ln)
            continue
        out.extend(sorted(run))
        run = []
        out.append(ln)
    out.extend(sorted(run))
    return out


def _propagate(clauses, assign):
    """unit propagation to a fixpoint; True on a conflict"""
    changed = True
    while changed:
        changed = False
        for c in clauses:
            free, sat = None, False
            nfree = 0
            for l in c:
                a = assign.get(l >> 1)
                if a is None:
                    nfree += 1
                    free = l
                    if nfree > 1:
                        break
                elif a == ((l & 1) == 0):
                    sat = True
                    break
            if sat or nfree > 1:
                continue
            if nfree == 0:
                return True
            assign[free >> 1] = (free & 1) == 0
            changed = True
    return False


def check_rup(formula, lines):
    """Every added line is implied by unit propagation over the formula, the lines
    added before it and not deleted since. Returns the number of lines checked."""
    active = {}
    for c in formula:
        key = tuple(sorted(c))
        active[key] = active.get(key, 0) + 1
    checked = 0
    for tag, lits in lines:
        key = tuple(sorted(lits))
        if tag == "d":
            if active.get(key, 0):
                active[key] -= 1
            continue
        assign = {}
        conflict = False
        for l in lits:
            v, val = l >> 1, (l & 1) == 1   # assume the literal false
            if assign.get(v, val) != val:
                conflict = True
                break
            assign[v] = val
        clauses = [c for c, k in active.items() for _ in range(k)]
        if not conflict:
            conflict = _propagate(clauses, assign)
        assert conflict, f"line {checked} a {lits} is not implied by unit propagation"
        active[key] = active.get(key, 0) + 1
        checked += 1
    return checked
