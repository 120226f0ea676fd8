"""Bulk-data decks written from the golden fixtures (test infrastructure): the
same model in small-field (8 columns, NASTRAN exponent shorthand, '+'
continuations), large-field (GRID*, '*' continuations) and free-field (comma)
form, with comments, so the deck reader is exercised on every field rule."""
import numpy as np

from jf_2026_05_25_amd.deck import parse_number


def fmt8(x):
    """An 8-column field that reads back as exactly x, or None."""
    if x is None:
        return ""
    if float(x).is_integer() and abs(x) < 1e7:
        s = f"{int(x)}."
        if len(s) <= 8:
            return s
    r = repr(float(x))
    if len(r) <= 8 and parse_number(r) == x:
        return r
    m, e = f"{x:.15e}".split("e")
    e = int(e)
    for digits in range(0, 8):
        mant = f"{float(m):.{digits}f}".rstrip("0")
        s = f"{mant}{e:+d}"
        if len(s) <= 8 and parse_number(s) == x:
            return s
    return None


def card_small(name, fields):
    """Small-field card; None when a value does not fit 8 columns exactly."""
    cells = []
    for f in fields:
        if isinstance(f, str):
            cells.append(f)
        else:
            s = fmt8(f)
            if s is None:
                return None
            cells.append(s)
    lines, k, cont = [], 0, 0
    head = f"{name:<8}"
    while True:
        chunk = cells[k:k + 8]
        lines.append(head + "".join(f"{c:>8}" for c in chunk))
        k += 8
        if k >= len(cells):
            break
        cont += 1
        head = f"+C{cont:<6}"
    return "\n".join(lines)


def card_free(name, fields):
    return name + "," + ",".join(f if isinstance(f, str) else repr(float(f)) for f in fields)


def card_large(name, fields):
    cells = [f if isinstance(f, str) else repr(float(f)) for f in fields]
    if any(len(c) > 16 for c in cells):
        return card_free(name, fields)
    lines, k = [], 0
    head = f"{name + '*':<8}"
    while True:
        chunk = cells[k:k + 4]
        lines.append(head + "".join(f"{c:>16}" for c in chunk))
        k += 4
        if k >= len(cells):
            break
        head = f"{'*':<8}"
    return "\n".join(lines)


def fixture_deck(case, style):
    """The deck of a golden case in one style ('small' falls back to free field
    per card when a value needs more than 8 columns)."""
    write = {"small": lambda n, f: card_small(n, f) or card_free(n, f), "free": card_free,
             "large": card_large}[style]
    out = ["$ written from the golden fixture " + case["case"], "SOL 101", "CEND",
           "TITLE = golden deck", "SPC = 1"]
    for s in case["subcases"]:
        out += [f"SUBCASE {s['subcase']}", f"  LABEL = load set {s['load_sid']}", f"  LOAD = {s['load_sid']}"]
    out.append("BEGIN BULK")
    out.append(write("PARAM", ["K6ROT", float(case["k6rot"])]))
    for gid in sorted(int(k) for k in case["grids"]):
        x, y, z = case["grids"][str(gid)]
        out.append(write("GRID", [str(gid), "", x, y, z]))
    for q in case["cquad4"]:
        out.append(write("CQUAD4", [str(q["eid"]), str(q["pid"])] + [str(n) for n in q["nodes"]]))
    for pid, ps in case["pshell"].items():
        out.append(write("PSHELL", [pid, str(ps["mid1"]), ps["t"], str(ps["mid2"]) if ps["mid2"] else "",
                                    ps["bend_ratio"] if ps["mid2"] else "", str(ps["mid3"]) if ps["mid3"] else "",
                                    ps["ts_t"]] + ([ps["nsm"]] if ps.get("nsm") else [])))
    for mid, mt in case["mat1"].items():
        out.append(write("MAT1", [mid, mt["E"], mt["G"] if mt["G"] else "", mt["nu"]] +
                         ([mt["rho"]] if mt.get("rho") else [])))
        out.append("$ inline comments are stripped $ like this one")
    comps = {}
    for s in case["spc1"]:
        comps.setdefault("".join(str(c) for c in s["comps"]), []).append(s["grid"])
    for c, grids in comps.items():
        grids = sorted(grids)
        # a contiguous run as THRU, the rest listed
        if len(grids) > 2 and grids == list(range(grids[0], grids[-1] + 1)):
            out.append(write("SPC1", ["1", c, str(grids[0]), "THRU", str(grids[-1])]))
        else:
            out.append(write("SPC1", ["1", c] + [str(g) for g in grids]))
    for s in case["subcases"]:
        for f in s["force"]:
            out.append(write("FORCE", [str(s["load_sid"]), str(f["grid"]), "", f["mag"]] + list(f["dir"])))
        for p in s["pload4"]:
            out.append(write("PLOAD4", [str(s["load_sid"]), str(p["eid"]), p["p"]]))
        for gr in s.get("grav", []):
            out.append(write("GRAV", [str(s["load_sid"]), "0", gr["a"]] + list(gr["n"])))
    out.append("ENDDATA")
    return "\n".join(out) + "\n"
