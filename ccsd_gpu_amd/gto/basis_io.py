"""Basis-set files in NWChem format (the format PySCF's bundled basis sets use).

    parse_nwchem(text)        -> {element: [(l, exponents, coefficient columns), ...]}
    load(name)                -> the parsed table of gto/basis/<name>.nw (cached)
    load_file(path)           -> the parsed table of any NWChem-format file

Format: lines `El  L` open a shell block (L one of S P D F G H I, or SP / L for a
Pople shell that carries an s and a p contraction over the same exponents); the
following numeric lines hold an exponent and one coefficient per column (Fortran
D exponents accepted). A block with several columns is one generally contracted
shell (nctr = columns), as PySCF stores it. Lines starting with '#', `BASIS`,
`END` and blank lines are skipped; `SPHERICAL` / `CARTESIAN` markers are ignored
(the molecule's `cart` flag decides).
"""
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
BASIS_DIR = os.path.join(_HERE, "basis")
_L = {c: i for i, c in enumerate("SPDFGHI")}
_HEAD = re.compile(r"^([A-Z][a-z]?)\s+(SP|L|[SPDFGHI])\s*$", re.IGNORECASE)


def _num(tok):
    return float(tok.replace("D", "E").replace("d", "e"))


def parse_nwchem(text):
    table = {}
    block = None                         # (element, kind, rows)

    def flush():
        if block is None or not block[2]:
            return
        el, kind, rows = block
        exps = [r[0] for r in rows]
        ncol = len(rows[0]) - 1
        if any(len(r) - 1 != ncol for r in rows):
            raise ValueError(f"{el} {kind}: rows with different column counts")
        cols = [[r[1 + c] for r in rows] for c in range(ncol)]
        shells = table.setdefault(el, [])
        if kind in ("SP", "L"):
            if ncol != 2:
                raise ValueError(f"{el} {kind}: an SP shell needs two coefficient columns")
            shells.append((0, exps, [cols[0]]))
            shells.append((1, exps, [cols[1]]))
        else:
            shells.append((_L[kind], exps, cols))

    for raw in text.splitlines():
        line = raw.split("#", 1)[0].strip()
        if not line:
            continue
        up = line.upper()
        if up.startswith("BASIS") or up == "END" or up in ("SPHERICAL", "CARTESIAN"):
            flush()
            block = None
            continue
        m = _HEAD.match(line)
        if m:
            flush()
            el = m.group(1)[0].upper() + m.group(1)[1:].lower()
            block = (el, m.group(2).upper(), [])
            continue
        if block is None:
            raise ValueError(f"numbers outside a shell block: {raw!r}")
        block[2].append([_num(t) for t in line.split()])
    flush()
    return table


_CACHE = {}


def load_file(path):
    path = os.path.abspath(path)
    if path not in _CACHE:
        with open(path) as f:
            _CACHE[path] = parse_nwchem(f.read())
    return _CACHE[path]


def available():
    return sorted(f[:-3] for f in os.listdir(BASIS_DIR) if f.endswith(".nw"))


def load(name):
    key = name.lower().replace("_", "-")
    key = {"ccpvdz": "cc-pvdz", "ccpvtz": "cc-pvtz", "ccpvqz": "cc-pvqz", "sto3g": "sto-3g",
           "def2svp": "def2-svp"}.get(key, key)
    path = os.path.join(BASIS_DIR, key + ".nw")
    if not os.path.exists(path):
        raise KeyError(f"basis set {name!r} is not available (have: {', '.join(available())})")
    return load_file(path)
