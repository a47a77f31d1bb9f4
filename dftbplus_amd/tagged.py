"""DFTB+'s tagged output files (autotest.tag, results.tag): writer and reader.

The layout is that of ``TTaggedWriter`` (src/dftbp/io/taggedoutput.F90:262-379): a header line
``label:type:rank:dims`` with the label padded or cut to 20 characters and the dimensions
comma-separated (``label:type:0:`` for a scalar), then the values in Fortran (column-major)
order in fixed-width fields:

* real: ``(3E24.15E003)`` -- three per line, 15 significant digits, three-digit exponent;
* complex: ``(1(2E24.15E003))`` -- one (re, im) pair per line;
* integer: ``(2I33)``;
* logical: ``(40L2)``.

The writers of the tags themselves follow ``writeAutotestTag`` and ``writeResultsTag``
(src/dftbp/dftbplus/mainio.F90:1873-2340) for the quantities this backend produces.
"""
from __future__ import annotations

import re
from typing import List, Sequence, Tuple, Union

import numpy as np

LABEL_LEN = 20
_HEADER = re.compile(r"^(.{%d}):(real|integer|logical|complex):(\d+):(.*)$" % LABEL_LEN)

TagValue = Union[float, int, bool, complex, np.ndarray]


def _label(tag: str) -> str:
    return tag[:LABEL_LEN].ljust(LABEL_LEN)


def format_real(x: float) -> str:
    """One ``E24.15E003`` field: sign, ``0.`` and 15 digits, ``E``, sign, 3 exponent digits."""
    x = float(x)
    if x == 0.0:
        body = "0.000000000000000E+000"
        return (("-" if np.signbit(x) else "") + body).rjust(24)
    mant, ex = f"{abs(x):.14e}".split("e")       # d.dddddddddddddd, decimal exponent
    e = int(ex) + 1                                # the mantissa moves behind the point
    body = "0." + mant.replace(".", "") + ("E+" if e >= 0 else "E-") + f"{abs(e):03d}"
    return (("-" if x < 0 else "") + body).rjust(24)


def _lines(kind: str, flat: Sequence) -> List[str]:
    if kind == "real":
        per, fmt = 3, format_real
    elif kind == "complex":
        per, fmt = 1, lambda v: format_real(v.real) + format_real(v.imag)
    elif kind == "integer":
        per, fmt = 2, lambda v: f"{int(v):33d}"
    else:
        per, fmt = 40, lambda v: " T" if v else " F"
    return ["".join(fmt(v) for v in flat[i:i + per]) for i in range(0, len(flat), per)]


def _kind(value) -> str:
    a = np.asarray(value)
    if a.dtype == bool:
        return "logical"
    if np.issubdtype(a.dtype, np.integer):
        return "integer"
    if np.issubdtype(a.dtype, np.complexfloating):
        return "complex"
    return "real"


def format_tag(tag: str, value: TagValue) -> str:
    """Header and value lines of one tag (numpy arrays are written in Fortran order)."""
    a = np.asarray(value)
    kind = _kind(a)
    if a.ndim == 0:
        header = f"{_label(tag)}:{kind}:0:"
    else:
        header = f"{_label(tag)}:{kind}:{a.ndim}:" + ",".join(str(d) for d in a.shape)
    flat = list(a.reshape(-1, order="F")) if a.ndim else [a.item()]
    return "\n".join([header] + _lines(kind, flat)) + "\n"


def write_tags(path: str, items: Sequence[Tuple[str, TagValue]], append: bool = False) -> None:
    with open(path, "a" if append else "w") as fh:
        for tag, value in items:
            fh.write(format_tag(tag, value))


def parse_tags(text: str) -> List[Tuple[str, np.ndarray]]:
    """Tags of a tagged file in order: (label without padding, value array in its shape)."""
    lines = text.split("\n")
    out = []
    i = 0
    while i < len(lines):
        m = _HEADER.match(lines[i])
        if not m:
            if lines[i].strip():
                raise ValueError(f"tagged file line {i + 1}: not a tag header: {lines[i]!r}")
            i += 1
            continue
        label, kind, rank, dims = m.group(1).rstrip(), m.group(2), int(m.group(3)), m.group(4)
        shape = tuple(int(x) for x in dims.split(",")) if rank else ()
        count = int(np.prod(shape)) if shape else 1
        vals: list = []
        i += 1
        while len(vals) < count:
            tok = lines[i].split()
            i += 1
            if kind == "real":
                vals += [float(t) for t in tok]
            elif kind == "integer":
                vals += [int(t) for t in tok]
            elif kind == "logical":
                vals += [t == "T" for t in tok]
            else:
                f = [float(t) for t in tok]
                vals += [complex(f[k], f[k + 1]) for k in range(0, len(f), 2)]
        a = np.array(vals[:count])
        out.append((label, a.reshape(shape, order="F") if shape else a.reshape(())))
    return out


def format_tags(items: Sequence[Tuple[str, TagValue]]) -> str:
    return "".join(format_tag(t, v) for t, v in items)
