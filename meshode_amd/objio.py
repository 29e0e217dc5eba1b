"""OBJ text I/O with the reference's parsing rules.

read_obj restates Mesh::ReadOBJ (mesh.cc:14-44):
  * lines are read with a 256-byte buffer (``is.getline(buffer, 256)``); a
    line of 256 or more characters stops reading, as the stream's failbit
    ends the loop;
  * the first whitespace token selects the record; ``v`` reads three
    numbers, ``f`` reads three tokens and keeps the digits before the first
    ``/`` of each (``a``, ``a/b``, ``a//b``, ``a/b/c``), 1-based;
  * a missing ``f`` token reuses the previous token (the reference reads
    into the same buffer); a missing ``v`` coordinate is 0;
  * every other record (``vn``, ``vt``, ``#``, ``mtllib`` ...) is ignored.

write_obj restates Mesh::WriteOBJ (mesh.cc:46-64) with the default
ostream formatting of doubles (6 significant digits, ``%g``).
"""
from __future__ import annotations

import numpy as np


def _parse_index(tok: str) -> int:
    idx = 0
    for ch in tok:
        if ch == "/":
            break
        idx = idx * 10 + (ord(ch) - ord("0"))
    return idx - 1


def _float(tok):
    try:
        return float(tok)
    except ValueError:
        return 0.0


def read_obj(path):
    V, F = [], []
    with open(path, "rb") as fh:
        for raw in fh:
            line = raw.rstrip(b"\n")
            if len(line) >= 256:
                break
            toks = line.decode("latin-1").split()
            if not toks:
                continue
            if toks[0] == "v":
                xyz = [_float(t) for t in toks[1:4]] + [0.0] * max(0, 3 - len(toks[1:4]))
                V.append(xyz[:3])
            elif toks[0] == "f":
                buf = toks[0]
                face = []
                for j in range(3):
                    if 1 + j < len(toks):
                        buf = toks[1 + j]
                    face.append(_parse_index(buf))
                F.append(face)
    V = np.asarray(V, dtype=np.float64).reshape(-1, 3)
    F = np.asarray(F, dtype=np.int64).reshape(-1, 3).astype(np.int32)
    return V, F


def _g(x: float) -> str:
    return "%g" % x


def write_obj(path, V, F, scale: float = 1.0, pos=(0.0, 0.0, 0.0)):
    V = np.asarray(V, dtype=np.float64)
    F = np.asarray(F)
    pos = np.asarray(pos, dtype=np.float64)
    out = []
    for v in V:
        w = v * scale + pos
        out.append("v %s %s %s\n" % (_g(w[0]), _g(w[1]), _g(w[2])))
    for f in F:
        out.append("f %d %d %d\n" % (int(f[0]) + 1, int(f[1]) + 1, int(f[2]) + 1))
    with open(path, "w") as fh:
        fh.write("".join(out))
