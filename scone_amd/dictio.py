"""SCONE dictionary input format: reader and writer.

The format (docs/Dictionary Input.rst; DataStructures/dictParser_func.f90) is a tree of
``key value;``, ``key (v1 v2 ...);`` and ``key { ... }`` entries with ``!`` and ``//`` line
comments.  Numbers become ``int``/``float``; everything else stays a string.  Lists stay lists
even with one element, as ``dictionary%get`` on an array key expects.
"""
from __future__ import annotations

import re

_TOKEN = re.compile(r"\s*(?:([{}();])|([^\s{}();]+))")


def _strip_comments(text: str) -> str:
    out = []
    for line in text.splitlines():
        cut = len(line)
        for mark in ("!", "//"):
            i = line.find(mark)
            if i >= 0:
                cut = min(cut, i)
        out.append(line[:cut])
    return "\n".join(out)


def _convert(tok: str):
    try:
        return int(tok)
    except ValueError:
        pass
    try:
        return float(tok)
    except ValueError:
        return tok


def parse(text: str) -> dict:
    """Parse dictionary text into nested Python dicts (insertion order = file order)."""
    toks = []
    pos = 0
    text = _strip_comments(text)
    while True:
        m = _TOKEN.match(text, pos)
        if not m or m.end() == pos:
            break
        pos = m.end()
        toks.append(m.group(1) or m.group(2))
    it = iter(toks)

    def parse_block(end_tok):
        d = {}
        for tok in it:
            if tok == end_tok:
                return d
            if tok in "{}();":
                raise ValueError(f"unexpected token {tok!r}")
            key = tok
            nxt = next(it)
            if nxt == "{":
                d[key] = parse_block("}")
            elif nxt == "(":
                vals = []
                for v in it:
                    if v == ")":
                        break
                    vals.append(_convert(v))
                if next(it) != ";":
                    raise ValueError(f"missing ';' after list {key}")
                d[key] = vals
            else:
                val = _convert(nxt)
                if next(it) != ";":
                    raise ValueError(f"missing ';' after {key}")
                d[key] = val
        if end_tok is not None:
            raise ValueError("unterminated block")
        return d

    return parse_block(None)


def read(path: str) -> dict:
    with open(path) as f:
        return parse(f.read())


def _fmt(v) -> str:
    if isinstance(v, float):
        return repr(v)
    return str(v)


def dumps(d: dict, indent: int = 0) -> str:
    """Write nested dicts back in the dictionary format (floats with repr, i.e. round-trip exact)."""
    pad = "  " * indent
    lines = []
    for k, v in d.items():
        if isinstance(v, dict):
            lines.append(f"{pad}{k} {{")
            lines.append(dumps(v, indent + 1))
            lines.append(f"{pad}}}")
        elif isinstance(v, (list, tuple)):
            lines.append(f"{pad}{k} (" + " ".join(_fmt(x) for x in v) + ");")
        else:
            lines.append(f"{pad}{k} {_fmt(v)};")
    return "\n".join(lines)
