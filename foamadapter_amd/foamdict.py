"""OpenFOAM ASCII dictionaries (system/fvSolution, fvSchemes, controlDict) as
nested Python dicts: the part of FoamAdapter::convert (src/compatibility/
convert.cpp) that turns a case's dictionaries into NeoN dictionaries.

Entries `key value...;` become strings, ints or floats (a multi-token value a
list of them), `key { ... }` a nested dict; comments and the FoamFile header
are dropped.
"""
from __future__ import annotations

import re
from pathlib import Path

_TOKEN = re.compile(r'"[^"]*"|[{};]|[^\s{};]+')


def _strip_comments(text: str) -> str:
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def _value(tok: str):
    if tok.startswith('"') and tok.endswith('"'):
        return tok[1:-1]
    try:
        return int(tok)
    except ValueError:
        pass
    try:
        return float(tok)
    except ValueError:
        return tok


def parse_foam_dict(text: str) -> dict:
    toks = _TOKEN.findall(_strip_comments(text))
    pos = 0

    def block() -> dict:
        nonlocal pos
        out: dict = {}
        while pos < len(toks):
            key = toks[pos]
            pos += 1
            if key == "}":
                return out
            if key == ";":
                continue
            if key.startswith('"') and key.endswith('"'):
                key = key[1:-1]
            if pos < len(toks) and toks[pos] == "{":
                pos += 1
                out[key] = block()
                continue
            vals = []
            while pos < len(toks) and toks[pos] != ";":
                if toks[pos] in "{}":
                    raise ValueError(f"unexpected {toks[pos]!r} after {key!r}")
                vals.append(_value(toks[pos]))
                pos += 1
            pos += 1  # ';'
            out[key] = vals[0] if len(vals) == 1 else vals
        return out

    d = block()
    d.pop("FoamFile", None)
    return d


def read_foam_dict(path) -> dict:
    return parse_foam_dict(Path(path).read_text())
