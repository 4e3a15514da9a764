"""Fortran namelist reader (the subset the reference's model.namelist files use).

The reference reads its run-time configuration from Fortran namelist groups
(``&model_nl``, ``&domain_nl``, ``&swm_forcing_nl`` ... ``/``), several of which
may repeat (forcing streams, damping terms, tracers, diag tasks).  Groups are
read in file order, as ``read(unit, nml=...)`` does (e.g.
src/swm_forcing_module.f90:107-121, src/swm_damping_module.f90:175-201).

Real values are parsed with Python's correctly rounded ``float``, which is what
gfortran's list-directed namelist read of a ``real(8)`` gives.
"""
from __future__ import annotations

import re
from typing import Any

__all__ = ["parse_namelist", "read_namelist_file", "groups"]

_TOKEN = re.compile(
    r"""\s*(?:
        (?P<str>'(?:[^']|'')*'|"(?:[^"]|"")*")     |
        (?P<logic>\.(?:true|false|t|f)\.)          |
        (?P<num>[+-]?(?:\d+\.?\d*|\.\d+)(?:[eEdD][+-]?\d+)?) |
        (?P<name>[A-Za-z_][A-Za-z0-9_%]*)          |
        (?P<eq>=) | (?P<comma>,) | (?P<end>/)
    )""",
    re.VERBOSE | re.IGNORECASE,
)


def _strip_comment(line: str) -> str:
    out, quote = [], None
    for ch in line:
        if quote:
            out.append(ch)
            if ch == quote:
                quote = None
        elif ch in "'\"":
            quote = ch
            out.append(ch)
        elif ch == "!":
            break
        else:
            out.append(ch)
    return "".join(out)


def _value(tok: re.Match) -> Any:
    if tok.group("str") is not None:
        s = tok.group("str")
        q = s[0]
        return s[1:-1].replace(q + q, q)
    if tok.group("logic") is not None:
        return tok.group("logic").lower() in (".true.", ".t.")
    s = tok.group("num")
    if re.fullmatch(r"[+-]?\d+", s):
        return int(s)
    return float(s.replace("d", "e").replace("D", "e"))


def parse_namelist(text: str) -> list[tuple[str, dict[str, Any]]]:
    """Return the namelist groups of ``text`` as ``[(group, {key: value})]``.

    Keys are lower-cased (Fortran is case-insensitive).  A key given several
    values (``nc_cmode = "a", "b"``) maps to a list.
    """
    body = "\n".join(_strip_comment(line) for line in text.splitlines())
    out: list[tuple[str, dict[str, Any]]] = []
    pos = 0
    while True:
        start = body.find("&", pos)
        if start < 0:
            break
        m = re.match(r"&([A-Za-z_][A-Za-z0-9_]*)", body[start:])
        if not m:
            pos = start + 1
            continue
        group = m.group(1).lower()
        pos = start + m.end()
        values: dict[str, Any] = {}
        key = None
        pending: list[Any] = []
        while pos < len(body):
            tok = _TOKEN.match(body, pos)
            if tok is None:
                if body[pos:].strip() == "":
                    break
                raise ValueError(f"namelist {group}: cannot parse near {body[pos:pos + 30]!r}")
            pos = tok.end()
            if tok.group("end") is not None:
                break
            if tok.group("name") is not None:
                nxt = _TOKEN.match(body, pos)
                if nxt is not None and nxt.group("eq") is not None:
                    if key is not None:
                        values[key] = pending[0] if len(pending) == 1 else pending
                    key, pending = tok.group("name").lower(), []
                    pos = nxt.end()
                    continue
                raise ValueError(f"namelist {group}: unexpected name {tok.group('name')!r}")
            if tok.group("comma") is not None or tok.group("eq") is not None:
                continue
            pending.append(_value(tok))
        if key is not None:
            values[key] = pending[0] if len(pending) == 1 else pending
        out.append((group, values))
    return out


def read_namelist_file(path: str) -> list[tuple[str, dict[str, Any]]]:
    with open(path, "r", encoding="utf-8", errors="replace") as fh:
        return parse_namelist(fh.read())


def groups(nml: list[tuple[str, dict[str, Any]]], name: str) -> list[dict[str, Any]]:
    """All instances of group ``name`` in file order."""
    return [v for g, v in nml if g == name.lower()]
