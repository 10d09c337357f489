"""A reader for the subset of RON (Rusty Object Notation) that cenote scene
files use (format.rs:1-18): plain RON with explicit ``Some``, externally
tagged enums and no extensions.

Values map to Python as:

* ``(a: 1, b: 2)`` — a struct — to ``dict``; ``(1, 2)`` — a tuple — to ``tuple``;
* ``[..]`` to ``list``; strings, numbers and ``true``/``false`` to themselves;
* ``None`` to ``None`` and ``Some(x)`` to ``Some(x)`` (so ``Some(None)``
  stays distinct from ``None``);
* an enum variant ``Name`` / ``Name(x)`` / ``Name(x, y)`` / ``Name(a: x)``
  to ``Variant(name, value)``, where value is ``None``, ``x``, a tuple, or a
  dict of the struct variant's fields.

Comments (``//`` and ``/* */``) and trailing commas are accepted.
"""

from __future__ import annotations

import dataclasses
import re


class RonError(ValueError):
    """A parse error, with its line and column."""


@dataclasses.dataclass(frozen=True)
class Some:
    value: object


@dataclasses.dataclass(frozen=True)
class Variant:
    name: str
    value: object = None


_TOKEN = re.compile(r"""
    (?P<ws>\s+)
  | (?P<line>//[^\n]*)
  | (?P<block>/\*.*?\*/)
  | (?P<num>[-+]?(?:\d[\d_]*\.?[\d_]*(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?|inf|NaN))
  | (?P<str>"(?:[^"\\]|\\.)*")
  | (?P<ident>[A-Za-z_][A-Za-z0-9_]*)
  | (?P<punct>[()\[\]{}:,])
""", re.VERBOSE | re.DOTALL)

_ESCAPES = {"n": "\n", "t": "\t", "r": "\r", "0": "\0", "\\": "\\", '"': '"', "'": "'"}


def _unescape(body: str) -> str:
    out, i = [], 0
    while i < len(body):
        c = body[i]
        if c != "\\":
            out.append(c)
            i += 1
            continue
        n = body[i + 1]
        if n == "u":
            close = body.index("}", i)
            out.append(chr(int(body[i + 3:close], 16)))
            i = close + 1
        else:
            out.append(_ESCAPES.get(n, n))
            i += 2
    return "".join(out)


def _tokens(text: str):
    pos = 0
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise RonError(f"unexpected character {text[pos]!r} at {_where(text, pos)}")
        kind = m.lastgroup
        if kind not in ("ws", "line", "block"):
            yield kind, m.group(kind), pos
        pos = m.end()
    yield "end", "", pos


def _where(text: str, pos: int) -> str:
    line = text.count("\n", 0, pos) + 1
    col = pos - (text.rfind("\n", 0, pos) + 1) + 1
    return f"line {line}, column {col}"


class _Parser:
    def __init__(self, text: str):
        self.text = text
        self.toks = list(_tokens(text))
        self.i = 0

    def peek(self, k=0):
        return self.toks[min(self.i + k, len(self.toks) - 1)]

    def take(self):
        t = self.toks[self.i]
        self.i += 1
        return t

    def expect(self, value):
        kind, v, pos = self.take()
        if v != value:
            raise RonError(f"expected {value!r}, found {v or 'end of file'!r} at {_where(self.text, pos)}")

    def fail(self, what):
        _, v, pos = self.peek()
        raise RonError(f"expected {what}, found {v or 'end of file'!r} at {_where(self.text, pos)}")

    def value(self):
        kind, v, pos = self.peek()
        if kind == "num":
            self.take()
            s = v.replace("_", "")
            if s in ("inf", "+inf"):
                return float("inf")
            if s == "-inf":
                return float("-inf")
            if s == "NaN":
                return float("nan")
            if re.fullmatch(r"[-+]?\d+", s):
                return int(s)
            return float(s)
        if kind == "str":
            self.take()
            return _unescape(v[1:-1])
        if v == "(":
            return self.paren()
        if v == "[":
            self.take()
            items = []
            while self.peek()[1] != "]":
                items.append(self.value())
                if self.peek()[1] == ",":
                    self.take()
                elif self.peek()[1] != "]":
                    self.fail("',' or ']'")
            self.take()
            return items
        if kind == "ident":
            self.take()
            if v == "true":
                return True
            if v == "false":
                return False
            if v == "None":
                return None
            if v == "Some":
                self.expect("(")
                inner = self.value()
                if self.peek()[1] == ",":
                    self.take()
                self.expect(")")
                return Some(inner)
            if self.peek()[1] == "(":
                payload = self.paren(variant=True)
                return Variant(v, payload)
            return Variant(v)
        self.fail("a value")

    def paren(self, variant=False):
        """A parenthesised group: a struct (named fields), a tuple, or — for a
        variant's payload — a single value."""
        self.expect("(")
        if self.peek()[1] == ")":
            self.take()
            return {} if not variant else ()
        if self.peek()[0] == "ident" and self.peek(1)[1] == ":":
            fields = {}
            while self.peek()[1] != ")":
                kind, name, pos = self.take()
                if kind != "ident":
                    raise RonError(f"expected a field name at {_where(self.text, pos)}")
                self.expect(":")
                if name in fields:
                    raise RonError(f"duplicate field {name!r} at {_where(self.text, pos)}")
                fields[name] = self.value()
                if self.peek()[1] == ",":
                    self.take()
                elif self.peek()[1] != ")":
                    self.fail("',' or ')'")
            self.take()
            return fields
        items = []
        while self.peek()[1] != ")":
            items.append(self.value())
            if self.peek()[1] == ",":
                self.take()
            elif self.peek()[1] != ")":
                self.fail("',' or ')'")
        self.take()
        if variant and len(items) == 1:
            return items[0]
        return tuple(items)


def loads(text: str):
    """Parse one RON value (the whole text)."""
    p = _Parser(text)
    v = p.value()
    if p.peek()[0] != "end":
        p.fail("end of file")
    return v


def _dump_str(text: str) -> str:
    out = []
    for ch in text:
        if ch in ('"', "\\"):
            out.append("\\" + ch)
        elif ch == "\n":
            out.append("\\n")
        elif ch == "\t":
            out.append("\\t")
        elif ch == "\r":
            out.append("\\r")
        elif ch == "\0":
            out.append("\\0")
        else:
            out.append(ch)
    return '"' + "".join(out) + '"'


def _dump_float(x: float) -> str:
    if x != x:
        return "NaN"
    if x in (float("inf"), float("-inf")):
        return "inf" if x > 0 else "-inf"
    text = repr(float(x))
    return text if any(c in text for c in ".eE") else text + ".0"


def dumps(value, indent: int = 0) -> str:
    """One RON value as text that ``loads`` reads back to an equal value
    (structs as named-field groups, variants with their payload, floats in
    their shortest round-tripping form)."""
    pad, inner = "    " * indent, "    " * (indent + 1)
    if value is None:
        return "None"
    if value is True:
        return "true"
    if value is False:
        return "false"
    if isinstance(value, int):
        return str(value)
    if isinstance(value, float):
        return _dump_float(value)
    if isinstance(value, str):
        return _dump_str(value)
    if isinstance(value, Some):
        return f"Some({dumps(value.value, indent)})"
    if isinstance(value, Variant):
        if value.value is None:
            return value.name
        if isinstance(value.value, dict):
            return value.name + dumps(value.value, indent)
        if isinstance(value.value, tuple):
            return value.name + "(" + ", ".join(dumps(v, indent) for v in value.value) + ")"
        return f"{value.name}({dumps(value.value, indent)})"
    if isinstance(value, dict):
        if not value:
            return "()"
        fields = [f"{inner}{k}: {dumps(v, indent + 1)}," for k, v in value.items()]
        return "(\n" + "\n".join(fields) + "\n" + pad + ")"
    if isinstance(value, tuple):
        if len(value) == 1:
            return f"({dumps(value[0], indent)},)"
        return "(" + ", ".join(dumps(v, indent) for v in value) + ")"
    if isinstance(value, list):
        if not value:
            return "[]"
        if all(not isinstance(v, (dict, Variant, Some)) for v in value):
            return "[" + ", ".join(dumps(v, indent) for v in value) + "]"
        return "[\n" + "\n".join(f"{inner}{dumps(v, indent + 1)}," for v in value) + "\n" + pad + "]"
    raise RonError(f"cannot write {type(value).__name__} as RON")

