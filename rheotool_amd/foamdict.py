"""A reader for OpenFOAM dictionary files (case I/O of the stress step).

Reads the entries the stress-step tools need from `constant/` and `system/`
files and `0/` fields: `key value;`, `key name [dims] value;`, nested
`name { ... }` dictionaries, quoted regex keys ("tau.*|A.*"), lists
`( ... )` of numbers, vectors/tensors and named dictionaries (multiMode's
`models`), `uniform` field values, C/C++ comments, and `#{ ... #}` code
blocks (kept as the single word `#codeBlock`, never run).  The FoamFile header
is read like any other sub-dictionary.

A dictionary is a plain dict: key -> list of tokens (strings, numbers, nested
lists) or key -> dict.  `text` of a sub-dictionary is kept too, so it can be
handed to the C-ABI as OpenFOAM text unchanged.
"""
from __future__ import annotations

import re

# A word that starts with a letter keeps balanced parentheses written without
# spaces (OpenFOAM keywords such as div(phi,theta) or div(etaP,grad(U))).
_WORD = r'[A-Za-z_][^\s{}()\[\];"]*(?:\((?:[^\s(){};"]|\([^\s(){};"]*\))*\)[^\s{}()\[\];"]*)*'
_TOKEN = re.compile(r'"[^"]*"|' + _WORD + r'|[{}()\[\];]|[^\s{}()\[\];"]+')


def _strip_comments(s: str) -> str:
    """Drops comments; a verbatim `#{ ... #}` code block becomes one word."""
    s = re.sub(r"#\{.*?#\}", " #codeBlock ", s, flags=re.S)
    s = re.sub(r"/\*.*?\*/", " ", s, flags=re.S)
    return re.sub(r"//[^\n]*", " ", s)


def _num(t):
    try:
        return float(t)
    except ValueError:
        return t


class FoamDict(dict):
    """dict of entries with the source text of this (sub-)dictionary."""

    def __init__(self, *a, text="", **k):
        super().__init__(*a, **k)
        self.text = text

    def sub(self, key) -> "FoamDict":
        v = self.get(key)
        if not isinstance(v, FoamDict):
            raise KeyError(f"sub-dictionary {key} not found")
        return v

    def scalar(self, key, default=None):
        """The last number of the entry outside [...] (skips "name [dims]")."""
        v = self.get(key)
        if v is None or isinstance(v, FoamDict):
            if default is None:
                raise KeyError(f"keyword {key} is undefined")
            return default
        nums = [x for x in v if isinstance(x, float)]
        if not nums:
            if default is None:
                raise KeyError(f"keyword {key} has no value")
            return default
        return nums[-1]

    def word(self, key, default=None):
        v = self.get(key)
        if not v or isinstance(v, FoamDict):
            if default is None:
                raise KeyError(f"keyword {key} is undefined")
            return default
        return v[0]

    def lookup_regex(self, name):
        """The entry whose key (possibly a quoted regex) matches name."""
        if name in self:
            return self[name]
        for k, v in self.items():
            if k.startswith('"') and re.fullmatch(k.strip('"'), name):
                return v
        return None


def _parse_list(toks, i):
    """toks[i] == '(' -> (list, next index).  Items: numbers, words, nested
    lists, and `name { ... }` dictionaries as (name, FoamDict) pairs."""
    out = []
    i += 1
    while toks[i] != ")":
        t = toks[i]
        if t == "(":
            lst, i = _parse_list(toks, i)
            out.append(lst)
        elif i + 1 < len(toks) and toks[i + 1] == "{":
            d, i = _parse_dict(toks, i + 2, toks[i])
            out.append((t, d))
        else:
            out.append(_num(t))
            i += 1
    return out, i + 1


def _parse_dict(toks, i, name=None):
    """Parses entries until the matching '}' (or the end when name is None)."""
    d = FoamDict()
    start = i
    while i < len(toks):
        t = toks[i]
        if t == "}":
            d.text = " ".join(toks[start:i])
            return d, i + 1
        if t == ";":
            i += 1
            continue
        key = t
        i += 1
        if i < len(toks) and toks[i] == "{":
            sub, i = _parse_dict(toks, i + 1, key)
            d[key] = sub
            continue
        vals = []
        while i < len(toks) and toks[i] not in (";", "}"):
            if toks[i] == "(":
                lst, i = _parse_list(toks, i)
                vals.append(lst)
            elif toks[i] == "[":
                j = toks.index("]", i)
                vals.append(("dims", toks[i + 1:j]))
                i = j + 1
            else:
                vals.append(_num(toks[i]))
                i += 1
        if i < len(toks) and toks[i] == ";":
            i += 1
        d[key] = vals
    if name is not None:
        raise ValueError(f"unbalanced '{{' in dictionary {name}")
    d.text = " ".join(toks[start:])
    return d, i


def parse(text: str) -> FoamDict:
    toks = _TOKEN.findall(_strip_comments(text))
    d, _ = _parse_dict(toks, 0)
    return d


def read(path: str) -> FoamDict:
    with open(path) as fh:
        return parse(fh.read())


def uniform_value(field: FoamDict):
    """The internalField of a `uniform` field file: a float or a tuple."""
    v = field.get("internalField")
    if not v or v[0] != "uniform":
        raise ValueError("only uniform internal fields are read")
    x = v[1]
    return tuple(x) if isinstance(x, list) else x
