"""Variable trees: ``{A, B, {C, D}, {E}}`` -- the variables of a node
followed by independent subtrees (the grammar of the reference's
Autotuning/tuner/vartree.py, parsed here by a small recursive descent)."""
from __future__ import annotations

import re
from dataclasses import dataclass, field
from typing import List


@dataclass
class VarTree:
    vars: List[str] = field(default_factory=list)
    subtrees: List["VarTree"] = field(default_factory=list)

    def flatten(self) -> List[str]:
        out = list(self.vars)
        for st in self.subtrees:
            out += st.flatten()
        return out

    def flatten_children(self) -> List[str]:
        out = []
        for st in self.subtrees:
            out += st.flatten()
        return out

    def __str__(self):
        parts = list(self.vars) + [str(s) for s in self.subtrees]
        return "{" + ", ".join(parts) + "}"


_TOKEN = re.compile(r"\s*(\{|\}|,|[A-Za-z_][A-Za-z0-9_]*)")


def parse_vartree(text: str) -> VarTree:
    toks = []
    pos = 0
    text = text.strip()
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise ValueError(f"bad variable tree near {text[pos:pos + 10]!r}")
        toks.append(m.group(1))
        pos = m.end()
    i = 0

    def node():
        nonlocal i
        if toks[i] != "{":
            raise ValueError("variable tree must start with '{'")
        i += 1
        t = VarTree()
        while True:
            if toks[i] == "}":
                i += 1
                return t
            if toks[i] == "{":
                t.subtrees.append(node())
            else:
                if t.subtrees:
                    raise ValueError("variables must precede subtrees in a node")
                t.vars.append(toks[i])
                i += 1
            if toks[i] == ",":
                i += 1
            elif toks[i] != "}":
                raise ValueError(f"expected ',' or '}}', got {toks[i]!r}")

    tree = node()
    if i != len(toks):
        raise ValueError("trailing text after variable tree")
    names = tree.flatten()
    if len(names) != len(set(names)):
        raise ValueError("a variable appears twice in the tree")
    return tree
