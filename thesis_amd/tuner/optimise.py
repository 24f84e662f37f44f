"""Variable-tree search (Autotuning/tuner/optimisation.py:134-257) and the
exhaustive search it is checked against (optimisation_bf.py:136-165)."""
from __future__ import annotations

import itertools
from typing import Callable, Dict, List, Optional

from .vartree import VarTree


def crossproduct(vt: VarTree, values) -> List[Dict[str, str]]:
    names = vt.vars
    return [dict(zip(names, combo)) for combo in itertools.product(*[values[v] for v in names])]


def wave64_block_filter(valuation: Dict[str, str]) -> bool:
    """The thesis' block-shape filter, re-targeted to 64-lane waves: with
    BLOCK_X/Y/Z present, the workgroup must hold whole waves and at most
    1024 threads, and keep BLOCK_X >= BLOCK_Y >= BLOCK_Z and BLOCK_X >= 8
    (the reference's ordering rule).  Valuations without block variables
    pass."""
    if not all(k in valuation for k in ("BLOCK_X", "BLOCK_Y", "BLOCK_Z")):
        return True
    bx, by, bz = (int(valuation[k]) for k in ("BLOCK_X", "BLOCK_Y", "BLOCK_Z"))
    t = bx * by * bz
    return t % 64 == 0 and 64 <= t <= 1024 and bx >= by >= bz and bx >= 8


class TreeOptimiser:
    def __init__(self, vartree: VarTree, values: Dict[str, List[str]], evaluator,
                 minimise: bool = True, accept: Optional[Callable[[Dict[str, str]], bool]] = None):
        self.tree = vartree
        self.values = values
        self.ev = evaluator
        self.best = min if minimise else max
        self.accept = accept or (lambda v: True)

    def tests_required(self, vt: Optional[VarTree] = None) -> int:
        vt = vt or self.tree
        top = 1
        for v in vt.vars:
            top *= len(self.values[v])
        if not vt.subtrees:
            return top
        child = [self.tests_required(st) for st in vt.subtrees]
        return top * (sum(child) - (len(vt.subtrees) - 1))

    def optimise(self):
        """Returns (valuation, score) or None when no test succeeded."""
        return self._opt(self.tree, {})

    def _arbitrary(self, names):
        return {v: self.values[v][0] for v in names}

    def _opt(self, vt: VarTree, presets: Dict[str, str]):
        tests = []
        for d in crossproduct(vt, self.values):
            d.update(presets)
            tests.append(d)
        if vt.subtrees:
            candidates = []
            for val in tests:
                val = dict(val)
                val.update(self._arbitrary(vt.flatten_children()))
                ok = True
                for st in vt.subtrees:
                    local = dict(val)
                    for v in st.flatten():
                        del local[v]
                    r = self._opt(st, local)
                    if r is None:
                        ok = False
                        break
                    val, _ = r
                if ok:
                    candidates.append(val)
            candidates = [c for c in candidates if self.ev.score(c) is not None]
            if not candidates:
                return None
            opt = self.best(candidates, key=self.ev.score)
            return opt, self.ev.score(opt)
        runnable = [t for t in tests if self.accept(t)]
        self.ev.evaluate(runnable)
        ok = [t for t in runnable if self.ev.score(t) is not None]
        if not ok:
            return None
        opt = self.best(ok, key=self.ev.score)
        return opt, self.ev.score(opt)


class BruteForceOptimiser:
    def __init__(self, vartree: VarTree, values, evaluator, minimise: bool = True,
                 accept: Optional[Callable] = None):
        self.names = vartree.flatten()
        self.values = values
        self.ev = evaluator
        self.best = min if minimise else max
        self.accept = accept or (lambda v: True)

    def optimise(self):
        tests = [dict(zip(self.names, c)) for c in itertools.product(*[self.values[n] for n in self.names])]
        tests = [t for t in tests if self.accept(t)]
        self.ev.evaluate(tests)
        ok = [t for t in tests if self.ev.score(t) is not None]
        if not ok:
            return None
        opt = self.best(ok, key=self.ev.score)
        return opt, self.ev.score(opt)
