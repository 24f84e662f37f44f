"""Evaluators: score valuations, repeat them, aggregate, and log
(behaviour of Autotuning/tuner/evaluator.py:131-396 and logging.py:15-61)."""
from __future__ import annotations

import csv
import subprocess
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

from .conf import substitute


def _key(valuation: Dict[str, str]):
    return tuple(sorted((k, str(v)) for k, v in valuation.items()))


@dataclass
class TestRecord:
    test_id: int
    valuation: Dict[str, str]
    results: List[float] = field(default_factory=list)
    overall: Optional[float] = None
    failed: bool = False


class _Base:
    def __init__(self, repeat: int = 1, aggregator: Callable = min):
        self.repeat = repeat
        self.aggregator = aggregator
        self.log: Dict[tuple, TestRecord] = {}
        self.order: List[TestRecord] = []

    def clear(self):
        self.log.clear()
        self.order.clear()

    def score(self, valuation) -> Optional[float]:
        rec = self.log.get(_key(valuation))
        return None if rec is None or rec.failed else rec.overall

    def evaluate(self, valuations):
        for v in valuations:
            k = _key(v)
            if k in self.log:
                continue  # each distinct valuation is tested once
            rec = TestRecord(len(self.order) + 1, dict(v))
            self.log[k] = rec
            self.order.append(rec)
            self._run(rec)
            if not rec.failed and rec.results:
                rec.overall = rec.results[0] if len(rec.results) == 1 else self.aggregator(rec.results)
            else:
                rec.failed = True

    def write_csv(self, path: str, variables: List[str]):
        n = max((len(r.results) for r in self.order), default=0)
        with open(path, "w", newline="") as f:
            w = csv.writer(f)
            w.writerow(["TestNo"] + variables + [f"Score_{i + 1}" for i in range(n)] + ["Score_Overall"])
            for r in self.order:
                res = [str(x) for x in r.results] + [""] * (n - len(r.results))
                w.writerow([r.test_id] + [r.valuation.get(v, "") for v in variables] + res
                           + ["" if r.overall is None else str(r.overall)])

    def _run(self, rec: TestRecord):
        raise NotImplementedError


class ShellEvaluator(_Base):
    """Runs the compile/test/clean commands of a .conf file."""

    def __init__(self, settings, cwd: Optional[str] = None, verbose: bool = False):
        super().__init__(settings.repeat, settings.aggregator)
        self.s = settings
        self.cwd = cwd
        self.verbose = verbose

    def _sh(self, cmd):
        return subprocess.run(cmd, shell=True, cwd=self.cwd, capture_output=True, text=True)

    def _run(self, rec: TestRecord):
        s = self.s
        if s.compile:
            p = self._sh(substitute(s.compile, rec.test_id, rec.valuation))
            if p.returncode != 0:
                rec.failed = True
                return
        try:
            for _ in range(s.repeat):
                t0 = time.perf_counter()
                p = self._sh(substitute(s.test, rec.test_id, rec.valuation))
                el = time.perf_counter() - t0
                if p.returncode != 0:
                    rec.failed = True
                    return
                if s.custom_fom:
                    lines = [l for l in p.stdout.strip().splitlines() if l.strip()]
                    try:
                        rec.results.append(float(lines[-1]))
                    except (IndexError, ValueError):
                        rec.failed = True
                        return
                else:
                    rec.results.append(el)
                if self.verbose:
                    print(f"test {rec.test_id} {rec.valuation}: {rec.results[-1]}", flush=True)
        finally:
            if s.clean:
                self._sh(substitute(s.clean, rec.test_id, rec.valuation))


class FunctionEvaluator(_Base):
    """Scores a valuation with fn(valuation) -> float; an exception or a
    None result marks the test as failed."""

    def __init__(self, fn: Callable, repeat: int = 1, aggregator: Callable = min):
        super().__init__(repeat, aggregator)
        self.fn = fn

    def _run(self, rec: TestRecord):
        for _ in range(self.repeat):
            try:
                x = self.fn(rec.valuation)
            except Exception:
                rec.failed = True
                return
            if x is None:
                rec.failed = True
                return
            rec.results.append(float(x))
