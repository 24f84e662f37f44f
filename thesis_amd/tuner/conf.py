"""Reading tuner .conf files (format of Autotuning/tuner/tune_conf.py:18-209)."""
from __future__ import annotations

import configparser
import statistics
from dataclasses import dataclass
from typing import Callable, Dict, List, Optional

from .vartree import VarTree, parse_vartree


def _med(xs):
    ys = sorted(xs)
    n = len(ys)
    return ys[n // 2] if n % 2 else (ys[n // 2] + ys[n // 2 - 1]) / 2.0


AGGREGATORS = {"min": min, "max": max, "med": _med, "avg": statistics.fmean}


def substitute(template: str, test_id: int, valuation: Dict[str, str]) -> str:
    s = template.replace("%%ID%%", str(test_id))
    for k, v in valuation.items():
        s = s.replace(f"%{k}%", str(v))
    return s


@dataclass
class Settings:
    vartree: VarTree
    values: Dict[str, List[str]]
    compile: Optional[str]
    test: str
    clean: Optional[str]
    optimal: str            # 'min' or 'max'
    custom_fom: bool        # score from the last stdout line (else wall time)
    repeat: int
    overall: str
    aggregator: Callable
    log: Optional[str]


class ConfError(ValueError):
    pass


def load_settings(path: str) -> Settings:
    cp = configparser.RawConfigParser()
    cp.optionxform = str  # variable names are case sensitive
    if not cp.read(path):
        raise ConfError(f"cannot read {path}")
    for sec in ("variables", "values", "testing", "scoring", "output"):
        if not cp.has_section(sec):
            raise ConfError("config file must have sections [variables], [values], [testing], "
                            "[scoring], [output]")
    if not cp.has_option("variables", "variables"):
        raise ConfError("option 'variables' missing from [variables]")
    tree = parse_vartree(cp.get("variables", "variables"))
    values = {}
    for v in tree.flatten():
        if not cp.has_option("values", v):
            raise ConfError(f"no possible values for {v} in [values]")
        values[v] = [x.strip() for x in cp.get("values", v).split(",")]
    if not cp.has_option("testing", "test"):
        raise ConfError("option 'test' missing from [testing]")
    optimal, custom = "min", False
    if cp.has_option("scoring", "optimal"):
        o = cp.get("scoring", "optimal").strip().lower()
        if o not in ("min", "max", "min_time", "max_time"):
            raise ConfError(f"invalid 'optimal' {o!r}")
        optimal, custom = o[:3], len(o) == 3
    repeat, overall = 1, "min"
    if cp.has_option("scoring", "repeat"):
        head, _, tail = cp.get("scoring", "repeat").partition(",")
        try:
            repeat = int(head)
        except ValueError:
            raise ConfError("invalid 'repeat'")
        if repeat < 1:
            raise ConfError("'repeat' must be at least 1")
        if tail.strip():
            overall = tail.strip().lower()
            if overall not in AGGREGATORS:
                raise ConfError(f"invalid aggregate {overall!r}")
    return Settings(
        vartree=tree, values=values,
        compile=cp.get("testing", "compile") if cp.has_option("testing", "compile") else None,
        test=cp.get("testing", "test"),
        clean=cp.get("testing", "clean") if cp.has_option("testing", "clean") else None,
        optimal=optimal, custom_fom=custom, repeat=repeat, overall=overall,
        aggregator=AGGREGATORS[overall],
        log=cp.get("output", "log") if cp.has_option("output", "log") else None,
    )
