"""The autotuner: .conf parsing, variable-tree search against brute force
with a mock evaluator that honours the tree's independence structure (the
reference's own self-test idea, Autotuning/tuner/testing.py:101-200 with
test_evaluations.py:71-175), shell evaluation and failure handling."""
import hashlib
import itertools
import os
import sys

import pytest

from thesis_amd.tuner import (BruteForceOptimiser, FunctionEvaluator, ShellEvaluator, TreeOptimiser,
                              load_settings, parse_vartree, wave64_block_filter)
from thesis_amd.tuner.conf import substitute


def test_parse_vartree():
    t = parse_vartree("{A, B, {C, D}, {E, {F}}}")
    assert t.vars == ["A", "B"] and len(t.subtrees) == 2
    assert t.flatten() == ["A", "B", "C", "D", "E", "F"]
    assert str(t) == "{A, B, {C, D}, {E, {F}}}"
    for bad in ("A, B", "{A, {B}, C}", "{A, A}", "{A B}"):
        with pytest.raises(ValueError):
            parse_vartree(bad)


def _h(*parts):
    return int(hashlib.sha256("|".join(map(str, parts)).encode()).hexdigest()[:8], 16) / 2 ** 32


def independent_fn(tree):
    """score = sum over tree nodes of h(node id, values of the node's vars and
    of its ancestors): subtrees contribute independently given their parents."""
    nodes = []

    def walk(t, anc, path):
        nodes.append((path, anc + t.vars))
        for i, st in enumerate(t.subtrees):
            walk(st, anc + t.vars, path + (i,))

    walk(tree, [], ())

    def f(v):
        return sum(_h(path, *[v[n] for n in names]) for path, names in nodes)

    return f


@pytest.mark.parametrize("spec", ["{A, B}", "{A, {B, C}, {D}}", "{A, {B, {C, D}}, {E, F}}",
                                  "{{A, B}, {C}, {D, E}}"])
@pytest.mark.parametrize("minimise", [True, False])
def test_tree_search_matches_brute_force(spec, minimise):
    tree = parse_vartree(spec)
    values = {n: [str(i) for i in range(3)] for n in tree.flatten()}
    f = independent_fn(tree)
    ev1, ev2 = FunctionEvaluator(f), FunctionEvaluator(f)
    t = TreeOptimiser(tree, values, ev1, minimise=minimise)
    b = BruteForceOptimiser(tree, values, ev2, minimise=minimise)
    rt, rb = t.optimise(), b.optimise()
    assert abs(rt[1] - rb[1]) < 1e-12
    assert len(ev1.order) <= t.tests_required()
    if tree.subtrees:
        assert len(ev1.order) < len(ev2.order)  # the tree search saves tests


def test_failed_tests_are_skipped():
    tree = parse_vartree("{A, B}")
    values = {"A": ["1", "2", "3"], "B": ["1", "2"]}

    def f(v):
        if v["A"] == "1":
            raise RuntimeError("does not run")
        return int(v["A"]) * 10 + int(v["B"])

    ev = FunctionEvaluator(f, repeat=2)
    val, score = TreeOptimiser(tree, values, ev).optimise()
    assert (val["A"], val["B"], score) == ("2", "1", 21)
    assert sum(r.failed for r in ev.order) == 2


def test_wave64_filter():
    assert wave64_block_filter({"BLOCK_X": "64", "BLOCK_Y": "4", "BLOCK_Z": "1"})
    assert not wave64_block_filter({"BLOCK_X": "32", "BLOCK_Y": "1", "BLOCK_Z": "1"})
    assert not wave64_block_filter({"BLOCK_X": "64", "BLOCK_Y": "32", "BLOCK_Z": "1"})
    assert not wave64_block_filter({"BLOCK_X": "8", "BLOCK_Y": "16", "BLOCK_Z": "1"})
    assert wave64_block_filter({"TY": "2"})


def test_conf_and_shell_evaluator(tmp_path):
    conf = tmp_path / "t.conf"
    log = tmp_path / "log.csv"
    conf.write_text(f"""[variables]
variables = {{X, {{Y}}}}
[values]
X = 1, 2, 3
Y = 5, 7
[testing]
test = {sys.executable} -c "import sys; x, y = %X%, %Y%; print('run %%ID%%'); sys.exit(1 if x == 3 else 0) if x == 3 else print((x - 2) ** 2 + y)"
[scoring]
repeat = 3, med
optimal = min
[output]
log = {log}
""")
    s = load_settings(str(conf))
    assert s.repeat == 3 and s.overall == "med" and s.custom_fom and s.optimal == "min"
    assert substitute("a %X% b %%ID%%", 4, {"X": "9"}) == "a 9 b 4"
    ev = ShellEvaluator(s)
    val, score = TreeOptimiser(s.vartree, s.values, ev).optimise()
    assert (val, score) == ({"X": "2", "Y": "5"}, 5.0)
    ev.write_csv(str(log), s.vartree.flatten())
    rows = log.read_text().splitlines()
    assert rows[0] == "TestNo,X,Y,Score_1,Score_2,Score_3,Score_Overall"
    assert any(r.split(",")[1] == "3" and r.endswith(",") for r in rows[1:])  # failed test logged


def test_repo_confs_parse():
    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "thesis_amd",
                        "tuner", "confs")
    for name in os.listdir(here):
        s = load_settings(os.path.join(here, name))
        assert s.vartree.flatten()
