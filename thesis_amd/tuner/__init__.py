"""Autotuner: a Python 3 restatement and extension of the reference's
vendored tuner (mrfarstad/thesis Autotuning/tuner, Ben Spencer's Autotuning
System v0.16, BSD-2).

Behaviour kept from the reference:
  * the .conf format: sections [variables] [values] [testing] [scoring]
    [output]; %VAR% and %%ID%% substitution in the compile/test/clean
    commands (Autotuning/tuner/tune_conf.py:18-209);
  * the variable-tree search: the variables of a node are searched
    exhaustively, each subtree is optimised for every valuation of its
    parents with the other subtrees held at their first listed value
    (Autotuning/tuner/optimisation.py:134-257);
  * scoring: `repeat = n, min|med|avg|max`, `optimal = min|max|min_time|max_time`,
    the score of a run is the last stdout line, a non-zero exit is a failed
    test that is skipped (Autotuning/tuner/evaluator.py:131-396);
  * the CSV log  TestNo,<vars>,Score_1..n,Score_Overall
    (Autotuning/tuner/logging.py:15-61).
Extensions:
  * the thesis' block filter (Autotuning/tuner/optimisation.py:149-159) is
    re-targeted to 64-lane waves and asks the library's own knob validation
    (stencil_validate) whether a valuation can run at all;
  * an in-process evaluator scores valuations through a Python callable
    (the batched tile tuner times HIP launches without spawning processes).
"""
from .conf import load_settings, Settings
from .vartree import VarTree, parse_vartree
from .optimise import TreeOptimiser, BruteForceOptimiser, wave64_block_filter
from .evaluate import ShellEvaluator, FunctionEvaluator, TestRecord

__all__ = ["load_settings", "Settings", "VarTree", "parse_vartree", "TreeOptimiser",
           "BruteForceOptimiser", "wave64_block_filter", "ShellEvaluator", "FunctionEvaluator",
           "TestRecord"]
