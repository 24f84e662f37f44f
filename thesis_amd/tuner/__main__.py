"""python -m thesis_amd.tuner CONF  -- tune a configuration file (tune.py's
command-line behaviour: print the optimal valuation and its score)."""
from __future__ import annotations

import sys

from . import ShellEvaluator, TreeOptimiser, load_settings, wave64_block_filter


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 1:
        print("usage: python -m thesis_amd.tuner CONFIG_FILE", file=sys.stderr)
        return 1
    s = load_settings(argv[0])
    ev = ShellEvaluator(s, verbose=True)
    opt = TreeOptimiser(s.vartree, s.values, ev, minimise=s.optimal == "min",
                        accept=wave64_block_filter)
    print(f"Variable tree: {s.vartree}; tests required: {opt.tests_required()}", flush=True)
    r = opt.optimise()
    if s.log:
        ev.write_csv(s.log, s.vartree.flatten())
    if r is None:
        print("No valuations were successfully evaluated.")
        return 1
    val, score = r
    print(f"Number of tests run: {len(ev.order)}")
    print(("Minimal" if s.optimal == "min" else "Maximal") + " valuation:")
    for k in s.vartree.flatten():
        print(f"{k} = {val[k]}")
    print(f"{'Minimal' if s.optimal == 'min' else 'Maximal'} Score:")
    print(score)
    return 0


if __name__ == "__main__":
    sys.exit(main())
