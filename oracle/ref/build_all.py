#!/usr/bin/env python3
"""Build every reference configuration the fixtures use (tests/golden/make_ref_fixtures.py),
in parallel, into oracle/_ref/.  Needs the reference sources and flang (development
container only)."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, os.path.join(REPO, "tests", "golden"))

from make_ref_fixtures import CASES  # noqa: E402


def main():
    specs = {}
    for spec in CASES.values():
        key = spec[:8]
        specs[key] = spec
    cmds = [[os.path.join(HERE, "build_ref.sh")] + [str(x) for x in key] for key in specs]
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        rcs = list(ex.map(lambda c: subprocess.run(c, stdout=subprocess.DEVNULL,
                                                   stderr=subprocess.DEVNULL).returncode, cmds))
    bad = [c for c, r in zip(cmds, rcs) if r]
    if bad:
        raise SystemExit(f"reference build failed for {bad}")
    print(f"built {len(cmds)} reference configurations in oracle/_ref/")


if __name__ == "__main__":
    main()
