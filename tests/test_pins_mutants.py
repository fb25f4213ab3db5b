"""The oracle's pins fail when a plausible mistake is planted in the oracle (runs without a GPU).

Each planted mistake is one `#if VO_MUTANT == n` site of the oracle's sources; `make mutants`
builds every variant as its own library. The pin tests of that part of the oracle run against
each variant (VELO_ORACLE_LIB) and must fail; a crash of the variant counts as a failure too.
"""
import concurrent.futures
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

CLUSTERING = "tests/test_pins_clustering.py"
SEARCH = "tests/test_pins_search.py"
MUTANTS = {
    1: ("estimate: the neighbour scan tests the other word's isolation flag", CLUSTERING),
    2: ("estimate: the neighbour scan filters by sensor", CLUSTERING),
    3: ("estimate: candidate pixel encoded as the bit, not clz - 24", CLUSTERING),
    4: ("clustering: a cluster in the window's lowest row is kept", CLUSTERING),
    5: ("clustering: the west extension also rejects the lowest row", CLUSTERING),
    6: ("clustering: textbook diagonal growth masks", CLUSTERING),
    7: ("clustering: no precedence of the east column", CLUSTERING),
    8: ("clustering: the window loads isolated words", CLUSTERING),
    10: ("seeding: 4 h0 candidates at the first step", SEARCH),
    11: ("seeding: 8 h0 candidates after the first step", SEARCH),
    12: ("seeding: pendulum visits -1 before +1", SEARCH),
    13: ("forwarding: one more skipped module pair allowed", SEARCH),
    14: ("forwarding: used hits are not candidates", SEARCH),
    15: ("seeding: equal scatters go to the last candidate", SEARCH),
    16: ("seeding: phi extrapolated to the other module of the next pair", SEARCH),
    17: ("forwarding: module parity from the last hit", SEARCH),
    18: ("seeding: seeds past 2048 dropped instead of wrapping the ring", SEARCH),
    19: ("lock-step seeding: equal scatters of one round go to the higher lane", SEARCH),
    20: ("pruned seeding: a group with no seed above the threshold keeps none", SEARCH),
    21: ("pruned seeding: groups cut from the module pair's whole list, not per run of positions", SEARCH),
}


@pytest.fixture(scope="module")
def mutant_libs():
    subprocess.run(["make", "-s", "-j8", "-C", ORACLE, "mutants"], check=True, capture_output=True)
    return {m: os.path.join(ORACLE, "build", "mutants", f"libvelo_oracle_m{m}.so") for m in MUTANTS}


def _run(lib, test_file):
    env = dict(os.environ, VELO_ORACLE_LIB=lib)
    p = subprocess.run([sys.executable, "-m", "pytest", test_file, "-x", "-q", "-p", "no:cacheprovider"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    failed = [line.split(" ")[1] for line in p.stdout.splitlines() if line.startswith("FAILED")]
    return p.returncode, failed


def test_every_planted_mistake_fails_a_pin(mutant_libs):
    with concurrent.futures.ThreadPoolExecutor(max_workers=6) as pool:
        runs = {m: pool.submit(_run, mutant_libs[m], f) for m, (_, f) in MUTANTS.items()}
        results = {m: r.result() for m, r in runs.items()}
    survivors = []
    for m, (rc, failed) in sorted(results.items()):
        what = MUTANTS[m][0]
        print(f"mutant {m:2d} ({what}): rc {rc}, first failure {failed[:1] or 'crash'}")
        if rc == 0:
            survivors.append(f"{m}: {what}")
    assert not survivors, "pins that no planted mistake fails: " + "; ".join(survivors)
