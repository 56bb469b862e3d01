"""The oracle's pins catch plausible mistakes: the oracle built with one planted mistake at a time (ORC_MUTANT, on the
CPU) fails at least one of the reference-held checks of test_oracle_physics.py, while the oracle itself passes them
all. Each mistake is one a restatement could plausibly make: a table entry, a dropped term, a wrong sign, a wrong
constant, an index one level off."""
import os
import subprocess

import pytest

import test_oracle_physics as P
from oracle import binding as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "stormlab_oracle.cpp")
MUTANTS = {1: "Koenig table peak halved", 2: "evaporated charge dropped", 3: "wet growth freezes all",
           4: "sublimation grows the ice", 5: "stability sign in K", 6: "no cloud screening of conduction",
           7: "Takahashi table sign", 8: "L_v heating the deposition below -38 degC", 9: "sedimentation inflow density index"}


@pytest.fixture(scope="module")
def mutants(tmp_path_factory):
    out = tmp_path_factory.mktemp("mutants")
    jobs = {}
    for k in MUTANTS:
        path = str(out / f"liboracle_m{k}.so")
        cmd = ["g++", "-O1", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", f"-DORC_MUTANT={k}", "-shared", "-o", path, SRC]
        jobs[k] = (path, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT))
    libs = {}
    for k, (path, proc) in jobs.items():
        log, _ = proc.communicate(timeout=300)
        assert proc.returncode == 0, log.decode()
        libs[k] = B.load(path)
    return libs


def failing_checks(L):
    failed = []
    for check in P.CHECKS:
        try:
            check(L)
        except AssertionError:
            failed.append(check.__name__)
    return failed


def test_the_oracle_passes_every_check():
    assert failing_checks(B.lib()) == []


@pytest.mark.parametrize("mutant", sorted(MUTANTS), ids=[MUTANTS[k] for k in sorted(MUTANTS)])
def test_a_planted_mistake_fails_a_check(mutants, mutant):
    failed = failing_checks(mutants[mutant])
    print(MUTANTS[mutant], "caught by", failed)
    assert failed, f"mutant {mutant} ({MUTANTS[mutant]}) passes every check"
