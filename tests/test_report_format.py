"""Known-answer tests of the reference's output: the pattern lines of report /
print_dfs_code (gspan.cu:67-77), their order (gspan.cu:1142-1168, 1176-1190) and the
timing line of the fsm command (fsm.cu:40).

The expected lines below were derived by hand from the reference's rules for a 4-vertex
graph -- a triangle A-B-C with a pendant vertex D on C:

    vertex labels  A = 7, B = 3, C = 3, D = 9
    edge labels    A-B = 0, A-C = 0, B-C = 4, C-D = 0

Single-edge codes are the CSR entries with label(from) <= label(to) (gspan.cu:421),
ordered by (from_label, edge_label, to_label) (gspan.cu:41-50): (3,0,7) rows B->A, C->A;
(3,0,9) row C->D; (3,4,3) rows B->C, C->B.  MNI support = the smallest number of distinct
images of a pattern vertex (gspan.cu:282-288): (3,4,3) has 2, every other pattern 1.
Children come backward first (ascending v_i, el), then forward with descending v_i, then
ascending el and to_label; a child is reported only if its code is the minimum DFS code.
Worked through for k = 3 (min_support 1):

* (3,0,7) + first forward from v1 = A to the other label-3 vertex: the path 3-7-3.  Its
  children: the backward edge v2 -> v0 with label 4 (the triangle; the rule at
  gspan.cu:436-440 needs 0 < 4), the forward edge v2 -> D, and the forward edge v0 -> D,
  whose code is not minimal (the same path is reached by v2 -> D).
* (3,0,7) + v0 -> D (label 9 passes 7 <= 9 at equal edge labels) and then v0 -> B (edge
  label 4): the star around C.
* (3,0,7) + v0 -> the other label-3 vertex by edge label 4, then v2 -> D: the path
  7-3-3-9 with edge labels 0, 4, 0.
* (3,0,9) + v0 -> B by edge label 4; its child v2 -> A spells the same path 9-3-3-7 from
  the other end, and (3,0,7) < (3,0,9) makes that code non-minimal.
* (3,4,3): its children start with a smaller edge ((3,0,7) or (3,0,9)), so none is
  minimal.

Labels are printed as given (not as ranks), backward edges print -1 for both labels and
forward edges after the first print -1 as from_label (gspan.cu:1146,1161).
"""
import os
import re
import subprocess

import numpy as np
import pytest

from gspan_amd.graph import csr_from_undirected, synthetic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FSM_BIN = os.path.join(ROOT, "gspan_amd", "bin", "fsm")
ORACLE_CLI = os.path.join(ROOT, "oracle", "build", "gspan_oracle_cli")

A, B, C, D = 0, 1, 2, 3
TRIANGLE_PENDANT = dict(
    u=np.array([A, A, B, C]), v=np.array([B, C, C, D]), el=np.array([0, 0, 4, 0]),
    vl=np.array([7, 3, 3, 9]))

EXPECTED_K3 = """\
(0 1 3 0 7): 1
(0 1 3 0 7)(1 2 -1 0 3): 1
(0 1 3 0 7)(1 2 -1 0 3)(2 0 -1 4 -1): 1
(0 1 3 0 7)(1 2 -1 0 3)(2 3 -1 0 9): 1
(0 1 3 0 7)(0 2 -1 0 9): 1
(0 1 3 0 7)(0 2 -1 0 9)(0 3 -1 4 3): 1
(0 1 3 0 7)(0 2 -1 4 3): 1
(0 1 3 0 7)(0 2 -1 4 3)(2 3 -1 0 9): 1
(0 1 3 0 9): 1
(0 1 3 0 9)(0 2 -1 4 3): 1
(0 1 3 4 3): 2
"""


def expected(k: int, min_sup: int) -> str:
    """The lines of EXPECTED_K3 with at most k edges and support >= min_sup, in order
    (a pattern below the threshold is not reported, and neither are its descendants)."""
    out, cut = [], None
    for line in EXPECTED_K3.splitlines():
        code, sup = line.rsplit(": ", 1)
        m = code.count("(")
        if cut is not None and m > cut:
            continue
        cut = None
        if int(sup) < min_sup:
            cut = m
            continue
        if m <= k:
            out.append(line)
    return "".join(l + "\n" for l in out)


def triangle_pendant():
    t = TRIANGLE_PENDANT
    return csr_from_undirected(4, t["u"], t["v"], t["el"], t["vl"])


TIMING = re.compile(r"^load: \d+\.\d{6}s prepare: \d+\.\d{6}s fsm: \d+\.\d{6}s$")


@pytest.mark.parametrize("k,min_sup", [(1, 1), (2, 1), (3, 1), (3, 2), (2, 3)])
def test_oracle_report_literal(oracle_mod, k, min_sup):
    want, _ = oracle_mod.fsm(triangle_pendant(), k, min_sup)
    assert oracle_mod.format_report(want) == expected(k, min_sup)


def test_expected_helper():
    assert expected(1, 1) == "(0 1 3 0 7): 1\n(0 1 3 0 9): 1\n(0 1 3 4 3): 2\n"
    assert expected(3, 2) == "(0 1 3 4 3): 2\n"
    assert expected(2, 3) == ""


def test_oracle_cli_prints_reference_lines(tmp_path):
    """The oracle's command line prints the pattern lines and then the fsm.cu:40 line."""
    oracle_bin = ORACLE_CLI
    if not os.path.exists(oracle_bin):
        pytest.skip("oracle CLI not built")
    p = tmp_path / "g.bin"
    triangle_pendant().save(p)
    r = subprocess.run([oracle_bin, str(p), "1", "3"], capture_output=True, text=True,
                       check=True)
    lines = r.stdout.splitlines()
    assert "".join(l + "\n" for l in lines[:-1]) == EXPECTED_K3
    assert TIMING.match(lines[-1]), lines[-1]


# ---------------------------------------------------------------- on the GPU

@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_gpu_report_literal(mode):
    from gspan_amd.miner import Miner, format_report
    print(f"[gpu-case] report literal mode={mode}", flush=True)
    g = triangle_pendant()
    with Miner(g, mode=mode) as mn:
        for k, ms in ((1, 1), (2, 1), (3, 1), (3, 2), (2, 3)):
            assert format_report(mn.mine(k, ms)) == expected(k, ms), (k, ms)


def _run_fsm(path, min_sup, k, *extra):
    r = subprocess.run([FSM_BIN, str(path), str(min_sup), str(k), *extra], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r.stdout.splitlines()


@pytest.mark.gpu
def test_fsm_cli_literal(tmp_path):
    """The C++ drop-in (gspan_amd/bin/fsm: the reference's main and fsm() over the C-ABI)
    prints the hand-derived lines and the timing line."""
    print("[gpu-case] fsm CLI literal", flush=True)
    p = tmp_path / "g.bin"
    triangle_pendant().save(p)
    for extra in ((), ("--mode=plain",)):
        lines = _run_fsm(p, 1, 3, *extra)
        assert "".join(l + "\n" for l in lines[:-1]) == EXPECTED_K3, extra
        assert TIMING.match(lines[-1]), lines[-1]


@pytest.mark.gpu
def test_fsm_cli_matches_oracle_cli_er(tmp_path):
    """The C++ drop-in against the oracle's command line on the Erdos-Renyi configuration
    (10,000 vertices / 40,000 edges, 5 labels, 1 edge label, min_support 200, k = 4):
    every pattern line equal, in order."""
    if not os.path.exists(ORACLE_CLI):
        pytest.skip("oracle CLI not built")
    print("[gpu-case] fsm CLI vs oracle CLI on ER 10k/40k k=4", flush=True)
    g = synthetic("er", n=10000, m=40000, n_vlabels=5, n_elabels=1, seed=1)
    p = tmp_path / "er.bin"
    g.save(p)
    want = subprocess.run([ORACLE_CLI, str(p), "200", "4"], capture_output=True, text=True,
                          check=True, timeout=600).stdout.splitlines()
    for extra in ((), ("--mode=plain",)):
        got = _run_fsm(p, 200, 4, *extra)
        assert len(got) == len(want) == 4416, extra
        assert got[:-1] == want[:-1], extra
        assert TIMING.match(got[-1]), got[-1]
