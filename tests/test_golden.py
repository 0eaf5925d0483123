"""Golden fixtures written by the reference's own code (tests/golden/make_golden.py, from
oracle/_ref built out of the reference's graph.cc / convert.cc): the host converters and
the oracle must reproduce them exactly, here and wherever the reference is absent."""
import glob
import json
import os
import subprocess

import numpy as np
import pytest

from gspan_amd import graph as G

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CONVERT = os.path.join(os.path.dirname(GOLD), "..", "gspan_amd", "bin", "convert")


def convert_cases():
    return sorted(glob.glob(os.path.join(GOLD, "convert_*.txt")))


def prune_cases():
    out = []
    for p in glob.glob(os.path.join(GOLD, "prune_*.bin")):
        tag = os.path.basename(p)[len("prune_"):-len(".bin")]
        if tag.isdigit():
            out.append((int(tag), p))
    return sorted(out)


def test_fixtures_present():
    assert len(convert_cases()) == 3 and len(prune_cases()) >= 4
    assert os.path.exists(os.path.join(GOLD, "pattern_graphs.json"))


@pytest.mark.parametrize("txt", convert_cases(), ids=os.path.basename)
def test_host_loader_matches_reference_convert(txt, tmp_path):
    want = open(txt[:-4] + ".bin", "rb").read()
    G.from_txt(txt).save(tmp_path / "py.bin")
    assert (tmp_path / "py.bin").read_bytes() == want
    if os.path.exists(CONVERT):
        subprocess.run([CONVERT, txt, str(tmp_path / "cli.bin")], check=True, capture_output=True)
        assert (tmp_path / "cli.bin").read_bytes() == want


def test_oracle_prune_matches_reference(oracle_mod):
    g = G.CSRGraph.load(os.path.join(GOLD, "prune_in.bin"))
    run = oracle_mod.prune_graph(g)
    for ms, path in prune_cases():
        want = G.CSRGraph.load(path)
        got = run(ms)
        np.testing.assert_array_equal(got.offsets, want.offsets)
        np.testing.assert_array_equal(got.vlabels, want.vlabels)
        np.testing.assert_array_equal(got.edges, want.edges)


def test_oracle_pattern_graph_matches_reference(oracle_mod):
    cases = json.load(open(os.path.join(GOLD, "pattern_graphs.json")))
    assert len(cases) >= 100
    for c in cases:
        off, vl, ed = oracle_mod.pattern_graph(c["code"])
        assert off.tolist() == c["offsets"], c["code"]
        assert vl.tolist() == c["vlabels"], c["code"]
        assert np.asarray(ed).reshape(-1, 3).tolist() == c["edges"], c["code"]
