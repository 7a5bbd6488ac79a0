"""The command-line driver (patternmatching_binary's CLI) on the GPU writes
the reference's result files: results.totem rows equal the oracle's step
trace, vertex_%04d.totem / graph_%04d.totem hold the oracle's final state."""
import shutil
import subprocess

import numpy as np
import pytest

from conftest import FIXTURES, fixture_paths

pytestmark = pytest.mark.gpu

STEP = {"LCC0": 2, "LCC": 3, "CC": 4, "PC": 5, "TDS": 6, "ENUMERATION": 7, "END": 1}


@pytest.mark.parametrize("name", FIXTURES)
def test_cli_result_files(pj, orc, name, tmp_path):
    f = fixture_paths(name)
    graph = tmp_path / "graph.totem"
    shutil.copy(f["graph"], graph)
    out = tmp_path / "out"
    cli = pj.native.CLI
    res = subprocess.run([str(cli), "-i", str(graph), "-v", f["vertex"], "-p", f["pattern_dir"] + "/", "-o",
                          str(out)], capture_output=True, text=True, timeout=120)
    print(res.stdout[-2000:], res.stderr[-2000:], flush=True)
    assert res.returncode == 0
    assert (tmp_path / "graph.totem.bin").exists()  # binary cache written as the reference does

    lines = (out / "results.totem").read_text().splitlines()
    assert lines[0].startswith("Iteration; Step; Step time")
    rows = [[x.strip() for x in l.split(";")] for l in lines[1:]]
    g = pj.Graph.load(f["graph"], f["vertex"])
    p = pj.Graph.load(f["pattern"], f["pattern_vertex"])
    c = pj.Constraints(p)
    o = orc.Oracle(g, c)
    orows, _ = o.run(orc.COMPLETE)
    assert len(rows) == len(orows) + 1 and rows[-1][1] == "END"
    for r, q in zip(rows, orows):
        assert int(r[0]) == q.iteration and STEP[r[1]] == q.step
        assert [int(x) for x in (r[4], r[5], r[6], r[7], r[8], r[9], r[10], r[11])] == [
            q.vertex_eliminated, q.active_vertices, g.vertex_count, q.edge_eliminated, q.active_edges, g.edge_count,
            q.match_eliminated, q.enumeration]

    cand, vf, ea, *_ = o.state()
    last = len(orows)
    vert = (out / f"vertex_{last:04d}.totem").read_text().split("\n")
    expect_v = [f"{v} " + "".join(f"{b} " for b in range(64) if (int(cand[v]) >> b) & 1)
                for v in range(g.vertex_count) if vf[v] & 1]
    assert vert[: len(expect_v)] == expect_v and all(l == "" for l in vert[len(expect_v):])
    bits = np.unpackbits(ea.view(np.uint8), bitorder="little")
    off, ed, lab = g.offsets(), g.edges(), g.labels()
    expect_g = [f"{v} {int(ed[e])} {int(lab[v])} {int(lab[ed[e]])}" for v in range(g.vertex_count) if vf[v] & 1
                for e in range(off[v], off[v + 1]) if bits[e] and vf[ed[e]] & 1]
    gl = (out / f"graph_{last:04d}.totem").read_text().split("\n")
    assert gl[: len(expect_g)] == expect_g
    assert "Circular constraint number" in (out / "log.totem").read_text()


@pytest.mark.parametrize("name", ["test_3", "test_4"])
def test_cli_second_run_streams_the_binary_cache(pj, name, tmp_path):
    """A second run finds the binary cache the first wrote and streams it
    into HBM (pj_upload_totem_file); its result files equal the first's
    apart from the timing columns."""
    f = fixture_paths(name)
    graph = tmp_path / "graph.totem"
    shutil.copy(f["graph"], graph)
    outs = []
    for run in range(2):
        out = tmp_path / f"out{run}"
        res = subprocess.run([str(pj.native.CLI), "-i", str(graph), "-v", f["vertex"], "-p", f["pattern_dir"] + "/",
                              "-o", str(out)], capture_output=True, text=True, timeout=120)
        assert res.returncode == 0, res.stderr[-2000:]
        assert ("Saved graph as" in res.stdout) == (run == 0)
        outs.append(out)
    strip = lambda l: [x.strip() for i, x in enumerate(l.split(";")) if i not in (2, 3)]
    a = [strip(l) for l in (outs[0] / "results.totem").read_text().splitlines()]
    b = [strip(l) for l in (outs[1] / "results.totem").read_text().splitlines()]
    assert a == b
    for dump in sorted(p.name for p in outs[0].iterdir() if p.name.startswith(("graph_", "vertex_"))):
        assert (outs[0] / dump).read_text() == (outs[1] / dump).read_text(), dump
