"""TOTEM formats, vertex files and the synthetic generators."""
import struct

import numpy as np
import pytest

from conftest import FIXTURES, fixture_paths


def parse_text_edges(path):
    lines = [l.split() for l in open(path) if l.strip() and not l.startswith("#")]
    return [(int(a), int(b)) for a, b, *_ in lines]


@pytest.mark.parametrize("name", ["test_1", "test_2", "test_4"])
def test_text_fixture_matches_file(pj, name):
    f = fixture_paths(name)
    g = pj.Graph.load(f["graph"])
    pairs = parse_text_edges(f["graph"])
    off, ed = g.offsets(), g.edges()
    assert g.edge_count == len(pairs)
    got = [(v, int(u)) for v in range(g.vertex_count) for u in ed[off[v]:off[v + 1]]]
    assert got == pairs  # row order = file order
    assert g.labels() is None


def test_binary_fixture_decodes(pj):
    f = fixture_paths("test_3")
    raw = open(f["graph"], "rb").read()
    magic, vid, eid, V, E = struct.unpack_from("<5I", raw, 0)
    assert (magic, vid, eid) == (0x10102048, 4, 4)
    valued = raw[20]
    off = struct.unpack_from(f"<{V + 1}I", raw, 23)
    vals = struct.unpack_from(f"<{V}Q", raw, 23 + 4 * (V + 1))
    ed = struct.unpack_from(f"<{E}I", raw, 23 + 4 * (V + 1) + 8 * V)
    g = pj.Graph.load(f["graph"])
    assert valued == 1
    assert g.vertex_count == V and g.edge_count == E
    assert list(g.offsets()) == list(off)
    assert list(g.edges()) == list(ed)
    assert list(g.labels()) == list(vals)


def test_vertex_file_fills_gaps_and_tail_with_zero(pj, tmp_path):
    gpath = fixture_paths("test_1")["graph"]
    vf = tmp_path / "v.txt"
    vf.write_text("# comment\n1 7\n4 9\n")
    g = pj.Graph.load(gpath, str(vf))
    lab = g.labels()
    assert list(lab[:6]) == [0, 7, 0, 0, 9, 0]
    assert (lab[6:] == 0).all()


@pytest.mark.parametrize("eid_bytes", [4, 8])
def test_binary_round_trip(pj, tmp_path, eid_bytes):
    g = pj.Graph.rmat(8, 4, seed=3).labels_uniform(5, seed=1)
    p = tmp_path / "g.bin"
    g.store_binary(str(p), eid_bytes)
    h = pj.Graph.load(str(p))
    assert np.array_equal(g.offsets(), h.offsets())
    assert np.array_equal(g.edges(), h.edges())
    assert np.array_equal(g.labels(), h.labels())
    raw = open(p, "rb").read()
    assert struct.unpack_from("<3I", raw, 0) == (0x10102048, 4, eid_bytes)


def test_text_round_trip(pj, tmp_path):
    g = pj.Graph.rmat(7, 4, seed=9).labels_uniform(3, seed=2)
    p = tmp_path / "g.txt"
    g.store_text(str(p))
    h = pj.Graph.load(str(p))
    assert np.array_equal(g.offsets(), h.offsets())
    assert np.array_equal(g.edges(), h.edges())
    assert np.array_equal(g.labels(), h.labels())


@pytest.mark.parametrize("content,msg", [
    ("", "empty"),
    ("#Nodes: 3\n#Edges: 1\n", "metadata"),
    ("#Nodes: 3\n#Edges: 2\n#Directed\n1 0\n0 1\n", "sorted"),
    ("#Nodes: 3\n#Edges: 3\n#Directed\n0 1\n1 0\n", "fewer"),
    ("#Nodes: 2\n#Bogus: 2\n#Directed\n", "keyword"),
])
def test_bad_text_files_are_rejected(pj, tmp_path, content, msg):
    p = tmp_path / "bad.txt"
    p.write_text(content)
    with pytest.raises(ValueError, match=msg):
        pj.Graph.load(str(p))


def test_missing_file_is_an_error(pj, tmp_path):
    with pytest.raises(ValueError, match="cannot open"):
        pj.Graph.load(str(tmp_path / "nope"))


def test_truncated_binary_is_an_error(pj, tmp_path):
    g = pj.Graph.rmat(6, 4, seed=1)
    p = tmp_path / "g.bin"
    g.store_binary(str(p), 4)
    data = open(p, "rb").read()
    open(p, "wb").write(data[: len(data) - 7])
    with pytest.raises(ValueError, match="truncated"):
        pj.Graph.load(str(p))



@pytest.mark.parametrize("off,ed,why", [
    ([0, 2, 1, 3], [1, 2, 0], "offsets"),      # offsets decrease
    ([0, 1, 2, 2], [1, 3], "not a vertex"),     # neighbour id 3 with V = 3
    ([1, 2, 3, 3], [1, 2, 0], "offsets"),      # offsets do not start at 0
])
def test_malformed_csr_is_rejected(pj, off, ed, why):
    with pytest.raises(ValueError):
        pj.Graph.from_csr(off, ed, [0, 0, 0])


def test_malformed_binary_file_is_rejected(pj, tmp_path):
    """A binary TOTEM file whose edge list names a vertex beyond the count
    fails to load instead of reaching the device."""
    g = pj.Graph.from_csr([0, 1, 2, 2], [1, 0], [0, 0, 0])
    p = tmp_path / "g.bin"
    g.store_binary(str(p), 4)
    data = bytearray(open(p, "rb").read())
    # last 4 bytes hold the second neighbour id (no labels follow the edges)
    data[-4:] = (7).to_bytes(4, "little")
    open(p, "wb").write(bytes(data))
    with pytest.raises(ValueError, match="not a vertex"):
        pj.Graph.load(str(p))

# --- R-MAT -----------------------------------------------------------------

MASK = (1 << 64) - 1


def splitmix64(x):
    z = (x + 0x9E3779B97F4A7C15) & MASK
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def rmat_python(scale, ef, a, b, c, seed):
    t = [int(a * 2 ** 32), int(a * 2 ** 32) + int(b * 2 ** 32), int(a * 2 ** 32) + int(b * 2 ** 32) + int(c * 2 ** 32)]
    V = 1 << scale
    adj = [set() for _ in range(V)]
    for i in range(ef << scale):
        h = splitmix64(seed ^ splitmix64(i))
        s = d = 0
        for l in range(scale):
            u = splitmix64((h + 0x632BE59BD9B4E019 * (l >> 1)) & MASK)
            r = (u & 0xFFFFFFFF) if l & 1 else (u >> 32)
            bs, bd = (0, 0) if r < t[0] else (0, 1) if r < t[1] else (1, 0) if r < t[2] else (1, 1)
            s, d = (s << 1) | bs, (d << 1) | bd
        if s != d:
            adj[s].add(d)
            adj[d].add(s)
    off = [0]
    ed = []
    for v in range(V):
        ed += sorted(adj[v])
        off.append(len(ed))
    return off, ed


def test_rmat_matches_its_specification(pj):
    g = pj.Graph.rmat(7, 8, 0.57, 0.19, 0.19, seed=42)
    off, ed = rmat_python(7, 8, 0.57, 0.19, 0.19, 42)
    assert list(g.offsets()) == off
    assert list(g.edges()) == ed


def test_rmat_invariants(pj):
    g = pj.Graph.rmat(12, 16, seed=5)
    off, ed = g.offsets(), g.edges()
    assert g.vertex_count == 4096
    assert g.edge_count <= 2 * 16 * 4096
    assert g.asymmetric_slots() == 0
    for v in range(g.vertex_count):
        row = ed[off[v]:off[v + 1]]
        assert (np.diff(row.astype(np.int64)) > 0).all()  # ascending, no duplicates
        assert v not in row  # no self loops
    h = pj.Graph.rmat(12, 16, seed=5)
    assert np.array_equal(h.edges(), ed)
    k = pj.Graph.rmat(12, 16, seed=6)
    assert not np.array_equal(k.offsets(), off)
    # skew: quadrant a = 0.57 concentrates edges on low ids
    deg = np.diff(off)
    assert deg[:64].sum() > deg[-64:].sum() * 10


def test_label_samplers(pj):
    g = pj.Graph.rmat(14, 1, seed=1)
    lab = g.labels_uniform(8, seed=3).labels()
    assert lab.min() >= 0 and lab.max() <= 7
    counts = np.bincount(lab.astype(np.int64), minlength=8)
    assert (abs(counts - len(lab) / 8) < 0.05 * len(lab) / 8 + 50).all()
    for v in range(0, 100):
        w = splitmix64(3 ^ splitmix64(v)) >> 32
        assert lab[v] == (w * 8) >> 32
    z = g.labels_zipf(32, 1.2, seed=4).labels()
    p = np.arange(1, 33) ** -1.2
    p /= p.sum()
    counts = np.bincount(z.astype(np.int64), minlength=32) / len(z)
    assert np.allclose(counts, p, atol=0.01)
    assert z.max() <= 31
