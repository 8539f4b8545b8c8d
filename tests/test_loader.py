"""Host-side network files against the oracle's loader (a restatement of
FileIO.cpp:32-265, pinned by the reference's own outputs in
test_oracle.py): the multi-threaded parser must give exactly the sequential
result at any thread count, including the reference's quirks (blank lines
reuse the previous line's weight token, '#' lines in edge lists, missing
weights, duplicates summed in file order, self-loops and w <= 0 dropped),
and the binary CSR must read back identically."""
import numpy as np
import pytest

from conftest import INPUTS, data_path


def _oracle(path):
    from oracle.pyoracle import OracleNet
    o = OracleNet.load(path)
    ptr, idx, _ = o.out_csr()
    return {"n": o.n, "m": o.m, "ptr": ptr, "idx": idx, "w": o.raw_out_w(), "ns": o.n_selflinks,
            "nd": o.n_double_links, "names": [o.name(i) for i in range(o.n)]}


def _ours(path, threads):
    from relaxmap_amd import NetworkFile
    f = NetworkFile(path, threads=threads)
    ptr, idx, w, _ = f.arrays()
    return f, {"n": f.n, "m": f.m, "ptr": ptr, "idx": idx, "w": w, "ns": f.n_self_links, "nd": f.n_double_links,
               "names": [f.name(i) for i in range(f.n)]}


def _same(a, b):
    assert a["n"] == b["n"] and a["m"] == b["m"]
    assert np.array_equal(a["ptr"], b["ptr"]) and np.array_equal(a["idx"], b["idx"])
    assert np.array_equal(a["w"], b["w"])          # bit-exact sums
    assert a["ns"] == b["ns"] and a["nd"] == b["nd"]
    assert a["names"] == b["names"]


@pytest.mark.parametrize("key", list(INPUTS))
@pytest.mark.parametrize("threads", [1, 3, 8])
def test_inputs_match_oracle(key, threads, tmp_path):
    path = data_path(INPUTS[key])
    ref = _oracle(path)
    f, got = _ours(path, threads)
    _same(got, ref)
    binp = tmp_path / "net.rmxcsr"
    f.write(binp)
    _, back = _ours(binp, 0)
    _same(back, ref)


QUIRKY_TXT = """# a comment
1 2 0.5
2 3

3 1 2.0
# another comment
1 2 0.25
4 4 1.0
5 1 -1.0
2 5 3.0
7
6 7 1e-3
6 7 2e-3
"""

QUIRKY_NET = """*Vertices 6
1 "alpha" 2.0
2 "beta gamma"
3 delta 0
4 "eps" -1
5 zeta 3.5
6 "eta"
*Arcs
1 2 0.5
2 3

3 1
1 2 0.25
4 4 1.0
5 1 -1.0
6 2 2.0
6 2 1e-3
"""


@pytest.mark.parametrize("text,ext", [(QUIRKY_TXT, "txt"), (QUIRKY_NET, "net")])
@pytest.mark.parametrize("threads", [1, 2, 5, 13])
def test_quirks_match_oracle_at_every_cut(text, ext, threads, tmp_path):
    # repeated so that thread cuts fall on every kind of line
    body = text if ext == "net" else text * 50
    path = tmp_path / f"quirky.{ext}"
    path.write_text(body)
    ref = _oracle(str(path))
    _, got = _ours(str(path), threads)
    _same(got, ref)


def test_node_weights_and_upload_roundtrip(tmp_path):
    from relaxmap_amd import NetworkFile
    path = tmp_path / "w.net"
    path.write_text(QUIRKY_NET)
    f = NetworkFile(str(path), threads=2)
    _, _, _, nw = f.arrays()
    assert nw.tolist() == [2.0, 1.0, 1.0, 1.0, 3.5, 1.0]   # <= 0 -> 1 (FileIO.cpp:114-119)
    assert [f.name(i) for i in range(6)] == ["alpha", "beta gamma", "delta", "eps", "zeta", "eta"]


def test_bad_files(tmp_path):
    from relaxmap_amd import NetworkFile
    with pytest.raises(IOError):
        NetworkFile(str(tmp_path / "missing.txt"))
    bad = tmp_path / "bad.net"
    bad.write_text("hello\n")
    with pytest.raises(ValueError):
        NetworkFile(str(bad))
    junk = tmp_path / "junk.rmxcsr"
    junk.write_bytes(b"RMXCSR01" + b"\x00" * 3)
    with pytest.raises(ValueError):
        NetworkFile(str(junk))
