"""The .tree and .clu writers (rmx_write_tree, rmx_write_clu) against the
reference's own writers, byte for byte:
- the reference's print_twoLevel_Cluster (OmpRelaxmap.cpp:785-808), run by
  the reference harness on an installed partition (tests/golden/*.writers);
- a whole reference command-line run on one thread: the partition of the
  .clu it wrote (OmpRelaxmap.cpp:236-249) installed on the GPU gives the same
  .clu bytes, and the same .tree member lines and module count (the header's
  codelength is the reference's tracked value after its coarse tunes, equal
  to ours up to its printed 6 digits and its incremental rounding).
"""
import pytest

from conftest import INPUTS, data_path, golden

pytestmark = pytest.mark.gpu

KEYS = ["flow_net", "lfr1k_mu03", "sbm1k"]


def _net(key, part):
    from relaxmap_amd import Network
    net = Network()
    net.load(data_path(INPUTS[key]))
    net.initiate()
    net.set_partition(part)
    return net


def _read(path):
    with open(path, newline="") as f:
        return f.read()


@pytest.mark.parametrize("key", KEYS)
def test_tree_bytes_equal_the_reference_writer(key, tmp_path):
    g = golden(f"{key}.writers")
    with _net(key, g["partition"]) as net:
        net.write_tree(tmp_path / "g.tree")
    got, want = _read(tmp_path / "g.tree"), g["tree"]
    print(f"[gpu] writers {key}: .tree {len(got)} bytes, reference {len(want)} bytes", flush=True)
    assert got == want


@pytest.mark.parametrize("key", KEYS)
def test_files_of_a_reference_run(key, tmp_path):
    cli = golden(f"{key}.writers")["cli_one_thread"]
    assert cli["returncode"] == 0
    lines = cli["clu"].split("\r\n")
    n = int(lines[0].split()[1])
    part = [int(x) - 1 for x in lines[1:1 + n]]
    with _net(key, part) as net:
        net.write_clu(tmp_path / "g.clu")
        net.write_tree(tmp_path / "g.tree")
    assert _read(tmp_path / "g.clu") == cli["clu"]
    got, want = _read(tmp_path / "g.tree").split("\n"), cli["tree"].split("\n")
    assert got[1:] == want[1:]                       # every member line
    hg, hw = got[0].split(), want[0].split()         # "# Code length L in K modules."
    assert hg[:3] == hw[:3] and hg[4:] == hw[4:]     # same module count
    assert abs(float(hg[3]) - float(hw[3])) <= 1e-5 * abs(float(hw[3]))
