"""The FP32 op schedule (netrax_amd/csrc/schedule.cpp) checked on the CPU.

tests/schedule_check.cpp builds the grouped op order and the reverse
sweep's flags for a network, then replays the sweep's memory semantics:
the plain form's level runs (ops sharing no child, delivered concurrently:
own adjoint complete, first deliveries) and, for the FP32 order,
every variant once and after its children, OP_GROUP_INDEP only where the
group before is not consumed, first deliveries, early read-backs at least
two groups after the previous delivery, and LDS pass slots (a slot read sees
its own child's complete adjoint, no live slot is overwritten, later
deliveries add into the child's own slot).  Run on the reference's sample
networks and on random networks, for every group width and slot count the
engine can choose."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from netrax_amd.network import parse_newick, random_network

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SAMPLES = os.path.join(HERE, "golden", "sample_networks")
HIP_INCLUDE = "/opt/rocm/include"
PAIRS = [(g, s) for g in (2, 3, 4, 8) for s in (0, 1, 2, 4, 5)]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None or not os.path.isdir(HIP_INCLUDE):
        pytest.skip("g++ or the HIP headers are not available")
    exe = str(tmp_path_factory.mktemp("schedule") / "schedule_check")
    csrc = os.path.join(ROOT, "netrax_amd", "csrc")
    cmd = [gxx, "-std=c++17", "-O1", f"-I{HIP_INCLUDE}", "-D__HIP_PLATFORM_AMD__",
           os.path.join(HERE, "schedule_check.cpp"), os.path.join(csrc, "schedule.cpp"),
           os.path.join(csrc, "dag.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def _run(exe, net):
    head = [net.n_nodes, net.n_tips, net.root, len(net.edge_parent), len(net.retic_first_edge)]
    body = [*head, *net.edge_parent, *net.edge_child, *net.retic_first_edge, *net.retic_second_edge]
    text = " ".join(str(int(x)) for x in body) + "\n" + "\n".join(f"{g} {s}" for g, s in PAIRS) + "\n"
    r = subprocess.run([exe], input=text, capture_output=True, text=True, timeout=60)
    lines = r.stdout.strip().split("\n")
    assert r.returncode == 0 and all(l.startswith("ok") for l in lines), r.stdout
    assert len(lines) == len(PAIRS) + 1  # the plain form's runs, then every pair
    return [dict(zip(l.split()[1::2], map(int, l.split()[2::2]))) for l in lines[1:]]


@pytest.mark.parametrize("name", sorted(f for f in os.listdir(SAMPLES) if f.endswith(".nw")))
def test_schedule_of_reference_sample_networks(checker, name):
    with open(os.path.join(SAMPLES, name)) as f:
        net = parse_newick(f.read().strip())
    _run(checker, net)


def test_schedule_of_random_networks(checker):
    rng = np.random.default_rng(2024)
    slot_deliveries = 0
    for _ in range(40):
        net = random_network(int(rng.integers(3, 90)), int(rng.integers(0, 7)), rng)
        res = _run(checker, net)
        slot_deliveries += sum(r["slot_deliveries"] for r in res)
    assert slot_deliveries > 0  # the slot paths were exercised
