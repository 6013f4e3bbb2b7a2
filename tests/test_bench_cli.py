"""CPU checks of bench.py's host logic: the one-GPU contract (--gpus above 1
exits with a message and status 2 before touching a GPU) and the workload
builders of jd_lepic_amd.workloads (particle counts, species loads, wall
labels, multigrid depth)."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench as B  # noqa: E402
from jd_lepic_amd import workloads as WL  # noqa: E402


def test_more_than_one_gpu_is_refused():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 2 and "one GPU" in out.stderr and out.stdout == ""


def loaded(W):
    return sum(k * W["n2"] * int(round((xr - xl) / W["h"])) for s in W["species"] for k, xl, xr in s["loads"])


def test_argon_workloads():
    for name, n, ppc, ng in (("argon-128", 128, 200, 7), ("argon-512", 512, 100, 9), ("argon-1024", 1024, 100, 10),
                             ("argon-2048", 2048, 120, 11)):
        W = WL.CONFIGS[name]()
        assert W["name"] == name and W["ng"] == ng and loaded(W) == 2 * n * n * ppc == sum(WL.load_counts(W))
        assert W["Nm"] == 1e16 * (5e-5) ** 2 / ppc
    W = WL.CONFIGS["argon-512"]()
    assert W["kind"] == "e-Ar" and W["ng"] == 9 and loaded(W) == 2 * 512 * 512 * 100
    assert [s["name"] for s in W["species"]] == ["e", "i"] and len(W["vt0"]) == 3
    bc = W["bcnd"]
    assert bc.shape == (515, 515) and set(np.unique(bc)) == {-1, 1}
    assert np.all(bc[2:513, 2:513] == -1) and np.all(bc[:2] == 1) and np.all(bc[:, 513:] == 1)


def test_ion_source_workload():
    W = WL.CONFIGS["ion-source"]()
    assert (W["n1"], W["n2"], W["ng"], W["ngrid"], W["tag_neg"]) == (1024, 512, 9, 3, 4)
    assert [s["name"] for s in W["species"]] == ["e", "Hp", "H2p", "Hm"] and len(W["vt0"]) == 6
    assert 7.5e7 < loaded(W) < 8.5e7
    bc = W["bcnd"]
    assert bc.shape == (515, 1027) and set(np.unique(bc)) == {-1, 1, 2, 3}
    assert np.all(bc[:, 1025:] == 3)                     # the extraction wall
    # strips cover [0, L) without overlap, densest over the driver half
    for s in W["species"]:
        edges = [(xl, xr) for _, xl, xr in s["loads"]]
        assert all(abs(a[1] - b[0]) < 1e-12 for a, b in zip(edges, edges[1:]))
        ks = [k for k, _, _ in s["loads"]]
        assert ks[0] == max(ks) and ks[-1] == min(ks)
    # quasi-neutral where the density is 1e17: e + H- = H+ + H2+
    k0 = {s["name"]: s["loads"][0][0] for s in W["species"]}
    assert abs(k0["e"] + k0["Hm"] - k0["Hp"] - k0["H2p"]) <= 2


def test_default_bench_workload_is_the_1024_box():
    import argparse  # noqa: F401
    src = open(os.path.join(ROOT, "bench.py")).read()
    assert 'default="argon-1024"' in src
    assert B.host_threads() >= 1
