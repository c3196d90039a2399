"""Host-side checks that need no GPU: the C-ABI library loads and exports
every entry point declared in include/mdft_gpu.h; workloads and grid helpers."""
import os
import re

import numpy as np

from mdft import grid, workloads
from mdft.gpu import MDFT_EXPORTS, load_library

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "mdft_gpu.h")).read()
    return sorted(set(re.findall(r"\b(mdft_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = load_library()
    names = declared_symbols()
    assert set(names) == set(MDFT_EXPORTS)
    for n in names:
        assert hasattr(lib, n), n


def test_grid_helpers_match_oracle(oracle_mod):
    for mmax in range(0, 6):
        g = grid.angular_grid(mmax, 2)
        o = oracle_mod.Oracle((4, 4, 4), (8.0, 8.0, 8.0), mmax, 2, 2.4, 0.0333)
        assert g["no"] == o.no and g["np"] == o.np_
        assert np.allclose(g["w"], o.w, atol=1e-15)
        assert np.allclose(g["thetaofntheta"], o.theta, atol=1e-15)


def test_workload_solutes():
    w = workloads.WORKLOADS["solute20_128"]
    s = w.solute()
    assert s.xyz.shape == (20, 3) and abs(s.net_charge) < 1e-12
    d = np.sqrt(((s.xyz[:, None] - s.xyz[None]) ** 2).sum(-1)) + np.eye(20) * 9
    assert d.min() >= 1.2
    assert np.all(np.linalg.norm(s.xyz - np.array(w.length) / 2, axis=1) <= 5.0)
    s0 = workloads.WORKLOADS["lj_sphere_48"].solute()
    assert s0.sig[0] == 3.73 and s0.eps[0] == 1.23
    assert np.array_equal(w.solute().xyz, s.xyz)  # seeded


def test_vext_orientation_independent_lj_only():
    w = workloads.small("lj_sphere_48", 8)
    v = workloads.vext_numpy(w, w.solute())
    # a neutral LJ sphere acts on the oxygen site only, which sits at the molecule origin
    assert np.allclose(v, v[:, :1])
    assert v.max() <= 694.7 + 1e-9


def test_library_exports_every_warm_start_symbol():
    from mdft.warmstart import WSNET_EXPORTS
    lib = load_library()
    txt = open(os.path.join(ROOT, "include", "mdft_wsnet.h")).read()
    names = sorted(set(re.findall(r"\b(mdft_[a-z0-9_]+)\s*\(", txt)))
    assert set(names) == set(WSNET_EXPORTS)
    for n in names:
        assert hasattr(lib, n), n
