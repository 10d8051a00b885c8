"""C ABI of liborbc_hip.so and the host-side code that runs without a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import openrbc_amd as orbc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "orbc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(orbc_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = orbc.load_library()
    names = header_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n


def test_python_binding_covers_the_header():
    bound = {name for name, _, _ in orbc.SIGNATURES}
    assert set(header_functions()) <= bound


def test_config_defaults_are_the_reference_defaults():
    cfg = orbc.default_config()
    # RTParameter, src/runtime_parameter.h:47-80
    assert cfg.kBT == pytest.approx(0.22) and cfg.eta == pytest.approx(0.01) and cfg.dt == pytest.approx(0.01)
    assert cfg.rho == pytest.approx(1.05) and cfg.voronoi_cell_size == 14
    assert (cfg.freq_voronoi, cfg.freq_sort_ctrd, cfg.freq_sort_bond, cfg.freq_cleanup) == (2, 24, 120, 60)
    assert cfg.seed == 0xBAD5EED and list(cfg.box_lo) == [-1000] * 3 and list(cfg.box_hi) == [1000] * 3


def test_invalid_configuration_is_an_error_not_an_exit():
    # src/runtime_parameter.h:118: freq_sort_ctrd must be a multiple of freq_voronoi
    L = orbc.load_library()
    cfg = orbc.default_config(freq_voronoi=7)
    h = C.c_void_p()
    assert L.orbc_create(C.byref(cfg), C.byref(h)) == 1        # ORBC_ERR_INVALID
    cfg = orbc.default_config(dt=-1.0)
    assert L.orbc_create(C.byref(cfg), C.byref(h)) == 1
    # lipids per Voronoi cell must be positive (it divides the lipid count)
    for bad in (0, -3):
        cfg = orbc.default_config(voronoi_cell_size=bad)
        assert L.orbc_create(C.byref(cfg), C.byref(h)) == 1
    # the kernels are built for workgroups of 256
    cfg = orbc.default_config(block_size=128)
    assert L.orbc_create(C.byref(cfg), C.byref(h)) == 1


def test_create_without_gpu_fails_cleanly():
    L = orbc.load_library()
    cfg = orbc.default_config()
    h = C.c_void_p()
    rc = L.orbc_create(C.byref(cfg), C.byref(h))
    assert rc in (0, 2)          # OK on a GPU host, ORBC_ERR_HIP without one; never a crash
    if rc == 0:
        L.orbc_destroy(h)


def test_init_rbc_counts_icosphere_level3():
    """C1: icosphere level 3 (642 V, 1920 E, 1280 F) populated by the init_rbc
    rules gives 185,700 lipids + 39,684 proteins and 40,962 bonds."""
    v, e, f = orbc.icosphere(3)
    assert (len(v), len(e), len(f)) == (642, 1920, 1280)
    s = orbc.init_rbc(v, e, f, seed=1)
    assert len(s["lipid_x"]) == 185700
    assert len(s["protein_x"]) == 39684 and len(s["bond_type"]) == 40962
    counts = np.bincount(s["protein_type"], minlength=6)
    assert list(counts) == [0, 0, 1920, 642, 642, 1920 * 19]
    np_ = len(s["protein_x"])
    assert s["protein_id"].min() == 1 and s["protein_id"].max() == np_
    assert s["lipid_id"].min() == np_ + 1                # src/init_rbc.h:306
    # bonds: actin-glycophorin (0), spectrin-band3 (1), actin-spectrin (2), spectrin chain (3)
    assert list(np.bincount(s["bond_type"], minlength=4)) == [642, 1920, 2 * 1920, 18 * 1920]
    assert np.allclose(np.linalg.norm(s["lipid_n"], axis=1), 1.0, atol=1e-6)


def test_init_rbc_example_large_counts():
    """The full-size RBC mesh: 2,639,690 lipids + 565,816 proteins, 584,066 bonds."""
    v, e, f = orbc.load_example_large()
    assert (len(v), len(e), len(f)) == (9128, 27378, 18252)
    s = orbc.init_rbc(v, e, f, seed=7)
    assert len(s["lipid_x"]) == 2639690 and len(s["protein_x"]) == 565816 and len(s["bond_type"]) == 584066


def test_init_rbc_geometry_rules():
    v, e, f = orbc.icosphere(2)
    s = orbc.init_rbc(v, e, f, seed=3)
    nv = len(v)
    px, pt = s["protein_x"], s["protein_type"]
    gly, act = px[:nv], px[nv:2 * nv]
    # actin 2.25 below glycophorin along the vertex normal
    assert np.allclose(np.linalg.norm(gly - act, axis=1), 2.25, atol=1e-5)
    # mean mesh edge rescaled to 80 nm / 4.4545 nm
    assert abs(np.linalg.norm(gly[e[:, 0]] - gly[e[:, 1]], axis=1).mean() - 80 / 4.4545) < 1e-3
    # no lipid within 1.685 of a protein (src/init_rbc.h:275-279)
    from scipy.spatial import cKDTree
    d, _ = cKDTree(px).query(s["lipid_x"])
    assert d.min() >= 1.685 - 1e-6
