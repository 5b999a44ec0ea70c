"""CPU tests of the boundary and host logic (no GPU compute).

* libpt_gravity.so loads and exports every symbol include/pt_gravity.h declares;
* configuration handling mirrors the reference (CLI letters, -x param files,
  only the octree path accepted) and fails with error codes, not aborts;
* the GravityVisitor factors and the input generators / Tipsy I/O.
"""
import ctypes
import os
import re

import numpy as np
import pytest

from paratreet_amd import _lib, ics
from paratreet_amd.gravity import Configuration, GravityVisitor, TreeSpec

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "pt_gravity.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pt_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    declared = _declared_symbols()
    assert len(declared) >= 30
    for name in declared:
        assert hasattr(L, name), name
    assert set(declared) == set(_lib.EXPORTS)


def test_config_default_matches_reference_defaults():
    c = _lib.PtConfig()
    _lib.load().pt_config_default(ctypes.byref(c))
    assert c.max_particles_per_leaf == 12          # examples/Main.C:29
    assert c.tree_type == _lib.PT_TREE_OCT and c.decomp_type == _lib.PT_DECOMP_OCT
    assert c.min_n_subtrees == 1 and c.min_n_partitions == 1 and c.has_dsoft == 0


@pytest.mark.parametrize("tree,decomp", [(_lib.PT_TREE_KD, _lib.PT_DECOMP_OCT),
                                         (_lib.PT_TREE_OCT, _lib.PT_DECOMP_SFC),
                                         (_lib.PT_TREE_BINARY_OCT, _lib.PT_DECOMP_BINARY_OCT)])
def test_unsupported_trees_return_error_code(tree, decomp):
    L = _lib.load()
    c = _lib.PtConfig()
    L.pt_config_default(ctypes.byref(c))
    c.tree_type, c.decomp_type = tree, decomp
    ctx = ctypes.c_void_p()
    assert L.pt_ctx_create(ctypes.byref(c), 0, None, ctypes.byref(ctx)) == _lib.PT_ERR_UNSUPPORTED
    assert not ctx.value


def test_invalid_bucket_and_periodic_rejected():
    L = _lib.load()
    c = _lib.PtConfig()
    L.pt_config_default(ctypes.byref(c))
    c.max_particles_per_leaf = 0
    ctx = ctypes.c_void_p()
    assert L.pt_ctx_create(ctypes.byref(c), 0, None, ctypes.byref(ctx)) == _lib.PT_ERR_ARG
    # periodic needs a positive finite period and 0 <= nReplicas <= 16
    L.pt_config_default(ctypes.byref(c))
    c.periodic = 1
    assert L.pt_ctx_create(ctypes.byref(c), 0, None, ctypes.byref(ctx)) == _lib.PT_ERR_ARG
    c.period[0] = c.period[1] = c.period[2] = 4.0
    c.n_replicas = 17
    assert L.pt_ctx_create(ctypes.byref(c), 0, None, ctypes.byref(ctx)) == _lib.PT_ERR_ARG
    c.n_replicas = 1
    c.period[2] = float("inf")
    assert L.pt_ctx_create(ctypes.byref(c), 0, None, ctypes.byref(ctx)) == _lib.PT_ERR_ARG
    assert L.pt_ctx_create(None, 0, None, ctypes.byref(ctx)) == _lib.PT_ERR_ARG


def test_periodic_parameter_file(tmp_path):
    """bPeriodic / dxPeriod / dyPeriod / dzPeriod / nReplicas are parameter-file
    fields without command-line letters (src/Configuration.h:145-149)."""
    from paratreet_amd.gravity import Configuration
    f = tmp_path / "p.param"
    f.write_text("bPeriodic = 1\ndxPeriod = 2.5\ndyPeriod = 3\ndzPeriod = 4\nnReplicas = 2\n")
    cfg = Configuration()
    cfg.parse(["-x", str(f)])
    c = cfg.to_c()
    assert c.periodic == 1 and c.n_replicas == 2
    assert tuple(c.period) == (2.5, 3.0, 4.0)


def test_null_context_calls_fail_cleanly():
    L = _lib.load()
    assert L.pt_assign_keys(None) == _lib.PT_ERR_ARG
    assert L.pt_sort(None) == _lib.PT_ERR_ARG
    assert L.pt_last_error(None) == b"null context"
    L.pt_ctx_destroy(None)


def test_configuration_parse_cli_and_param_file(tmp_path):
    cfg = Configuration()
    rest = cfg.parse(["-l", "16", "-e", "0.005", "-t", "oct", "-d", "oct", "-n", "1", "-p", "1",
                      "-i", "2", "-m", "-f", "in.tipsy"])
    assert cfg.max_particles_per_leaf == 16 and cfg.dSoft == 0.005
    assert cfg.num_iterations == 2 and cfg.input_file == "in.tipsy"
    assert rest == ["-m"]  # app-level flags (Morton) are left to the caller
    c = cfg.to_c()
    assert c.has_dsoft == 1 and c.dsoft == 0.005 and c.max_particles_per_leaf == 16
    pf = tmp_path / "run.param"
    pf.write_text("nParticlesPerLeafMax = 32\n# comment\ndSoft = 0.01\nachTreeType = oct\n")
    cfg2 = Configuration()
    cfg2.parse(["-x", str(pf)])
    assert cfg2.max_particles_per_leaf == 32 and cfg2.dSoft == 0.01
    assert Configuration().to_c().has_dsoft == 0


def test_treespec_only_octree():
    TreeSpec(Configuration())
    with pytest.raises(_lib.PtError):
        TreeSpec(Configuration(tree_type="kd"))
    assert Configuration(tree_type="binoct").branch_factor() == 2


def test_gravity_visitor_factors():
    v = GravityVisitor((0.0, 0.0, 0.0), 0.7)
    assert v.gravity_factor == (4.0 / 3.0) / (0.7 * 0.7)
    assert np.isclose(v.monopole_gravity_factor, np.sqrt(4.0 / 3.0) / 0.7 ** 4)
    assert v.CallSelfLeaf and v.nMinParticleNode == 6


def test_plummer_generator_properties():
    n = 20000
    pos, vel, m = ics.plummer(n, seed=1)
    r = np.linalg.norm(pos, axis=1)
    # truncated at 10a before the (small) centre-of-mass shift
    assert r.max() <= 10.05
    assert np.allclose(pos.mean(0), 0, atol=1e-12) and np.allclose(vel.mean(0), 0, atol=1e-12)
    # half-mass radius of a Plummer sphere (a = 1, truncation at 10a shifts it slightly)
    assert abs(np.median(r) - 1.305) < 0.04
    # virial ratio 2K/|W| ~ 1 (W = -3 pi M^2 / (32 a) for the untruncated model)
    K = 0.5 * (m * (vel ** 2).sum(1)).sum()
    assert abs(2 * K / (3 * np.pi / 32) - 1) < 0.05
    p2, _, _ = ics.plummer(n, seed=1)
    assert np.array_equal(pos, p2)


def test_halo_mixture_shape():
    pos, vel, m = ics.halo_mixture(50000, n_halos=16, seed=2)
    assert pos.shape == (50000, 3) and np.allclose(m, 1 / 50000)
    inside = np.all((pos >= 0) & (pos <= 100), axis=1)
    assert inside.mean() > 0.9


def test_tipsy_and_acc_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    n = 7
    dark = rng.standard_normal((n, 9)).astype(">f4")
    hdr = np.array([(1.5, n, 3, 0, n, 0, 0)], dtype=np.dtype(">f8,>i4,>i4,>i4,>i4,>i4,>i4"))
    f = tmp_path / "t.tipsy"
    f.write_bytes(hdr.tobytes() + dark.tobytes())
    t = ics.read_tipsy(str(f))
    assert t["time"] == 1.5
    assert np.array_equal(t["pos"], dark[:, 1:4].astype(np.float64))
    assert np.array_equal(t["soft"], dark[:, 7].astype(np.float64))
    acc = rng.standard_normal((n, 3))
    ics.write_acc(str(tmp_path / "x.acc"), acc)
    back = ics.read_acc(str(tmp_path / "x.acc"))
    assert np.allclose(back, acc, rtol=1e-13)
