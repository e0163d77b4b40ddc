"""CPU-side checks of the C ABI library and the host data layer (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import REPO
from fock_amd import abi, molecule as m

HEADER = os.path.join(REPO, "include", "fock.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(fock_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(abi.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 14
    for name in names:
        assert hasattr(lib, name), name
    assert set(abi.EXPORTED) <= set(names)


def test_version_string():
    assert b"gfx950" in abi.load().fock_version()


def test_default_options_are_the_reference_constants():
    o = abi.default_opts()
    assert o.max_iters == 100 and o.epsilon == 1e-6 and o.molecule_charge == 0
    assert o.integral_policy == abi.POLICY_REF_COMPAT and o.screen == abi.SCREEN_REF_BF_PAIR
    assert o.precision == abi.PRECISION_FP64 and o.incremental == 0
    assert o.diis_start == 5 and o.diis_window == 16 and o.level_shift == 0.0  # scf/src/lib.rs:197-207
    assert o.autotune == 0
    assert o.ediis == 0 and o.damping == 1.0          # the reference's mixing factor F = 1.0 (:218)


def test_struct_sizes_match_header_layout():
    # offsets of the C structs (x86-64 SysV): keeps the ctypes mirror honest
    assert ctypes.sizeof(abi.BasisDesc) == 48
    assert ctypes.sizeof(abi.NucleiDesc) == 24
    assert ctypes.sizeof(abi.ScfOpts) == 120
    assert ctypes.sizeof(abi.ScfIter) == 32
    assert ctypes.sizeof(abi.JKStats) == 12 + 4 + 6 * 8 * 5 + 8 + 8 + 8
    assert ctypes.sizeof(abi.TuneReport) == 6 * 4 + 6 * 7 * 8 + 4 + 4 + 2 * 8 + 8


def test_component_order_is_the_references():
    # basis_set/src/lib.rs:87-99: p as (z, y, x); d as zz, yz, yy, xz, xy, xx
    assert m.cartesian_components(1) == [(0, 0, 1), (0, 1, 0), (1, 0, 0)]
    assert m.cartesian_components(2) == [(0, 0, 2), (0, 1, 1), (0, 2, 0), (1, 0, 1), (1, 1, 0), (2, 0, 0)]


def test_primitive_norm_unit_self_overlap():
    # a normalised primitive has unit self-overlap: N^2 (pi/2a)^{3/2} prod (2l-1)!!/(4a)^l = 1
    for ang in [(0, 0, 0), (1, 0, 0), (0, 1, 0), (2, 0, 0), (1, 1, 0)]:
        for a in [0.1, 1.3, 27.0]:
            N = m.primitive_norm(ang, a)
            dfact = lambda k: np.prod(np.arange(2 * k - 1, 0, -2)) if k > 0 else 1.0
            s = (np.pi / (2 * a)) ** 1.5 * np.prod([dfact(l) / (4 * a) ** l for l in ang])
            assert abs(N * N * s - 1.0) < 1e-13


def test_sizes_of_named_workloads():
    fb, fn = m.flatten(m.h2o_sto3g())
    assert (fb.n, len(fb.exponents), m.n_electrons(m.h2o_sto3g())) == (7, 21, 10)
    fb, _ = m.flatten(m.octane_631g())
    assert (fb.n, len(fb.exponents)) == (108, 248)
    fb, _ = m.flatten(m.eicosane_631g())
    assert (fb.n, len(fb.exponents)) == (264, 608)


def test_xyz_reader_uses_1p89(tmp_path):
    p = tmp_path / "w.xyz"
    p.write_text("3\nWater\nH 0.75 0.0 0.44\nO 0.0 0.0 -0.165\nH -0.75 0.0 0.44\n")
    atoms = m.read_xyz(str(p), m.BasisSet.load("STO-3G"))
    assert [a.z for a in atoms] == [1, 8, 1]
    assert atoms[0].position[0] == 0.75 * 1.89


def test_water_cluster_generator_is_seeded_and_valid():
    a = m.water_cluster_batch(8, seed=7)
    b = m.water_cluster_batch(8, seed=7)
    assert a == b
    for g in a:
        O = np.array([[x, y, z] for s, x, y, z in g if s == "O"])
        k = len(O)
        assert 1 <= k <= 8 and len(g) == 3 * k
        if k > 1:
            d = np.linalg.norm(O[:, None] - O[None], axis=-1) + np.eye(k) * 99
            assert d.min() >= 2.8 - 1e-12
        assert np.all(np.linalg.norm(O, axis=1) <= 2.0 * k ** (1 / 3) + 1e-12)
