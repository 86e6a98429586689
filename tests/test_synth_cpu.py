"""Synthetic inputs in the reference formats (cb_synth_gr_chains,
cb_synth_go_trees, cb_synth_io_gj; host only, no GPU):

* granule chains: Gr.txt rows (17 columns), one block of comp rows per cell,
  an unbranched chain rooted at its last compartment; compartment k takes
  Ra, Cm, type and the ten densities of the template's first row of type
  types[k], and that row's diameter and length times a factor in
  [0.8, 1.25); the oracle reads a block as a population of comp-compartment
  cells with the cylinder areas of that geometry;
* Golgi trees: Go_131.txt rows (20 columns), the template's topology, types
  and densities, every diameter and length times a factor in [0.8, 1.25);
* olive junctions: every cell has at least k distinct partners, no self
  pairs, each pair once with pre < post, g within [g_min, g_max];
* the same seed gives the same files, another seed other geometry.
"""
import ctypes
import os

import numpy as np

import cbrkc
import orc


def rows(path):
    return np.array([[float(x) for x in l.split()] for l in open(path) if l.strip()])


def first_of_type(tpl, t):
    return tpl[np.where(tpl[:, 6] == t)[0][0]]


def test_granule_chains(tmp_path, ref_data):
    tplf = os.path.join(ref_data, "Gr.txt")
    tpl = rows(tplf)
    f = str(tmp_path / "gr.txt")
    types = [3, 4, 6, 7]
    cbrkc.synth_gr_chains(f, tplf, 11, 50, types)
    r = rows(f)
    assert r.shape == (200, 17)
    for c in range(50):
        blk = r[4 * c:4 * c + 4]
        assert list(blk[:, 0]) == [0, 1, 2, 3]
        assert list(blk[:, 1]) == [1, 2, 3, -1]
        for k, t in enumerate(types):
            p = first_of_type(tpl, t)
            assert blk[k, 6] == t
            np.testing.assert_array_equal(blk[k, 4:6], p[4:6])       # Ra, Cm
            np.testing.assert_array_equal(blk[k, 7:], p[7:])         # densities
            for col in (2, 3):                                       # diameter, length
                fac = blk[k, col] / p[col]
                assert 0.8 - 1e-6 <= fac < 1.25 + 1e-6, (c, k, col, fac)
    fac = r[:, 2] / np.tile([first_of_type(tpl, t)[2] for t in types], 50)
    assert np.std(fac) > 0.05                                        # the factors do vary
    g = str(tmp_path / "gr2.txt")
    cbrkc.synth_gr_chains(g, tplf, 11, 50, types)
    assert open(f).read() == open(g).read()
    cbrkc.synth_gr_chains(g, tplf, 12, 50, types)
    assert open(f).read() != open(g).read()
    # the oracle reads the first block as a 4-compartment cell: cylinder areas
    one = str(tmp_path / "gr1.txt")
    with open(one, "w") as out:
        out.writelines(open(f).readlines()[:4])
    L = orc.lib()
    dp = ctypes.POINTER(ctypes.c_double)
    L.or_gr_create_comp.restype = ctypes.c_void_p
    L.or_gr_create_comp.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_int]
    v0 = np.array([-70.0])
    o = orc.Pop(L.or_gr_create_comp(one.encode(), 4, 1, 1, orc._dp(v0), 1e-9, 1e-9, 0), 1, 4)
    blk = rows(one)
    np.testing.assert_allclose(o.get(107), 2 * np.pi * (0.5 * blk[:, 2] * 1e-4) * (blk[:, 3] * 1e-4), rtol=1e-6)


def test_golgi_trees(tmp_path, ref_data):
    tplf = os.path.join(ref_data, "Go_131.txt")
    tpl = rows(tplf)
    f = str(tmp_path / "go.txt")
    cbrkc.synth_go_trees(f, tplf, 5, 7)
    r = rows(f)
    assert r.shape == (7 * 131, 20)
    for c in range(7):
        blk = r[131 * c:131 * (c + 1)]
        np.testing.assert_array_equal(blk[:, [0, 1, 4, 5, 6]], tpl[:, [0, 1, 4, 5, 6]])
        np.testing.assert_array_equal(blk[:, 7:], tpl[:, 7:])
        fac = blk[:, 2:4] / tpl[:, 2:4]
        assert np.all(fac >= 0.8 - 1e-6) and np.all(fac < 1.25 + 1e-6)


def test_olive_junction_list(tmp_path):
    f = str(tmp_path / "gj.csv")
    n, k = 300, 6
    cbrkc.synth_io_gj(f, 9, n, k=k, g_min_nS=0.5, g_max_nS=1.5)
    pairs = [l.strip().split(",") for l in open(f) if l.strip()]
    a = np.array([int(p[0]) for p in pairs])
    b = np.array([int(p[1]) for p in pairs])
    g = np.array([float(p[2]) for p in pairs])
    assert np.all(a < b) and np.all(a >= 0) and np.all(b < n)
    assert len(set(zip(a.tolist(), b.tolist()))) == len(pairs)
    assert np.all(g >= 0.5) and np.all(g <= 1.5) and np.std(g) > 0.2
    deg = np.bincount(np.concatenate([a, b]), minlength=n)
    assert deg.min() >= k and deg.mean() < 2 * k
