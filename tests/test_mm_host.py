"""Host side of the molecular-mechanics path (no GPU): the .cc1 bond list and
build_flat_topology of libfmd against the reference build's golden terms,
bitwise; the Python mirror's array layout."""
import numpy as np
import pytest

import oracle
from conftest import golden
from fuller_md_en_amd import mm, system


@pytest.mark.parametrize("name,spec", [("mm_c60_nc1.npz", "C60"), ("mm_c70_hcp.npz", "C70")])
def test_bonds_and_topology_equal_reference(name, spec):
    g = golden(name)
    path, _ = system.resolve_fullerene(spec)
    bonds = mm.load_cc1_bonds(path)
    assert np.array_equal(bonds, g["bonds"])
    mol = system.load_cc1(path)
    assert np.array_equal(mol.coords, g["coords"])
    nmol = int(g["mol_id"].max()) + 1
    ft = mm.build_flat_topology(mol.coords, bonds, nmol)
    natom = len(mol.coords)
    nb = len(g["t_b"])
    # molecule 0 bitwise, and every replica shifted by natom (MM:314-332)
    assert np.array_equal(np.c_[ft["b_i0"], ft["b_i1"]][:nb], g["t_b"])
    assert np.array_equal(np.c_[ft["b_kb"], ft["b_r0"]][:nb], g["t_bp"])
    na = len(g["t_a"])
    assert np.array_equal(np.c_[ft["ang_i0"], ft["ang_i1"], ft["ang_i2"]][:na], g["t_a"])
    assert np.array_equal(np.c_[ft["ang_kth"], ft["ang_th0"]][:na], g["t_ap"])
    nd = len(g["t_d"])
    d = np.c_[ft["dih_i0"], ft["dih_i1"], ft["dih_i2"], ft["dih_i3"]]
    assert np.array_equal(d[:nd], g["t_d"])
    assert np.array_equal(np.c_[ft["dih_Vn"], ft["dih_gamma"]][:nd], g["t_dp"])
    assert np.array_equal(ft["dih_mult"][:nd], g["t_dm"])
    ni = len(g["t_i"])
    im = np.c_[ft["imp_i0"], ft["imp_i1"], ft["imp_i2"], ft["imp_i3"]]
    assert np.array_equal(im[:ni], g["t_i"])
    assert np.array_equal(np.c_[ft["imp_ki"], ft["imp_gamma"]][:ni], g["t_ip"])
    shift = (np.arange(nmol) * natom).repeat(nd)[:, None]
    assert np.array_equal(d, np.tile(g["t_d"], (nmol, 1)) + shift)
    assert np.array_equal(ft["mol_id"], g["mol_id"])
    assert ft.counts == (nb * nmol, na * nmol, nd * nmol, ni * nmol)


def test_topology_counts_and_bad_input():
    g = golden("mm_c60_nc1.npz")
    cnt = np.zeros(4, np.int32)
    from fuller_md_en_amd import _native as N
    N.check(N.lib().fmd_mm_topology_counts(60, np.ascontiguousarray(g["bonds"]).reshape(-1),
                                           len(g["bonds"]), cnt))
    # the truncated icosahedron: 90 bonds, 3 angles and 1 improper per atom,
    # 4 dihedrals per bond
    assert list(cnt) == [90, 180, 360, 60]
    bad = np.ascontiguousarray(g["bonds"]).reshape(-1).copy()
    bad[5] = 60  # an atom index out of range
    assert N.lib().fmd_mm_topology_counts(60, bad, len(g["bonds"]), cnt) != 0


def test_lj_params_equal_oracle():
    p = mm.lj_params()
    o = oracle.LJParams()
    oracle.lib().orc_lj_params_mm(o)
    assert p.sigma ** 2 == o.sig2 and p.eps == o.eps and p.rcut ** 2 == o.rc2
    assert p.vshift == o.vshift
