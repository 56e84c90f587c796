"""The oracle against the reference's own known-answer tests (test/tests_ints.cpp).

Scalar KATs are printed in the reference with 12 decimals and checked there at absolute 1e-12
(tests_ints.cpp:15); they are checked the same way here.  The ERI checksums are sums of |x| over
10^4-10^7 values, which move at ~1e-13 relative with summation order alone (the reference's own
eri4 and deployERI4Kernel sums of the same tensor differ by 1.4e-13 relative), so they are checked
at 1e-12 relative: a dropped term, a wrong sign or a transposed operand changes them at >=1e-6.
"""
import numpy as np
import pytest

from lible_amd import basis as B
from tests.conftest import C2H4, CO2, H2O, O3, basis_path

ABS = 1e-12
REL = 1e-12


def mol(main, geom, aux=None):
    nrs, xyz = geom
    return B.molecule_from_basis_files(basis_path(main), nrs, xyz,
                                       basis_path(aux) if aux else None)


def test_calc_boys_f(oracle):                                   # tests_ints.cpp:1296-1313
    assert abs(oracle.boys(5, 2.5).sum() - 0.728561124180) < ABS


def test_calc_rints3d(oracle):                                  # tests_ints.cpp:1315-1335
    r = oracle.rints3d(5, 0.342412347, [0.5, -1.2, 3.4], oracle.boys(5, 2.5))
    assert abs(np.abs(r).sum() - 2.814391945942) < ABS


def test_pure_primitive_norm(oracle):                           # tests_ints.cpp:1359-1371
    assert abs(oracle.pure_primitive_norm(4, 1.70325241) - 4.813408368164) < ABS


def test_ecoeffs_recurrence1(oracle):                           # tests_ints.cpp:1150-1165
    s = np.abs(oracle.ecoeffs_recurrence1(19.0034512342, 4)).sum()
    assert abs(s - 181021.807437296084) <= REL * 181021.8


def test_ecoeffs_recurrence2(oracle):                           # tests_ints.cpp:1167-1181
    e = oracle.ecoeffs_recurrence2(0.0023445, 1.123621, 2, 3, 0.5113241, 1.2345123, 0.003451234)
    assert abs(np.abs(e).sum() - 0.218834692700) < ABS


def test_ecoeffs_recurrence2_n1(oracle):                        # tests_ints.cpp:1183-1208
    e = oracle.ecoeffs_recurrence2_n1(0.0023445, 1.123621, 2, 3, 0.5113241, 1.2345123,
                                      0.003451234, -0.571235891, 1.612357121)
    assert abs(np.abs(e).sum() - 0.053908505303) < ABS


def test_ecoeffs_primitive_pair(oracle):                        # tests_ints.cpp:1225-1242
    e = oracle.ecoeffs_primitive_pair(1230.02342162, 0.0023445, 4, 3, [0, 1.5, 0], [1, 0, -1])
    assert abs(np.abs(e).sum() - 16.100457827896) < ABS


def test_ecoeffs_primitive_pair_n1(oracle):                     # tests_ints.cpp:1244-1275
    e = oracle.ecoeffs_primitive_pair(1230.02342162, 0.0023445, 4, 3, [0, 1.5, 0], [1, 0, -1],
                                      [0.571235891, 1.612357121, -0.234512341],
                                      [1.2345123, -0.5113241, 0.003451234], deriv=1)
    assert abs(np.abs(e).sum() - 21.712096988812) < ABS


def test_ecoeffs_shell(oracle):                                 # tests_ints.cpp:1277-1294
    s = oracle.ecoeffs_shell_abs_sum(5, [0.95556865, 0.36628794, 0.14651517])
    assert abs(s - 112231.357645556767) <= REL * 112231.36


def test_spherical_trafo(oracle):                               # tests_ints.cpp:856-871
    t = oracle.sph_trafo(8)
    s = sum(abs(v) + i + j for i, j, v in t)
    assert abs(s - 7005.923248093791) < ABS


@pytest.mark.parametrize("name,main,geom,aux,ref", [
    ("eri4", "def2-svp", O3, None, 8246.711763196334),                     # :748-764
    ("eri4_diag", "def2-tzvp", O3, None, 306.806419250122),                # :712-728
    ("eri3", "def2-tzvp", O3, "def2-qzvp-rifit", 25854.105851585416531),   # :730-746
    ("eri2", "def2-svp", H2O, "def2-qzvp-rifit", 8801.334703460838),       # :694-710
    ("eri2_diag", "def2-svp", H2O, "def2-universal-jkfit", 830.788976485982),  # :149-165
])
def test_eri_checksums(oracle, name, main, geom, aux, ref):
    s = oracle.Structure(mol(main, geom, aux))
    val = np.abs(getattr(s, name)()).sum()
    assert abs(val - ref) <= REL * ref, (val, ref)


@pytest.mark.parametrize("which,symm,ref", [
    ("eri4", False, 8246.711763197583),        # deployERI4Kernel :873-900
    ("eri4d1", True, 30546.780073140635),      # deployERI4D1Kernel :961-990
])
def test_deploy_eri4(oracle, which, symm, ref):
    s = oracle.Structure(mol("def2-svp", O3))
    val = s.deploy_sum(which, symm)
    assert abs(val - ref) <= REL * ref, (val, ref)


@pytest.mark.parametrize("which,ref", [
    ("eri3", 12350.130354189232),     # deployERI3Kernel :902-930
    ("eri2", 21014.512841176023),     # deployERI2Kernel :932-959
    ("eri3d1", 78466.360848033844),   # deployERI3D1Kernel :992-1022
    ("eri2d1", 39949.218421684251),   # deployERI2D1Kernel :1024-1053
])
def test_deploy_ri(oracle, which, ref):
    s = oracle.Structure(mol("def2-svp", O3, "def2-universal-jkfit"))
    val = s.deploy_sum(which, False)
    assert abs(val - ref) <= REL * ref, (val, ref)


@pytest.mark.parametrize("main,geom,aux,which,ref", [
    ("def2-tzvp", O3, "def2-qzvp-rifit", 0, 25854.105851585416531),   # eri3 :730-746
    ("def2-svp", H2O, "def2-qzvp-rifit", 2, 8801.334703460838),       # eri2 :694-710
])
def test_ri_sums_kat(oracle, main, geom, aux, which, ref):
    """The dense-free checksums of eri3/eri2 (used for the 256-molecule fixtures) against the
    reference's eri3 and eri2 checksums."""
    s = oracle.Structure(mol(main, geom, aux))
    val = s.ri_sums()[which]
    assert abs(val - ref) <= REL * ref, (val, ref)


def test_ri_sums_match_dense(oracle):
    """All four sums against the dense tensors on a synthetic molecule with f main and g
    auxiliary shells (signed sums too: a wrong image weight changes them)."""
    m = B.synthetic_molecule(77, 3, 4.0, B.RECIPE_MAIN, B.RECIPE_AUX)
    s = oracle.Structure(m)
    e3, e2 = s.eri3(), s.eri2()
    ref = np.array([np.abs(e3).sum(), e3.sum(), np.abs(e2).sum(), e2.sum()])
    got = s.ri_sums()
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max()), (got, ref)


@pytest.mark.parametrize("which,aux,ref", [
    ("eri2d2", "def2-universal-jkfit", 175567.635699242965),   # deployERI2D2Kernel :1055-1085
    ("eri4soc", None, 12940.953099658431),                     # deployERI4SOCKernel :1087-1116
    ("eri3soc", "def2-universal-jkfit", 18124.713152990007),   # deployERI3SOCKernel :1118-1148
])
def test_deploy_d2_soc(oracle, which, aux, ref):
    s = oracle.Structure(mol("def2-svp", O3, aux))
    val = s.deploy_sum2(which)
    assert abs(val - ref) <= REL * ref, (val, ref)


@pytest.mark.parametrize("which,basis,geom,ref", [
    ("overlap", "cc-pvdz", C2H4, 226.099184393591),        # tests_ints.cpp:167-183
    ("kinetic", "cc-pvdz", C2H4, 252.755644635021),        # :236-252
    ("nuclear", "6-31+g", CO2, 2141.927982294058),         # :303-319
])
def test_one_electron_checksums(oracle, which, basis, geom, ref):
    m = mol(basis, geom)
    s = oracle.Structure(m)
    if which == "nuclear":
        charges = np.column_stack([m.coords, np.asarray(geom[0], np.float64)])
        val = np.abs(s.external_charges(charges)).sum()
    else:
        val = np.abs(getattr(s, which)()).sum()
    assert abs(val - ref) <= REL * ref, (val, ref)
