"""atom_grp2spf through the reference's struct API (mol_atom_group +
form-factor mapping file; pdb2spf.c:154-175) on the GPU, against the
reference's own output for its test receptor (tests/golden/ref_spf, written
by saxs_test.c:59-127 from 4g9s_r_native at L = 15, 50 q)."""
import ctypes as C
import os

import numpy as np
import pytest

from tests.test_front import G, _group

pytestmark = pytest.mark.gpu


def test_atom_grp2spf_struct_api_vs_ref_spf():
    import libfmftsaxs_amd as F
    from libfmftsaxs_amd import synthetic as S
    lib = F.lib()
    lib.default_ff_table.restype = C.c_void_p
    lib.default_ff_table.argtypes = [C.c_char_p]
    lib.atom_grp2spf.restype = C.POINTER(F.SpfFull)
    lib.atom_grp2spf.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int]
    spf = np.load(os.path.join(G, "ref_spf.npz"))
    L, Q = int(spf["L"]), int(spf["qnum"])
    t = np.load(os.path.join(G, "atoms_4g9s_r_native.npz"))
    m = S.Molecule(t["xyz"], ["C"] * len(t["xyz"]))
    xyz = m.translated(-m.center_of_extrema()).coords
    ag, keep = _group(xyz, np.nan_to_num(t["vdw"]), [str(x) for x in t["res"]],
                      [str(x) for x in t["name"]])
    table = lib.default_ff_table(os.path.join(G, "pdb_formfactor_mapping_clean.prm").encode())
    q = F.mkarray(0.0, 0.5, Q)
    p = lib.atom_grp2spf(C.byref(ag), table, q.ctypes.data, Q, L, 1)
    re, im, rm = F.spf_to_arrays(p)
    lib.sxs_spf_full_free(p)
    assert abs(rm - float(spf["rm"])) < 5e-5
    # vacuum and excluded volume: the fixture's 5 significant digits
    for X in (0, 1):
        for a, b in ((re[X], spf["re"][X]), (im[X], spf["im"][X])):
            sig = np.abs(b) > 1e-3 * np.abs(b).max()
            assert (np.abs(a - b)[sig] / np.abs(b)[sig]).max() < 5.01e-5
    # water shell through this library's SASA (libmol2's accs is not available)
    w0, w0r = re[2, 0, 0], spf["re"][2, 0, 0]
    print(f"W00 {w0:.6g} vs reference {w0r:.6g}")
    assert abs(w0 - w0r) / w0r < 1e-3
