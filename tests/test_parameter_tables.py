"""The committed parameter tables against the reference's parameter headers.

xtbloom_amd/data/gfn2_tables.h is generated from the reference's gfn2.json,
tblite_sto.hpp and d4.hpp.  The fixture tests/golden/gfn2_parameters_hcno.json
holds the H, C, N and O entries of the reference's own C++ parameter headers
(data/parameters/gfn2.hpp: kElements, kShells; data/parameters/d4.hpp:
kElements, kReferences, kReferenceC6), a second transcription of the same
parameters, for the elements every generated workload is made of (nitrogen
appears in no reference golden or SCC trace).  Every value must agree
exactly, including the shell and reference offsets the kernels index with.
"""
import json
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

TABLES = os.path.join(ROOT, "xtbloom_amd", "data", "gfn2_tables.h")


def _arrays():
    text = open(TABLES).read()
    out = {}
    for m in re.finditer(r"static const (double|int) (xtbp_\w+)\[(\d+)\] = \{(.*?)\};", text, re.S):
        kind, name, count, body = m.groups()
        vals = [float(v) for v in body.replace("\n", " ").split(",") if v.strip()]
        assert len(vals) == int(count), name
        out[name] = np.array(vals, dtype=float if kind == "double" else int)
    return out


@pytest.fixture(scope="module")
def tables():
    return _arrays()


@pytest.fixture(scope="module")
def ref():
    with open(os.path.join(GOLDEN, "gfn2_parameters_hcno.json")) as fh:
        return json.load(fh)


@pytest.mark.parametrize("z", [1, 6, 7, 8])
def test_element_and_shell_parameters(tables, ref, z):
    r = ref["elements"][str(z)]
    i = z - 1
    for key, name in (("gam", "xtbp_elem_gam"), ("gam3", "xtbp_elem_gam3"),
                      ("zeff", "xtbp_elem_zeff"), ("arep", "xtbp_elem_arep"),
                      ("en", "xtbp_elem_en"), ("dkernel", "xtbp_elem_dkernel"),
                      ("qkernel", "xtbp_elem_qkernel"), ("mprad", "xtbp_elem_mprad"),
                      ("mpvcn", "xtbp_elem_mpvcn")):
        assert tables[name][i] == r[key], (z, key)
    off, cnt = r["shell_offset"], r["shell_count"]
    assert tables["xtbp_elem_shell_offset"][i] == off and tables["xtbp_elem_shell_count"][i] == cnt
    for k, sh in enumerate(r["shells"]):
        p = off + k
        assert tables["xtbp_shell_n"][p] == sh["n"] and tables["xtbp_shell_l"][p] == sh["l"]
        assert tables["xtbp_shell_ngauss"][p] == sh["ngauss"]
        for key, name in (("level_ev", "xtbp_shell_level_ev"), ("slater", "xtbp_shell_slater"),
                          ("refocc", "xtbp_shell_refocc"), ("kpoly", "xtbp_shell_kpoly"),
                          ("kcn_ev", "xtbp_shell_kcn_ev"),
                          ("hubbard_scale", "xtbp_shell_hubbard_scale")):
            assert tables[name][p] == sh[key], (z, k, key)


@pytest.mark.parametrize("z", [1, 6, 7, 8])
def test_d4_element_and_reference_data(tables, ref, z):
    d = ref["elements"][str(z)]["d4"]
    i = z - 1
    r0, rn = d["ref_offset"], d["ref_count"]
    assert tables["xtbp_d4_ref_offset"][i] == r0 and tables["xtbp_d4_ref_count"][i] == rn
    for key, name in (("covalent_radius", "xtbp_d4_covalent_radius"), ("en", "xtbp_d4_en"),
                      ("zeff", "xtbp_d4_zeff"), ("hardness", "xtbp_d4_hardness"),
                      ("r4r2", "xtbp_d4_r4r2")):
        assert tables[name][i] == d[key], (z, key)
    assert np.array_equal(tables["xtbp_d4_ref_cn"][r0:r0 + rn], d["ref_cn"])
    assert np.array_equal(tables["xtbp_d4_ref_q"][r0:r0 + rn], d["ref_q"])
    assert np.array_equal(tables["xtbp_d4_ref_ngauss"][r0:r0 + rn], d["ref_ngauss"])


def test_d4_reference_c6_among_hcno(tables, ref):
    nref = int(round(np.sqrt(len(tables["xtbp_d4_c6"]))))
    assert nref * nref == len(tables["xtbp_d4_c6"])
    for i, j, v in ref["c6"]:
        assert tables["xtbp_d4_c6"][i * nref + j] == v, (i, j)
