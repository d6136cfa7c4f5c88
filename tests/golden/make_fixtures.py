#!/usr/bin/env python3
"""Write the committed parity fixtures from the reference's own test data.

Run where the reference tree is mounted (default /root/reference).  Nothing at
test time reads the reference; the JSON files written here are the vectors.

Fixtures (all data -- inputs and expected outputs):
  conformance.json   whole-path GFN2 goldens (tblite 0.7.0 / xTB 6.7.1 live
                     runs) from data/conformance/{inputs,golden}; gates
                     data/conformance/manifest.json:335-379 (5e-7 absolute).
  scc_traces.json    per-iteration SCC traces (patched-tblite observer) from
                     data/conformance/scc-traces: statics S/H0 and, per
                     iteration, the mixed input, raw Mulliken output, H, P and
                     free energy (replay contract of tests/scc_trace_replay.cpp).
  mixer_known_answer.json
                     the modified-Broyden known answer of
                     tests/scc_mixer_test.cpp:254-319 (memory 3, damping 0.4).
  gfn2_parameters_hcno.json
                     the H, C, N, O entries of the reference's parameter
                     headers data/parameters/gfn2.hpp (kElements, kShells,
                     kGlobal) and data/parameters/d4.hpp (kElements,
                     kReferences, kReferenceC6 among their references): the
                     elements of the generated workloads, read from a source
                     independent of the gfn2.json the tables are generated from.
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))
SYMBOLS = {"h": 1, "he": 2, "li": 3, "be": 4, "b": 5, "c": 6, "n": 7, "o": 8, "f": 9, "ne": 10,
           "na": 11, "mg": 12, "al": 13, "si": 14, "p": 15, "s": 16, "cl": 17, "ar": 18,
           "zn": 30, "i": 53, "rn": 86}


def read_coord(path):
    Z, X = [], []
    for line in open(path):
        t = line.split()
        if not t or t[0].startswith("$"):
            continue
        X.append([float(v) for v in t[:3]])
        Z.append(SYMBOLS[t[3].lower()])
    return Z, X


def conformance():
    manifest = json.load(open(os.path.join(REF, "data/conformance/manifest.json")))
    cases = []
    for case in manifest["cases"]:
        cid = case["id"]
        gold = json.load(open(os.path.join(REF, case["golden"])))
        props = gold["properties"]
        item = {"id": cid, "engine": case["reference_engine"],
                "charge": float(case["molecular_charge"]),
                "unpaired": int(case["unpaired_electrons"]),
                "energy": props["energy_hartree"],
                "forces": props["forces_hartree_per_bohr"]}
        inp = os.path.join(REF, case["input"])
        if inp.endswith(".qmmm.json"):
            q = json.load(open(inp))
            item["Z"] = q["qm"]["atomic_numbers"]
            item["xyz"] = q["qm"]["positions_bohr"]
            pc = q["external_point_charges"]
            item["pc_xyz"] = pc["positions_bohr"]
            item["pc_q"] = pc["charges_e"]
            item["pc_gamma"] = pc["gammas_hartree"]
            item["pc_forces"] = props["point_charge_forces_hartree_per_bohr"]
        else:
            item["Z"], item["xyz"] = read_coord(inp)
        if "partial_charges_e" in props:
            item["charges"] = props["partial_charges_e"]
        if "efield" in case:
            # uniform field in Hartree/(e bohr); the golden pins the energy only
            # (its analytic forces follow another convention; forces are checked
            # by finite differences of the energy)
            item["efield"] = case["efield"]
            item["oracle_properties"] = case.get("xtbloom_oracle_properties", ["energy_hartree"])
        cases.append(item)
    tol = manifest["tolerances"]
    return {"source": "data/conformance (reference)", "tolerances": {
        k: tol[k]["atol"] for k in ("energy", "forces", "charges", "point_charge_forces")},
        "cases": cases}


def flat(x):
    if isinstance(x, list):
        return [v for item in x for v in flat(item)]
    return [x]


def traces():
    root = os.path.join(REF, "data/conformance/scc-traces")
    index = json.load(open(os.path.join(root, "manifest.json")))
    out = []
    for cid, meta in sorted(index["cases"].items()):
        d = json.load(open(os.path.join(root, meta["path"])))
        inp = d["input"]
        item = {"id": cid, "Z": inp["atomic_numbers"], "xyz": flat(inp["positions"]),
                "charge": inp["molecular_charge"], "unpaired": inp["unpaired_electrons"],
                "temperature_kelvin": inp["temperature"],
                "overlap": flat(d["statics"]["overlap"][0]),
                "h0": flat(d["statics"]["core_hamiltonian"][0]),
                "terminal_iterations": d["terminal"]["iterations"], "iterations": []}
        pcs = inp.get("point_charges")
        if pcs:
            item["pc_xyz"] = flat(pcs["positions"])
            item["pc_q"] = pcs["charges"]
            item["pc_gamma"] = pcs["hardnesses"]
        for it in d["iterations"]:
            item["iterations"].append({
                "mixed": flat(it["mixed_qsh"][0]) + flat(it["mixed_dipoles"][0]) +
                flat(it["mixed_quadrupoles"][0]),
                "raw": flat(it["raw_qsh"][0]) + flat(it["raw_dipoles"][0]) +
                flat(it["raw_quadrupoles"][0]),
                "energy": it["energy"], "hamiltonian": flat(it["hamiltonian"][0]),
                "density": flat(it["density"][0])})
        out.append(item)
    return {"source": "data/conformance/scc-traces (reference)", "cases": out}


def traces_unrestricted():
    """The unrestricted (spin_channels = 2) traces of manifest-v2.json.  Fields
    keep the reference's packing: every multipole field holds the charge then
    the magnetization channel; H, P and eigenvalues hold alpha then beta.  H is
    the solver Hamiltonian, H0 + V_q +/- V_m (tblite doubles its half-converted
    assembly, so the assembled one is the same operator at half scale)."""
    root = os.path.join(REF, "data/conformance/scc-traces")
    index = json.load(open(os.path.join(root, "manifest-v2.json")))
    out = []
    for cid, meta in sorted(index["cases"].items()):
        d = json.load(open(os.path.join(root, meta["path"])))
        inp = d["input"]
        item = {"id": cid, "Z": inp["atomic_numbers"], "xyz": flat(inp["positions"]),
                "charge": inp["molecular_charge"], "unpaired": inp["unpaired_electrons"],
                "spin_channels": inp["spin_channels"], "temperature_kelvin": inp["temperature"],
                "overlap": flat(d["statics"]["overlap"][0]),
                "h0": flat(d["statics"]["core_hamiltonian"][0]),
                "terminal_iterations": d["terminal"]["iterations"], "iterations": []}
        for it in d["iterations"]:
            item["iterations"].append({
                "mixed": flat(it["mixed_qsh"]) + flat(it["mixed_dipoles"]) + flat(it["mixed_quadrupoles"]),
                "raw": flat(it["raw_qsh"]) + flat(it["raw_dipoles"]) + flat(it["raw_quadrupoles"]),
                "energy": it["energy"], "hamiltonian": flat(it["solver_hamiltonian"]),
                "density": flat(it["density"]), "eigenvalues": flat(it["eigenvalues"]),
                "occupations": flat(it["occupations"]), "residual_rms": it["residual_rms"]})
        out.append(item)
    return {"source": "data/conformance/scc-traces/manifest-v2.json (reference)", "cases": out}


def _cxx_rows(text, name):
    """Rows ``{...}`` (or ``Type{...}``) of ``name{{ ... }};`` as lists of numbers."""
    body = text[text.index(name + "{{") + len(name) + 2:]
    body = body[:body.index("}};")]
    return [[float(tok.strip().rstrip("u")) for tok in row.split(",") if tok.strip()]
            for row in re.findall(r"\{([^{}]*)\}", body)]


def parameters():
    gfn2 = open(os.path.join(REF, "data/parameters/gfn2.hpp")).read()
    d4 = open(os.path.join(REF, "data/parameters/d4.hpp")).read()
    elements = _cxx_rows(gfn2, "kElements")
    shells = _cxx_rows(gfn2, "kShells")
    d4el = _cxx_rows(d4, "kElements")
    d4ref = _cxx_rows(d4, "kReferences")
    c6body = d4[d4.index("kReferenceC6{{") + 14:]
    c6 = [float(x) for x in re.findall(r"[-+]?(?:\d+\.\d*|\d+)(?:[eE][-+]?\d+)?",
                                       c6body[:c6body.index("}};")])]
    nref = len(d4ref)
    assert len(elements) == 86 and len(shells) == 237 and len(d4el) == 86 and len(c6) == nref * nref
    out = {"source": "data/parameters/gfn2.hpp, data/parameters/d4.hpp (reference)", "elements": {}}
    for z in (1, 6, 7, 8):
        e = elements[z - 1]
        off, cnt = int(e[0]), int(e[2])
        assert int(e[1]) == z
        d = d4el[z - 1]
        r0, rn = int(d[0]), int(d[1])
        out["elements"][str(z)] = {
            "gam": e[3], "gam3": e[4], "zeff": e[5], "arep": e[6], "en": e[8],
            "dkernel": e[9], "qkernel": e[10], "mprad": e[11], "mpvcn": e[12],
            "shell_offset": off, "shell_count": cnt,
            "shells": [{"n": int(sh[0]), "l": int(sh[1]), "ngauss": int(sh[2]), "level_ev": sh[4],
                        "slater": sh[5], "refocc": sh[6], "kpoly": sh[7], "kcn_ev": sh[8],
                        "hubbard_scale": sh[9]} for sh in shells[off:off + cnt]],
            "d4": {"ref_offset": r0, "ref_count": rn, "covalent_radius": d[2], "en": d[3],
                   "zeff": d[4], "hardness": d[5], "r4r2": d[6],
                   "ref_cn": [d4ref[r][0] for r in range(r0, r0 + rn)],
                   "ref_q": [d4ref[r][1] for r in range(r0, r0 + rn)],
                   "ref_ngauss": [int(d4ref[r][2]) for r in range(r0, r0 + rn)]}}
    # C6 among the H/C/N/O references (row-major over all references)
    refs = [r for z in (1, 6, 7, 8) for r in range(int(d4el[z - 1][0]),
                                                   int(d4el[z - 1][0]) + int(d4el[z - 1][1]))]
    out["c6"] = [[i, j, c6[i * nref + j]] for i in refs for j in refs]
    return out


def mixer():
    text = open(os.path.join(REF, "tests/scc_mixer_test.cpp")).read()
    body = text[text.index("int test_tblite_trace_and_ring_history()"):]
    block = body[body.index("expected{{"):body.index("}};") + 3]
    nums = [float(x) for x in re.findall(r"[-+]?\d+\.\d+(?:[eE][-+]?\d+)?", block)]
    assert len(nums) == 60, len(nums)
    expected = [nums[i * 10:(i + 1) * 10] for i in range(6)]
    # Residual sequence and initial vector restated from the same test (lines 264-307).
    residuals = []
    for it in range(6):
        row = []
        for c in range(10):
            i = c + 1.0
            if it == 0:
                v = 0.01 * i
            elif it == 1:
                v = 0.008 * i + (0.003 if c % 2 == 0 else -0.003)
            elif it == 2:
                v = -0.004 * i + 0.002 * ((c % 3) - 1.0)
            elif it == 3:
                v = 0.003 * i - (0.0015 if c % 2 == 0 else -0.0015)
            elif it == 4:
                v = -0.002 * i + 0.0007 * ((c % 4) - 1.5)
            else:
                v = 0.001 * i + (0.0002 if c % 2 == 0 else -0.0002)
            row.append(v)
        residuals.append(row)
    return {"source": "tests/scc_mixer_test.cpp:254-319 (reference)", "memory": 3,
            "damping": 0.4, "tolerance": 8.0e-13,
            "initial": [0.05 * (c + 1) for c in range(10)], "residuals": residuals,
            "expected": expected}


def main():
    for name, data in (("conformance.json", conformance()), ("scc_traces.json", traces()),
                       ("scc_traces_unrestricted.json", traces_unrestricted()),
                       ("mixer_known_answer.json", mixer()),
                       ("gfn2_parameters_hcno.json", parameters())):
        path = os.path.join(OUT, name)
        with open(path, "w") as fh:
            json.dump(data, fh, separators=(",", ":"))
        print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
