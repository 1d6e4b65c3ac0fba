"""Local PDB / mmCIF reader and assembly expansion (emsim_amd/utils/pdb.py).

The reference reads structures through Biopython, which is not installed
here, so its behaviour is pinned by hand-made fixtures whose content is known
exactly: every model, one location per disordered atom (highest occupancy),
elements from the element columns or the atom name, BIOMT / oper_list
operators (frame operators skipped), and the same structure in both formats.
"""
import numpy as np
import pytest

from emsim_amd.utils import pdb

ATOMS = [
    # rec, name, alt, res, chain, seq, x, y, z, occ, element, model
    ("ATOM", " N  ", " ", "ALA", "A", 1, 1.0, 2.0, 3.0, 1.00, "N", 1),
    ("ATOM", " CA ", "A", "ALA", "A", 1, 2.0, 3.0, 4.0, 0.40, "C", 1),
    ("ATOM", " CA ", "B", "ALA", "A", 1, 2.5, 3.5, 4.5, 0.60, "C", 1),
    ("ATOM", " CB ", " ", "ALA", "A", 1, 3.0, 4.0, 5.0, 1.00, "", 1),     # element from name
    ("HETATM", "FE  ", " ", "HEM", "A", 2, 4.0, 5.0, 6.0, 1.00, "", 1),   # two-letter from name
    ("HETATM", " O  ", " ", "HOH", "A", 3, 5.0, 6.0, 7.0, 1.00, "O", 1),
    ("ATOM", " N  ", " ", "ALA", "A", 1, 1.1, 2.1, 3.1, 1.00, "N", 2),
]
EXPECT_Z = [7, 6, 6, 26, 8, 7]
EXPECT_XYZ = [(1, 2, 3), (2.5, 3.5, 4.5), (3, 4, 5), (4, 5, 6), (5, 6, 7), (1.1, 2.1, 3.1)]
OPS = [np.hstack([np.eye(3), np.zeros((3, 1))]),
       np.array([[-1, 0, 0, 10.0], [0, -1, 0, 0], [0, 0, 1, 5.0]])]


def _pdb_text():
    lines = ["HEADER    TEST STRUCTURE"]
    for k, op in enumerate(OPS, 1):
        for r in range(3):
            lines.append(f"REMARK 350   BIOMT{r + 1}{k:>4}{op[r, 0]:10.6f}{op[r, 1]:10.6f}"
                         f"{op[r, 2]:10.6f}{op[r, 3]:15.5f}")
    model = None
    for i, (rec, name, alt, res, ch, seq, x, y, z, occ, el, mdl) in enumerate(ATOMS, 1):
        if mdl != model:
            if model is not None:
                lines.append("ENDMDL")
            lines.append(f"MODEL     {mdl:>4}")
            model = mdl
        lines.append(f"{rec:<6}{i:>5} {name:<4}{alt}{res:>3} {ch}{seq:>4}    "
                     f"{x:8.3f}{y:8.3f}{z:8.3f}{occ:6.2f}{0.0:6.2f}          {el:>2}")
    lines += ["ENDMDL", "END"]
    return "\n".join(lines) + "\n"


def _cif_text():
    head = ["data_TEST", "#", "loop_"] + [f"_atom_site.{k}" for k in (
        "group_PDB", "id", "type_symbol", "label_atom_id", "label_alt_id", "label_comp_id",
        "auth_asym_id", "auth_seq_id", "pdbx_PDB_ins_code", "Cartn_x", "Cartn_y", "Cartn_z",
        "occupancy", "pdbx_PDB_model_num")]
    syms = ["N", "C", "C", "C", "FE", "O", "N"]
    rows = []
    for i, ((rec, name, alt, res, ch, seq, x, y, z, occ, _, mdl), s) in enumerate(
            zip(ATOMS, syms), 1):
        a = alt.strip() or "."
        rows.append(f"{rec} {i} {s} {name.strip()} {a} {res} {ch} {seq} ? {x:.3f} {y:.3f} "
                    f"{z:.3f} {occ:.2f} {mdl}")
    rows.append("HETATM 8 ZZ \"C1'\" . XXX B 9 ? 0 0 0 1.00 1")   # unknown element: skipped
    ops = ["#", "loop_"] + [f"_pdbx_struct_oper_list.{k}" for k in (
        ["id", "type", "name"] + [f"matrix[{r}][{c}]" for r in (1, 2, 3) for c in (1, 2, 3)] +
        [f"vector[{r}]" for r in (1, 2, 3)])]
    def row(oid, typ, op):
        m = " ".join(f"{op[r, c]:g}" for r in range(3) for c in range(3))
        v = " ".join(f"{op[r, 3]:g}" for r in range(3))
        return f"{oid} '{typ}' ? {m} {v}"
    ops += [row("1", "identity operation", OPS[0]), row("2", "point symmetry operation", OPS[1]),
            row("P", "transform to point frame", np.hstack([2 * np.eye(3), np.ones((3, 1))]))]
    return "\n".join(head + rows + ops + ["#"]) + "\n"


@pytest.fixture(params=["pdb", "cif"])
def structure(request, tmp_path):
    if request.param == "pdb":
        p = tmp_path / "pdb1abc.ent"
        p.write_text(_pdb_text())
    else:
        p = tmp_path / "1abc.cif"
        p.write_text(_cif_text())
    return p


def test_read_atoms(structure):
    fmt, code = pdb.check_file_format(structure)
    assert code == "1abc" and fmt in ("pdb", "mmcif")
    al = pdb.read_atoms(structure)
    assert al.elements.tolist() == EXPECT_Z
    np.testing.assert_allclose(al.coordinates, EXPECT_XYZ, atol=1e-6)
    assert al.label == "1abc"


def test_symmetries_and_biological_unit(structure):
    ops = pdb.read_symmetries(structure)
    assert ops.shape == (2, 3, 4)
    np.testing.assert_allclose(ops, np.stack(OPS))
    unit = pdb.build_biological_unit(structure)
    n = len(EXPECT_Z)
    assert len(unit) == 2 * n
    xyz = np.asarray(EXPECT_XYZ, dtype=np.float64)
    np.testing.assert_allclose(unit.coordinates[:n], xyz, atol=1e-5)
    mirrored = np.stack([10 - xyz[:, 0], -xyz[:, 1], xyz[:, 2] + 5], axis=1)
    np.testing.assert_allclose(unit.coordinates[n:], mirrored, atol=1e-5)
    assert unit.elements.tolist() == EXPECT_Z * 2


def test_no_operators_means_identity(tmp_path):
    p = tmp_path / "x.pdb"
    p.write_text("\n".join(l for l in _pdb_text().splitlines() if "BIOMT" not in l) + "\n")
    ops = pdb.read_symmetries(p)
    assert ops.shape == (1, 3, 4)
    np.testing.assert_array_equal(ops[0], OPS[0])
