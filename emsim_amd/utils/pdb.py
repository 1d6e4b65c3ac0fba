"""Local PDB / PDBx-mmCIF reading and biological-assembly expansion, with no
Biopython and no network.

Mirrors the atom-ingestion interface of ``emsim/utils/pdb.py``:
``read_atoms`` (``:73-118``), ``check_file_format`` (``:121-141``),
``read_symmetries`` (``:144-152``), ``read_mmcif_struct_oper_list``
(``:155-195``), ``read_pdb_remark_350`` (``:198-230``), ``apply_symmetry``
(``:233-263``) and ``build_biological_unit`` (``:52-56``).  The reference
parses with Biopython (absent here); this module reads the formats directly:

* PDB: ATOM / HETATM records of every MODEL (Biopython's
  ``structure.get_atoms()`` walks all models), coordinates from columns
  31-54, element from columns 77-78 or, when blank, from the atom name
  (columns 13-16) as the format specification describes.  For atoms with
  alternate locations the one with the highest occupancy is kept (first on
  ties), which is the location a Biopython disordered atom selects.
* mmCIF: the ``_atom_site`` loop (``type_symbol``, ``Cartn_x/y/z``,
  ``label_alt_id``, ``occupancy``, model and residue identifiers), with the
  same alternate-location rule; quoted values and ';' text fields are
  tokenised.
* Assemblies: REMARK 350 BIOMT records read from the same fixed columns as
  the reference, or ``_pdbx_struct_oper_list`` operators with numeric ids.

Coordinates keep the file's (x, y, z) column order, as the reference's
``read_atoms`` does.  The network download helpers of the reference
(``retrieve_pdb_file``, ``fetch_all_pdb_file``, ``retrieve_pdb_files``) are
out of scope (no network).
"""
import shlex
from pathlib import Path
from typing import Dict, List, Optional, Tuple, Union

import numpy as np

from .. import atoms as atm
from .. import elem


def check_file_format(pdb_file: Union[str, Path]) -> Tuple[str, str]:
    """('pdb' | 'mmcif', id code) from the file name, as the reference does;
    a file whose name says neither is recognised by its first record."""
    p = Path(pdb_file)
    code, fmt = p.stem, ""
    if p.stem.startswith("pdb") and p.suffix == ".ent":
        code, fmt = code[3:], "pdb"
    elif p.suffix == ".pdb":
        fmt = "pdb"
    elif p.suffix == ".cif":
        fmt = "mmcif"
    else:
        with open(p) as fh:
            head = fh.read(4096)
        fmt = "mmcif" if head.lstrip().startswith("data_") else "pdb"
    return fmt, code


# ---------------------------------------------------------------------------
# PDB
# ---------------------------------------------------------------------------
def _pdb_element(line: str) -> Optional[int]:
    sym = line[76:78].strip() if len(line) >= 78 else ""
    if not sym:
        # columns 13-14 hold the element right-justified in a 4-character
        # atom name: " CA " is carbon, "FE  " is iron, "1HB " hydrogen
        name = line[12:16]
        two = name[:2].strip()
        if len(two) == 2 and two.isalpha() and elem.number(two) is not None and name[0] != " ":
            sym = two
        else:
            sym = next((ch for ch in name if ch.isalpha()), "")
    return elem.number(sym) if sym else None


def _read_pdb(path: Path):
    atoms: Dict[tuple, tuple] = {}
    order: List[tuple] = []
    model = 0
    with open(path) as fh:
        for line in fh:
            rec = line[:6]
            if rec.startswith("MODEL"):
                try:
                    model = int(line[10:14])
                except ValueError:
                    model += 1
                continue
            if rec not in ("ATOM  ", "HETATM"):
                continue
            z = _pdb_element(line)
            if z is None:
                continue
            alt = line[16]
            key = (model, line[21], line[22:26], line[26], line[17:20], line[12:16])
            xyz = (float(line[30:38]), float(line[38:46]), float(line[46:54]))
            try:
                occ = float(line[54:60])
            except ValueError:
                occ = 1.0
            _keep(atoms, order, key, alt, occ, z, xyz)
    return atoms, order


def _keep(atoms, order, key, alt, occ, z, xyz):
    """One atom per key; among alternate locations the highest occupancy."""
    cur = atoms.get(key)
    if cur is None:
        atoms[key] = (occ, z, xyz)
        order.append(key)
    elif alt not in (" ", ".", "?", "") and occ > cur[0]:
        atoms[key] = (occ, z, xyz)


def read_pdb_remark_350(pdb_file: Union[str, Path]) -> np.ndarray:
    """(n, 3, 4) BIOMT operators (rotation | translation); the identity when
    the file has none."""
    ops = []
    with open(pdb_file) as fh:
        for line in fh:
            line = line.strip()
            if line.startswith("REMARK 350") and line[13:18] == "BIOMT":
                ops.append([float(x) for x in (line[24:33], line[34:43], line[44:53], line[58:68])])
    if not ops:
        return np.hstack((np.identity(3), np.zeros((3, 1))))[None]
    return np.array(ops, dtype=np.float64).reshape(-1, 3, 4)


# ---------------------------------------------------------------------------
# mmCIF
# ---------------------------------------------------------------------------
def _cif_tokens(text: str):
    """CIF tokens: bare words, '...' / "..." strings and ;-delimited text fields."""
    lines = text.splitlines()
    i = 0
    while i < len(lines):
        line = lines[i]
        if line.startswith(";"):
            buf = [line[1:]]
            i += 1
            while i < len(lines) and not lines[i].startswith(";"):
                buf.append(lines[i])
                i += 1
            yield "\n".join(buf).strip()
            i += 1
            continue
        s = line.split("#", 1)[0] if "'" not in line and '"' not in line else line
        if s.strip():
            lex = shlex.shlex(s, posix=True)
            lex.whitespace_split = True
            lex.commenters = "#"
            for tok in lex:
                yield tok
        i += 1


def _cif_tables(path: Path) -> Dict[str, Dict[str, List[str]]]:
    """category -> item -> values, from loops and key-value pairs."""
    tables: Dict[str, Dict[str, List[str]]] = {}
    toks = list(_cif_tokens(Path(path).read_text()))
    i = 0
    while i < len(toks):
        t = toks[i]
        if t == "loop_":
            i += 1
            items = []
            while i < len(toks) and toks[i].startswith("_"):
                items.append(toks[i])
                i += 1
            vals = []
            while i < len(toks) and not toks[i].startswith("_") and toks[i] != "loop_" and \
                    not toks[i].startswith("data_"):
                vals.append(toks[i])
                i += 1
            if items:
                n = len(items)
                for j, item in enumerate(items):
                    cat, key = item.split(".", 1)
                    tables.setdefault(cat, {})[key] = vals[j::n]
            continue
        if t.startswith("_") and "." in t and i + 1 < len(toks):
            cat, key = t.split(".", 1)
            tables.setdefault(cat, {})[key] = [toks[i + 1]]
            i += 2
            continue
        i += 1
    return tables


def _read_mmcif(path: Path):
    site = _cif_tables(path).get("_atom_site", {})
    if not site:
        return {}, []
    n = len(site["Cartn_x"])
    col = lambda k, d: site.get(k, [d] * n)
    groups, syms = col("group_PDB", "ATOM"), col("type_symbol", "")
    names = site.get("label_atom_id", site.get("auth_atom_id", [""] * n))
    alts, occs = col("label_alt_id", "."), col("occupancy", "1")
    models = col("pdbx_PDB_model_num", "1")
    chains = site.get("auth_asym_id", site.get("label_asym_id", ["A"] * n))
    seqs = site.get("auth_seq_id", site.get("label_seq_id", ["."] * n))
    icodes = col("pdbx_PDB_ins_code", "?")
    resn = site.get("label_comp_id", site.get("auth_comp_id", [""] * n))
    atoms: Dict[tuple, tuple] = {}
    order: List[tuple] = []
    for i in range(n):
        if groups[i] not in ("ATOM", "HETATM"):
            continue
        z = elem.number(syms[i])
        if z is None:
            continue
        try:
            occ = float(occs[i])
        except ValueError:
            occ = 1.0
        key = (models[i], chains[i], seqs[i], icodes[i], resn[i], names[i])
        xyz = (float(site["Cartn_x"][i]), float(site["Cartn_y"][i]), float(site["Cartn_z"][i]))
        _keep(atoms, order, key, alts[i], occ, z, xyz)
    return atoms, order


def read_mmcif_struct_oper_list(pdb_file: Union[str, Path]) -> np.ndarray:
    """(n, 3, 4) operators of ``_pdbx_struct_oper_list`` with numeric ids
    (the 'P' and 'X0' frame operators are skipped, as in the reference)."""
    ops = _cif_tables(Path(pdb_file)).get("_pdbx_struct_oper_list", {})
    if not ops:
        return np.hstack((np.identity(3), np.zeros((3, 1))))[None]
    keep = [i for i, d in enumerate(ops["id"]) if d.strip().isdigit()]
    out = np.zeros((len(keep), 3, 4))
    for r in range(3):
        for c in range(3):
            vals = ops[f"matrix[{r + 1}][{c + 1}]"]
            out[:, r, c] = [float(vals[i]) for i in keep]
        vals = ops[f"vector[{r + 1}]"]
        out[:, r, 3] = [float(vals[i]) for i in keep]
    return out


# ---------------------------------------------------------------------------
# public interface
# ---------------------------------------------------------------------------
def read_atoms(pdb_file: Union[str, Path], identifier: Optional[str] = None) -> atm.AtomList:
    """Atoms of a PDB (.pdb / pdbXXXX.ent) or mmCIF (.cif) file: every
    model, elements the library knows, one location per disordered atom."""
    path = Path(pdb_file)
    fmt, code = check_file_format(path)
    atoms, order = _read_pdb(path) if fmt == "pdb" else _read_mmcif(path)
    z = np.array([atoms[k][1] for k in order], dtype=np.int64)
    xyz = np.array([atoms[k][2] for k in order], dtype=np.float64).reshape(-1, 3)
    return atm.AtomList(elements=z, coordinates=xyz, label=identifier or code)


def read_symmetries(pdb_file: Union[str, Path]) -> np.ndarray:
    fmt, _ = check_file_format(pdb_file)
    return read_pdb_remark_350(pdb_file) if fmt == "pdb" else read_mmcif_struct_oper_list(pdb_file)


def apply_symmetry(atom_list: atm.AtomList, operations: np.ndarray) -> atm.AtomList:
    """Every operator (rotation | translation, shape (n, 3, 4)) applied to
    every atom; atoms of operator k follow those of operator k - 1."""
    ops = np.asarray(operations, dtype=np.float64)
    xyz = np.asarray(atom_list.coordinates, dtype=np.float64)
    out = np.einsum("kij,aj->kai", ops[:, :, :3], xyz) + ops[:, None, :, 3]
    return atm.AtomList(elements=np.tile(atom_list.elements, len(ops)),
                        coordinates=out.reshape(-1, 3), label=atom_list.label)


def build_biological_unit(pdb_file: Union[str, Path]) -> atm.AtomList:
    return apply_symmetry(read_atoms(pdb_file), read_symmetries(pdb_file))
