"""UPF pseudopotential files for the norm-conserving species of this path.

Reading: UPF version 1 (positional PP_HEADER, the format of the reference's
examples/data/pseudopotentials/nc/Si.pz-vbc.UPF) and version 2 (attribute PP_HEADER), both
norm-conserving; ultrasoft and PAW files are refused with UnsupportedPseudopotentialType, as
the path has S = I.  The result is the species dict the projector builder takes
(KPoint.kb_projectors): radial mesh r, rab, per-projector angular momentum l and cutoff index,
beta_r (r * beta(r), nbeta x nmesh), the D matrix dion_ry (Ry), z_val, and the local part
vloc_ry on the mesh.  What the reader takes from the file follows the reference's UPF reader
(io/pseudopotential/upf_reader/v2/init.hpp:60-230): the header's z_valence, mesh_size,
number_of_proj and l_max are required, a missing or unparseable one is an error that names the
attribute and the file; PP_BETA.i gives angular_momentum and cutoff_radius_index (default:
the whole mesh) and is zero-padded to the mesh; PP_DIJ is the full nproj x nproj matrix in Ry.

Writing: version 2, norm-conserving, for the synthetic species (workloads.synthetic_species):
values in 17 significant digits, so a written file reads back bit-exactly.  Its PP_LOCAL is
-z_val erf(r / (sqrt(2) r_loc)) / r, the real-space function whose transform is the synthetic
local form factor -4 pi z_val exp(-G^2 r_loc^2 / 2) / (Omega G^2) (workloads.local_potential).
"""
from __future__ import annotations

import math
import re

import numpy as np


class UPFError(ValueError):
    """A UPF file that cannot be read: the message names the attribute or section and the file."""


class UnsupportedPseudopotentialType(UPFError):
    """A readable file of a type this path does not take (ultrasoft, PAW)."""


_NUM = re.compile(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eEdD][-+]?\d+)?")


def _numbers(text: str) -> np.ndarray:
    return np.array([float(x.replace("D", "E").replace("d", "e")) for x in _NUM.findall(text)])


def _attrs(tag_text: str) -> dict:
    """Attributes of one start tag: name="value" pairs (names may follow newlines or tabs)."""
    return {m.group(1): m.group(2).strip() for m in re.finditer(r'([A-Za-z_][\w.]*)\s*=\s*"([^"]*)"', tag_text)}


def _section(content: str, name: str, source: str, required: bool = True):
    """(start-tag text, body) of <name ...>body</name>, or None when absent and not required."""
    m = re.search(r"<" + re.escape(name) + r"(\s[^>]*)?>", content)
    if not m:
        if required:
            raise UPFError(f"{source}: section <{name}> is missing")
        return None
    if m.group(0).endswith("/>"):
        return m.group(0), ""
    end = content.find("</" + name + ">", m.end())
    if end < 0:
        raise UPFError(f"{source}: section <{name}> is not closed")
    return m.group(0), content[m.end():end]


def _required_int(attrs: dict, name: str, source: str) -> int:
    if name not in attrs or attrs[name] == "":
        raise UPFError(f"{source}: PP_HEADER attribute {name} is missing")
    v = attrs[name]
    if not re.fullmatch(r"[-+]?\d+", v):
        raise UPFError(f"{source}: PP_HEADER attribute {name}=\"{v}\" is not an integer")
    return int(v)


def _required_real(attrs: dict, name: str, source: str) -> float:
    if name not in attrs or attrs[name] == "":
        raise UPFError(f"{source}: PP_HEADER attribute {name} is missing")
    v = attrs[name].replace("D", "E").replace("d", "e")
    try:
        return float(v)
    except ValueError:
        raise UPFError(f"{source}: PP_HEADER attribute {name}=\"{attrs[name]}\" is not a number") from None


def _check_type(kind: str, source: str):
    k = kind.strip().upper()
    if k in ("US", "USPP", "PAW"):
        raise UnsupportedPseudopotentialType(f"{source}: pseudo_type {kind.strip()} (only norm-conserving is taken)")
    if k not in ("NC", "SL", "1/R", ""):
        raise UPFError(f"{source}: unknown pseudo_type {kind.strip()}")


def read_upf(path: str) -> dict:
    """Species dict of a norm-conserving UPF file (version 1 or 2)."""
    try:
        with open(path) as f:
            content = f.read()
    except OSError as e:
        raise UPFError(f"{path}: cannot be read ({e.strerror})") from None
    return parse_upf(content, path)


def parse_upf(content: str, source: str = "<string>") -> dict:
    head = _section(content, "PP_HEADER", source)
    if "=" in head[0]:
        return _parse_v2(content, head[0], source)
    return _parse_v1(content, head[1], source)


def _parse_v2(content: str, header_tag: str, source: str) -> dict:
    h = _attrs(header_tag)
    _check_type(h.get("pseudo_type", ""), source)
    z = _required_real(h, "z_valence", source)
    mesh = _required_int(h, "mesh_size", source)
    nproj = _required_int(h, "number_of_proj", source)
    lmax = _required_int(h, "l_max", source)
    mesh_attr = _attrs(_section(content, "PP_MESH", source)[0]).get("mesh", "")
    if re.fullmatch(r"\d+", mesh_attr):  # PP_MESH's own point count, when given, wins
        mesh = int(mesh_attr)
    r = _numbers(_section(content, "PP_R", source)[1])
    rab = _numbers(_section(content, "PP_RAB", source)[1])
    vloc = _numbers(_section(content, "PP_LOCAL", source)[1])
    if len(r) < mesh or len(rab) < mesh:
        raise UPFError(f"{source}: PP_R / PP_RAB hold fewer than mesh_size={mesh} values")
    ls, cuts, betas = [], [], []
    for i in range(1, nproj + 1):
        sec = _section(content, f"PP_BETA.{i}", source, required=False)
        if sec is None:
            raise UPFError(f"{source}: section <PP_BETA.{i}> is missing (number_of_proj={nproj})")
        a = _attrs(sec[0])
        ls.append(int(a.get("angular_momentum", "0") or 0))
        cut = a.get("cutoff_radius_index", "")
        cuts.append(int(cut) if cut.strip().lstrip("+-").isdigit() else mesh)
        v = _numbers(sec[1])[:mesh]
        betas.append(np.pad(v, (0, mesh - len(v))))
    dion = np.zeros((nproj, nproj))
    if nproj:
        d = _numbers(_section(content, "PP_DIJ", source)[1])
        if len(d) < nproj * nproj:
            raise UPFError(f"{source}: PP_DIJ holds {len(d)} values for {nproj} x {nproj}")
        dion = d[:nproj * nproj].reshape(nproj, nproj)
    if ls and max(ls) > lmax:
        raise UPFError(f"{source}: a projector has l = {max(ls)} above l_max = {lmax}")
    return {"r": r[:mesh], "rab": rab[:mesh], "l": ls, "cutoff": cuts,
            "beta_r": np.array(betas).reshape(nproj, mesh), "dion_ry": dion, "z_val": z,
            "vloc_ry": vloc[:mesh], "element": h.get("element", "").strip(), "version": 2}


def _parse_v1(content: str, header: str, source: str) -> dict:
    lines = [ln.strip() for ln in header.strip().splitlines() if ln.strip()]
    # version, element, type, core correction, functional, Z valence, total energy, cutoffs,
    # l_max, mesh points, "wavefunctions projectors", then the wavefunction table
    if len(lines) < 11:
        raise UPFError(f"{source}: PP_HEADER (version 1) has {len(lines)} lines, fewer than 11")
    element = lines[1].split()[0]
    _check_type(lines[2].split()[0], source)
    try:
        z = float(lines[5].split()[0].replace("D", "E"))
        mesh = int(lines[9].split()[0])
        nproj = int(lines[10].split()[1])
    except (ValueError, IndexError):
        raise UPFError(f"{source}: PP_HEADER (version 1): Z valence, mesh or projector count unreadable") from None
    r = _numbers(_section(content, "PP_R", source)[1])[:mesh]
    rab = _numbers(_section(content, "PP_RAB", source)[1])[:mesh]
    vloc = _numbers(_section(content, "PP_LOCAL", source)[1])[:mesh]
    nonlocal_body = _section(content, "PP_NONLOCAL", source)[1]
    ls, cuts, betas = [], [], []
    for m in re.finditer(r"<PP_BETA>(.*?)</PP_BETA>", nonlocal_body, re.S):
        block = m.group(1).strip().splitlines()
        idx_l = block[0].split()
        npts = int(block[1].split()[0])
        ls.append(int(idx_l[1]))
        cuts.append(npts)
        v = _numbers("\n".join(block[2:]))[:npts]
        betas.append(np.pad(v, (0, mesh - len(v))))
    if len(betas) != nproj:
        raise UPFError(f"{source}: {len(betas)} PP_BETA sections for {nproj} projectors")
    dion = np.zeros((nproj, nproj))
    dij = re.search(r"<PP_DIJ>(.*?)</PP_DIJ>", nonlocal_body, re.S)
    if nproj and dij:
        rows = [ln.split() for ln in dij.group(1).strip().splitlines()]
        nnz = int(rows[0][0])
        for ln in rows[1:1 + nnz]:
            i, j, v = int(ln[0]) - 1, int(ln[1]) - 1, float(ln[2].replace("D", "E"))
            dion[i, j] = dion[j, i] = v
    return {"r": r, "rab": rab, "l": ls, "cutoff": cuts, "beta_r": np.array(betas).reshape(nproj, mesh),
            "dion_ry": dion, "z_val": z, "vloc_ry": vloc, "element": element, "version": 1}


def _block(values, columns: int = 4) -> str:
    v = np.asarray(values, dtype=np.float64).ravel()
    rows = [" ".join(f"{x: .16e}" for x in v[i:i + columns]) for i in range(0, len(v), columns)]
    return "\n".join(rows)


def write_upf_v2(path: str, species: dict, element: str = "Si", info: str = "") -> None:
    """Writes a norm-conserving species dict (workloads.synthetic_species) as UPF version 2."""
    r = np.asarray(species["r"], np.float64)
    rab = np.asarray(species["rab"], np.float64)
    n = len(r)
    beta = np.asarray(species["beta_r"], np.float64).reshape(len(species["l"]), n)
    dion = np.asarray(species["dion_ry"], np.float64)
    z = float(species["z_val"])
    if "vloc_ry" in species:
        vloc = np.asarray(species["vloc_ry"], np.float64)
    else:
        rl = float(species["r_loc"])
        vloc = np.array([-z * math.erf(x / (math.sqrt(2.0) * rl)) / x for x in r])
    nproj = len(species["l"])
    out = ['<UPF version="2.0.1">',
           "<PP_INFO>", info or "Synthetic norm-conserving species: Gaussian local part, Gaussian s and p "
           "projectors (r^l exp(-r^2 / 2 sigma^2)).", "</PP_INFO>",
           f'<PP_HEADER generated="q3_amd.upf" element="{element}" pseudo_type="NC" relativistic="no" '
           f'is_ultrasoft="F" is_paw="F" core_correction="F" functional="PZ" z_valence="{z!r}" '
           f'l_max="{max(species["l"]) if nproj else 0}" mesh_size="{n}" number_of_wfc="0" '
           f'number_of_proj="{nproj}"/>',
           "<PP_MESH>",
           f'<PP_R type="real" size="{n}" columns="4">', _block(r), "</PP_R>",
           f'<PP_RAB type="real" size="{n}" columns="4">', _block(rab), "</PP_RAB>",
           "</PP_MESH>",
           f'<PP_LOCAL type="real" size="{n}" columns="4">', _block(vloc), "</PP_LOCAL>",
           "<PP_NONLOCAL>"]
    for i in range(nproj):
        out += [f'<PP_BETA.{i + 1} type="real" size="{n}" columns="4" index="{i + 1}" '
                f'angular_momentum="{species["l"][i]}" cutoff_radius_index="{species["cutoff"][i]}">',
                _block(beta[i]), f"</PP_BETA.{i + 1}>"]
    out += [f'<PP_DIJ type="real" size="{nproj * nproj}" columns="4">', _block(dion), "</PP_DIJ>",
            "</PP_NONLOCAL>", "</UPF>", ""]
    with open(path, "w") as f:
        f.write("\n".join(out))
