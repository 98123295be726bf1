"""Data files of a Simulation: the `_dat.ndt` table of averages, the field
store and restart files.

* `_dat.ndt` follows the reference's writer (interface/nmag/main.py:2571-2639):
  a title line and a units line, each starting with '#', then one row per
  save; every column 25 characters wide ('% 25.13g', '% 25d', '% 25s'; titles
  '%-25s'); tensors spread over columns name_0, name_1, ... (nfem
  tensor2componentstrings, interface/nfem/main.py:503-580); units written as
  the SI dense strings '<A/m>', '<s>', '<1>' (interface/nsim/si_units/lib.py
  :304-341).
* The reference keeps fields and restart data in HDF5 (`_dat.h5`,
  `_restart.h5`; main.py:2641-2706, :2764-2885).  There is no HDF5 library
  here, so both are zip archives of .npy arrays (readable with numpy.load,
  no pickles): `<name>_dat.npz` gains one group of arrays per save, and
  `<name>_restart.npz` holds the fields and the clock (as JSON) of one
  moment, written to a temporary name and renamed, as the reference does.
"""
import io
import json
import os
import zipfile

import numpy as np

from .units import SI

_BASE = ("m", "kg", "s", "A")


def dens_str(q):
    """'<A/m>'-style unit string of an SI value or a plain number."""
    if isinstance(q, str):
        return q
    if not isinstance(q, SI):
        q = SI(q)

    def powers(positive):
        out = ""
        for name, p in zip(_BASE, q.dims):
            if p == 0 or (p > 0) != positive:
                continue
            out += name if abs(p) == 1 else "%s^%d" % (name, abs(p))
        return out

    pos, neg = powers(True), powers(False)
    units = pos + ("/" + neg if neg else "")
    if q.value in (1, 1.0):
        body = units if units else "1"
    else:
        body = "%g%s" % (q.value, units)
    return "<%s>" % body


def components(name, value, unit):
    """Split a scalar / vector / matrix entry into per-component columns."""
    arr = np.asarray(value, dtype=object if isinstance(value, str) else None)
    if arr.ndim == 0:
        return [(name, value, unit)]
    out = []
    for idx in np.ndindex(arr.shape):
        out.append(("%s_%s" % (name, "".join(str(i) for i in idx)), arr[idx].item(), unit))
    return out


def _fmt(v):
    if isinstance(v, (bool, np.bool_)):
        return "% 25d " % int(v)
    if isinstance(v, (int, np.integer)):
        return "% 25d " % v
    if isinstance(v, (float, np.floating)):
        return "% 25.13g " % v
    if isinstance(v, str):
        return "% 25s " % v
    raise ValueError("cannot write a value of type %s to an ndt file" % type(v))


def write_ndt_row(path, names, values, units, header):
    """Append one row (and, when `header`, the two title lines first)."""
    cols = []
    for n, v, u in zip(names, values, units):
        cols.extend(components(n, v, u))
    with open(path, "a") as f:
        if header:
            f.write("#" + "".join("%-25s " % c[0] for c in cols) + "\n")
            f.write("#" + "".join("%-25s " % dens_str(c[2]) for c in cols) + "\n")
        f.write("".join(_fmt(c[1]) for c in cols) + "\n")


def read_ndt(path):
    """(column names, units, rows) of an ndt file; numeric cells as floats."""
    with open(path) as f:
        lines = f.read().splitlines()
    names = lines[0][1:].split()
    units = lines[1][1:].split()
    rows = []
    for ln in lines[2:]:
        if not ln.strip():
            continue
        row = []
        for tok in ln.split():
            try:
                row.append(float(tok))
            except ValueError:
                row.append(tok)
        rows.append(row)
    return names, units, rows


def mesh_edges(tets):
    """Unique node pairs sharing a tetrahedron, (E, 2) with a < b."""
    t = np.asarray(tets, dtype=np.int64)
    pairs = np.concatenate([t[:, [a, b]] for a in range(4) for b in range(a + 1, 4)])
    pairs.sort(axis=1)
    return np.unique(pairs, axis=0)


def max_angle_deg(v, edges):
    """Largest angle in degrees between the vectors at the two ends of any
    mesh edge (the reference's maxangle column, nsimexec.ml:3477-3562)."""
    a, b = v[edges[:, 0]], v[edges[:, 1]]
    na, nb = np.linalg.norm(a, axis=1), np.linalg.norm(b, axis=1)
    c = np.einsum("ij,ij->i", a, b) / (na * nb)
    return float(np.degrees(np.arccos(np.clip(c, -1.0, 1.0)).max())) if len(c) else 0.0


# ------------------------------------------------------------- field store
def _put(zf, name, arr):
    buf = io.BytesIO()
    np.lib.format.write_array(buf, np.ascontiguousarray(arr), allow_pickle=False)
    zf.writestr(name + ".npy", buf.getvalue())


def append_fields(path, row, meta, fields, positions, tets):
    """Add the fields of one save (SI arrays by subfield name) as row `row`;
    the first save also stores the mesh (positions in m, tetrahedra)."""
    new = not os.path.exists(path)
    with zipfile.ZipFile(path, "a", compression=zipfile.ZIP_STORED) as zf:
        if new:
            _put(zf, "mesh/positions", np.asarray(positions, dtype=np.float64))
            _put(zf, "mesh/tets", np.asarray(tets, dtype=np.int32))
        for k, v in meta.items():
            _put(zf, "row%d/meta/%s" % (row, k), np.asarray(v))
        for k, v in fields.items():
            _put(zf, "row%d/field/%s" % (row, k), np.asarray(v, dtype=np.float64))


def _rows(z):
    rows = sorted({int(k.split("/")[0][3:]) for k in z.files if k.startswith("row")})
    return rows


def read_subfield(path, subfieldname, id=None, row=None):
    """SI array of a saved subfield: the save with clock id `id`, else the
    row-th save (default the first); the reference's get_subfield_from_h5file
    (interface/nmag/main.py:2915-2956)."""
    with np.load(path, allow_pickle=False) as z:
        rows = _rows(z)
        if id is not None:
            hit = [r for r in rows if "row%d/meta/id" % r in z.files
                   and int(z["row%d/meta/id" % r]) == id
                   and "row%d/field/%s" % (r, subfieldname) in z.files]
            if not hit:
                raise KeyError("no save with id %d holds '%s' in %s" % (id, subfieldname, path))
            r = hit[0]
        else:
            have = [r for r in rows if "row%d/field/%s" % (r, subfieldname) in z.files]
            idx = 0 if row is None else row
            if idx >= len(have):
                raise KeyError("'%s' was saved %d times in %s; row %d asked"
                               % (subfieldname, len(have), path, idx))
            r = have[idx]
        return np.array(z["row%d/field/%s" % (r, subfieldname)])


def read_meta(path):
    """List of the saves' clock values (dicts), in save order."""
    with np.load(path, allow_pickle=False) as z:
        out = []
        for r in _rows(z):
            pre = "row%d/meta/" % r
            out.append({k[len(pre):]: np.array(z[k]).item() for k in z.files
                        if k.startswith(pre)})
        return out


def read_positions(path):
    """Node positions (m) of the mesh a field store or restart file was saved on."""
    with np.load(path, allow_pickle=False) as z:
        return np.array(z["mesh/positions"])


# ----------------------------------------------------------------- restart
def write_restart(path, fields, clock, positions):
    tmp = path + ".internal"
    if os.path.exists(tmp):
        os.remove(tmp)
    with zipfile.ZipFile(tmp, "w", compression=zipfile.ZIP_STORED) as zf:
        _put(zf, "mesh/positions", np.asarray(positions, dtype=np.float64))
        for k, v in fields.items():
            _put(zf, "row0/field/%s" % k, np.asarray(v, dtype=np.float64))
        _put(zf, "clock", np.frombuffer(json.dumps(clock).encode(), dtype=np.uint8))
    os.replace(tmp, path)


def read_restart(path):
    """(fields by subfield name, clock dict, positions) of a restart file."""
    with np.load(path, allow_pickle=False) as z:
        fields = {k[len("row0/field/"):]: np.array(z[k]) for k in z.files
                  if k.startswith("row0/field/")}
        clock = json.loads(bytes(np.array(z["clock"])).decode())
        return fields, clock, np.array(z["mesh/positions"])
