"""Reader of EXP phase-space (PSP "OUT.") files, io/psp_io.py:34-230.

File layout (little-endian):
  master header  time f8, nbodies_tot u32, ncomp u32                 (16 B)
  per component  [magic u32 0xADBFABC4, rsize u32]  (float files only)
                 nbodies u32, niatr u32, ndatr u32, infostringlen u32
                 info string (infostringlen bytes, NUL padded)
                 nbodies records: [id i8 if indexing] m x y z vx vy vz pot
                 (f4 or f8) + niatr i4 + ndatr floats
The magic number selects float (f4) records, its absence double records
(io/psp_io.py:176-189).  The info string is YAML in current EXP files; older
files carry "name : force id : component params : force params"
(tests/initial001dump.dat), for which indexing is off.

Differences from the reference reader: old-style info strings are parsed
instead of failing in yaml, and component data is read with one np.fromfile
of a structured dtype.
"""
from __future__ import annotations

import numpy as np
import yaml

PSP_FLOAT_MAGIC = 0xADBFABC4   # 2915019716


def _parse_params(text):
    """'a=1, b=x, c=0.1' -> dict with numbers converted where possible."""
    out = {}
    for item in text.split(","):
        if "=" not in item:
            continue
        k, v = (s.strip() for s in item.split("=", 1))
        for conv in (int, float):
            try:
                v = conv(v)
                break
            except ValueError:
                pass
        out[k] = v
    return out


def parse_info_string(raw):
    """Component header dict from the info string (YAML or old style)."""
    text = raw.split(b"\x00", 1)[0].decode(errors="replace")
    try:
        d = yaml.safe_load(text)
    except yaml.YAMLError:
        d = None
    if isinstance(d, dict) and "name" in d:
        d.setdefault("parameters", {})
        if "indexing" not in d["parameters"]:
            d["parameters"]["indexing"] = str(d.get("indexing", "false")).lower() == "true"
        return d
    # "name : force id : component parameters : force parameters"
    parts = [p.strip() for p in text.split(":")]
    d = {"name": parts[0], "id": parts[1] if len(parts) > 1 else "",
         "parameters": _parse_params(parts[2]) if len(parts) > 2 else {},
         "force": {"id": parts[1] if len(parts) > 1 else "",
                   "parameters": _parse_params(parts[3]) if len(parts) > 3 else {}}}
    d["parameters"]["indexing"] = False
    d["style"] = "old"
    return d


class Input:
    """io/psp_io.py:34-101: header of every component; .data of `comp`
    (dict of id?, m, x, y, z, vx, vy, vz, potE, i_attr*, f_attr*)."""

    def __init__(self, filename, comp=None, verbose=0):
        self.filename = filename
        self.verbose = verbose
        self.header = {}
        self.comp = comp
        with open(filename, "rb") as f:
            self.time = float(np.fromfile(f, dtype="<f8", count=1)[0])
            self._nbodies_tot, self._ncomp = (int(v) for v in np.fromfile(f, dtype="<u4", count=2))
            f.seek(16)
            magic = np.fromfile(f, dtype="<u4", count=1)
            self._float_len = 4 if (magic.size and magic[0] == PSP_FLOAT_MAGIC) else 8
            pos = 16
            for _ in range(self._ncomp):
                f.seek(pos)
                pos = self._component_header(f)
        if comp is not None:
            if comp not in self.header:
                raise IOError("The specified component does not exist.")
            h = self.header[comp]
            self.data = self._component_data(h)

    def _component_header(self, f):
        if self._float_len == 4:
            v = np.fromfile(f, dtype="<u4", count=6)[2:]
        else:
            v = np.fromfile(f, dtype="<u4", count=4)
        nbodies, niatr, ndatr, ninfo = (int(a) for a in v)
        h = parse_info_string(f.read(ninfo))
        start = f.tell()
        idx = bool(h["parameters"].get("indexing", False))
        rec = 8 * idx + self._float_len * 8 + 4 * niatr + self._float_len * ndatr
        h.update(nint_attr=niatr, nfloat_attr=ndatr, nbodies=nbodies, data_start=start,
                 data_end=start + nbodies * rec)
        self.header[h["name"]] = h
        self.indexing = idx
        return h["data_end"]

    def _component_data(self, h):
        fl = "<f4" if self._float_len == 4 else "<f8"
        fields = []
        if h["parameters"].get("indexing", False):
            fields.append(("id", "<i8"))
        fields += [(k, fl) for k in ("m", "x", "y", "z", "vx", "vy", "vz", "potE")]
        fields += [(f"i_attr{i}", "<i4") for i in range(h["nint_attr"])]
        fields += [(f"f_attr{i}", fl) for i in range(h["nfloat_attr"])]
        dt = np.dtype(fields)
        with open(self.filename, "rb") as f:
            f.seek(h["data_start"])
            rec = np.fromfile(f, dtype=dt, count=h["nbodies"])
        if rec.size != h["nbodies"]:
            raise IOError(f"{self.filename}: component {h['name']} is truncated")
        return {name: np.array(rec[name]) for name in dt.names}


def write_psp(filename, time, comps, float_len=4):
    """Write a PSP file.  comps: list of (info_string_or_dict, data dict with
    m, x, y, z, vx, vy, vz, potE and optionally id).  Dict headers are written
    as YAML (with parameters.indexing set from the presence of 'id')."""
    fl = "<f4" if float_len == 4 else "<f8"
    with open(filename, "wb") as f:
        ntot = sum(int(np.asarray(d["x"]).size) for _, d in comps)
        np.array([time], dtype="<f8").tofile(f)
        np.array([ntot, len(comps)], dtype="<u4").tofile(f)
        for info, d in comps:
            n = int(np.asarray(d["x"]).size)
            idx = "id" in d
            if isinstance(info, dict):
                info = dict(info)
                info.setdefault("parameters", {})["indexing"] = idx
                text = yaml.safe_dump(info).encode()
            else:
                text = info.encode()
            ninfo = max(len(text) + 1, 1024)
            if float_len == 4:
                np.array([PSP_FLOAT_MAGIC, 4], dtype="<u4").tofile(f)
            np.array([n, 0, 0, ninfo], dtype="<u4").tofile(f)
            f.write(text + b"\x00" * (ninfo - len(text)))
            fields = ([("id", "<i8")] if idx else []) + \
                [(k, fl) for k in ("m", "x", "y", "z", "vx", "vy", "vz", "potE")]
            rec = np.zeros(n, dtype=np.dtype(fields))
            for k, _ in fields:
                rec[k] = np.asarray(d.get(k, np.zeros(n)))
            rec.tofile(f)
