"""Reader of EXP split phase-space ("SPL.") files: drop-in for exptool.io.spl_io.

  Input        io/spl_io.py:27-270

Layout (little-endian).  The primary file SPL.<run>.<NNNNN> holds
  master header  time f8, nbodies_tot u32, ncomp u32                 (16 B)
  per component  u32 magic (0xADBFABC4 for float records), u32 rsize,
                 u32 nprocs, nbodies, nint_attr, nfloat_attr, infostringlen,
                 the YAML info string, then nprocs 1024-byte NUL-padded
                 names of the subfiles holding the particles
and each subfile (in the primary file's directory) holds u32 nbodies and the
records [id i8 if indexing] m x y z vx vy vz potE (f4 or f8), nint_attr i4,
nfloat_attr floats.  Component data is the subfiles concatenated in order.
"""
from __future__ import annotations

import os

import numpy as np
import yaml

from .psp_io import PSP_FLOAT_MAGIC, parse_info_string

PBUF_SZ = 1024


class Input:
    """io/spl_io.py:27-103: headers of every component and, for `comp`,
    .data (dict of id?, m, x, y, z, vx, vy, vz, potE, i_attr*, f_attr*).
    With comp=None the components are listed and nothing is read."""

    def __init__(self, filename, comp=None, verbose=0):
        self.verbose = verbose
        self.filename = filename
        self.header = dict()
        self.component_map = dict()
        self.subfile_lists = dict()
        if not os.path.exists(filename):
            raise IOError('Failed to open "{}"'.format(filename))
        with open(filename, 'rb') as f:
            self._read_primary_header(f)
        self.comp = comp
        if comp is None:
            self._summarise_primary_header()
            return
        if comp not in self.header:
            raise IOError('The specified component does not exist.')
        # subfiles sit beside the primary file (spl_io.py:95)
        self.indir = filename.split('SPL')[0]
        self.subfiles = self.subfile_lists[comp]
        self._read_spl_component_data()

    def _read_primary_header(self, f):
        """spl_io.py:106-121 and :124-160."""
        self.time = float(np.fromfile(f, dtype='<f8', count=1)[0])
        self._nbodies_tot, self._ncomp = (int(v) for v in np.fromfile(f, dtype='<u4', count=2))
        magic = np.fromfile(f, dtype='<u4', count=1)
        self._float_len = 4 if (magic.size and magic[0] == PSP_FLOAT_MAGIC) else 8
        pos = 16
        for _ in range(self._ncomp):
            f.seek(pos)
            _, _, nprocs, nbodies, nint, nfloat, ninfo = (int(v) for v in
                                                          np.fromfile(f, dtype='<u4', count=7))
            h = parse_info_string(f.read(ninfo))
            h.update(nint_attr=nint, nfloat_attr=nfloat, nbodies=nbodies, nprocs=nprocs)
            start = pos + 4 * 7 + ninfo
            names = []
            for _ in range(nprocs):
                names.append(f.read(PBUF_SZ).split(b'\x00')[0].decode())
            self.component_map[h['name']] = start
            self.header[h['name']] = h
            self.subfile_lists[h['name']] = names
            self.nprocs = nprocs
            self.indexing = bool(h['parameters'].get('indexing', False))
            pos = start + nprocs * PBUF_SZ

    def _summarise_primary_header(self):
        comps = list(self.header.keys())
        print("Found {} components.".format(len(comps)))
        for n, c in enumerate(comps):
            print("Component {}: {}".format(n, c))

    def _dtype(self):
        h = self.header[self.comp]
        fl = '<f4' if self._float_len == 4 else '<f8'
        fields = [('id', '<i8')] if h['parameters'].get('indexing', False) else []
        fields += [(k, fl) for k in ('m', 'x', 'y', 'z', 'vx', 'vy', 'vz', 'potE')]
        fields += [('i_attr{}'.format(i), '<i4') for i in range(h['nint_attr'])]
        fields += [('f_attr{}'.format(i), fl) for i in range(h['nfloat_attr'])]
        return np.dtype(fields)

    def _handle_spl_subfile(self, subfilename):
        """spl_io.py:214-224 and :227-270: u32 nbodies, then the records."""
        dt = self._dtype()
        with open(self.indir + subfilename, 'rb') as f:
            n = int(np.fromfile(f, dtype='<u4', count=1)[0])
            rec = np.fromfile(f, dtype=dt, count=n)
        if rec.size != n:
            raise IOError('{}: subfile is truncated'.format(subfilename))
        return {k: np.array(rec[k]) for k in dt.names}

    def _read_spl_component_data(self):
        """spl_io.py:174-199: every subfile, concatenated."""
        parts = [self._handle_spl_subfile(s) for s in self.subfiles]
        self.data = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}


def write_spl(filename, time, comps, nprocs=2, float_len=4):
    """Write SPL.<run>.<NNNNN> and its subfiles <name>_<c>-<p> beside it.
    comps: list of (header dict, data dict with m, x, y, z, vx, vy, vz, potE
    and optionally id); each component is split over nprocs subfiles."""
    fl = '<f4' if float_len == 4 else '<f8'
    d0 = os.path.dirname(filename)
    base = os.path.basename(filename)
    with open(filename, 'wb') as f:
        ntot = sum(int(np.asarray(d['x']).size) for _, d in comps)
        np.array([time], dtype='<f8').tofile(f)
        np.array([ntot, len(comps)], dtype='<u4').tofile(f)
        for ci, (info, d) in enumerate(comps):
            n = int(np.asarray(d['x']).size)
            idx = 'id' in d
            info = dict(info)
            info.setdefault('parameters', {})['indexing'] = idx
            text = yaml.safe_dump(info).encode() + b'\x00'
            rsize = 4 if float_len == 4 else 8
            magic = PSP_FLOAT_MAGIC if float_len == 4 else 0
            np.array([magic, rsize, nprocs, n, 0, 0, len(text)], dtype='<u4').tofile(f)
            f.write(text)
            fields = ([('id', '<i8')] if idx else []) + \
                [(k, fl) for k in ('m', 'x', 'y', 'z', 'vx', 'vy', 'vz', 'potE')]
            cuts = np.linspace(0, n, nprocs + 1).astype(int)
            for p in range(nprocs):
                sub = '{}_{}-{}'.format(base, ci, p)
                f.write(sub.encode() + b'\x00' * (PBUF_SZ - len(sub)))
                m = int(cuts[p + 1] - cuts[p])
                rec = np.zeros(m, dtype=np.dtype(fields))
                for k, _ in fields:
                    rec[k] = np.asarray(d.get(k, np.zeros(n)))[cuts[p]:cuts[p + 1]]
                with open(os.path.join(d0, sub), 'wb') as g:
                    np.array([m], dtype='<u4').tofile(g)
                    rec.tofile(g)
