"""Reader of TELEMAC result files (SELAFIN / SERAFIN format).

A SELAFIN file is a sequence of big-endian Fortran unformatted records, each
framed by its byte length: title (80 characters), the counts of variables, one
record per variable name (32 characters: 16 name + 16 unit), ten integer
parameters (the 7th is the number of planes of a 3D file; the 10th set means a
date record follows), the element / point / nodes-per-element counts, the
connectivity (1-based), the boundary node numbering, the x and y coordinates,
then per time record the time and one record per variable.  Floats are 4 or 8
bytes (the coordinate records' lengths tell which).

This replaces the data_manip TelemacFile calls of
cylag/core/eulerian_field_set.pyx:317-483 (tri, times, npoin2, nplan,
get_data_value); the triangulation of a 3D file is the bottom plane's.
"""
import struct

import numpy as np


class SelafinFile:
    def __init__(self, file_name):
        self.file_name = file_name
        with open(file_name, "rb") as f:
            self._buf = f.read()
        self._pos = 0
        self.title = self._record().decode("latin-1")
        nbv1, nbv2 = self._ints(self._record())[:2]
        names = [self._record().decode("latin-1") for _ in range(nbv1 + nbv2)]
        self.varnames = [n[:16].strip() for n in names[:nbv1]]
        self.varunits = [n[16:].strip() for n in names[:nbv1]]
        self.iparam = self._ints(self._record())
        if self.iparam[9] == 1:
            self.date = self._ints(self._record())
        self.nelem, self.npoin, self.ndp = self._ints(self._record())[:3]
        self.ikle = self._ints(self._record()).reshape(self.nelem, self.ndp) - 1
        self.ipobo = self._ints(self._record())
        xrec = self._record()
        self.fsize = len(xrec) // self.npoin
        if self.fsize not in (4, 8):
            raise ValueError("{}: unexpected coordinate record length".format(file_name))
        self._fdt = ">f4" if self.fsize == 4 else ">f8"
        self.x = np.frombuffer(xrec, dtype=self._fdt).astype("d")
        self.y = np.frombuffer(self._record(), dtype=self._fdt).astype("d")
        self.nplan = max(1, int(self.iparam[6]))
        self.npoin2 = self.npoin // self.nplan
        if self.nplan > 1:
            nelem2 = self.nelem // (self.nplan - 1)
            self.ikle2 = self.ikle[:nelem2, :3].copy()
        else:
            self.ikle2 = self.ikle[:, :3].copy()
        # time records: offsets of each record's variable records
        self._data_pos = []
        times = []
        nv = nbv1
        while self._pos + 8 <= len(self._buf):  # (a stray trailing byte ends the file)
            t = np.frombuffer(self._record(), dtype=self._fdt)[0]
            times.append(float(t))
            pos = []
            for _ in range(nv):
                pos.append(self._pos)
                self._skip()
            self._data_pos.append(pos)
        self.times = np.array(times, dtype="d")
        self.ntimestep = len(times)

    # -- records -------------------------------------------------------------------
    def _record(self):
        (n,) = struct.unpack(">i", self._buf[self._pos:self._pos + 4])
        start = self._pos + 4
        end = start + n
        (n2,) = struct.unpack(">i", self._buf[end:end + 4])
        if n2 != n:
            raise ValueError("{}: corrupt record at byte {}".format(self.file_name, self._pos))
        self._pos = end + 4
        return self._buf[start:end]

    def _skip(self):
        (n,) = struct.unpack(">i", self._buf[self._pos:self._pos + 4])
        self._pos += n + 8

    @staticmethod
    def _ints(b):
        return np.frombuffer(b, dtype=">i4").astype(np.int64)

    # -- data ----------------------------------------------------------------------
    @property
    def tri(self):
        """matplotlib triangulation of the (bottom) 2D mesh."""
        from matplotlib.tri import Triangulation
        return Triangulation(self.x[:self.npoin2], self.y[:self.npoin2], self.ikle2.astype(np.int32))

    def get_data_value(self, varname, record):
        """Values of a variable (name, stripped) at a time record, all points."""
        if varname not in self.varnames:
            raise KeyError("{}: no variable {!r} (has {})".format(self.file_name, varname, self.varnames))
        p = self._data_pos[record][self.varnames.index(varname)]
        (n,) = struct.unpack(">i", self._buf[p:p + 4])
        return np.frombuffer(self._buf[p + 4:p + 4 + n], dtype=self._fdt).astype("d")
