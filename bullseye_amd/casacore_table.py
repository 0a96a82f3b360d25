"""Read casacore tables -- and the Measurement Set columns bullseye grids -- without casacore.

bullseye reads its data through python-casacore (bullseye/data/data_set_loader.py:69-174 for the
subtables, :183-286 for UVW, DATA, WEIGHT(_SPECTRUM), FLAG, FLAG_ROW, DATA_DESC_ID, ANTENNA1/2,
FIELD_ID and TIME).  python-casacore is not available here, so this module decodes the on-disk
format directly:

* ``table.dat``: an AipsIO stream (big-endian, "canonical"): objects are
  ``[magic 0xbebebebe at the outermost level] u32 length, String type, u32 version`` followed by
  their fields.  It holds the row count, the data endianness, the table description (columns:
  name, data type, option bits, shape, data manager, keywords as TableRecords), the column set
  (data manager list, column -> manager) and each manager's header blob.
* ``table.f<seq>``: the StandardStMan (SSM) file of manager seq: a 512-byte header (bucket size,
  bucket count, where the row index lives), then fixed-size buckets.  The row index maps the last
  row held by each data bucket to its bucket number; inside a data bucket a column starts at the
  offset the manager header gives and holds its rows contiguously (Bool as bits, least
  significant first; String as 12 bytes: up to 8 characters inline plus the length, or bucket,
  offset and length into a string bucket whose data follows a 16-byte header).

Only StandardStMan columns of fixed cell size are decoded (scalars and fixed-shape arrays).  A
column stored by another manager (IncrementalStMan, TiledColumnStMan, TiledShapeStMan, ...) or
with variable shape raises NotImplementedError naming it.

Pinned by the reference's own table ``misc/EVLA`` (the antenna table its simulation script
passes to makems, misc/gen_data.sh:3): its POSITION column equals the 27 coordinates
bullseye/mo/benchmarks/benchmark.cpp:63-93 copied from a Measurement Set, and its names and
stations read as the VLA's.  The bit order of Bool cells and string-bucket continuation are not
exercised by that table (parity unpinned).
"""
import os
import struct

import numpy as np

MAGIC = 0xBEBEBEBE

# casacore DataType enumeration (the values table.dat stores)
TP_BOOL, TP_CHAR, TP_UCHAR, TP_SHORT, TP_USHORT, TP_INT, TP_UINT, TP_FLOAT, TP_DOUBLE = range(9)
TP_COMPLEX, TP_DCOMPLEX, TP_STRING, TP_TABLE = 9, 10, 11, 12
TP_RECORD, TP_INT64, TP_ARRAY_INT64 = 25, 29, 30
_ARRAY_OF = {13 + k: k for k in range(12)}      # TpArrayBool .. TpArrayString -> element type
_ARRAY_OF[TP_ARRAY_INT64] = TP_INT64

# element (numpy dtype) of each scalar type; Bool and String are handled separately
_NP = {TP_CHAR: "i1", TP_UCHAR: "u1", TP_SHORT: "i2", TP_USHORT: "u2", TP_INT: "i4",
       TP_UINT: "u4", TP_FLOAT: "f4", TP_DOUBLE: "f8", TP_COMPLEX: "c8", TP_DCOMPLEX: "c16",
       TP_INT64: "i8"}

OPT_DIRECT, OPT_UNDEFINED, OPT_FIXED_SHAPE = 1, 2, 4


class AipsIO:
    """Sequential reader of an AipsIO byte stream of either byte order."""

    def __init__(self, data, big_endian=True, pos=0):
        self.d, self.pos = data, pos
        self.e = ">" if big_endian else "<"

    def _un(self, fmt, n):
        v = struct.unpack_from(self.e + fmt, self.d, self.pos)
        self.pos += n
        return v[0]

    def u8(self):
        return self._un("B", 1)

    def u16(self):
        return self._un("H", 2)

    def i16(self):
        return self._un("h", 2)

    def u32(self):
        return self._un("I", 4)

    def i32(self):
        return self._un("i", 4)

    def i64(self):
        return self._un("q", 8)

    def u64(self):
        return self._un("Q", 8)

    def f32(self):
        return self._un("f", 4)

    def f64(self):
        return self._un("d", 8)

    def string(self):
        n = self.u32()
        s = self.d[self.pos:self.pos + n].decode("latin-1")
        self.pos += n
        return s

    def peek_u32(self, at=0):
        return struct.unpack_from(self.e + "I", self.d, self.pos + at)[0]

    def getstart(self, expected=None, outer=False):
        """Open an object: returns (type, version, end offset)."""
        if outer:
            magic = struct.unpack_from(">I", self.d, self.pos)[0]
            if magic != MAGIC:
                raise ValueError("not an AipsIO stream (magic %08x)" % magic)
            self.pos += 4
        start = self.pos
        length = self.u32()
        typ = self.string()
        if expected is not None and typ != expected:
            raise ValueError("expected AipsIO object %r, found %r" % (expected, typ))
        version = self.u32()
        return typ, version, start + length

    def next_is(self, typ):
        """Whether the next object (no magic) is of the given type."""
        try:
            n = struct.unpack_from(self.e + "I", self.d, self.pos + 4)[0]
            return n == len(typ) and self.d[self.pos + 8:self.pos + 8 + n] == typ.encode()
        except struct.error:
            return False

    def iposition(self):
        _, version, end = self.getstart("IPosition")
        n = self.u32()
        vals = [self.i64() if version >= 2 else self.i32() for _ in range(n)]
        self.pos = end
        return tuple(vals)

    def block(self):
        _, _, end = self.getstart("Block")
        n = self.u32()
        vals = [self.u32() for _ in range(n)]
        self.pos = end
        return vals

    # -- values of a TableRecord ----------------------------------------------------------
    def scalar(self, tp):
        if tp == TP_BOOL:
            return bool(self.u8())
        if tp == TP_STRING:
            return self.string()
        if tp in (TP_COMPLEX, TP_DCOMPLEX):
            f = self.f32 if tp == TP_COMPLEX else self.f64
            re = f()
            return complex(re, f())
        if tp == TP_UCHAR:
            return self.u8()
        if tp == TP_CHAR:
            return struct.unpack("b", bytes([self.u8()]))[0]
        if tp == TP_SHORT:
            return self.i16()
        if tp == TP_USHORT:
            return self.u16()
        if tp == TP_INT:
            return self.i32()
        if tp == TP_UINT:
            return self.u32()
        if tp == TP_INT64:
            return self.i64()
        if tp == TP_FLOAT:
            return self.f32()
        if tp == TP_DOUBLE:
            return self.f64()
        raise NotImplementedError("record value of data type %d" % tp)

    def array(self, elem):
        typ, version, end = self.getstart()
        if not typ.startswith("Array<"):
            raise ValueError("expected an Array object, found %r" % typ)
        ndim = self.u32()
        shape = [self.u32() for _ in range(ndim)]
        n = self.u32()
        vals = [self.scalar(elem) for _ in range(n)]
        self.pos = end
        arr = np.array(vals, dtype=object if elem == TP_STRING else None)
        # casacore arrays are column-major (first axis fastest): axes reversed, as
        # python-casacore returns them
        return arr.reshape(shape[::-1]) if ndim else arr

    def record_desc(self):
        _, _, end = self.getstart("RecordDesc")
        fields = []
        for _ in range(self.u32()):
            name = self.string()
            tp = self.u32()
            sub = None
            if tp == TP_RECORD:
                sub = self.record_desc()
            elif tp in _ARRAY_OF:
                self.iposition()
            elif tp == TP_TABLE:
                self.string()
            self.string()   # comment
            fields.append((name, tp, sub))
        self.pos = end
        return fields

    def table_record(self):
        _, _, end = self.getstart("TableRecord")
        fields = self.record_desc()
        self.u32()          # record type (fixed / variable)
        out = {}
        for name, tp, _ in fields:
            if tp == TP_RECORD:
                out[name] = self.table_record()
            elif tp in _ARRAY_OF:
                out[name] = self.array(_ARRAY_OF[tp])
            elif tp == TP_TABLE:
                out[name] = self.string()
            else:
                out[name] = self.scalar(tp)
        self.pos = end
        return out


class Column:
    def __init__(self, name, dtype, option, ndim, shape, dm_type, dm_group, keywords):
        self.name, self.dtype, self.option = name, dtype, option
        self.ndim, self.shape = ndim, shape
        self.dm_type, self.dm_group, self.keywords = dm_type, dm_group, keywords
        self.dm_seq = None

    @property
    def is_array(self):
        return self.dtype in _ARRAY_OF or self.ndim != 0


def _column_desc(io):
    """One ColumnDesc of a TableDesc (ScalarColumnDesc<T> or ArrayColumnDesc<T>)."""
    io.u32()                        # ColumnDesc version
    cls = io.string()               # e.g. "ScalarColumnDesc<double  "
    io.u32()                        # BaseColumnDesc version
    name, _comment = io.string(), io.string()
    dm_type, dm_group = io.string(), io.string()
    dtype, option, ndim = io.u32(), io.u32(), io.i32()
    shape = io.iposition() if io.next_is("IPosition") else ()
    io.u32()                        # maximum string length
    keywords = io.table_record()
    if cls.startswith("ScalarColumnDesc"):
        io.u32()                    # class version
        io.scalar(dtype)            # default value
    elif cls.startswith("ArrayColumnDesc"):
        io.u32()
        io.u8()
    else:
        raise NotImplementedError("column description class %r" % cls.strip())
    return Column(name, dtype, option, ndim, shape, dm_type, dm_group, keywords)


class Table:
    """A casacore table directory: ``nrows``, ``colnames()``, ``getcol(name)``, ``getcell``,
    ``getcolkeyword``, ``keywords``."""

    def __init__(self, path):
        self.path = path
        raw = open(os.path.join(path, "table.dat"), "rb").read()
        io = AipsIO(raw, big_endian=True)
        _, version, _ = io.getstart("Table", outer=True)
        self.nrows = io.u64() if version >= 3 else io.u32()
        self.little_endian = io.u32() == 1
        self.table_type = io.string()
        _, _, desc_end = io.getstart("TableDesc")
        io.string(), io.string(), io.string()    # name, version, comment
        self.keywords = io.table_record()
        io.table_record()                         # private keywords
        self.columns = {}
        for _ in range(io.u32()):
            c = _column_desc(io)
            self.columns[c.name] = c
        io.pos = desc_end
        # column set: version (stored negated), rows, data managers, column -> manager
        cs_version = -io.i32()
        io.u64() if cs_version >= 3 else io.u32()
        io.u32()
        self.managers = []
        for _ in range(io.u32()):
            self.managers.append((io.string(), io.u32()))
        for _ in self.columns:
            io.u32()
            name = io.string()
            io.u32()
            col = self.columns[name]
            col.dm_seq = io.u32()
            if col.is_array and io.u8():
                col.shape = io.iposition()
        self.manager_headers = {}
        for _, seq in self.managers:
            n = io.u32()
            self.manager_headers[seq] = raw[io.pos:io.pos + n]
            io.pos += n
        self._ssm = {}

    def colnames(self):
        return list(self.columns)

    def getcolkeyword(self, column, keyword):
        return self.columns[column].keywords[keyword]

    def _manager(self, col):
        dm_type = dict((seq, t) for t, seq in self.managers)[col.dm_seq]
        if dm_type != "StandardStMan":
            raise NotImplementedError("column %s is stored by %s, which this reader does not "
                                      "decode (StandardStMan only)" % (col.name, dm_type))
        if col.dm_seq not in self._ssm:
            self._ssm[col.dm_seq] = _SSMFile(self, col.dm_seq)
        return self._ssm[col.dm_seq]

    def getcol(self, name, startrow=0, nrow=-1):
        col = self.columns[name]
        if nrow < 0:
            nrow = self.nrows - startrow
        return self._manager(col).read(col, startrow, nrow)

    def getcell(self, name, row):
        return self.getcol(name, row, 1)[0]


class _SSMFile:
    """The StandardStMan data file of one data manager."""

    def __init__(self, table, seq):
        hdr = AipsIO(table.manager_headers[seq], big_endian=True)
        hdr.getstart("SSM", outer=True)
        hdr.string()                                       # manager name
        self.col_offset = hdr.block()                      # byte offset of each column in a bucket
        self.col_index = hdr.block()                       # which row index each column uses
        self.col_order = [c for c in table.columns if table.columns[c].dm_seq == seq]
        self.raw = open(os.path.join(table.path, "table.f%d" % seq), "rb").read()
        f = AipsIO(self.raw, big_endian=False)
        _, version, _ = f.getstart("StandardStMan", outer=True)
        if version != 3:
            raise NotImplementedError("StandardStMan file version %d" % version)
        self.big_endian = bool(f.u8())
        f.e = ">" if self.big_endian else "<"
        self.bucket_size = f.u32()
        self.n_buckets = f.u32()
        f.u32()                                  # persistent cache size
        f.u32(), f.i32()                         # free-bucket count and first free bucket
        n_idx_buckets, first_idx_bucket = f.u32(), f.i32()
        idx_offset = f.u32()
        f.i32()                                  # last string bucket
        f.u32()                                  # index length
        n_indices = f.u32()
        if n_idx_buckets != 1:
            raise NotImplementedError("a row index spread over %d buckets" % n_idx_buckets)
        ix = AipsIO(self.raw, big_endian=self.big_endian,
                    pos=self._bucket(first_idx_bucket) + idx_offset)
        self.indices = []
        for k in range(n_indices):
            _, _, end = ix.getstart("SSMIndex", outer=(k == 0))
            n_used, rows_per_bucket = ix.u32(), ix.u32()
            ix.u32()                             # columns in the index
            _, _, map_end = ix.getstart("SimpleOrderedMap")
            ix.pos = map_end                     # free-space map
            last_rows, buckets = ix.block(), ix.block()
            self.indices.append((last_rows[:n_used], buckets[:n_used], rows_per_bucket))
            ix.pos = end
        self.e = ">" if self.big_endian else "<"

    def _bucket(self, b):
        return 512 + b * self.bucket_size

    def _cell_bytes(self, col, n_elem):
        tp = _ARRAY_OF.get(col.dtype, col.dtype)
        if tp == TP_STRING:
            return 12 * n_elem
        if tp == TP_BOOL:
            return None
        return np.dtype(_NP[tp]).itemsize * n_elem

    def _string(self, b, off, n):
        out = b""
        while n > 0:
            base = self._bucket(b)
            room = self.bucket_size - 16 - off
            take = min(n, room)
            out += self.raw[base + 16 + off:base + 16 + off + take]
            n -= take
            if n > 0:
                b = struct.unpack_from(">i", self.raw, base + 12)[0]
                if b < 0:
                    raise ValueError("string runs past the last string bucket")
                off = 0
        return out.decode("latin-1")

    def read(self, col, startrow, nrow):
        if col.is_array and not col.shape:
            raise NotImplementedError("column %s has no fixed shape (indirect arrays are not "
                                      "decoded)" % col.name)
        k = self.col_order.index(col.name)
        offset, (last_rows, buckets, _) = self.col_offset[k], self.indices[self.col_index[k]]
        tp = _ARRAY_OF.get(col.dtype, col.dtype)
        n_elem = int(np.prod(col.shape)) if col.is_array else 1
        cell = self._cell_bytes(col, n_elem)
        out = []
        row, end = startrow, startrow + nrow
        while row < end:
            j = int(np.searchsorted(last_rows, row))
            if j >= len(last_rows):
                raise ValueError("row %d is not in the row index" % row)
            first = last_rows[j - 1] + 1 if j else 0
            stop = min(end, last_rows[j] + 1)
            base = self._bucket(buckets[j]) + offset
            r0, r1 = row - first, stop - first
            if tp == TP_BOOL:
                nbits = (r1 - r0) * n_elem
                lo = r0 * n_elem
                byts = np.frombuffer(self.raw, np.uint8, (lo + nbits + 7) // 8 - lo // 8,
                                     base + lo // 8)
                bits = np.unpackbits(byts, bitorder="little")[lo % 8:lo % 8 + nbits]
                out.append(bits.astype(bool).reshape(r1 - r0, n_elem))
            elif tp == TP_STRING:
                vals = []
                for r in range(r0, r1):
                    for e in range(n_elem):
                        p = base + (r * n_elem + e) * 12
                        w = struct.unpack_from(self.e + "III", self.raw, p)
                        if w[2] <= 8:
                            vals.append(self.raw[p:p + w[2]].decode("latin-1"))
                        else:
                            vals.append(self._string(w[0], w[1], w[2]))
                out.append(np.array(vals, dtype=object).reshape(r1 - r0, n_elem))
            else:
                dt = np.dtype(_NP[tp]).newbyteorder(self.e)
                a = np.frombuffer(self.raw, dt, (r1 - r0) * n_elem, base + r0 * cell)
                out.append(a.astype(dt.newbyteorder("=")).reshape(r1 - r0, n_elem))
            row = stop
        a = np.concatenate(out) if out else np.zeros((0, n_elem))
        if not col.is_array:
            return a.reshape(-1)
        # casacore cells are column-major (first axis fastest): [row] + reversed cell axes, as
        # python-casacore's getcol returns them (DATA: [row][channel][correlation])
        return a.reshape((nrow,) + tuple(col.shape[::-1]))


def read_ms(path, data_column="DATA", start_row=0, no_rows=-1):
    """The columns and subtable values bullseye's loader reads from a Measurement Set
    (data_set_loader.py:69-174, 183-286), rows ordered by TIME as its TaQL ORDERBY TIME
    (stable).  Returns a dict of NumPy arrays; raises NotImplementedError when a needed column is
    stored by a manager other than StandardStMan."""
    ms = Table(path)
    n = ms.nrows - start_row if no_rows < 0 else no_rows
    get = lambda c: ms.getcol(c, start_row, n)
    out = {"uvw": get("UVW"), "data": get(data_column), "flag": get("FLAG"),
           "flag_row": get("FLAG_ROW"), "data_desc_id": get("DATA_DESC_ID"),
           "antenna1": get("ANTENNA1"), "antenna2": get("ANTENNA2"),
           "field_id": get("FIELD_ID"), "time": get("TIME")}
    if "WEIGHT_SPECTRUM" in ms.columns:
        out["weight"] = get("WEIGHT_SPECTRUM")
    else:   # one weight per correlation, for every channel (data_set_loader.py:239-246)
        w = get("WEIGHT")
        out["weight"] = np.broadcast_to(w[:, None, :], out["data"].shape).copy()
    order = np.argsort(out["time"], kind="stable")
    out = {k: v[order] for k, v in out.items()}
    sub = lambda name: Table(os.path.join(path, name))
    ant = sub("ANTENNA")
    out["antenna_positions"] = ant.getcol("POSITION")
    out["antenna_names"] = ant.getcol("NAME")
    spw = sub("SPECTRAL_WINDOW")
    out["chan_freq"] = spw.getcol("CHAN_FREQ")
    pol = sub("POLARIZATION")
    out["corr_type"] = pol.getcol("CORR_TYPE")
    field = sub("FIELD")
    out["phase_dir"] = field.getcol("PHASE_DIR")
    dd = sub("DATA_DESCRIPTION")
    out["dd_spw"] = dd.getcol("SPECTRAL_WINDOW_ID")
    return out
