"""Madagascar/SEP RSF files: a text header of key=value pairs plus a binary.

Follows the header semantics of SEFILESYSTEM/src/se_fs_sep.cpp:
  * every ``key=value`` field on every line is parsed, later ones override
    earlier ones (sep_headers_parse, :440-640);
  * ``n1..n9`` (default 1), ``o1..o9`` (default 0), ``d1..d9`` (default 1,
    a zero value is ignored), ``esize`` (default 4);
  * ``data_format="native_float"`` is little-endian, ``"xdr_float"`` is
    big-endian, and a header without the key is read as XDR (:456-457,
    :612-623);
  * ``in=`` names the binary; a relative path is taken relative to the
    header's directory (sep_expand_binary_path), written as ``<name>@``
    next to the header (:780-782).
Axis 1 is the fastest; arrays are returned in C order with shape
(n_ndim, ..., n2, n1).
"""

from __future__ import annotations

import os
import re
import shlex
from dataclasses import dataclass, field

import numpy as np

_FIELD = re.compile(r"^([A-Za-z_][A-Za-z0-9_]*)=(.*)$")


@dataclass
class Axis:
    n: int = 1
    o: float = 0.0
    d: float = 1.0


@dataclass
class Header:
    axes: list = field(default_factory=lambda: [Axis() for _ in range(9)])
    ndim: int = 1
    esize: int = 4
    little_endian: bool = False
    binary: str = ""
    extra: dict = field(default_factory=dict)

    @property
    def shape(self):
        """C-order shape (slowest first)."""
        return tuple(self.axes[i].n for i in reversed(range(self.ndim)))


def read_header(path: str) -> Header:
    h = Header()
    h.axes[0].n = 0
    with open(path, "rb") as f:
        raw = f.read()
    eot = raw.find(b"\x04")
    if eot >= 0:
        raw = raw[:eot]
    text = raw.decode("latin-1")
    for line in text.splitlines():
        line = line.split("#", 1)[0].strip()
        if not line:
            continue
        try:
            fields = shlex.split(line, posix=True)
        except ValueError:
            fields = line.split()
        for fld in fields:
            m = _FIELD.match(fld.strip())
            if not m:
                continue
            key, val = m.group(1), m.group(2).strip().strip('"')
            km = re.fullmatch(r"([nod])([1-9])", key)
            if km:
                idx = int(km.group(2)) - 1
                try:
                    dv = float(val)
                except ValueError:
                    continue
                if km.group(1) == "n":
                    if dv < 1.0:
                        continue
                    h.axes[idx].n = int(dv)
                elif km.group(1) == "o":
                    h.axes[idx].o = dv
                else:
                    if dv == 0.0:
                        continue
                    h.axes[idx].d = dv
                h.ndim = max(h.ndim, idx + 1)
            elif key == "in":
                if val:
                    h.binary = val
            elif key == "esize":
                try:
                    e = int(val)
                except ValueError:
                    continue
                if e in (1, 2, 4, 8, 12):
                    h.esize = e
            elif key == "data_format":
                if val == "native_float":
                    h.little_endian = True
                elif val == "xdr_float":
                    h.little_endian = False
            else:
                h.extra[key] = val
    return h


def _binary_path(path: str, h: Header) -> str:
    b = h.binary
    if not b:
        raise ValueError(f"RSF header {path} has no in=")
    if os.path.isabs(b):
        return b
    cand = os.path.join(os.path.dirname(os.path.abspath(path)), b)
    if os.path.exists(cand):
        return cand
    return os.path.abspath(b)


def read(path: str, mmap: bool = False):
    """Return (array float32, Header)."""
    h = read_header(path)
    if h.esize != 4:
        raise ValueError(f"{path}: only esize=4 float data is supported")
    dt = np.dtype("<f4" if h.little_endian else ">f4")
    bpath = _binary_path(path, h)
    count = int(np.prod(h.shape))
    if mmap:
        arr = np.memmap(bpath, dtype=dt, mode="r", shape=h.shape)
    else:
        arr = np.fromfile(bpath, dtype=dt, count=count)
        if arr.size != count:
            raise ValueError(f"{path}: binary holds {arr.size} values, header needs {count}")
        arr = arr.reshape(h.shape)
    if not h.little_endian and not mmap:
        arr = arr.astype("<f4")
    return arr, h


def _fmt(v: float) -> str:
    return repr(float(v))


def write(path: str, arr, o=None, d=None, labels=None, units=None, extra=None) -> None:
    """Write float32 ``arr`` (C order, slowest axis first) as native_float RSF."""
    a = np.ascontiguousarray(np.asarray(arr, dtype="<f4"))
    shape = a.shape if a.ndim else (1,)
    ndim = len(shape)
    ns = list(reversed(shape))
    o = list(o) if o is not None else [0.0] * ndim
    d = list(d) if d is not None else [1.0] * ndim
    bname = os.path.basename(path) + "@"
    bpath = os.path.join(os.path.dirname(os.path.abspath(path)), bname)
    a.tofile(bpath)
    lines = [f'\tin="{bname}"', '\tdata_format="native_float"']
    lines += [f"\tn{i + 1}={ns[i]}" for i in range(ndim)]
    lines += [f"\td{i + 1}={_fmt(d[i])}" for i in range(ndim)]
    lines += [f"\to{i + 1}={_fmt(o[i])}" for i in range(ndim)]
    lines.append("\tesize=4")
    for i, lab in enumerate(labels or []):
        if lab:
            lines.append(f'\tlabel{i + 1}="{lab}"')
    for i, un in enumerate(units or []):
        if un:
            lines.append(f'\tunit{i + 1}="{un}"')
    for k, v in (extra or {}).items():
        lines.append(f'\t{k}="{v}"' if isinstance(v, str) else f"\t{k}={v}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
