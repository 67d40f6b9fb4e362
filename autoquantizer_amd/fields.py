"""Quantized fields placed in bit containers: the drop-in for the part of
Taichi's SNode/field API the quantized MPM path uses.

Mirrors (TS/ = taichi_submission/):
  SNode.dense/bit_struct/bit_pack/place(shared_exponent)  TS/python/taichi/lang/snode.py:108-180
  ti.field / ti.Vector.field / ti.Matrix.field            TS/python/taichi/lang/impl.py, matrix.py
  ti.bit_pack_assign(parent, index, values)               TS/python/taichi/lang/impl.py:753-756
  field[i] reads and writes, to_numpy/from_numpy

Storage is a torch int64 tensor of shape [nwords, n] (word planes) on the GPU;
every encode/decode runs in libaqmpm (autoquantizer_amd/csrc).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .quant import place_record

i = 0  # axis tags, as ti.i / ti.j
j = 1


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class Field:
    """A scalar, vector or matrix field whose members are quantized scalars."""

    def __init__(self, shape, dtypes, name=None):
        self.elem_shape = tuple(shape)
        n = int(np.prod(self.elem_shape)) if self.elem_shape else 1
        if not isinstance(dtypes, (list, tuple, np.ndarray)):
            dtypes = [dtypes] * n
        flat = list(np.asarray(dtypes, dtype=object).reshape(-1))
        if len(flat) != n:
            raise ValueError("one dtype or one per member")
        self.dtypes = flat
        self.name = name
        self.members = [None] * n  # (layout, component index)
        self.container = None

    # --- placement ---
    def _members(self):
        return [(self, k) for k in range(len(self.dtypes))]

    def get_scalar_field(self, *idx):
        k = int(np.ravel_multi_index(idx, self.elem_shape)) if self.elem_shape else 0
        return _ScalarMember(self, k)

    def parent(self):
        return self.container

    @property
    def layout(self):
        return self.container.snode.layout()

    # --- access ---
    def _comp_ids(self):
        self.layout  # materialize
        return [m[1] for m in self.members]

    def __getitem__(self, idx):
        lay = self.layout
        vals = lay.read_records(np.atleast_1d(np.asarray(idx, dtype=np.int64)))
        out = vals[self._comp_ids(), 0]
        if not self.elem_shape:
            return float(out[0])
        return out.reshape(self.elem_shape)

    def __setitem__(self, idx, value):
        lay = self.layout
        v = np.asarray(value, dtype=np.float32).reshape(-1)
        if v.size == 1 and len(self.dtypes) > 1:
            v = np.full(len(self.dtypes), v[0], dtype=np.float32)
        lay.write_members(self._comp_ids(), np.atleast_1d(np.asarray(idx, dtype=np.int64)),
                          v.reshape(-1, 1))

    def to_numpy(self):
        vals = self.layout.unpack()
        return vals[self._comp_ids()].T.reshape((-1,) + self.elem_shape).copy()

    def from_numpy(self, arr):
        lay = self.layout
        a = np.asarray(arr, dtype=np.float32).reshape(lay.n, -1).T
        lay.write_members(self._comp_ids(), np.arange(lay.n, dtype=np.int64), a)

    def fill(self, value):
        lay = self.layout
        a = np.full((len(self.dtypes), lay.n), value, dtype=np.float32)
        lay.write_members(self._comp_ids(), np.arange(lay.n, dtype=np.int64), a)


class _ScalarMember:
    def __init__(self, field, k):
        self.field, self.k = field, k

    def _members(self):
        return [(self.field, self.k)]


def field(dtype, shape=None, name=None):
    f = Field((), dtype, name)
    if shape is not None:
        root.dense(i, shape).place(f)
    return f


class Vector:
    @staticmethod
    def field(n, dtype, shape=None, name=None):
        f = Field((n,), dtype, name)
        if shape is not None:
            root.dense(i, shape).place(f)
        return f


class Matrix:
    @staticmethod
    def field(n, m, dtype, shape=None, name=None):
        f = Field((n, m), dtype, name)
        if shape is not None:
            root.dense(i, shape).place(f)
        return f


class Container:
    """A bit_struct(num_bits) or bit_pack() container under a dense node."""

    def __init__(self, snode, kind, num_bits):
        self.snode, self.kind, self.num_bits = snode, kind, num_bits
        self.entries = []  # (field, member index, shared group)

    def place(self, *fields, shared_exponent=False):
        if self.snode.materialized:
            raise RuntimeError("cannot place into a materialized layout")
        gid = self.snode.new_group() if shared_exponent else -1
        for f in fields:
            for (fld, k) in f._members():
                if fld.members[k] is not None:
                    raise RuntimeError("This variable has been placed.")
                fld.members[k] = (self, None)
                fld.container = self
                self.entries.append((fld, k, gid))
        return self


class DenseSNode:
    """ti.root.dense(ti.i, n): the particle axis of a set of bit containers."""

    def __init__(self, n):
        self.n = int(n)
        self.containers = []
        self._groups = 0
        self._layout = None

    @property
    def materialized(self):
        return self._layout is not None

    def new_group(self):
        self._groups += 1
        return self._groups - 1

    def dense(self, axis, n):
        return DenseSNode(self.n * int(n))

    def bit_pack(self):
        c = Container(self, _lib.AQ_BIT_PACK, 0)
        self.containers.append(c)
        return c

    def bit_struct(self, num_bits):
        c = Container(self, _lib.AQ_BIT_STRUCT, int(num_bits))
        self.containers.append(c)
        return c

    def place(self, *fields):
        """Unquantized members: each goes to its own 32-bit slot."""
        for f in fields:
            for (fld, k) in f._members():
                self.bit_struct(32).place(_ScalarMember(fld, k))
        return self

    def layout(self):
        if self._layout is None:
            self._layout = RecordLayout.from_snode(self)
        return self._layout


class _Root:
    def dense(self, axis, n):
        return DenseSNode(n)


root = _Root()


def plan(containers, places):
    """aq_layout_plan on python records; returns (list of component dicts, nwords)."""
    L = _lib.lib()
    ca = (_lib.Container * len(containers))()
    for ci, (kind, nb) in enumerate(containers):
        ca[ci].kind, ca[ci].num_bits = kind, nb
    pa = (_lib.Place * len(places))()
    for pi, rec in enumerate(places):
        for key, val in rec.items():
            setattr(pa[pi], key, val)
    out = (_lib.Component * len(places))()
    nw = ctypes.c_int()
    _lib.check(L.aq_layout_plan(ca, len(containers), pa, len(places), out, ctypes.byref(nw)))
    comps = []
    for c in out:
        comps.append({f: getattr(c, f) for f, _ in _lib.Component._fields_})
    return comps, nw.value


def comps_array(comps):
    arr = (_lib.Component * len(comps))()
    for k, c in enumerate(comps):
        for key, val in c.items():
            setattr(arr[k], key, val)
    return arr


class RecordLayout:
    """A planned particle record plus its word-plane storage on the GPU."""

    def __init__(self, comps, nwords, n, device="cuda", seed=0):
        if len(comps) > _lib.AQ_MAX_COMPS:
            raise ValueError(f"at most {_lib.AQ_MAX_COMPS} members per record")
        self.comps = comps
        self.nwords = nwords
        self.n = n
        self.seed = seed
        self.step = 0
        self.words = torch.zeros((nwords, n), dtype=torch.int64, device=device)
        self._carr = comps_array(comps)

    @classmethod
    def from_snode(cls, snode):
        containers, places, owners = [], [], []
        for ci, c in enumerate(snode.containers):
            containers.append((c.kind, c.num_bits))
            for (fld, k, gid) in c.entries:
                rec = place_record(fld.dtypes[k])
                rec.update(container=ci, shared_group=gid)
                places.append(rec)
                owners.append((fld, k))
        comps, nw = plan(containers, places)
        lay = cls(comps, nw, snode.n)
        for idx, (fld, k) in enumerate(owners):
            fld.members[k] = (fld.container, idx)
        return lay

    # --- bulk codec ---
    def unpack(self):
        vals = torch.empty((len(self.comps), self.n), dtype=torch.float32, device=self.words.device)
        _lib.check(_lib.lib().aq_unpack(self._carr, len(self.comps), self.nwords, _ptr(self.words),
                                        _ptr(vals), self.n, _stream()))
        return vals.cpu().numpy()

    def pack(self, vals, noise=None, step=None):
        v = torch.as_tensor(np.ascontiguousarray(vals, dtype=np.float32), device=self.words.device)
        nz = None if noise is None else torch.as_tensor(np.ascontiguousarray(noise, dtype=np.float32),
                                                        device=self.words.device)
        st = self.step if step is None else step
        _lib.check(_lib.lib().aq_pack(self._carr, len(self.comps), self.nwords, _ptr(v),
                                      None if nz is None else _ptr(nz), self.seed, st,
                                      _ptr(self.words), self.n, _stream()))
        torch.cuda.current_stream().synchronize()

    def read_records(self, idx):
        w = self.words[:, torch.as_tensor(idx, device=self.words.device)].contiguous()
        vals = torch.empty((len(self.comps), w.shape[1]), dtype=torch.float32, device=w.device)
        _lib.check(_lib.lib().aq_unpack(self._carr, len(self.comps), self.nwords, _ptr(w), _ptr(vals),
                                        w.shape[1], _stream()))
        return vals.cpu().numpy()

    def write_members(self, comp_ids, idx, vals):
        """vals[j, t]: value of comp_ids[j] for record idx[t] (masked store)."""
        order = np.argsort(comp_ids)
        mask = 0
        for c in comp_ids:
            mask |= 1 << int(c)
        v = np.asarray(vals, dtype=np.float32).reshape(len(comp_ids), -1)[order]
        count = v.shape[1]
        dv = torch.as_tensor(np.ascontiguousarray(v), device=self.words.device)
        di = torch.as_tensor(np.asarray(idx, dtype=np.int64), device=self.words.device)
        _lib.check(_lib.lib().aq_store_members(self._carr, len(self.comps), self.nwords,
                                               _ptr(self.words), self.n, mask, _ptr(di), _ptr(dv),
                                               count, self.seed, self.step, _stream()))
        torch.cuda.current_stream().synchronize()


def bit_pack_assign(container, index, values):
    """ti.bit_pack_assign (TS/python/taichi/lang/impl.py:753-756): write every
    member of `container` at record `index`; values in placement order."""
    lay = container.snode.layout()
    ids = [fld.members[k][1] for (fld, k, _) in container.entries]
    vals = np.asarray(values, dtype=np.float32).reshape(-1)
    if len(vals) != len(ids):
        raise ValueError("bit_pack_assign needs a value for every member of the container")
    lay.write_members(ids, np.atleast_1d(np.asarray(index, dtype=np.int64)), vals.reshape(-1, 1))
