"""Device-resident dftable on one rank's share.

Mirrors the reference's dftable interface (src/frovedis/dataframe/
dftable.hpp:32-260) for numeric columns, with every column held in HBM
between calls, so a chain of operations moves no data over PCIe:

- ``filter`` / ``sort`` / ``sort_desc`` return views that keep the base
  columns and a row index, as the reference's filtered_dftable
  (dfoperator.hpp:2443) and sorted_dftable (dftable.hpp:784-855) do (late
  materialization);
- ``group_by(cols).select(cols, aggs)``, ``hash_join``, ``outer_hash_join``
  and ``star_join`` produce materialised tables (dftable.cc:164-333,
  :346-487, :3724-3772);
- ``materialize`` gathers every column at the view's rows
  (global_extract on one rank).

NULL is numeric_limits<T>::max() with a per-column list of null row
positions, the reference's representation (typed_dfcolumn<T>::val and
nulls, dfcolumn.hpp:1012).  Every step calls one of the library's entry
points; nothing is computed in Python.
"""
import numpy as np

from . import FhipError
from .core import DeviceArray
from . import dataframe as _df

__all__ = ["DFTable", "GroupedDFTable"]


def _as_dev(a, device):
    if isinstance(a, DeviceArray):
        return a
    return DeviceArray.from_numpy(np.ascontiguousarray(a), device=device)


class DFTable:
    """A table of named numeric device columns, optionally viewed through a
    row index (the rows of a filtered or sorted table, in order)."""

    def __init__(self, columns=None, nulls=None, device="cuda", _idx=None):
        self._cols = {}
        self._nulls = {}
        self._order = []
        self._idx = _idx
        self._device = device
        n = None
        for name, col in (columns or {}).items():
            c = _as_dev(col, device)
            if n is not None and len(c) != n:
                raise FhipError("dftable: columns of different sizes")
            n = len(c)
            self._cols[name] = c
            self._order.append(name)
        for name, nl in (nulls or {}).items():
            if name not in self._cols:
                raise FhipError("no such column: " + name)
            if nl is not None and len(nl):
                self._nulls[name] = _as_dev(np.asarray(nl, dtype=np.uint64), device)

    # ---- shape ---------------------------------------------------------------
    def columns(self):
        return list(self._order)

    def num_row(self):
        if self._idx is not None:
            return len(self._idx)
        return len(self._cols[self._order[0]]) if self._order else 0

    def dtypes(self):
        return [(c, self._cols[c].dtype) for c in self._order]

    def _base_len(self):
        return len(self._cols[self._order[0]]) if self._order else 0

    def _view(self, idx, names=None):
        t = DFTable(device=self._device, _idx=idx)
        for c in (names if names is not None else self._order):
            t._cols[c] = self._cols[c]
            if c in self._nulls:
                t._nulls[c] = self._nulls[c]
            t._order.append(c)
        return t

    def _row_ids(self):
        """The base rows of this table, in order (a fresh device array)."""
        if self._idx is not None:
            return DeviceArray(self._idx.t.clone(), np.uint64)
        import torch
        return DeviceArray(torch.arange(self._base_len(), dtype=torch.int64,
                                        device=self._device), np.uint64)

    # ---- columns ---------------------------------------------------------------
    def _check(self, name):
        if name not in self._cols:
            raise FhipError("no such column: " + name)

    def column(self, name):
        """The column in this table's row order (a device array)."""
        self._check(name)
        c = self._cols[name]
        return c if self._idx is None else _df.global_extract(c, self._idx)

    def column_nulls(self, name):
        """Positions of the NULL rows of the column, in this table's order."""
        self._check(name)
        nl = self._nulls.get(name)
        if nl is None:
            return None
        if self._idx is None:
            return nl
        r = _df.extract_nulls(self._idx, nl, len(self._cols[name]))
        return r if len(r) else None

    def select(self, names):
        for c in names:
            self._check(c)
        return self._view(self._idx, list(names))

    def materialize(self):
        t = DFTable(device=self._device)
        for c in self._order:
            t._cols[c] = self.column(c)
            nl = self.column_nulls(c)
            if nl is not None:
                t._nulls[c] = nl
            t._order.append(c)
        return t

    def to_numpy(self):
        return {c: self.column(c).numpy() for c in self._order}

    def null_rows(self):
        return {c: (self.column_nulls(c).numpy() if self.column_nulls(c) is not None
                    else np.zeros(0, np.uint64)) for c in self._order}

    def show(self, limit=20):
        d = self.to_numpy()
        nulls = {c: set(v.tolist()) for c, v in self.null_rows().items()}
        print("\t".join(self._order))
        for i in range(min(self.num_row(), limit)):
            print("\t".join("NULL" if i in nulls[c] else str(d[c][i]) for c in self._order))

    # ---- filter / sort ---------------------------------------------------------
    def filter(self, name, op, other):
        """filter(eq/neq/lt/le/gt/ge(name, other)): `other` is a column
        name (filter_<op>) or a numpy scalar (filter_<op>_immed).  Returns
        a view of the rows where the comparison holds, NULL rows removed."""
        col = self.column(name)
        nl = self.column_nulls(name)
        if isinstance(other, str):
            pos = _df.filter(col, op, self.column(other), nulls=nl,
                             other_nulls=self.column_nulls(other))
        else:
            pos = _df.filter(col, op, other, nulls=nl)
        idx = pos if self._idx is None else _df.global_extract(self._idx, pos)
        return self._view(idx)

    def _sort(self, name, desc):
        self._check(name)
        if self._idx is None:
            _, order = (_df.sort_desc if desc else _df.sort)(self._cols[name])
        else:
            _, order = (_df.sort_desc if desc else _df.sort)(self._cols[name], self._idx)
        return self._view(order)

    def sort(self, name):
        """Stable sort by one column (sort again by another column for a
        multi-column order, the last call being the primary key)."""
        return self._sort(name, False)

    def sort_desc(self, name):
        return self._sort(name, True)

    # ---- group by ---------------------------------------------------------------
    def group_by(self, names):
        if isinstance(names, str):
            names = [names]
        for c in names:
            self._check(c)
        return GroupedDFTable(self, list(names))

    # ---- joins ------------------------------------------------------------------
    def _joined(self, right, lo, ro, only=None):
        clash = set(self._order) & set(right._order)
        if clash:
            raise FhipError("joined table has duplicate column names: " +
                            ", ".join(sorted(clash)))
        import torch
        if only is not None and len(only):
            lids = DeviceArray(torch.cat([lo.t, only.t]), np.uint64)
            null = torch.full((len(only),), -1, dtype=torch.int64, device=lo.t.device)
            rids = DeviceArray(torch.cat([ro.t, null]), np.uint64)
        else:
            lids, rids = lo, ro
        t = DFTable(device=self._device)
        for side, ids in ((self, lids), (right, rids)):
            for c in side._order:
                t._cols[c] = _df.global_extract(side._cols[c], ids)
                nl = side._nulls.get(c)
                if side is right and only is not None and len(only):
                    n = _df.extract_nulls(ids, nl, len(side._cols[c]))
                elif nl is not None:
                    n = _df.extract_nulls(ids, nl, len(side._cols[c]))
                else:
                    n = None
                if n is not None and len(n):
                    t._nulls[c] = n
                t._order.append(c)
        return t

    def _join_keys(self, right, left_on, right_on):
        lk, rk = self.column(left_on), right.column(right_on)
        if lk.dtype != rk.dtype:
            raise FhipError("column types are different")
        if self.column_nulls(left_on) is not None or right.column_nulls(right_on) is not None:
            raise FhipError("join keys with NULL rows are not supported")
        return lk, self._row_ids(), rk, right._row_ids()

    def hash_join(self, right, left_on, right_on, strategy="auto"):
        """hash_join(right, eq(left_on, right_on)) on one rank: the
        reference's equi_join (unique right keys: probe order; otherwise
        sort-merge order), materialised."""
        lk, li, rk, ri = self._join_keys(right, left_on, right_on)
        lo, ro, _ = _df.equi_join(lk, li, rk, ri, strategy=strategy)
        return self._joined(right, lo, ro)

    def outer_hash_join(self, right, left_on, right_on, strategy="auto"):
        """LEFT outer join: the left rows without a partner follow the
        pairs with NULL right columns (hash_joined_dftable with right_nulls,
        dftable.hpp:874-888)."""
        lk, li, rk, ri = self._join_keys(right, left_on, right_on)
        lo, ro, only, _ = _df.outer_equi_join(lk, li, rk, ri, strategy=strategy)
        return self._joined(right, lo, ro, only)

    def star_join(self, rights, ons):
        """star_join(rights, eq(left_on, right_on) per right table): every
        right key column must be unique."""
        if len(rights) != len(ons) or not rights:
            raise FhipError("star_join: size of table is different from size of ops")
        names = set(self._order)
        for r in rights:
            if names & set(r._order):
                raise FhipError("joined table has duplicate column names")
            names |= set(r._order)
        lcols = [self._cols[lo] for lo, _ in ons]
        for lo, _ in ons:
            if lo in self._nulls:
                raise FhipError("join keys with NULL rows are not supported")
        rk = [(r.column(ro), r._row_ids()) for r, (_, ro) in zip(rights, ons)]
        lo, ros = _df.star_join(lcols, rk, left_idx=self._row_ids())
        t = DFTable(device=self._device)
        for side, ids in [(self, lo)] + list(zip(rights, ros)):
            for c in side._order:
                t._cols[c] = _df.global_extract(side._cols[c], ids)
                nl = side._nulls.get(c)
                if nl is not None:
                    n = _df.extract_nulls(ids, nl, len(side._cols[c]))
                    if len(n):
                        t._nulls[c] = n
                t._order.append(c)
        return t


class GroupedDFTable:
    """group_by(cols) of a DFTable (grouped_dftable, dftable.hpp:1120-1160)."""

    AGGS = ("sum", "count", "avg", "max", "min")

    def __init__(self, table, keys):
        self.table = table
        self.keys = keys

    def select(self, names, aggs=()):
        """select(key columns, [(op, column, as_name), ...]): one row per
        group, groups in lexicographic key order."""
        for c in names:
            if c not in self.keys:
                raise FhipError("select of grouped table: not a group-by column: " + c)
        t = self.table
        for op, c, _ in aggs:
            if op not in self.AGGS:
                raise FhipError(f"unknown aggregation {op!r}")
            t._check(c)
        for k in self.keys:
            if t._nulls.get(k) is not None:
                raise FhipError("group_by keys with NULL rows are not supported")
        gi, split, keys = _df.multi_group_by([t._cols[k] for k in self.keys],
                                             local_idx=t._row_ids())
        out = DFTable(device=t._device)
        for c in names:
            out._cols[c] = keys[self.keys.index(c)]
            out._order.append(c)
        for op, c, as_name in aggs:
            nl = t._nulls.get(c)
            r = _df.grouped_agg(op, t._cols[c], gi, split, nulls=nl)
            out._cols[as_name] = r
            out._order.append(as_name)
            if nl is None or op == "count":
                continue
            # the NULL groups of the result (dfcolumn_grouped_agg.cc): sum
            # and avg where a group has no non-NULL value (:232-270); max
            # and min where the result is the NULL value (:492-502,
            # :708-712; the reference finds max's NULLs before turning
            # lowest() into max(), so a group whose true maximum is max()
            # differs here)
            if op in ("sum", "avg"):
                cnt = _df.grouped_agg("count", t._cols[c], gi, split, nulls=nl)
                rn = _df.filter(cnt, "eq", np.uint64(0))
            else:
                dt = r.dtype
                mx = np.finfo(dt).max if dt.kind == "f" else np.iinfo(dt).max
                rn = _df.filter(r, "eq", dt.type(mx))
            if len(rn):
                out._nulls[as_name] = rn
        return out
