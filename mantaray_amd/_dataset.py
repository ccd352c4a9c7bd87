"""A minimal stand-in for xarray.Dataset, used when xarray is not installed.

mantaray.core returns an xarray.Dataset (python/mantaray/core.py:56-63,
:114-132).  This class keeps the parts callers of that API use: named
variables with dimension names, `sizes`, `coords` / `data_vars`, `attrs`,
attribute and item access, and numpy arrays underneath.
"""
from __future__ import annotations

import numpy as np


class DataArray:
    def __init__(self, name, dims, values):
        self.name = name
        self.dims = tuple(dims)
        self.values = np.asarray(values)

    @property
    def shape(self):
        return self.values.shape

    @property
    def sizes(self):
        return dict(zip(self.dims, self.values.shape))

    def __array__(self, dtype=None, copy=None):
        return self.values if dtype is None else self.values.astype(dtype)

    def __len__(self):
        return len(self.values)

    def __getitem__(self, k):
        return self.values[k]

    def _wrap(self, v):
        return DataArray(self.name, self.dims, v)

    def __eq__(self, o):
        return self._wrap(self.values == np.asarray(o))

    def __ne__(self, o):
        return self._wrap(self.values != np.asarray(o))

    def __lt__(self, o):
        return self._wrap(self.values < np.asarray(o))

    def __gt__(self, o):
        return self._wrap(self.values > np.asarray(o))

    def all(self):
        return bool(np.all(self.values))

    def any(self):
        return bool(np.any(self.values))

    def isnull(self):
        return self._wrap(np.isnan(self.values))

    def __repr__(self):
        return f"<DataArray {self.name} {dict(self.sizes)}>"


class Dataset:
    def __init__(self, data_vars=None, coords=(), attrs=None):
        self._vars = {}
        for name, (dims, vals) in (data_vars or {}).items():
            self._vars[name] = DataArray(name, dims, vals)
        self._coords = list(coords)
        self.attrs = dict(attrs or {})

    def set_coords(self, names):
        for n in names:
            if n not in self._coords:
                self._coords.append(n)
        return self

    @property
    def sizes(self):
        out = {}
        for v in self._vars.values():
            out.update(v.sizes)
        return out

    dims = sizes

    @property
    def coords(self):
        return {n: self._vars[n] for n in self._coords if n in self._vars}

    @property
    def data_vars(self):
        return {n: v for n, v in self._vars.items() if n not in self._coords}

    @property
    def variables(self):
        return dict(self._vars)

    def __contains__(self, name):
        return name in self._vars

    def __getitem__(self, name):
        return self._vars[name]

    def __setitem__(self, name, value):
        if isinstance(value, DataArray):
            self._vars[name] = value
        else:
            vals = np.asarray(list(value) if isinstance(value, range) else value)
            self._vars[name] = DataArray(name, (name,), vals)
            if name not in self._coords:
                self._coords.append(name)

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        try:
            return self._vars[name]
        except KeyError:
            raise AttributeError(name) from None

    def __repr__(self):
        return f"<Dataset sizes={self.sizes} vars={list(self._vars)}>"
