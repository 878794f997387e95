"""Minimal histogram object for monitor outputs.

The reference returns `histogram.histogram(name, axes, data=..., errors=...)`
objects and dumps them with `histogram.hdf` (acc/components/monitors/
MonitorBase.py:8-12); that package is not available here.  This class keeps
the same fields (name, axes as (name, centers, unit), I = data, E2 = errors,
i.e. variances) and writes `.npz` files next to the requested file name.
"""
import os

import numpy as np


class Histogram:

    def __init__(self, name, axes, data, errors):
        self.name = name
        self.axes = [(a[0], np.asarray(a[1]), a[2] if len(a) > 2 else "") for a in axes]
        self.I = np.asarray(data, dtype=float)
        self.E2 = np.asarray(errors, dtype=float)

    def shape(self):
        return self.I.shape

    def axis(self, name):
        for a in self.axes:
            if a[0] == name:
                return a[1]
        raise KeyError(name)

    def __truediv__(self, other):
        if isinstance(other, Histogram):
            I = self.I / other.I
            E2 = self.E2 / other.I ** 2 + (self.I ** 2 / other.I ** 4) * other.E2
        else:
            I = self.I / other
            E2 = self.E2 / other ** 2
        return Histogram(self.name, self.axes, I, E2)

    def save(self, filename):
        root, ext = os.path.splitext(filename)
        path = root + ".npz" if ext in (".h5", ".hdf5", "") else filename
        arrays = {"I": self.I, "E2": self.E2, "name": np.array(self.name)}
        for k, (n, c, u) in enumerate(self.axes):
            arrays[f"axis{k}_name"] = np.array(n)
            arrays[f"axis{k}_centers"] = c
            arrays[f"axis{k}_unit"] = np.array(u)
        np.savez(path, **arrays)
        return path


def load(path):
    z = np.load(path, allow_pickle=False)
    axes = []
    k = 0
    while f"axis{k}_name" in z:
        axes.append((str(z[f"axis{k}_name"]), z[f"axis{k}_centers"], str(z[f"axis{k}_unit"])))
        k += 1
    return Histogram(str(z["name"]), axes, z["I"], z["E2"])
