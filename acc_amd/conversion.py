"""Neutron unit conversions (mcni.utils.conversion, used by acc/neutron.py:69-83).

mcni is not vendored with the reference; the values are McStas' published
constants.  K2V is taken as 1/V2K.
"""
import math

V2K = 1.58825361e-3   # v [m/s] -> k [1/AA]
K2V = 1.0 / V2K       # k [1/AA] -> v [m/s]
SE2V = 437.393377     # sqrt(E [meV]) -> v [m/s]
VS2E = 5.22703725e-6  # v^2 [m^2/s^2] -> E [meV]


def e2v(e):
    return math.sqrt(e) * SE2V


def v2e(v):
    return v * v * VS2E


def v2k(v):
    return v * V2K


def k2v(k):
    return k * K2V
