"""The external potentials of the workloads as host callables V(xxx, yyy, zzz,
psi, t) for the oracle (the form of the reference's V callback,
src/solver/KDK.jl:64), built from the same configuration the device models
take: the Hernquist host -G M / (|x - c| + a) and the rotating quadrupole
1/2 (x - c)^T R(t) T0 R(t)^T (x - c), R(t) a rotation by 2 pi t / period about
the configured axis (Rodrigues' formula)."""
import math

import numpy as np


def rotation(axis, angle):
    n = np.asarray(axis, dtype=np.float64)
    n = n / np.linalg.norm(n)
    K = np.array([[0, -n[2], n[1]], [n[2], 0, -n[0]], [-n[1], n[0], 0]])
    return np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * K @ K


def oracle_V(cfg):
    pot = cfg["potential"]
    if pot is None:
        return None
    kind, kw = pot
    if kind == "hernquist":
        G = cfg["four_pi_G"] / (4 * math.pi)
        M, a, c = kw["M"], kw["a"], kw["center"]

        def V(xxx, yyy, zzz, psi, t):
            return -G * M / (np.sqrt((xxx - c[0]) ** 2 + (yyy - c[1]) ** 2 + (zzz - c[2]) ** 2) + a)
        return V
    if kind == "quadrupole":
        T0, axis, period, c = np.asarray(kw["T0"]), kw["axis"], kw["period"], kw["center"]

        def V(xxx, yyy, zzz, psi, t):
            R = rotation(axis, 2 * math.pi * t / period)
            T = R @ T0 @ R.T
            d = (xxx - c[0], yyy - c[1], zzz - c[2])
            out = np.zeros(np.shape(xxx))
            for i in range(3):
                for j in range(3):
                    out += 0.5 * T[i, j] * d[i] * d[j]
            return out
        return V
    raise KeyError(kind)
