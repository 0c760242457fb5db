"""Shared helpers of the parity tests."""
import numpy as np

import nairnmpm as nm

F = nm.F


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    num = np.linalg.norm(a - b)
    if den == 0.0:
        return num
    return num / den


def full_stress(d, dim):
    """rho*(sp - P I) up to the constant rho0 factor: the archived stress (ArchiveData.cpp:889-921)."""
    comps = ["sp_xx", "sp_yy", "sp_zz", "sp_xy"] + (["sp_xz", "sp_yz"] if dim == 3 else [])
    out = []
    for c in comps:
        s = d[F[c]].copy()
        if c in ("sp_xx", "sp_yy", "sp_zz"):
            s = s - d[F["pressure"]]
        out.append(s)
    return np.stack(out)


def compare(gpu, ref, dim):
    """relative L2 of positions, displacements, velocities and stresses"""
    ax = ["x", "y", "z"][:dim]
    pos_g = np.stack([gpu[F["pos_" + a]] for a in ax])
    pos_r = np.stack([ref[F["pos_" + a]] for a in ax])
    org = np.stack([ref[F["orig_" + a]] for a in ax])
    vel_g = np.stack([gpu[F["vel_" + a]] for a in ax])
    vel_r = np.stack([ref[F["vel_" + a]] for a in ax])
    return {
        "pos": rel_l2(pos_g, pos_r),
        "disp": rel_l2(pos_g - org, pos_r - org),
        "vel": rel_l2(vel_g, vel_r),
        "stress": rel_l2(full_stress(gpu, dim), full_stress(ref, dim)),
    }
