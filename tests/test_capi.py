"""The C-ABI library loads and exports every entry point its header declares; host-side errors
come back as status codes with a message (no GPU needed for any of these)."""
import ctypes
import os
import re

import numpy as np
import pytest

import nairnmpm as nm
from nairnmpm import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nairnmpm_gpu.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(mpmgpu_\w+)\s*\(", text)))


def test_header_symbols_are_exported():
    names = declared()
    assert len(names) >= 25
    lib = nm.library()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared()) <= set(nm._SIGNATURES) | {"mpmgpu_problem_destroy", "mpmgpu_destroy"}


def test_kernel_group_names():
    lib = nm.library()
    got = [lib.mpmgpu_kernel_group_name(i).decode() for i in range(len(nm.KERNEL_GROUPS))]
    assert got == nm.KERNEL_GROUPS
    assert lib.mpmgpu_kernel_group_name(len(nm.KERNEL_GROUPS)) is None or \
        lib.mpmgpu_kernel_group_name(len(nm.KERNEL_GROUPS)) == b""


def test_problem_errors_carry_messages():
    p = nm.Problem.create(nm.THREED)
    with pytest.raises(nm.MPMError, match="."):
        p.add_point(0, [1., 1., 1.])                 # no grid yet
    p.set_grid([0, 4, 0, 4, 0, 4], [1, 1, 1])
    mid = p.add_material(nm.MAT_ISOTROPIC, "a", E=10., nu=0.3, rho=1.)
    with pytest.raises(nm.MPMError, match="not within the grid"):
        p.add_point(mid, [40., 1., 1.])              # outside the grid
    with pytest.raises(nm.MPMError, match="."):
        p.set("NoSuchSetting", 1.)
    with pytest.raises(nm.MPMError, match="."):
        nm.Problem.create(11)                        # plane stress is off this path


def test_bad_material_constants_rejected():
    p = nm.Problem.create(nm.THREED)
    p.set_grid([0, 4, 0, 4, 0, 4], [1, 1, 1])
    mid = p.add_material(nm.MAT_ISOTROPIC, "a", E=10., nu=0.6, rho=1.)
    p.add_point(mid, [1.5, 1.5, 1.5])
    with pytest.raises(nm.MPMError, match="Poisson"):
        p.describe()


def test_undefined_material_rejected_at_setup():
    """as in the reference, a point's material is checked when the particles are initialized
    (NairnMPM::PreliminaryParticleCalcs), not when it is read"""
    p = nm.Problem.create(nm.THREED)
    p.set_grid([0, 4, 0, 4, 0, 4], [1, 1, 1])
    p.add_material(nm.MAT_ISOTROPIC, "a", E=10., nu=0.3, rho=1.)
    p.add_point(5, [1.5, 1.5, 1.5])
    with pytest.raises(nm.MPMError, match="undefined material"):
        p.describe()


def test_create_fails_loudly_without_a_device():
    """Without a GPU the product path reports an error; it never falls back to a CPU step."""
    prob = nm.Problem.from_xml(configs.elastic_block_2d(cells=8, block=4))
    try:
        sim = nm.Simulation.from_problem(prob)
    except nm.MPMError as e:
        assert str(e)
        return
    # a device is present: the simulation is a GPU one
    sim.close()


def test_static_solve_refuses_sizes_beyond_its_scratch():
    """mpmgpu_solve_static checks its size against the static array before any device call."""
    prob = nm.Problem.from_xml(configs.j2_block_3d(cells=128, lo=16, hi=112))
    g, mats, run, n, _ = prob.describe()
    with pytest.raises(nm.MPMError, match="bytes of static scratch"):
        nm.solve_static(g, mats, run, n, 0x1000, 0x1000, 1)
    with pytest.raises(nm.MPMError, match="bad step count"):
        nm.solve_static(g, mats, run, n, 0x1000, 0x1000, -1)
