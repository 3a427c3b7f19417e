"""Host logic of the solver-step mirror (wavedm.jl_amd/kdk.py) that needs no
device: the gravity configuration's baryon modes and the reference's
time-step rule on host arrays (src/physics/poisson.jl:35-50)."""
import math

import numpy as np
import pytest


def test_gravity_config_baryon_modes(wdm):
    K = wdm.kdk
    assert K.GravityConfig().baryon_potential() is None
    assert K.GravityConfig(baryon_mode=":ignored", Phi_b=np.ones((2, 2, 2))).baryon_potential() is None
    phi_b = (np.arange(24.0) + 0.5j).reshape(2, 3, 4)
    cg = K.GravityConfig(baryon_mode=":mesh", Phi_b=phi_b)
    assert cg.baryon_mode == "mesh"
    got = cg.baryon_potential()
    assert got.dtype == np.float64 and np.array_equal(got, phi_b.real)
    with pytest.raises(ValueError, match="needs the baryon potential"):
        K.GravityConfig(baryon_mode="mesh")
    with pytest.raises(ValueError, match="3-D"):
        K.GravityConfig(baryon_mode="mesh", Phi_b=np.ones(8))
    # the reference's particle baryons (tree gravity) are not on this path
    with pytest.raises(ValueError, match="Unsupported baryon_mode"):
        K.GravityConfig(baryon_mode=":particles_static")


def test_compute_timestep_host_arrays(wdm):
    K = wdm.kdk
    dx, Tmax = 0.1, 0.05
    phi = np.array([[[-3.0, 2.0]]])
    psi = np.array([[[1.0 + 1.0j, 0.5]]])
    t, dt, Nt = K.compute_timestep(dx, phi, 2.0, psi, Tmax, 10, True, 0.5)
    lims = (2 * dx * dx / math.pi, 2 * math.pi / 3.0, 2 * math.pi / 2.0 / math.sqrt(2.0))
    assert Nt == math.ceil(Tmax / (0.5 * min(lims)))
    assert dt == float(np.float32(Tmax / Nt))
    assert len(t) == Nt and t[0] == 0.0 and t[-1] == Tmax
    # autoset off keeps the caller's Nt
    assert K.compute_timestep(dx, phi, 0.0, psi, Tmax, 10, False, 0.5)[2] == 10


def test_baryon_mesh_potential_coefficient(wdm):
    """An explicit four_pi_G = 0 gives a zero potential (the C ABI's scale 0
    would mean the plan's coefficient); None passes the plan's coefficient
    explicitly, a value passes through."""
    K = wdm.kdk

    class FakePlan:
        shape = (2, 3, 4)
        four_pi_G = 7.5

        def __init__(self):
            self.scales = []

        def poisson(self, rho, scale=0.0):
            self.scales.append(scale)
            return np.full(self.shape, scale)

    p = FakePlan()
    rho = np.ones(p.shape)
    out = K.baryon_mesh_potential(p, rho, four_pi_G=0.0)
    assert out.shape == p.shape and not np.any(out) and p.scales == []
    K.baryon_mesh_potential(p, rho)
    K.baryon_mesh_potential(p, rho, four_pi_G=2.0)
    assert p.scales == [7.5, 2.0]
    with pytest.raises(ValueError, match="shape"):
        K.baryon_mesh_potential(p, np.ones((2, 2, 2)))
