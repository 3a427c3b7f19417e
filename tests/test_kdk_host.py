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


def test_split_energy_is_the_conserved_one(wdm):
    """The reference split K(dt/2) D(dt) conserves K + (dt_kick/dt)(W + V_ext),
    not K + W + V_ext: 300 oracle steps of a self-gravitating blob in a static
    Hernquist field (src/simulation_main.jl:536-547 restated by the oracle).
    The first varies by the O(dt) splitting error only; the second by tens of
    per cent.  With dt_kick = dt (Lie split) both coincide."""
    from oracle import ssfm_oracle as O
    K_ = wdm.kdk
    shape, X = (32, 32, 32), (4.0, 4.0, 4.0)
    x, y, z, _, _ = O.setup_grid(*X, *shape)
    xxx, yyy, zzz, _ = O.setup_coordinates(x, y, z)
    psi0 = (3.0 * np.exp(-((xxx - 0.5) ** 2 + yyy ** 2 + (zzz + 0.3) ** 2) / 1.5) * np.exp(0.8j * xxx)
            + 0.3 * np.exp(-((xxx + 1.5) ** 2 + (yyy - 1) ** 2 + zzz ** 2)))
    G4 = 2 * np.pi
    V = -20.0 / (np.sqrt((xxx - 6.0) ** 2 + yyy ** 2 + zzz ** 2) + 2.0)
    dt = 2e-3
    rec = []

    def red(psi):
        K, W, Vx = O.energies(psi, X, four_pi_G=G4, V_field=V)
        return {"kinetic": K, "w_self": W, "v_ext": Vx}

    def cb(s, psi):
        if (s + 1) % 30 == 0:
            rec.append(red(psi))

    rec.append(red(psi0))
    O.run_ssfm(psi0, X, 300, dt, four_pi_G=G4, V=V, callback=cb)
    e_split = np.array([K_.split_energy(r, dt, 0.5 * dt) for r in rec])
    e_naive = np.array([r["kinetic"] + r["w_self"] + r["v_ext"] for r in rec])
    assert np.ptp(e_split) / abs(e_split[0]) < 5e-3
    assert np.ptp(e_naive) / abs(e_naive[0]) > 0.1
    assert math.isclose(K_.split_energy(rec[0], dt, dt), e_naive[0], rel_tol=1e-14)
    assert math.isclose(K_.split_energy(rec[0], dt, 0.5 * dt, split="strang"), e_naive[0], rel_tol=1e-14)
