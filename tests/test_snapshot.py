"""Snapshot file format and the Prop.csv table (host side, no GPU)."""
import numpy as np
import pytest

from conftest import load_wavedm

S = load_wavedm().snapshot


class _FakePlan:
    """Duck-typed stand-in for Plan: the fields and getters snapshot.save reads."""

    def __init__(self, shape, precision):
        rng = np.random.default_rng(1)
        self.shape, self.xmax, self.precision = shape, (3.0, 2.0, 1.5), precision
        self.consistent_grid, self.hbar_over_m, self.four_pi_G, self.self_gravity = False, 0.5, 7.0, True
        dt = np.complex64 if precision == "f32" else np.complex128
        self.psi = np.asfortranarray((rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(dt))
        self.phi = np.asfortranarray(rng.normal(size=shape))
        self.avg = np.asfortranarray(rng.uniform(size=shape))

    def get_psi(self, dtype=np.complex128):
        return self.psi.astype(dtype)

    def get_phi(self):
        return self.phi

    def density_sum(self):
        return self.avg * 4.0, 4

    def mass_ledger_state(self):
        return (self.precision == "f32"), (123.5 if self.precision == "f32" else None)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_snapshot_round_trip(tmp_path, precision):
    p = _FakePlan((6, 5, 4), precision)
    f = tmp_path / "snap.npz"
    S.save(f, p, t=1.25, step=40, dt=0.03125, phi_all=True, average=True, meta={"run": "x"})
    z = S.load(f)
    assert z["format"] == S.FORMAT and z["t"] == 1.25 and z["step"] == 40 and z["dt"] == 0.03125
    assert tuple(z["shape"]) == p.shape and z["precision"] == precision and z["average_count"] == 4
    assert z["psi"].dtype == p.psi.dtype and np.array_equal(z["psi"], p.psi) and z["psi"].flags.f_contiguous
    assert np.array_equal(z["phi_all"], p.phi) and np.array_equal(z["psi2_sum"], p.avg * 4.0)
    assert np.array_equal(z["psi2_averaged"], (p.avg * 4.0) / 4.0)
    assert z["meta"] == {"run": "x"} and z["hbar_over_m"] == 0.5 and z["four_pi_G"] == 7.0
    assert z["mass_ledger"] == (precision == "f32")
    assert z["mass_reference"] == (123.5 if precision == "f32" else None)


def test_snapshot_rejects_foreign_files(tmp_path):
    f = tmp_path / "other.npz"
    np.savez(f, header=np.frombuffer(b'{"format": "something else"}', dtype=np.uint8))
    with pytest.raises(ValueError):
        S.load(f)


def test_prop_table_columns_and_values(tmp_path):
    # save_property_dataframe's column order: t, R, R1..R9, TotalMass, then the
    # virial block PE_abs, KE, QE, Virial, MomentumX..Z
    tab = S.PropTable(virial=True, scale={"R": 2.0, "R1": 2.0})
    diag = {k: float(i) for i, k in enumerate(S.PROP_COLUMNS[1:] + S.VIRIAL_COLUMNS)}
    tab.append(0.5, diag)
    tab.append(0.75, {k: v + 1 for k, v in diag.items()})
    f = tmp_path / "Prop.csv"
    tab.write(f)
    cols, rows = S.PropTable.read(f)
    assert cols == ["t", "R"] + [f"R{i}" for i in range(1, 10)] + ["TotalMass", "PE_abs", "KE", "QE", "Virial",
                                                                    "MomentumX", "MomentumY", "MomentumZ"]
    assert rows[0][0] == 0.5 and rows[0][1] == 0.0 and rows[0][2] == 2.0 and rows[1][2] == 4.0
    assert rows[1][cols.index("MomentumZ")] == diag["MomentumZ"] + 1
    assert S.PropTable(virial=False).columns == S.PROP_COLUMNS


def test_prop_table_total_mass_lags_one_row():
    """The reference pushes total_halo_mass before recomputing it from the new
    state (src/simulation_main.jl:557 vs :652; seeded with the IC mass at
    :349), so TotalMass row k is the mass of row k-1's state, and an infinite
    mass is recorded as 0."""
    cols = S.PROP_COLUMNS[1:]
    states = [dict({c: 1.0 for c in cols}, TotalMass=m) for m in (10.0, float("inf"), 12.0)]
    tab = S.PropTable(virial=False, initial_mass=9.0, scale={"TotalMass": 2.0})
    for k, d in enumerate(states):
        tab.append(0.1 * k, d)
    assert [r["TotalMass"] for r in tab.rows] == [18.0, 20.0, 0.0]
    # without the IC mass the first row carries its own state's mass
    tab = S.PropTable(virial=False)
    for k, d in enumerate(states):
        tab.append(0.1 * k, d)
    assert [r["TotalMass"] for r in tab.rows] == [10.0, 10.0, 0.0]
