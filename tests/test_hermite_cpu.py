"""The Hermite harness with the FP64 oracle as regular-force provider
conserves energy on the N=1024 Plummer sphere (pins the harness itself)."""
from abyss_amd import workloads

from tests import hermite_util


def test_oracle_driven_hermite_conserves_energy(tmp_path):
    js = workloads.plummer_equal(1024)
    ic = str(tmp_path / "ic.bin")
    hermite_util.write_ic(ic, js)
    out = hermite_util.run("oracle", ic, 0.125)
    assert abs(out["dE_over_E"]) < 1e-8, out
    assert out["blocks"] > 10 and out["particle_steps"] > js.n
    assert max(abs(c) for c in out["com"]) < 1e-7, out


def test_oracle_driven_ahmad_cohen_conserves_energy(tmp_path):
    # two time scales: irregular steps over resident lists, regular steps
    # with the neighbour-set correction and the regular corrector, radius
    # adaptation towards 48 neighbours -- all steps by the oracle
    js = workloads.plummer_equal(1024)
    ic = str(tmp_path / "ic.bin")
    hermite_util.write_ic(ic, js)
    out = hermite_util.run_ac("oracle", ic, 1.0, 0.02, 0.001, 48)
    assert abs(out["dE_over_E"]) < 5e-7, out
    assert out["overflows"] == 0
    assert out["regular_steps"] < 0.7 * out["irregular_steps"], out
    assert 35 < out["mean_nnb"] < 60, out
    assert max(abs(c) for c in out["com"]) < 1e-6, out
