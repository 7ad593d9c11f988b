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
