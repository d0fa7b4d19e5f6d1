"""The C oracle against the reference's own outputs (golden fixtures).

The fixtures were written by the compiled reference (tests/golden/make_golden.py);
the oracle restates the reference in the same evaluation order, so agreement is
expected to the last bit.
"""
import numpy as np
import pytest

import golden_cases
from pyoracle import Oracle


@pytest.mark.parametrize("name", golden_cases.NAMES)
def test_tempin_rhs_and_steps_bitexact(name, oracle_built):
    c, fx = golden_cases.load(name)
    o = Oracle(c)
    o.set_timestep(c.dump.tstep)
    o.set_state(c.dump.state, 0.0)
    _, T, P, _ = o.get()
    assert np.array_equal(T, fx["T0"])
    assert np.array_equal(o.itndex(), fx["itndex0"])
    rhs = o.rhs(1)
    assert np.array_equal(rhs, fx["rhs"])
    _, T1, P1, rate1 = o.get()
    assert np.array_equal(T1, fx["T1"])
    assert np.array_equal(P1, fx["P1"])
    assert np.array_equal(rate1, fx["rate1"])
    o.set_state(c.dump.state, 0.0)
    o.step(int(fx["nsteps"]))
    run, T2, P2, rate2 = o.get()
    assert np.array_equal(run, fx["run2"])
    assert np.array_equal(T2, fx["T2"])
    assert np.array_equal(rate2, fx["rate2"])


@pytest.mark.parametrize("name", sorted(golden_cases.OPS))
def test_operators_bitexact(name, oracle_built):
    fx = dict(np.load(golden_cases.os.path.join(golden_cases.HERE, name + ".npz")))
    c = golden_cases.build(golden_cases.OPS[name])
    o = Oracle(c)
    out = o.derivatives(fx["field"])
    assert np.array_equal(out, fx["ops"])


def test_adaptive_step_bitexact(oracle_built):
    """ADAPTT: error register, global error norm and PID controller, against
    the reference run with the adaptive step switch on."""
    c, fx = golden_cases.load(golden_cases.ADAPT)
    o = Oracle(c)
    o.set_timestep(c.dump.tstep)
    o.set_adaptive(True, float(fx["errold"]), float(fx["errldr"]))
    o.set_state(c.dump.state, 0.0)
    o.step(int(fx["nsteps"]))
    run, T2, _, _ = o.get()
    tstep, _, _ = o.timestep()
    assert tstep == float(fx["tstep2"])
    assert tstep != float(fx["tstep"])
    assert np.array_equal(run, fx["run2"])
    assert np.array_equal(T2, fx["T2"])


@pytest.mark.parametrize("name,flag", [("h2_nscbc_mixwp_12x11x11", "nfmixw"), ("h2_nscbc_mixwp_12x11x11", "nfmixp"),
                                       ("h2_nscbc_sordf_12x11x11", "nfmsor"), ("h2_nscbc_sordf_12x11x11", "nfmduf"),
                                       ("h2_per_duf_11x12x11", "nfmduf"), ("h2_nscbc_alltr_12x11x11", "nfmixw"),
                                       ("h2_nscbc_alltr_12x11x11", "nfmsor"), ("ch4_16sp_nscbc_12x11x11", "nfmduf"),
                                       ("ch4_16sp_per_11x12x11", "nfmixp")])
def test_transport_fixture_exercises_each_term(name, flag, oracle_built):
    """The molecular-transport fixtures pin each term they switch on: with
    that switch off the oracle no longer reproduces the reference's
    right-hand side (the density and momentum are untouched by them)."""
    c, fx = golden_cases.load(name)
    setattr(c.control, flag, 0)
    o = Oracle(c)
    o.set_timestep(c.dump.tstep)
    o.set_state(c.dump.state, 0.0)
    rhs = o.rhs(1)
    assert np.array_equal(rhs[:4], fx["rhs"][:4])
    assert not np.array_equal(rhs[4:], fx["rhs"][4:])


@pytest.mark.parametrize("name", ["h2_nscbc_radn_12x11x11", "h2_per_radn_cle_11x12x11", "ch4_16sp_nscbc_12x11x11"])
def test_radiation_fixture_exercises_the_loss(name, oracle_built, tmp_path):
    """The radiation fixtures pin the loss term: without it only the energy
    right-hand side changes, by far more than the parity tolerance; with the
    H2O absorption polynomial's constant term doubled in radn.dat (a planted
    mistake in the data the oracle reads) the energy no longer matches."""
    c, fx = golden_cases.load(name)
    ref = fx["rhs"]
    c.control.nfradn = 0
    o = Oracle(c)
    o.set_timestep(c.dump.tstep)
    o.set_state(c.dump.state, 0.0)
    rhs = o.rhs(1)
    assert np.array_equal(np.delete(rhs, 4, axis=0), np.delete(ref, 4, axis=0))
    assert np.abs(rhs[4] - ref[4]).max() / np.abs(ref[4]).max() > 1e-8
    c.control.nfradn = 1
    txt = open(c.radn_path).read().replace("6.869480D-04", "1.373896D-03")
    p = tmp_path / "radn.dat"
    p.write_text(txt)
    c.radn_path = str(p)
    o = Oracle(c)
    o.set_timestep(c.dump.tstep)
    o.set_state(c.dump.state, 0.0)
    rhs = o.rhs(1)
    assert np.array_equal(rhs[:4], ref[:4])
    assert not np.array_equal(rhs[4], ref[4])
