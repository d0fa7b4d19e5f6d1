import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "senga-versiona_amd"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests", "golden"), os.path.join(ROOT, "senga-versiona_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


@pytest.fixture(scope="session")
def oracle_built():
    """The oracle shared library (built here if missing)."""
    import build as _build
    path = os.path.join(ROOT, "oracle", "lib", "liboracle.so")
    if not os.path.exists(path):
        _build.build_oracle()
    return path


def rel_err(a, b):
    """max|a-b| / max|b| (max-norm relative error of a field)."""
    import numpy as np
    den = float(np.max(np.abs(b)))
    return float(np.max(np.abs(a - b))) / (den if den > 0 else 1.0)


def pw_err(a, b, floor=1e-6):
    """Pointwise relative error with a floor: max_i |a_i-b_i| / max(|b_i|, floor*max|b|).
    Unlike rel_err it sees a large local error in a small value (a minor
    species far below its peak, a velocity component near zero) down to the
    floor, given as a fraction of the field's maximum."""
    import numpy as np
    b = np.asarray(b)
    den = np.maximum(np.abs(b), floor * float(np.max(np.abs(b))))
    den = np.where(den > 0, den, 1.0)
    return float(np.max(np.abs(np.asarray(a) - b) / den))


def field_errors(run_a, T_a, run_b, T_b, floor=1e-6):
    """(max-norm errors per conserved field, pointwise-floored errors per
    field, T max-norm, T pointwise) of state a against reference state b."""
    n = run_b.shape[0]
    return ([rel_err(run_a[v], run_b[v]) for v in range(n)],
            [pw_err(run_a[v], run_b[v], floor) for v in range(n)],
            rel_err(T_a, T_b), pw_err(T_a, T_b, floor))


def assert_state_close(tag, run_a, T_a, run_b, T_b, tol, tol_T, tol_pw=None, floor=1e-6):
    """Print both metrics per field and check them: max-norm relative error
    <= tol on every conserved field, <= tol_T on T, and (when tol_pw is given)
    the pointwise-floored error <= tol_pw on every field and on T."""
    mx, pw, eT, pT = field_errors(run_a, T_a, run_b, T_b, floor)
    print(tag, "max-norm", " ".join(f"{e:.1e}" for e in mx), f"T {eT:.1e}", flush=True)
    print(tag, f"pointwise(floor {floor:g})", " ".join(f"{e:.1e}" for e in pw), f"T {pT:.1e}", flush=True)
    assert max(mx) <= tol, f"{tag}: max-norm error {max(mx):.2e} > {tol:g}"
    assert eT <= tol_T, f"{tag}: T error {eT:.2e} > {tol_T:g}"
    if tol_pw is not None:
        assert max(pw) <= tol_pw and pT <= tol_pw, f"{tag}: pointwise error {max(max(pw), pT):.2e} > {tol_pw:g}"
    return mx, pw, eT, pT
