"""Host-side pieces of the learned corrector (no GPU)."""
from fvm_solver_amd.learned.corrector import base_coefficients


def test_base_coefficients_integrate_the_line():
    """The base integrates the straight line through (-h, fh) and (0, f0) exactly: checked
    on f(t) = 1 + t against the integrals over [0, W/2] and [0, W]."""
    for k, settle in ((8, 0), (8, 1), (4, 2)):
        (am, bm), (ae, be) = base_coefficients(k, settle)
        W = 1.0
        h = W / 2 + settle * W / k
        f0, fh = 1.0, 1.0 - h
        assert abs(W * (am * f0 + bm * fh) - (W / 2 + W * W / 8)) < 1e-12
        assert abs(W * (ae * f0 + be * fh) - (W + W * W / 2)) < 1e-12
    assert base_coefficients(8, 0) == ((0.75, -0.25), (2.0, -1.0))
