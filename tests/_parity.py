"""Parity bars shared by the GPU tests (the library through its C ABI against
the oracle).

  * RHS: velocity components within tol x v0; each particle's phase derivative
    within tol x max(|a_ref|, term scale), the term scale being the mean
    absolute term of that particle's sum (sigma (1/n) sum |sin| for the
    chimera, (1/(n0+n1)) sum |W sin| + (1/n2) sum |kappa sin| for the vortex
    arrays), so that cancellation in the sum does not loosen or tighten the
    bar.  A particle without partners must keep its input value exactly.
  * single steps: each component within tol x the largest increment of that
    component over the step, plus 2 ulp of the largest state value (the
    rounding of the update itself);
  * trajectories: the single-step bar summed over the steps taken.
tol is 1e-12 for the FP64 forms and 1e-5 for the FP32-compensated form.
"""
import numpy as np

TOL = {"fp64": 1e-12, "fp64f": 1e-12, "fp32k": 1e-5}


def periodic_diff(a, b, L):
    d = a - b
    d[0::3] -= np.round(d[0::3] / L) * L
    d[1::3] -= np.round(d[1::3] / L) * L
    return d


def assert_rhs_terms(gpu, ref, tscale, v0, tol, what, untouched=None):
    """Per-particle bars on a k-form RHS (AoS [vx, vy, a_phi]); returns the
    largest error-to-bar ratio and the plain relative error of a_phi."""
    dv = np.abs(gpu - ref)
    vel = max(dv[0::3].max(), dv[1::3].max())
    assert vel <= tol * abs(v0), f"{what}: velocity error {vel:.3e} > {tol} x {v0}"
    scale = np.maximum(np.abs(ref[2::3]), tscale)
    err = dv[2::3]
    ratio = np.where(scale > 0, err / np.where(scale > 0, scale, 1.0), np.where(err > 0, np.inf, 0.0))
    worst = int(np.argmax(ratio))
    assert ratio[worst] <= tol, (f"{what}: particle {worst} phase-derivative error {err[worst]:.3e} > {tol} x "
                                 f"max(|a|, term scale) = {tol * scale[worst]:.3e}; "
                                 f"{int((ratio > tol).sum())} particles beyond the bar")
    if untouched is not None:
        assert np.array_equal(gpu[2::3][untouched], ref[2::3][untouched]), f"{what}: untouched entries changed"
    plain = float((err / np.maximum(np.abs(ref[2::3]), 1e-300)).max())
    return float(ratio[worst]), plain


def assert_state(gpu, ref, before, L, tol, what=""):
    """Step parity on the increments: |gpu - ref| <= tol x |increment| scale + 2 ulp of the state."""
    inc = periodic_diff(ref, before, L)
    d = np.abs(periodic_diff(gpu, ref, L))
    worst = 0.0
    for c, name in ((0, "x"), (1, "y"), (2, "phi")):
        scale = max(np.abs(inc[c::3]).max(), 1e-300)
        ulp = 2 * np.spacing(np.abs(ref[c::3]).max())
        err = d[c::3].max()
        assert err <= tol * scale + ulp, f"{what}: {name} error {err:.3e} > {tol} x {scale:.3e} + {ulp:.1e}"
        worst = max(worst, err / (tol * scale + ulp))
    return worst


class TrajectoryBar:
    """The single-step bar accumulated over a trajectory: after k steps each
    component may be off by tol x the sum of its increment scales over those
    steps plus 2 ulp of the state per step, the most that k steps each within
    the single-step bar add up to when errors are not amplified."""

    def __init__(self, tol):
        self.tol = tol
        self.inc = np.zeros(3)
        self.steps = 0

    def check(self, gpu, ref, before, L, what=""):
        inc = np.abs(periodic_diff(ref, before, L))
        self.inc += [inc[c::3].max() for c in range(3)]
        self.steps += 1
        d = np.abs(periodic_diff(gpu, ref, L))
        worst = 0.0
        for c, name in ((0, "x"), (1, "y"), (2, "phi")):
            bar = self.tol * self.inc[c] + 2 * self.steps * np.spacing(np.abs(ref[c::3]).max())
            err = d[c::3].max()
            assert err <= bar, f"{what}: {name} error {err:.3e} > accumulated bar {bar:.3e} ({self.steps} steps)"
            worst = max(worst, err / bar)
        return worst
