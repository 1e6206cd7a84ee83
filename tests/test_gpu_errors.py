"""Error paths of the entry points added for the X+ potential, P/D internal
states, real time and the moving impurity: each refusal returns its status
code (the reference STOPs at the matching points) and leaves the context
usable."""
import numpy as np
import pytest

import eff_pot_amd as ep
from eff_pot_amd.abi import make_params, make_morse, SolveOpts, PREC_MIXED, Plus

pytestmark = pytest.mark.gpu

ERR_ARG, ERR_POT, ERR_STATE = 1, 4, 7


def _code(fn):
    with pytest.raises(ep.EffpotError) as e:
        fn()
    return e.value.code


def test_refusals(gpu_available):
    print("[errors] refusals", flush=True)
    real = make_params(32, xmax=12.0, n4=100.0, psi_complex=0)
    with ep.Context(real) as c:
        assert _code(lambda: c.set_real_time(True)) == ERR_ARG           # real psi
        assert _code(lambda: c.set_impurity_motion([0, 0, 1], 40.0)) == ERR_STATE  # no impurity yet
        bad = Plus()
        bad.kind = 9
        assert _code(lambda: c.set_plus(bad)) == ERR_POT
        st = ep.make_state_imp(ep.STATE_P, instate=3)                    # P has 3 Kramers pairs
        assert _code(lambda: c.set_impurity_state(st)) == ERR_ARG
        st = ep.make_state_imp(ep.STATE_D, instate=0)
        st.v_pi.kind = 9
        assert _code(lambda: c.set_impurity_state(st)) == ERR_POT
        # instates' choices of invar: an Ljz outside the j manifold, P-only
        # modes and fixed (exciplex) moments on a D state, a negative amplitude
        for state, instate, sel in ((ep.STATE_P, 0, dict(mode=ep.SELECT_JZ, ljz="+3/2")),
                                    (ep.STATE_D, 2, dict(mode=ep.SELECT_JZ, ljz="+7/2")),
                                    (ep.STATE_D, 0, dict(mode=ep.SELECT_FIX)),
                                    (ep.STATE_D, 0, dict(fixed_moments=True)),
                                    (ep.STATE_P, 0, dict(mode=ep.SELECT_MIXTURE, a_sig_1o2=-1.0)),
                                    (ep.STATE_P, 0, dict(mode=7))):
            st = ep.make_state_imp(state, instate=instate, select=ep.make_select(**sel))
            assert _code(lambda: c.set_impurity_state(st)) == ERR_ARG, sel
        # a P-state impurity cannot move (its invar equation of motion is not built)
        c.init_fermi(0.0218, 2.22 * 100 ** (1 / 3), 0.6, 1e-6, 1)
        c.set_impurity_state(ep.make_state_imp(ep.STATE_P, instate=0))
        assert _code(lambda: c.set_impurity_motion([0, 0, 1], 40.0)) == ERR_STATE
        # evolving invar needs a P/D state; relaxation refuses it
        c.set_state_dynamics(True)
        opts = SolveOpts(dt=1e-4, tol=1e-8, n_rk=3, pener=50, max_steps=10, polish_steps=0, tol_fp32=0.0,
                         check_finite=1)
        assert _code(lambda: c.solve(opts)) == ERR_STATE
        c.set_state_dynamics(False)
        # still usable
        c.set_impurity(make_morse())
        assert _code(lambda: c.set_state_dynamics(True)) == ERR_STATE  # isotropic impurity
        c.poten()
        assert np.isfinite(c.energy()[0]["etot"])
    cplx = make_params(32, xmax=12.0, n4=100.0, psi_complex=1)
    with ep.Context(cplx) as c:
        c.init_fermi(0.0218, 2.22 * 100 ** (1 / 3), 0.6, 1e-6, 1)
        c.set_impurity(make_morse())
        c.set_real_time(True)
        opts = SolveOpts(dt=1e-4, tol=1e-8, n_rk=3, pener=50, max_steps=100, polish_steps=0, tol_fp32=0.0,
                         check_finite=1)
        assert _code(lambda: c.solve(opts)) == ERR_STATE                 # relaxation refuses real time
        c.set_real_time(False)
        c.set_impurity_motion([0, 0, 1], 40.0)
        assert _code(lambda: c.solve(opts)) == ERR_STATE                 # ... and a moving impurity
        c.set_impurity_motion([0, 0, 0], 0.0)                            # mass 0 stops it
        r = c.solve(opts)[0]
        assert r.steps == 100
    mixed = make_params(32, xmax=12.0, n4=100.0, precision=PREC_MIXED, psi_complex=1)
    with ep.Context(mixed) as c:
        c.init_fermi(0.0218, 2.22 * 100 ** (1 / 3), 0.6, 1e-6, 1)
        c.set_impurity(make_morse())
        assert _code(lambda: c.set_real_time(True)) == ERR_ARG
        assert _code(lambda: c.set_impurity_motion([0, 0, 1], 40.0)) == ERR_ARG
        c.set_impurity_state(ep.make_state_imp(ep.STATE_D, instate=0))
        assert _code(lambda: c.set_state_dynamics(True)) == ERR_ARG
