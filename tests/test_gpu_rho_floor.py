"""The Chebyshev restart rule doubles 1 - rho every third restart, down to
rho = 0 (plain Jacobi). The 256x256 workload from rest with a plain-Jacobi
delay of 8 iterations restarts at nearly every check in its third frame (the
C++ oracle counts 252 restarts in 20,000 iterations there); without the floor
rho from 0.999999 passes -1 at the 63rd restart, the weights 2 / (2 - rho^2)
pass 2 and the frame ends in a non-finite residual. With the floor the frame
stays finite, whether or not it converges. A restart is a check whose
residual grew since the previous one, so the residual history counts them:
the third frame must restart more than 63 times, past the point where rho
would have fallen below -1."""
import numpy as np
import pytest

import clothhip as ch
from helpers import CFG2, CFG2_WIND

pytestmark = pytest.mark.gpu


def test_repeated_restarts_stay_finite():
    s = ch.Scene(**CFG2)
    s.upload()
    sol = ch.Solver(ch.config(mode=ch.MODE_PD, gravity=-9.81, dt=CFG2["dt"], rho=0.999999,
                              cheb_delay=8, ext_accel=CFG2_WIND, max_iterations=20000,
                              tolerance=1e-5, check_every=50), s.N, s.T, s.E)
    args = s.step_args()
    for frame in range(3):
        sol.step(args)
        st = sol.stats()
        hist = sol.residual_history()
        pos, vel = s.get_state()
        assert np.isfinite(pos).all() and np.isfinite(vel).all(), f"frame {frame}"
        assert np.isfinite(hist).all() and len(hist) > 0, f"frame {frame}"
        assert 0 < st.iterations <= 20000
        if st.converged:
            assert hist[-1] <= 1e-5
        else:
            assert st.iterations == 20000
        if frame < 2:
            # the first two frames converge (the oracle: 7,150 iterations
            # with 15 restarts, then 1,400 with none)
            assert st.converged, f"frame {frame}"
        else:
            growth = int(np.sum(hist[1:] > hist[:-1]))
            print(f"frame 3: {st.iterations} iterations, {growth} restarts", flush=True)
            assert growth > 63
