"""Non-finite forcing on the GPU: the step flags it, wfh_sync reports WFH_ERR_NONFINITE, the handle
refuses further steps, and loading a saved state repairs it (the continued run then equals a run
that never saw the bad forcing, bitwise). The ensemble reports it as well."""
import numpy as np
import pytest

from test_gpu_parity import STATE_VARS
from wflowhip import Ensemble, Model, synth
from wflowhip.model import WflowError

pytestmark = pytest.mark.gpu


def test_nonfinite_forcing_is_reported_and_recoverable():
    c = synth.make_catchment(32, cell=250.0, dt=3600.0, river_threshold=30, seed=701)
    m = Model(c)
    ref = Model(c)
    for k in range(2):
        f = synth.forcing_numpy(c, k)
        m.run_timestep(*f)
        ref.run_timestep(*f)
    saved = m.save_state()
    for bad_var, value in ((0, np.nan), (1, np.inf), (2, -np.inf)):
        f = list(synth.forcing_numpy(c, 2))
        f[bad_var] = f[bad_var].copy()
        f[bad_var][c.n // 2] = value
        m.run_timestep(*f)
        with pytest.raises(WflowError) as e:
            m.sync()
        assert e.value.code == -4
        with pytest.raises(WflowError) as e:
            m.run_timestep(*synth.forcing_numpy(c, 2))
        assert e.value.code == -6
        m.load_state(saved)
    for k in range(2, 5):
        f = synth.forcing_numpy(c, k)
        m.run_timestep(*f)
        ref.run_timestep(*f)
    for v in STATE_VARS:
        a, b = m.get(v), ref.get(v)
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), v.name
    m.close()
    ref.close()


def test_ensemble_reports_nonfinite_forcing():
    c = synth.make_catchment(24, cell=250.0, dt=3600.0, river_threshold=20, seed=702)
    f1 = np.ones(64)
    ens = Ensemble(c, f1, f1, f1)
    P, E, T = synth.forcing_numpy(c, 0)
    T = T.copy()
    T[3] = np.nan
    ens.run_timestep(P, E, T)
    with pytest.raises(WflowError) as e:
        ens.sync()
    assert e.value.code == -4
    ens.close()
