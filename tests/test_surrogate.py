"""The learned surrogate end to end on a small scale: spectra from the GPU solver, a trained
ensemble, the plain solver as fallback and judge (what the trained network gains: spectra/s at a
bounded error, with out-of-box requests still exact)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_surrogate_gain_and_fallback():
    from symboltz import surrogate as sg
    ks = np.logspace(-4, np.log10(5.0), 32)
    X, Xh = sg.lhs(256, seed=11), sg.lhs(32, seed=12)
    Y, kern_train, _ = sg.solve_spectra(X, ks_mpc=ks)
    Yh, kern_h, _ = sg.solve_spectra(Xh, ks_mpc=ks)
    S = sg.Surrogate(n_members=2, hidden=(128, 128), n_out=len(ks), max_batch=4096)
    hist, _ = S.fit(X, Y, epochs=400, batch=64, lr=2e-3)
    assert hist[0][-1] < 0.02 * hist[0][0]
    mean, disagree, inbox, _ = S.predict(Xh)
    err = np.abs(np.exp(mean - Yh) - 1.0).max(axis=1)
    assert np.all(inbox) and err.max() < 0.05          # a few percent at this small scale
    # speed: a batch of 4096 spectra from the networks against the plain solver's rate
    S.predict(Xh)
    _, _, _, ms = S.predict(sg.lhs(4096, seed=13))
    assert 4096 / (ms * 1e-3) > 100 * len(Xh) / kern_h
    # out-of-box requests fall back to the solver, which is exact by definition
    Xo = Xh[:4].copy()
    Xo[:, 0] = sg.HI[0] + 0.02
    yo, fb = S.predict_with_fallback(Xo, ks_mpc=ks)
    assert np.all(fb)
    assert np.max(np.abs(yo - sg.solve_spectra(Xo, ks_mpc=ks)[0])) < 1e-12
    S.close()
