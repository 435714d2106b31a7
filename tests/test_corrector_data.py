"""The trained corrector parameters shipped with the package (CPU only): present, finite,
of the network's size, and trained with the configuration the package declares."""
import importlib.util
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _cnn_oracle():
    spec = importlib.util.spec_from_file_location("cnn_oracle", os.path.join(ROOT, "oracle", "cnn_oracle.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_trained_parameters_present(gc):
    from gridcalc_amd import corrector as cr
    p, meta = cr.load_params(256, 256)
    assert p is not None, "trained parameters missing from the package data"
    assert p.dtype == np.float32 and p.shape == (gc.N_CNN_PARAMS,)
    assert np.isfinite(p).all()
    assert meta["scale"] > 0
    assert meta["config"] == cr.CONFIG
    # the loss went down over training
    assert meta["history"][-1] < 0.1 * meta["history"][0]


def test_init_matches_reference_network_layout(gc):
    from gridcalc_amd import corrector as cr
    cnn = _cnn_oracle()
    assert cnn.N_PARAMS == gc.N_CNN_PARAMS
    np.testing.assert_array_equal(cr.init_params(3), cnn.init_params(3).astype(np.float32))


def test_phase_parameters_present(gc):
    """One fine-tuned copy per phase of a run: present, finite, of the network's size, with
    the base network's output scale, and made by the fine-tuning recipe the package declares
    (phase ranges ordered, inside the run, starting after the plain warm-up)."""
    from gridcalc_amd import corrector as cr
    base, bmeta = cr.load_params(256, 256)
    end = 0
    for name, lo, hi, stages in cr.PHASES:
        p, meta = cr.load_params(256, 256, name)
        assert p is not None, f"phase {name} parameters missing from the package data"
        assert p.dtype == np.float32 and p.shape == (gc.N_CNN_PARAMS,)
        assert np.isfinite(p).all()
        assert not np.array_equal(p, base)
        assert meta["scale"] == bmeta["scale"]
        assert meta["macro_steps"] == [lo, hi]
        assert meta["stages"] == [list(s) for s in stages]
        assert meta["config"] == cr.CONFIG
        assert len(meta["history"]) == sum(s[0] for s in stages)
        assert lo >= max(end, cr.WARM) and hi > lo
        end = hi
