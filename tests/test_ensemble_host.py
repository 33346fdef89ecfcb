"""Parameter checks that run before any device work (CPU): options an ensemble cannot
honour, and out-of-range options of a single context, are refused, not ignored."""
import pytest

from rednative import generate


@pytest.mark.parametrize("opt", [dict(radii_enhance_factor=1.0), dict(encounter_hill_factor=3.0),
                                 dict(close_encounter=True), dict(ejection=100.0), dict(hit_centrum=0.1),
                                 dict(frame_center=0), dict(precision=1)])
def test_ensemble_refuses_unsupported_options(native_mod, opt):
    disks = [generate.to_internal(generate.disk_small(n_pl=20, seed=5, giants=False))]
    with pytest.raises(native_mod.RedError) as e:
        native_mod.Ensemble(disks, integrator=4, adaptive=True, tolerance=1e-11, **opt)
    assert e.value.code == 5   # RED_ERR_UNSUPPORTED


@pytest.mark.parametrize("opt", [dict(precision=7), dict(strategy=9), dict(frame_center=5), dict(integrator=9)])
def test_context_refuses_invalid_options(native_mod, opt):
    b = generate.to_internal(generate.disk_small(n_pl=20, seed=5))
    with pytest.raises(native_mod.RedError) as e:
        native_mod.Disk(b, **opt)
    assert e.value.code in (1, 5)   # RED_ERR_ARG, RED_ERR_UNSUPPORTED
