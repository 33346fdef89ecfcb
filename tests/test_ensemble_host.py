"""Parameter checks that run before any device work (CPU): options an ensemble cannot
honour, and out-of-range options of a single context, are refused, not ignored."""
import numpy as np
import pytest

from rednative import generate


@pytest.mark.parametrize("opt", [dict(frame_center=0), dict(precision=1),
                                 dict(inner_steps=False, radii_enhance_factor=1.0)])
def test_ensemble_refuses_unsupported_options(native_mod, opt):
    disks = [generate.to_internal(generate.disk_small(n_pl=20, seed=5))]
    with pytest.raises(native_mod.RedError) as e:
        native_mod.Ensemble(disks, integrator=4, adaptive=True, tolerance=1e-11, **opt)
    assert e.value.code == 5   # RED_ERR_UNSUPPORTED


def test_ensemble_hill_test_needs_a_star(native_mod):
    b = generate.to_internal(generate.disk_small(n_pl=20, seed=5))
    b.types = np.where(b.types == 0, 3, b.types).astype(np.int32)   # the star typed as a protoplanet
    b.counts = [0, b.counts[1], b.counts[2], b.counts[3] + 1, b.counts[4], b.counts[5], b.counts[6]]
    order = np.argsort(b.types, kind="stable")
    for f in ("ids", "types", "epoch", "mass", "radius", "density", "cd", "mig_type", "mig_stop_at", "r", "v"):
        setattr(b, f, getattr(b, f)[order])
    with pytest.raises(native_mod.RedError) as e:
        native_mod.Ensemble([b], integrator=4, adaptive=True, tolerance=1e-11, encounter_hill_factor=3.0)
    assert e.value.code == 6   # RED_ERR_NO_STAR


@pytest.mark.parametrize("opt", [dict(precision=7), dict(strategy=9), dict(frame_center=5), dict(integrator=9)])
def test_context_refuses_invalid_options(native_mod, opt):
    b = generate.to_internal(generate.disk_small(n_pl=20, seed=5))
    with pytest.raises(native_mod.RedError) as e:
        native_mod.Disk(b, **opt)
    assert e.value.code in (1, 5)   # RED_ERR_ARG, RED_ERR_UNSUPPORTED
