"""angelo_amd -- MI355X-native NeuS / Neuralangelo hot path of Instant-angelo.

Drop-in for the reference's model plugin API (hugoycj/Instant-angelo
models/__init__.py:4-16): `register` / `make` with the names 'neus',
'volume-sdf' and 'volume-radiance'; `network_utils.get_encoding` /
`get_mlp`; `VolumeSDF`, `VolumeRadiance`, `NeuSModel`.  Every kernel is a
hand-written HIP (gfx950) launch behind the C-ABI in include/angelo_amd.h.
"""
models = {}


def register(name):
    """models/__init__.py:4-8."""
    def decorator(cls):
        models[name] = cls
        return cls
    return decorator


def make(name, config):
    """models/__init__.py:11-13."""
    return models[name](config)


def _register_all():
    from . import geometry, neus, texture  # noqa: F401  (registration side effects)


try:  # modules register themselves; importing them needs only torch
    _register_all()
except ImportError:  # pragma: no cover - partial checkouts
    pass
