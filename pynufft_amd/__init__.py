"""pynufft_amd: MI355X-native min-max NUFFT with pynufft's NUFFT class API."""
__version__ = "0.1.0"

_LAZY = {
    'NUFFT': ('.nufft', 'NUFFT'),
    'helper': ('.helper', None),
}


def __getattr__(name):
    if name in _LAZY:
        import importlib
        mod_name, attr = _LAZY[name]
        mod = importlib.import_module(mod_name, __name__)
        return mod if attr is None else getattr(mod, attr)
    raise AttributeError(name)
