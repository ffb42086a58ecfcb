"""The pybind11 module _core: the reference's standalone LibRawWrapper methods
(core/python_bindings.cpp:320-351) with the same names, argument names and
defaults, and the same argument error (LibRawInvalidArgument for an image that
is not (H, W, 3), core/libraw_wrapper.cpp:3066-3069, :3100-3104, :3137-3141,
:3181-3185).  No device work here: the shape checks run before the HIP
context is made."""
import numpy as np
import pytest

core = pytest.importorskip("libraw_enhanced_amd._core", reason="_core not built (run __graft_entry__.build())")


def _sig(fn):
    return fn.__doc__.splitlines()[0]


def test_methods_and_defaults():
    w = core.LibRawWrapper()
    assert "radius: typing.SupportsFloat | typing.SupportsIndex = 10.0" in _sig(w.defringe)
    assert "strength: typing.SupportsFloat | typing.SupportsIndex = 10.0" in _sig(w.defringe)
    assert "defringe_green: bool = False" in _sig(w.defringe)
    assert "after_scale: typing.SupportsFloat | typing.SupportsIndex = 1.0" in _sig(w.tone_mapping)
    s = _sig(w.enhance_micro_contrast)
    assert "threshold: typing.SupportsFloat | typing.SupportsIndex = -1.0" in s
    assert "strength: typing.SupportsFloat | typing.SupportsIndex = 8.0" in s
    assert "target_contrast: typing.SupportsFloat | typing.SupportsIndex = 0.05999999865889549" in s  # 0.06f
    assert "threshold: typing.SupportsFloat | typing.SupportsIndex = -1.0" in _sig(w.recover_highlights)
    assert w.get_threshold() == 1.0  # no RAW loaded: the reference's fallback


@pytest.mark.parametrize("name", ["tone_mapping", "defringe", "enhance_micro_contrast", "recover_highlights"])
@pytest.mark.parametrize("shape", [(8, 8), (8, 8, 4), (8, 8, 3, 1), (0, 8, 3)])
def test_bad_shape_raises_invalid_argument(name, shape):
    w = core.LibRawWrapper()
    with pytest.raises(core.LibRawInvalidArgument):
        getattr(w, name)(np.zeros(shape, np.float32))


def test_exceptions_exported():
    assert issubclass(core.LibRawInvalidArgument, Exception)
    assert issubclass(core.LibRawError, Exception)
    assert issubclass(core.LibRawOutOfRange, Exception)
