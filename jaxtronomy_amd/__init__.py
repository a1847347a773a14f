"""MI355X-native strong-lensing image simulation (ray shooting -> Sersic light ->
supersample pooling -> PSF convolution -> chi^2 likelihood) behind the
JAXtronomy LensModel / LightModel / ImageModel interface.

Compute runs in hand-written HIP kernels for gfx950 (``_lib/liblensim.so``,
C ABI in ``include/lensim.h``); torch provides device buffers and streams.
"""

__version__ = "0.1.0"


def __getattr__(name):
    # lazy imports keep ``import jaxtronomy_amd`` cheap and free of device work
    if name == "LensModel":
        from .LensModel.lens_model import LensModel
        return LensModel
    if name == "LightModel":
        from .LightModel.light_model import LightModel
        return LightModel
    if name == "ImageModel":
        from .ImSim.image_model import ImageModel
        return ImageModel
    if name == "ImageData":
        from .Data.imaging_data import ImageData
        return ImageData
    if name == "PSF":
        from .Data.psf import PSF
        return PSF
    raise AttributeError(name)
