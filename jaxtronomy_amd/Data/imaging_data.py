"""ImageData: pixel grid + data image + noise model, held on the device.

Mirrors the reference's ``ImageData`` (Data/imaging_data.py:50-133) together
with the lenstronomy ``PixelGrid`` it derives from and ``ImageNoise``
(Data/image_noise.py:16-97).  The data, noise map, exposure map (floored at
1e-10 as at image_noise.py:52-54) live in device memory from construction on.
"""

import numpy as np
import torch

from .. import backend


class ImageData(object):
    def __init__(self, image_data, exposure_time=None, background_rms=None, noise_map=None,
                 gradient_boost_factor=None, ra_at_xy_0=0, dec_at_xy_0=0, transform_pix2angle=None, ra_shift=0,
                 dec_shift=0, phi_rot=0, log_likelihood_constant=0, antenna_primary_beam=None,
                 likelihood_method="diagonal", flux_scaling=1):
        if antenna_primary_beam is not None:
            raise ValueError("primary beam is not supported")
        if gradient_boost_factor is not None:
            raise ValueError("gradient_boost_factor not supported")
        if likelihood_method != "diagonal":
            raise ValueError("likelihood_method %s not supported! likelihood_method can only be 'diagonal'"
                             % likelihood_method)
        image_data = np.asarray(image_data, dtype=np.float64)
        nx, ny = np.shape(image_data)  # the reference's convention (imaging_data.py:96)
        if transform_pix2angle is None:
            transform_pix2angle = np.array([[1, 0], [0, 1]])
        cos_phi, sin_phi = np.cos(phi_rot), np.sin(phi_rot)
        rot = np.array([[cos_phi, -sin_phi], [sin_phi, cos_phi]])
        self._nx, self._ny = int(nx), int(ny)
        self.transform_pix2angle = np.dot(np.asarray(transform_pix2angle, dtype=np.float64), rot)
        self._ra_at_xy_0 = float(ra_at_xy_0 + ra_shift)
        self._dec_at_xy_0 = float(dec_at_xy_0 + dec_shift)
        self._logL_constant = log_likelihood_constant
        self._logL_method = likelihood_method
        self.flux_scaling = flux_scaling

        # ImageNoise (image_noise.py:44-92)
        if exposure_time is None:
            if noise_map is None:
                raise ValueError("Exposure map has not been specified in Noise() class!")
            self.exposure_map = None
        else:
            exposure_time = np.asarray(exposure_time, dtype=np.float64)
            self.exposure_map = np.where(exposure_time <= 10 ** (-10), 10 ** (-10), exposure_time)
        if background_rms is None:
            if noise_map is None:
                raise ValueError("rms background value as 'background_rms' not specified!")
            self.background_rms = float(np.median(noise_map))
        else:
            self.background_rms = float(background_rms)
        self.data = image_data
        if noise_map is not None:
            assert np.shape(noise_map) == np.shape(image_data)
            self._noise_map = np.asarray(noise_map, dtype=np.float64)
        else:
            self._noise_map = None
        self._device_cache = None

    # ---- PixelGrid -----------------------------------------------------------
    @property
    def num_pixel_axes(self):
        return self._nx, self._ny

    @property
    def num_pixel(self):
        return self._nx * self._ny

    @property
    def pixel_width(self):
        """sqrt(|det M|) of the pixel-to-angle transform (lenstronomy PixelGrid)."""
        return float(np.sqrt(abs(np.linalg.det(self.transform_pix2angle))))

    @property
    def radec_at_xy_0(self):
        return self._ra_at_xy_0, self._dec_at_xy_0

    def map_pix2coord(self, x, y):
        M = self.transform_pix2angle
        return M[0, 0] * x + M[0, 1] * y + self._ra_at_xy_0, M[1, 0] * x + M[1, 1] * y + self._dec_at_xy_0

    @property
    def center(self):
        return self.map_pix2coord((self._nx - 1) / 2.0, (self._ny - 1) / 2.0)

    @property
    def width(self):
        return self.pixel_width * np.array([self._nx, self._ny])

    @property
    def primary_beam(self):
        return None

    # ---- ImageNoise ----------------------------------------------------------
    @property
    def noise_map(self):
        return self._noise_map

    @property
    def C_D(self):
        """Covariance diagonal from the data, image_noise.py:82-92."""
        if self._noise_map is not None:
            return self._noise_map ** 2
        d_pos = np.where(self.data >= 0, self.data, 0)
        return d_pos / self.exposure_map + self.background_rms ** 2

    def device_buffers(self):
        """Device copies of data / noise map / exposure map (created once)."""
        if self._device_cache is None:
            dev = backend.device()
            data = torch.as_tensor(np.ascontiguousarray(self.data), device=dev)
            noise = None if self._noise_map is None else torch.as_tensor(np.ascontiguousarray(self._noise_map),
                                                                         device=dev)
            expo_map, expo_time = None, 0.0
            if self.exposure_map is not None:
                if np.ndim(self.exposure_map) == 0:
                    expo_time = float(self.exposure_map)
                else:
                    expo_map = torch.as_tensor(np.ascontiguousarray(np.broadcast_to(self.exposure_map,
                                                                                    self.data.shape)),
                                               device=dev, dtype=torch.float64)
            self._device_cache = dict(data=data, noise=noise, expo_map=expo_map, expo_time=expo_time)
        return self._device_cache

    def update_data(self, image_data):
        """imaging_data.py:135-157: replaces the data (same shape)."""
        image_data = np.asarray(image_data, dtype=np.float64)
        if image_data.shape != self.data.shape:
            raise ValueError("New data shape %s should match old data shape %s" % (image_data.shape,
                                                                                   self.data.shape))
        self.data = image_data
        self._device_cache = None
