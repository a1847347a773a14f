"""Helpers to set up simulated data (lenstronomy SimulationAPI conventions used
by the reference's tests, e.g. test/test_ImSim/test_image_model.py:33-37)."""

import numpy as np


def make_grid_with_coordtransform(numPix, deltapix, center_ra=0, center_dec=0, inverse=True):
    """Pixel-to-angle matrix and the coordinates of pixel (0, 0) of a numPix^2 grid
    centred on (center_ra, center_dec); inverse=True makes RA grow to the left."""
    sign = -1.0 if inverse else 1.0
    M = np.array([[sign * deltapix, 0.0], [0.0, deltapix]])
    c = (numPix - 1) / 2.0
    ra_at_xy_0 = center_ra - (M[0, 0] * c + M[0, 1] * c)
    dec_at_xy_0 = center_dec - (M[1, 0] * c + M[1, 1] * c)
    return ra_at_xy_0, dec_at_xy_0, M


def data_configure_simple(numPix, deltaPix, exposure_time=None, background_rms=None, center_ra=0, center_dec=0,
                          inverse=False):
    """kwargs for ImageData of an empty numPix^2 image."""
    ra0, dec0, M = make_grid_with_coordtransform(numPix, deltaPix, center_ra, center_dec, inverse)
    return {
        "background_rms": background_rms,
        "exposure_time": exposure_time,
        "ra_at_xy_0": ra0,
        "dec_at_xy_0": dec0,
        "transform_pix2angle": M,
        "image_data": np.zeros((numPix, numPix)),
    }
