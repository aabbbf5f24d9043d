"""Oracle (test infrastructure only): restatement of the reference's PNM
snapshot colouring, seismic_CPML_2D_isotropic_second_order.f90:861-999
(create_color_image).  Used by the tests as the checker of the HIP image
kernel; never called by the library.

Pixels run top row first (Fortran iy = NY down to 1), x left to right.  Per
pixel, first match wins:
  * source cross (|ix - ISOURCE| <= 5 and |iy - JSOURCE| <= 1, or the
    transposed bar): orange 255 157 0;
  * two-pixel frame (ix <= 2, ix >= NX-1, iy <= 2, iy >= NY-1): black;
  * the inner edge of each PML layer (ix == NPOINTS_PML, ix == NX-NPOINTS_PML,
    iy == NPOINTS_PML, iy == NY-NPOINTS_PML, for the layers in use): 255 150 0;
  * |v| <= 0.01 max|v|: white;
  * v / max|v| clamped to [-1, 1], red nint(255 x^0.3) if >= 0, else blue
    nint(255 |x|^0.3);
then a receiver square (|ix - ix_rec| <= 3 and |iy - iy_rec| <= 3) paints
dark green 30 180 60 over any of them.  Arithmetic in the precision of the
field (the single-precision build of the reference computes the image in
real(4)); nint rounds half away from zero (the values are non-negative).
"""
from __future__ import annotations

import numpy as np

POWER_DISPLAY = 0.30
CUTVECT = 0.01
WIDTH_CROSS, THICKNESS_CROSS, SIZE_SQUARE = 5, 1, 3


def color_image(field, isrc, jsrc, ix_rec, iy_rec, npml, use_pml=(True, True, True, True),
                ghost=0):
    """field: (NY, NX) array, row j = Fortran iy - 1.  isrc, jsrc, ix_rec,
    iy_rec: 1-based Fortran indices.  use_pml: (xmin, xmax, ymin, ymax).
    ghost = 1: the image is the ghost-padded (NY+2, NX+2) array that the
    4th-order program passes (seismic_CPML_2D_isotropic_fourth_order.f90:
    729-731), zero on the border, with the source, receivers and PML lines at
    their interior indices.  Returns uint8 (rows, columns, 3), top row first."""
    v = np.asarray(field)
    R = np.float32 if v.dtype == np.float32 else np.float64
    v = v.astype(R)
    if ghost:
        v = np.pad(v, 1)
    ny, nx = v.shape
    amax = np.abs(v).max()
    iy, ix = np.meshgrid(np.arange(ny, 0, -1), np.arange(1, nx + 1), indexing="ij")
    vv = v[::-1, :]  # top row first
    with np.errstate(divide="ignore", invalid="ignore"):
        nv = vv / amax
    nv = np.where(nv < R(-1), R(-1), nv)
    nv = np.where(nv > R(1), R(1), nv)
    rgb = np.zeros((ny, nx, 3), np.int64)
    x = (R(255) * np.abs(nv) ** R(POWER_DISPLAY)).astype(np.float64)
    fl = np.floor(x)
    red = fl + (x - fl >= 0.5)  # nint, half away from zero (x >= 0)
    pos = nv >= 0
    rgb[..., 0] = np.where(pos, red, 0)
    rgb[..., 2] = np.where(pos, 0, red)
    below = np.abs(vv) <= amax * R(CUTVECT)
    rgb[below] = (255, 255, 255)
    xmin, xmax, ymin, ymax = use_pml
    edge = ((xmin & (ix == npml)) | (xmax & (ix == nx - npml)) | (ymin & (iy == npml))
            | (ymax & (iy == ny - npml)))
    rgb[edge] = (255, 150, 0)
    frame = (ix <= 2) | (ix >= nx - 1) | (iy <= 2) | (iy >= ny - 1)
    rgb[frame] = (0, 0, 0)
    cross = (((np.abs(ix - isrc) <= WIDTH_CROSS) & (np.abs(iy - jsrc) <= THICKNESS_CROSS))
             | ((np.abs(ix - isrc) <= THICKNESS_CROSS) & (np.abs(iy - jsrc) <= WIDTH_CROSS)))
    rgb[cross] = (255, 157, 0)
    for xr, yr in zip(np.asarray(ix_rec).ravel(), np.asarray(iy_rec).ravel()):
        sq = (np.abs(ix - xr) <= SIZE_SQUARE) & (np.abs(iy - yr) <= SIZE_SQUARE)
        rgb[sq] = (30, 180, 60)
    return rgb.astype(np.uint8)
