"""CPU restatement of the depth-layer stitch of angle gathers.  TEST
INFRASTRUCTURE ONLY: only tests/ import it, to check kmig_batch_stitch.

Follows src/layercom_cig.cpp: the lateral edge scale (:445-456), the
overlap / keep_upper / blend bookkeeping (:484-492) and the per-sample loop
(:508-545), in float32 where the reference computes in float (the scaled
value, the blend) and double where it computes in double (the weight
0.5 (1 - cos(pi u)) and the lateral ramp).  Layers are (ng, nx, n1) arrays
(RSF (z, x, angle) with axis 1 fastest), top to bottom.
"""

from __future__ import annotations

import math

import numpy as np


def lateral_scale(nx, scalex_aper, scalex_scale):
    lat = np.ones(nx, np.float32)
    if scalex_aper > 0:
        for i2 in range(nx):
            edge = min(i2, nx - 1 - i2)
            if edge < scalex_aper:
                taper = 1.0 if scalex_aper == 1 else 1.0 - edge / (scalex_aper - 1)
                lat[i2] = np.float32(1.0 + (scalex_scale - 1.0) * taper)
    return lat


def stitch(layers, start, end, scale=None, stitch_blend=15, scalex_aper=0, scalex_scale=1.0,
           angles=None):
    ng, nx, _ = layers[0].shape
    full_n1 = end[-1] + 1
    lat = lateral_scale(nx, scalex_aper, scalex_scale)
    lscale = [np.float32(1.0)] * len(layers) if scale is None else [np.float32(v) for v in scale]
    out = np.zeros((ng, nx, full_n1), np.float32)
    previous_end = -1
    nonfinite = 0
    for il, lay in enumerate(layers):
        s0, e0 = start[il], end[il]
        n1 = e0 - s0 + 1
        assert lay.shape == (ng, nx, n1)
        sx = lat if il > 0 else np.ones(nx, np.float32)
        sc = (lscale[il] * sx).astype(np.float32)
        v = (lay.astype(np.float32) * sc[None, :, None]).astype(np.float32)
        bad = ~np.isfinite(v)
        nonfinite += int(bad.sum())
        v[bad] = 0.0
        overlap = previous_end - s0 + 1 if il > 0 else 0
        blend = overlap
        if overlap > 0 and stitch_blend >= 0:
            blend = min(stitch_blend, overlap)
        keep_upper = overlap - blend
        for k in range(n1):
            g1 = s0 + k
            if overlap > 0 and g1 <= previous_end:
                if k < keep_upper or blend == 0:
                    continue
                bi = k - keep_upper
                u = bi / (blend - 1) if blend > 1 else 1.0
                w = np.float32(0.5 * (1.0 - math.cos(math.pi * u)))
                out[:, :, g1] = out[:, :, g1] * (np.float32(1.0) - w) + v[:, :, k] * w
            else:
                out[:, :, g1] = v[:, :, k]
        previous_end = e0
    if angles is not None:
        g0, gn = angles
        out = out[g0:g0 + gn]
    return out, nonfinite
