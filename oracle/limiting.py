"""SPICE voltage limiting for Newton iterations (oracle).

Restated from vajax/analysis/limiting.py:

* ``pnjlim`` (limiting.py:22-82, VACASK DEVpnjlim): when vnew > vcrit and the
  step exceeds 2*vt, logarithmic compression -- from a forward-biased old
  value vold + vt*log(1 + d/vt) for a rise, vold - vt*log(1 - d/vt) for a fall;
  from vold <= 0, vt*log(vnew/vt) if vnew > 0 else vcrit; otherwise for
  vnew < 0 a clamp to max(vnew, -vold - 1) (vold > 0) or max(vnew, 2*vold - 1);
  logs take max(., 1e-30);
* ``fetlim`` (limiting.py:85-107): vold + clip(vnew - vold, -s, s) with
  s = 2*|vold - vto| + 2 when vold >= vto, else 0.5.

Elementwise on numpy arrays (or scalars), each branch as written there.
"""

from __future__ import annotations

import numpy as np


def pnjlim(vnew, vold, vt, vcrit):
    vnew = np.asarray(vnew, dtype=np.float64)
    vold = np.asarray(vold, dtype=np.float64)
    delta = vnew - vold
    arg = delta / vt
    fwd = (vnew > vcrit) & (np.abs(delta) > 2 * vt)
    with np.errstate(divide="ignore", invalid="ignore"):
        up = vold + vt * np.log(np.maximum(1.0 + arg, 1e-30))
        down = vold - vt * np.log(np.maximum(1.0 - arg, 1e-30))
        from_off = np.where(vnew > 0, vt * np.log(np.maximum(vnew / vt, 1e-30)), vcrit)
    forward = np.where(arg >= 0, up, down)
    limited_fwd = np.where(vold > 0, forward, from_off)
    clamp = np.where(vold > 0, -vold - 1.0, 2.0 * vold - 1.0)
    limited_neg = np.maximum(vnew, clamp)
    return np.where(fwd, limited_fwd, np.where(vnew < 0, limited_neg, vnew))


def fetlim(vnew, vold, vto):
    vnew = np.asarray(vnew, dtype=np.float64)
    vold = np.asarray(vold, dtype=np.float64)
    step = np.where(vold >= vto, 2 * np.abs(vold - vto) + 2, 0.5)
    return vold + np.clip(vnew - vold, -step, step)
