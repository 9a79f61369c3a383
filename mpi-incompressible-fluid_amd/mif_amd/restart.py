"""Text format of the reference's restart files (host-side, no device).

Tensor::dump (include/Tensor.h:465-476) streams a tensor through its
operator<< (:488-505): the line "Tensor (row major). Dimensions: [ sx sy sz ]",
then "[ ", every value in storage order (x fastest) followed by a space, and
"]".  Values print with the stream's default formatting, which is printf's
%g (6 significant digits).
"""
from __future__ import annotations

import numpy as np


def format_text_dump(values: np.ndarray) -> str:
    """The dump text of a field given as a (sz, sy, sx) array in storage order."""
    sz, sy, sx = values.shape
    body = "".join("%g " % v for v in np.asarray(values, dtype=np.float64).ravel())
    return "Tensor (row major). Dimensions: [ %d %d %d ]\n[ %s]\n" % (sx, sy, sz, body)
