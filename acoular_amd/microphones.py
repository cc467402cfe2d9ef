"""Microphone array geometry (``acoular/microphones.py:34-275``)."""

from __future__ import annotations

import numpy as np

from ._base import HasParams


class MicGeom(HasParams):
    """Microphone positions; ``pos`` drops the ``invalid_channels`` (``microphones.py:175-180``)."""

    _params = ('pos_total', 'invalid_channels')

    def __init__(self, **kwargs):
        self.pos_total = np.zeros((3, 0))
        self.invalid_channels = []
        super().__init__(**kwargs)

    def __setattr__(self, name, value):
        if name == 'pos_total':
            value = np.asarray(value, dtype=float)
            if value.ndim != 2 or value.shape[0] != 3:
                msg = 'pos_total must have shape (3, num_mics)'
                raise ValueError(msg)
        object.__setattr__(self, name, value)

    @property
    def pos(self) -> np.ndarray:
        if len(self.invalid_channels) == 0:
            return self.pos_total
        keep = [i for i in range(self.pos_total.shape[-1]) if i not in self.invalid_channels]
        return self.pos_total[:, np.array(keep, dtype=int)]

    @property
    def num_mics(self) -> int:
        return self.pos.shape[-1]

    @property
    def center(self):
        """Mean position of the valid microphones (``microphones.py:186-193``)."""
        if self.pos.any():
            c = np.average(self.pos, axis=1)
            c[abs(c) < 1e-16] = 0.0
            return c
        return None

    @property
    def aperture(self):
        if self.pos.any():
            d = self.pos[:, :, None] - self.pos[:, None, :]
            return float(np.sqrt((d**2).sum(0)).max())
        return None
