"""Time-data source used by :class:`~acoular_amd.spectra.PowerSpectra` (``acoular/sources.py:299``).

``TimeSamples`` holds a ``(num_samples, num_channels)`` array and yields it block-wise like the
reference's ``result(num)`` generator (``sources.py:417-470``), the last block possibly shorter.
"""

from __future__ import annotations

import numpy as np

from ._base import HasParams, digest_of


class TimeSamples(HasParams):
    _params = ('data', 'sample_freq')

    def __init__(self, **kwargs):
        self.data = np.zeros((0, 0))
        self.sample_freq = 1.0
        super().__init__(**kwargs)

    def __setattr__(self, name, value):
        if name == 'data':
            value = np.asarray(value)
            if value.ndim != 2:
                msg = 'data must have shape (num_samples, num_channels)'
                raise ValueError(msg)
        if name == 'sample_freq':
            value = float(value)
        object.__setattr__(self, name, value)

    @property
    def digest(self) -> str:
        # the sample array is identified by object identity and shape, not hashed (it can be GBs)
        return digest_of('TimeSamples', id(self.data), self.data.shape, self.data.dtype.str, self.sample_freq)

    @property
    def num_samples(self) -> int:
        return self.data.shape[0]

    @property
    def num_channels(self) -> int:
        return self.data.shape[1]

    def result(self, num=128):
        if self.num_samples == 0:
            msg = 'no samples available'
            raise OSError(msg)
        for i in range(0, self.num_samples, num):
            yield self.data[i : i + num]
