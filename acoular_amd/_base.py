"""Minimal parameter/caching base used instead of ``traits`` (which the reference builds on).

Objects keep their public parameters as plain attributes; derived values are cached per
``digest``, a hash of the parameters and of the digests of the objects they depend on, in the
spirit of Acoular's ``digest`` properties (``acoular/internal.py:29-40``).
"""

from __future__ import annotations

import hashlib

import numpy as np


def digest_of(*items) -> str:
    h = hashlib.blake2b(digest_size=16)
    for it in items:
        if isinstance(it, HasParams):
            h.update(it.digest.encode())
        elif isinstance(it, np.ndarray):
            h.update(str((it.dtype, it.shape)).encode())
            h.update(np.ascontiguousarray(it).tobytes())
        else:
            h.update(repr(it).encode())
        h.update(b'|')
    return h.hexdigest()


class HasParams:
    """Keyword-constructible object with a per-digest cache of derived values."""

    _params: tuple = ()

    def __init__(self, **kwargs):
        self.__dict__['_cache'] = {}
        for name, value in kwargs.items():
            if name not in self._params and not isinstance(getattr(type(self), name, None), property):
                msg = f'{type(self).__name__} has no parameter {name!r}'
                raise TypeError(msg)
            setattr(self, name, value)

    def _digest_items(self):
        return tuple(getattr(self, p) for p in self._params)

    @property
    def digest(self) -> str:
        return digest_of(type(self).__name__, *self._digest_items())

    def _cached(self, key, fn):
        d = self.digest
        hit = self._cache.get(key)
        if hit is not None and hit[0] == d:
            return hit[1]
        val = fn()
        self._cache[key] = (d, val)
        return val
