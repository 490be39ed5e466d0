"""Per-type (or per-type-pair) parameter dictionaries with a schema.

Stands in for the host engine's ``TypeParameterDict`` as the reference uses it
(reference oxdna/forces.py:33-85, 131-144): keys are one type name (``len_keys=1``)
or an unordered pair of type names (``len_keys=2``; ``(a, b)`` and ``(b, a)`` are
the same entry); values are dicts validated against a schema of
``float`` / ``int`` / ``bool`` / tuples of those, with optional defaults. A key can
also be a list of names (or a pair of lists), which sets every combination.
"""

import itertools
import numbers


def _coerce(spec, value, name):
    if isinstance(spec, tuple):
        try:
            seq = tuple(value)
        except TypeError:
            raise TypeError(f"parameter '{name}' must be a sequence of length {len(spec)}")
        if len(seq) != len(spec):
            raise ValueError(f"parameter '{name}' must have length {len(spec)}, got {len(seq)}")
        return tuple(_coerce(s, v, name) for s, v in zip(spec, seq))
    if spec is bool:
        if not isinstance(value, (bool, numbers.Integral)):
            raise TypeError(f"parameter '{name}' must be a bool")
        return bool(value)
    if spec is int:
        if isinstance(value, bool) or not isinstance(value, numbers.Integral):
            raise TypeError(f"parameter '{name}' must be an int")
        return int(value)
    if spec is float:
        if not isinstance(value, numbers.Real):
            raise TypeError(f"parameter '{name}' must be a real number")
        return float(value)
    raise TypeError(f"bad schema entry for '{name}'")


class TypeParameterDict:
    """Schema-checked mapping from type keys to parameter dicts.

    ``schema`` maps parameter names to specs; a bare spec (``float``) as schema makes
    a scalar-valued dict (used for ``r_cut``).
    """

    def __init__(self, schema=None, len_keys=1, defaults=None, scalar=None, default_value=None):
        self._schema = dict(schema or {})
        self._len = int(len_keys)
        self._defaults = dict(defaults or {})
        self._scalar = scalar
        self._default_value = default_value
        self._data = {}

    # -- keys -------------------------------------------------------------
    def _norm_key(self, key):
        if self._len == 1:
            if isinstance(key, tuple) and len(key) == 1:
                key = key[0]
            if not isinstance(key, str):
                raise KeyError(f"expected a type name, got {key!r}")
            return key
        if not (isinstance(key, tuple) and len(key) == 2 and all(isinstance(k, str) for k in key)):
            raise KeyError(f"expected a pair of type names, got {key!r}")
        return tuple(sorted(key))

    def _expand(self, key):
        if self._len == 1:
            if isinstance(key, (list, set)):
                return [self._norm_key(k) for k in key]
            return [self._norm_key(key)]
        if isinstance(key, tuple) and len(key) == 2:
            a, b = key
            la = list(a) if isinstance(a, (list, set)) else [a]
            lb = list(b) if isinstance(b, (list, set)) else [b]
            return [self._norm_key((x, y)) for x, y in itertools.product(la, lb)]
        if isinstance(key, list):
            return [k for sub in key for k in self._expand(sub)]
        raise KeyError(f"expected a pair of type names, got {key!r}")

    # -- values -----------------------------------------------------------
    def _validate(self, value):
        if self._scalar is not None:
            return _coerce(self._scalar, value, "value")
        if not isinstance(value, dict):
            raise TypeError("parameters must be given as a dict")
        unknown = set(value) - set(self._schema)
        if unknown:
            raise KeyError(f"unknown parameter(s): {sorted(unknown)}")
        out = {}
        for name, spec in self._schema.items():
            if name in value:
                out[name] = _coerce(spec, value[name], name)
            elif name in self._defaults:
                out[name] = _coerce(spec, self._defaults[name], name)
            else:
                raise KeyError(f"missing required parameter '{name}'")
        return out

    def __setitem__(self, key, value):
        v = self._validate(value)
        for k in self._expand(key):
            self._data[k] = dict(v) if isinstance(v, dict) else v

    def __getitem__(self, key):
        k = self._norm_key(key)
        if k in self._data:
            return self._data[k]
        if self._scalar is not None and self._default_value is not None:
            return self._default_value
        raise KeyError(f"no parameters set for {key!r}")

    def __contains__(self, key):
        try:
            k = self._norm_key(key)
        except KeyError:
            return False
        return k in self._data or (self._scalar is not None and self._default_value is not None)

    def keys(self):
        return list(self._data.keys())

    def items(self):
        return list(self._data.items())

    def __len__(self):
        return len(self._data)

    def require(self, types):
        """Raise if any needed key is missing for the given type names."""
        if self._len == 1:
            keys = [t for t in types]
        else:
            keys = [tuple(sorted(p)) for p in itertools.combinations_with_replacement(types, 2)]
        for k in keys:
            if k not in self._data and not (self._scalar is not None and self._default_value is not None):
                raise KeyError(f"parameters for {k!r} are not set")
