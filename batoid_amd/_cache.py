"""Mixin for objects that describe themselves as descriptor records.

``_records()`` returns (records, arrays) on the host; ``_dev`` uploads them
once to device memory and keeps them for the object's lifetime (the role of
the reference's ``getDevPtr()`` object cloning, e.g. src/plane.cpp:37-51).
Pickling drops the device copy."""


class DeviceCached:
    def _records(self):
        raise NotImplementedError

    @property
    def _dev(self):
        d = self.__dict__.get("_dev_cache")
        if d is None:
            from ._device import Materialized
            recs, arrs = self._records()
            d = Materialized(recs, arrs)
            self.__dict__["_dev_cache"] = d
        return d

    def __getstate__(self):
        return {k: v for k, v in self.__dict__.items() if not k.startswith("_dev")}

    def __setstate__(self, state):
        self.__dict__.update(state)

    def __ne__(self, rhs):
        return not (self == rhs)
